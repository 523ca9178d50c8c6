// k_query_mfma_inst.hip -- instantiations of the fused MFMA query kernel,
// built once per (operand type, lobe family, numerics) with
//   -DNPMG_INST_{bf16,f16,fp8,bf16x3} -DNPMG_INST_{vmf,asg} -DNPMG_INST_{exact,fast}
// so that the sixteen objects compile in parallel.
#include "query_mfma.h"

namespace npmg {

#if defined(NPMG_INST_bf16)
#define NPMG_TYPES __bf16, __bf16
#elif defined(NPMG_INST_f16)
#define NPMG_TYPES _Float16, _Float16
#elif defined(NPMG_INST_fp8)
#define NPMG_TYPES __bf16, mf::Fp8  // BF16 grid table and features, E4M3 MLP operands
#elif defined(NPMG_INST_bf16x3)
#define NPMG_TYPES _Float16, mf::Bf16x3  // FP16 grid table and features, split-BF16 MLP operands
#define NPMG_NARROW_ONLY                 // two weight planes of a 128-wide network exceed the LDS
#else
#error "operand type: define NPMG_INST_bf16, NPMG_INST_f16, NPMG_INST_fp8 or NPMG_INST_bf16x3"
#endif

#if defined(NPMG_INST_fast)
#define NPMG_FAST true
#elif defined(NPMG_INST_exact)
#define NPMG_FAST false
#else
#error "numerics: define NPMG_INST_exact or NPMG_INST_fast"
#endif

#define NPMG_INST(WW, KK, PP)                                                                                      \
    template hipError_t launch_mfma_t<NPMG_TYPES, WW, KK, PP, NPMG_FAST>(                                          \
        const DevCfg&, const void*, const void*, const void*, const QueryIO&, uint32_t, uint32_t, uint32_t, int, \
        int, const float*, hipStream_t);                                                                           \
    template hipError_t launch_mfma_queues_t<NPMG_TYPES, WW, KK, PP, NPMG_FAST>(                                   \
        const DevCfg&, const void*, const void*, const QueueSet&, int, int, const float*, hipStream_t);

#if defined(NPMG_INST_vmf)
NPMG_INST(64, 7, 4)
NPMG_INST(64, 8, 4)
NPMG_INST(64, 16, 4)
#if !defined(NPMG_NARROW_ONLY)
NPMG_INST(128, 7, 4)
NPMG_INST(128, 8, 4)
NPMG_INST(128, 16, 4)
#endif
#elif defined(NPMG_INST_asg)
NPMG_INST(64, 16, 6)
#if !defined(NPMG_NARROW_ONLY)
NPMG_INST(128, 16, 6)
#endif
#else
#error "lobe family: define NPMG_INST_vmf or NPMG_INST_asg"
#endif

}  // namespace npmg
