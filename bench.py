"""Benchmark of the guiding hot path on one MI355X.

Headline precision "bf16x3": BF16 MFMA on split operands (hi + lo, three
products, FP32 accumulation) with FP16 tables, the mode whose pdfs stay within
1e-3 of FP64; plain BF16, FP16, FP8 and FP32 are timed beside it.

Workload (full-frame wavefront size): hash grid 16 levels x 2^19 entries x 2
features, OneBlob(16) over the 6 direction/normal/roughness inputs, MLP
3 x 64 (bias-free, ReLU), K = 8 vMF lobes.
  * infer + sample + pdf: 2^21 queries per step, each query encoded, run
    through the network, mapped to its mixture (with the cosine product),
    sampled once and evaluated at the sample and at 4 given directions;
  * train step: 2^18 samples (forward, KL adjoint, MLP + grid backward, Adam).
Inputs are synthetic (pcg32 keyed by query index) and resident in HBM before
the timed region. One step of the headline metric is one launch of the fused
query kernel, timed with HIP events on the stream the library launches on.

Prints one JSON line (see the repository's bench contract).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "neural-mixture-guiding_amd"))

METRIC = "guided queries/s (infer+sample+pdf) & train samples/s at 2M batch, BF16 vs FP32"
N_QUERIES = 1 << 21
N_TRAIN = 1 << 18
FRAME_PATHS = 1920 * 1080  # one wavefront bounce of a 1920x1080 frame
FRAME_BOUNCES = 8
HBM_PEAK_GBS = 8000.0
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4  # wave-instructions/s: 256 CUs x 4 SIMDs, 64 lanes over 16-wide SIMDs
FP32_PEAK_TFLOPS = 157.3
MFMA_BF16_PEAK_TFLOPS = 2500.0
# in, u, normal, 4 eval dirs; dir, pdf, 4 pdfs, lobe
BYTES_PER_QUERY = 36 + 12 + 12 + 48 + 12 + 4 + 16 + 4
GRID_FEATURES = 16 * 2  # levels x features of the full configuration
PRECISION_NOTE = {
    "fp32": "FP32 VALU, bit-exact with the FP32 oracle",
    "bf16": "BF16 MFMA operands, FP32 accumulation",
    "fp16": "FP16 MFMA operands, FP32 accumulation",
    "bf16x3": "BF16 MFMA on split operands (hi = bf16(x), lo = bf16(x - hi)), three products per tile "
              "(hi hi + hi lo + lo hi), FP32 accumulation; FP16 grid table and level-major features",
}


def profiled(precision, key, queries_per_launch, kernel="k_query"):
    """A per-launch figure of `kernel` (k_query: the fused query kernel;
    k_grid_encode: the level-major grid encode) from the committed rocprofv3
    counter summary: "<key>_per_query" scaled to the launch, or "<key>_per_launch"
    as written; None when the summary lacks it."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{precision}_full_counters.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + "_per_query:"):
            return float(line.split(":", 1)[1]) * queries_per_launch
        if line.startswith(key + "_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_region(fn, steps, warmup, stream):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    end.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(end) / steps  # ms per step


def _host_info():
    """Host cores available to this process and the CPU model (nproc / lscpu)."""
    import subprocess

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", avail) or avail)
    model = None
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"nproc": avail, "os_cpu_count": os.cpu_count(), "omp_num_threads": share, "model": model,
            "threads_used": max(1, min(avail, share))}


def cpu_baseline(cfg, params, q_host, train_host):
    """The oracle's FP32 path (the restatement of the reference's algorithm,
    stable numerics) on the host cores, in the same run: infer + sample + pdf
    (cosine product, 1 sample, 4 evaluation pdfs) on 65,536 queries (1 warm-up,
    median of 5) and on the 2^21 benchmark queries (median of 3), and one train
    step (forward, adjoint, MLP and grid backward, Adam over all parameters)
    on the 2^18 benchmark records (1 warm-up on 2^14, median of 3). Rates per
    query / per record."""
    import statistics

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    o = pyoracle.Oracle()
    oc = pyoracle.make_cfg(cfg)
    info = _host_info()
    th = info["threads_used"]

    def query_rate(n, runs):
        args = [np.ascontiguousarray(q_host[k][..., :n]) for k in ("inputs", "normal", "u", "eval_dirs")]
        times = []
        for _ in range(runs):
            t0 = time.perf_counter()
            o.query(oc, params, *args, n_threads=th)
            times.append(time.perf_counter() - t0)
        return n / statistics.median(times), times

    query_rate(4096, 1)  # warm-up
    r64, t64 = query_rate(65536, 5)
    r21, t21 = query_rate(N_QUERIES, 3)

    def train_rate(n, runs):
        t = {k: np.ascontiguousarray(v[..., :n]) for k, v in train_host.items()}
        times = []
        for _ in range(runs):
            p = params.copy()
            m, v = np.zeros_like(p), np.zeros_like(p)
            st = np.zeros(p.size, np.uint32)
            t0 = time.perf_counter()
            o.train_step(oc, p, m, v, st, t, n_threads=th)
            times.append(time.perf_counter() - t0)
        return n / statistics.median(times), times

    train_rate(1 << 14, 1)  # warm-up
    rt, tt = train_rate(N_TRAIN, 3)
    return {"value": r21, "unit": "queries/s", "cores": th, "kind": "port",
            "sample": (f"oracle FP32 (stable numerics) infer+sample+pdf on the 2^21 benchmark queries, "
                       f"median of 3 runs ({', '.join(f'{x:.2f}' for x in t21)} s) on {th} threads"),
            "queries_per_s_65536": r64, "queries_per_s_2e21": r21, "train_samples_per_s": rt,
            "train_sample": (f"one FP32 train step (forward, adjoint, MLP and grid backward, Adam over "
                             f"{params.size} parameters) on the 2^18 benchmark records, median of 3 "
                             f"({', '.join(f'{x:.2f}' for x in tt)} s)"),
            "host": info}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="bf16x3", choices=["fp32", "bf16", "fp16", "bf16x3"],
                    help="headline precision; bf16x3: BF16 MFMA on split (hi + lo) operands, three products, "
                         "the mode that meets the 1e-3 pdf error bar")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", default="all", choices=["all", "query", "frame", "train"],
                    help="query: time only the fused query kernel (for counter runs); "
                         "frame: only the anisotropic 8-bounce frame; train: only the train step of --precision")
    ap.add_argument("--tuning", default=None, metavar="CHUNKS,BLOCKS_PER_CU",
                    help="use this query launch shape instead of autotuning (counter runs: one launch "
                         "of each kernel per step)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a plain 1 GiB device copy 5 times (HBM-counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch

        src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the guiding path would shard a frame's queries by index, one parameter replica per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import npmg

    torch.cuda.set_device(0)
    stream = torch.cuda.current_stream()
    q = npmg.synth_queries(2024, N_QUERIES)
    out = {
        "dir": torch.empty(3, N_QUERIES, device="cuda"), "pdf": torch.empty(N_QUERIES, device="cuda"),
        "eval_pdf": torch.empty(4, N_QUERIES, device="cuda"), "lobe": torch.empty(N_QUERIES, dtype=torch.int32, device="cuda"),
    }
    lib = npmg.lib()
    P = npmg._ptr

    def full_cfg(precision, numerics="fast"):
        """The full configuration; "fp8" is BF16 training with FP8 E4M3 inference.
        Every precision's mixture stage uses the same numerics ("fast": hardware
        transcendentals; "stable": the bit-exact functions of the FP32 parity path)."""
        c = npmg.benchmark_config("full", precision="bf16" if precision == "fp8" else precision)
        if precision == "fp8":
            c["inference_precision"] = "fp8"
        c["mixture"] = dict(c["mixture"], numerics=numerics)
        return c

    def measure_queries(precision, numerics="fast"):
        net = npmg.GuidingNetwork(full_cfg(precision, numerics), seed=7272)
        # launch shape from the library's autotuner (pipeline chunks x workgroups per CU), untimed
        if args.tuning and precision != "fp32":
            net.set_tuning(*[int(v) for v in args.tuning.split(",")])
            tuned = net.tuning()
        else:
            tuned = net.autotune(N_QUERIES) if precision != "fp32" else net.tuning()

        def query_step():
            st = lib.npmg_sample_pdf(net._h, P(q["inputs"]), P(q["normal"]), P(q["u"]), P(q["eval_dirs"]), 4,
                                     P(out["dir"]), P(out["pdf"]), P(out["eval_pdf"]), P(out["lobe"]), None,
                                     N_QUERIES, npmg.NPMG_COSINE_AWARE, stream.cuda_stream)
            if st != 0:
                raise npmg.NpmgError(st, lib.npmg_last_error(net._h).decode())

        ms = time_region(query_step, args.steps, max(1, args.warmup), stream)  # FP8: the first call calibrates
        if numerics != "fast":
            log(f"{precision} ({numerics} numerics) infer+sample+pdf: {ms:.3f} ms per 2^21 queries")
            net.close()
            return None, ms
        log(f"{precision} infer+sample+pdf: {ms:.3f} ms per 2^21 queries ({N_QUERIES / ms / 1e3:.1f} M queries/s), "
            f"tuning {tuned}")
        # per-kernel launch durations: HIP events recorded by the library around each
        # launch, on the stream that launch runs on (a separate pass, not the timed one)
        net.timing(True)
        for _ in range(args.steps):
            query_step()
        kern = {}
        for tag, name in ((npmg.NPMG_TIME_QUERY, "query"), (npmg.NPMG_TIME_GRID_ENCODE, "grid_encode")):
            tot, cnt = net.timing_read(tag)
            if cnt:
                kern[name] = {"avg_ms_per_launch": tot / cnt, "launches_per_step": cnt / args.steps}
        net.timing(False)
        kern["tuning"] = {"pipeline_chunks": tuned[0], "blocks_per_cu": tuned[1]}
        per_kernel[precision] = kern
        return net, ms

    t = npmg.synth_train(77, N_TRAIN)
    loss = torch.empty(1, device="cuda")

    def measure_train(netp, precision, graph=False):  # precision: label for the log
        fn = lib.npmg_train_step_graph if graph else lib.npmg_train_step

        def train_step():
            st = fn(netp._h, P(t["inputs"]), P(t["tgt_dir"]), P(t["tgt_rgb"]), P(t["tgt_pdf"]), N_TRAIN, P(loss),
                    stream.cuda_stream)
            if st != 0:
                raise npmg.NpmgError(st, lib.npmg_last_error(netp._h).decode())

        ms = time_region(train_step, max(2, args.steps // 2), 2, stream)
        log(f"{precision} train step{' (graph)' if graph else ''}: {ms:.3f} ms per 2^18 samples "
            f"({N_TRAIN / ms / 1e3:.2f} M samples/s), loss {loss.item():.4f}")
        return ms

    def loss_after(precision, steps=500):
        """Train a fresh network for `steps` steps on fresh batches of one analytic
        field (same seeds for every precision), then the mean loss of unseen
        records of that field with the training weights."""
        netp = npmg.GuidingNetwork(npmg.benchmark_config("full", precision=precision), seed=7272)
        batch = npmg.synth_train(123, N_TRAIN)
        for k in range(steps):
            lib.npmg_synth_train(123, k * N_TRAIN, N_TRAIN, P(batch["inputs"]), P(batch["tgt_dir"]),
                                 P(batch["tgt_rgb"]), P(batch["tgt_pdf"]), stream.cuda_stream)
            netp.train_step_graph(batch["inputs"], batch["tgt_dir"], batch["tgt_rgb"], batch["tgt_pdf"], loss)
        held = npmg.synth_train(123, N_TRAIN, first=(steps + 1) * N_TRAIN)  # same field, unseen records
        raw = netp.forward(held["inputs"])
        _, per = netp.mixture_adjoint(raw, held["tgt_dir"], held["tgt_rgb"], held["tgt_pdf"], 1.0)
        value = per.mean().item()
        netp.close()
        log(f"{precision}: held-out loss after {steps} steps {value:.4f}")
        return value

    def pdf_agreement(make_cfg, precisions, cosine, train_steps=200, n=1 << 16, n_grad=1 << 14):
        """Each precision against the FP32 path on identical parameters (an FP32
        network trained for `train_steps` steps on the analytic field, copied
        into the other; the FP32 path is within 1e-5 of FP64 on the same
        parameters, so these are also the errors against FP64):
          pdf: relative difference of the 4 evaluation pdfs of n queries
               (median, 99th percentile, max);
          mixture: the mixture parameters the query kernel's raw outputs map to
               (softmax mapping, in FP64: lobe weights, absolute; concentrations,
               relative; mean directions, angle in radians), median and p99;
          gradient: dL/dparams of n_grad training records (the library's
               forward, adjoint and backward in that precision), relative
               2-norm error of the MLP block and of the grid block.
        FP8 is inference only: no gradient."""
        ref = npmg.GuidingNetwork(make_cfg("fp32"), seed=7272)
        batch = npmg.synth_train(123, N_TRAIN)
        for k in range(train_steps):
            lib.npmg_synth_train(123, k * N_TRAIN, N_TRAIN, P(batch["inputs"]), P(batch["tgt_dir"]),
                                 P(batch["tgt_rgb"]), P(batch["tgt_pdf"]), stream.cuda_stream)
            ref.train_step_graph(batch["inputs"], batch["tgt_dir"], batch["tgt_rgb"], batch["tgt_pdf"], loss)
        params = ref.get_params(1)
        ref.set_params(params)
        qq = npmg.synth_queries(555, n)
        nrm = qq["normal"] if cosine else None
        tt = npmg.synth_train(777, n_grad)
        mix_ok = make_cfg("fp32")["mixture"].get("otype", "vMF") == "vMF" and \
            make_cfg("fp32")["mixture"].get("mapping", "softmax") == "softmax"

        def run(net, grads):
            o = net.sample_pdf(qq["inputs"], nrm, None, qq["eval_dirs"], cosine_aware=cosine, raw=True)
            g = None
            if grads:
                raw = net.forward(tt["inputs"])
                dout, _ = net.mixture_adjoint(raw, tt["tgt_dir"], tt["tgt_rgb"], tt["tgt_pdf"], 128.0)
                g, _ = net.backward(tt["inputs"], dout)
                g = g.double() / 128.0
            return o["eval_pdf"].double(), o["raw"].double(), g

        def mixture(raw):
            K = raw.shape[0] // 4
            x = raw[:4 * K].reshape(K, 4, -1)
            th = torch.pi * torch.sigmoid(x[:, 0])
            ph = 2 * torch.pi * torch.sigmoid(x[:, 1])
            kap = torch.exp(x[:, 2]).clamp(1e-2, 1e3)
            w = torch.softmax(x[:, 3], dim=0)
            mu = torch.stack((torch.sin(th) * torch.cos(ph), torch.sin(th) * torch.sin(ph), torch.cos(th)))
            return w, kap, mu

        def q(t):
            t = t.flatten()
            return {"median": t.median().item(), "p99": t.quantile(0.99).item() if t.numel() <= 1 << 24 else None}

        a, ra, ga = run(ref, True)
        res = {}
        for p in precisions:
            other = npmg.GuidingNetwork(make_cfg(p), seed=7272)
            other.set_params(params)
            b, rb, gb = run(other, p != "fp8")
            rel = ((b - a).abs() / a).flatten()
            r = {"median": rel.median().item(), "p99": rel.quantile(0.99).item(), "max": rel.max().item(),
                 "queries": n, "trained_steps": train_steps}
            if mix_ok:
                wa, ka, ma = mixture(ra)
                wb, kb, mb = mixture(rb)
                ang = torch.acos((ma * mb).sum(0).clamp(-1.0, 1.0))
                r["mixture"] = {"weight_abs": q((wb - wa).abs()), "kappa_rel": q(((kb - ka) / ka).abs()),
                                "mean_angle_rad": q(ang)}
            if gb is not None:
                nm = other.n_mlp
                r["gradient_rel_l2"] = {
                    "mlp": ((gb[:nm] - ga[:nm]).norm() / ga[:nm].norm()).item(),
                    "grid": ((gb[nm:] - ga[nm:]).norm() / ga[nm:].norm()).item(), "records": n_grad}
            other.close()
            res[p] = r
        ref.close()
        return res

    def asg_cfg(precision):
        c = npmg.benchmark_config("full", precision=precision, width=128, hidden=3)
        c["mixture"] = {"otype": "ASG", "n_lobes": 16, "numerics": "fast"}
        return c

    def measure_frame(precision):
        """One wavefront frame of the anisotropic network: 8 bounce queues of
        floor(2,073,600 * 0.68^b) queries (capacity 2,073,600 each) served by one
        persistent npmg_sample_pdf_queues launch; then the per-frame train step
        (2^18 samples) as a graph."""
        netf = npmg.GuidingNetwork(asg_cfg(precision), seed=7272)
        tuned = netf.autotune(1 << 21) if precision != "fp32" else netf.tuning()
        counts = [int(FRAME_PATHS * 0.68 ** b) for b in range(FRAME_BOUNCES)]
        cap = FRAME_PATHS
        qin = torch.empty(FRAME_BOUNCES, 9, cap, device="cuda")
        qu = torch.empty(FRAME_BOUNCES, 3, cap, device="cuda")
        qe = torch.empty(FRAME_BOUNCES, 4, 3, cap, device="cuda")
        scratch_n = torch.empty(3, cap, device="cuda")
        for b in range(FRAME_BOUNCES):
            lib.npmg_synth_queries(900 + b, 0, cap, P(qin[b]), P(scratch_n), P(qu[b]), P(qe[b]), stream.cuda_stream)
        odir = torch.empty(FRAME_BOUNCES, 3, cap, device="cuda")
        opdf = torch.empty(FRAME_BOUNCES, cap, device="cuda")
        oev = torch.empty(FRAME_BOUNCES, 4, cap, device="cuda")
        olobe = torch.empty(FRAME_BOUNCES, cap, dtype=torch.int32, device="cuda")
        arr = (npmg.NpmgQueue * FRAME_BOUNCES)()
        for b in range(FRAME_BOUNCES):
            arr[b] = npmg.NpmgQueue(P(qin[b]), None, P(qu[b]), P(qe[b]), P(odir[b]), P(opdf[b]), P(oev[b]),
                                    P(olobe[b]), cap)
        cnt = torch.tensor(counts, dtype=torch.int32, device="cuda")

        def frame_step():
            st = lib.npmg_sample_pdf_queues(netf._h, arr, P(cnt), FRAME_BOUNCES, 4, 0, stream.cuda_stream)
            if st != 0:
                raise npmg.NpmgError(st, lib.npmg_last_error(netf._h).decode())

        ms = time_region(frame_step, max(2, args.steps // 2), 2, stream)
        total = sum(counts)
        log(f"ASG frame {precision}: {ms:.3f} ms per frame of {total} queries in 8 queues "
            f"({total / ms / 1e3:.1f} M queries/s), tuning {tuned}")
        ms_train = measure_train(netf, precision + " ASG", graph=True)
        netf.close()
        del qin, qu, qe, odir, opdf, oev, olobe
        return {"ms_per_frame": ms, "queries_per_frame": total, "queries_per_s": total / (ms * 1e-3),
                "train_graph_ms": ms_train, "train_samples_per_s": N_TRAIN / (ms_train * 1e-3),
                "tuning": {"pipeline_chunks": tuned[0], "blocks_per_cu": tuned[1]}}

    per_prec, per_prec_train, per_prec_graph, per_kernel = {}, {}, {}, {}
    # the anisotropic frame's 128-wide network: split operands do not fit its weights in LDS
    fprec = "bf16" if args.precision == "bf16x3" else args.precision
    if args.only == "frame":
        frame = {p: measure_frame(p) for p in ("fp32", fprec)}
        frame["pdf_rel_diff_vs_fp32_same_params"] = pdf_agreement(asg_cfg, [fprec], False, train_steps=100)[fprec]
        print(json.dumps(frame), flush=True)
        return
    if args.only == "train":
        nt = npmg.GuidingNetwork(full_cfg(args.precision), seed=7272)
        ms_e = measure_train(nt, args.precision)
        ms_g = measure_train(nt, args.precision, graph=True)
        nt.close()
        print(json.dumps({"precision": args.precision, "train_ms_per_step": ms_e, "train_graph_ms_per_step": ms_g,
                          "train_samples_per_s": N_TRAIN / (ms_e * 1e-3)}), flush=True)
        return
    precs = [args.precision] if args.precision == "fp32" else ["fp32", args.precision]
    if args.precision != "fp32" and args.only != "query":
        for extra in ("bf16", "fp16", "bf16x3"):
            if extra != args.precision:
                precs.append(extra)
        precs.append("fp8")  # inference only
    for prec in precs:
        netp, ms = measure_queries(prec)
        per_prec[prec] = ms
        if args.only != "query" and prec != "fp8":
            per_prec_train[prec] = measure_train(netp, prec)
            per_prec_graph[prec] = measure_train(netp, prec, graph=True)
        if prec != args.precision:
            netp.close()
        else:
            net = netp
    ms_exact = {p: measure_queries(p, "stable")[1] for p in (["fp32"] + ([args.precision] if args.precision != "fp32" else []))}
    log(f"network: {net.n_mlp} MLP + {net.n_grid} grid parameters")
    ms_q = per_prec[args.precision]
    qps = N_QUERIES / (ms_q * 1e-3)

    if args.only == "query":
        print(json.dumps({"ms_per_step": ms_q, "queries_per_s": qps, "precision": args.precision}))
        return
    ms_t = per_prec_train[args.precision]
    sps = N_TRAIN / (ms_t * 1e-3)

    half = args.precision != "fp32"
    # Per-launch figures of the two kernels of the step: algorithmic bytes of one
    # launch (its share of the queries) over its average launch duration.
    kq = per_kernel[args.precision]["query"]
    q_per_launch = N_QUERIES / kq["launches_per_step"]
    genc_bytes = (GRID_FEATURES * 2) if half else 0  # level-major grid features read back (BF16)
    table_bytes = 0 if half else net.n_grid * 4       # FP32 path gathers the table itself
    pos_bytes = 12 if half else 0                     # MFMA path: the positions are the encode's input
    io_bytes = (BYTES_PER_QUERY - pos_bytes + genc_bytes) * q_per_launch + table_bytes  # what the kernel moves
    # algorithmic bytes of the launch: the path's 132 B per query over the queries the
    # launch serves, plus the grid table once (16-bit tables for the MFMA paths)
    alg_bytes = 132 * q_per_launch + net.n_grid * (2 if half else 4)
    achieved = alg_bytes / (kq["avg_ms_per_launch"] * 1e-3) / 1e9
    flops = 2 * net.n_mlp * q_per_launch
    traffic = profiled(args.precision, "traffic_bytes", q_per_launch)  # FETCH_SIZE + WRITE_SIZE, calibrated
    fused_roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                      "kernel": ("k_query_mfma" if half else "k_query_fp32") +
                                " (fused OneBlob encode + MLP + mixture + sample + pdf)",
                      "algorithmic_bytes_per_launch": alg_bytes, "kernel_io_bytes_per_launch": io_bytes,
                      "avg_launch_ms": kq["avg_ms_per_launch"], "launches_per_step": kq["launches_per_step"]}
    roofline = fused_roofline
    ke = per_kernel[args.precision].get("grid_encode")
    encode_roofline = None
    if ke:
        # level-major grid encode: positions (12 B) in and L F BF16 features (64 B) out per
        # query, the BF16 table (26 MB) read once
        qe = N_QUERIES / ke["launches_per_step"]
        e_io = (12 + GRID_FEATURES * 2) * qe + net.n_grid * 2  # positions in, features out, table once
        e_bytes = 132 * qe + net.n_grid * 2
        e_ach = e_bytes / (ke["avg_ms_per_launch"] * 1e-3) / 1e9
        encode_roofline = {"bound": "hbm", "achieved": e_ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": e_ach / HBM_PEAK_GBS,
                           "traffic": profiled(args.precision, "traffic_bytes", qe, "k_grid_encode"),
                           "kernel": "k_grid_encode_lds + k_grid_encode (level-major hash-grid gathers, 16 levels "
                                     "x 8 corners; levels 0-1 from LDS-staged tables at 2^20 queries and more, "
                                     "timed as launches of the same tag)",
                           "algorithmic_bytes_per_launch": e_bytes, "kernel_io_bytes_per_launch": e_io,
                           "avg_launch_ms": ke["avg_ms_per_launch"], "launches_per_step": ke["launches_per_step"],
                           "limiter": "L1 cache-line lookups of random gathers (about 5 per query and level), "
                                      "not HBM bytes",
                           "l1_accesses_per_launch": profiled(args.precision, "l1_accesses", qe, "k_grid_encode")}
        if ke["avg_ms_per_launch"] * ke["launches_per_step"] > kq["avg_ms_per_launch"] * kq["launches_per_step"]:
            roofline = encode_roofline  # the dominant kernel of the step
    # Whole-step figures on the per-unit algorithmic work of the step: infer +
    # sample + pdf 132 B of query I/O per query plus the grid table once (16-bit
    # or FP32); train step 52 B of record I/O per sample, Adam + EMA at 42 B per
    # parameter, gradient zero and scatter 4 B per parameter, the 16-bit table
    # read once; MLP flops 2 n_mlp per query (forward) and 6 n_mlp per sample
    # (forward, input and weight gradients).
    def step_figures(ms, half_table):
        b = 132 * N_QUERIES + net.n_grid * (2 if half_table else 4)
        return {"bound": "hbm", "bytes_per_step": b, "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "mlp_flops_per_step": 2 * net.n_mlp * N_QUERIES,
                "mlp_tflops": 2 * net.n_mlp * N_QUERIES / (ms * 1e-3) / 1e12}

    step_roofline = {p: step_figures(per_prec[p], p != "fp32") for p in per_prec if p in ("fp32", args.precision)}
    n_all = net.n_mlp + net.n_grid
    t_bytes = 52 * N_TRAIN + 42 * n_all + 4 * n_all + 2 * net.n_grid
    t_flops = 6 * net.n_mlp * N_TRAIN

    def train_figures(ms, mfma):
        return {"bound": "hbm", "bytes_per_step": t_bytes, "achieved": t_bytes / (ms * 1e-3) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": t_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "mlp_flops_per_step": t_flops, "mlp_tflops": t_flops / (ms * 1e-3) / 1e12,
                "mlp_peak_tflops": MFMA_BF16_PEAK_TFLOPS if mfma else FP32_PEAK_TFLOPS,
                "ms_per_step": ms}

    train_roofline = {p: train_figures(v, p != "fp32") for p, v in per_prec_train.items()}
    result = {
        "metric": METRIC,
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_q,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16" if args.precision == "bf16x3" else args.precision,
        "data": "synthetic (pcg32-keyed hit records, analytic radiance field)",
        "config": {"workload": "infer+sample+pdf of 2^21 queries, hash grid 16x2^19x2 + OneBlob(16)x6, "
                               "MLP 3x64, 8 vMF lobes, cosine product, 4 eval directions, fast mixture numerics",
                   "queries_per_step": N_QUERIES, "train_batch": N_TRAIN, "precision": args.precision,
                   "operands": PRECISION_NOTE[args.precision]},
        "step_roofline": step_roofline,
        "train_roofline": train_roofline,
        "train_samples_per_s": sps,
        "train_ms_per_step": ms_t,
        "roofline": roofline,
        "fused_kernel_roofline": fused_roofline,
        "grid_encode_roofline": encode_roofline,
        "kernel_launches": per_kernel,
        # the kernel's actual limiter: VALU issue (4 cycles per 64-lane FP32 instruction per SIMD,
        # 1024 SIMDs at 2.4 GHz = 614.4 G wave-instructions/s; transcendentals issue at quarter rate)
        "valu_issue": (lambda v: None if v is None else {
            "wave_instructions_per_launch": v,
            "achieved_per_s": v / (kq["avg_ms_per_launch"] * 1e-3),
            "peak_per_s": VALU_ISSUE_PEAK,
            "frac": v / (kq["avg_ms_per_launch"] * 1e-3) / VALU_ISSUE_PEAK,
            "source": "profiles/k_query_%s_full_counters.md (SQ_INSTS_VALU)" % args.precision})(
            profiled(args.precision, "valu_wave_instructions", q_per_launch)),
        "mlp_compute": {"flops_per_launch": flops,
                        "achieved_tflops": flops / (kq["avg_ms_per_launch"] * 1e-3) / 1e12,
                        "peak_tflops": MFMA_BF16_PEAK_TFLOPS if half else FP32_PEAK_TFLOPS},
        "queries_per_s_by_precision": {k: N_QUERIES / (v * 1e-3) for k, v in per_prec.items()},
        "numerics": "fast (hardware transcendentals in the mixture stage, pdfs within 2e-5 of FP64 on the same "
                    "network outputs); the bit-exact stable numerics beside it",
        "queries_per_s_stable_numerics": {k: N_QUERIES / (v * 1e-3) for k, v in ms_exact.items()},
        "train_samples_per_s_by_precision": {k: N_TRAIN / (v * 1e-3) for k, v in per_prec_train.items()},
        "train_samples_per_s_graph_by_precision": {k: N_TRAIN / (v * 1e-3) for k, v in per_prec_graph.items()},
        "heldout_loss_after_500_steps": {p: loss_after(p) for p in ("fp32", "bf16", "fp16", "bf16x3")},
    }
    if half:
        result["pdf_rel_diff_vs_fp32_same_params"] = pdf_agreement(full_cfg, ["bf16", "fp16", "bf16x3", "fp8"], True)
        result["fp8_inference_speedup_vs_fp32"] = per_prec["fp32"] / per_prec["fp8"]
        frame = {p: measure_frame(p) for p in ("fp32", fprec)}
        frame["workload"] = ("one frame of 8 bounce queues of floor(2,073,600 * 0.68^b) queries, one persistent "
                             "launch; anisotropic mixture 16 lobes x 6, MLP 3x128, hash grid 16x2^19x2, "
                             "sample + 4 eval pdfs; train step 2^18 samples as one graph")
        frame[f"{fprec}_speedup_vs_fp32"] = frame["fp32"]["ms_per_frame"] / frame[fprec]["ms_per_frame"]
        frame["pdf_rel_diff_vs_fp32_same_params"] = pdf_agreement(asg_cfg, [fprec], False, train_steps=100)[fprec]
        result["anisotropic_frame"] = frame
    if "fp32" in per_prec and half:
        result[f"{args.precision}_speedup_vs_fp32"] = per_prec["fp32"] / per_prec[args.precision]
        result[f"{args.precision}_train_speedup_vs_fp32"] = per_prec_train["fp32"] / per_prec_train[args.precision]
    if not args.no_cpu:
        params = net.get_params(1).cpu().numpy()
        q_host = {k: v.cpu().numpy() for k, v in q.items()}
        t_host = {k: v.cpu().numpy() for k, v in t.items()}
        result["cpu_baseline"] = cpu_baseline(net.config, params, q_host, t_host)
        cb = result["cpu_baseline"]
        log(f"cpu baseline ({cb['cores']} threads): {cb['queries_per_s_65536']:.0f} queries/s at 65536, "
            f"{cb['value']:.0f} at 2^21, train {cb['train_samples_per_s']:.0f} samples/s")
    net.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
