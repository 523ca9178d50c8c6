"""torch.autograd wrapping of the library's forward/backward passes.

Follows the pattern of tiny-cuda-nn's PyTorch binding
(src/ext/tcnn/bindings/torch/tinycudann/modules.py:82-151): the parameters are
a flat float32 tensor in the library's order (MLP matrices, then the grid
table); forward copies them into the context's training master (the EMA /
inference copy is left alone) and runs the HIP forward; backward copies the
parameters saved at forward time back into the master (another forward, a
set_params or an optimiser step may have changed it in between, as
modules.py:82-151 passes the saved params to bwd) and runs the hand-written
HIP backward, returning dL/dparams. Inputs (query records) receive no
gradient, as in the reference where the network input is data.
"""
from __future__ import annotations

import torch


class NetworkFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, net, inputs, params):
        p = params.detach()
        net.set_params(p, which=0)
        out = net.forward(inputs.detach().contiguous(), use_inference_params=False)
        ctx.net = net
        ctx.save_for_backward(inputs, p)
        return out

    @staticmethod
    def backward(ctx, dout):
        inputs, params = ctx.saved_tensors
        ctx.net.set_params(params, which=0)
        dparams, _ = ctx.net.backward(inputs.contiguous(), dout.contiguous().float())
        return None, None, dparams


def network_apply(net, inputs, params):
    """raw outputs [n_out][n] = net(inputs; params), differentiable in params."""
    return NetworkFunction.apply(net, inputs, params)


class MixtureLossFunction(torch.autograd.Function):
    """Mean KL loss -f/q log(p + eps) of the mixture over a batch
    (compute_dL_doutput_divergence, train.cu:69-124) with its hand-written
    adjoint as the backward: dL/draw from npmg_mixture_adjoint with loss
    scale 1 (so already divided by the batch size). The adjoint zeroes
    non-finite gradients and applies the configuration's upper clamp, as the
    reference does; set "loss": {"grad_clamp": 0} for an unclamped gradient."""

    @staticmethod
    def forward(ctx, net, raw, tgt_dir, tgt_rgb, tgt_pdf):
        dout, loss = net.mixture_adjoint(raw.detach().contiguous(), tgt_dir, tgt_rgb, tgt_pdf, 1.0)
        ctx.save_for_backward(dout)
        return loss.mean()

    @staticmethod
    def backward(ctx, grad):
        (dout,) = ctx.saved_tensors
        return None, dout * grad, None, None, None


def mixture_loss(net, raw, tgt_dir, tgt_rgb, tgt_pdf):
    """Mean mixture KL loss of raw outputs [n_out][n], differentiable in raw."""
    return MixtureLossFunction.apply(net, raw, tgt_dir, tgt_rgb, tgt_pdf)
