"""torch.autograd wrapping of the library's forward/backward passes.

Follows the pattern of tiny-cuda-nn's PyTorch binding
(src/ext/tcnn/bindings/torch/tinycudann/modules.py:82-151): the parameters are
a flat float32 tensor in the library's order (MLP matrices, then the grid
table); forward copies them into the context and runs the HIP forward;
backward runs the hand-written HIP backward and returns dL/dparams. Inputs
(query records) receive no gradient, as in the reference where the network
input is data.
"""
from __future__ import annotations

import torch


class NetworkFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, net, inputs, params):
        net.set_params(params.detach())
        out = net.forward(inputs.detach().contiguous(), use_inference_params=False)
        ctx.net = net
        ctx.save_for_backward(inputs)
        return out

    @staticmethod
    def backward(ctx, dout):
        (inputs,) = ctx.saved_tensors
        dparams, _ = ctx.net.backward(inputs.contiguous(), dout.contiguous().float())
        return None, None, dparams


def network_apply(net, inputs, params):
    """raw outputs [n_out][n] = net(inputs; params), differentiable in params."""
    return NetworkFunction.apply(net, inputs, params)
