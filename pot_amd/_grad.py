"""Gradient plumbing of the drop-in API.

The reference attaches gradients in two ways: ``nx.set_gradients`` (the
envelope gradient of ``ot.solve``, ot/backend.py TorchBackend.set_gradients,
used at ot/solvers.py:483-495) and plain autodiff through its torch
operations.  The iterations here run in the HIP library, which torch cannot
trace, so a value computed there gets its gradient attached explicitly.
"""

import torch


class SetGradients(torch.autograd.Function):
    """``value`` whose gradient with respect to ``inputs`` is ``grads``.

    ``SetGradients.apply(value, *inputs, *grads)``: ``inputs`` and ``grads``
    have the same length; ``grads[k]`` has the shape of ``inputs[k]``.
    """

    @staticmethod
    def forward(ctx, value, *inputs_and_grads):
        k = len(inputs_and_grads) // 2
        ctx.grads = inputs_and_grads[k:]
        return value.detach().clone()

    @staticmethod
    def backward(ctx, grad_out):
        outs = [None]
        outs += [grad_out * g for g in ctx.grads]
        outs += [None] * len(ctx.grads)
        return tuple(outs)


def requires_grad(*xs):
    """True if any argument is a torch tensor that requires grad."""
    return any(isinstance(x, torch.Tensor) and x.requires_grad for x in xs)


def envelope(value, pairs, dtype, device):
    """0-d tensor ``value`` (a float) carrying ``d value / d x = g`` for each
    ``(x, g)`` of ``pairs`` whose ``x`` is a tensor that requires grad."""
    inputs, grads = [], []
    for x, g in pairs:
        if isinstance(x, torch.Tensor) and x.requires_grad:
            inputs.append(x)
            grads.append(g.detach().to(device=x.device, dtype=x.dtype).reshape(x.shape))
    val = torch.tensor(value, dtype=dtype, device=device)
    if not inputs:
        return val
    return SetGradients.apply(val, *inputs, *grads)


class SetBatchGradients(torch.autograd.Function):
    """Per-problem values ``(B,)`` whose gradient with respect to
    ``inputs[k]`` (leading dimension B) is ``grads[k]``: ``d value_p / d x[p] =
    g[p]`` and no cross-problem terms."""

    @staticmethod
    def forward(ctx, value, *inputs_and_grads):
        k = len(inputs_and_grads) // 2
        ctx.grads = inputs_and_grads[k:]
        return value.detach().clone()

    @staticmethod
    def backward(ctx, grad_out):
        outs = [None]
        for g in ctx.grads:
            outs.append(grad_out.reshape((-1,) + (1,) * (g.dim() - 1)) * g)
        outs += [None] * len(ctx.grads)
        return tuple(outs)


def batch_envelope(values, pairs):
    """``values`` (B,) tensor carrying ``d values[p] / d x[p] = g[p]`` for each
    ``(x, g)`` of ``pairs`` whose ``x`` is a tensor that requires grad."""
    inputs, grads = [], []
    for x, g in pairs:
        if isinstance(x, torch.Tensor) and x.requires_grad:
            inputs.append(x)
            grads.append(g.detach().to(device=x.device, dtype=x.dtype).reshape(x.shape))
    if not inputs:
        return values
    return SetBatchGradients.apply(values, *inputs, *grads)
