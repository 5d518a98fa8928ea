"""Drop-in ``ot.batch`` quadratic (Gromov-Wasserstein) solvers.

Same names, arguments and outputs as ot/batch/_quadratic.py:

* ``tensor_batch``                  :17-152 (sqeuclidean and KL losses)
* ``loss_quadratic_batch``          :155-198
* ``loss_quadratic_samples_batch``  :201-266
* ``solve_gromov_batch``            :269-501 (proximal gradient: each outer
  step projects ``K = -M_k / reg + log T`` with the batched log Bregman
  projection, :458-473)

The inner projection is ``pot_amd.batch.bregman_log_projection_batch``, i.e.
the gfx950 dense solver (log domain, rows first, kernel given).  The tensor
product ``L x T = const - hC1 T hC2^T`` (:535-572) is two plain batched GEMMs
per feature channel, done with ``torch.matmul`` on the device.  With torch
inputs that require grad, ``grad="envelope"`` (the default) differentiates
the values through the cost tensor with the plan fixed; ``"autodiff"``
through the iterations is not supported and raises.
"""

import numpy as np
import torch

from . import batch as _batch
from ._device import Ingress, require_gpu
from .utils import OTResult


def _dev(x, dtype, device):
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=dtype)
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


def _transpose(C):
    return C.transpose(1, 2) if C.dim() == 3 else C.transpose(1, 2)


def _bmv(A, v):
    return torch.matmul(A, v.unsqueeze(-1)).squeeze(-1)


def compute_const_from_marginals(fC1, fC2, a, b, nx=None):
    """``fC1 a 1^T + 1 (fC2 b)^T`` per problem (:581-590)."""
    return _bmv(fC1, a)[:, :, None] + _bmv(fC2, b)[:, None, :]


def compute_tensor_batch(f1, f2, h1, h2, a, b, C1, C2, symmetric=True):
    """The matrices that define the cost tensor (:504-532)."""
    fC1 = f1(C1)
    fC2 = f2(C2)
    if not symmetric:
        fC1 = 0.5 * (fC1 + _transpose(fC1))
        fC2 = 0.5 * (fC2 + _transpose(fC2))
    hC1 = h1(C1)
    hC2 = h2(C2)
    constC = compute_const_from_marginals(fC1, fC2, a, b)
    return {"constC": constC, "hC1": hC1, "hC2": hC2, "fC1": fC1, "fC2": fC2}


def tensor_batch(a, b, C1, C2, symmetric=True, nx=None, loss="sqeuclidean", logits=None):
    """Gromov-Wasserstein cost tensor of a batch (:17-152); torch tensors in."""
    if loss == "sqeuclidean":
        def f1(C):
            return torch.sum(C ** 2, dim=-1) if C.dim() == 4 else C ** 2

        def f2(C):
            return torch.sum(C ** 2, dim=-1) if C.dim() == 4 else C ** 2

        def h1(C):
            return 2 * (C.unsqueeze(-1) if C.dim() == 3 else C)

        def h2(C):
            return C.unsqueeze(-1) if C.dim() == 3 else C
    elif loss == "kl":
        assert logits in [True, False], "logits must be either True or False for KL loss"

        def f1(C):
            return torch.zeros(C.shape[:3], dtype=C.dtype, device=C.device)

        def f2(C):
            assert C.dim() == 4, "C2 must be a bxnxnxd tensor"
            return torch.sum(C * torch.log(C + 1e-15), dim=-1)

        def h1(C):
            return C if logits else torch.log(C + 1e-15)

        def h2(C):
            return C
    else:
        raise ValueError(f"Unknown loss function: {loss}")
    return compute_tensor_batch(f1, f2, h1, h2, a, b, C1, C2, symmetric=symmetric)


def _hTh(hC1, T, hC2):
    """sum_d hC1[..., d] T hC2[..., d]^T  (B, n, m)."""
    h1 = hC1.permute(0, 3, 1, 2).contiguous()          # B d n n
    h2t = hC2.permute(0, 3, 2, 1).contiguous()         # B d m m (transposed)
    out = torch.matmul(torch.matmul(h1, T.unsqueeze(1)), h2t)  # B d n m
    return out.sum(1)


def tensor_product_batch(L, T, nx=None, recompute_const=False, symmetric=True):
    """``L x T = const - hC1 T hC2^T`` (:535-572)."""
    if recompute_const:
        const = compute_const_from_marginals(L["fC1"], L["fC2"], T.sum(2), T.sum(1))
    else:
        const = L["constC"]
    dot = _hTh(L["hC1"], T, L["hC2"])
    if not symmetric:
        dot_t = _hTh(L["hC1"].transpose(1, 2), T, L["hC2"].transpose(1, 2))
        dot = (dot + dot_t) / 2
    return const - dot


def loss_quadratic_batch(L, T, recompute_const=False, symmetric=True, nx=None):
    """``<L x T, T>`` per problem (:155-198)."""
    LT = tensor_product_batch(L, T, recompute_const=recompute_const, symmetric=symmetric)
    return torch.sum(LT * T, dim=(1, 2))


def loss_quadratic_samples_batch(a, b, C1, C2, T, loss="sqeuclidean", symmetric=None,
                                 nx=None, logits=None, recompute_const=False):
    """``tensor_batch`` then ``loss_quadratic_batch`` (:201-266)."""
    conv = _Conv(a, b, C1, C2, T)
    a, b, C1, C2, T = conv.to_dev(a, b, C1, C2, T)
    L = tensor_batch(a, b, C1, C2, symmetric=symmetric if symmetric is not None else True,
                     loss=loss, logits=logits)
    return conv.out(loss_quadratic_batch(L, T, recompute_const=recompute_const,
                                         symmetric=symmetric if symmetric is not None
                                         else True))


class _Conv:
    """numpy or torch in; device torch inside; the caller's kind out."""

    def __init__(self, *xs):
        ref = next((x for x in xs if x is not None), None)
        self.is_torch = isinstance(ref, torch.Tensor)
        self.device = require_gpu()
        if self.is_torch:
            self.dtype = ref.dtype if ref.dtype in (torch.float32, torch.float64) \
                else torch.float32
            self.out_device = ref.device
        else:
            self.dtype = torch.float64 if np.asarray(ref).dtype == np.float64 else torch.float32

    def to_dev(self, *xs):
        return [None if x is None else _dev(x, self.dtype, self.device) for x in xs]

    def out(self, t):
        if t is None:
            return None
        if self.is_torch:
            return t.to(self.out_device)
        return t.detach().cpu().numpy()


def solve_gromov_batch(C1, C2, reg=1e-2, a=None, b=None, loss="sqeuclidean", symmetric=None,
                       M=None, alpha=None, T_init=None, max_iter=50, tol=1e-5,
                       max_iter_inner=50, tol_inner=1e-5, grad="envelope", logits=None):
    """Batched (fused) Gromov-Wasserstein by proximal gradient (:269-501)."""
    needs_grad = any(isinstance(x, torch.Tensor) and x.requires_grad
                     for x in (C1, C2, a, b, M))
    if needs_grad and grad == "autodiff":
        raise NotImplementedError("grad='autodiff' would differentiate through the GPU "
                                  "projections; use grad='envelope' or 'detach'")
    conv = _Conv(C1, C2, a, b, M, T_init)
    C1, C2, a, b, M, T_init = conv.to_dev(C1, C2, a, b, M, T_init)
    B, n, m = C1.shape[0], C1.shape[1], C2.shape[1]
    if a is None:
        a = torch.ones((B, n), dtype=C1.dtype, device=C1.device) / n
    if b is None:
        b = torch.ones((B, m), dtype=C2.dtype, device=C2.device) / m
    if symmetric is None:
        symmetric = bool(torch.allclose(C1, _transpose(C1), atol=1e-10)
                         and torch.allclose(C2, _transpose(C2), atol=1e-10))
    if isinstance(loss, str):
        L = tensor_batch(a, b, C1, C2, symmetric=symmetric, loss=loss, logits=logits)
    else:
        raise ValueError(f"Unknown loss function: {loss}")
    if M is None and alpha is None:
        M = torch.zeros((B, n, m), dtype=C1.dtype, device=C1.device)
        alpha = 1.0
    elif M is not None and alpha is None:
        raise ValueError("If M is provided, alpha must also be provided for Fused "
                         "Gromov-Wasserstein problem")
    elif M is None and alpha is not None:
        raise ValueError("If alpha is provided, M must also be provided for Fused "
                         "Gromov-Wasserstein problem")
    else:
        if M.shape[0] != B or M.shape[1] != n or M.shape[2] != m:
            raise ValueError(f"Shape of M {tuple(M.shape)} does not match the batch size "
                             f"{B} and dimensions {n}, {m}")
        if alpha < 0 or alpha > 1:
            raise ValueError("Alpha must be in [0, 1] for Fused Gromov-Wasserstein problem")

    with torch.no_grad():
        if T_init is None:
            const = torch.sqrt(a.sum(1) * b.sum(1))
            T_init = a[:, :, None] * b[:, None, :] / const[:, None, None]
        T = T_init.detach()
        log_T = torch.log(T + 1e-15)
        Ld = {k: v.detach() for k, v in L.items()}
        Md = M.detach()
        it = -1
        out = None
        for it in range(max_iter):
            T_prev = T
            Mk = (1 - alpha) * Md + 2 * alpha * tensor_product_batch(Ld, T, symmetric=symmetric)
            K = -Mk / reg + log_T
            out = _batch.bregman_log_projection_batch(K, a.detach(), b.detach(),
                                                      max_iter=max_iter_inner, tol=tol_inner)
            T = out["T"]
            log_T = out["log_T"]
            max_err = torch.max(torch.mean(torch.abs(T_prev - T), dim=(1, 2)))
            if max_err < tol:
                break
    if grad == "detach":
        L = {k: v.detach() for k, v in L.items()}
        M = M.detach()
    value_linear = torch.sum(M * T, dim=(1, 2))
    value_quadratic = loss_quadratic_batch(L, T, recompute_const=True, symmetric=symmetric)
    value = (1 - alpha) * value_linear + alpha * value_quadratic
    potentials = None if out is None else tuple(conv.out(p) for p in out["potentials"])
    return OTResult(value=conv.out(value), value_linear=conv.out(value_linear * (1 - alpha)),
                    value_quad=conv.out(value_quadratic * alpha), plan=conv.out(T),
                    potentials=potentials, log={"n_iter": it})
