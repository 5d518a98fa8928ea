"""Drop-in ``ot.batch`` quadratic (Gromov-Wasserstein) solvers.

Public names and arguments follow ot/batch/_quadratic.py: ``tensor_batch``
(:17-152), ``loss_quadratic_batch`` (:155-198),
``loss_quadratic_samples_batch`` (:201-266) and ``solve_gromov_batch``
(:269-501).  The cost-tensor dictionary keeps the reference's keys
(``constC``, ``hC1``, ``hC2``, ``fC1``, ``fC2``) so user code that builds or
reads one keeps working.

How the work is laid out here:

* **Loss split.** A quadratic loss ``l(c1, c2) = f1(c1) + f2(c2) -
  <h1(c1), h2(c2)>`` (Peyre, Cuturi and Solomon, ICML 2016, Prop. 1) turns
  the 4-D product ``(L x T)_ik = sum_jl l(C1_ij, C2_kl) T_jl`` into
  ``f1(C1) p 1^T + 1 (f2(C2) q)^T - sum_d H1_d T H2_d^T`` with ``p, q`` the
  marginals.  ``_LOSSES`` holds the four maps of each supported loss; every
  structure is handled as channel-last ``(B, n, n, D)`` (a 3-D one is one
  channel).
* **Product on the library.** ``_TensorProduct`` runs the two mat-vecs and
  the two GEMMs per channel on ``potamd_gemm_batched``, which reads a
  channel of the channel-last tensors in place through its strides; its
  backward (the envelope gradient of the quadratic value) uses the same
  GEMM.
* **Outer loop.** Mirror descent on the linearised cost (Peyre et al.,
  Alg. 2): step k projects ``exp(log T_k - ((1 - alpha) M + 2 alpha
  L x T_k) / reg)`` onto the couplings with the library's batched log
  projection (the dense solver, rows first, kernel given), and stops when
  the largest per-problem mean change of the plan drops below ``tol``.

Gradients: ``grad="envelope"`` (default) differentiates the returned values
with the plan held fixed, through ``M`` and the cost tensor (so through
``C1`` and ``C2``); ``"detach"`` returns plain values; ``"autodiff"`` would
differentiate through the projections, which run in the HIP library, and
raises when a gradient is needed.
"""

import numpy as np
import torch

from . import _solve
from . import batch as _batch
from ._device import require_gpu
from ._grad import requires_grad
from .utils import OTResult

EPS_LOG = 1e-15  # the guard inside the logs of the KL loss and of log T


def _as_channels(C):
    """Channel-last view: (B, n, n) -> (B, n, n, 1); (B, n, n, D) unchanged."""
    return C.unsqueeze(-1) if C.dim() == 3 else C


def _sq_parts(C1, C2, logits):
    return (C1 ** 2).sum(-1), (C2 ** 2).sum(-1), 2.0 * C1, C2


def _kl_parts(C1, C2, logits):
    f1 = torch.zeros(C1.shape[:3], dtype=C1.dtype, device=C1.device)
    f2 = (C2 * torch.log(C2 + EPS_LOG)).sum(-1)
    return f1, f2, (C1 if logits else torch.log(C1 + EPS_LOG)), C2


# loss name -> (f1(C1), f2(C2), h1(C1), h2(C2)) on channel-last structures
_LOSSES = {"sqeuclidean": _sq_parts, "kl": _kl_parts}


class _TensorProduct(torch.autograd.Function):
    """``f1 p 1^T + 1 (f2 q)^T - S`` with ``S = sum_d H1_d T H2_d^T``, averaged
    with ``sum_d H1_d^T T H2_d`` when the structures are not symmetric.
    Inputs: f1 (B, n, n), f2 (B, m, m), p (B, n), q (B, m), H1 (B, n, n, D),
    H2 (B, m, m, D), T (B, n, m), one float dtype."""

    @staticmethod
    def forward(ctx, f1, f2, p, q, H1, H2, T, symmetric):
        f1, f2, p, q, H1, H2, T = (t.detach() for t in (f1, f2, p, q, H1, H2, T))
        T = T.contiguous()
        out = (_solve.bmm(f1, p.unsqueeze(-1)) + _solve.bmm(f2, q.unsqueeze(-1)).transpose(1, 2))
        out = out.contiguous()
        w = 1.0 if symmetric else 0.5
        for d in range(H1.shape[-1]):
            h1, h2 = H1[..., d], H2[..., d]
            _solve.gemm(_solve.bmm(h1, T), h2, out, alpha=-w, beta=1.0, transB=True)
            if not symmetric:
                _solve.gemm(_solve.bmm(h1, T, transA=True), h2, out, alpha=-w, beta=1.0)
        ctx.save_for_backward(f1, f2, p, q, H1, H2, T)
        ctx.symmetric = symmetric
        return out

    @staticmethod
    def backward(ctx, G):
        f1, f2, p, q, H1, H2, T = ctx.saved_tensors
        sym = ctx.symmetric
        G = G.contiguous().to(T.dtype)
        need = ctx.needs_input_grad
        rg, cg = G.sum(2), G.sum(1)
        df1 = rg.unsqueeze(2) * p.unsqueeze(1) if need[0] else None
        df2 = cg.unsqueeze(2) * q.unsqueeze(1) if need[1] else None
        dp = _solve.bmm(f1, rg.unsqueeze(-1), transA=True)[..., 0] if need[2] else None
        dq = _solve.bmm(f2, cg.unsqueeze(-1), transA=True)[..., 0] if need[3] else None
        dH1 = torch.zeros_like(H1) if need[4] else None
        dH2 = torch.zeros_like(H2) if need[5] else None
        dT = torch.zeros_like(T) if need[6] else None
        w = 1.0 if sym else 0.5
        for d in range(H1.shape[-1]):
            h1, h2 = H1[..., d].contiguous(), H2[..., d].contiguous()
            # S1 = h1 T h2^T: dh1 = G h2 T^T, dh2 = G^T h1 T, dT = h1^T G h2
            if dH1 is not None:
                dH1[..., d] -= w * _solve.bmm(_solve.bmm(G, h2), T, transB=True)
            if dH2 is not None:
                dH2[..., d] -= w * _solve.bmm(_solve.bmm(G, h1, transA=True), T)
            if dT is not None:
                dT -= w * _solve.bmm(_solve.bmm(h1, G, transA=True), h2)
            if not sym:
                # S2 = h1^T T h2: dh1 = T h2 G^T, dh2 = T^T h1 G, dT = h1 G h2^T
                if dH1 is not None:
                    dH1[..., d] -= w * _solve.bmm(_solve.bmm(T, h2), G, transB=True)
                if dH2 is not None:
                    dH2[..., d] -= w * _solve.bmm(_solve.bmm(T, h1, transA=True), G)
                if dT is not None:
                    dT -= w * _solve.bmm(_solve.bmm(h1, G), h2, transB=True)
        return df1, df2, dp, dq, dH1, dH2, dT, None


def _symmetrised(F, symmetric):
    return F if symmetric else 0.5 * (F + F.transpose(1, 2))


def _product(L, T, p, q, symmetric):
    return _TensorProduct.apply(L["fC1"], L["fC2"], p, q, _as_channels(L["hC1"]),
                                _as_channels(L["hC2"]), T, bool(symmetric))


def _separable(f1, f2, p, q):
    """``f1 p 1^T + 1 (f2 q)^T`` (the constant part of the product)."""
    B, n, m = f1.shape[0], f1.shape[1], f2.shape[1]
    none1 = f1.new_zeros((B, n, n, 0))
    none2 = f2.new_zeros((B, m, m, 0))
    return _TensorProduct.apply(f1, f2, p, q, none1, none2, f1.new_zeros((B, n, m)), True)


def tensor_batch(a, b, C1, C2, symmetric=True, nx=None, loss="sqeuclidean", logits=None):
    """Cost-tensor dictionary of a batch (:17-152).  NumPy inputs give a
    dictionary of NumPy arrays, torch inputs one of device tensors."""
    if not isinstance(C1, torch.Tensor):
        conv = _Conv(a, b, C1, C2)
        L = tensor_batch(*conv.to_dev(a, b, C1, C2), symmetric=symmetric, loss=loss,
                         logits=logits)
        return {k: conv.out(v) for k, v in L.items()}
    if loss not in _LOSSES:
        raise ValueError(f"Unknown loss function: {loss}")
    if loss == "kl":
        assert logits in [True, False], "logits must be either True or False for KL loss"
        assert C2.dim() == 4, "C2 must be a bxnxnxd tensor"
    f1, f2, h1, h2 = _LOSSES[loss](_as_channels(C1), _as_channels(C2), logits)
    f1, f2 = _symmetrised(f1, symmetric), _symmetrised(f2, symmetric)
    return {"constC": _separable(f1, f2, a, b), "hC1": h1, "hC2": h2, "fC1": f1, "fC2": f2}


def tensor_product_batch(L, T, nx=None, recompute_const=False, symmetric=True):
    """``L x T`` (:535-572): with ``recompute_const`` the separable part uses
    the plan's own marginals instead of the stored ``constC``."""
    if recompute_const:
        return _product(L, T, T.sum(2), T.sum(1), symmetric)
    B, n, m = T.shape
    zp = torch.zeros((B, n), dtype=T.dtype, device=T.device)
    zq = torch.zeros((B, m), dtype=T.dtype, device=T.device)
    return L["constC"] + _product(L, T, zp, zq, symmetric)  # zero marginals: no f part


class _Conv:
    """NumPy or torch in; device torch inside; the caller's kind out."""

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
        out = []
        for x in xs:
            if x is None:
                out.append(None)
            elif isinstance(x, torch.Tensor):
                out.append(x.to(device=self.device, dtype=self.dtype))
            else:
                out.append(torch.as_tensor(np.asarray(x), dtype=self.dtype, device=self.device))
        return out

    def out(self, t):
        if t is None:
            return None
        if self.is_torch:
            return t.to(self.out_device)
        return t.detach().cpu().numpy()


def loss_quadratic_batch(L, T, recompute_const=False, symmetric=True, nx=None):
    """``<L x T, T>`` per problem (:155-198)."""
    conv = _Conv(T)
    Ld = {k: conv.to_dev(v)[0] for k, v in L.items()}
    Td = conv.to_dev(T)[0]
    LT = tensor_product_batch(Ld, Td, recompute_const=recompute_const, symmetric=symmetric)
    return conv.out(torch.sum(LT * Td, dim=(1, 2)))


def loss_quadratic_samples_batch(a, b, C1, C2, T, loss="sqeuclidean", symmetric=None,
                                 nx=None, logits=None, recompute_const=False):
    """``tensor_batch`` then ``loss_quadratic_batch`` (:201-266)."""
    conv = _Conv(a, b, C1, C2, T)
    a, b, C1, C2, T = conv.to_dev(a, b, C1, C2, T)
    # symmetric=None is passed on as is, and reads as "not symmetric"
    sym = bool(symmetric)
    L = tensor_batch(a, b, C1, C2, symmetric=sym, loss=loss, logits=logits)
    LT = tensor_product_batch(L, T, recompute_const=recompute_const, symmetric=sym)
    return conv.out(torch.sum(LT * T, dim=(1, 2)))


def _check_fused(M, alpha, B, n, m, like):
    """The linear term of the fused problem, with the reference's checks
    (:420-446): neither M nor alpha means plain Gromov-Wasserstein."""
    if M is None and alpha is None:
        return torch.zeros((B, n, m), dtype=like.dtype, device=like.device), 1.0
    if alpha is None:
        raise ValueError("If M is provided, alpha must also be provided for Fused "
                         "Gromov-Wasserstein problem")
    if M is None:
        raise ValueError("If alpha is provided, M must also be provided for Fused "
                         "Gromov-Wasserstein problem")
    if tuple(M.shape) != (B, n, m):
        raise ValueError(f"Shape of M {tuple(M.shape)} does not match the batch size "
                         f"{B} and dimensions {n}, {m}")
    if alpha < 0 or alpha > 1:
        raise ValueError("Alpha must be in [0, 1] for Fused Gromov-Wasserstein problem")
    return M, alpha


def _is_symmetric(C):
    return bool(torch.allclose(C, C.transpose(1, 2), atol=1e-10))


def solve_gromov_batch(C1, C2, reg=1e-2, a=None, b=None, loss="sqeuclidean", symmetric=None,
                       M=None, alpha=None, T_init=None, max_iter=50, tol=1e-5,
                       max_iter_inner=50, tol_inner=1e-5, grad="envelope", logits=None):
    """Batched (fused) Gromov-Wasserstein; see the module docstring."""
    if grad == "autodiff" and requires_grad(C1, C2, a, b, M):
        raise NotImplementedError("grad='autodiff' would differentiate through the GPU "
                                  "projections; use grad='envelope' or 'detach'")
    conv = _Conv(C1, C2, a, b, M, T_init)
    C1, C2, a, b, M, T_init = conv.to_dev(C1, C2, a, b, M, T_init)
    B, n, m = C1.shape[0], C1.shape[1], C2.shape[1]
    a = torch.full((B, n), 1.0 / n, dtype=C1.dtype, device=C1.device) if a is None else a
    b = torch.full((B, m), 1.0 / m, dtype=C1.dtype, device=C1.device) if b is None else b
    if symmetric is None:
        symmetric = _is_symmetric(C1) and _is_symmetric(C2)
    if not isinstance(loss, str):
        raise ValueError(f"Unknown loss function: {loss}")
    L = tensor_batch(a, b, C1, C2, symmetric=symmetric, loss=loss, logits=logits)
    M, alpha = _check_fused(M, alpha, B, n, m, C1)

    # the outer loop sees detached data only
    Ld = {k: v.detach() for k, v in L.items()}
    Md, ad, bd = M.detach(), a.detach(), b.detach()
    if T_init is None:
        scale = torch.sqrt(ad.sum(1) * bd.sum(1))
        T = ad[:, :, None] * bd[:, None, :] / scale[:, None, None]
    else:
        T = T_init.detach()
    log_T = torch.log(T + EPS_LOG)
    proj, step = None, -1
    with torch.no_grad():
        for step in range(max_iter):
            lin = (1 - alpha) * Md + (2 * alpha) * tensor_product_batch(Ld, T,
                                                                         symmetric=symmetric)
            proj = _batch.bregman_log_projection_batch(log_T - lin / reg, ad, bd,
                                                       max_iter=max_iter_inner, tol=tol_inner)
            change = torch.max(torch.mean(torch.abs(proj["T"] - T), dim=(1, 2)))
            T, log_T = proj["T"], proj["log_T"]
            if change < tol:
                break

    if grad == "detach":
        L = Ld
        M = Md
    value_linear = torch.sum(M * T, dim=(1, 2))
    # the quadratic value always uses the plan's own marginals (:485-487)
    value_quad = torch.sum(tensor_product_batch(L, T, recompute_const=True,
                                                symmetric=symmetric) * T, dim=(1, 2))
    value = (1 - alpha) * value_linear + alpha * value_quad
    potentials = None if proj is None else tuple(conv.out(x) for x in proj["potentials"])
    return OTResult(value=conv.out(value), value_linear=conv.out(value_linear * (1 - alpha)),
                    value_quad=conv.out(value_quad * alpha), plan=conv.out(T),
                    potentials=potentials, log={"n_iter": step})
