"""Drop-in ``ot.batch`` linear solvers.

Same names, arguments and outputs as ot/batch/_linear.py and
ot/batch/_utils.py; the Sinkhorn iterations, plans, per-problem sums and
cost matrices run in the gfx950 HIP library.

* ``bregman_log_projection_batch``  _utils.py:165-290 (rows first, one stop on
  the largest row-marginal error of the batch, every 10th iteration)
* ``bregman_projection_batch``      _utils.py:46-162
* ``solve_batch``                   _linear.py:236-357
* ``solve_sample_batch``            _linear.py:360-435
* ``dist_batch`` and its metrics    _linear.py:20-233
* ``loss_linear_batch``, ``loss_linear_samples_batch``, ``entropy_batch``,
  ``norm_batch``                    _linear.py:123-175, _utils.py:4-29

Gradient modes (``grad``), for torch inputs that require grad:

* ``"detach"``: no gradient;
* ``"envelope"`` (``solve_batch``'s default): the plan is held fixed, so
  ``d value / d M = T`` (and, with ``reg_type="kl"``, the KL term's
  derivative in ``a`` and ``b``), _linear.py:332-336;
* ``"last_step"``: the solve runs as usual, then one more update (u then v,
  or f then g) is taken in torch on the caller's cost, so the gradient flows
  through that update (_utils.py:263-266, :280-284, :134-135, :150-154);
  without gradients the same extra update runs on the device;
* ``"autodiff"`` would differentiate through every iteration, which run in
  the HIP library: it raises ``NotImplementedError`` when a gradient is
  needed (without one it is the same as ``"detach"``).

The cost matrices carry gradients to the points: the squared Euclidean
kernel's backward is ``dX = 2 (rowsum(G) X - G Y)``, ``dY = 2 (colsum(G) Y -
G^T X)``, computed with the library's batched GEMM; the Minkowski and KL
costs have closed-form backward passes in torch.
"""

import ctypes

import numpy as np
import torch

from . import _device, _lib, _solve
from ._device import Ingress
from ._grad import batch_envelope, requires_grad
from .utils import OTResult


def entropy_batch(T, nx=None, eps=1e-16):
    """``-sum T log(T + eps)`` per plan (_utils.py:4-22)."""
    if isinstance(T, torch.Tensor):
        return -torch.sum(T * torch.log(T + eps), dim=(1, 2))
    return -np.sum(T * np.log(T + eps), axis=(1, 2))


def norm_batch(u, p=2, nx=None):
    """``(sum u^p)^(1/p)`` per row (_utils.py:25-29)."""
    return (u ** p).sum(1) ** (1 / p)


def loss_linear_batch(M, T, nx=None):
    """``<M_b, T_b>`` per problem (_linear.py:123-147)."""
    if isinstance(M, torch.Tensor):
        return torch.sum(M * T, dim=(1, 2))
    return np.sum(M * T, axis=(1, 2))


def _autodiff_error():
    return NotImplementedError(
        "grad='autodiff' would differentiate through every Sinkhorn iteration, which run "
        "in the HIP library; use grad='envelope', 'last_step' or 'detach'")


def _weights_dev(w, ing, shape):
    if w is None:
        return torch.full(shape, 1.0 / shape[1], dtype=torch.float64, device=ing.device)
    return ing.dev(w, torch.float64)


def _caller_weights(w, wdev, like):
    """The weights as the caller's tensor (keeps their gradient) or the
    device copy in the cost's dtype."""
    if isinstance(w, torch.Tensor):
        return w
    return wdev.to(device=like.device, dtype=like.dtype)


def _last_step(Kc, u, v, a_c, b_c, domain):
    """One update from the detached solver state on the caller's (not
    detached) kernel ``Kc``: log domain ``u = log a - LSE_j(K + v)``, ``v =
    log b - LSE_i(K + u)``, ``log T = K.detach() + u + v`` (_utils.py:280-286);
    scaling ``f = a / (K g)``, ``g = b / (K^T f)``, ``T = K.detach() f g``
    (:150-156).  Returns (T, log_T or None, u, v)."""
    u = u.to(device=Kc.device, dtype=Kc.dtype)
    v = v.to(device=Kc.device, dtype=Kc.dtype)
    if domain == _lib.LOG:
        un = torch.log(a_c) - torch.logsumexp(Kc + v[:, None, :], dim=2)
        vn = torch.log(b_c) - torch.logsumexp(Kc + un[:, :, None], dim=1)
        log_T = Kc.detach() + un[:, :, None] + vn[:, None, :]
        return torch.exp(log_T), log_T, un, vn
    f = a_c / torch.sum(Kc * v[:, None, :], dim=2)
    g = b_c / torch.sum(Kc * f[:, :, None], dim=1)
    return Kc.detach() * f[:, :, None] * g[:, None, :], None, f, g


def _projection(K, a, b, max_iter, tol, domain, grad):
    if grad not in ("detach", "autodiff", "last_step", None):
        raise ValueError(f"Unknown gradient mode: {grad}")
    if grad == "autodiff" and requires_grad(K, a, b):
        raise _autodiff_error()
    ing = Ingress(K, a, b)
    Kt = ing.dev(K)
    B, n, m = Kt.shape
    at = _weights_dev(a, ing, (B, n))
    bt = _weights_dev(b, ing, (B, m))
    init = 0.0 if domain == _lib.LOG else 1.0
    u0 = torch.full((B, n), init, dtype=torch.float64, device=ing.device)
    v0 = torch.full((B, m), init, dtype=torch.float64, device=ing.device)
    res = _solve.dense_solve(Kt, at, bt, u0, v0, None, domain=domain,
                             order=_lib.ROW_FIRST, kernel_given=True,
                             max_iter=max_iter, stop_thr=tol, err_norm=0)
    u, v = res.u, res.v
    if grad == "last_step":
        if requires_grad(K, a, b):
            a_c = _caller_weights(a, at, K)
            b_c = _caller_weights(b, bt, K)
            T, L, un, vn = _last_step(K, u, v, a_c, b_c, domain)
            return ing, res.niter, T, L, (un, vn), True
        # the same update on the device: one more row-first iteration
        res1 = _solve.dense_solve(Kt, at, bt, u, v, None, domain=domain,
                                  order=_lib.ROW_FIRST, kernel_given=True, max_iter=1,
                                  stop_thr=0.0, err_norm=0)
        u, v = res1.u, res1.v
    P, L, _ = _solve.plan(Kt, u, v, None, domain=domain, kernel_given=True,
                          want_log=domain == _lib.LOG, want_sums=False)
    T = ing.out(P)
    log_T = None if L is None else ing.out(L)
    return ing, res.niter, T, log_T, (ing.out(u), ing.out(v)), False


def bregman_log_projection_batch(K, a=None, b=None, nx=None, max_iter=10000, tol=1e-5,
                                 grad="detach"):
    """Batched log-domain Bregman projection of ``exp(K)`` (_utils.py:165-290).

    Returns ``{"T", "log_T", "n_iters", "potentials": (u, v)}``.
    """
    ing, niter, T, L, pots, _ = _projection(K, a, b, max_iter, tol, _lib.LOG, grad)
    return {"T": T, "log_T": L, "n_iters": niter, "potentials": pots}


def bregman_projection_batch(K, a=None, b=None, nx=None, max_iter=10000, tol=1e-5,
                             grad="detach"):
    """Batched Bregman projection of ``K`` (_utils.py:46-162).

    Returns ``{"T", "n_iters", "potentials": (f, g)}``.
    """
    ing, niter, T, _, pots, _ = _projection(K, a, b, max_iter, tol, _lib.SCALING, grad)
    return {"T": T, "n_iters": niter, "potentials": pots}


# --------------------------------------------------------------- costs

class _SqEuclidean(torch.autograd.Function):
    """``|x|^2 + |y|^2 - 2 x.y`` per pair on the library kernel, with the
    closed-form backward ``dX = 2 (rowsum(G) X - G Y)``, ``dY = 2 (colsum(G) Y
    - G^T X)`` on the library's batched GEMM."""

    @staticmethod
    def forward(ctx, X, Y):
        Xc, Yc = X.detach().contiguous(), Y.detach().contiguous()
        ctx.save_for_backward(Xc, Yc)
        return _solve.sqeuclidean(Xc, Yc, squared=True, clamp=False)

    @staticmethod
    def backward(ctx, G):
        X, Y = ctx.saved_tensors
        G = G.contiguous().to(X.dtype)
        dX = dY = None
        if ctx.needs_input_grad[0]:
            dX = (2.0 * G.sum(2, keepdim=True)) * X
            _solve.gemm(G, Y, dX, alpha=-2.0, beta=1.0)
        if ctx.needs_input_grad[1]:
            dY = (2.0 * G.sum(1).unsqueeze(-1)) * Y
            _solve.gemm(G, X, dY, alpha=-2.0, beta=1.0, transA=True)
        return dX, dY


METRIC_MINKOWSKI, METRIC_KL, METRIC_KL_LOGITS = 0, 1, 2


def _pairwise(X, Y, metric, p=2.0, q=1.0, eps=1e-10):
    """Batched pairwise cost on the library kernel (potamd_pairwise_batch)."""
    lib = _lib.load()
    X = X.contiguous()
    Y = Y.contiguous()
    B, n, d = X.shape
    m = Y.shape[1]
    if (Y.dim() != 3 or Y.shape[0] != B or Y.shape[2] != d or X.dtype != Y.dtype
            or X.dtype not in (torch.float32, torch.float64) or X.device != Y.device):
        raise ValueError(f"pairwise cost: X {tuple(X.shape)} {X.dtype} and Y "
                         f"{tuple(Y.shape)} {Y.dtype} do not agree")
    M = torch.empty((B, n, m), dtype=X.dtype, device=X.device)
    code = _lib.F32 if X.dtype == torch.float32 else _lib.F64
    rc = lib.potamd_pairwise_batch(
        code, metric, ctypes.c_void_p(X.data_ptr()), ctypes.c_void_p(Y.data_ptr()), B, n, m,
        d, float(p), float(q), float(eps), ctypes.c_void_p(M.data_ptr()),
        ctypes.c_void_p(_device.stream_ptr(X.device)))
    _lib.check(rc, "potamd_pairwise_batch")
    return M


class _Pairwise(torch.autograd.Function):
    """Minkowski / KL costs on the library kernel; the backward passes are the
    closed-form derivatives of the formulas of _linear.py:20-51 and :88-120."""

    @staticmethod
    def forward(ctx, X, Y, metric, p, q, eps):
        Xc, Yc = X.detach().contiguous(), Y.detach().contiguous()
        M = _pairwise(Xc, Yc, metric, p, q, eps)
        ctx.save_for_backward(Xc, Yc, M)
        ctx.args = (metric, p, q, eps)
        return M

    @staticmethod
    def backward(ctx, G):
        X, Y, M = ctx.saved_tensors
        metric, p, q, eps = ctx.args
        dX = dY = None
        if metric == METRIC_MINKOWSKI:
            # M = S^(q/p), S = sum_d |x - y|^p: dM/dx = q S^(q/p - 1) |x-y|^(p-1) sign(x-y)
            diff = X[:, :, None, :] - Y[:, None, :, :]
            S = M ** (p / q) if q != p else M
            coef = G * q * torch.where(S > 0, S ** (q / p - 1.0), torch.zeros_like(S))
            w = coef[..., None] * torch.abs(diff) ** (p - 1.0) * torch.sign(diff)
            if ctx.needs_input_grad[0]:
                dX = w.sum(2)
            if ctx.needs_input_grad[1]:
                dY = -w.sum(1)
        else:
            # M = sum_d y log(y + eps) - sum_d y h(x), h(x) = log(x + eps) or x
            if ctx.needs_input_grad[0]:
                gy = torch.matmul(G, Y)                        # (B, n, d)
                dh = 1.0 / (X + eps) if metric == METRIC_KL else torch.ones_like(X)
                dX = -gy * dh
            if ctx.needs_input_grad[1]:
                hX = torch.log(X + eps) if metric == METRIC_KL else X
                cs = G.sum(1).unsqueeze(-1)                   # (B, m, 1)
                dY = cs * (torch.log(Y + eps) + Y / (Y + eps)) \
                    - torch.matmul(G.transpose(1, 2).contiguous(), hX)
        return dX, dY, None, None, None, None


def _cost(X, Y, fn):
    """Run a cost on the device and return it in the caller's kind; torch
    inputs keep their autograd graph."""
    ing = Ingress(X, Y)
    if ing.is_torch and requires_grad(X, Y):
        Xd = X if X.device == ing.device and X.dtype == ing.dtype else X.to(ing.device,
                                                                            ing.dtype)
        Yd = Y if Y.device == ing.device and Y.dtype == ing.dtype else Y.to(ing.device,
                                                                            ing.dtype)
        return ing.out(fn(Xd, Yd))
    return ing.out(fn(ing.dev(X), ing.dev(Y)))


def dist_euclidean_batch(X, Y, squared=True, nx=None):
    """``|x|^2 + |y|^2 - 2 x.y`` per pair, not clamped (_linear.py:54-85)."""
    def fn(Xd, Yd):
        M = _SqEuclidean.apply(Xd, Yd)
        return M if squared else torch.sqrt(M)
    return _cost(X, Y, fn)


def dist_lp_batch(X, Y, p=2, q=1, nx=None):
    """``(sum_d |x - y|^p)^(q/p)`` per pair (_linear.py:20-51)."""
    return _cost(X, Y, lambda Xd, Yd: _Pairwise.apply(Xd, Yd, METRIC_MINKOWSKI, float(p),
                                                      float(q), 0.0))


def dist_kl_batch(X, Y, logits_X=False, nx=None, eps=1e-10):
    """``KL(y_j || x_i) = sum_d y log(y + eps) - y log(x + eps)`` per pair
    (_linear.py:88-120).

    With ``logits_X`` the second term is ``sum_d y x`` (X holds log
    probabilities).  The reference's logits branch (:115-116) omits that sum
    over d, so its (B, 1, m) minus (B, n, m, d) broadcast raises unless the
    sizes happen to line up; the sum here is the formula of its docstring.
    """
    metric = METRIC_KL_LOGITS if logits_X else METRIC_KL
    return _cost(X, Y, lambda Xd, Yd: _Pairwise.apply(Xd, Yd, metric, 0.0, 0.0, float(eps)))


def dist_batch(X1, X2=None, metric="sqeuclidean", p=2, nx=None):
    """Batched cost matrices (_linear.py:178-233)."""
    X2 = X2 if X2 is not None else X1
    metric = metric.lower()
    if metric == "sqeuclidean":
        return dist_euclidean_batch(X1, X2, squared=True)
    if metric == "euclidean":
        return dist_euclidean_batch(X1, X2, squared=False)
    if metric == "minkowski":
        return dist_lp_batch(X1, X2, p=p, q=1)
    if metric == "kl":
        return dist_kl_batch(X1, X2, logits_X=False)
    raise ValueError(f"Unknown metric: {metric}")


def loss_linear_samples_batch(X, Y, T, metric="l2"):
    """``dist_batch`` then ``loss_linear_batch`` (_linear.py:150-175)."""
    return loss_linear_batch(dist_batch(X, Y, metric=metric), T)


# --------------------------------------------------------------- solver

def solve_batch(M, reg, a=None, b=None, max_iter=1000, tol=1e-5, solver="log_sinkhorn",
                reg_type="entropy", grad="envelope"):
    """Batched entropic OT (_linear.py:236-357); returns an ``OTResult``.

    ``value_linear = <M, T>``; ``value`` adds ``reg * sum T log(T + 1e-16)``
    (``reg_type="entropy"``) or ``reg * KL(T | a b^T)`` (``"kl"``).  Gradient
    modes as in the module docstring.
    """
    if solver == "log_sinkhorn":
        domain = _lib.LOG
    elif solver == "sinkhorn":
        domain = _lib.SCALING
    else:
        raise ValueError(f"Unknown solver: {solver}")
    kl = reg_type.lower() == "kl"
    if not kl and reg_type.lower() != "entropy":
        raise ValueError(f"Unknown reg_type: {reg_type}")
    has_grad = requires_grad(M, a, b)
    if grad == "autodiff" and has_grad:
        raise _autodiff_error()
    ing = Ingress(M, a, b)
    Mt = ing.dev(M)
    B, n, m = Mt.shape
    at = _weights_dev(a, ing, (B, n))
    bt = _weights_dev(b, ing, (B, m))
    init = 0.0 if domain == _lib.LOG else 1.0
    u0 = torch.full((B, n), init, dtype=torch.float64, device=ing.device)
    v0 = torch.full((B, m), init, dtype=torch.float64, device=ing.device)
    res = _solve.dense_solve(Mt, at, bt, u0, v0, reg, domain=domain, order=_lib.ROW_FIRST,
                             max_iter=max_iter, stop_thr=tol, err_norm=0)
    u, v = res.u, res.v
    log = {"n_iter": res.niter}

    if grad == "last_step":
        if has_grad:
            Mc = M if isinstance(M, torch.Tensor) else torch.as_tensor(M, device=ing.device)
            Kc = -Mc / reg if domain == _lib.LOG else torch.exp(-Mc / reg)
            a_c = _caller_weights(a, at, Mc)
            b_c = _caller_weights(b, bt, Mc)
            T, _, un, vn = _last_step(Kc, u, v, a_c, b_c, domain)
            value_linear = torch.sum(Mc * T, dim=(1, 2))
            if kl:
                ref = a_c[:, :, None] * b_c[:, None, :]
                value = value_linear + reg * torch.sum(T * torch.log(T / ref + 1e-16), dim=(1, 2))
            else:
                value = value_linear + reg * torch.sum(T * torch.log(T + 1e-16), dim=(1, 2))
            return OTResult(value=value, value_linear=value_linear, potentials=(un, vn),
                            plan=T, log=log)
        res1 = _solve.dense_solve(Mt, at, bt, u, v, reg, domain=domain, order=_lib.ROW_FIRST,
                                  max_iter=1, stop_thr=0.0, err_norm=0)
        u, v = res1.u, res1.v

    P, _, sums = _solve.plan(Mt, u, v, reg, domain=domain, a=at, b=bt)
    value_linear = sums[:, 0]
    value = value_linear + reg * (sums[:, 2] if kl else sums[:, 1])
    potentials = (ing.out(u), ing.out(v))
    T = ing.out(P)
    if has_grad and grad == "envelope":
        # the plan is held fixed (_linear.py:334-336): d<M, T>/dM = T; the KL
        # reference measure a b^T stays differentiable
        odt, odev = ing.out_dtype, ing.out_device
        vl = batch_envelope(value_linear.to(device=odev, dtype=odt), ((M, P),))
        pairs = [(M, P)]
        if kl and requires_grad(a, b):
            w = P / (at[:, :, None] * bt[:, None, :])
            t = P * w / (w + 1e-16)
            pairs += [(a, -reg * t.sum(2) / at), (b, -reg * t.sum(1) / bt)]
        val = batch_envelope(value.to(device=odev, dtype=odt), pairs)
        return OTResult(value=val, value_linear=vl, potentials=potentials, plan=T, log=log)
    return OTResult(value=ing.out(value), value_linear=ing.out(value_linear),
                    potentials=potentials, plan=T, log=log)


def solve_sample_batch(X_a, X_b, reg, a=None, b=None, metric="sqeuclidean", p=2,
                       max_iter=1000, tol=1e-5, solver="log_sinkhorn", reg_type="entropy",
                       grad="envelope"):
    """``dist_batch`` then ``solve_batch`` (_linear.py:360-435)."""
    M = dist_batch(X_a, X_b, metric=metric, p=p)
    return solve_batch(M, reg, a=a, b=b, max_iter=max_iter, tol=tol, solver=solver,
                       reg_type=reg_type, grad=grad)


# Quadratic (Gromov-Wasserstein) batch solvers share this namespace, as in
# ot.batch (ot/batch/__init__.py).
from .quadratic import (  # noqa: E402
    loss_quadratic_batch,
    loss_quadratic_samples_batch,
    solve_gromov_batch,
    tensor_batch,
)
