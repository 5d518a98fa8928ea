"""Drop-in ``ot.batch`` linear solvers.

Same names, arguments and outputs as ot/batch/_linear.py and
ot/batch/_utils.py; the Sinkhorn iterations, plans and per-problem sums run
in the gfx950 HIP library.

* ``bregman_log_projection_batch``  _utils.py:165-290 (rows first, one stop on
  the largest row-marginal error of the batch, every 10th iteration)
* ``bregman_projection_batch``      _utils.py:46-162
* ``solve_batch``                   _linear.py:236-357
* ``solve_sample_batch``            _linear.py:360-435
* ``dist_batch`` and its metrics    _linear.py:20-233
* ``loss_linear_batch``, ``loss_linear_samples_batch``, ``entropy_batch``,
  ``norm_batch``                    _linear.py:123-175, _utils.py:4-29
"""

import numpy as np
import torch

from . import _lib, _solve
from ._device import Ingress
from .utils import OTResult


def _backend_of(*xs):
    for x in xs:
        if isinstance(x, torch.Tensor):
            return "torch"
    return "numpy"


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


def _projection(K, a, b, max_iter, tol, domain, grad):
    ing = Ingress(K, a, b)
    Kt = ing.dev(K)
    B, n, m = Kt.shape
    at = torch.full((B, n), 1.0 / n, dtype=torch.float64, device=ing.device) \
        if a is None else ing.dev(a, torch.float64)
    bt = torch.full((B, m), 1.0 / m, dtype=torch.float64, device=ing.device) \
        if b is None else ing.dev(b, torch.float64)
    init = 0.0 if domain == _lib.LOG else 1.0
    u0 = torch.full((B, n), init, dtype=torch.float64, device=ing.device)
    v0 = torch.full((B, m), init, dtype=torch.float64, device=ing.device)
    res = _solve.dense_solve(Kt, at, bt, u0, v0, None, domain=domain,
                             order=_lib.ROW_FIRST, kernel_given=True,
                             max_iter=max_iter, stop_thr=tol, err_norm=0)
    P, L, _ = _solve.plan(Kt, res.u, res.v, None, domain=domain, kernel_given=True,
                          want_log=domain == _lib.LOG, want_sums=False)
    return ing, res, P, L


def bregman_log_projection_batch(K, a=None, b=None, nx=None, max_iter=10000, tol=1e-5,
                                 grad="detach"):
    """Batched log-domain Bregman projection of ``exp(K)`` (_utils.py:165-290).

    Returns ``{"T", "log_T", "n_iters", "potentials": (u, v)}``.
    """
    ing, res, P, L = _projection(K, a, b, max_iter, tol, _lib.LOG, grad)
    return {"T": ing.out(P), "log_T": ing.out(L), "n_iters": res.niter,
            "potentials": (ing.out(res.u), ing.out(res.v))}


def bregman_projection_batch(K, a=None, b=None, nx=None, max_iter=10000, tol=1e-5,
                             grad="detach"):
    """Batched Bregman projection of ``K`` (_utils.py:46-162).

    Returns ``{"T", "n_iters", "potentials": (f, g)}``.
    """
    ing, res, P, _ = _projection(K, a, b, max_iter, tol, _lib.SCALING, grad)
    return {"T": ing.out(P), "n_iters": res.niter,
            "potentials": (ing.out(res.u), ing.out(res.v))}


def dist_euclidean_batch(X, Y, squared=True, nx=None):
    """``|x|^2 + |y|^2 - 2 x.y`` per pair, not clamped (_linear.py:54-85)."""
    ing = Ingress(X, Y)
    M = _solve.sqeuclidean(ing.dev(X), ing.dev(Y), squared=True, clamp=False)
    if not squared:
        M = torch.sqrt(M)
    return ing.out(M)


def dist_lp_batch(X, Y, p=2, q=1, nx=None):
    """``(sum |x - y|^p)^(q/p)`` (_linear.py:20-51), on the device."""
    ing = Ingress(X, Y)
    Xt, Yt = ing.dev(X), ing.dev(Y)
    M = (torch.abs(Xt[:, :, None] - Yt[:, None, :]) ** p).sum(-1)
    if q != p:
        M = M ** (q / p)
    return ing.out(M)


def dist_kl_batch(X, Y, logits_X=False, nx=None, eps=1e-10):
    """``KL(y_j || x_i)`` per pair (_linear.py:88-120), on the device."""
    ing = Ingress(X, Y)
    Xt, Yt = ing.dev(X), ing.dev(Y)
    entr_y = torch.sum(Yt * torch.log(Yt + eps), dim=-1)
    if logits_X:
        M = entr_y[:, None, :] - torch.sum(Yt[:, None, :, :] * Xt[:, :, None, :], dim=-1)
    else:
        M = entr_y[:, None, :] - torch.sum(
            Yt[:, None, :, :] * torch.log(Xt + eps)[:, :, None, :], dim=-1)
    return ing.out(M)


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


def solve_batch(M, reg, a=None, b=None, max_iter=1000, tol=1e-5, solver="log_sinkhorn",
                reg_type="entropy", grad="envelope"):
    """Batched entropic OT (_linear.py:236-357); returns an ``OTResult``.

    ``value_linear = <M, T>``; ``value`` adds ``reg * sum T log(T + 1e-16)``
    (``reg_type="entropy"``) or ``reg * KL(T | a b^T)`` (``"kl"``).  With torch
    inputs that require grad and ``grad="envelope"``, ``value`` and
    ``value_linear`` carry the envelope gradient ``d/dM = T`` (T detached).
    """
    if solver == "log_sinkhorn":
        domain = _lib.LOG
    elif solver == "sinkhorn":
        domain = _lib.SCALING
    else:
        raise ValueError(f"Unknown solver: {solver}")
    if reg_type.lower() not in ("entropy", "kl"):
        raise ValueError(f"Unknown reg_type: {reg_type}")
    ing = Ingress(M, a, b)
    Mt = ing.dev(M)
    B, n, m = Mt.shape
    at = torch.full((B, n), 1.0 / n, dtype=torch.float64, device=ing.device) \
        if a is None else ing.dev(a, torch.float64)
    bt = torch.full((B, m), 1.0 / m, dtype=torch.float64, device=ing.device) \
        if b is None else ing.dev(b, torch.float64)
    init = 0.0 if domain == _lib.LOG else 1.0
    u0 = torch.full((B, n), init, dtype=torch.float64, device=ing.device)
    v0 = torch.full((B, m), init, dtype=torch.float64, device=ing.device)
    res = _solve.dense_solve(Mt, at, bt, u0, v0, reg, domain=domain, order=_lib.ROW_FIRST,
                             max_iter=max_iter, stop_thr=tol, err_norm=0)
    P, _, sums = _solve.plan(Mt, res.u, res.v, reg, domain=domain, a=at, b=bt)
    T = ing.out(P)
    value_linear = sums[:, 0]
    reg_term = sums[:, 1] if reg_type.lower() == "entropy" else sums[:, 2]
    value = value_linear + reg * reg_term
    if ing.is_torch and grad == "envelope" and isinstance(M, torch.Tensor) and M.requires_grad:
        vl = torch.sum(M * T.detach(), dim=(1, 2))
        value_linear = vl
        value = vl + (reg * reg_term).to(dtype=vl.dtype, device=vl.device)
    else:
        value_linear = ing.out(value_linear)
        value = ing.out(value)
    potentials = (ing.out(res.u), ing.out(res.v))
    return OTResult(value=value, value_linear=value_linear, potentials=potentials,
                    plan=T, log={"n_iter": res.niter})


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
