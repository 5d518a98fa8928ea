"""``ot.solve`` / ``ot.solve_sample`` on the balanced entropic route.

Reference: ot/solvers.py:54-604 (``solve``): with ``reg > 0``, no
``unbalanced`` and ``reg_type`` "entropy" or "KL" (case-insensitive) it runs
``sinkhorn_log`` with ``max_iter = 1000``, ``tol = 1e-9`` by default
(:428-470) and returns an ``OTResult`` whose

* ``potentials`` are ``(log_u, log_v)``,
* ``value_linear = <M, P>``,
* ``value = value_linear + reg * sum P log(P + 1e-16)`` ("entropy") or
  ``value_linear + reg * sum P log(P / (a b^T) + 1e-16)`` ("KL"),
* ``status`` is None,

with the gradient modes of :429-495: "autodiff" (plain), "detach",
"envelope" (``d value / d(M, a, b) = (P, reg (u - mean u), reg (v - mean v))``)
and "last_step" (``max_iter - 1`` iterations, then one more differentiable
update).  ``solve_sample`` (:1354-1751, non-lazy branch) computes
``M = dist(X_a, X_b, metric)`` and calls ``solve``.

The iterations, the plan and the fp64 sums run in the HIP library; the only
torch arithmetic is the one differentiable update of "last_step" on inputs
that require grad.  Other routes of the reference (exact EMD, unbalanced,
L2 and general regularisers, lazy/low-rank/factored/Gaussian/geomloss
methods) raise ``NotImplementedError``.
"""

import warnings

import numpy as np
import torch

from . import _lib, _solve
from ._grad import envelope, requires_grad as _requires_grad
from ._device import Ingress
from .bregman import NOT_CONVERGED, _print_errs
from .utils import OTResult, dist

LAZY_METHODS = ("1d", "gaussian", "lowrank", "nystroem", "factored", "geomloss",
                "geomloss_auto", "geomloss_tensorized", "geomloss_online",
                "geomloss_multiscale")


def solve(M, a=None, b=None, reg=None, c=None, reg_type="KL", unbalanced=None,
          unbalanced_type="KL", method=None, n_threads=1, max_iter=None, plan_init=None,
          potentials_init=None, tol=None, verbose=False, grad="autodiff"):
    """Entropic OT through ``sinkhorn_log`` on the GPU; see the module docstring."""
    if reg is None or reg == 0:
        raise NotImplementedError("exact OT (reg = 0, the EMD solver) is not part of this "
                                  "build; use reg > 0")
    if unbalanced is not None:
        raise NotImplementedError("unbalanced OT is not part of this build")
    if isinstance(reg_type, tuple) or reg_type.lower() not in ("entropy", "kl"):
        raise NotImplementedError(f'Not implemented reg_type="{reg_type}"')
    if c is not None:
        raise NotImplementedError("a reference measure c is not supported on this route")
    if max_iter is None:
        max_iter = 1000
    if tol is None:
        tol = 1e-9
    has_grad = _requires_grad(M, a, b)
    if has_grad and grad == "autodiff":
        raise NotImplementedError(
            "grad='autodiff' would differentiate through the GPU iterations; use "
            "grad='envelope', 'last_step' or 'detach'")
    if grad == "last_step":
        if max_iter == 0:
            raise ValueError("The maximum number of iterations must be greater than 0 "
                             "when using grad=last_step.")
        max_iter = max_iter - 1

    ing = Ingress(M, a, b)
    Mt = ing.dev(M)
    n, m = Mt.shape
    at = torch.full((n,), 1.0 / n, dtype=torch.float64, device=ing.device) \
        if a is None else ing.dev(a, torch.float64).reshape(n)
    bt = torch.full((m,), 1.0 / m, dtype=torch.float64, device=ing.device) \
        if b is None else ing.dev(b, torch.float64).reshape(m)
    u = torch.zeros((1, n), dtype=torch.float64, device=ing.device)
    v = torch.zeros((1, m), dtype=torch.float64, device=ing.device)
    niter = -1
    if max_iter > 0:
        res = _solve.dense_solve(Mt, at.view(1, n), bt.view(1, m), u, v, reg,
                                 domain=_lib.LOG, order=_lib.COL_FIRST, max_iter=max_iter,
                                 stop_thr=tol)
        if verbose:
            _print_errs(res.errs)
        if res.status == _lib.MAX_ITER:
            warnings.warn(NOT_CONVERGED)
        u, v, niter = res.u, res.v, res.niter

    if grad == "last_step":
        # One more update from the detached potentials (ot/solvers.py:463-475):
        # v from the last u, u from the last v (not from the new v), so the
        # forward value of u is the last u itself and only v moves.
        if has_grad:
            M0 = M if isinstance(M, torch.Tensor) else torch.as_tensor(M)
            a0 = a if a is not None else at.to(device=M0.device, dtype=M0.dtype)
            b0 = b if b is not None else bt.to(device=M0.device, dtype=M0.dtype)
            u_d = u[0].to(device=M0.device, dtype=M0.dtype)
            v_d = v[0].to(device=M0.device, dtype=M0.dtype)
            vn = torch.log(b0) - torch.logsumexp(-M0 / reg + u_d[:, None], 0)
            un = torch.log(a0) - torch.logsumexp(-M0 / reg + v_d[None, :], 1)
            plan = torch.exp(-M0 / reg + un[:, None] + vn[None, :])
            # the values use the detached inputs (:435, :476-481): the
            # gradient reaches M, a and b through the plan only
            Md = M0.detach()
            value_linear = torch.sum(Md * plan)
            if reg_type.lower() == "entropy":
                value = value_linear + reg * torch.sum(plan * torch.log(plan + 1e-16))
            else:
                ab = a0.detach()[:, None] * b0.detach()[None, :]
                value = value_linear + reg * torch.sum(plan * torch.log(plan / ab + 1e-16))
            return OTResult(potentials=(un, vn), value=value, value_linear=value_linear,
                            plan=plan, status=None)
        # Without gradients the same step is the column half-step of one more
        # iteration: a one-iteration solve from u returns v = c-transform(u)
        # (the solver updates v first); its u is discarded.
        res = _solve.dense_solve(Mt, at.view(1, n), bt.view(1, m), u, v, reg,
                                 domain=_lib.LOG, order=_lib.COL_FIRST, max_iter=1,
                                 stop_thr=0.0)
        v = res.v
        niter = max_iter + 1

    P, _, sums = _solve.plan(Mt, u, v, reg, domain=_lib.LOG, a=at.view(1, n),
                             b=bt.view(1, m))
    value_linear = float(sums[0, 0])
    reg_term = float(sums[0, 1] if reg_type.lower() == "entropy" else sums[0, 2])
    value = value_linear + reg * reg_term
    plan = ing.out(P[0])
    log_u, log_v = ing.out(u[0]), ing.out(v[0])
    if ing.is_torch and grad == "envelope" and has_grad:
        # d value / d(M, a, b) = (P, reg (u - mean u), reg (v - mean v)), :483-495
        pairs = ((M, P[0]), (a, reg * (u[0] - u[0].mean())), (b, reg * (v[0] - v[0].mean())))
        value_t = envelope(value, pairs, ing.out_dtype, ing.out_device)
        vl_t = torch.tensor(value_linear, dtype=ing.out_dtype, device=ing.out_device)
        return OTResult(potentials=(log_u, log_v), value=value_t, value_linear=vl_t,
                        plan=plan, status=None)
    return OTResult(potentials=(log_u, log_v), value=ing.scalar(value),
                    value_linear=ing.scalar(value_linear), plan=plan, status=None)


def solve_sample(X_a, X_b, a=None, b=None, metric="sqeuclidean", reg=None, c=None,
                 reg_type="KL", unbalanced=None, unbalanced_type="KL", lazy=False,
                 batch_size=None, method=None, n_threads=1, max_iter=None, plan_init=None,
                 rank=100, scaling=0.95, potentials_init=None, X_init=None, tol=None,
                 verbose=False, grad="autodiff", random_state=None):
    """``solve`` on ``M = dist(X_a, X_b, metric)`` (ot/solvers.py:1729-1751)."""
    if lazy or (method is not None and method.lower() in LAZY_METHODS):
        raise NotImplementedError("lazy and low-rank sample solvers are not part of this "
                                  "build; pot_amd.online.sinkhorn_online runs sinkhorn_log "
                                  "on the squared Euclidean cost without storing it")
    M = dist(X_a, X_b, metric)
    return solve(M, a, b, reg, c, reg_type, unbalanced, unbalanced_type, method, n_threads,
                 max_iter, plan_init, potentials_init, tol, verbose, grad)
