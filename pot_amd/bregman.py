"""Drop-in entropic Sinkhorn solvers (the ``ot.bregman`` / ``ot.sinkhorn`` API).

Same names, arguments, defaults, warnings and ``log`` contents as
ot/bregman/_sinkhorn.py:21-918; the iterations run in the gfx950 HIP library.

* ``sinkhorn``        _sinkhorn.py:21-234   method dispatch
* ``sinkhorn2``       _sinkhorn.py:237-484  loss variant
* ``sinkhorn_knopp``  _sinkhorn.py:487-687  scaling form
* ``sinkhorn_log``    _sinkhorn.py:690-918  log-domain form
* ``greenkhorn``      _sinkhorn.py:921-1102 greedy form

``sinkhorn_stabilized`` and ``sinkhorn_epsilon_scaling`` live in
``pot_amd.stabilized``; unknown methods raise ``ValueError`` like the
reference (:233-234).
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _device, _lib, _solve
from ._device import Ingress, is_empty
from ._grad import envelope, requires_grad
from .utils import list_to_array

NOT_CONVERGED = (
    "Sinkhorn did not converge. You might want to "
    "increase the number of iterations `numItermax` "
    "or the regularization parameter `reg`."
)


def _print_errs(errs, start=0, step=10):
    """The verbose table of _sinkhorn.py:655-658 / :894-897, printed after the
    solve (the device does not stream it)."""
    for k, e in enumerate(errs):
        ii = start + step * k
        if ii % 200 == 0:
            print("{:5s}|{:12s}".format("It.", "Err") + "\n" + "-" * 19)
        print("{:5d}|{:8e}|".format(ii, e))


def _method_error(method):
    return ValueError("Unknown method '%s'." % method)


def sinkhorn(a, b, M, reg, method="sinkhorn", numItermax=1000, stopThr=1e-9,
             verbose=False, log=False, warn=True, warmstart=None, **kwargs):
    """Entropic OT plan; dispatches on ``method.lower()`` (_sinkhorn.py:164-234)."""
    name = method.lower()
    if name == "sinkhorn":
        return sinkhorn_knopp(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                              verbose=verbose, log=log, warn=warn,
                              warmstart=warmstart, **kwargs)
    if name == "sinkhorn_log":
        return sinkhorn_log(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                            verbose=verbose, log=log, warn=warn,
                            warmstart=warmstart, **kwargs)
    if name == "greenkhorn":
        return greenkhorn(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                          verbose=verbose, log=log, warn=warn, warmstart=warmstart)
    if name == "sinkhorn_stabilized":
        from .stabilized import sinkhorn_stabilized
        return sinkhorn_stabilized(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                                   warmstart=warmstart, verbose=verbose, log=log, warn=warn,
                                   **kwargs)
    if name == "sinkhorn_epsilon_scaling":
        from .stabilized import sinkhorn_epsilon_scaling
        return sinkhorn_epsilon_scaling(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                                        warmstart=warmstart, verbose=verbose, log=log,
                                        warn=warn, **kwargs)
    raise _method_error(method)


def sinkhorn2(a, b, M, reg, method="sinkhorn", numItermax=1000, stopThr=1e-9,
              verbose=False, log=False, warn=False, warmstart=None, **kwargs):
    """Entropic OT loss ``<P, M>`` (_sinkhorn.py:387-484); ``warn`` defaults to False.

    Gradients: with torch inputs that require grad (one histogram ``b``), the
    loss carries the envelope gradient of the entropic problem at the
    returned potentials, ``d/dM = P``, ``d/da = reg (log u - mean)``,
    ``d/db = reg (log v - mean)`` (the ``set_gradients`` of ot/solvers.py:
    483-495).  The reference instead differentiates through its torch
    iterations (test/test_bregman.py:203-220 only checks that the gradients
    exist and have the inputs' shapes); the iterations here run in the HIP
    library, so that derivative is not available.
    """
    M, a, b = list_to_array(M, a, b)
    name = method.lower()
    if name == "sinkhorn_stabilized":
        from .stabilized import sinkhorn_stabilized
        return sinkhorn_stabilized(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                                   warmstart=warmstart, verbose=verbose, log=log, warn=warn,
                                   _loss_only=True, **kwargs)
    if name not in ("sinkhorn", "sinkhorn_log"):
        raise _method_error(method)
    solver = sinkhorn_knopp if name == "sinkhorn" else sinkhorn_log
    if requires_grad(a, b, M):
        if len(b.shape) >= 2:
            raise NotImplementedError("gradients of sinkhorn2 with several histograms are "
                                      "not part of this build")
        parts = solver(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                       verbose=verbose, log=log, warn=warn, warmstart=warmstart,
                       _grad_parts=True, **kwargs)
        ing, loss, P, lu, lv, logd = parts
        pairs = ((M, P), (a, reg * (lu - lu.mean())), (b, reg * (lv - lv.mean())))
        val = envelope(loss, pairs, ing.out_dtype, ing.out_device)
        return (val, logd) if log else val
    return solver(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                  verbose=verbose, log=log, warn=warn, warmstart=warmstart,
                  _loss_only=True, **kwargs)


def _prep(a, b, M):
    a, b, M = list_to_array(a, b, M)
    ing = Ingress(M, a, b)
    Mt = ing.dev(M)
    n, m = Mt.shape
    at = torch.full((n,), 1.0 / n, dtype=torch.float64, device=ing.device) \
        if is_empty(a) else ing.dev(a, torch.float64)
    bt = torch.full((m,), 1.0 / m, dtype=torch.float64, device=ing.device) \
        if is_empty(b) else ing.dev(b, torch.float64)
    return ing, Mt, at, bt


def greenkhorn(a, b, M, reg, numItermax=10000, stopThr=1e-9, verbose=False, log=False,
               warn=True, warmstart=None):
    """Greedy Sinkhorn (_sinkhorn.py:921-1102), the whole loop in one GPU launch.

    Each iteration updates the one row or column scaling whose marginal
    violates its weight the most (the first index on a tie), keeps both
    violation vectors up to date incrementally, and stops once the largest
    violation read before the update is ``<= stopThr`` (after that update, as
    the reference).  Computed in float64 whatever the cost's dtype; returns
    the plan ``u_i K_ij v_j`` in M's dtype, and with ``log`` the dict
    ``{"u", "v", "n_iter"}`` (scalings, not their logarithms).  ``verbose`` is
    accepted and, as in the reference, prints nothing.
    """
    ing, Mt, at, bt = _prep(a, b, M)
    n, m = Mt.shape
    if bt.dim() > 1:
        raise ValueError("greenkhorn takes one target histogram b (the reference's loop "
                         "indexes b as a vector)")
    dev = ing.device
    if warmstart is None:
        u = torch.full((n,), 1.0 / n, dtype=torch.float64, device=dev)
        v = torch.full((m,), 1.0 / m, dtype=torch.float64, device=dev)
    else:
        u = torch.exp(ing.dev(warmstart[0], torch.float64).reshape(n))
        v = torch.exp(ing.dev(warmstart[1], torch.float64).reshape(m))
    lib = _lib.load()
    code = _lib.F32 if Mt.dtype == torch.float32 else _lib.F64
    ws_bytes = lib.potamd_greenkhorn_workspace_size(n, m)
    if ws_bytes == 0:
        raise ValueError(f"greenkhorn: unsupported size {n} x {m}")
    ws = _device.workspace(ws_bytes, dev)
    P = torch.empty((n, m), dtype=torch.float64, device=dev)
    info = (ctypes.c_int * 4)()
    p = _solve._ptr
    rc = lib.potamd_greenkhorn(code, p(Mt), Mt.stride(0), n, m, float(reg), p(at), p(bt), p(u),
                               p(v), int(numItermax), float(stopThr), p(P), p(ws), ws_bytes,
                               ctypes.c_void_p(_device.stream_ptr(dev)), info)
    _lib.check(rc, "potamd_greenkhorn")
    if rc == _lib.MAX_ITER and warn:
        warnings.warn(NOT_CONVERGED)
    G = ing.out(P)
    if log:
        if info[0] < 0:
            raise UnboundLocalError("numItermax must be positive when log=True")
        return G, {"u": ing.out(u), "v": ing.out(v), "n_iter": int(info[0])}
    return G


def sinkhorn_log(a, b, M, reg, numItermax=1000, stopThr=1e-9, verbose=False,
                 log=False, warn=True, warmstart=None, _loss_only=False, _grad_parts=False,
                 **kwargs):
    """Log-domain Sinkhorn (_sinkhorn.py:799-918), iterations on the GPU.

    Each iteration: ``v = log b - LSE_i(-M/reg + u)``, ``u = log a -
    LSE_j(-M/reg + v)``; every 10th the column-marginal error is logged and
    compared (strict ``<``) with ``stopThr``.  Returns the plan (or, from
    ``sinkhorn2``, the loss ``<P, M>``) and optionally the log dict with
    ``err``, ``niter``, ``log_u``, ``log_v``, ``u``, ``v``.
    """
    ing, Mt, at, bt = _prep(a, b, M)
    n, m = Mt.shape
    if bt.dim() > 1:
        nh = bt.shape[1]
        if nh == 1 and warmstart is None:
            # the reference builds per-column warm starts only for n_hists > 1
            # and then indexes None (_sinkhorn.py:816, :834)
            raise TypeError("'NoneType' object is not subscriptable")
        ws = warmstart if warmstart is not None else [None] * nh
        if numItermax > 0 and not verbose:
            return _sinkhorn_log_columns(ing, Mt, at, bt, reg, numItermax, stopThr, log, ws)
        # Multi-histogram b, verbose or numItermax <= 0: one solve per column,
        # warn not forwarded (_sinkhorn.py:816-854).
        b_np = b if not isinstance(b, list) else np.asarray(b)
        losses, lus, lvs = [], [], []
        for k in range(nh):
            bk = b_np[:, k]
            res = sinkhorn_log(a, bk, M, reg, numItermax=numItermax, stopThr=stopThr,
                               verbose=verbose, log=log, warmstart=ws[k],
                               _loss_only=True)
            if log:
                losses.append(res[0])
                lus.append(res[1]["log_u"])
                lvs.append(res[1]["log_v"])
            else:
                losses.append(res)
        if ing.is_torch:
            out = torch.stack([torch.as_tensor(x) for x in losses])
        else:
            out = np.stack(losses)
        if log:
            if ing.is_torch:
                lu, lv = torch.stack(lus, 1), torch.stack(lvs, 1)
                logd = {"log_u": lu, "log_v": lv, "u": torch.exp(lu), "v": torch.exp(lv)}
            else:
                lu, lv = np.stack(lus, 1), np.stack(lvs, 1)
                logd = {"log_u": lu, "log_v": lv, "u": np.exp(lu), "v": np.exp(lv)}
            return out, logd
        return out

    if warmstart is None:
        u0 = torch.zeros((1, n), dtype=torch.float64, device=ing.device)
        v0 = torch.zeros((1, m), dtype=torch.float64, device=ing.device)
    else:
        u0 = ing.dev(warmstart[0], torch.float64).view(1, n)
        v0 = ing.dev(warmstart[1], torch.float64).view(1, m)

    if numItermax <= 0:
        if warn:
            warnings.warn(NOT_CONVERGED)
        if log:
            raise UnboundLocalError("numItermax must be positive when log=True")
        u, v, errs, niter = u0, v0, [], -1
    else:
        res = _solve.dense_solve(Mt, at.view(1, n), bt.view(1, m), u0, v0, reg,
                                 domain=_lib.LOG, order=_lib.COL_FIRST,
                                 max_iter=numItermax, stop_thr=stopThr)
        u, v, errs, niter = res.u, res.v, res.errs, res.niter
        if verbose:
            _print_errs(errs)
        if res.status == _lib.MAX_ITER and warn:
            warnings.warn(NOT_CONVERGED)

    P, _, sums = _solve.plan(Mt, u, v, reg, domain=_lib.LOG,
                             want_plan=_grad_parts or not _loss_only)
    out = ing.scalar(sums[0, 0].item()) if _loss_only else ing.out(P[0])
    logd = None
    if log:
        log_u = ing.out(u[0])
        log_v = ing.out(v[0])
        logd = {
            "err": [ing.scalar(e) for e in errs],
            "niter": niter,
            "log_u": log_u,
            "log_v": log_v,
            "u": ing.out(torch.exp(u[0])),
            "v": ing.out(torch.exp(v[0])),
        }
    if _grad_parts:
        return ing, sums[0, 0].item(), P[0], u[0], v[0], logd
    if log:
        return out, logd
    return out


def _sinkhorn_log_columns(ing, Mt, at, bt, reg, numItermax, stopThr, log, ws):
    """Multi-histogram ``sinkhorn_log`` (_sinkhorn.py:816-854) in one launch:
    the histograms are a batch sharing M (bstride 0), and each stops on its
    own check exactly as the reference's per-column call does, so potentials
    and losses equal those of the column loop.  Each column that reaches
    numItermax warns, as the reference's per-column calls do (their ``warn``
    is left at its default)."""
    n, m = Mt.shape
    nh = bt.shape[1]
    u0 = torch.zeros((nh, n), dtype=torch.float64, device=ing.device)
    v0 = torch.zeros((nh, m), dtype=torch.float64, device=ing.device)
    for k in range(nh):
        if ws[k] is not None:
            u0[k] = ing.dev(ws[k][0], torch.float64).reshape(n)
            v0[k] = ing.dev(ws[k][1], torch.float64).reshape(m)
    a_b = at.view(1, n).expand(nh, n).contiguous()
    b_b = bt.t().contiguous()
    res = _solve.dense_solve(Mt, a_b, b_b, u0, v0, reg, domain=_lib.LOG, order=_lib.COL_FIRST,
                             max_iter=numItermax, stop_thr=stopThr, batch=nh, bstride=0,
                             each=True)
    for st in res.statuses.cpu().tolist():
        if st == _lib.MAX_ITER:
            warnings.warn(NOT_CONVERGED)
    _, _, sums = _solve.plan(Mt, res.u, res.v, reg, domain=_lib.LOG, want_plan=False,
                             batch=nh, bstride=0)
    out = ing.out(sums[:, 0])
    if not log:
        return out
    lu, lv = res.u.t(), res.v.t()
    return out, {"log_u": ing.out(lu), "log_v": ing.out(lv), "u": ing.out(torch.exp(lu)),
                 "v": ing.out(torch.exp(lv))}


def sinkhorn_knopp(a, b, M, reg, numItermax=1000, stopThr=1e-9, verbose=False,
                   log=False, warn=True, warmstart=None, _loss_only=False, _grad_parts=False,
                   **kwargs):
    """Scaling-form Sinkhorn-Knopp (_sinkhorn.py:597-687), iterations on the GPU.

    ``v = b / (K^T u)``, ``u = 1 / (K/a v)`` with ``K = exp(-M/reg)``; a zero
    in ``K^T u`` or a non-finite scaling reverts to the previous pair and
    warns ``"Warning: numerical errors at iteration %d"``.  A 2-D ``b`` solves
    all histograms together with one joint stop (Frobenius error) and returns
    the per-histogram losses.
    """
    ing, Mt, at, bt = _prep(a, b, M)
    n, m = Mt.shape
    nh = bt.shape[1] if bt.dim() > 1 else 0
    B = max(nh, 1)
    a_b = at.view(1, n).expand(B, n).contiguous()
    b_b = bt.t().contiguous() if nh else bt.view(1, m)
    if warmstart is None:
        u0 = torch.full((B, n), 1.0 / n, dtype=torch.float64, device=ing.device)
        v0 = torch.full((B, m), 1.0 / m, dtype=torch.float64, device=ing.device)
    else:
        w0 = ing.dev(warmstart[0], torch.float64)
        w1 = ing.dev(warmstart[1], torch.float64)
        u0 = torch.exp(w0.t() if w0.dim() > 1 else w0.view(1, n)).expand(B, n).contiguous()
        v0 = torch.exp(w1.t() if w1.dim() > 1 else w1.view(1, m)).expand(B, m).contiguous()

    if numItermax <= 0:
        if warn:
            warnings.warn(NOT_CONVERGED)
        if log:
            raise UnboundLocalError("numItermax must be positive when log=True")
        u, v, errs, niter = u0, v0, [], -1
    else:
        res = _solve.dense_solve(Mt, a_b, b_b, u0, v0, reg, domain=_lib.SCALING,
                                 order=_lib.COL_FIRST, max_iter=numItermax,
                                 stop_thr=stopThr, err_norm=1, batch=B, bstride=0)
        u, v, errs, niter = res.u, res.v, res.errs, res.niter
        if res.status == _lib.NUMERICAL:
            warnings.warn("Warning: numerical errors at iteration %d" % res.bad_iter)
        if verbose:
            shown = errs
            if res.status == _lib.CONVERGED:
                shown = errs[:-1]  # the reference prints after the break test
            _print_errs(shown)
        if res.status == _lib.MAX_ITER and warn:
            warnings.warn(NOT_CONVERGED)

    want_plan = _grad_parts or not (nh or _loss_only)
    P, _, sums = _solve.plan(Mt, u, v, reg, domain=_lib.SCALING, want_plan=want_plan,
                             batch=B, bstride=0)
    if nh:
        out = ing.out(sums[:, 0])
    elif _loss_only:
        out = ing.scalar(sums[0, 0].item())
    else:
        out = ing.out(P[0])
    logd = None
    if log:
        uu = u.t() if nh else u[0]
        vv = v.t() if nh else v[0]
        logd = {"err": [ing.scalar(e) for e in errs], "niter": niter,
                "u": ing.out(uu), "v": ing.out(vv)}
    if _grad_parts:
        return ing, sums[0, 0].item(), P[0], torch.log(u[0]), torch.log(v[0]), logd
    if log:
        return out, logd
    return out


# ot.bregman also exposes the stabilised solvers (ot/bregman/_sinkhorn.py:1105, :1371)
# and the lazy empirical ones (ot/bregman/__init__.py:37-43)
from .stabilized import sinkhorn_epsilon_scaling, sinkhorn_stabilized  # noqa: E402
from .empirical import empirical_sinkhorn, empirical_sinkhorn2  # noqa: E402
