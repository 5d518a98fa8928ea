"""``sinkhorn_stabilized`` and ``sinkhorn_epsilon_scaling`` (ot/bregman/_sinkhorn.py).

``sinkhorn_stabilized`` (:1105-1367) runs entirely on the device
(``potamd_sinkhorn_stabilized``): the scaling iterations, the absorption of
large scalings into ``alpha``, ``beta`` with the kernel rebuilt, the
``print_period`` error checks, the stop and the NaN guard.  The plan is
``exp(-M/reg + logu + logv)`` from the fp64 plan kernel, with ``logu =
alpha/reg + log u`` as in the reference's log.  ``sinkhorn_epsilon_scaling``
(:1371-1566) is the reference's outer loop over decreasing ``reg`` around it,
warm started by ``(alpha, beta)``; its error uses the marginals that the
stabilized call returns.

A 2-D ``b`` (several target histograms) runs the reference's
several-histogram branch on the device (``potamd_sinkhorn_stabilized_multi``):
scalings of shape (n, H) and (m, H), absorption ``alpha += reg max_h log u``
and ``beta += reg max log v`` without resetting the scalings, and the error
``(max|u - u'| / max(max|u|, max|u'|, 1) + the same for v) / 2`` (:1230-1296);
the result is the vector of per-histogram losses (:1330-1367).
``sinkhorn_epsilon_scaling`` with a 2-D ``b`` raises: the reference's loop
cannot run it either (its first check sums the loss vector along axis 1).
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _device, _lib, _solve
from .bregman import NOT_CONVERGED, _prep
from .utils import list_to_array


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _print_checks(errs, print_period):
    for k, e in enumerate(errs):
        ii = k * print_period
        if ii % (print_period * 20) == 0:
            print("{:5s}|{:12s}".format("It.", "Err") + "\n" + "-" * 19)
        print("{:5d}|{:8e}|".format(ii, e))


def _stabilized_device(ing, Mt, at, bt, reg, numItermax, tau, stopThr, warmstart, print_period,
                       warn, want_marg=False):
    lib = _lib.load()
    n, m = Mt.shape
    dev = ing.device
    if warmstart is None:
        alpha = torch.zeros(n, dtype=torch.float64, device=dev)
        beta = torch.zeros(m, dtype=torch.float64, device=dev)
    else:
        alpha = ing.dev(warmstart[0], torch.float64).reshape(n).clone()
        beta = ing.dev(warmstart[1], torch.float64).reshape(m).clone()
    u = torch.empty(n, dtype=torch.float64, device=dev)
    v = torch.empty(m, dtype=torch.float64, device=dev)
    err_cap = max(1, numItermax // print_period + 2)
    err_hist = torch.zeros(err_cap, dtype=torch.float64, device=dev)
    marg = torch.zeros(2, dtype=torch.float64, device=dev) if want_marg else None
    code = _lib.F32 if Mt.dtype == torch.float32 else _lib.F64
    ws_bytes = lib.potamd_stabilized_workspace_size(code, n, m)
    ws = _device.workspace(ws_bytes, dev)
    info = (ctypes.c_int * 4)()
    hc = _device.host_ctrl()
    rc = lib.potamd_sinkhorn_stabilized(
        code, _ptr(Mt), Mt.stride(0), n, m, float(reg), _ptr(at), _ptr(bt), _ptr(alpha),
        _ptr(beta), _ptr(u), _ptr(v), float(tau), int(numItermax), float(stopThr),
        int(print_period), _ptr(err_hist), err_cap, _ptr(marg), _ptr(ws), ws_bytes,
        ctypes.c_void_p(hc.data_ptr()), ctypes.c_void_p(_device.stream_ptr(dev)), info)
    _lib.check(rc, "potamd_sinkhorn_stabilized")
    if rc == _lib.NUMERICAL:
        warnings.warn("Numerical errors at iteration %d" % info[3])
    elif rc == _lib.MAX_ITER and warn:
        warnings.warn(NOT_CONVERGED)
    n_err = min(info[1], err_cap)
    errs = err_hist[:n_err].cpu().tolist() if n_err else []
    return alpha, beta, u, v, errs, info[0], marg


def _stabilized_multi(ing, Mt, at, bt, reg, numItermax, tau, stopThr, warmstart, verbose,
                      print_period, log, warn):
    """Several target histograms (b: (m, H)); see the module docstring."""
    lib = _lib.load()
    n, m = Mt.shape
    H = bt.shape[1]
    dev = ing.device
    if warmstart is None:
        alpha = torch.zeros(n, dtype=torch.float64, device=dev)
        beta = torch.zeros(m, dtype=torch.float64, device=dev)
    else:
        alpha = ing.dev(warmstart[0], torch.float64).reshape(n).clone()
        beta = ing.dev(warmstart[1], torch.float64).reshape(m).clone()
    b_h = bt.t().contiguous()  # histogram-major (H, m)
    u = torch.empty((H, n), dtype=torch.float64, device=dev)
    v = torch.empty((H, m), dtype=torch.float64, device=dev)
    err_cap = max(1, numItermax // print_period + 2)
    err_hist = torch.zeros(err_cap, dtype=torch.float64, device=dev)
    code = _lib.F32 if Mt.dtype == torch.float32 else _lib.F64
    ws_bytes = lib.potamd_stabilized_multi_workspace_size(code, n, m, H)
    ws = _device.workspace(ws_bytes, dev)
    info = (ctypes.c_int * 4)()
    hc = _device.host_ctrl()
    rc = lib.potamd_sinkhorn_stabilized_multi(
        code, _ptr(Mt), Mt.stride(0), n, m, H, float(reg), _ptr(at), _ptr(b_h), _ptr(alpha),
        _ptr(beta), _ptr(u), _ptr(v), float(tau), int(numItermax), float(stopThr),
        int(print_period), _ptr(err_hist), err_cap, _ptr(ws), ws_bytes,
        ctypes.c_void_p(hc.data_ptr()), ctypes.c_void_p(_device.stream_ptr(dev)), info)
    _lib.check(rc, "potamd_sinkhorn_stabilized_multi")
    if rc == _lib.NUMERICAL:
        warnings.warn("Numerical errors at iteration %d" % info[3])
    elif rc == _lib.MAX_ITER and warn:
        warnings.warn(NOT_CONVERGED)
    n_err = min(info[1], err_cap)
    errs = err_hist[:n_err].cpu().tolist() if n_err else []
    if verbose:
        _print_checks(errs, print_period)
    logu = alpha[None, :] / reg + torch.log(u)  # (H, n)
    logv = beta[None, :] / reg + torch.log(v)
    _, _, sums = _solve.plan(Mt, logu, logv, reg, domain=_lib.LOG, want_plan=False,
                             batch=H, bstride=0)
    out = ing.out(sums[:, 0])
    if not log:
        return out
    al = alpha[:, None] + reg * torch.log(u.t())
    be = beta[:, None] + reg * torch.log(v.t())
    logd = {"err": [ing.scalar(e) for e in errs], "n_iter": info[0], "logu": ing.out(logu.t()),
            "logv": ing.out(logv.t()), "alpha": ing.out(al), "beta": ing.out(be)}
    logd["warmstart"] = (logd["alpha"], logd["beta"])
    return out, logd


def sinkhorn_stabilized(a, b, M, reg, numItermax=1000, tau=1e3, stopThr=1e-9, warmstart=None,
                        verbose=False, print_period=20, log=False, warn=True, _loss_only=False,
                        **kwargs):
    """Log-stabilised Sinkhorn; see the module docstring."""
    a, b, M = list_to_array(a, b, M)
    ing, Mt, at, bt = _prep(a, b, M)
    if numItermax < 1:
        raise ValueError("numItermax must be positive")
    if bt.dim() > 1:
        return _stabilized_multi(ing, Mt, at, bt, reg, numItermax, tau, stopThr, warmstart,
                                 verbose, print_period, log, warn)
    alpha, beta, u, v, errs, niter, _ = _stabilized_device(
        ing, Mt, at, bt, reg, numItermax, tau, stopThr, warmstart, print_period, warn)
    if verbose:
        _print_checks(errs, print_period)
    logu = alpha / reg + torch.log(u)
    logv = beta / reg + torch.log(v)
    P, _, sums = _solve.plan(Mt, logu.view(1, -1), logv.view(1, -1), reg, domain=_lib.LOG,
                             want_plan=not _loss_only)
    out = ing.scalar(float(sums[0, 0])) if _loss_only else ing.out(P[0])
    if log:
        al = alpha + reg * torch.log(u)
        be = beta + reg * torch.log(v)
        logd = {"err": [ing.scalar(e) for e in errs], "n_iter": niter, "logu": ing.out(logu),
                "logv": ing.out(logv), "alpha": ing.out(al), "beta": ing.out(be)}
        logd["warmstart"] = (logd["alpha"], logd["beta"])
        return out, logd
    return out


def sinkhorn_epsilon_scaling(a, b, M, reg, numItermax=100, epsilon0=1e4, numInnerItermax=100,
                             tau=1e3, stopThr=1e-9, warmstart=None, verbose=False,
                             print_period=10, log=False, warn=True, **kwargs):
    """ε-scaling around ``sinkhorn_stabilized`` (_sinkhorn.py:1371-1566)."""
    a, b, M = list_to_array(a, b, M)
    ing, Mt, at, bt = _prep(a, b, M)
    if bt.dim() > 1:
        raise NotImplementedError("sinkhorn_epsilon_scaling with several histograms: the "
                                  "reference's loop sums the loss vector along axis 1 at its "
                                  "first check and fails")
    numItermin = 35
    numItermax = max(numItermin, numItermax)
    n, m = Mt.shape
    if warmstart is None:
        alpha = torch.zeros(n, dtype=torch.float64, device=ing.device)
        beta = torch.zeros(m, dtype=torch.float64, device=ing.device)
    else:
        alpha = ing.dev(warmstart[0], torch.float64).reshape(n)
        beta = ing.dev(warmstart[1], torch.float64).reshape(m)
    err = 1.0
    errs = []
    ii = -1
    state = None
    for ii in range(numItermax):
        regi = (epsilon0 - reg) * np.exp(-ii) + reg
        al, be, u, v, _, _, marg = _stabilized_device(
            ing, Mt, at, bt, regi, numInnerItermax, tau, stopThr, (alpha, beta), 20, True,
            want_marg=True)
        state = (al, be, u, v, regi)
        alpha = al + regi * torch.log(u)
        beta = be + regi * torch.log(v)
        if ii % print_period == 0:
            sq = marg.cpu().numpy()
            err = float(sq[0] + sq[1])
            errs.append(err)
            if verbose:
                if ii % (print_period * 10) == 0:
                    print("{:5s}|{:12s}".format("It.", "Err") + "\n" + "-" * 19)
                print("{:5d}|{:8e}|".format(ii, err))
        if err <= stopThr and ii > numItermin:
            break
    else:
        if warn:
            warnings.warn(NOT_CONVERGED)
    al, be, u, v, regi = state
    logu = al / regi + torch.log(u)
    logv = be / regi + torch.log(v)
    P, _, _ = _solve.plan(Mt, logu.view(1, -1), logv.view(1, -1), regi, domain=_lib.LOG,
                          want_sums=False)
    G = ing.out(P[0])
    if log:
        logd = {"err": [ing.scalar(e) for e in errs], "alpha": ing.out(alpha),
                "beta": ing.out(beta), "niter": ii}
        logd["warmstart"] = (logd["alpha"], logd["beta"])
        return G, logd
    return G
