"""Sinkhorn on the online squared Euclidean cost (the cost matrix is never stored).

``sinkhorn_online`` runs the iterations of ``sinkhorn_log``
(ot/bregman/_sinkhorn.py:880-906: column update first, marginal error
``||P^T 1 - b||_2`` every 10 iterations, strict ``err < stopThr``, the
for-else warning) on ``M = dist(X, Y) / scale`` with ``dist`` the squared
Euclidean distance of ot/utils.py:270-307 and ``scale`` its maximum by
default (the ``M / M.max()`` normalisation of the benchmark problems).  The
``x . y^T`` term is computed on the GPU in the requested precision:

* ``"f32"``    fp32 MFMA
* ``"bf16"``   bf16 MFMA, fp32 accumulation
* ``"bf16x3"`` bf16 hi + lo split, three MFMA products (compensated)
* ``"fp8"``    fp8 e4m3 MFMA with a per-point scale
* ``"f64"``    fp64 vector ALU (the full-size reference)

Potentials, sums and the stop test are fp64 in every variant.  The plan,
the cost ``<P, M>`` and both marginal errors come from ``online_plan``, which
evaluates the cost in fp64 from the fp32 points.
"""

import ctypes
import warnings

import numpy as np
import torch

from . import _device, _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _points(X, dev, dtype=torch.float32):
    X = torch.as_tensor(X)
    if X.device != dev:
        X = X.to(dev)
    X = X.to(dtype)
    if X.dim() != 2:
        raise ValueError("point clouds must be 2-D (n, d)")
    if X.stride(-1) != 1:
        X = X.contiguous()
    return X


def _weights(w, k, dev):
    if w is None or (hasattr(w, "__len__") and len(w) == 0):
        return torch.full((k,), 1.0 / k, dtype=torch.float64, device=dev)
    w = torch.as_tensor(w).to(device=dev, dtype=torch.float64).reshape(-1).contiguous()
    if w.numel() != k:
        raise ValueError("weights and points differ in length")
    return w


def sinkhorn_online(X, Y, reg, a=None, b=None, numItermax=1000, stopThr=1e-9, *,
                    precision="f32", cost_scale=None, warmstart=None, log=False, warn=True,
                    splits=0, device=None, _semantics=0):
    """Online-cost ``sinkhorn_log``; returns ``(log_u, log_v, log_dict)``.

    ``X`` (n, d) and ``Y`` (m, d) are point clouds (NumPy or torch; computed
    in fp32 on the GPU).  ``cost_scale`` divides the cost (``None``: its
    maximum, computed on the GPU and returned as ``log_dict["cost_scale"]``).
    ``warmstart`` is ``(u0, v0)`` in log units, as sinkhorn_log's; only
    ``u0`` affects the result (the first update is the column update).
    """
    if precision not in _lib.ONLINE_PRECISIONS:
        raise ValueError(f"unknown precision {precision!r}; one of "
                         f"{sorted(_lib.ONLINE_PRECISIONS)}")
    dev = _device.require_gpu() if device is None else torch.device(device)
    lib = _lib.load()
    # fp64 points keep their precision on the fp64 pass; the MFMA passes pack fp32
    pdt = torch.float64 if precision == "f64" else torch.float32
    Xt = _points(X, dev, pdt)
    Yt = _points(Y, dev, pdt)
    n, d = Xt.shape
    m = Yt.shape[0]
    if Yt.shape[1] != d:
        raise ValueError("point clouds differ in dimension")
    at = _weights(a, n, dev)
    bt = _weights(b, m, dev)
    prec = _lib.ONLINE_PRECISIONS[precision]
    ws_bytes = lib.potamd_online_workspace_size(prec, n, m, d)
    if ws_bytes == 0:
        raise _lib.LibraryError(f"precision {precision!r} does not support d = {d}")
    ws = _device.workspace(ws_bytes, dev)
    if warmstart is None:
        u = torch.zeros(n, dtype=torch.float64, device=dev)
    else:
        u = torch.as_tensor(warmstart[0]).to(device=dev, dtype=torch.float64).reshape(-1).clone()
    v = torch.zeros(m, dtype=torch.float64, device=dev)
    err_cap = max(1, numItermax // 10 + 2)
    err_hist = torch.zeros(err_cap, dtype=torch.float64, device=dev)
    info = (ctypes.c_int * 4)()
    scale_out = ctypes.c_double(0.0)
    hc = _device.host_ctrl()
    rc = lib.potamd_sinkhorn_online(
        prec, _lib.F64 if pdt == torch.float64 else _lib.F32, _ptr(Xt), Xt.stride(0), _ptr(Yt),
        Yt.stride(0), n, m, d, float(reg),
        0.0 if cost_scale is None else float(cost_scale), _ptr(at), _ptr(bt), _ptr(u), _ptr(v),
        int(numItermax), float(stopThr), int(_semantics), _ptr(err_hist), err_cap, int(splits),
        _ptr(ws),
        ws_bytes, ctypes.c_void_p(hc.data_ptr()), ctypes.c_void_p(_device.stream_ptr(dev)), info,
        ctypes.byref(scale_out))
    _lib.check(rc, "potamd_sinkhorn_online")
    n_err = min(info[1], err_cap)
    if rc == _lib.MAX_ITER and warn:
        warnings.warn("Sinkhorn did not converge. You might want to "
                      "increase the number of iterations `numItermax` "
                      "or the regularization parameter `reg`.")
    out = {"niter": info[0], "err": err_hist[:n_err].cpu().tolist() if n_err else [],
           "status": rc, "cost_scale": scale_out.value, "precision": precision,
           "a": at, "b": bt}
    return u, v, out


def online_plan(X, Y, reg, log_u, log_v, cost_scale, a=None, b=None, *, want_plan=False,
                device=None, fp64_points=False):
    """fp64 plan / cost / marginal errors of the online cost.

    ``fp64_points`` evaluates the cost from float64 coordinates (otherwise the
    points are taken as float32).
    Returns ``(P or None, {"cost", "mass", "err_a", "err_b"})`` with
    ``P_ij = exp(-M_ij / reg + log_u_i + log_v_j)`` (ot/bregman/_sinkhorn.py:915-918)
    and ``cost = <P, M>`` (the loss of sinkhorn2, :436).
    """
    dev = _device.require_gpu() if device is None else torch.device(device)
    lib = _lib.load()
    pdt = torch.float64 if fp64_points else torch.float32
    Xt = _points(X, dev, pdt)
    Yt = _points(Y, dev, pdt)
    n, d = Xt.shape
    m = Yt.shape[0]
    at = _weights(a, n, dev)
    bt = _weights(b, m, dev)
    u = torch.as_tensor(log_u).to(device=dev, dtype=torch.float64).contiguous()
    v = torch.as_tensor(log_v).to(device=dev, dtype=torch.float64).contiguous()
    P = torch.empty((n, m), dtype=torch.float64, device=dev) if want_plan else None
    sums = torch.zeros(4, dtype=torch.float64, device=dev)
    ws_bytes = lib.potamd_online_plan_workspace_size(n, m)
    ws = _device.workspace(ws_bytes, dev)
    rc = lib.potamd_online_plan(
        _lib.F64 if fp64_points else _lib.F32, _ptr(Xt), Xt.stride(0), _ptr(Yt), Yt.stride(0),
        n, m, d, float(reg), float(cost_scale),
        _ptr(u), _ptr(v), _ptr(at), _ptr(bt), _ptr(P), m, _ptr(sums), _ptr(ws), ws_bytes,
        ctypes.c_void_p(_device.stream_ptr(dev)))
    _lib.check(rc, "potamd_online_plan")
    s = sums.cpu().numpy()
    return P, {"cost": float(s[0]), "mass": float(s[1]), "err_a": float(s[2]),
               "err_b": float(s[3])}
