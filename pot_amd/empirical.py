"""``empirical_sinkhorn`` / ``empirical_sinkhorn2`` (ot/bregman/_empirical.py).

``isLazy=False`` (the default) forms ``M = dist(X_s, X_t, metric)`` and calls
``sinkhorn`` / ``sinkhorn2`` (:272-302, :484-517).  ``isLazy=True`` runs the
reference's blocked loop (:193-270) on the online-cost GPU solver: the squared
Euclidean cost is never stored, f is updated before g, the 1-norm error of
the source marginal is checked when ``(i_ot + 1) % 10 == 0`` with a non-strict
``err <= stopThr``, ``log`` holds ``err``, ``u`` (= f), ``v`` (= g) and
``niter``.  float64 points run the fp64 pass on float64 coordinates,
float32 points the fp32 MFMA pass.  The lazy path supports
``metric="sqeuclidean"``; ``batchSize`` is validated as in the reference but
the GPU pass tiles the cost itself.  With ``log`` the lazy path also returns
``log["lazy_plan"]``, the plan as a ``LazyTensor`` whose blocks
``exp(f_i + g_j - C_ij / reg)`` are computed on the GPU from the points when
indexed (``get_sinkhorn_lazytensor``, _empirical.py:26-71).
"""

import numpy as np
import torch

from . import online
from ._device import Ingress
from .bregman import sinkhorn, sinkhorn2
from .utils import LazyTensor, dist, list_to_array


def _index(k, size, dev):
    """A LazyTensor index as a device index tensor (ints count as a slice of one)."""
    if isinstance(k, (int, np.integer)):
        k = int(k)
        if not -size <= k < size:
            raise IndexError(f"index {k} is out of bounds for size {size}")
        k = slice(k % size, k % size + 1)
    if isinstance(k, slice):
        return torch.arange(size, device=dev)[k]
    idx = torch.as_tensor(np.asarray(k) if not isinstance(k, torch.Tensor) else k)
    if idx.dtype == torch.bool:
        idx = torch.nonzero(idx).reshape(-1)
    idx = idx.to(device=dev, dtype=torch.int64).reshape(-1)
    if idx.numel() and (int(idx.max()) >= size or int(idx.min()) < -size):
        raise IndexError(f"index out of bounds for size {size}")
    return idx % size if idx.numel() else idx


def get_sinkhorn_lazytensor(X_a, X_b, f, g, metric="sqeuclidean", reg=1e-1, nx=None):
    """The Sinkhorn plan ``T = exp(f 1^T + 1 g^T - C / reg)`` of the dual
    potentials as a LazyTensor (ot/bregman/_empirical.py:26-71): indexing it
    evaluates the requested block on the GPU (the squared Euclidean cost
    recomputed from the points in fp64 arithmetic, ``potamd_online_plan``)
    and returns it in the kind and dtype of the inputs; integers select one
    row or column as a slice of one, as in the reference."""
    if metric != "sqeuclidean":
        raise NotImplementedError("the GPU lazy plan computes the 'sqeuclidean' cost")
    shape = (X_a.shape[0], X_b.shape[0])
    ing = Ingress(X_a, X_b, f, g)
    f64 = ing.dtype == torch.float64
    pdt = torch.float64 if f64 else torch.float32
    dev = {}

    def resident():
        # device copies made on first use and kept with the tensor
        if not dev:
            dev["Xa"] = ing.dev(X_a, pdt)
            dev["Xb"] = ing.dev(X_b, pdt)
            dev["f"] = ing.dev(f, torch.float64).reshape(-1)
            dev["g"] = ing.dev(g, torch.float64).reshape(-1)
        return dev

    def func(i, j, X_a, X_b, f, g, metric, reg):
        d = resident()
        ii = _index(i, shape[0], ing.device)
        jj = _index(j, shape[1], ing.device)
        if ii.numel() == 0 or jj.numel() == 0:
            return ing.out(torch.zeros((ii.numel(), jj.numel()), dtype=torch.float64,
                                       device=ing.device))
        P, _ = online.online_plan(d["Xa"].index_select(0, ii).contiguous(),
                                  d["Xb"].index_select(0, jj).contiguous(), reg,
                                  d["f"].index_select(0, ii), d["g"].index_select(0, jj), 1.0,
                                  want_plan=True, device=ing.device, fp64_points=f64)
        return ing.out(P)

    return LazyTensor(shape, func, X_a=X_a, X_b=X_b, f=f, g=g, metric=metric, reg=reg)


def _lazy(X_s, X_t, reg, a, b, metric, numIterMax, stopThr, batchSize, verbose, warn,
          warmstart):
    if metric != "sqeuclidean":
        raise NotImplementedError("the lazy GPU path computes the 'sqeuclidean' cost")
    if not (isinstance(batchSize, int) or (isinstance(batchSize, tuple) and len(batchSize) == 2)):
        raise ValueError("Batch size must be in integer or a tuple of two integers")
    ing = Ingress(X_s, X_t, a, b)
    f64 = ing.dtype == torch.float64
    pdt = torch.float64 if f64 else torch.float32
    at = None if a is None else ing.dev(a, torch.float64)
    bt = None if b is None else ing.dev(b, torch.float64)
    ws = None if warmstart is None else (ing.dev(warmstart[1], torch.float64), None)
    # Seen from the target set: its potential g is the driver's u, the
    # source potential f its v, and the checked marginal is the source's.
    g, f, lg = online.sinkhorn_online(
        ing.dev(X_t, pdt), ing.dev(X_s, pdt), reg, a=bt, b=at, numItermax=numIterMax,
        stopThr=stopThr, precision="f64" if f64 else "f32", cost_scale=1.0, warmstart=ws,
        warn=warn, _semantics=1)
    if verbose:
        for k, e in enumerate(lg["err"]):
            if (10 * k + 10) % 100 == 0:
                print("Error in marginal at iteration {} = {}".format(10 * k + 10, e))
    return ing, f, g, lg


def empirical_sinkhorn(X_s, X_t, reg, a=None, b=None, metric="sqeuclidean", numIterMax=10000,
                       stopThr=1e-9, isLazy=False, batchSize=100, verbose=False, log=False,
                       warn=True, warmstart=None, **kwargs):
    """Entropic OT between samples; see the module docstring."""
    X_s, X_t = list_to_array(X_s, X_t)
    if isLazy:
        ing, f, g, lg = _lazy(X_s, X_t, reg, a, b, metric, numIterMax, stopThr, batchSize,
                              verbose, warn, warmstart)
        f_o, g_o = ing.out(f), ing.out(g)
        if log:
            return f_o, g_o, {"err": [ing.scalar(e) for e in lg["err"]], "u": f_o, "v": g_o,
                              "niter": lg["niter"],
                              "lazy_plan": get_sinkhorn_lazytensor(X_s, X_t, f_o, g_o, metric,
                                                                   reg)}
        return f_o, g_o
    M = dist(X_s, X_t, metric=metric)
    if a is None:
        a = np.full(len(X_s), 1.0 / len(X_s)) if not isinstance(X_s, torch.Tensor) else \
            torch.full((len(X_s),), 1.0 / len(X_s), dtype=X_s.dtype, device=X_s.device)
    if b is None:
        b = np.full(len(X_t), 1.0 / len(X_t)) if not isinstance(X_t, torch.Tensor) else \
            torch.full((len(X_t),), 1.0 / len(X_t), dtype=X_t.dtype, device=X_t.device)
    return sinkhorn(a, b, M, reg, numItermax=numIterMax, stopThr=stopThr, verbose=verbose,
                    log=log, warmstart=warmstart, **kwargs)


def empirical_sinkhorn2(X_s, X_t, reg, a=None, b=None, metric="sqeuclidean",
                        numIterMax=10000, stopThr=1e-9, isLazy=False, batchSize=100,
                        verbose=False, log=False, warn=True, warmstart=None, **kwargs):
    """Entropic OT loss between samples (``<P, M>``); see the module docstring."""
    X_s, X_t = list_to_array(X_s, X_t)
    if isLazy:
        ing, f, g, lg = _lazy(X_s, X_t, reg, a, b, metric, numIterMax, stopThr, batchSize,
                              verbose, warn, warmstart)
        f64 = ing.dtype == torch.float64
        pdt = torch.float64 if f64 else torch.float32
        _, sums = online.online_plan(ing.dev(X_s, pdt), ing.dev(X_t, pdt), reg, f, g, 1.0,
                                     a=lg["b"], b=lg["a"], fp64_points=f64)
        loss = ing.scalar(sums["cost"])
        if log:
            f_o, g_o = ing.out(f), ing.out(g)
            return loss, {"err": [ing.scalar(e) for e in lg["err"]], "u": f_o, "v": g_o,
                          "niter": lg["niter"],
                          "lazy_plan": get_sinkhorn_lazytensor(X_s, X_t, f_o, g_o, metric, reg)}
        return loss
    M = dist(X_s, X_t, metric=metric)
    if a is None:
        a = np.full(len(X_s), 1.0 / len(X_s)) if not isinstance(X_s, torch.Tensor) else \
            torch.full((len(X_s),), 1.0 / len(X_s), dtype=X_s.dtype, device=X_s.device)
    if b is None:
        b = np.full(len(X_t), 1.0 / len(X_t)) if not isinstance(X_t, torch.Tensor) else \
            torch.full((len(X_t),), 1.0 / len(X_t), dtype=X_t.dtype, device=X_t.device)
    return sinkhorn2(a, b, M, reg, numItermax=numIterMax, stopThr=stopThr, verbose=verbose,
                     log=log, warn=warn, warmstart=warmstart, **kwargs)
