"""Float64 NumPy restatement of POT's balanced entropic Sinkhorn solvers.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, step by step:

* ``sinkhorn``        ot/bregman/_sinkhorn.py:21-234   (method dispatch; greenkhorn
  is ``oracle.greenkhorn``)
* ``sinkhorn2``       ot/bregman/_sinkhorn.py:237-484  (loss variant)
* ``sinkhorn_knopp``  ot/bregman/_sinkhorn.py:487-687  (scaling form)
* ``sinkhorn_log``    ot/bregman/_sinkhorn.py:690-918  (log-domain form)

The log-sum-exp is ``scipy.special.logsumexp``, the very routine the reference
NumPy backend calls (ot/backend.py:1306-1307).

Pinning (tests/test_oracle.py):

* closed form of the symmetric 2x2 problem (the doctests at
  _sinkhorn.py:123-129, :342-347, :563-569, :766-772) and of the general 2x2
  problem (solved independently with a scalar root finder);
* separable costs ``M_ij = alpha_i + beta_j`` give ``P = a b^T``;
* dual fixed point, gauge invariance, marginals, knopp == log;
* brute force: L-BFGS on the smooth entropic dual (scipy.optimize) at small n;
* iteration semantics (``niter``, ``err`` cadence, strict ``<``, warnings)
  against hand-computed sequences at numItermax in {1, 2, 11}.
"""

import warnings

import numpy as np
from scipy.special import logsumexp

_NOT_CONVERGED = (
    "Sinkhorn did not converge. You might want to "
    "increase the number of iterations `numItermax` "
    "or the regularization parameter `reg`."
)


def _as_arrays(*arrs):
    """Lists become float64 arrays (ot/utils.py:59-94, numpy case)."""
    return [np.asarray(x, dtype=np.float64) if isinstance(x, list) else x for x in arrs]


def _default_weights(a, b, M):
    """Empty weights default to uniform (_sinkhorn.py:801-804, :603-606)."""
    if len(a) == 0:
        a = np.full((M.shape[0],), 1.0 / M.shape[0], dtype=M.dtype)
    if len(b) == 0:
        b = np.full((M.shape[1],), 1.0 / M.shape[1], dtype=M.dtype)
    return a, b


def sinkhorn(a, b, M, reg, method="sinkhorn", numItermax=1000, stopThr=1e-9,
             verbose=False, log=False, warn=True, warmstart=None, **kwargs):
    """Dispatch on ``method.lower()`` (_sinkhorn.py:164-234)."""
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
        from .greenkhorn import greenkhorn
        return greenkhorn(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                          verbose=verbose, log=log, warn=warn, warmstart=warmstart)
    if name in ("sinkhorn_stabilized", "sinkhorn_epsilon_scaling"):
        raise NotImplementedError("oracle.sinkhorn covers sinkhorn / sinkhorn_log / greenkhorn; "
                                  "see oracle.stabilized")
    raise ValueError("Unknown method '%s'." % method)


def sinkhorn2(a, b, M, reg, method="sinkhorn", numItermax=1000, stopThr=1e-9,
              verbose=False, log=False, warn=False, warmstart=None, **kwargs):
    """Loss variant (_sinkhorn.py:387-484).

    1-D ``b``: ``sum(M * P)``; 2-D ``b``: the solver's own per-histogram loss.
    """
    M, a, b = _as_arrays(M, a, b)
    name = method.lower()
    if name not in ("sinkhorn", "sinkhorn_log"):
        if name == "sinkhorn_stabilized":
            raise NotImplementedError("oracle covers sinkhorn / sinkhorn_log only")
        raise ValueError("Unknown method '%s'." % method)
    solver = sinkhorn_knopp if name == "sinkhorn" else sinkhorn_log
    res = solver(a, b, M, reg, numItermax=numItermax, stopThr=stopThr,
                 verbose=verbose, log=log, warn=warn, warmstart=warmstart, **kwargs)
    if b.ndim >= 2:
        return res
    if log:
        return np.sum(M * res[0]), res[1]
    return np.sum(M * res)


def _print_err_row(ii, err):
    """Verbose table of _sinkhorn.py:655-658 / :894-897."""
    if ii % 200 == 0:
        print("{:5s}|{:12s}".format("It.", "Err") + "\n" + "-" * 19)
    print("{:5d}|{:8e}|".format(ii, err))


def sinkhorn_knopp(a, b, M, reg, numItermax=1000, stopThr=1e-9, verbose=False,
                   log=False, warn=True, warmstart=None, **kwargs):
    """Scaling-form Sinkhorn (_sinkhorn.py:597-687).

    Per iteration ii: ``v = b / (K^T u)``, ``u = 1 / (Kp v)`` with
    ``K = exp(-M/reg)`` and ``Kp = K / a[:, None]``.  A zero in ``K^T u`` or a
    non-finite scaling reverts to the previous pair and stops with a warning
    (:633-645).  Every 10th iteration the column-marginal error is checked
    before the strict ``err < stopThr`` test (:646-659).
    """
    a, b, M = _as_arrays(a, b, M)
    a, b = _default_weights(a, b, M)
    na, nb = len(a), b.shape[0]
    nh = b.shape[1] if b.ndim > 1 else 0
    logd = {"err": []} if log else None

    if warmstart is None:
        shape_u = (na, nh) if nh else (na,)
        shape_v = (nb, nh) if nh else (nb,)
        u = np.ones(shape_u, dtype=M.dtype) / na
        v = np.ones(shape_v, dtype=M.dtype) / nb
    else:
        u, v = np.exp(warmstart[0]), np.exp(warmstart[1])

    K = np.exp(M / (-reg))
    Kp = (1.0 / a).reshape(-1, 1) * K

    ii = -1
    completed = True
    for ii in range(numItermax):
        u_prev, v_prev = u, v
        ktu = K.T @ u
        v = b / ktu
        u = 1.0 / (Kp @ v)
        bad = (np.any(ktu == 0) or np.any(np.isnan(u)) or np.any(np.isnan(v))
               or np.any(np.isinf(u)) or np.any(np.isinf(v)))
        if bad:
            warnings.warn("Warning: numerical errors at iteration %d" % ii)
            u, v = u_prev, v_prev
            completed = False
            break
        if ii % 10 == 0:
            if nh:
                marg = np.einsum("ik,ij,jk->jk", u, K, v)
            else:
                marg = np.einsum("i,ij,j->j", u, K, v)
            err = np.linalg.norm(marg - b)
            if log:
                logd["err"].append(err)
            if err < stopThr:
                completed = False
                break
            if verbose:
                _print_err_row(ii, err)
    if completed and warn:
        warnings.warn(_NOT_CONVERGED)

    if log:
        if ii < 0:
            raise UnboundLocalError("numItermax must be positive when log=True")
        logd["niter"] = ii
        logd["u"] = u
        logd["v"] = v

    if nh:
        res = np.einsum("ik,ij,jk,ij->k", u, K, v, M)
        return (res, logd) if log else res
    P = u.reshape((-1, 1)) * K * v.reshape((1, -1))
    return (P, logd) if log else P


def sinkhorn_log(a, b, M, reg, numItermax=1000, stopThr=1e-9, verbose=False,
                 log=False, warn=True, warmstart=None, **kwargs):
    """Log-domain Sinkhorn (_sinkhorn.py:799-918).

    Per iteration ii (column update first, :881-882)::

        v = log b - LSE_rows(Mr + u)        (Mr = -M / reg)
        u = log a - LSE_cols(Mr + v)

    Every 10th iteration: ``err = || sum_i exp(Mr + u + v) - b ||_2``, logged,
    then ``break`` on the strict ``err < stopThr`` (:884-899).  Exhausting the
    loop warns (for-else, :900-906).  ``log`` holds ``err``, ``niter`` (the last
    ``ii``), ``log_u``, ``log_v``, ``u = exp(log_u)``, ``v = exp(log_v)``.

    A 2-D ``b`` loops over its columns and returns the stacked losses
    ``sum(M * P_k)``; the per-column call does not forward ``warn``
    (:819-854).
    """
    a, b, M = _as_arrays(a, b, M)
    a, b = _default_weights(a, b, M)
    na, nb = len(a), b.shape[0]
    nh = b.shape[1] if b.ndim > 1 else 0

    if nh > 1 and warmstart is None:
        warmstart = [None] * nh
    if nh:
        losses, lus, lvs = [], [], []
        for k in range(nh):
            res = sinkhorn_log(a, b[:, k], M, reg, numItermax=numItermax,
                               stopThr=stopThr, verbose=verbose, log=log,
                               warmstart=warmstart[k], **kwargs)
            if log:
                losses.append(np.sum(M * res[0]))
                lus.append(res[1]["log_u"])
                lvs.append(res[1]["log_v"])
            else:
                losses.append(np.sum(M * res))
        out = np.stack(losses)
        if log:
            logd = {"log_u": np.stack(lus, 1), "log_v": np.stack(lvs, 1)}
            logd["u"] = np.exp(logd["log_u"])
            logd["v"] = np.exp(logd["log_v"])
            return out, logd
        return out

    logd = {"err": []} if log else None
    Mr = -M / reg
    if warmstart is None:
        u = np.zeros(na, dtype=M.dtype)
        v = np.zeros(nb, dtype=M.dtype)
    else:
        u, v = warmstart
    loga = np.log(a)
    logb = np.log(b)

    ii = -1
    completed = True
    for ii in range(numItermax):
        v = logb - logsumexp(Mr + u[:, None], 0)
        u = loga - logsumexp(Mr + v[None, :], 1)
        if ii % 10 == 0:
            marg = np.sum(np.exp(Mr + u[:, None] + v[None, :]), 0)
            err = np.linalg.norm(marg - b)
            if log:
                logd["err"].append(err)
            if verbose:
                _print_err_row(ii, err)
            if err < stopThr:
                completed = False
                break
    if completed and warn:
        warnings.warn(_NOT_CONVERGED)

    P = np.exp(Mr + u[:, None] + v[None, :])
    if log:
        if ii < 0:
            raise UnboundLocalError("numItermax must be positive when log=True")
        logd["niter"] = ii
        logd["log_u"] = u
        logd["log_v"] = v
        logd["u"] = np.exp(u)
        logd["v"] = np.exp(v)
        return P, logd
    return P
