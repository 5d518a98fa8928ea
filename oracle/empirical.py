"""Float64 NumPy restatement of the lazy ``empirical_sinkhorn`` loop.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ot/bregman/_empirical.py:193-270 (``isLazy=True``) and :430-482
(``empirical_sinkhorn2``'s lazy loss):

* f, g log-scalings, zero or the warm start; the cost is ``dist(X_s, X_t)``
  (squared Euclidean, clamped at 0; not normalised);
* each iteration ``f = log a - LSE_j(g_j - M_ij / reg)`` then
  ``g = log b - LSE_i(f_i - M_ij / reg)``;
* when ``(i_ot + 1) % 10 == 0``: ``err = sum_i |sum_j exp(f_i + g_j -
  M_ij / reg) - a_i|``, appended to the log, and ``err <= stopThr`` stops;
* the for-else warning; ``niter = i_ot``.

The blocking over ``batchSize`` only changes how the sums are formed, so it
is not restated.  Pinned (tests/test_oracle.py) by the doctest plan of
_empirical.py:152-157 and by the fixed point of the two updates.
"""

import warnings

import numpy as np
from scipy.special import logsumexp


def _sqeuclid(x, y):
    M = (x ** 2).sum(1)[:, None] + (y ** 2).sum(1)[None, :] - 2.0 * x @ y.T
    return np.maximum(M, 0.0)


def empirical_sinkhorn_lazy(X_s, X_t, reg, a=None, b=None, numIterMax=10000, stopThr=1e-9,
                            warn=True, warmstart=None):
    """Returns (f, g, log) with log = {"err", "u", "v", "niter"}."""
    X_s = np.asarray(X_s, dtype=np.float64)
    X_t = np.asarray(X_t, dtype=np.float64)
    ns, nt = len(X_s), len(X_t)
    a = np.full(ns, 1.0 / ns) if a is None else np.asarray(a, dtype=np.float64)
    b = np.full(nt, 1.0 / nt) if b is None else np.asarray(b, dtype=np.float64)
    M = _sqeuclid(X_s, X_t)
    log_a, log_b = np.log(a), np.log(b)
    if warmstart is None:
        f, g = np.zeros(ns), np.zeros(nt)
    else:
        f, g = (np.asarray(w, dtype=np.float64) for w in warmstart)
    errs = []
    i_ot = -1
    for i_ot in range(numIterMax):
        f = log_a - logsumexp(g[None, :] - M / reg, axis=1)
        g = log_b - logsumexp(f[:, None] - M / reg, axis=0)
        if (i_ot + 1) % 10 == 0:
            m1 = np.exp(f[:, None] + g[None, :] - M / reg).sum(1)
            err = float(np.sum(np.abs(m1 - a)))
            errs.append(err)
            if err <= stopThr:
                break
    else:
        if warn:
            warnings.warn("Sinkhorn did not converge. You might want to "
                          "increase the number of iterations `numItermax` "
                          "or the regularization parameter `reg`.")
    return f, g, {"err": errs, "u": f, "v": g, "niter": i_ot}


def empirical_sinkhorn2_lazy(X_s, X_t, reg, a=None, b=None, numIterMax=10000, stopThr=1e-9,
                             warn=True, warmstart=None):
    f, g, log = empirical_sinkhorn_lazy(X_s, X_t, reg, a, b, numIterMax, stopThr, warn,
                                        warmstart)
    M = _sqeuclid(np.asarray(X_s, dtype=np.float64), np.asarray(X_t, dtype=np.float64))
    loss = float(np.sum(M * np.exp(f[:, None] + g[None, :] - M / reg)))
    return loss, log
