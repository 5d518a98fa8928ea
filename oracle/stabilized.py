"""Float64 NumPy restatement of ``sinkhorn_stabilized`` and
``sinkhorn_epsilon_scaling``.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ot/bregman/_sinkhorn.py:1105-1367 (stabilized: u, v start at
1/dim; v = b / K^T u, u = a / K v; absorption into alpha, beta when
max|u| or max|v| exceeds tau, with u, v reset and K = exp(-(M - alpha -
beta) / reg) rebuilt; every print_period-th iteration err = ||sum_i Gamma -
b||_2; ``err <= stopThr`` stops; a NaN restores the previous u, v and stops;
log n_iter, logu, logv, alpha, beta, warmstart) and :1371-1566 (epsilon
scaling: reg_i = (epsilon0 - reg) e^-i + reg, inner stabilized solves warm
started by (alpha, beta), err = ||G^T 1 - b||^2 + ||G 1 - a||^2 every
print_period-th outer iteration, stop when err <= stopThr after 35 outer
iterations).  ``sinkhorn_stabilized_multi`` is the several-histogram branch
(a 2-D b, :1230-1367): scalings (n, H) and (m, H) starting at 1/dim,
absorption ``alpha += reg max_h log u``, ``beta += reg max log v`` without
resetting u, v, error ``(max|u - u'| / max(max|u|, max|u'|, 1) + the same
for v) / 2`` every print_period-th iteration, per-histogram losses
``sum(Gamma_h * M)`` and (n, H) / (m, H) logs.

Pinned (tests/test_oracle.py) by the doctests (_sinkhorn.py:1191-1193,
:1458-1460), by stabilized == sinkhorn_log at convergence (test_bregman.py:
282-314), and by absorption invariance (tau small enough to absorb at every
iteration gives the same plan).  The several-histogram branch is pinned by
its fixed point: at a tight tolerance each histogram's loss equals the
single-histogram loss of the same column, with and without absorption, and
identical columns give identical scalings.
"""

import warnings

import numpy as np


def sinkhorn_stabilized(a, b, M, reg, numItermax=1000, tau=1e3, stopThr=1e-9, warmstart=None,
                        print_period=20, log=False, warn=True):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    dim_a, dim_b = M.shape
    if len(a) == 0:
        a = np.full(dim_a, 1.0 / dim_a)
    if len(b) == 0:
        b = np.full(dim_b, 1.0 / dim_b)
    errs = []
    if warmstart is None:
        alpha, beta = np.zeros(dim_a), np.zeros(dim_b)
    else:
        alpha, beta = (np.asarray(w, dtype=np.float64).copy() for w in warmstart)
    u = np.full(dim_a, 1.0 / dim_a)
    v = np.full(dim_b, 1.0 / dim_b)

    def get_K(al, be):
        return np.exp(-(M - al[:, None] - be[None, :]) / reg)

    def get_Gamma(al, be, uu, vv):
        return np.exp(-(M - al[:, None] - be[None, :]) / reg + np.log(uu)[:, None]
                      + np.log(vv)[None, :])

    K = get_K(alpha, beta)
    err = 1
    ii = -1
    with np.errstate(all="ignore"):
        for ii in range(numItermax):
            uprev, vprev = u, v
            v = b / (K.T @ u)
            u = a / (K @ v)
            if np.max(np.abs(u)) > tau or np.max(np.abs(v)) > tau:
                alpha, beta = alpha + reg * np.log(u), beta + reg * np.log(v)
                u = np.full(dim_a, 1.0 / dim_a)
                v = np.full(dim_b, 1.0 / dim_b)
                K = get_K(alpha, beta)
            if ii % print_period == 0:
                err = np.linalg.norm(get_Gamma(alpha, beta, u, v).sum(0) - b)
                errs.append(err)
            if err <= stopThr:
                break
            if np.any(np.isnan(u)) or np.any(np.isnan(v)):
                warnings.warn("Numerical errors at iteration %d" % ii)
                u, v = uprev, vprev
                break
        else:
            if warn:
                warnings.warn("Sinkhorn did not converge. You might want to "
                              "increase the number of iterations `numItermax` "
                              "or the regularization parameter `reg`.")
        G = get_Gamma(alpha, beta, u, v)
        logd = {"err": errs, "n_iter": ii, "logu": alpha / reg + np.log(u),
                "logv": beta / reg + np.log(v), "alpha": alpha + reg * np.log(u),
                "beta": beta + reg * np.log(v)}
    logd["warmstart"] = (logd["alpha"], logd["beta"])
    return (G, logd) if log else G


def sinkhorn_stabilized_multi(a, b, M, reg, numItermax=1000, tau=1e3, stopThr=1e-9,
                              warmstart=None, print_period=20, log=False, warn=True):
    """The several-histogram branch of _sinkhorn.py:1230-1367 (b: (m, H))."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    dim_a, dim_b = M.shape
    if len(a) == 0:
        a = np.full(dim_a, 1.0 / dim_a)
    H = b.shape[1]
    a2 = a[:, None]
    errs = []
    if warmstart is None:
        alpha, beta = np.zeros(dim_a), np.zeros(dim_b)
    else:
        alpha, beta = (np.asarray(w, dtype=np.float64).copy() for w in warmstart)
    u = np.full((dim_a, H), 1.0 / dim_a)
    v = np.full((dim_b, H), 1.0 / dim_b)

    def get_K(al, be):
        return np.exp(-(M - al[:, None] - be[None, :]) / reg)

    K = get_K(alpha, beta)
    err = 1
    ii = -1
    with np.errstate(all="ignore"):
        for ii in range(numItermax):
            uprev, vprev = u, v
            v = b / (K.T @ u)
            u = a2 / (K @ v)
            if np.max(np.abs(u)) > tau or np.max(np.abs(v)) > tau:
                # the reference resets u, v only in its one-histogram branch
                alpha = alpha + reg * np.max(np.log(u), 1)
                beta = beta + reg * np.max(np.log(v))
                K = get_K(alpha, beta)
            if ii % print_period == 0:
                err_u = np.max(np.abs(u - uprev))
                err_u /= max(np.max(np.abs(u)), np.max(np.abs(uprev)), 1.0)
                err_v = np.max(np.abs(v - vprev))
                err_v /= max(np.max(np.abs(v)), np.max(np.abs(vprev)), 1.0)
                err = 0.5 * (err_u + err_v)
                errs.append(err)
            if err <= stopThr:
                break
            if np.any(np.isnan(u)) or np.any(np.isnan(v)):
                warnings.warn("Numerical errors at iteration %d" % ii)
                u, v = uprev, vprev
                break
        else:
            if warn:
                warnings.warn("Sinkhorn did not converge. You might want to "
                              "increase the number of iterations `numItermax` "
                              "or the regularization parameter `reg`.")
        al, be = alpha[:, None], beta[:, None]
        losses = np.array([np.sum(np.exp(-(M - al - be.T) / reg + np.log(u[:, [h]])
                                         + np.log(v[:, [h]]).T) * M) for h in range(H)])
        logd = {"err": errs, "n_iter": ii, "logu": al / reg + np.log(u),
                "logv": be / reg + np.log(v), "alpha": al + reg * np.log(u),
                "beta": be + reg * np.log(v)}
    logd["warmstart"] = (logd["alpha"], logd["beta"])
    return (losses, logd) if log else losses


def sinkhorn_epsilon_scaling(a, b, M, reg, numItermax=100, epsilon0=1e4, numInnerItermax=100,
                             tau=1e3, stopThr=1e-9, warmstart=None, print_period=10,
                             log=False, warn=True):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    M = np.asarray(M, dtype=np.float64)
    dim_a, dim_b = M.shape
    if len(a) == 0:
        a = np.full(dim_a, 1.0 / dim_a)
    if len(b) == 0:
        b = np.full(dim_b, 1.0 / dim_b)
    numItermin = 35
    numItermax = max(numItermin, numItermax)
    errs = []
    if warmstart is None:
        alpha, beta = np.zeros(dim_a), np.zeros(dim_b)
    else:
        alpha, beta = warmstart
    err = 1
    ii = -1
    G = None
    for ii in range(numItermax):
        regi = (epsilon0 - reg) * np.exp(-ii) + reg
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            G, logi = sinkhorn_stabilized(a, b, M, regi, numItermax=numInnerItermax,
                                          stopThr=stopThr, warmstart=(alpha, beta),
                                          print_period=20, tau=tau, log=True)
        alpha, beta = logi["alpha"], logi["beta"]
        if ii % print_period == 0:
            err = (np.linalg.norm(G.sum(0) - b) ** 2 + np.linalg.norm(G.sum(1) - a) ** 2)
            errs.append(err)
        if err <= stopThr and ii > numItermin:
            break
    else:
        if warn:
            warnings.warn("Sinkhorn did not converge. You might want to "
                          "increase the number of iterations `numItermax` "
                          "or the regularization parameter `reg`.")
    logd = {"err": errs, "alpha": alpha, "beta": beta, "warmstart": (alpha, beta), "niter": ii}
    return (G, logd) if log else G
