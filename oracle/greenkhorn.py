"""Float64 NumPy restatement of POT's greedy Sinkhorn (``ot.bregman.greenkhorn``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ot/bregman/_sinkhorn.py:921-1102 step by step: uniform weights for
empty ``a``/``b`` (:1034-1037), ``K = exp(-M / reg)`` (:1041), the start
scalings (:1043-1047), ``G = u K v`` and the violations (:1048-1051), the
greedy loop (:1058-1086: both argmaxes, the stop value read before the
update, the row or column update with the incremental violation updates, the
stop after the update) and the warning and log (:1087-1102).

Pinning (tests/test_oracle.py):

* the reference's doctest (_sinkhorn.py:995-1001): the symmetric 2x2
  problem's plan to 8 digits;
* the converged plan equals the sinkhorn_log plan of the same problem (both
  solve the same entropic problem; tight stopThr);
* the incrementally updated violations equal the recomputed marginals of
  ``u K v`` minus the weights at every stage (a dropped or mis-signed term of
  the incremental update fails this);
* the first two iterations by hand on a 2x3 problem (which index is chosen,
  which scaling moves, and the stop taken after the update).
"""

import warnings

import numpy as np

_NOT_CONVERGED = (
    "Sinkhorn did not converge. You might want to "
    "increase the number of iterations `numItermax` "
    "or the regularization parameter `reg`."
)


def greenkhorn(a, b, M, reg, numItermax=10000, stopThr=1e-9, verbose=False, log=False,
               warn=True, warmstart=None, _trace=None):
    """Greedy Sinkhorn (_sinkhorn.py:1025-1102).  ``_trace`` (a list, tests
    only) receives ``(ii, row, index, u, v, viol, viol_2)`` after each update."""
    a = np.asarray(a, dtype=np.float64) if isinstance(a, list) else a
    b = np.asarray(b, dtype=np.float64) if isinstance(b, list) else b
    M = np.asarray(M, dtype=np.float64) if isinstance(M, list) else M
    if len(a) == 0:
        a = np.ones((M.shape[0],), dtype=M.dtype) / M.shape[0]
    if len(b) == 0:
        b = np.ones((M.shape[1],), dtype=M.dtype) / M.shape[1]
    dim_a, dim_b = a.shape[0], b.shape[0]

    K = np.exp(-M / reg)
    if warmstart is None:
        u = np.full((dim_a,), 1.0 / dim_a, dtype=K.dtype)
        v = np.full((dim_b,), 1.0 / dim_b, dtype=K.dtype)
    else:
        u, v = np.exp(warmstart[0]), np.exp(warmstart[1])
    G = u[:, None] * K * v[None, :]

    viol = np.sum(G, axis=1) - a
    viol_2 = np.sum(G, axis=0) - b
    logd = {"u": u, "v": v} if log else None

    ii = -1
    stopped = False
    for ii in range(numItermax):
        i_1 = int(np.argmax(np.abs(viol)))
        i_2 = int(np.argmax(np.abs(viol_2)))
        m_viol_1 = np.abs(viol[i_1])
        m_viol_2 = np.abs(viol_2[i_2])
        stop_val = np.maximum(m_viol_1, m_viol_2)
        if m_viol_1 > m_viol_2:
            old_u = u[i_1]
            new_u = a[i_1] / np.dot(K[i_1, :], v)
            G[i_1, :] = new_u * K[i_1, :] * v
            viol[i_1] = np.dot(new_u * K[i_1, :], v) - a[i_1]
            viol_2 += K[i_1, :].T * (new_u - old_u) * v
            u[i_1] = new_u
            if _trace is not None:
                _trace.append((ii, True, i_1, u.copy(), v.copy(), viol.copy(), viol_2.copy()))
        else:
            old_v = v[i_2]
            new_v = b[i_2] / np.dot(K[:, i_2].T, u)
            G[:, i_2] = u * K[:, i_2] * new_v
            viol += (-old_v + new_v) * K[:, i_2] * u
            viol_2[i_2] = new_v * np.dot(K[:, i_2], u) - b[i_2]
            v[i_2] = new_v
            if _trace is not None:
                _trace.append((ii, False, i_2, u.copy(), v.copy(), viol.copy(), viol_2.copy()))
        if stop_val <= stopThr:
            stopped = True
            break
    if not stopped and warn:
        warnings.warn(_NOT_CONVERGED)

    if log:
        if ii < 0:
            raise UnboundLocalError("numItermax must be positive when log=True")
        logd["n_iter"] = ii
        logd["u"] = u
        logd["v"] = v
        return G, logd
    return G
