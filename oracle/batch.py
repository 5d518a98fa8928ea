"""Float64 NumPy restatement of POT's batched linear OT solvers (``ot.batch``).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows:

* ``entropy_batch``, ``norm_batch``             ot/batch/_utils.py:4-29
* ``bregman_projection_batch``                  ot/batch/_utils.py:46-162
* ``bregman_log_projection_batch``              ot/batch/_utils.py:165-290
* ``dist_euclidean_batch``, ``dist_lp_batch``,
  ``dist_kl_batch``, ``dist_batch``             ot/batch/_linear.py:20-233
* ``loss_linear_batch``                         ot/batch/_linear.py:123-147
* ``solve_batch``, ``solve_sample_batch``       ot/batch/_linear.py:236-435
* ``sinkhorn_log_ragged``: per-problem ``sinkhorn_log`` (oracle.sinkhorn) over a
  list of problems of uneven sizes -- the checker of the ragged batched solver,
  which has no reference counterpart (the reference batch needs one (n, m)).

Pinning: ``test_solve_batch.py:26-62`` (each problem equals ``ot.solve``, i.e.
``sinkhorn_log``), ``:65-78`` (plain and log projections agree), ``:81-110``
(the three loss routes agree), plus per-problem invariants of sinkhorn_log.
"""

import numpy as np
from scipy.special import logsumexp

from . import sinkhorn as _sk


def entropy_batch(T, eps=1e-16):
    """``-sum T log(T + eps)`` over each plan (_utils.py:4-22)."""
    return -np.sum(T * np.log(T + eps), axis=(1, 2))


def norm_batch(u, p=2):
    """``(sum u^p)^(1/p)`` per row, no absolute value (_utils.py:25-29)."""
    return np.sum(u ** p, axis=1) ** (1.0 / p)


def bregman_projection_batch(K, a=None, b=None, max_iter=10000, tol=1e-5):
    """Batched scaling Sinkhorn (_utils.py:127-162), rows first.

    ``f = a / (K g)``, ``g = b / (K^T f)``; every 10th iteration the largest
    row-marginal error over the batch is compared with ``tol``.
    """
    B, n, m = K.shape
    a = np.ones((B, n)) / n if a is None else a
    b = np.ones((B, m)) / m if b is None else b
    f = np.ones((B, n))
    g = np.ones((B, m))
    it = -1
    for it in range(max_iter):
        f = a / np.sum(K * g[:, None, :], axis=2)
        g = b / np.sum(K * f[:, :, None], axis=1)
        if it % 10 == 0:
            T = K * f[:, :, None] * g[:, None, :]
            err = np.max(norm_batch(np.sum(T, axis=2) - a))
            if err < tol:
                break
    T = K * f[:, :, None] * g[:, None, :]
    return {"T": T, "n_iters": it, "potentials": (f, g)}


def bregman_log_projection_batch(K, a=None, b=None, max_iter=10000, tol=1e-5):
    """Batched log-domain Sinkhorn (_utils.py:249-290), rows first.

    ``u = log a - LSE_j(K + v)``, ``v = log b - LSE_i(K + u)``; every 10th
    iteration the largest row-marginal error over the batch is compared with
    ``tol`` (one stop for the whole batch).
    """
    B, n, m = K.shape
    a = np.ones((B, n)) / n if a is None else a
    b = np.ones((B, m)) / m if b is None else b
    u = np.zeros((B, n), dtype=K.dtype)
    v = np.zeros((B, m), dtype=K.dtype)
    loga, logb = np.log(a), np.log(b)
    it = -1
    for it in range(max_iter):
        u = loga - logsumexp(K + v[:, None, :], axis=2)
        v = logb - logsumexp(K + u[:, :, None], axis=1)
        if it % 10 == 0:
            T = np.exp(K + u[:, :, None] + v[:, None, :])
            err = np.max(norm_batch(np.sum(T, axis=2) - a))
            if err < tol:
                break
    log_T = K + u[:, :, None] + v[:, None, :]
    return {"T": np.exp(log_T), "log_T": log_T, "n_iters": it, "potentials": (u, v)}


def dist_euclidean_batch(X, Y, squared=True):
    """``|x|^2 + |y|^2 - 2 x.y`` per pair, not clamped (_linear.py:54-85)."""
    XX = np.sum(X ** 2, axis=-1, keepdims=True)
    YY = np.transpose(np.sum(Y ** 2, axis=-1, keepdims=True), (0, 2, 1))
    M = XX + YY - 2 * np.einsum("bid,bjd->bij", X, Y)
    return M if squared else np.sqrt(M)


def dist_lp_batch(X, Y, p=2, q=1):
    """``(sum |x - y|^p)^(q/p)`` (_linear.py:20-51)."""
    M = np.sum(np.abs(X[:, :, None] - Y[:, None, :]) ** p, axis=-1)
    return M ** (q / p) if q != p else M


def dist_kl_batch(X, Y, eps=1e-10, logits_X=False):
    """``KL(y_j || x_i)`` per pair with probability-vector rows (_linear.py:88-120).

    ``logits_X``: X holds log probabilities, the cross term is ``sum_d y x``
    (the docstring formula; the reference's logits branch, :115-116, omits
    the sum over d)."""
    entr_y = np.sum(Y * np.log(Y + eps), axis=-1)
    hX = X if logits_X else np.log(X + eps)
    cross = np.sum(Y[:, None, :, :] * hX[:, :, None, :], axis=-1)
    return entr_y[:, None, :] - cross


def dist_batch(X1, X2=None, metric="sqeuclidean", p=2):
    """Metric dispatch (_linear.py:178-233)."""
    X2 = X1 if X2 is None else X2
    metric = metric.lower()
    if metric == "sqeuclidean":
        return dist_euclidean_batch(X1, X2, squared=True)
    if metric == "euclidean":
        return dist_euclidean_batch(X1, X2, squared=False)
    if metric == "minkowski":
        return dist_lp_batch(X1, X2, p=p, q=1)
    if metric == "kl":
        return dist_kl_batch(X1, X2)
    raise ValueError(f"Unknown metric: {metric}")


def loss_linear_batch(M, T):
    """``<M_b, T_b>`` per problem (_linear.py:123-147)."""
    return np.sum(M * T, axis=(1, 2))


def solve_batch(M, reg, a=None, b=None, max_iter=1000, tol=1e-5,
                solver="log_sinkhorn", reg_type="entropy"):
    """Batched entropic solve (_linear.py:304-357), gradients ignored.

    Returns a dict with the fields the reference stores in its OTResult.
    """
    B, n, m = M.shape
    a = np.ones((B, n)) / n if a is None else a
    b = np.ones((B, m)) / m if b is None else b
    if solver == "log_sinkhorn":
        out = bregman_log_projection_batch(-M / reg, a, b, max_iter=max_iter, tol=tol)
    elif solver == "sinkhorn":
        out = bregman_projection_batch(np.exp(-M / reg), a, b, max_iter=max_iter, tol=tol)
    else:
        raise ValueError(f"Unknown solver: {solver}")
    T = out["T"]
    value_linear = loss_linear_batch(M, T)
    if reg_type.lower() == "entropy":
        value = value_linear + reg * (-entropy_batch(T))
    elif reg_type.lower() == "kl":
        ref = np.einsum("bi,bj->bij", a, b)
        value = value_linear + reg * np.sum(T * np.log(T / ref + 1e-16), axis=(1, 2))
    else:
        raise ValueError(f"Unknown reg_type: {reg_type}")
    return {"value": value, "value_linear": value_linear,
            "potentials": out["potentials"], "plan": T,
            "log": {"n_iter": out["n_iters"]}}


def solve_sample_batch(X_a, X_b, reg, a=None, b=None, metric="sqeuclidean", p=2,
                       max_iter=1000, tol=1e-5, solver="log_sinkhorn",
                       reg_type="entropy"):
    """``dist_batch`` then ``solve_batch`` (_linear.py:421-435)."""
    M = dist_batch(X_a, X_b, metric=metric, p=p)
    return solve_batch(M, reg, a=a, b=b, max_iter=max_iter, tol=tol,
                       solver=solver, reg_type=reg_type)


def sinkhorn_log_ragged(problems, numItermax=1000, stopThr=1e-9):
    """Per-problem ``sinkhorn_log`` over ``(a, b, M, reg)`` tuples.

    Returns a list of ``(P, log)`` pairs; each problem stops on its own.
    """
    out = []
    for a, b, M, reg in problems:
        out.append(_sk.sinkhorn_log(a, b, M, reg, numItermax=numItermax,
                                    stopThr=stopThr, log=True, warn=False))
    return out
