"""Float64 NumPy restatement of ``ot.batch``'s Gromov-Wasserstein solver.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ot/batch/_quadratic.py: ``tensor_batch`` (:17-152; sqeuclidean loss
f1 = C^2, f2 = C^2, h1 = 2C, h2 = C, summed over a trailing feature axis for
(B, n, n, D) structures; KL loss f1 = 0, f2 = C2 log(C2 + 1e-15), h1 =
log(C1 + 1e-15) or C1 for logits, h2 = C2), ``tensor_product_batch`` (:535-572,
``const - hC1 T hC2^T``, averaged with the transposed product when not
symmetric), ``compute_const_from_marginals`` (:581-590) and
``solve_gromov_batch`` (:269-501: T0 = a b^T / sqrt(|a||b|), each step
``K = -((1 - alpha) M + 2 alpha L x T) / reg + log(T + 1e-15)`` projected by
``bregman_log_projection_batch``, stop when the largest mean |T - T_prev|
< tol; values recomputed with the plan's marginals).

Pinned (tests/test_oracle.py) by the brute-force 4-D tensor
``sum_jl L(C1_ij, C2_kl) T_jl`` on small problems (squared loss, symmetric
and not, with and without feature channels; KL loss), and the outer loop by
its closed form at alpha = 0: each step then projects ``exp(-M/reg) * T_prev``
and the projection absorbs T_prev's scalings, so after K steps with converged
projections T is the Sinkhorn plan of (M, reg / K).
"""

import numpy as np

from .batch import bregman_log_projection_batch


def _channels(C):
    return C if C.ndim == 4 else C[..., None]


def tensor_batch(a, b, C1, C2, symmetric=True, loss="sqeuclidean", logits=False):
    C1c, C2c = _channels(C1), _channels(C2)
    if loss == "sqeuclidean":
        fC1, fC2 = (C1c ** 2).sum(-1), (C2c ** 2).sum(-1)
        hC1, hC2 = 2 * C1c, C2c
    else:  # "kl"
        fC1 = np.zeros(C1.shape[:3])
        fC2 = (C2c * np.log(C2c + 1e-15)).sum(-1)
        hC1 = C1c if logits else np.log(C1c + 1e-15)
        hC2 = C2c
    if not symmetric:
        fC1 = 0.5 * (fC1 + fC1.transpose(0, 2, 1))
        fC2 = 0.5 * (fC2 + fC2.transpose(0, 2, 1))
    constC = np.einsum("bij,bj->bi", fC1, a)[:, :, None] + np.einsum("bij,bj->bi", fC2, b)[:, None, :]
    return {"constC": constC, "hC1": hC1, "hC2": hC2, "fC1": fC1, "fC2": fC2}


def tensor_product_batch(L, T, recompute_const=False, symmetric=True):
    if recompute_const:
        const = (np.einsum("bij,bj->bi", L["fC1"], T.sum(2))[:, :, None]
                 + np.einsum("bij,bj->bi", L["fC2"], T.sum(1))[:, None, :])
    else:
        const = L["constC"]
    dot = np.einsum("bijd,bjk,blkd->bil", L["hC1"], T, L["hC2"])
    if not symmetric:
        dot_t = np.einsum("bjid,bjk,bkld->bil", L["hC1"], T, L["hC2"])
        dot = (dot + dot_t) / 2
    return const - dot


def loss_quadratic_batch(L, T, recompute_const=False, symmetric=True):
    return np.sum(tensor_product_batch(L, T, recompute_const, symmetric) * T, axis=(1, 2))


def solve_gromov_batch(C1, C2, reg=1e-2, a=None, b=None, symmetric=None, M=None, alpha=None,
                       max_iter=50, tol=1e-5, max_iter_inner=50, tol_inner=1e-5,
                       loss="sqeuclidean", logits=False):
    B, n, m = C1.shape[0], C1.shape[1], C2.shape[1]
    a = np.ones((B, n)) / n if a is None else a
    b = np.ones((B, m)) / m if b is None else b
    if symmetric is None:
        # transpose of the two structure axes only, channels stay last
        # (ot/batch/_quadratic.py:575-578)
        symmetric = (np.allclose(C1, np.swapaxes(C1, 1, 2), atol=1e-10)
                     and np.allclose(C2, np.swapaxes(C2, 1, 2), atol=1e-10))
    L = tensor_batch(a, b, C1, C2, symmetric, loss=loss, logits=logits)
    if M is None:
        M, alpha = np.zeros((B, n, m)), 1.0
    const = np.sqrt(a.sum(1) * b.sum(1))
    T = a[:, :, None] * b[:, None, :] / const[:, None, None]
    log_T = np.log(T + 1e-15)
    it = -1
    for it in range(max_iter):
        T_prev = T
        Mk = (1 - alpha) * M + 2 * alpha * tensor_product_batch(L, T, symmetric=symmetric)
        out = bregman_log_projection_batch(-Mk / reg + log_T, a, b, max_iter=max_iter_inner,
                                           tol=tol_inner)
        T, log_T = out["T"], out["log_T"]
        if np.max(np.mean(np.abs(T_prev - T), axis=(1, 2))) < tol:
            break
    vl = np.sum(M * T, axis=(1, 2))
    vq = loss_quadratic_batch(L, T, recompute_const=True, symmetric=symmetric)
    return {"T": T, "value": (1 - alpha) * vl + alpha * vq, "value_linear": vl * (1 - alpha),
            "value_quad": vq * alpha, "n_iter": it, "potentials": out["potentials"]}
