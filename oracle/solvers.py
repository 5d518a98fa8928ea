"""Float64 NumPy restatement of ``ot.solve``'s balanced entropic route.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows ot/solvers.py:428-486: defaults ``max_iter = 1000``, ``tol = 1e-9``;
``sinkhorn_log`` with ``log=True``; ``potentials = (log_u, log_v)``;
``value_linear = sum(M * P)``; ``value = value_linear + reg * sum(P log(P +
1e-16))`` for ``reg_type="entropy"`` or ``+ reg * kl_div(P, a b^T)`` with
``kl_div(p, q) = sum p log(p / q + 1e-16)`` (ot/backend.py:1542-1546) for
"KL"; "last_step" runs ``max_iter - 1`` iterations, then (:463-475)
``v' = log b - LSE_i(-M/reg + u)`` and ``u' = log a - LSE_j(-M/reg + v)``
from the pre-step pair ``(u, v)`` (u' is computed from the old v).

Pinning (tests/test_oracle.py): for any plan of the form
``P = exp(-M/reg + u_i + v_j)`` the two values satisfy, up to the 1e-16
guards, ``value_KL = reg (sum_i (u_i - log a_i) r_i + sum_j (v_j - log b_j)
c_j)`` and ``value_entropy = reg (sum_i u_i r_i + sum_j v_j c_j)`` with
``r, c`` the plan's marginals; and "last_step" returns the ``log_u`` of a
``max_iter - 1``-iteration solve with the ``log_v`` of a ``max_iter``-iteration
one (both run with stopThr = 0, so neither stops early): the last iteration
of sinkhorn_log sets ``u = c-transform(v)``, so ``u'`` equals ``u``.
"""

import numpy as np
from scipy.special import logsumexp

from .sinkhorn import sinkhorn_log


def solve_entropic(M, a=None, b=None, reg=1.0, reg_type="KL", max_iter=None, tol=None,
                   grad="autodiff"):
    """Returns (plan, (log_u, log_v), value, value_linear, niter)."""
    M = np.asarray(M, dtype=np.float64)
    n, m = M.shape
    a = np.full(n, 1.0 / n) if a is None else np.asarray(a, dtype=np.float64)
    b = np.full(m, 1.0 / m) if b is None else np.asarray(b, dtype=np.float64)
    if max_iter is None:
        max_iter = 1000
    if tol is None:
        tol = 1e-9
    if grad == "last_step":
        if max_iter == 0:
            raise ValueError("The maximum number of iterations must be greater than 0 "
                             "when using grad=last_step.")
        max_iter = max_iter - 1
    plan, log = sinkhorn_log(a, b, M, reg, numItermax=max_iter, stopThr=tol, log=True)
    u, v = log["log_u"], log["log_v"]
    niter = log["niter"]
    if grad == "last_step":
        v_new = np.log(b) - logsumexp(-M / reg + u[:, None], 0)
        u = np.log(a) - logsumexp(-M / reg + v[None, :], 1)
        v = v_new
        plan = np.exp(-M / reg + u[:, None] + v[None, :])
        niter = max_iter + 1
    value_linear = float(np.sum(M * plan))
    if reg_type.lower() == "entropy":
        value = value_linear + reg * float(np.sum(plan * np.log(plan + 1e-16)))
    else:
        value = value_linear + reg * float(np.sum(plan * np.log(plan / (a[:, None] * b[None, :])
                                                                + 1e-16)))
    return plan, (u, v), value, value_linear, niter
