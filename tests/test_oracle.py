"""Pins the float64 oracle (CPU only).

Each check is independent of the oracle's own loop: closed forms, separable
costs, an L-BFGS solve of the entropic dual, a pure-Python first iterate, and
the invariants the reference tests hold (test/test_bregman.py,
test/batch/test_solve_batch.py).
"""

import math
import warnings

import numpy as np
import pytest
from scipy.optimize import minimize

import oracle.batch as ob
import oracle.sinkhorn as osk
import oracle.utils as ou


def _problem(n, m, seed, reg_scale=1.0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 2))
    y = rng.standard_normal((m, 2)) + 0.7
    M = ((x[:, None] - y[None]) ** 2).sum(-1)
    return rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m)), M / M.max() * reg_scale


# --- closed forms ----------------------------------------------------------

def test_symmetric_2x2_doctest_values():
    # ot/bregman/_sinkhorn.py:123-129, :342-347, :563-569, :766-772
    a = [0.5, 0.5]
    M = [[0.0, 1.0], [1.0, 0.0]]
    exp = np.array([[0.36552929, 0.13447071], [0.13447071, 0.36552929]])
    for method in ("sinkhorn", "sinkhorn_log"):
        P = osk.sinkhorn(a, a, M, 1, method=method)
        np.testing.assert_allclose(P, exp, atol=5e-9)
        np.testing.assert_allclose(P[0, 0], math.e / (2 * (1 + math.e)), rtol=1e-12)
        loss = osk.sinkhorn2(a, a, M, 1, method=method)
        np.testing.assert_allclose(loss, 0.26894142136999516, rtol=1e-12)
        np.testing.assert_allclose(loss, 1 / (1 + math.e), rtol=1e-12)


def _general_2x2(a, b, M, reg):
    """Entropic 2x2 plan from its one free entry: P11 P22 / (P12 P21) = kappa."""
    kappa = math.exp(-(M[0][0] + M[1][1] - M[0][1] - M[1][0]) / reg)
    a1, b1 = a[0], b[0]
    # p (1 - a1 - b1 + p) = kappa (a1 - p)(b1 - p)
    A = 1 - kappa
    B = (1 - a1 - b1) + kappa * (a1 + b1)
    C = -kappa * a1 * b1
    if abs(A) < 1e-15:
        p = -C / B
    else:
        disc = math.sqrt(B * B - 4 * A * C)
        roots = [(-B + disc) / (2 * A), (-B - disc) / (2 * A)]
        lo, hi = max(0.0, a1 + b1 - 1), min(a1, b1)
        p = [r for r in roots if lo - 1e-15 <= r <= hi + 1e-15][0]
    return np.array([[p, a1 - p], [b1 - p, 1 - a1 - b1 + p]])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_general_2x2_closed_form(seed):
    rng = np.random.default_rng(seed)
    a = rng.dirichlet([1, 1])
    b = rng.dirichlet([1, 1])
    M = rng.random((2, 2))
    reg = 0.3
    exact = _general_2x2(a, b, M, reg)
    for method in ("sinkhorn", "sinkhorn_log"):
        P = osk.sinkhorn(a, b, M, reg, method=method, stopThr=1e-15, numItermax=5000,
                         warn=False)
        np.testing.assert_allclose(P, exact, rtol=1e-9, atol=1e-14)


def test_separable_cost_gives_product_plan():
    rng = np.random.default_rng(3)
    a, b = rng.dirichlet(np.ones(7)), rng.dirichlet(np.ones(5))
    M = rng.random(7)[:, None] + rng.random(5)[None, :]
    for method in ("sinkhorn", "sinkhorn_log"):
        with pytest.warns(UserWarning):
            P = osk.sinkhorn(a, b, M, 0.05, method=method, numItermax=1, stopThr=0)
        np.testing.assert_allclose(P, np.outer(a, b), rtol=1e-10)


def test_large_reg_tends_to_product_plan():
    a, b, M = _problem(9, 11, 4)
    P = osk.sinkhorn_log(a, b, M, 1e4, stopThr=1e-14)
    np.testing.assert_allclose(P, np.outer(a, b), rtol=1e-3)


# --- brute force: L-BFGS on the smooth entropic dual -----------------------

def _dual_plan(a, b, M, reg):
    n, m = M.shape

    def negdual(z):
        f, g = z[:n], z[n:]
        P = np.exp((f[:, None] + g[None, :] - M) / reg - 1.0)
        val = f @ a + g @ b - reg * P.sum()
        grad = np.concatenate([a - P.sum(1), b - P.sum(0)])
        return -val, -grad

    r = minimize(negdual, np.zeros(n + m), jac=True, method="L-BFGS-B",
                 options={"gtol": 1e-14, "ftol": 1e-16, "maxiter": 20000, "maxcor": 50})
    f, g = r.x[:n], r.x[n:]
    return np.exp((f[:, None] + g[None, :] - M) / reg - 1.0)


@pytest.mark.parametrize("method", ["sinkhorn", "sinkhorn_log"])
def test_matches_dual_bruteforce(method):
    a, b, M = _problem(8, 6, 5)
    reg = 0.2
    P_dual = _dual_plan(a, b, M, reg)
    P = osk.sinkhorn(a, b, M, reg, method=method, stopThr=1e-14, numItermax=20000,
                     warn=False)
    np.testing.assert_allclose(P, P_dual, rtol=1e-6, atol=1e-12)


# --- invariants --------------------------------------------------------------

def test_fixed_point_and_marginals():
    a, b, M = _problem(40, 30, 6)
    reg = 0.05
    P, lg = osk.sinkhorn_log(a, b, M, reg, stopThr=1e-12, numItermax=20000, log=True)
    Mr = -M / reg
    lu, lv = lg["log_u"], lg["log_v"]
    # u is the last update: exact row fixed point, exact row marginal
    ref_u = np.log(a) - np.log(np.exp(Mr + lv[None, :]).sum(1))
    np.testing.assert_allclose(lu, ref_u, atol=1e-12)
    np.testing.assert_allclose(P.sum(1), a, rtol=1e-12)
    assert np.linalg.norm(P.sum(0) - b) < 1e-12
    # primal value <P,M> + reg sum P log P equals reg (<log_u, a> + <log_v, b>)
    # at the fixed point (log P = -M/reg + u + v with marginals a, b)
    primal = np.sum(P * M) + reg * np.sum(P * np.log(P))
    np.testing.assert_allclose(primal, reg * (lu @ a + lv @ b), rtol=1e-9)


def test_gauge_invariance_of_warmstart():
    a, b, M = _problem(12, 15, 7)
    P0 = osk.sinkhorn_log(a, b, M, 0.1, numItermax=37, stopThr=0, warn=False)
    P1 = osk.sinkhorn_log(a, b, M, 0.1, numItermax=37, stopThr=0, warn=False,
                          warmstart=(np.full(12, 3.25), np.full(15, -3.25)))
    np.testing.assert_allclose(P0, P1, rtol=1e-11)
    # shifting M by row/column constants leaves the plan unchanged
    P2 = osk.sinkhorn_log(a, b, M + np.arange(12)[:, None] * 0.01 + 0.02 * np.arange(15),
                          0.1, stopThr=1e-13, numItermax=20000)
    P3 = osk.sinkhorn_log(a, b, M, 0.1, stopThr=1e-13, numItermax=20000)
    np.testing.assert_allclose(P2, P3, rtol=1e-9)


def test_knopp_equals_log():
    # test_bregman.py:282-314
    a, b, M = _problem(25, 20, 8)
    P1 = osk.sinkhorn_knopp(a, b, M, 0.1, stopThr=1e-13, numItermax=20000)
    P2 = osk.sinkhorn_log(a, b, M, 0.1, stopThr=1e-13, numItermax=20000)
    np.testing.assert_allclose(P1, P2, rtol=1e-8, atol=1e-16)


# --- iteration semantics ----------------------------------------------------

def _first_iterate_pure_python(a, b, M, reg):
    n, m = len(a), len(b)
    v = []
    for j in range(m):
        v.append(math.log(b[j]) - math.log(sum(math.exp(-M[i][j] / reg) for i in range(n))))
    u = []
    for i in range(n):
        u.append(math.log(a[i]) - math.log(sum(math.exp(-M[i][j] / reg + v[j])
                                               for j in range(m))))
    colsum = [sum(math.exp(-M[i][j] / reg + u[i] + v[j]) for i in range(n)) for j in range(m)]
    err = math.sqrt(sum((colsum[j] - b[j]) ** 2 for j in range(m)))
    return u, v, err


def test_first_iterate_and_check_cadence():
    a, b, M = _problem(6, 5, 9)
    reg = 0.07
    u1, v1, err0 = _first_iterate_pure_python(a, b, M, reg)
    with pytest.warns(UserWarning, match="did not converge"):
        P, lg = osk.sinkhorn_log(a, b, M, reg, numItermax=1, stopThr=0, log=True)
    assert lg["niter"] == 0 and len(lg["err"]) == 1
    np.testing.assert_allclose(lg["err"][0], err0, rtol=1e-12)
    np.testing.assert_allclose(lg["log_u"], u1, atol=1e-13)
    np.testing.assert_allclose(lg["log_v"], v1, atol=1e-13)
    # 11 iterations: checks at ii = 0 and 10
    with pytest.warns(UserWarning):
        _, lg = osk.sinkhorn_log(a, b, M, reg, numItermax=11, stopThr=0, log=True)
    assert lg["niter"] == 10 and len(lg["err"]) == 2
    # a generous threshold stops at the first check, strict '<', no warning
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        _, lg = osk.sinkhorn_log(a, b, M, reg, stopThr=err0 * 1.0000001, log=True)
    assert lg["niter"] == 0
    with pytest.warns(UserWarning):
        _, lg = osk.sinkhorn_log(a, b, M, reg, stopThr=err0, numItermax=5, log=True)
    assert lg["niter"] == 4  # err0 < err0 is false


def test_knopp_numerical_guard_and_warnings():
    a = ou.make_1D_gauss(100, m=30, s=10)
    b = ou.make_1D_gauss(100, m=40, s=10)
    M = ou.dist0(100)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        with pytest.warns(UserWarning, match="numerical errors"):
            osk.sinkhorn(a, b, M, 0, method="sinkhorn", warn=False)
    for method in ("sinkhorn", "sinkhorn_log"):
        with pytest.warns(UserWarning):
            osk.sinkhorn(a, b, M, 1.0, method=method, stopThr=0, numItermax=1)
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            osk.sinkhorn2(a, b, M, 1, method=method, stopThr=0, numItermax=1, warn=False)
    with pytest.raises(ValueError):
        osk.sinkhorn(a, b, M, 1.0, method="new_method")


def test_empty_weights_default_uniform():
    _, _, M = _problem(10, 12, 10)
    P0 = osk.sinkhorn_log([], [], M, 0.1)
    P1 = osk.sinkhorn_log(np.ones(10) / 10, np.ones(12) / 12, M, 0.1)
    np.testing.assert_allclose(P0, P1)


def test_multi_histogram_log_equals_per_column():
    rng = np.random.RandomState(0)
    x = rng.randn(10, 2)
    u = np.ones(10) / 10
    b = rng.rand(10, 3)
    b = b / b.sum(0, keepdims=True)
    M = ou.dist(x, x)
    for method in ("sinkhorn", "sinkhorn_log"):
        loss0 = osk.sinkhorn(u, b, M, 0.1, method=method, stopThr=1e-10)
        loss = [osk.sinkhorn2(u, b[:, k], M, 0.1, method=method, stopThr=1e-10)
                for k in range(3)]
        np.testing.assert_allclose(loss0, loss, atol=1e-4)


# --- batch --------------------------------------------------------------------

@pytest.mark.parametrize("solver", ["sinkhorn", "log_sinkhorn"])
@pytest.mark.parametrize("reg_type", ["kl", "entropy"])
def test_solve_batch_equals_per_problem_sinkhorn_log(solver, reg_type):
    # test/batch/test_solve_batch.py:26-62 with the oracle's sinkhorn_log as ot.solve
    rng = np.random.RandomState(0)
    M = rng.rand(4, 16, 16)
    res = ob.solve_batch(M, 0.1, max_iter=10000, tol=1e-5, solver=solver, reg_type=reg_type)
    for i in range(4):
        P = osk.sinkhorn_log(np.ones(16) / 16, np.ones(16) / 16, M[i], 0.1,
                             numItermax=10000, stopThr=1e-5)
        np.testing.assert_allclose(res["plan"][i], P, atol=1e-5)
        np.testing.assert_allclose(res["value_linear"][i], np.sum(M[i] * P), atol=1e-4)


def test_batch_log_and_plain_agree_and_losses():
    rng = np.random.RandomState(0)
    X = rng.rand(4, 4, 2)
    M = ob.dist_batch(X, X)
    r1 = ob.bregman_projection_batch(np.exp(-M / 0.01), max_iter=50, tol=1e-10)
    r2 = ob.bregman_log_projection_batch(-M / 0.01, max_iter=50, tol=1e-10)
    np.testing.assert_allclose(r1["T"], r2["T"], atol=1e-3)
    res = ob.solve_batch(M, 0.1, max_iter=10)
    np.testing.assert_allclose(res["value_linear"], ob.loss_linear_batch(M, res["plan"]),
                               atol=1e-12)


def test_dist_bruteforce():
    # test/test_utils.py:207-227
    rng = np.random.default_rng(11)
    x = rng.standard_normal((13, 4))
    y = rng.standard_normal((9, 4))
    D = ou.dist(x, y)
    brute = np.array([[np.sum((xi - yj) ** 2) for yj in y] for xi in x])
    np.testing.assert_allclose(D, brute, atol=1e-14)
    Db = ob.dist_batch(x[None], y[None])[0]
    np.testing.assert_allclose(Db, brute, atol=1e-13)
    np.testing.assert_allclose(ou.dist(x, y, "euclidean"), np.sqrt(brute), atol=1e-13)


def test_solve_entropic_value_identities():
    # value formulas of ot/solvers.py:476-486 against the exact identities
    # value_KL = reg (sum (u - log a) r + sum (v - log b) c),
    # value_ent = reg (sum u r + sum v c), r, c the plan's marginals.
    from oracle.solvers import solve_entropic
    rng = np.random.default_rng(4)
    n, m, reg = 30, 25, 0.2
    M = rng.random((n, m))
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(m))
    for reg_type in ("KL", "entropy"):
        P, (u, v), val, vl, _ = solve_entropic(M, a, b, reg, reg_type=reg_type, max_iter=37,
                                              tol=0.0)
        r, c = P.sum(1), P.sum(0)
        assert abs(vl - np.sum(M * P)) < 1e-14
        if reg_type == "KL":
            ident = reg * (np.sum((u - np.log(a)) * r) + np.sum((v - np.log(b)) * c))
        else:
            ident = reg * (np.sum(u * r) + np.sum(v * c))
        assert abs(val - ident) < 1e-9 * max(1.0, abs(ident))


def test_solve_entropic_last_step_pins():
    # Not converged (small reg, Dirichlet weights, few iterations), so a wrong
    # update order shows.  The reference's last step (ot/solvers.py:463-475)
    # computes v' from the last u and u' from the last v; since sinkhorn_log's
    # last iteration already set u = c-transform(v), u' is the (max_iter - 1)
    # run's log_u and v' the max_iter run's log_v.
    from oracle.sinkhorn import sinkhorn_log
    from oracle.solvers import solve_entropic
    rng = np.random.default_rng(5)
    n, m, reg, it = 40, 30, 0.01, 12
    M = rng.random((n, m))
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(m))
    P1, (u1, v1), val1, vl1, n1 = solve_entropic(M, a, b, reg=reg, max_iter=it, tol=0.0,
                                                 grad="last_step", reg_type="entropy")
    _, lg_a = sinkhorn_log(a, b, M, reg, numItermax=it - 1, stopThr=0.0, log=True, warn=False)
    _, lg_b = sinkhorn_log(a, b, M, reg, numItermax=it, stopThr=0.0, log=True, warn=False)
    np.testing.assert_allclose(u1, lg_a["log_u"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(v1, lg_b["log_v"], rtol=0, atol=1e-12)
    assert n1 == it
    # v' is a column half-step: the plan's column marginal is b exactly, its
    # row marginal is not (the solve has not converged)
    np.testing.assert_allclose(P1.sum(0), b, rtol=1e-12)
    assert np.abs(P1.sum(1) - a).max() > 1e-4
    # ... and differs from the plan of a plain max_iter solve
    P2, _, _, _, _ = solve_entropic(M, a, b, reg=reg, max_iter=it, tol=0.0)
    assert np.abs(P1 - P2).max() > 1e-6
    np.testing.assert_allclose(vl1, np.sum(M * P1), rtol=1e-14)
    np.testing.assert_allclose(val1, vl1 + reg * np.sum(P1 * np.log(P1 + 1e-16)), rtol=1e-14)


def test_empirical_lazy_doctest_and_fixed_point():
    # doctest of ot/bregman/_empirical.py:152-157 (reg = 0.1, two points)
    from oracle.empirical import empirical_sinkhorn_lazy
    X = np.arange(2, dtype=np.float64).reshape(2, 1)
    f, g, log = empirical_sinkhorn_lazy(X, X, 0.1)
    M = (X - X.T) ** 2
    P = np.exp(f[:, None] + g[None, :] - M / 0.1)
    np.testing.assert_allclose(P, [[4.99977301e-01, 2.26989344e-05],
                                   [2.26989344e-05, 4.99977301e-01]], rtol=1e-8)
    assert log["niter"] == 9 and len(log["err"]) == 1   # converges at the first check
    rng = np.random.default_rng(0)
    xs, xt = rng.standard_normal((20, 2)), rng.standard_normal((15, 2)) + 1
    f, g, log = empirical_sinkhorn_lazy(xs, xt, 0.5, stopThr=1e-12)
    M = ((xs[:, None] - xt[None]) ** 2).sum(-1)
    from scipy.special import logsumexp
    np.testing.assert_allclose(f, np.log(1 / 20) - logsumexp(g[None] - M / 0.5, 1), atol=1e-10)
    np.testing.assert_allclose(g, np.log(1 / 15) - logsumexp(f[:, None] - M / 0.5, 0),
                               atol=1e-12)
    assert (log["niter"] + 1) % 10 == 0 and log["err"][-1] <= 1e-12


@pytest.mark.parametrize("symmetric", [True, False])
def test_gw_tensor_product_bruteforce(symmetric):
    # L x T of ot/batch/_quadratic.py against the 4-D tensor
    # sum_jl l(C1_ij, C2_kl) T_jl (averaged with l(C1_ji, C2_lk) when the
    # costs are not symmetric), l the squared difference.
    from oracle.quadratic import tensor_batch, tensor_product_batch
    rng = np.random.default_rng(7)
    B, n, m = 2, 5, 4
    C1 = rng.random((B, n, n))
    C2 = rng.random((B, m, m))
    if symmetric:
        C1 = C1 + C1.transpose(0, 2, 1)
        C2 = C2 + C2.transpose(0, 2, 1)
    T = rng.random((B, n, m))
    T /= T.sum(axis=(1, 2), keepdims=True)
    L = tensor_batch(T.sum(2), T.sum(1), C1, C2, symmetric=symmetric)
    got = tensor_product_batch(L, T, symmetric=symmetric)
    loss = (C1[:, :, :, None, None] - C2[:, None, None, :, :]) ** 2   # b i j k l
    ref = np.einsum("bijkl,bjl->bik", loss, T)
    if not symmetric:
        loss_t = (C1.transpose(0, 2, 1)[:, :, :, None, None]
                  - C2.transpose(0, 2, 1)[:, None, None, :, :]) ** 2
        ref = 0.5 * (ref + np.einsum("bijkl,bjl->bik", loss_t, T))
    np.testing.assert_allclose(got, ref, rtol=1e-12)


def test_gw_tensor_product_bruteforce_channels_and_kl():
    # (B, n, n, D) structures: l(c1, c2) = sum_d (c1_d - c2_d)^2; KL loss:
    # l(c1, c2) = sum_d c2_d log(c2_d + 1e-15) - c2_d h(c1_d), h = log(. + 1e-15)
    # (or the identity for logits).
    from oracle.quadratic import tensor_batch, tensor_product_batch
    rng = np.random.default_rng(11)
    B, n, m, D = 2, 5, 4, 3
    C1 = rng.random((B, n, n, D)) + 0.1
    C2 = rng.random((B, m, m, D)) + 0.1
    T = rng.random((B, n, m))
    T /= T.sum(axis=(1, 2), keepdims=True)
    a, b = T.sum(2), T.sum(1)
    cases = {
        ("sqeuclidean", False): ((C1[:, :, :, None, None] - C2[:, None, None]) ** 2).sum(-1),
        ("kl", False): (C2[:, None, None] * np.log(C2[:, None, None] + 1e-15)
                        - C2[:, None, None] * np.log(C1[:, :, :, None, None] + 1e-15)).sum(-1),
        ("kl", True): (C2[:, None, None] * np.log(C2[:, None, None] + 1e-15)
                       - C2[:, None, None] * C1[:, :, :, None, None]).sum(-1),
    }
    for (loss, logits), tens in cases.items():
        L = tensor_batch(a, b, C1, C2, symmetric=False, loss=loss, logits=logits)
        got = tensor_product_batch(L, T, symmetric=False)
        ref = np.einsum("bijkl,bjl->bik", tens, T)
        tens_t = tens.transpose(0, 2, 1, 4, 3)
        ref = 0.5 * (ref + np.einsum("bijkl,bjl->bik", tens_t, T))
        # the const of fC1 / fC2 uses the symmetrised f: with T's own
        # marginals as a, b it equals the brute force
        np.testing.assert_allclose(got, ref, rtol=1e-12, err_msg=f"{loss} logits={logits}")


def test_gw_outer_loop_alpha0_closed_form():
    # alpha = 0 with M given: each outer step projects exp(-M/reg) * T_prev;
    # the projection absorbs T_prev's row and column scalings, so after K
    # steps with converged projections T is sinkhorn_log(a, b, M, reg / K).
    from oracle.quadratic import solve_gromov_batch
    rng = np.random.default_rng(12)
    B, n, m, K, reg = 2, 9, 7, 4, 1.0
    C1 = rng.random((B, n, n))
    C2 = rng.random((B, m, m))
    M = rng.random((B, n, m))
    a = rng.dirichlet(np.ones(n), size=B)
    b = rng.dirichlet(np.ones(m), size=B)
    out = solve_gromov_batch(C1, C2, reg=reg, a=a, b=b, M=M, alpha=0.0, max_iter=K, tol=0.0,
                             max_iter_inner=5000, tol_inner=1e-14)
    assert out["n_iter"] == K - 1
    for p in range(B):
        P = osk.sinkhorn_log(a[p], b[p], M[p], reg / K, numItermax=5000, stopThr=1e-14)
        np.testing.assert_allclose(out["T"][p], P, rtol=1e-9, atol=1e-15)


def test_dist_lp_and_kl_batch_pins():
    # Minkowski against scipy's cdist, raised to q; KL against scipy.stats
    # entropy(y, x) = sum y log(y / x) on probability vectors; the logits form
    # on log X equals the plain form up to the eps guard.
    from scipy.spatial.distance import cdist
    from scipy.stats import entropy
    from oracle.batch import dist_kl_batch, dist_lp_batch
    rng = np.random.default_rng(13)
    B, n, m, d = 2, 6, 5, 4
    X = rng.standard_normal((B, n, d))
    Y = rng.standard_normal((B, m, d))
    for p, q in ((1, 1), (2, 1), (3, 1), (1.5, 2.0), (2, 2)):
        got = dist_lp_batch(X, Y, p=p, q=q)
        for k in range(B):
            ref = cdist(X[k], Y[k], metric="minkowski", p=p) ** q
            np.testing.assert_allclose(got[k], ref, rtol=1e-12, err_msg=f"p={p} q={q}")
    Xp = rng.dirichlet(np.ones(d), size=(B, n))
    Yp = rng.dirichlet(np.ones(d), size=(B, m))
    got = dist_kl_batch(Xp, Yp)
    for k in range(B):
        ref = np.array([[entropy(Yp[k, j], Xp[k, i]) for j in range(m)] for i in range(n)])
        np.testing.assert_allclose(got[k], ref, rtol=1e-8, atol=1e-9)
    np.testing.assert_allclose(dist_kl_batch(np.log(Xp), Yp, logits_X=True), got, rtol=1e-8,
                               atol=1e-9)


def test_stabilized_and_eps_scaling_oracle():
    # doctests _sinkhorn.py:1191-1193 and :1458-1460
    from oracle.stabilized import sinkhorn_epsilon_scaling, sinkhorn_stabilized
    want = [[0.36552929, 0.13447071], [0.13447071, 0.36552929]]
    np.testing.assert_allclose(sinkhorn_stabilized([.5, .5], [.5, .5], [[0., 1.], [1., 0.]], 1),
                               want, atol=5e-9)
    np.testing.assert_allclose(
        sinkhorn_epsilon_scaling([.5, .5], [.5, .5], [[0., 1.], [1., 0.]], 1), want, atol=5e-9)
    # stabilized == sinkhorn_log at convergence (test_bregman.py:282-314), with
    # the default tau and with tau = 100; both absorb during the early
    # iterations.  (A tau so small that every iteration absorbs never
    # converges: u, v restart at 1/dim after an absorption.)
    rng = np.random.default_rng(8)
    n, m = 40, 30
    M = rng.random((n, m))
    a, b = rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m))
    P_log = osk.sinkhorn_log(a, b, M, 0.05, stopThr=1e-12, numItermax=5000)
    P1, l1 = sinkhorn_stabilized(a, b, M, 0.05, stopThr=1e-12, numItermax=5000, log=True)
    P2, l2 = sinkhorn_stabilized(a, b, M, 0.05, stopThr=1e-12, numItermax=5000, tau=100.0,
                                 log=True)
    np.testing.assert_allclose(P1, P_log, atol=1e-10)
    np.testing.assert_allclose(P2, P_log, atol=1e-10)
    assert l1["n_iter"] % 20 == 0 and l1["err"][-1] <= 1e-12
    assert np.abs(l2["alpha"]).max() > 0  # absorbed at least once


def test_stabilized_multi_histogram_oracle_fixed_point():
    # the several-histogram branch (_sinkhorn.py:1230-1367) at a tight
    # tolerance: each histogram's loss is the single-histogram loss of its
    # column, with and without absorption; identical columns give identical
    # scalings; a 2x2 doctest problem in two equal columns gives the doctest loss
    from oracle.stabilized import sinkhorn_stabilized_multi
    rng = np.random.default_rng(9)
    n, m, H = 30, 25, 3
    M = rng.random((n, m))
    a = rng.dirichlet(np.ones(n))
    B = np.stack([rng.dirichlet(np.ones(m)) for _ in range(H)], 1)
    want = np.array([np.sum(osk.sinkhorn_log(a, B[:, h], M, 0.05, stopThr=1e-13,
                                             numItermax=20000) * M) for h in range(H)])
    errs = {}
    for tau in (1e3, 30.0):
        loss, lg = sinkhorn_stabilized_multi(a, B, M, 0.05, tau=tau, stopThr=1e-14,
                                             numItermax=20000, log=True)
        np.testing.assert_allclose(loss, want, rtol=1e-9)
        assert lg["logu"].shape == (n, H) and lg["alpha"].shape == (n, H)
        errs[tau] = lg["err"]
    # tau = 30 absorbs (max |u| reaches the hundreds): another path, same end
    assert errs[30.0][1] != errs[1e3][1]
    Bsame = np.stack([B[:, 0], B[:, 0]], 1)
    _, lg = sinkhorn_stabilized_multi(a, Bsame, M, 0.05, stopThr=1e-12, log=True)
    np.testing.assert_array_equal(lg["logu"][:, 0], lg["logu"][:, 1])
    l2 = sinkhorn_stabilized_multi([.5, .5], np.array([[.5, .5], [.5, .5]]),
                                   [[0., 1.], [1., 0.]], 1, stopThr=1e-14)
    np.testing.assert_allclose(l2, [0.26894142] * 2, atol=1e-8)  # 1 / (1 + e)


# ------------------------------------------------------------- greenkhorn

def test_greenkhorn_doctest_value():
    import oracle.greenkhorn as ogk
    # ot/bregman/_sinkhorn.py:995-1001, printed to 8 digits
    G = ogk.greenkhorn([0.5, 0.5], [0.5, 0.5], [[0.0, 1.0], [1.0, 0.0]], 1)
    np.testing.assert_allclose(G, [[0.36552929, 0.13447071], [0.13447071, 0.36552929]],
                               atol=5e-9)


def test_greenkhorn_converges_to_the_sinkhorn_plan():
    # Both solve the same entropic problem: at a tight stop the plans agree.
    import oracle.greenkhorn as ogk
    rng = np.random.default_rng(4)
    x, y = rng.normal(size=(12, 2)), rng.normal(size=(15, 2)) + 0.5
    M = ou.dist(x, y)
    a = rng.dirichlet(np.ones(12))
    b = rng.dirichlet(np.ones(15))
    G, lg = ogk.greenkhorn(a, b, M, 0.5, numItermax=200000, stopThr=1e-13, log=True)
    P = osk.sinkhorn_log(a, b, M, 0.5, numItermax=10000, stopThr=1e-13)
    assert lg["n_iter"] < 200000 - 1
    np.testing.assert_allclose(G, P, atol=1e-12)
    np.testing.assert_allclose(G, lg["u"][:, None] * np.exp(-M / 0.5) * lg["v"][None, :],
                               rtol=1e-12)


def test_greenkhorn_violations_and_greedy_choice():
    # The incrementally kept violations equal the recomputed marginals of u K v
    # minus the weights after every update; each update moves the row or
    # column of largest |violation| (the first index on a tie) and leaves its
    # own marginal equal to its weight.
    import oracle.greenkhorn as ogk
    rng = np.random.default_rng(9)
    M = rng.uniform(0.0, 1.0, size=(7, 5))
    a = rng.dirichlet(np.ones(7))
    b = rng.dirichlet(np.ones(5))
    K = np.exp(-M / 0.3)
    trace = []
    ogk.greenkhorn(a, b, M, 0.3, numItermax=60, stopThr=0.0, warn=False, _trace=trace)
    assert len(trace) == 60
    u_prev = np.full(7, 1 / 7)
    v_prev = np.full(5, 1 / 5)
    for ii, row, idx, u, v, viol, viol_2 in trace:
        # the choice, from the marginals before this update
        G0 = u_prev[:, None] * K * v_prev[None, :]
        r0, c0 = np.abs(G0.sum(1) - a), np.abs(G0.sum(0) - b)
        if row:
            assert idx == int(np.argmax(r0)) and r0.max() > c0.max()
        else:
            assert idx == int(np.argmax(c0)) and r0.max() <= c0.max()
        G = u[:, None] * K * v[None, :]
        np.testing.assert_allclose(viol, G.sum(1) - a, atol=1e-15)
        np.testing.assert_allclose(viol_2, G.sum(0) - b, atol=1e-15)
        if row:
            assert abs(G[idx].sum() - a[idx]) < 1e-15
            np.testing.assert_array_equal(v, v_prev)
        else:
            assert abs(G[:, idx].sum() - b[idx]) < 1e-15
            np.testing.assert_array_equal(u, u_prev)
        u_prev, v_prev = u, v


def test_greenkhorn_stop_after_the_update_and_warnings():
    # The stop value is read before the update, the update still happens, and
    # the loop ends at that iteration: with a huge stopThr exactly one update
    # (n_iter 0), no warning; with numItermax reached, a warning and
    # n_iter = numItermax - 1.
    import oracle.greenkhorn as ogk
    rng = np.random.default_rng(2)
    M = rng.uniform(size=(4, 6))
    a, b = np.full(4, 0.25), np.full(6, 1 / 6)
    trace = []
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        G, lg = ogk.greenkhorn(a, b, M, 1.0, stopThr=10.0, log=True, _trace=trace)
    assert lg["n_iter"] == 0 and len(trace) == 1
    _, _, _, u, v, _, _ = trace[0]
    np.testing.assert_allclose(G, u[:, None] * np.exp(-M) * v[None, :], rtol=1e-15)
    with pytest.warns(UserWarning, match="did not converge"):
        _, lg = ogk.greenkhorn(a, b, M, 1.0, numItermax=3, stopThr=0.0, log=True)
    assert lg["n_iter"] == 2
    # empty weights are uniform (_sinkhorn.py:1034-1037)
    G = ogk.greenkhorn([], [], M, 1.0, stopThr=1e-12, numItermax=100000)
    np.testing.assert_allclose(G.sum(1), a, atol=1e-12)
    np.testing.assert_allclose(G.sum(0), b, atol=1e-12)
