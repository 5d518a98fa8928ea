"""GPU parity of the drop-in solvers against the float64 oracle.

Tolerances (stated per test):
* float64 inputs: the GPU path runs in fp64 end to end; plans agree with the
  oracle to 1e-10 relative on entries >= 1e-30 and iteration counts agree.
* float32 inputs: plan entries >= 1e-30 within 1e-5 relative, cost within
  1e-6 relative, against the oracle run on the same (float32-valued) inputs
  in float64.
"""

import warnings

import numpy as np
import pytest
import torch

import oracle.batch as obatch
import oracle.sinkhorn as osk
import pot_amd
from pot_amd import batch as pbatch
from pot_amd import bregman

pytestmark = pytest.mark.gpu


def _gauss_problem(n, m, seed, d=2, normalize=True, dtype=np.float64):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    y = rng.standard_normal((m, d)) + 1.0
    M = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    if normalize:
        M = M / M.max()
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(m))
    return a, b, M.astype(dtype)


def _rel_plan_err(P, Pref, floor=1e-30):
    mask = Pref >= floor
    return float(np.max(np.abs(P[mask] - Pref[mask]) / Pref[mask]))


def test_closed_form_2x2():
    # Doctest of ot/bregman/_sinkhorn.py:123-129 and :342-347.
    e = np.e
    P_exact = np.array([[e, 1.0], [1.0, e]]) / (2 * (1 + e))
    for method in ("sinkhorn", "sinkhorn_log"):
        P = pot_amd.sinkhorn([0.5, 0.5], [0.5, 0.5], [[0.0, 1.0], [1.0, 0.0]], 1,
                             method=method)
        np.testing.assert_allclose(P, P_exact, rtol=1e-12)
        loss = pot_amd.sinkhorn2([0.5, 0.5], [0.5, 0.5], [[0.0, 1.0], [1.0, 0.0]], 1,
                                 method=method)
        np.testing.assert_allclose(loss, 1.0 / (1.0 + e), rtol=1e-12)


@pytest.mark.parametrize("n,m,reg", [(64, 80, 0.1), (200, 150, 0.05), (33, 257, 0.2)])
def test_sinkhorn_log_f64_matches_oracle(n, m, reg):
    a, b, M = _gauss_problem(n, m, seed=n + m)
    P, lg = bregman.sinkhorn_log(a, b, M, reg, stopThr=1e-9, log=True)
    Pr, lr = osk.sinkhorn_log(a, b, M, reg, stopThr=1e-9, log=True)
    assert lg["niter"] == lr["niter"]
    assert len(lg["err"]) == len(lr["err"])
    np.testing.assert_allclose(lg["err"], lr["err"], rtol=1e-6, atol=1e-15)
    assert _rel_plan_err(P, Pr) < 1e-10
    np.testing.assert_allclose(lg["log_u"], lr["log_u"], rtol=0, atol=1e-9)
    np.testing.assert_allclose(lg["log_v"], lr["log_v"], rtol=0, atol=1e-9)
    assert P.dtype == np.float64


@pytest.mark.parametrize("n,m,reg", [(128, 96, 0.1), (512, 512, 1e-2)])
def test_sinkhorn_log_f32_matches_oracle(n, m, reg):
    a, b, M = _gauss_problem(n, m, seed=7 * n + m, dtype=np.float32)
    it = 300
    P, lg = bregman.sinkhorn_log(a, b, M, reg, numItermax=it, stopThr=0.0, log=True,
                                 warn=False)
    Pr, lr = osk.sinkhorn_log(a, b, M.astype(np.float64), reg, numItermax=it, stopThr=0.0,
                              log=True, warn=False)
    assert P.dtype == np.float32
    assert lg["niter"] == lr["niter"] == it - 1
    # fp32 plan: 1e-5 relative on entries >= 1e-30 (below, fp32 has no normals)
    assert _rel_plan_err(P.astype(np.float64), Pr) < 1e-5
    cost = float(np.sum(M.astype(np.float64) * Pr))
    gcost = float(pot_amd.sinkhorn2(a, b, M, reg, method="sinkhorn_log", numItermax=it,
                                    stopThr=0.0))
    assert abs(gcost - cost) / cost < 1e-6


def test_sinkhorn_log_torch_roundtrip():
    a, b, M = _gauss_problem(50, 60, seed=3)
    Mt = torch.tensor(M, dtype=torch.float64, device="cuda")
    at = torch.tensor(a, dtype=torch.float64, device="cuda")
    bt = torch.tensor(b, dtype=torch.float64, device="cuda")
    P = pot_amd.sinkhorn(at, bt, Mt, 0.1, method="sinkhorn_log")
    assert isinstance(P, torch.Tensor) and P.device.type == "cuda" and P.dtype == torch.float64
    Pr = osk.sinkhorn_log(a, b, M, 0.1)
    assert _rel_plan_err(P.cpu().numpy(), Pr) < 1e-10


@pytest.mark.parametrize("n,m,reg", [(64, 80, 0.1), (100, 100, 0.5)])
def test_sinkhorn_knopp_f64_matches_oracle(n, m, reg):
    a, b, M = _gauss_problem(n, m, seed=11 * n + m)
    P, lg = bregman.sinkhorn_knopp(a, b, M, reg, stopThr=1e-9, log=True)
    Pr, lr = osk.sinkhorn_knopp(a, b, M, reg, stopThr=1e-9, log=True)
    assert lg["niter"] == lr["niter"]
    np.testing.assert_allclose(lg["err"], lr["err"], rtol=1e-6, atol=1e-15)
    assert _rel_plan_err(P, Pr) < 1e-9
    np.testing.assert_allclose(lg["u"], lr["u"], rtol=1e-9)
    np.testing.assert_allclose(lg["v"], lr["v"], rtol=1e-9)


def test_knopp_multi_histogram():
    rng = np.random.RandomState(0)
    n = 10
    x = rng.randn(n, 2)
    u = np.ones(n) / n
    b = rng.rand(n, 3)
    b = b / np.sum(b, 0, keepdims=True)
    M = ((x[:, None] - x[None]) ** 2).sum(-1)
    loss, lg = pot_amd.sinkhorn(u, b, M, 0.1, stopThr=1e-10, log=True)
    lr, lgr = osk.sinkhorn(u, b, M, 0.1, stopThr=1e-10, log=True)
    np.testing.assert_allclose(loss, lr, rtol=1e-9)
    assert lg["niter"] == lgr["niter"]
    for method in ("sinkhorn", "sinkhorn_log"):
        l0 = pot_amd.sinkhorn(u, b, M, 0.1, method=method, stopThr=1e-10)
        per = [pot_amd.sinkhorn2(u, b[:, k], M, 0.1, method=method, stopThr=1e-10)
               for k in range(3)]
        np.testing.assert_allclose(l0, per, atol=1e-4)


def test_convergence_warning_and_nan_guard():
    a, b, M = _gauss_problem(30, 30, seed=1)
    for method in ("sinkhorn", "sinkhorn_log"):
        with pytest.warns(UserWarning):
            pot_amd.sinkhorn(a, b, M, 1.0, method=method, stopThr=0, numItermax=1)
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            pot_amd.sinkhorn2(a, b, M, 1.0, method=method, stopThr=0, numItermax=1)
    # reg = 0 in the scaling form: numerical-error warning (test_bregman.py:104-115)
    with pytest.warns(UserWarning, match="numerical errors"):
        pot_amd.sinkhorn(a, b, M, 0, method="sinkhorn", warn=False)


def test_marginals_reference_test():
    # test_bregman.py:21-39
    n = 100
    rng = np.random.RandomState(0)
    x = rng.randn(n, 2)
    u = np.ones(n) / n
    M = pot_amd.dist(x, x)
    G = pot_amd.sinkhorn(u, u, M, 1, stopThr=1e-10)
    np.testing.assert_allclose(u, G.sum(1), atol=1e-05)
    np.testing.assert_allclose(u, G.sum(0), atol=1e-05)


@pytest.mark.parametrize("solver", ["sinkhorn", "log_sinkhorn"])
@pytest.mark.parametrize("reg_type", ["kl", "entropy"])
def test_solve_batch_matches_oracle(solver, reg_type):
    # test/batch/test_solve_batch.py:26-62 shape, checked against the oracle batch
    rng = np.random.RandomState(0)
    M = rng.rand(4, 16, 16)
    res = pbatch.solve_batch(M, reg=0.1, max_iter=10000, tol=1e-5, solver=solver,
                             reg_type=reg_type, grad="detach")
    ref = obatch.solve_batch(M, 0.1, max_iter=10000, tol=1e-5, solver=solver,
                             reg_type=reg_type)
    assert res.log["n_iter"] == ref["log"]["n_iter"]
    np.testing.assert_allclose(res.plan, ref["plan"], rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(res.value_linear, ref["value_linear"], rtol=1e-9)
    np.testing.assert_allclose(res.value, ref["value"], rtol=1e-9)
    # and each problem equals sinkhorn_log at tol (the reference test's oracle)
    for i in range(4):
        Pi = osk.sinkhorn_log(np.ones(16) / 16, np.ones(16) / 16, M[i], 0.1,
                              numItermax=10000, stopThr=1e-5)
        np.testing.assert_allclose(res.plan[i], Pi, atol=1e-5)


def test_bregman_batch_projections():
    rng = np.random.RandomState(0)
    X = rng.rand(4, 4, 2)
    M = pbatch.dist_batch(X, X)
    Mr = obatch.dist_batch(X, X)
    np.testing.assert_allclose(M, Mr, rtol=1e-12, atol=1e-14)
    K = np.exp(-M / 0.01)
    logK = -M / 0.01
    r1 = pbatch.bregman_projection_batch(K, max_iter=50, tol=1e-10)
    r2 = pbatch.bregman_log_projection_batch(logK, max_iter=50, tol=1e-10)
    o1 = obatch.bregman_projection_batch(K, max_iter=50, tol=1e-10)
    o2 = obatch.bregman_log_projection_batch(logK, max_iter=50, tol=1e-10)
    assert r1["n_iters"] == o1["n_iters"] and r2["n_iters"] == o2["n_iters"]
    np.testing.assert_allclose(r1["T"], o1["T"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(r2["T"], o2["T"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(r2["log_T"], o2["log_T"], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(r1["T"], r2["T"], atol=1e-3)


def test_dist_matches_bruteforce():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((37, 5))
    y = rng.standard_normal((29, 5))
    D = pot_amd.dist(x, y)
    brute = np.array([[np.sum((xi - yj) ** 2) for yj in y] for xi in x])
    np.testing.assert_allclose(D, brute, atol=1e-12)
    Dxx = pot_amd.dist(x)
    assert np.all(np.diag(Dxx) == 0)


def test_reference_cpu_case_fixed_iterations_match_oracle():
    # the reference's CPU-runnable configuration (n = m = 2048, float64, M / max,
    # reg 1e-2) at 30 fixed iterations: exact niter, plan to 1e-10 relative
    from pot_amd import problems
    a, b, M = problems.config1()
    G, lg = pot_amd.sinkhorn(a, b, M, 1e-2, method="sinkhorn_log", stopThr=0.0, numItermax=30,
                             log=True, warn=False)
    print("gpu done", flush=True)
    Go, lo = osk.sinkhorn_log(a, b, M, 1e-2, numItermax=30, stopThr=0.0, log=True, warn=False)
    assert lg["niter"] == lo["niter"]
    assert _rel_plan_err(G, Go) < 1e-10
    np.testing.assert_allclose(lg["err"], lo["err"], rtol=1e-9)


def test_sinkhorn_log_multi_histogram_one_launch():
    # _sinkhorn.py:816-854: each column is its own sinkhorn_log solve; here
    # they run as one batch sharing M, each stopping on its own check
    rng = np.random.default_rng(7)
    n, m, nh = 70, 55, 4
    x = rng.standard_normal((n, 2))
    y = rng.standard_normal((m, 2)) + 0.7
    M = ((x[:, None] - y[None]) ** 2).sum(-1)
    M /= M.max()
    a = rng.dirichlet(np.ones(n))
    b = np.stack([rng.dirichlet(np.full(m, c)) for c in (0.3, 1.0, 3.0, 30.0)], 1)
    REG = 0.01  # the columns stop at 130, 110, 110 and 100 iterations
    loss, lg = pot_amd.sinkhorn(a, b, M, REG, method="sinkhorn_log", stopThr=1e-9,
                                log=True)
    lr, lgr = osk.sinkhorn_log(a, b, M, REG, stopThr=1e-9, log=True)
    iters = [osk.sinkhorn_log(a, b[:, k], M, REG, stopThr=1e-9, log=True)[1]["niter"]
             for k in range(nh)]
    assert len(set(iters)) > 1, iters  # the columns stop at different iterations
    np.testing.assert_allclose(loss, lr, rtol=1e-10)
    for key in ("log_u", "log_v"):
        assert lg[key].shape == lgr[key].shape
        np.testing.assert_allclose(lg[key], lgr[key], atol=1e-9)
    np.testing.assert_allclose(lg["u"], lgr["u"], rtol=1e-8)
    # sinkhorn2 and the torch path return the same losses
    l2 = pot_amd.sinkhorn2(a, b, M, REG, method="sinkhorn_log", stopThr=1e-9)
    np.testing.assert_allclose(l2, lr, rtol=1e-10)
    lt = pot_amd.sinkhorn2(torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda"),
                           torch.tensor(M, device="cuda"), REG, method="sinkhorn_log",
                           stopThr=1e-9)
    assert isinstance(lt, torch.Tensor) and lt.shape == (nh,)
    np.testing.assert_allclose(lt.cpu().numpy(), lr, rtol=1e-10)
    # warm starts per column
    ws = [(lgr["log_u"][:, k], lgr["log_v"][:, k]) for k in range(nh)]
    lw = pot_amd.sinkhorn2(a, b, M, REG, method="sinkhorn_log", stopThr=1e-9, warmstart=ws)
    lwr = osk.sinkhorn2(a, b, M, REG, method="sinkhorn_log", stopThr=1e-9, warmstart=ws)
    np.testing.assert_allclose(lw, lwr, rtol=1e-10)
    # one warning per column that does not converge (the per-column calls warn)
    with warnings.catch_warnings(record=True) as rec:
        warnings.simplefilter("always")
        pot_amd.sinkhorn(a, b, M, REG, method="sinkhorn_log", stopThr=1e-9, numItermax=105)
    with warnings.catch_warnings(record=True) as rec_r:
        warnings.simplefilter("always")
        osk.sinkhorn_log(a, b, M, REG, stopThr=1e-9, numItermax=105)
    assert len(rec) == len(rec_r) == 3  # only the column stopping at 100 converges
    # b of shape (m, 1) without a warm start raises, as in the reference
    with pytest.raises(TypeError):
        pot_amd.sinkhorn(a, b[:, :1], M, REG, method="sinkhorn_log")
