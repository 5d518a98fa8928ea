"""The reference's own hot-path test cases (test/test_bregman.py), run on the
GPU path with the reference's inputs and tolerances.

* test_bregman.py:282-314  every method gives the same plan (atol 1e-5);
* :228-279                 empty a, b default to uniform (atol 1e-5);
* :330-387                 outputs keep the dtype and device of M;
* :1450-1486               a warm start gives the same plan (atol 1e-5); the
                           reference builds it from EMD duals, which are not
                           part of this build: arbitrary dual vectors stand in
                           (any start converges to the same plan);
* :118-127                 sinkhorn_log and sinkhorn_stabilized (tau = 1) agree
                           at reg = 1e-5 (loss atol 1e-6);
* :42-74                   stopThr = 0, numItermax = 1 warns for every method.

greenkhorn joins the cases the reference runs it in (:42-74, :228-314,
:330-387); the warm-start case (:1450-1486) does not list it.
"""

import warnings

import numpy as np
import pytest
import torch

import pot_amd

pytestmark = pytest.mark.gpu

METHODS = ["sinkhorn", "sinkhorn_log", "sinkhorn_stabilized", "sinkhorn_epsilon_scaling"]
ALL_METHODS = METHODS + ["greenkhorn"]


def _gauss_1d(n, m, s):
    # ot.datasets.make_1D_gauss: a normalised Gaussian histogram on 0 .. n-1
    x = np.arange(n, dtype=np.float64)
    h = np.exp(-(x - m) ** 2 / (2 * s ** 2))
    return h / h.sum()


def _dist0(n):
    # ot.utils.dist0(n): squared distances between the points 0 .. n-1
    x = np.arange(n, dtype=np.float64).reshape(-1, 1)
    return (x - x.T) ** 2


def _problem100(seed=0):
    rng = np.random.RandomState(seed)
    x = rng.randn(100, 2)
    u = np.full(100, 1.0 / 100)
    M = pot_amd.dist(x, x)
    return u, M


def test_every_method_gives_the_same_plan():
    u, M = _problem100()
    plans = {}
    for method in ALL_METHODS:
        plans[method] = pot_amd.sinkhorn(u, u, M, 1, method=method, stopThr=1e-10)
        print(method, "done", flush=True)
    ub = torch.tensor(u, device="cuda")
    Mb = torch.tensor(M, device="cuda")
    Gt = pot_amd.sinkhorn(ub, ub, Mb, 1, method="sinkhorn_log", stopThr=1e-10).cpu().numpy()
    G0 = plans["sinkhorn"]
    np.testing.assert_allclose(G0, plans["sinkhorn_log"], atol=1e-5)
    np.testing.assert_allclose(G0, plans["sinkhorn_stabilized"], atol=1e-5)
    np.testing.assert_allclose(G0, plans["sinkhorn_epsilon_scaling"], atol=1e-5)
    np.testing.assert_allclose(G0, plans["greenkhorn"], atol=1e-5)
    np.testing.assert_allclose(G0, Gt, atol=1e-5)


def test_empty_weights_default_to_uniform():
    u, M = _problem100()
    for method in METHODS:
        G, log = pot_amd.sinkhorn([], [], M, 1, stopThr=1e-10, method=method, log=True)
        print(method, "done", flush=True)
        np.testing.assert_allclose(u, G.sum(1), atol=1e-5)
        np.testing.assert_allclose(u, G.sum(0), atol=1e-5)
    # greenkhorn: the reference only runs it (its 1000 single-row updates are
    # not held to the marginals); here the plan must be the uniform one's
    G, log = pot_amd.sinkhorn([], [], M, 1, method="greenkhorn", stopThr=1e-10, log=True)
    Gu = pot_amd.sinkhorn(u, u, M, 1, method="greenkhorn", stopThr=1e-10)
    np.testing.assert_array_equal(G, Gu)
    assert log["n_iter"] == 999


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_outputs_keep_dtype_and_device(dtype):
    rng = np.random.RandomState(42)
    x = rng.randn(100, 2)
    u = np.full(100, 1.0 / 100)
    M = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1)
    ub = torch.tensor(u, dtype=dtype, device="cuda")
    Mb = torch.tensor(M, dtype=dtype, device="cuda")
    for method in ALL_METHODS:
        Gb = pot_amd.sinkhorn(ub, ub, Mb, 1, method=method, stopThr=1e-10)
        assert isinstance(Gb, torch.Tensor) and Gb.dtype == dtype and Gb.device == Mb.device
    for method in ("sinkhorn", "sinkhorn_stabilized", "sinkhorn_log"):
        lb = pot_amd.sinkhorn2(ub, ub, Mb, 1, method=method, stopThr=1e-10)
        assert isinstance(lb, torch.Tensor) and lb.dtype == dtype and lb.device == Mb.device
    print("dtype", dtype, "done", flush=True)
    # NumPy in, NumPy out, same dtype
    G = pot_amd.sinkhorn(u, u, M, 1, method="sinkhorn_log", stopThr=1e-10)
    assert isinstance(G, np.ndarray) and G.dtype == np.float64


def test_warm_start_gives_the_same_plan():
    m, n = 10, 20
    a = np.full(m, 1.0 / m)
    b = np.full(n, 1.0 / n)
    Xs = np.arange(m) * 1.0
    Xt = np.arange(n) * 1.0
    M = (Xs.reshape(-1, 1) - Xt.reshape(1, -1)) ** 2
    rng = np.random.default_rng(3)
    warmstart = (rng.normal(size=m) * 5.0, rng.normal(size=n) * 5.0)
    reg = 1
    pi_unif, _ = pot_amd.sinkhorn(a, b, M, reg, method="sinkhorn", log=True, warmstart=None)
    for method in METHODS:
        pi, _ = pot_amd.sinkhorn(a, b, M, reg, method=method, log=True, warmstart=warmstart)
        print(method, "done", flush=True)
        np.testing.assert_allclose(pi_unif, pi, atol=1e-5)


def test_log_and_stabilized_agree_at_small_reg():
    n = 100
    a1 = _gauss_1d(n, m=30, s=10)
    a2 = _gauss_1d(n, m=40, s=10)
    M = _dist0(n)
    reg = 1e-5
    loss1 = pot_amd.sinkhorn2(a1, a2, M, reg, method="sinkhorn_log")
    loss2 = pot_amd.sinkhorn2(a1, a2, M, reg, tau=1, method="sinkhorn_stabilized")
    np.testing.assert_allclose(loss1, loss2, atol=1e-6)


@pytest.mark.parametrize("method", ALL_METHODS)
def test_convergence_warning_every_method(method):
    n = 100
    a1 = _gauss_1d(n, m=30, s=10)
    a2 = _gauss_1d(n, m=40, s=10)
    M = _dist0(n)
    with pytest.warns(UserWarning):
        pot_amd.sinkhorn(a1, a2, M, 1.0, method=method, stopThr=0, numItermax=1)
    if method in ("sinkhorn", "sinkhorn_stabilized", "sinkhorn_log"):
        with pytest.warns(UserWarning):
            pot_amd.sinkhorn2(a1, a2, M, 1, method=method, stopThr=0, numItermax=1, warn=True)
        with warnings.catch_warnings():
            warnings.simplefilter("error")
            pot_amd.sinkhorn2(a1, a2, M, 1, method=method, stopThr=0, numItermax=1, warn=False)
