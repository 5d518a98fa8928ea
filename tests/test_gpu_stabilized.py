"""``sinkhorn_stabilized`` / ``sinkhorn_epsilon_scaling`` against the float64
restatement of ot/bregman/_sinkhorn.py:1105-1566.

float64 inputs: same ``n_iter`` and ``err`` list (rtol 1e-6), plan within
1e-10 relative on entries >= 1e-30, ``alpha``/``beta`` within 1e-9.
float32 inputs: plan within 1e-5 relative against the oracle on the same
float32 values.
"""

import warnings

import numpy as np
import pytest
import torch

import oracle.stabilized as ostab
import pot_amd
from pot_amd import bregman

pytestmark = pytest.mark.gpu


def _problem(n, m, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 2))
    y = rng.standard_normal((m, 2)) + 1.0
    M = ((x[:, None] - y[None]) ** 2).sum(-1)
    M /= M.max()
    return rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m)), M


def _rel(P, Pr, floor=1e-30):
    mask = Pr >= floor
    return float(np.max(np.abs(P[mask] - Pr[mask]) / Pr[mask]))


def test_stabilized_doctest_and_dispatch():
    want = [[0.36552929, 0.13447071], [0.13447071, 0.36552929]]
    P = pot_amd.sinkhorn([.5, .5], [.5, .5], [[0., 1.], [1., 0.]], 1, method="sinkhorn_stabilized")
    np.testing.assert_allclose(P, want, atol=5e-9)
    P = pot_amd.sinkhorn([.5, .5], [.5, .5], [[0., 1.], [1., 0.]], 1,
                         method="sinkhorn_epsilon_scaling")
    np.testing.assert_allclose(P, want, atol=5e-9)
    loss = pot_amd.sinkhorn2([.5, .5], [.5, .5], [[0., 1.], [1., 0.]], 1,
                             method="sinkhorn_stabilized")
    np.testing.assert_allclose(loss, 1.0 / (1.0 + np.e), rtol=1e-8)


@pytest.mark.parametrize("reg,tau", [(0.05, 1e3), (0.01, 1e3), (0.05, 100.0)])
def test_stabilized_f64_matches_oracle(reg, tau):
    a, b, M = _problem(70, 50, seed=3)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        P, lg = bregman.sinkhorn_stabilized(a, b, M, reg, tau=tau, numItermax=3000, log=True)
        Pr, lr = ostab.sinkhorn_stabilized(a, b, M, reg, tau=tau, numItermax=3000, log=True)
    assert lg["n_iter"] == lr["n_iter"]
    np.testing.assert_allclose(lg["err"], lr["err"], rtol=1e-6, atol=1e-15)
    assert _rel(P, Pr) < 1e-10
    np.testing.assert_allclose(lg["alpha"], lr["alpha"], atol=1e-9)
    np.testing.assert_allclose(lg["beta"], lr["beta"], atol=1e-9)


def test_stabilized_f32_and_eps_scaling():
    a, b, M = _problem(120, 90, seed=4)
    M32 = M.astype(np.float32)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        P = bregman.sinkhorn_stabilized(torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda"),
                                        torch.tensor(M32, device="cuda"), 0.05, numItermax=2000)
        Pr = ostab.sinkhorn_stabilized(a, b, M32.astype(np.float64), 0.05, numItermax=2000)
    assert P.dtype == torch.float32
    assert _rel(P.double().cpu().numpy(), Pr) < 1e-5
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        G, lg = bregman.sinkhorn_epsilon_scaling(a, b, M, 0.05, log=True)
        Gr, lr = ostab.sinkhorn_epsilon_scaling(a, b, M, 0.05, log=True)
    assert lg["niter"] == lr["niter"]
    np.testing.assert_allclose(lg["err"], lr["err"], rtol=1e-6, atol=1e-15)
    assert _rel(G, Gr) < 1e-9


@pytest.mark.parametrize("tau", [1e3, 30.0])
def test_stabilized_multi_histogram_matches_oracle(tau):
    # the several-histogram branch (_sinkhorn.py:1230-1367) on the device:
    # same n_iter, err list, losses and logs as the float64 restatement
    from oracle.stabilized import sinkhorn_stabilized_multi
    rng = np.random.default_rng(11)
    n, m, H = 60, 45, 3
    M = rng.random((n, m))
    a = rng.dirichlet(np.ones(n))
    B = np.stack([rng.dirichlet(np.ones(m)) for _ in range(H)], 1)
    loss, lg = pot_amd.sinkhorn(a, B, M, 0.02, method="sinkhorn_stabilized", tau=tau,
                                stopThr=1e-9, numItermax=3000, log=True)
    lr, lgr = sinkhorn_stabilized_multi(a, B, M, 0.02, tau=tau, stopThr=1e-9, numItermax=3000,
                                        log=True)
    assert lg["n_iter"] == lgr["n_iter"]
    np.testing.assert_allclose(lg["err"], lgr["err"], rtol=1e-8, atol=1e-15)
    np.testing.assert_allclose(loss, lr, rtol=1e-10)
    for key in ("logu", "logv", "alpha", "beta"):
        assert lg[key].shape == lgr[key].shape, key
        np.testing.assert_allclose(lg[key], lgr[key], rtol=1e-9, atol=1e-9)
    if tau == 30.0:  # absorbed: the error path differs from the run without absorption
        _, l_free = sinkhorn_stabilized_multi(a, B, M, 0.02, tau=1e3, stopThr=1e-9,
                                              numItermax=3000, log=True)
        assert l_free["err"][1] != lgr["err"][1]
    l2 = pot_amd.sinkhorn2(a, B, M, 0.02, method="sinkhorn_stabilized", tau=tau, stopThr=1e-9,
                           numItermax=3000)
    np.testing.assert_allclose(l2, lr, rtol=1e-10)
