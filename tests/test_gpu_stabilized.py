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
