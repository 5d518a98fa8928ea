"""GPU greedy Sinkhorn (``pot_amd.bregman.greenkhorn``, one launch running
ot/bregman/_sinkhorn.py:1058-1086) against the float64 oracle restatement.

Both run in float64; only the order of the dot products' sums differs, so
the greedy choices coincide and the iteration counts agree except where two
violations tie to the last bit (the bars below allow none in these cases).
"""

import warnings

import numpy as np
import pytest
import torch

import oracle.greenkhorn as ogk
import oracle.utils as ou
import pot_amd
from pot_amd import bregman

pytestmark = pytest.mark.gpu


def _problem(seed, n, m, d=2):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    y = rng.normal(size=(m, d)) + 0.7
    M = ou.dist(x, y)
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(m))
    return a, b, M


def test_greenkhorn_doctest():
    # ot/bregman/_sinkhorn.py:995-1001
    G = bregman.greenkhorn([0.5, 0.5], [0.5, 0.5], [[0.0, 1.0], [1.0, 0.0]], 1)
    np.testing.assert_allclose(G, [[0.36552929, 0.13447071], [0.13447071, 0.36552929]],
                               atol=5e-9)


# (up to 2048 points a side the loop runs on 512 threads, up to 4096 on 1024
# with its state in LDS; the last case keeps its state in HBM)
@pytest.mark.parametrize("n,m,reg,it", [(30, 40, 0.5, 20000), (64, 17, 0.2, 3000),
                                        (200, 150, 1.0, 2000), (2048, 33, 0.5, 600),
                                        (33, 2500, 0.5, 600), (4200, 9, 0.5, 400)])
def test_greenkhorn_matches_oracle(n, m, reg, it):
    a, b, M = _problem(n + m, n, m)
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        G, lg = bregman.greenkhorn(a, b, M, reg, numItermax=it, stopThr=1e-9, log=True)
        Gr, lr = ogk.greenkhorn(a, b, M, reg, numItermax=it, stopThr=1e-9, log=True)
    print(f"{n}x{m}: n_iter {lg['n_iter']} (oracle {lr['n_iter']})", flush=True)
    assert lg["n_iter"] == lr["n_iter"]
    np.testing.assert_allclose(G, Gr, rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(lg["u"], lr["u"], rtol=1e-10)
    np.testing.assert_allclose(lg["v"], lr["v"], rtol=1e-10)


def test_greenkhorn_converged_marginals_warm_start_and_dtype():
    a, b, M = _problem(7, 50, 60)
    reg = 0.5
    with warnings.catch_warnings():
        warnings.simplefilter("error")  # converges: no warning
        G, lg = bregman.greenkhorn(a, b, M, reg, numItermax=200000, stopThr=1e-10, log=True)
    assert lg["n_iter"] < 200000 - 1
    np.testing.assert_allclose(G.sum(1), a, atol=1e-10)
    np.testing.assert_allclose(G.sum(0), b, atol=1e-10)
    # a warm start from the converged scalings stops within a few updates (the
    # update taken at the stop leaves a violation just above stopThr), same
    # plan
    ws = (np.log(lg["u"]), np.log(lg["v"]))
    Gw, lw = bregman.greenkhorn(a, b, M, reg, numItermax=100, stopThr=1e-10, log=True,
                                warmstart=ws)
    assert lw["n_iter"] < 20
    np.testing.assert_allclose(Gw, G, atol=1e-9)
    # the oracle from the same warm start
    Gr, lr = ogk.greenkhorn(a, b, M, reg, numItermax=100, stopThr=1e-10, log=True,
                            warmstart=ws)
    assert lw["n_iter"] == lr["n_iter"]
    np.testing.assert_allclose(Gw, Gr, rtol=1e-10, atol=1e-16)
    # torch float32 in, float32 out on the device (computed in float64)
    at = torch.tensor(a, dtype=torch.float32, device="cuda")
    bt = torch.tensor(b, dtype=torch.float32, device="cuda")
    Mt = torch.tensor(M, dtype=torch.float32, device="cuda")
    Gt = pot_amd.sinkhorn(at, bt, Mt, reg, method="greenkhorn", numItermax=200000,
                          stopThr=1e-7)
    assert Gt.dtype == torch.float32 and Gt.device == Mt.device
    np.testing.assert_allclose(Gt.cpu().numpy(), G, atol=1e-6)


def test_greenkhorn_edge_cases():
    a, b, M = _problem(3, 8, 5)
    # numItermax <= 0: the start plan, a warning, and no log (as the reference)
    with pytest.warns(UserWarning, match="did not converge"):
        G0 = bregman.greenkhorn(a, b, M, 1.0, numItermax=0)
    K = np.exp(-M / 1.0)
    np.testing.assert_allclose(G0, (np.full(8, 1 / 8)[:, None] * K) * np.full(5, 1 / 5)[None, :],
                               rtol=1e-14)
    with pytest.raises(UnboundLocalError):
        with warnings.catch_warnings():
            warnings.simplefilter("ignore")
            bregman.greenkhorn(a, b, M, 1.0, numItermax=0, log=True)
    # one update per iteration; a huge stopThr stops after the first
    G1, l1 = bregman.greenkhorn(a, b, M, 1.0, stopThr=10.0, log=True)
    Gr, lr = ogk.greenkhorn(a, b, M, 1.0, stopThr=10.0, log=True)
    assert l1["n_iter"] == 0 == lr["n_iter"]
    np.testing.assert_allclose(G1, Gr, rtol=1e-13)
    # several histograms are refused; sinkhorn2 does not route greenkhorn
    with pytest.raises(ValueError):
        bregman.greenkhorn(a, np.stack([b, b], 1), M, 1.0)
    with pytest.raises(ValueError):
        pot_amd.sinkhorn2(a, b, M, 1.0, method="greenkhorn")
