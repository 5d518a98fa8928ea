"""Lazy ``empirical_sinkhorn`` / ``empirical_sinkhorn2`` (online cost) against
the float64 restatement of ot/bregman/_empirical.py:193-270.

float64 points (fp64 pass): same ``niter`` and ``err`` list (rtol 1e-6),
potentials within 1e-9.  float32 points (fp32 MFMA pass): potentials within
2e-4 of the oracle run on the same float32 values, loss within 1e-5.
"""

import numpy as np
import pytest
import torch

import oracle.empirical as oemp
import pot_amd

pytestmark = pytest.mark.gpu


def _samples(ns, nt, d, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((ns, d)), rng.standard_normal((nt, d)) * 0.8 + 0.5


def test_empirical_lazy_f64_matches_oracle():
    xs, xt = _samples(150, 120, 4, seed=1)
    f, g, lg = pot_amd.empirical_sinkhorn(xs, xt, 2.0, isLazy=True, log=True, stopThr=1e-9)
    fr, gr, lr = oemp.empirical_sinkhorn_lazy(xs, xt, 2.0, stopThr=1e-9)
    assert lg["niter"] == lr["niter"]
    np.testing.assert_allclose(lg["err"], lr["err"], rtol=1e-6, atol=1e-13)
    np.testing.assert_allclose(f + g[0], fr + gr[0], atol=1e-9)
    np.testing.assert_allclose(f - f[0], fr - fr[0], atol=1e-9)
    loss = pot_amd.empirical_sinkhorn2(xs, xt, 2.0, isLazy=True, stopThr=1e-9)
    loss_r, _ = oemp.empirical_sinkhorn2_lazy(xs, xt, 2.0, stopThr=1e-9)
    np.testing.assert_allclose(loss, loss_r, rtol=1e-10)


def test_empirical_lazy_f32_and_warmstart():
    xs, xt = _samples(300, 260, 8, seed=2)
    xs32, xt32 = xs.astype(np.float32), xt.astype(np.float32)
    X = torch.tensor(xs32, device="cuda")
    Y = torch.tensor(xt32, device="cuda")
    f, g = pot_amd.empirical_sinkhorn(X, Y, 5.0, isLazy=True, stopThr=1e-7)
    assert f.dtype == torch.float32 and f.is_cuda
    fr, gr, _ = oemp.empirical_sinkhorn_lazy(xs32, xt32, 5.0, stopThr=1e-7)
    fd, gd = f.double().cpu().numpy(), g.double().cpu().numpy()
    np.testing.assert_allclose(fd + gd[0], fr + gr[0], atol=2e-4)
    # warm start from the converged potentials stops at the first check
    _, _, lg = pot_amd.empirical_sinkhorn(xs, xt, 5.0, isLazy=True, log=True, stopThr=1e-7,
                                          warmstart=(fr, gr))
    assert lg["niter"] == 9
    loss = pot_amd.empirical_sinkhorn2(X, Y, 5.0, isLazy=True, stopThr=1e-7)
    loss_r, _ = oemp.empirical_sinkhorn2_lazy(xs32, xt32, 5.0, stopThr=1e-7)
    np.testing.assert_allclose(float(loss), loss_r, rtol=1e-5)


def test_empirical_non_lazy_and_errors():
    xs, xt = _samples(40, 30, 2, seed=3)
    P = pot_amd.empirical_sinkhorn(xs, xt, 1.0)
    M = ((xs[:, None] - xt[None]) ** 2).sum(-1)
    import oracle.sinkhorn as osk
    Pr = osk.sinkhorn(np.full(40, 1 / 40), np.full(30, 1 / 30), M, 1.0)
    np.testing.assert_allclose(P, Pr, rtol=1e-9)
    with pytest.raises(ValueError):
        pot_amd.empirical_sinkhorn(xs, xt, 1.0, isLazy=True, batchSize=(1, 2, 3))
    with pytest.raises(NotImplementedError):
        pot_amd.empirical_sinkhorn(xs, xt, 1.0, isLazy=True, metric="cityblock")


def test_lazy_plan_blocks_match_the_dense_plan():
    # test_bregman.py:1197-1277: log["lazy_plan"][:] equals exp(f + g - M / reg)
    n = 10
    X_s = np.reshape(np.arange(n, dtype=np.float64), (n, 1))
    X_t = np.reshape(np.arange(0, n, dtype=np.float64), (n, 1))
    M = (X_s - X_t.T) ** 2
    f, g, lg = pot_amd.empirical_sinkhorn(X_s, X_t, 1, numIterMax=1000, isLazy=True,
                                          batchSize=5, log=True)
    G_log = np.exp(f[:, None] + g[None, :] - M)
    T = lg["lazy_plan"]
    assert T.shape == (n, n)
    assert repr(T) == "LazyTensor(shape=(10, 10),attributes=(X_a,X_b,f,g,metric,reg))"
    np.testing.assert_allclose(T[:], G_log, rtol=1e-12, atol=1e-300)
    loss, lg2 = pot_amd.empirical_sinkhorn2(X_s, X_t, 1, numIterMax=1000, isLazy=True,
                                            batchSize=5, log=True)
    np.testing.assert_allclose(lg2["lazy_plan"][:], G_log, rtol=1e-12, atol=1e-300)
    np.testing.assert_allclose(loss, np.sum(G_log * M), rtol=1e-10)
    # an integer selects a row or column as a slice of one; slices and strides
    assert T[1].shape == (1, n) and T[1, 1].shape == (1, 1)
    np.testing.assert_allclose(T[1, 1][0, 0], G_log[1, 1], rtol=1e-12)
    np.testing.assert_allclose(T[::3, 5], G_log[::3, 5:6], rtol=1e-12)
    np.testing.assert_allclose(T[2:7, 1:9], G_log[2:7, 1:9], rtol=1e-12)
    with pytest.raises(NotImplementedError):
        T["error"]


def test_lazy_plan_torch_f32_points():
    xs, xt = _samples(300, 260, 8, seed=4)
    X = torch.tensor(xs.astype(np.float32), device="cuda")
    Y = torch.tensor(xt.astype(np.float32), device="cuda")
    f, g, lg = pot_amd.empirical_sinkhorn(X, Y, 2.0, isLazy=True, log=True, stopThr=1e-7)
    P = lg["lazy_plan"][:]
    assert isinstance(P, torch.Tensor) and P.dtype == torch.float32 and P.shape == (300, 260)
    x64, y64 = X.double().cpu().numpy(), Y.double().cpu().numpy()
    M = ((x64[:, None, :] - y64[None]) ** 2).sum(-1)
    Pr = np.exp(f.double().cpu().numpy()[:, None] + g.double().cpu().numpy()[None] - M / 2.0)
    np.testing.assert_allclose(P.cpu().numpy(), Pr, rtol=2e-6, atol=1e-30)
    rows = lg["lazy_plan"][40:45]
    np.testing.assert_allclose(rows.cpu().numpy(), Pr[40:45], rtol=2e-6, atol=1e-30)
