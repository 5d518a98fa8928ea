"""GPU parity of the ragged batched solver against per-problem oracle runs.

Every problem of the batch must behave like a lone ``sinkhorn_log`` call:
the same check iteration at which it stops (allowing one check of slack for
float32 errors that sit on the threshold), and float32-level agreement of
the plan (1e-5 relative on entries >= 1e-30 at a fixed iteration count;
1e-4 relative to the largest entry at convergence).
"""

import numpy as np
import pytest
import torch

import oracle.sinkhorn as osk
from pot_amd import problems, ragged

pytestmark = pytest.mark.gpu


def _oracle_inputs(R, p, normalize=True):
    M = R.cost(p).double().cpu().numpy()
    if normalize:
        M = M / float(R.cmax[p].item())
    return M


def _plan(M, u, v, reg):
    return np.exp(-M / reg + u[:, None] + v[None, :])


def test_ragged_fixed_iterations_match_oracle():
    fam = problems.ragged_family(12, seed=123, n_range=(20, 300), d=5, kmin=1, kmax=3,
                                 reg_range=(0.02, 0.2))
    R = ragged.from_clouds(fam.xs, fam.ys, fam.regs)
    it = 60
    res = ragged.sinkhorn_log_ragged(R, numItermax=it, stopThr=0.0)
    torch.cuda.synchronize()
    assert (res.niter.cpu().numpy() == it - 1).all()
    assert (res.status.cpu().numpy() == 1).all()
    for p in range(R.batch):
        M = _oracle_inputs(R, p)
        n, m = M.shape
        reg = fam.regs[p]
        a = np.full(n, 1.0 / n)
        b = np.full(m, 1.0 / m)
        Pr, lr = osk.sinkhorn_log(a, b, M, reg, numItermax=it, stopThr=0.0, log=True,
                                  warn=False)
        u, v = res.split(R, p)
        P = _plan(M, u.cpu().numpy(), v.cpu().numpy(), reg)
        mask = Pr >= 1e-30
        rel = np.max(np.abs(P[mask] - Pr[mask]) / Pr[mask])
        assert rel < 1e-5, (p, rel)
        np.testing.assert_allclose(res.err_first[p].item(), lr["err"][0], rtol=1e-4)


def test_ragged_converged_matches_oracle():
    fam = problems.ragged_family(16, seed=7, n_range=(64, 257), d=16, kmin=1, kmax=4,
                                 reg_range=(0.02, 0.05))
    R = ragged.from_clouds(fam.xs, fam.ys, fam.regs)
    res = ragged.sinkhorn_log_ragged(R, numItermax=3000, stopThr=1e-6)
    torch.cuda.synchronize()
    niter = res.niter.cpu().numpy()
    for p in range(R.batch):
        M = _oracle_inputs(R, p)
        n, m = M.shape
        reg = fam.regs[p]
        a = np.full(n, 1.0 / n)
        b = np.full(m, 1.0 / m)
        Pr, lr = osk.sinkhorn_log(a, b, M, reg, numItermax=3000, stopThr=1e-6, log=True,
                                  warn=False)
        assert abs(int(niter[p]) - lr["niter"]) <= 10, (p, niter[p], lr["niter"])
        u, v = res.split(R, p)
        P = _plan(M, u.cpu().numpy(), v.cpu().numpy(), reg)
        assert np.max(np.abs(P - Pr)) / Pr.max() < 1e-4
        # the GPU's own marginal error at its stop
        assert np.linalg.norm(P.sum(0) - b) < 2e-6


def test_ragged_warm_start_and_uneven_shapes():
    # m not a multiple of 4 exercises the padded-row masking; a warm start
    # from the converged potentials stops at the first check.
    rng = np.random.default_rng(0)
    Ms = [rng.random((n, m)).astype(np.float32) for n, m in [(13, 7), (64, 131), (200, 1024)]]
    regs = [0.1, 0.05, 0.2]
    R = ragged.from_costs(Ms, regs)
    res = ragged.sinkhorn_log_ragged(R, numItermax=5000, stopThr=1e-6)
    res2 = ragged.sinkhorn_log_ragged(R, u0=res.u, numItermax=5000, stopThr=1e-6)
    torch.cuda.synchronize()
    assert (res2.niter.cpu().numpy() == 0).all()
    for p, M in enumerate(Ms):
        n, m = M.shape
        Pr = osk.sinkhorn_log(np.full(n, 1 / n), np.full(m, 1 / m), M.astype(np.float64),
                              regs[p], numItermax=5000, stopThr=1e-6)
        u, v = res.split(R, p)
        P = _plan(M.astype(np.float64), u.cpu().numpy(), v.cpu().numpy(), regs[p])
        assert np.max(np.abs(P - Pr)) / Pr.max() < 1e-4


def test_ragged_graph_replays_match_eager():
    # One captured graph (ragged distances + batched solve), replayed on three
    # different batches of the same layout, with eager work and syncs between
    # the replays; every replay must equal the eager path bit for bit.
    sizes = problems.ragged_family(10, seed=5, n_range=(40, 400), d=6)
    ns = [len(x) for x in sizes.xs]
    ms = [len(y) for y in sizes.ys]
    G = ragged.RaggedGraph(ns, ms, 6, numItermax=500, stopThr=1e-7, check_every=10)
    for rep in range(3):
        rng = np.random.default_rng(100 + rep)
        xs = [rng.standard_normal((n, 6)).astype(np.float32) for n in ns]
        ys = [(rng.standard_normal((m, 6)) + 0.5).astype(np.float32) for m in ms]
        regs = np.exp(rng.uniform(np.log(0.02), np.log(0.2), size=len(ns)))
        R = ragged.from_clouds(xs, ys, regs)
        ref = ragged.sinkhorn_log_ragged(R, numItermax=500, stopThr=1e-7, check_every=10)
        torch.cuda.synchronize()
        ref_u, ref_v = ref.u.cpu().numpy(), ref.v.cpu().numpy()
        ref_n = ref.niter.cpu().numpy()
        out = G.solve(xs, ys, regs)
        torch.cuda.synchronize()
        assert np.array_equal(out.niter.cpu().numpy(), ref_n)
        assert np.array_equal(out.u.cpu().numpy(), ref_u)
        assert np.array_equal(out.v.cpu().numpy(), ref_v)
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
