"""GPU parity of the point-cloud batch solver (cost computed on the fly).

Each problem must behave like a lone ``sinkhorn_log`` call on
``ot.dist(x, y) / M.max()`` evaluated in float64 from the same float32
points: the plan within 1e-5 relative (entries >= 1e-30) and the cost within
1e-6 at a fixed iteration count and at convergence (the stop at the
oracle's check; one check apart only when the oracle's error there is within
1% of stopThr), and the cost maximum that the kernel finds equal to the
float64 maximum.
"""

import numpy as np
import pytest
import torch

import oracle.sinkhorn as osk
from oracle.utils import dist
from pot_amd import problems, ragged

pytestmark = pytest.mark.gpu


def _cost(x, y, normalize=True):
    M = dist(np.asarray(x, dtype=np.float32).astype(np.float64),
             np.asarray(y, dtype=np.float32).astype(np.float64))
    return M / M.max() if normalize else M


def _plan(M, u, v, reg):
    return np.exp(-M / reg + u[:, None] + v[None, :])


@pytest.mark.parametrize("waves", [8, 16])
def test_points_fixed_iterations_match_oracle(waves):
    # Uneven n, m (not multiples of 32), d < 16.
    fam = problems.ragged_family(12, seed=321, n_range=(20, 600), d=5, kmin=1, kmax=3,
                                 reg_range=(0.02, 0.2))
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    it = 60
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=it, stopThr=0.0, waves=waves)
    torch.cuda.synchronize()
    print("fixed-iteration batch done", flush=True)
    assert (res.niter.cpu().numpy() == it - 1).all()
    assert (res.status.cpu().numpy() == 1).all()
    cmax = res.cmax.cpu().numpy()
    for p in range(P.batch):
        Mraw = _cost(fam.xs[p], fam.ys[p], normalize=False)
        assert abs(cmax[p] - Mraw.max()) <= 1e-12 * Mraw.max(), (p, cmax[p], Mraw.max())
        M = Mraw / Mraw.max()
        n, m = M.shape
        reg = fam.regs[p]
        a = np.full(n, 1.0 / n)
        b = np.full(m, 1.0 / m)
        Pr, lr = osk.sinkhorn_log(a, b, M, reg, numItermax=it, stopThr=0.0, log=True,
                                  warn=False)
        u, v = res.split(P, p)
        Pl = _plan(M, u.cpu().numpy(), v.cpu().numpy(), reg)
        mask = Pr >= 1e-30
        rel = np.max(np.abs(Pl[mask] - Pr[mask]) / Pr[mask])
        assert rel < 1e-5, (p, rel)
        cost_rel = abs(np.sum(Pl * M) - np.sum(Pr * M)) / np.sum(Pr * M)
        assert cost_rel < 1e-6, (p, cost_rel)
        np.testing.assert_allclose(res.err_first[p].item(), lr["err"][0], rtol=1e-4)


def test_points_converged_matches_oracle_small_reg():
    # The warm-start family's regime: GMM clouds, d = 16, reg down to 5e-3.
    fam = problems.ragged_family(6, seed=77, n_range=(64, 320), d=16, kmin=1, kmax=8,
                                 reg_range=(5e-3, 5e-2))
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=3000, stopThr=1e-6)
    torch.cuda.synchronize()
    print("converged batch done", flush=True)
    niter = res.niter.cpu().numpy()
    assert ((res.status.cpu().numpy() & 3) == 0).all()
    for p in range(P.batch):
        M = _cost(fam.xs[p], fam.ys[p])
        n, m = M.shape
        reg = fam.regs[p]
        a = np.full(n, 1.0 / n)
        b = np.full(m, 1.0 / m)
        Pr, lr = osk.sinkhorn_log(a, b, M, reg, numItermax=3000, stopThr=1e-6, log=True,
                                  warn=False)
        u, v = res.split(P, p)
        Pl = _plan(M, u.cpu().numpy(), v.cpu().numpy(), reg)
        assert np.linalg.norm(Pl.sum(0) - b) < 1.1e-6
        if int(niter[p]) == lr["niter"]:
            # the same stop: the fp32 bar on the plan and the cost
            mask = Pr >= 1e-30
            rel = np.max(np.abs(Pl[mask] - Pr[mask]) / Pr[mask])
            cost_rel = abs(np.sum(Pl * M) - np.sum(Pr * M)) / np.sum(Pr * M)
            assert rel < 1e-5 and cost_rel < 1e-6, (p, rel, cost_rel)
        else:
            # one check apart only when the oracle's error at the earlier
            # stop is within 1% of stopThr
            assert abs(int(niter[p]) - lr["niter"]) == 10, (p, niter[p], lr["niter"])
            e = lr["err"][min(int(niter[p]), lr["niter"]) // 10]
            assert abs(e - 1e-6) < 1e-8, (p, e)


def test_points_full_size_problem_matches_oracle():
    fam = problems.config5(seed=5, count=1)
    x, y, reg = fam.xs[0], fam.ys[0], float(fam.regs[0])
    P = ragged.points_from_clouds([x], [y], [reg])
    it = 30
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=it, stopThr=0.0)
    torch.cuda.synchronize()
    print("1024 x 1024 problem done", flush=True)
    M = _cost(x, y)
    n, m = M.shape
    Pr = osk.sinkhorn_log(np.full(n, 1 / n), np.full(m, 1 / m), M, reg, numItermax=it,
                          stopThr=0.0, warn=False)
    Pl = _plan(M, res.u.cpu().numpy(), res.v.cpu().numpy(), reg)
    mask = Pr >= 1e-30
    rel = np.max(np.abs(Pl[mask] - Pr[mask]) / Pr[mask])
    assert rel < 1e-5, rel


@pytest.mark.parametrize("waves", [8, 16])
def test_points_agree_with_materialised_batch(waves):
    fam = problems.ragged_family(24, seed=11, n_range=(64, 1024), d=16, kmin=1, kmax=4,
                                 reg_range=(5e-3, 5e-2))
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    R = ragged.from_clouds(fam.xs, fam.ys, fam.regs)
    rp = ragged.sinkhorn_log_ragged_points(P, numItermax=2000, stopThr=1e-6, check_every=1,
                                           waves=waves)
    rd = ragged.sinkhorn_log_ragged(R, numItermax=2000, stopThr=1e-6, check_every=1)
    torch.cuda.synchronize()
    print("points vs materialised done", flush=True)
    assert ((rp.status.cpu().numpy() & 3) == 0).all()
    assert ((rd.status.cpu().numpy() & 3) == 0).all()
    for p in range(P.batch):
        M = _cost(fam.xs[p], fam.ys[p])
        reg = fam.regs[p]
        up, vp = rp.split(P, p)
        ud, vd = rd.split(R, p)
        Pp = _plan(M, up.cpu().numpy(), vp.cpu().numpy(), reg)
        Pd = _plan(M, ud.cpu().numpy(), vd.cpu().numpy(), reg)
        assert np.max(np.abs(Pp - Pd)) / Pd.max() < 1e-4, p


def test_points_warm_start_fallback_and_edge_cases():
    rng = np.random.default_rng(3)
    shapes = [(1, 7, 3), (13, 1, 3), (40, 33, 1), (200, 1024, 16), (1024, 65, 8)]
    xs = [rng.standard_normal((n, d)).astype(np.float32) for n, m, d in shapes]
    ys = [(rng.standard_normal((m, d)) + 0.3).astype(np.float32) for n, m, d in shapes]
    # One d for the whole batch: pad the clouds with zero coordinates.
    d = max(s[2] for s in shapes)
    xs = [np.pad(x, ((0, 0), (0, d - x.shape[1]))) for x in xs]
    ys = [np.pad(y, ((0, 0), (0, d - y.shape[1]))) for y in ys]
    regs = [0.1, 0.05, 0.2, 0.02, 0.05]
    a = [np.full(n, 1.0 / n) for n, _, _ in shapes]
    b = [np.full(m, 1.0 / m) for _, m, _ in shapes]
    # zero weights in one problem
    a[3] = a[3].copy()
    a[3][:10] = 0.0
    a[3] /= a[3].sum()
    b[3] = b[3].copy()
    b[3][5:9] = 0.0
    b[3] /= b[3].sum()
    P = ragged.points_from_clouds(xs, ys, regs, a=a, b=b)
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=5000, stopThr=1e-6)
    torch.cuda.synchronize()
    print("edge-case batch done", flush=True)
    assert ((res.status.cpu().numpy() & 3) == 0).all()
    niter = res.niter.cpu().numpy()
    for p in range(P.batch):
        M = _cost(xs[p], ys[p])
        Pr, lr = osk.sinkhorn_log(a[p], b[p], M, regs[p], numItermax=5000, stopThr=1e-6,
                                  warn=False, log=True)
        u, v = res.split(P, p)
        Pl = _plan(M, u.cpu().numpy(), v.cpu().numpy(), regs[p])
        assert np.all(Pl[Pr == 0] == 0), p  # zero-weight rows and columns
        if int(niter[p]) != lr["niter"]:
            # one check apart only when the oracle's error at the earlier
            # stop is within 1% of stopThr
            assert abs(int(niter[p]) - lr["niter"]) == 10, (p, niter[p], lr["niter"])
            e = lr["err"][min(int(niter[p]), lr["niter"]) // 10]
            assert abs(e - 1e-6) < 1e-8, (p, e)
            continue
        # the oracle's stop: the fp32 bar (plan 1e-5 relative on entries >=
        # 1e-30, cost 1e-6)
        mask = Pr >= 1e-30
        rel = np.max(np.abs(Pl[mask] - Pr[mask]) / Pr[mask])
        cost_rel = abs(np.sum(Pl * M) - np.sum(Pr * M)) / np.sum(Pr * M)
        assert rel < 1e-5 and cost_rel < 1e-6, (p, rel, cost_rel)
    u_p, _ = res.split(P, 3)
    assert np.isneginf(u_p.cpu().numpy()[:10]).all()
    # Warm start from the converged potentials: stops at the first check.
    res2 = ragged.sinkhorn_log_ragged_points(P, u0=res.u, numItermax=5000, stopThr=1e-6)
    # A wild warm start with the fallback enabled restarts cold.
    wild = torch.as_tensor(20.0 * rng.standard_normal(res.u.numel()), device=res.u.device)
    res3 = ragged.sinkhorn_log_ragged_points(P, u0=wild, numItermax=5000, stopThr=1e-6,
                                             fallback_thr=1e-2)
    torch.cuda.synchronize()
    print("warm-start batches done", flush=True)
    assert (res2.niter.cpu().numpy() == 0).all()
    st3 = res3.status.cpu().numpy()
    assert ((st3 & 3) == 0).all()
    # With a single row or column the first sweep is exact whatever u0 is, so
    # only problems 2-4 can reject the wild start; after the restart they
    # repeat the cold solve exactly.
    assert ((st3[2:] & 4) > 0).all()
    np.testing.assert_array_equal(res3.niter.cpu().numpy()[2:], res.niter.cpu().numpy()[2:])
    for p in (2, 3, 4):
        np.testing.assert_array_equal(res3.split(P, p)[0].cpu().numpy(),
                                      res.split(P, p)[0].cpu().numpy())


def test_points_unnormalised_and_out_of_range_reg():
    # reg = 1e-8 on an unnormalised cost puts the scaled coordinates outside
    # the kernel's f16 operand range: with unsolved="nan" the kernel reports
    # status 2 and NaN potentials; by default that problem is solved again on
    # the float64 dense solver, as the reference solves it.
    rng = np.random.default_rng(4)
    xs = [rng.standard_normal((50, 4)).astype(np.float32) for _ in range(3)]
    ys = [(rng.standard_normal((60, 4)) + 1.0).astype(np.float32) for _ in range(3)]
    regs = [2.0, 1e-8, 5.0]
    P = ragged.points_from_clouds(xs, ys, regs, normalize=False)
    raw = ragged.sinkhorn_log_ragged_points(P, numItermax=3000, stopThr=1e-7, unsolved="nan")
    torch.cuda.synchronize()
    st = raw.status.cpu().numpy()
    assert st[1] == 2 and (st[[0, 2]] == 0).all()
    u1, _ = raw.split(P, 1)
    assert np.isnan(u1.cpu().numpy()).all()
    with pytest.raises(ValueError):
        ragged.sinkhorn_log_ragged_points(P, unsolved="skip")
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=3000, stopThr=1e-7)
    torch.cuda.synchronize()
    print("unnormalised batch done", flush=True)
    np.testing.assert_array_equal(res.cmax.cpu().numpy(), [1.0, 1.0, 1.0])
    for p in (0, 1, 2):
        M = _cost(xs[p], ys[p], normalize=False)
        Pr, lr = osk.sinkhorn_log(np.full(50, 1 / 50), np.full(60, 1 / 60), M, regs[p],
                                  numItermax=3000, stopThr=1e-7, warn=False, log=True)
        u, v = res.split(P, p)
        Pl = _plan(M, u.cpu().numpy(), v.cpu().numpy(), regs[p])
        assert np.max(np.abs(Pl - Pr)) / Pr.max() < (1e-4 if p != 1 else 1e-6), p
        if p == 1:
            # the float64 dense solve: the oracle's loop, stop and errors
            assert int(res.niter[1]) == lr["niter"]
            assert int(res.status[1]) == (0 if lr["err"][-1] < 1e-7 else 1)
            np.testing.assert_allclose(float(res.err_last[1]), lr["err"][-1], rtol=1e-6)
        else:
            assert int(res.status[p]) == 0


def test_points_graph_replays_match_eager():
    sizes = problems.ragged_family(10, seed=5, n_range=(40, 400), d=6)
    ns = [len(x) for x in sizes.xs]
    ms = [len(y) for y in sizes.ys]
    G = ragged.RaggedGraph(ns, ms, 6, numItermax=500, stopThr=1e-7, check_every=10,
                           cost="points")
    for rep in range(3):
        rng = np.random.default_rng(200 + rep)
        xs = [rng.standard_normal((n, 6)).astype(np.float32) for n in ns]
        ys = [(rng.standard_normal((m, 6)) + 0.5).astype(np.float32) for m in ms]
        regs = np.exp(rng.uniform(np.log(0.02), np.log(0.2), size=len(ns)))
        P = ragged.points_from_clouds(xs, ys, regs)
        ref = ragged.sinkhorn_log_ragged_points(P, numItermax=500, stopThr=1e-7,
                                                check_every=10)
        torch.cuda.synchronize()
        ref_u, ref_v = ref.u.cpu().numpy(), ref.v.cpu().numpy()
        ref_n = ref.niter.cpu().numpy()
        out = G.solve(xs, ys, regs)
        torch.cuda.synchronize()
        print("graph replay", rep, "done", flush=True)
        assert np.array_equal(out.niter.cpu().numpy(), ref_n)
        assert np.array_equal(out.u.cpu().numpy(), ref_u)
        assert np.array_equal(out.v.cpu().numpy(), ref_v)
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays


def test_batch_autotuner_picks_a_compiled_config(tmp_path, monkeypatch):
    from pot_amd import autotune
    monkeypatch.setattr(autotune, "CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_mem", None)
    fam = problems.ragged_family(12, seed=21, n_range=(64, 300), d=8)
    ns = [len(x) for x in fam.xs]
    ms = [len(y) for y in fam.ys]
    for cost in ("points", "dense"):
        print("tuning", cost, flush=True)
        cfg, _ = autotune.tune_batch(ns, ms, d=8, cost=cost, iters=5, reps=1)
        assert tuple(cfg) in autotune.BATCH_CONFIGS[cost]
        assert autotune.lookup_batch(cost, ns, ms, 8 if cost == "points" else 0) == tuple(cfg)
    # a solve that leaves the configuration at 0 runs the tuned one; its plans
    # agree with the library default's
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    rt = ragged.sinkhorn_log_ragged_points(P, numItermax=300, stopThr=1e-7, check_every=10)
    r8 = ragged.sinkhorn_log_ragged_points(P, numItermax=300, stopThr=1e-7, check_every=10,
                                           waves=8)
    torch.cuda.synchronize()
    print("tuned and default solves done", flush=True)
    for p in range(P.batch):
        M = _cost(fam.xs[p], fam.ys[p])
        ut, vt = rt.split(P, p)
        u8, v8 = r8.split(P, p)
        Pt = _plan(M, ut.cpu().numpy(), vt.cpu().numpy(), fam.regs[p])
        P8 = _plan(M, u8.cpu().numpy(), v8.cpu().numpy(), fam.regs[p])
        assert np.max(np.abs(Pt - P8)) / P8.max() < 1e-5, p


def test_split_problems_fixed_iterations_match_oracle():
    # Problems shared by 2, 4 and 8 workgroups (uneven n, m, d < 16), against
    # the float64 oracle at a fixed iteration count; the cost maxima are
    # combined across the parts.
    fam = problems.ragged_family(6, seed=77, n_range=(130, 700), d=5, kmin=1, kmax=3,
                                 reg_range=(0.02, 0.2))
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    ns = np.array([len(x) for x in fam.xs])
    parts = np.where(ns >= 128, 4, 1)
    parts[::3] = np.where(ns[::3] >= 64, 2, 1)
    parts[4:] = np.where(ns[4:] >= 256, 8, parts[4:])  # 487 and 694 rows: 8 parts
    assert (parts == 8).sum() == 2 and (parts > 1).sum() >= 4
    it = 40
    res = ragged.sinkhorn_log_ragged_points(P, numItermax=it, stopThr=0.0, split=parts)
    torch.cuda.synchronize()
    print("split fixed-iteration batch done", flush=True)
    assert (res.niter.cpu().numpy() == it - 1).all()
    assert (res.status.cpu().numpy() == 1).all()
    cmax = res.cmax.cpu().numpy()
    for p in range(P.batch):
        Mraw = _cost(fam.xs[p], fam.ys[p], normalize=False)
        assert abs(cmax[p] - Mraw.max()) <= 1e-12 * Mraw.max(), (p, cmax[p], Mraw.max())
        M = Mraw / Mraw.max()
        n, m = M.shape
        a = np.full(n, 1.0 / n)
        b = np.full(m, 1.0 / m)
        Pr, lr = osk.sinkhorn_log(a, b, M, fam.regs[p], numItermax=it, stopThr=0.0, log=True,
                                  warn=False)
        u, v = res.split(P, p)
        Pl = _plan(M, u.cpu().numpy(), v.cpu().numpy(), fam.regs[p])
        mask = Pr >= 1e-30
        rel = np.max(np.abs(Pl[mask] - Pr[mask]) / Pr[mask])
        assert rel < 1e-5, (p, parts[p], rel)
        np.testing.assert_allclose(res.err_first[p].item(), lr["err"][0], rtol=1e-4)


def test_split_converged_warm_fallback_and_graph():
    # "auto" split of a small uneven batch: converged plans agree with the
    # one-workgroup-per-problem solve; a rejected warm start restarts cold in
    # every part; a captured graph replays the split launch exactly.
    fam = problems.ragged_family(16, seed=12, n_range=(64, 1024), d=16, kmin=1, kmax=4,
                                 reg_range=(1e-2, 5e-2))
    P = ragged.points_from_clouds(fam.xs, fam.ys, fam.regs)
    assert (ragged.split_plan(P.ns, P.ms) > 1).sum() >= 4
    r1 = ragged.sinkhorn_log_ragged_points(P, numItermax=2000, stopThr=1e-6, check_every=1)
    rs = ragged.sinkhorn_log_ragged_points(P, numItermax=2000, stopThr=1e-6, check_every=1,
                                           split="auto")
    torch.cuda.synchronize()
    print("split converged batch done", flush=True)
    assert ((rs.status.cpu().numpy() & 3) == 0).all()
    for p in range(P.batch):
        M = _cost(fam.xs[p], fam.ys[p])
        u1, v1 = r1.split(P, p)
        us, vs = rs.split(P, p)
        P1 = _plan(M, u1.cpu().numpy(), v1.cpu().numpy(), fam.regs[p])
        Ps = _plan(M, us.cpu().numpy(), vs.cpu().numpy(), fam.regs[p])
        assert np.max(np.abs(Ps - P1)) / P1.max() < 1e-4, p
    # warm start far off: every problem falls back to the cold start
    rng = np.random.default_rng(5)
    u0 = torch.tensor(rng.normal(size=int(P.ns.sum())) * 50.0, device="cuda")
    rw = ragged.sinkhorn_log_ragged_points(P, u0=u0, numItermax=2000, stopThr=1e-6,
                                           check_every=1, fallback_thr=1e-3, split="auto")
    torch.cuda.synchronize()
    print("split fallback batch done", flush=True)
    st = rw.status.cpu().numpy()
    assert ((st & 3) == 0).all() and ((st & 4) > 0).all()
    np.testing.assert_array_equal(rw.niter.cpu().numpy(), rs.niter.cpu().numpy())
    ns = [len(x) for x in fam.xs]
    ms = [len(y) for y in fam.ys]
    G = ragged.RaggedGraph(ns, ms, 16, numItermax=2000, stopThr=1e-6, check_every=1,
                           cost="points", split="auto")
    for rep in range(2):
        out = G.solve(P.X, P.Y, fam.regs)
        torch.cuda.synchronize()
        print("split graph replay", rep, "done", flush=True)
        assert np.array_equal(out.niter.cpu().numpy(), rs.niter.cpu().numpy())
        assert np.array_equal(out.u.cpu().numpy(), rs.u.cpu().numpy())
        assert np.array_equal(out.v.cpu().numpy(), rs.v.cpu().numpy())
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays


def test_points_batch_tables():
    """One upload carries offsets, sizes, regs and the launch order; a scalar
    reg applies to every problem and a reg count that does not match the
    batch is rejected (the kernel reads one reg per problem)."""
    rng = np.random.default_rng(5)
    xs = [rng.standard_normal((n, 3)).astype(np.float32) for n in (40, 70, 33)]
    ys = [rng.standard_normal((m, 3)).astype(np.float32) for m in (50, 20, 64)]
    P = ragged.points_from_clouds(xs, ys, [0.1, 0.05, 0.2])
    np.testing.assert_array_equal(P.aoff.cpu().numpy(), [0, 40, 110])
    np.testing.assert_array_equal(P.boff.cpu().numpy(), [0, 50, 70])
    np.testing.assert_array_equal(P.ns_d.cpu().numpy(), [40, 70, 33])
    np.testing.assert_array_equal(P.ms_d.cpu().numpy(), [50, 20, 64])
    np.testing.assert_array_equal(P.regs.cpu().numpy(), [0.1, 0.05, 0.2])
    # longest first by n m / reg: 40*50/0.1 = 2e4, 70*20/0.05 = 2.8e4, 33*64/0.2 = 1.056e4
    np.testing.assert_array_equal(P.order.cpu().numpy(), [1, 0, 2])
    assert P.ns_d.dtype == torch.int32 and P.order.dtype == torch.int32
    np.testing.assert_allclose(P.a.cpu().numpy()[:40], 1 / 40)
    np.testing.assert_allclose(P.a.cpu().numpy()[40:110], 1 / 70)
    Ps = ragged.points_from_clouds(xs, ys, 0.1)
    np.testing.assert_array_equal(Ps.regs.cpu().numpy(), [0.1, 0.1, 0.1])
    with pytest.raises(ValueError):
        ragged.points_from_clouds(xs, ys, [0.1, 0.2])
    # equal sizes: uniform weights from one fill, the same values
    Pe = ragged.points_from_clouds(xs[:1] * 2, ys[:1] * 2, 0.1)
    np.testing.assert_array_equal(Pe.a.cpu().numpy(), np.full(80, 1.0 / 40))
    np.testing.assert_array_equal(Pe.b.cpu().numpy(), np.full(100, 1.0 / 50))
    # (B, n, d) tensors: the same tables, each in its own type
    Xt = torch.tensor(np.stack([xs[0], xs[0]]), device="cuda")
    Yt = torch.tensor(np.stack([ys[0], ys[0], ys[0]])[:2], device="cuda")
    Pt = ragged.points_from_clouds(Xt, Yt, [0.2, 0.1])
    np.testing.assert_array_equal(Pt.aoff.cpu().numpy(), [0, 40])
    np.testing.assert_array_equal(Pt.boff.cpu().numpy(), [0, 50])
    np.testing.assert_array_equal(Pt.ns_d.cpu().numpy(), [40, 40])
    np.testing.assert_array_equal(Pt.ms_d.cpu().numpy(), [50, 50])
    np.testing.assert_array_equal(Pt.regs.cpu().numpy(), [0.2, 0.1])
    np.testing.assert_array_equal(Pt.order.cpu().numpy(), [1, 0])
    assert Pt.aoff.dtype == torch.int64 and Pt.regs.dtype == torch.float64
    assert Pt.ms_d.dtype == torch.int32 and Pt.order.dtype == torch.int32


def test_points_cost_maximum_shapes_and_ties():
    """The cost maximum found on the device equals the float64 maximum of
    ``ot.dist`` (to 1e-12 relative) for sizes from 1 point to 1024, every
    column-tile count, d from 1 to 16, and clouds whose largest distance is
    attained by many tied pairs (grid points, duplicated points), through the
    separate maximum launch and through the solves with 8 and 16 waves."""
    rng = np.random.default_rng(11)
    shapes = [(1, 1, 3), (1, 1024, 16), (1024, 1, 16), (33, 31, 1), (64, 257, 7),
              (300, 520, 16), (1024, 1024, 16), (97, 800, 2), (512, 96, 11), (1000, 1000, 16)]
    xs, ys = [], []
    for n, m, d in shapes:
        xs.append(rng.standard_normal((n, d)).astype(np.float32))
        ys.append((rng.standard_normal((m, d)) * 0.5 + 1.0).astype(np.float32))
    # ties: integer grid points, so many pairs share the largest distance
    g = np.stack(np.meshgrid(*[np.arange(4.0)] * 3), -1).reshape(-1, 3).astype(np.float32)
    xs.append(g)
    ys.append(g[::-1].copy())
    # ties: one cloud is a single point repeated, the other two points repeated
    xs.append(np.repeat(np.ones((1, 16), np.float32), 700, axis=0))
    ys.append(np.repeat(np.array([[0.0] * 16, [3.0] * 16], np.float32), 450, axis=0))
    want = np.array([_cost(x, y, normalize=False).max() for x, y in zip(xs, ys)])
    for d in sorted({x.shape[1] for x in xs}):
        idx = [i for i, x in enumerate(xs) if x.shape[1] == d]
        P = ragged.points_from_clouds([xs[i] for i in idx], [ys[i] for i in idx], 0.1)
        got = ragged.points_cmax(P).cpu().numpy()
        torch.cuda.synchronize()
        for k, i in enumerate(idx):
            assert abs(got[k] - want[i]) <= 1e-12 * want[i], (shapes + ["grid", "rep"])[i]
        for waves in (8, 16):
            res = ragged.sinkhorn_log_ragged_points(P, numItermax=2, stopThr=0.0, waves=waves,
                                                    split=None)
            cm = res.cmax.cpu().numpy()
            for k, i in enumerate(idx):
                assert abs(cm[k] - want[i]) <= 1e-12 * want[i], (waves, i)
    print("cost maxima match on", len(xs), "problems", flush=True)
