"""Benchmark of the entropic Sinkhorn hot path on one MI355X.

Default workload (the headline metric): the learned-warm-start family --
n = m = 1024 point clouds in d = 16 (Gaussian mixtures of 1-8 components),
reg ~ logU[5e-3, 5e-2], uniform weights, cost = squared distance / max.  A
step solves one batch of fresh problems to marginal error <= 1e-6
(sinkhorn_log semantics, numItermax 2000): the warm-start network when its
weights are present, then the batched solver, all timed with the inputs
already resident in HBM.  ``--cost points`` (default) runs the point-cloud
solver, which evaluates the cost from the points on MFMA in every sweep;
``--cost dense`` materialises the normalised costs and streams them.

Other workloads (``--workload``): ``c1`` (the reference's CPU case, n = m =
2048 in float64 through the drop-in ``sinkhorn``), ``dense`` (one 16384 x 16384 problem,
d = 64, 200 fixed iterations, materialised 1 GiB cost), ``ragged`` (512
uneven problems, n_i, m_i in [64, 1024]), ``online`` (the 16384 x 16384,
d = 64 problem with the cost computed on the fly, 200 fixed iterations, one
``--precision`` of the x . y^T term), ``precision`` (n = m = 65536, d = 128,
reg = 5e-2, 100 iterations on the online cost in every precision against the
full-size fp64 GPU run, and on a 4096-point subsample against the float64
NumPy oracle), ``copy`` (1 GiB device copy, the HBM reference),
``greenkhorn`` (the greedy Sinkhorn at n = m = 2048: 10000 updates, reg 0.05).

Prints one JSON line; only N = 1 is supported.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "problems/sec to marginal err≤1e-6 (learned warm-start vs plain GPU Sinkhorn)"
HBM_PEAK_GBS = 8000.0
# HBM bytes per iteration of the 16384^2 fp32 dense solve, from rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE per dispatch (KiB) of the iteration's three kernels:
# fetch 540584.37 (row sweep) + 8429.00 (merge) + 2.66 (decide), write
# 16911.19 + 136.00 + 0.04, each scaled by the counters' reading on a 1 GiB
# device copy (FETCH_SIZE 524315 KiB: x 2.000; WRITE_SIZE 1048576 KiB: x 1.000).
DENSE_C2_TRAFFIC = (540584.37 + 8429.00 + 2.66) * 1024 * 2.000 \
    + (16911.19 + 136.00 + 0.04) * 1024 * 1.000
# HBM bytes of one cold launch of the point-cloud solver on the 4096-problem
# family (rocprofv3 --pmc, the same scaling): fetch 853253.25 KiB x 2.000,
# write 1136088.28 KiB x 1.000.  The points are read three times (centring,
# operand packing for the cost-maximum search, packing for the solve); the
# writes are mostly the once-per-problem register saves of the out-of-line
# per-problem function.  Against 0.70 GB of algorithmic bytes (points, weights,
# start and potentials once), at about 94 GB/s: far from bounding the kernel.
POINTS_COLD_TRAFFIC = 853253.25 * 1024 * 2.000 + 1136088.28 * 1024 * 1.000
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f16": 2500.0, "bf16": 2500.0, "bf16x3": 2500.0, "fp8": 5000.0, "f64": 78.6}


def _args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="warmstart", choices=["warmstart", "dense", "c1", "ragged", "copy", "greenkhorn",
                             "online", "precision"])
    ap.add_argument("--precision", default="bf16x3",
                    choices=["f32", "bf16", "bf16x3", "fp8", "f64"])
    ap.add_argument("--subsample", type=int, default=4096)
    ap.add_argument("--graph", type=int, default=1, help="ragged: replay a captured hipGraph")
    ap.add_argument("--batch", type=int, default=4096, help="problems per step")
    ap.add_argument("--waves", type=int, default=0, help="wavefronts per problem (0 = default)")
    ap.add_argument("--prefetch", type=int, default=0, help="rows prefetched per wave (0 = 1)")
    ap.add_argument("--rotate", type=int, default=0, help="batch kernel: rotate row order")
    ap.add_argument("--occ", type=int, default=0,
                    help="batch kernel register cap in waves/SIMD (0 default, 4, 255 none)")
    ap.add_argument("--check-every", type=int, default=1,
                    help="iterations between marginal checks (1: stop at the first iterate "
                         "with err <= stopThr; 10: the reference's cadence)")
    ap.add_argument("--stop", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=2000)
    ap.add_argument("--cost", default="points", choices=["dense", "points"],
                    help="warmstart / ragged: cost computed on the fly from the points "
                         "(points) or materialised in HBM (dense)")
    ap.add_argument("--split", type=int, default=1,
                    help="ragged: share large problems among workgroups when the batch is small")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cold-only", action="store_true")
    ap.add_argument("--no-c2", action="store_true",
                    help="warmstart: skip the 16384^2 dense and 512-problem ragged figures "
                         "(profiling runs)")
    return ap.parse_args()


def cpu_baseline_problem(x, y, reg, stop, max_iter, budget_s=30.0):
    """Time the float64 NumPy oracle (sinkhorn_log) on one problem."""
    from oracle.sinkhorn import sinkhorn_log
    from pot_amd.problems import cost_host
    M = cost_host(x, y)
    n, m = M.shape
    t0 = time.perf_counter()
    _, lg = sinkhorn_log(np.full(n, 1.0 / n), np.full(m, 1.0 / m), M, reg,
                         numItermax=max_iter, stopThr=stop, log=True, warn=False)
    dt = time.perf_counter() - t0
    return dt, lg["niter"]


def _cpu_worker_init():
    # one thread per worker process: the pool supplies the parallelism
    from threadpoolctl import threadpool_limits
    global _POOL_LIMITS
    _POOL_LIMITS = threadpool_limits(1)


def _cpu_worker(job):
    if job is None:  # start-up task: imports done before the timed map
        import oracle.sinkhorn  # noqa: F401
        import scipy.special  # noqa: F401
        return 0.0, 0
    x, y, reg, stop, max_iter = job
    return cpu_baseline_problem(x, y, reg, stop, max_iter)


def host_info():
    """CPU model, core count and thread settings of the host running the CPU leg."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    import scipy
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "usable_cpus": usable,
            "threads_env": {k: os.environ.get(k) for k in
                            ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")},
            "numpy": np.__version__, "scipy": scipy.__version__}


def cpu_baseline_family(fam, stop, max_iter, single=24, pooled=96):
    """The float64 NumPy sinkhorn_log oracle on the first problems of the
    evaluation set: ``single`` problems in this process (one core), then
    ``pooled`` problems over a process pool of one worker per usable host CPU
    (at most 16), one problem per task."""
    import multiprocessing as mp
    info = host_info()
    tot, its = 0.0, 0
    for q in range(single):
        dt, nit = cpu_baseline_problem(fam.xs[q], fam.ys[q], fam.regs[q], stop, max_iter)
        tot += dt
        its += nit + 1
    workers = max(1, min(16, info["usable_cpus"] or 1))
    jobs = [(fam.xs[q], fam.ys[q], fam.regs[q], stop, max_iter) for q in range(pooled)]
    # fresh interpreters: the workers never touch the GPU this process holds
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers, initializer=_cpu_worker_init) as pool:
        pool.map(_cpu_worker, [None] * workers, chunksize=1)
        t0 = time.perf_counter()
        res = pool.map(_cpu_worker, jobs, chunksize=1)
        wall = time.perf_counter() - t0
    return {"value": pooled / wall, "unit": "problems/s", "cores": workers, "kind": "port",
            "sample": f"problems 0-{pooled - 1} of the evaluation set, float64 NumPy "
                      f"sinkhorn_log (stopThr {stop}), one problem per task over a pool of "
                      f"{workers} single-threaded processes (started before the timing), "
                      f"{sum(r[1] + 1 for r in res)} iterations in {wall:.1f} s wall",
            "single_core": {"value": single / tot, "unit": "problems/s", "cores": 1,
                            "sample": f"problems 0-{single - 1}, {its} iterations in "
                                      f"{tot:.1f} s"},
            "host": info}


def dense_c2_figure(torch, dev, steps=2):
    """One 16384 x 16384 problem (d = 64, Dirichlet weights, reg 1e-2) on its
    materialised fp32 cost (1 GiB): 200 fixed sinkhorn_log iterations, timed
    with events on the solve's stream.  Algorithmic bytes: one read of the
    cost per iteration (the column update is fused into the row sweep), plus
    the prologue's column pass."""
    from pot_amd import _lib, _solve, problems
    x, y, a, b = problems.config2_clouds()
    n = x.shape[0]
    X = torch.tensor(x, device=dev)
    Y = torch.tensor(y, device=dev)
    M = _solve.sqeuclidean(X[None], Y[None], clamp=True)[0]
    M /= M.max()
    at = torch.tensor(a, device=dev).view(1, n)
    bt = torch.tensor(b, device=dev).view(1, n)
    z = torch.zeros(1, n, dtype=torch.float64, device=dev)
    its = 200

    def solve():
        return _solve.dense_solve(M, at, bt, z, z, 1e-2, domain=_lib.LOG,
                                  order=_lib.COL_FIRST, max_iter=its, stop_thr=0.0)

    solve()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        solve()
    e1.record()
    torch.cuda.synchronize()
    dt = e0.elapsed_time(e1) * 1e-3 / steps
    alg = float(n) * n * 4 * (its + 1)
    gbs = alg / dt / 1e9
    del M
    return {"workload": "dense n=m=16384 d=64 Dirichlet weights reg=1e-2, 200 fixed "
                        "iterations, materialised fp32 cost (1 GiB)",
            "iterations_per_s": its / dt, "ms_per_iteration": 1e3 * dt / its,
            "two_pass_equivalent_gbs": 2 * gbs,
            "roofline": {"bound": "hbm", "kernel": "row_sweep_kernel + col_merge_kernel "
                         "(whole iteration)", "achieved": gbs, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                         "alg_bytes_per_iteration": float(n) * n * 4,
                         "traffic": DENSE_C2_TRAFFIC,
                         "traffic_over_alg": DENSE_C2_TRAFFIC / (float(n) * n * 4)}}


def ragged_c3_figure(torch, dev, steps=5, warmup=2, check=4, cpu=True):
    """The uneven batch: 512 problems, n_i, m_i ~ U[64, 1024], d = 16, GMM(1-4),
    reg_i ~ logU[5e-3, 5e-2], stopThr 1e-6 (checked every 10 iterations, the
    reference's cadence), numItermax 2000, solved as one captured point-cloud
    hipGraph with the automatic split; replays timed with events on the
    solve's stream.  With ``cpu`` (the CPU-baseline leg) the float64 oracle
    solves the ``check`` largest problems on one core, timed, and the GPU
    solution of each is compared with it at its own stop (plan relative
    error on entries >= 1e-30, cost relative error)."""
    from pot_amd import problems, ragged
    fam = problems.config3(seed=2, B=512)
    ns = np.array([len(x) for x in fam.xs])
    ms = np.array([len(y) for y in fam.ys])
    Xp = torch.tensor(np.concatenate(fam.xs), device=dev)
    Yp = torch.tensor(np.concatenate(fam.ys), device=dev)
    G = ragged.RaggedGraph(ns, ms, 16, numItermax=2000, stopThr=1e-6, check_every=10,
                           cost="points", split="auto")
    for _ in range(warmup):
        G.solve(Xp, Yp, fam.regs)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        r = G.solve(Xp, Yp, fam.regs)
    e1.record()
    torch.cuda.synchronize()
    dt = e0.elapsed_time(e1) * 1e-3 / steps
    nit = r.niter.cpu().numpy().astype(np.int64)
    st = r.status.cpu().numpy()
    elems = float(np.sum((nit + 1) * ns * ms))
    tf = 64.0 * elems / dt / 1e12
    out = {"workload": "ragged B=512 n_i,m_i~U[64,1024] d=16 GMM(1-4) reg~logU[5e-3,5e-2] "
                       "stopThr=1e-6, one captured point-cloud hipGraph, split auto",
           "problems_per_s": 512 / dt, "ms_per_step": 1e3 * dt,
           "iters": {"mean": float(nit.mean()), "max": int(nit.max())},
           "converged_frac": float(((st & 3) == 0).mean()),
           "roofline": {"bound": "mfma", "kernel": "batch_points_kernel (split, graph replay)",
                        "achieved": tf, "peak": MFMA_PEAK_TFLOPS["f16"], "unit": "TFLOP/s",
                        "frac": tf / MFMA_PEAK_TFLOPS["f16"], "traffic": None,
                        "note": "64 algorithmic flops per element-iteration over the replay"}}
    if not cpu:
        return out
    from oracle.sinkhorn import sinkhorn_log
    u_all, v_all = r.u.cpu().numpy(), r.v.cpu().numpy()
    ao = np.concatenate([[0], np.cumsum(ns)[:-1]])
    bo = np.concatenate([[0], np.cumsum(ms)[:-1]])
    worst_rel, worst_cost, exact, t_cpu = 0.0, 0.0, 0, 0.0
    for q in np.argsort(-(ns * ms))[:check]:
        x64 = fam.xs[q].astype(np.float64)
        y64 = fam.ys[q].astype(np.float64)
        M = np.maximum((x64 ** 2).sum(1)[:, None] + (y64 ** 2).sum(1)[None]
                       - 2 * x64 @ y64.T, 0.0)
        M /= M.max()
        a = np.full(ns[q], 1.0 / ns[q])
        b = np.full(ms[q], 1.0 / ms[q])
        t0 = time.perf_counter()
        Pr, lr = sinkhorn_log(a, b, M, fam.regs[q], numItermax=2000, stopThr=1e-6, log=True,
                              warn=False)
        t_cpu += time.perf_counter() - t0
        if lr["niter"] != int(nit[q]):
            continue
        exact += 1
        P = np.exp(-M / fam.regs[q] + u_all[ao[q]:ao[q] + ns[q]][:, None]
                   + v_all[bo[q]:bo[q] + ms[q]][None, :])
        mask = Pr >= 1e-30
        worst_rel = max(worst_rel, float(np.max(np.abs(P[mask] - Pr[mask]) / Pr[mask])))
        worst_cost = max(worst_cost, abs(float(np.sum(P * M) - np.sum(Pr * M)))
                         / float(np.sum(Pr * M)))
    out["cpu_baseline"] = {"value": check / t_cpu, "unit": "problems/s", "cores": 1,
                           "kind": "port",
                           "sample": f"the {check} largest problems, float64 NumPy "
                                     f"sinkhorn_log, one core, {t_cpu:.1f} s"}
    out["vs_oracle"] = {"problems": check, "same_stop": exact, "max_plan_rel_err": worst_rel,
                        "max_cost_rel_err": worst_cost}
    return out


def run_warmstart(args, torch, dev):
    from pot_amd import problems, ragged
    from pot_amd import warmstart as ws_mod

    B = args.batch
    nsteps = args.warmup + args.steps
    fam = problems.config5(seed=1_000_000, count=B)
    X = torch.tensor(np.stack(fam.xs), device=dev)
    Y = torch.tensor(np.stack(fam.ys), device=dev)
    regs = fam.regs
    net = None if args.cold_only else ws_mod.load_default(dev)

    def measure(warm):
        t_total = 0.0
        kern_ms = 0.0
        alg_bytes = 0.0
        elems = 0.0        # elements of the fused sweeps (one exp2 each)
        sweeps_tot = 0.0   # elements of every pass over the cost
        iters, conv, fb = [], [], []
        for k in range(nsteps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            fb_thr = ws_mod.FALLBACK_THR if warm else 0.0
            if args.cost == "points":
                R = ragged.points_from_clouds(X, Y, regs)
                cm = ragged.points_cmax(R) if warm else None
                u0 = ws_mod.predict_u0(net, R, X, Y, regs) if warm else None
                e0.record()
                # unsolved="nan": no status read-back inside the timed region; a
                # problem the kernel could not take counts as not converged
                res = ragged.sinkhorn_log_ragged_points(R, u0=u0, numItermax=args.max_iter,
                                                        stopThr=args.stop,
                                                        check_every=args.check_every,
                                                        fallback_thr=fb_thr, cmax=cm,
                                                        waves=args.waves, unsolved="nan")
            else:
                R = ragged.from_clouds(X, Y, regs)
                u0 = ws_mod.predict_u0(net, R, X, Y, regs) if warm else None
                e0.record()
                res = ragged.sinkhorn_log_ragged(R, u0=u0, numItermax=args.max_iter,
                                                 stopThr=args.stop, waves=args.waves,
                                                 prefetch=args.prefetch,
                                                 check_every=args.check_every,
                                                 occupancy=args.occ, rotate=args.rotate,
                                                 fallback_thr=fb_thr)
            e1.record()
            torch.cuda.synchronize()
            if k < args.warmup:
                continue
            t_total += time.perf_counter() - t0
            kern_ms += e0.elapsed_time(e1)
            nit = res.niter.cpu().numpy().astype(np.int64)
            st = res.status.cpu().numpy()
            sweeps = nit + 2  # prologue + (niter + 1) fused sweeps
            if warm:
                sweeps = sweeps + (st & 4 > 0) * 2  # a rejected warm start cost 2 sweeps
            if args.cost == "points":
                # over the points: the cost-maximum pass (cold) and a two-pass prologue
                sweeps = sweeps + (0 if warm else 1) + 1
            alg_bytes += float(np.sum(sweeps * R.ns.astype(np.int64) * R.ms * 4))
            elems += float(np.sum((nit + 1) * R.ns.astype(np.int64) * R.ms))
            sweeps_tot += float(np.sum(sweeps * R.ns.astype(np.int64) * R.ms))
            iters.append(nit)
            conv.append((st & 3) == 0)
            fb.append((st & 4) > 0)
        iters = np.concatenate(iters)
        conv = np.concatenate(conv)
        nprob = args.steps * B
        return {"problems_per_s": nprob / t_total, "ms_per_step": 1e3 * t_total / args.steps,
                "kernel_ms": kern_ms / args.steps, "alg_bytes": alg_bytes / args.steps,
                "sweep_elements": elems / args.steps, "pass_elements": sweeps_tot / args.steps,
                "iters": {"mean": float(iters.mean()), "p50": float(np.percentile(iters, 50)),
                          "p90": float(np.percentile(iters, 90)), "max": int(iters.max())},
                "converged_frac": float(conv.mean()),
                "fallback_frac": float(np.concatenate(fb).mean()), "problems": nprob}

    cold = measure(False)
    warm = measure(True) if net is not None else None
    head = warm or cold
    # The roofline describes the cold launch of batch_log_kernel: the kernel
    # is the same for both starts, and its HBM traffic was measured with
    # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this exact (seeded,
    # deterministic) cold workload: 3.3735e8 KiB fetched x 2.000 (the counter
    # scale of a 1 GiB device copy) + 6.610e4 KiB written x 1.000.
    gbs = cold["alg_bytes"] / (cold["kernel_ms"] * 1e-3) / 1e9
    traffic = 3.3735e8 * 1024 * 2.000 + 6.610e4 * 1024 * 1.000 \
        if (B == 4096 and args.stop == 1e-6 and args.max_iter == 2000) else None
    if args.cost == "points":
        # Algorithmic work of the point-cloud solve: every iteration forms
        # x . y^T twice (once per half-step), 2 d flops per element each, so
        # 4 d = 64 flops per element-iteration at d = 16, counted over
        # (niter + 1) sweeps of every problem (the last, fused check sweep
        # included).  The kernel executes more on MFMA: 4 products of
        # 32x32x16 per element of every pass (f16 hi.lo, lo.hi, hi.hi and the
        # bf16 potential slot), 128 flops, over niter + 4 passes.
        secs = cold["kernel_ms"] * 1e-3
        d = 16
        alg_flops = 4.0 * d * cold["sweep_elements"]
        tf = alg_flops / secs / 1e12
        tf_exec = 128.0 * cold["pass_elements"] / secs / 1e12
        # The binding unit is the VALU: per element of a fused sweep one exp2,
        # one add (row sum) and one fma (column sum), measured on this GPU
        # model at 8.2 + 2.6 + 2.7 cycles per wave64 instruction per SIMD
        # (transcendentals do not co-issue with other VALU work), i.e. at most
        # 4 SIMDs x 64 / 14.6 elements per clock per CU.
        valu_peak = 256 * 4 * 64 / 14.6 * 2.4e9
        roof = {"bound": "mfma", "kernel": "batch_points_kernel (cold launch)",
                "achieved": tf, "peak": MFMA_PEAK_TFLOPS["f16"], "unit": "TFLOP/s",
                "frac": tf / MFMA_PEAK_TFLOPS["f16"], "traffic": POINTS_COLD_TRAFFIC,
                "alg_flops_per_launch": alg_flops,
                "flops_per_element_iteration": 4 * d,
                "avg_launch_ms": cold["kernel_ms"],
                "executed_mfma": {"flops_per_launch": 128.0 * cold["pass_elements"],
                                  "achieved": tf_exec,
                                  "frac": tf_exec / MFMA_PEAK_TFLOPS["f16"]},
                "valu_bound": {"achieved": cold["sweep_elements"] / secs,
                               "peak": valu_peak, "unit": "elements/s",
                               "frac": cold["sweep_elements"] / secs / valu_peak},
                "note": "achieved = algorithmic x.y^T flops (4 d per element-iteration) over "
                        "the cold launch; the points stay in LDS and registers, so HBM "
                        "carries only points, weights, potentials and register saves "
                        "(traffic: measured bytes per launch, about 0.09 TB/s); the kernel "
                        "is bound by the VALU (exp2 + add + fma per element), see valu_bound"}
    else:
        roof = {"bound": "hbm", "kernel": "batch_log_kernel (cold launch)",
                "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": traffic,
                "alg_bytes_per_launch": cold["alg_bytes"],
                "avg_launch_ms": cold["kernel_ms"]}
    out = {
        "value": head["problems_per_s"],
        "unit": "problems/s",
        "ms_per_step": head["ms_per_step"],
        "dtype": "f32",
        "config": {"workload": "warmstart-family n=m=1024 d=16 GMM(1-8) reg~logU[5e-3,5e-2]",
                   "problems_per_step": B, "stopThr": args.stop, "numItermax": args.max_iter,
                   "start": "learned" if warm else "cold (network weights absent)",
                   "check_every": args.check_every, "cost": args.cost,
                   "seed_family": 1_000_000},
        "cold": cold,
        "warm": warm,
        "speedup_warm_vs_cold": (warm["problems_per_s"] / cold["problems_per_s"]) if warm else None,
        "roofline": roof,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_family(fam, args.stop, args.max_iter)
    if not args.no_c2:
        out["dense_c2"] = dense_c2_figure(torch, dev, steps=2)
        out["ragged_c3"] = ragged_c3_figure(torch, dev, cpu=not args.no_cpu)
    return out


def run_dense(args, torch, dev):
    from pot_amd import _lib, _solve, problems
    x, y, a, b = problems.config2_clouds()
    n = x.shape[0]
    X = torch.tensor(x, device=dev)
    Y = torch.tensor(y, device=dev)
    M = _solve.sqeuclidean(X[None], Y[None], clamp=True)[0]
    M /= M.max()
    at = torch.tensor(a, device=dev).view(1, n)
    bt = torch.tensor(b, device=dev).view(1, n)
    z = torch.zeros(1, n, dtype=torch.float64, device=dev)
    its = 200

    def solve():
        return _solve.dense_solve(M, at, bt, z, z, 1e-2, domain=_lib.LOG,
                                  order=_lib.COL_FIRST, max_iter=its, stop_thr=0.0)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    bytes_it = float(n) * n * 4
    gbs = bytes_it * its / dt / 1e9
    out = {"value": its / dt, "unit": "iterations/s", "ms_per_step": dt * 1e3, "dtype": "f32",
           "config": {"workload": "dense n=m=16384 d=64 reg=1e-2 200 fixed iterations",
                      "cost_bytes": bytes_it},
           "roofline": {"bound": "hbm", "kernel": "row_sweep_kernel (+col_merge)",
                        "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                        "note": "whole-iteration time, one cost sweep per iteration"}}
    if not args.no_cpu:
        from scipy.special import logsumexp
        Mh = (M[:2048].double().cpu().numpy())
        t0 = time.perf_counter()
        u = np.zeros(Mh.shape[0])
        for _ in range(3):
            v = np.log(b) - logsumexp(-Mh / 1e-2 + u[:, None], 0)
            u = np.log(a[:2048]) - logsumexp(-Mh / 1e-2 + v[None, :], 1)
        t = (time.perf_counter() - t0) / 3 * (n / 2048)
        out["cpu_baseline"] = {"value": 1.0 / t, "unit": "iterations/s", "cores": 1,
                               "kind": "port", "sample": "3 float64 sinkhorn_log iterations "
                               "on a 2048 x 16384 row slice, scaled to 16384 rows"}
    return out


def run_c1(args, torch, dev):
    """The reference's CPU-runnable case, through the drop-in on the GPU:
    ot.sinkhorn(method="sinkhorn_log") in float64, n = m = 2048,
    x ~ N(0, I2), y ~ N((2, 2), [[1, .5], [.5, 1]]), uniform weights,
    M = dist / max, reg = 1e-2, stopThr = 1e-9, numItermax = 1000.  A step is
    one whole call (plan and log included) on inputs resident in HBM."""
    import pot_amd
    from pot_amd import problems
    a, b, M = problems.config1()
    n = M.shape[0]
    at, bt, Mt = (torch.tensor(v, device=dev) for v in (a, b, M))

    def solve():
        return pot_amd.sinkhorn(at, bt, Mt, 1e-2, method="sinkhorn_log", stopThr=1e-9,
                                numItermax=1000, log=True, warn=False)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        G, lg = solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    niter = int(lg["niter"])
    # one fp64 sweep of M per iteration plus the prologue column pass
    alg = float(n) * n * 8 * (niter + 2)
    gbs = alg / dt / 1e9
    out = {"value": 1.0 / dt, "unit": "problems/s", "ms_per_step": dt * 1e3, "dtype": "f64",
           "config": {"workload": "reference CPU case n=m=2048 d=2 fp64 reg=1e-2 stopThr=1e-9 "
                                  "numItermax=1000", "niter": niter,
                      "iterations_per_s": niter / dt,
                      "cost": float((G * Mt).sum())},
           "roofline": {"bound": "hbm", "kernel": "row_sweep_kernel (+col_merge, whole call)",
                        "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                        "note": "the 32 MiB cost fits the 256 MB last-level cache"}}
    if not args.no_cpu:
        from oracle.sinkhorn import sinkhorn_log
        its = 50
        t0 = time.perf_counter()
        sinkhorn_log(a, b, M, 1e-2, numItermax=its, stopThr=0.0, log=True, warn=False)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": its / t, "unit": "iterations/s", "cores": 1,
                               "kind": "port", "sample": f"{its} iterations of the float64 "
                               "NumPy sinkhorn_log on the same problem"}
    return out


def run_ragged(args, torch, dev):
    """B = 512 uneven problems (n_i, m_i ~ U[64, 1024], d = 16, GMM(1-4)),
    stopThr 1e-6, numItermax 2000: ragged distances + batched solver, replayed
    as one captured hipGraph (``--graph``, default) or launched eagerly."""
    from pot_amd import problems, ragged
    fam = problems.config3(seed=2, B=512)
    ns = [len(x) for x in fam.xs]
    ms = [len(y) for y in fam.ys]
    X = [torch.tensor(x, device=dev) for x in fam.xs]
    Y = [torch.tensor(y, device=dev) for y in fam.ys]
    Xp = torch.cat(X)
    Yp = torch.cat(Y)
    G = ragged.RaggedGraph(ns, ms, 16, numItermax=2000, stopThr=1e-6, check_every=10,
                           waves=args.waves, prefetch=args.prefetch,
                           occupancy=args.occ, cost=args.cost,
                           split="auto" if args.split else None) if args.graph else None
    times, kms, abytes = [], [], []
    res = None
    for k in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        if G is not None:
            e0.record()
            res = G.solve(Xp, Yp, fam.regs)
            e1.record()
        elif args.cost == "points":
            R = ragged.points_from_clouds(X, Y, fam.regs)
            e0.record()
            res = ragged.sinkhorn_log_ragged_points(R, numItermax=2000, stopThr=1e-6,
                                                    unsolved="nan")
            e1.record()
        else:
            R = ragged.from_clouds(X, Y, fam.regs)
            e0.record()
            res = ragged.sinkhorn_log_ragged(R, numItermax=2000, stopThr=1e-6,
                                             waves=args.waves, prefetch=args.prefetch,
                                             occupancy=args.occ)
            e1.record()
        torch.cuda.synchronize()
        if k >= args.warmup:
            times.append(time.perf_counter() - t0)
            kms.append(e0.elapsed_time(e1))
            nit = res.niter.cpu().numpy().astype(np.int64)
            passes = nit + (4 if args.cost == "points" else 2)
            abytes.append(float(np.sum(passes * np.asarray(ns, np.int64) * np.asarray(ms))
                                * 4))
    dt = float(np.mean(times))
    gbs = np.mean(abytes) / (np.mean(kms) * 1e-3) / 1e9
    if args.cost == "points":
        # executed MFMA flops: 128 per element of every pass (see run_warmstart)
        tf = np.mean(abytes) / 4 * 128.0 / (np.mean(kms) * 1e-3) / 1e12
        roof = {"bound": "mfma", "kernel": "batch_points_kernel", "achieved": tf,
                "peak": MFMA_PEAK_TFLOPS["f16"], "unit": "TFLOP/s",
                "frac": tf / MFMA_PEAK_TFLOPS["f16"], "traffic": None}
    else:
        roof = {"bound": "hbm", "kernel": "batch_log_kernel (+ragged dist in the graph)",
                "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": None}
    out = {"value": 512 / dt, "unit": "problems/s", "ms_per_step": dt * 1e3, "dtype": "f32",
           "config": {"workload": "ragged B=512 n_i,m_i~U[64,1024] d=16 GMM(1-4) "
                                  "reg~logU[5e-3,5e-2] stopThr=1e-6",
                      "launch": "hipGraph replay" if G is not None else "eager",
                      "cost": args.cost, "seed": 2},
           "iters": {"mean": float(nit.mean()), "max": int(nit.max())},
           "roofline": roof}
    if not args.no_cpu:
        # bounded sample: problems 0-7 by the float64 oracle; the same run gives
        # the per-problem plan error of the GPU solution
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        import oracle.sinkhorn as osk
        u_all, v_all = res.u.cpu().numpy(), res.v.cpu().numpy()
        ao = np.concatenate([[0], np.cumsum(ns)[:-1]])
        bo = np.concatenate([[0], np.cumsum(ms)[:-1]])
        tot, worst = 0.0, 0.0
        for q in range(8):
            x64 = fam.xs[q].astype(np.float64)
            y64 = fam.ys[q].astype(np.float64)
            M = np.maximum((x64 ** 2).sum(1)[:, None] + (y64 ** 2).sum(1)[None]
                           - 2 * x64 @ y64.T, 0.0)
            M /= M.max()
            a = np.full(ns[q], 1.0 / ns[q])
            b = np.full(ms[q], 1.0 / ms[q])
            t0 = time.perf_counter()
            Pr = osk.sinkhorn_log(a, b, M, fam.regs[q], numItermax=2000, stopThr=1e-6,
                                  warn=False)
            tot += time.perf_counter() - t0
            u = u_all[ao[q]:ao[q] + ns[q]]
            v = v_all[bo[q]:bo[q] + ms[q]]
            P = np.exp(-M / fam.regs[q] + u[:, None] + v[None, :])
            worst = max(worst, float(np.max(np.abs(P - Pr)) / Pr.max()))
        out["cpu_baseline"] = {"value": 8 / tot, "unit": "problems/s", "cores": 1,
                               "kind": "port",
                               "sample": "problems 0-7, float64 NumPy sinkhorn_log"}
        out["plan_err_vs_cpu"] = {"max_abs_over_max_entry": worst, "problems": 8}
    return out


def run_online(args, torch, dev):
    """16384 x 16384, d = 64, reg = 1e-2: the dense problem with the cost on the fly."""
    from pot_amd import online, problems
    x, y, a, b = problems.config2_clouds()
    n, d = x.shape
    X = torch.tensor(x, device=dev)
    Y = torch.tensor(y, device=dev)
    at = torch.tensor(a, device=dev)
    bt = torch.tensor(b, device=dev)
    its = 200
    # the normalisation (max cost) once, outside the timed region
    _, _, lg = online.sinkhorn_online(X, Y, 1e-2, at, bt, numItermax=1, stopThr=0.0,
                                      precision=args.precision, warn=False)
    cs = lg["cost_scale"]

    def solve():
        return online.sinkhorn_online(X, Y, 1e-2, at, bt, numItermax=its, stopThr=0.0,
                                      precision=args.precision, cost_scale=cs, warn=False)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        u, v, lg = solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    passes = 2 * its + 1
    flops = 2.0 * n * n * d * passes * (3 if args.precision == "bf16x3" else 1)
    tf = flops / dt / 1e12
    peak = MFMA_PEAK_TFLOPS[args.precision]
    _, sums = online.online_plan(X, Y, 1e-2, u, v, cs, at, bt)
    out_cpu = None
    if not args.no_cpu:
        # the reference's on-the-fly CPU path (blocked cost recomputed every
        # half-iteration, float64 NumPy, one thread) on a 2048-row slice,
        # 2 iterations, scaled to 16384 rows
        from scipy.special import logsumexp
        from threadpoolctl import threadpool_limits
        rows = 2048
        xs = x[:rows].astype(np.float64)
        yt = y.astype(np.float64)
        la, lb = np.log(a[:rows]), np.log(b)
        g = np.zeros(n)
        with threadpool_limits(1):
            t0 = time.perf_counter()
            for _ in range(2):
                M = np.maximum((xs ** 2).sum(1)[:, None] + (yt ** 2).sum(1)[None] - 2 * xs @ yt.T,
                               0.0) / cs
                f = la - logsumexp(g[None, :] - M / 1e-2, axis=1)
                M = np.maximum((xs ** 2).sum(1)[:, None] + (yt ** 2).sum(1)[None] - 2 * xs @ yt.T,
                               0.0) / cs
                g = lb - logsumexp(f[:, None] - M / 1e-2, axis=0)
            t_it = (time.perf_counter() - t0) / 2 * (n / rows)
        out_cpu = {"value": 1.0 / t_it, "unit": "iterations/s", "cores": 1, "kind": "port",
                   "sample": "2 float64 NumPy iterations with the cost recomputed per "
                             "half-iteration (the lazy empirical_sinkhorn CPU path) on a "
                             "2048 x 16384 slice, scaled to 16384 rows, 1 BLAS thread"}
    return {"value": its / dt, "unit": "iterations/s", "ms_per_step": dt * 1e3,
            "cpu_baseline": out_cpu,
            "dtype": args.precision,
            "config": {"workload": "online n=m=16384 d=64 reg=1e-2 200 fixed iterations",
                       "precision": args.precision, "cost_scale": cs},
            "exp_per_s": float(n) * n * passes / dt,
            "check_fp64": sums,
            "roofline": {"bound": "mfma", "kernel": "online_pass_kernel (+merge)",
                         "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak,
                         "traffic": None,
                         "note": "whole-iteration time; 2 n m d flops per pass (x3 for the "
                                 "split), 2 passes per iteration + 1 prologue; the pass is "
                                 "bound by exp2/VALU work (n m exp2 per pass), not MFMA"}}


def run_precision(args, torch, dev):
    """Precision study: n = m = 65536, d = 128, reg = 5e-2, 100 iterations."""
    from pot_amd import online, problems
    x, y = problems.config4_clouds()
    n, d = x.shape
    X = torch.tensor(x, device=dev)
    Y = torch.tensor(y, device=dev)
    reg, its = 5e-2, 100
    t_all = time.perf_counter()
    rows = {}
    ref = None
    for prec in ["f64", "f32", "bf16x3", "bf16", "fp8"]:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        u, v, lg = online.sinkhorn_online(X, Y, reg, numItermax=its, stopThr=0.0,
                                          precision=prec, warn=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        _, sums = online.online_plan(X, Y, reg, u, v, lg["cost_scale"])
        row = {"s": dt, "iterations_per_s": its / dt, "cost": sums["cost"],
               "err_a": sums["err_a"], "err_b": sums["err_b"], "cost_scale": lg["cost_scale"]}
        if ref is None:
            ref = (u, v, sums["cost"])
        else:
            du = (u - ref[0]).cpu().numpy()
            dv = (v - ref[1]).cpu().numpy()
            g = du.mean()
            row["max_logplan_err"] = float(np.max(np.abs(du - g)) + np.max(np.abs(dv + g)))
            row["cost_rel_err"] = abs(sums["cost"] - ref[2]) / ref[2]
        rows[prec] = row
        print(f"[precision] {prec}: {json.dumps(row)}", file=sys.stderr, flush=True)
    # subsample against the float64 NumPy oracle
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import oracle.sinkhorn as osk
    k = args.subsample
    xs, ys = x[:k], y[:k]
    M = ((xs.astype(np.float64) ** 2).sum(1)[:, None] + (ys.astype(np.float64) ** 2).sum(1)[None]
         - 2.0 * xs.astype(np.float64) @ ys.astype(np.float64).T)
    M = np.maximum(M, 0.0)
    M /= M.max()
    ab = np.full(k, 1.0 / k)
    t0 = time.perf_counter()
    Pr = osk.sinkhorn_log(ab, ab, M, reg, numItermax=its, stopThr=0.0, warn=False)
    t_cpu = time.perf_counter() - t0
    cost_ref = float(np.sum(Pr * M))
    sub = {}
    for prec in ["f64", "f32", "bf16x3", "bf16", "fp8"]:
        u, v, lg = online.sinkhorn_online(xs, ys, reg, numItermax=its, stopThr=0.0,
                                          precision=prec, warn=False)
        P, sums = online.online_plan(xs, ys, reg, u, v, lg["cost_scale"], want_plan=True)
        Pn = P.cpu().numpy()
        mask = Pr >= 1e-30
        sub[prec] = {"plan_rel_err": float(np.max(np.abs(Pn[mask] - Pr[mask]) / Pr[mask])),
                     "cost_rel_err": abs(sums["cost"] - cost_ref) / cost_ref,
                     "err_b": sums["err_b"]}
        print(f"[precision] subsample {prec}: {json.dumps(sub[prec])}", file=sys.stderr,
              flush=True)
    best = rows["bf16x3"]
    return {"value": best["iterations_per_s"], "unit": "iterations/s",
            "ms_per_step": 1e3 * best["s"] / its, "dtype": "bf16x3",
            "config": {"workload": "precision n=m=65536 d=128 reg=5e-2 100 iterations, online cost",
                       "subsample": k},
            "full_size_vs_f64_gpu": rows, "subsample_vs_numpy_oracle": sub,
            "cpu_baseline": {"value": its / t_cpu, "unit": "iterations/s", "cores": 1,
                             "kind": "port",
                             "sample": f"float64 NumPy sinkhorn_log on the {k}-point subsample"},
            "wall_s": time.perf_counter() - t_all}


def run_greenkhorn(args, torch, dev):
    """The greedy Sinkhorn (ot.bregman.greenkhorn) through the drop-in: n = m =
    2048 points in d = 2 (y shifted by 1), cost / max, uniform weights, reg
    0.05, 10000 updates (stopThr 0).  A step is one whole call (K, K^T, the
    violations, the loop and the plan) on inputs resident in HBM.  With the
    CPU leg, 1000 updates of the float64 NumPy oracle on the same problem."""
    from pot_amd import bregman
    rng = np.random.default_rng(0)
    n, reg, its = 2048, 0.05, 10000
    x = rng.normal(size=(n, 2))
    y = rng.normal(size=(n, 2)) + 1.0
    M = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    M /= M.max()
    a = np.full(n, 1.0 / n)
    at, Mt = torch.tensor(a, device=dev), torch.tensor(M, device=dev)

    def solve():
        return bregman.greenkhorn(at, at, Mt, reg, numItermax=its, stopThr=0.0, warn=False)

    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solve()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    out = {"value": its / dt, "unit": "updates/s", "ms_per_step": dt * 1e3, "dtype": "f64",
           "config": {"workload": f"greenkhorn n=m={n} d=2 reg={reg} {its} updates"},
           "roofline": {"bound": "hbm", "kernel": "gk_loop_kernel (whole call)",
                        "achieved": its * n * 8 / dt / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": its * n * 8 / dt / 1e9 / HBM_PEAK_GBS,
                        "traffic": None,
                        "note": "one row of K (n doubles) per update; the loop is bound by "
                                "the latency of its reductions and barriers, not by HBM"}}
    if not args.no_cpu:
        from oracle.greenkhorn import greenkhorn
        k = 1000
        t0 = time.perf_counter()
        greenkhorn(a, a, M, reg, numItermax=k, stopThr=0.0, warn=False)
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": k / t, "unit": "updates/s", "cores": 1, "kind": "port",
                               "sample": f"{k} updates of the float64 NumPy greenkhorn on the "
                                         "same problem"}
    return out


def run_copy(args, torch, dev):
    """Device-to-device copy of 1 GiB: the HBM reference of the stream kernels and
    the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters."""
    nbytes = 1 << 30
    src = torch.ones(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * nbytes / (ms * 1e-3) / 1e9
    return {"value": gbs, "unit": "GB/s", "ms_per_step": ms, "dtype": "f32",
            "config": {"workload": "device copy 1 GiB (read 1 GiB + write 1 GiB)"},
            "roofline": {"bound": "hbm", "kernel": "copy", "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                         "traffic": 2 * nbytes}}


def main():
    args = _args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; use --gpus 1"}))
        sys.exit(2)
    import torch
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    dev = torch.device("cuda", 0)
    runner = {"warmstart": run_warmstart, "dense": run_dense, "c1": run_c1, "ragged": run_ragged,
              "copy": run_copy, "online": run_online, "greenkhorn": run_greenkhorn,
              "precision": run_precision}[args.workload]
    res = runner(args, torch, dev)
    # The default workload carries the headline metric; the others name
    # their own measure (their config says which workload it is).
    metric = METRIC if args.workload == "warmstart" else {
        "dense": "sinkhorn_log iterations/s, dense 16384^2 fp32 cost",
        "c1": "sinkhorn_log problems/s to stopThr, the reference's 2048^2 float64 case",
        "ragged": "problems/s to marginal err<=1e-6, 512 uneven point-cloud problems",
        "copy": "device copy GB/s",
        "online": "sinkhorn_log iterations/s, online cost (never stored)",
        "precision": "online-cost precision study",
        "greenkhorn": "greenkhorn greedy updates/s, 2048^2 float64",
    }[args.workload]
    line = {"metric": metric, "value": res.pop("value"), "unit": res.pop("unit"), "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": res.pop("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": res.pop("dtype"), "data": "synthetic (seeded NumPy point clouds)"}
    line.update(res)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
