"""Tile-shape autotuner of the dense solver and of the ragged batch solvers.

Dense: for a (dtype, n, m) shape class it times every compiled row-sweep
configuration that covers m -- threads per workgroup, columns per thread,
rows prefetched, rows per workgroup -- on a few fixed iterations of a
synthetic problem of that class, and keeps the fastest.

Ragged batches: for a (cost kind, largest n, largest m, d) class it times
every compiled batch configuration -- wavefronts per problem, and for
materialised costs rows prefetched and the register cap -- on a few fixed
iterations of a synthetic batch with the caller's sizes.

Results are cached in a JSON file keyed by GPU architecture and shape class;
the solvers look the cache up and fall back to the library's default shape
when it has no entry.
"""

import json
import os
import threading
import time

import numpy as np

import torch

from . import _lib

CACHE = os.environ.get("POT_AMD_TUNE_CACHE",
                       os.path.join(os.path.dirname(os.path.abspath(__file__)), "tune_cache.json"))

# (threads, cols_per_thread, prefetch) compiled per dtype (csrc/dense.hip)
ROW_CONFIGS = {
    "f32": [(64, 4, 1), (256, 4, 1), (256, 8, 1), (512, 8, 1), (1024, 8, 1), (1024, 16, 1),
            (1024, 8, 2), (1024, 16, 2), (512, 32, 1), (512, 32, 2)],
    "f64": [(64, 2, 1), (256, 2, 1), (256, 4, 1), (512, 4, 1), (1024, 4, 1), (1024, 8, 1)],
}
RPB_CHOICES = (8, 16, 32, 64, 128)

_lock = threading.RLock()
_mem = None


def _arch():
    try:
        return torch.cuda.get_device_properties(0).gcnArchName.split(":")[0]
    except Exception:  # noqa: BLE001 - no GPU: no arch
        return "unknown"


def _class(n, m):
    """Shape class: powers of two of n and m."""
    return f"{1 << max(0, (int(n) - 1).bit_length())}x{1 << max(0, (int(m) - 1).bit_length())}"


def _key(dtype, n, m):
    return f"{_arch()}|{dtype}|{_class(n, m)}"


def _load():
    global _mem
    if _mem is None:
        with _lock:
            if _mem is None:
                try:
                    with open(CACHE) as fh:
                        _mem = json.load(fh)
                except (OSError, ValueError):
                    _mem = {}
    return _mem


def lookup(dtype, n, m):
    """Cached (threads, cols_per_thread, rows_per_block, prefetch) or None."""
    ent = _load().get(_key(dtype, n, m))
    return tuple(ent["tile"]) if ent else None


def candidates(dtype, m):
    return [c for c in ROW_CONFIGS[dtype] if c[0] * c[1] >= m]


def tune(n, m, dtype="f32", iters=20, reps=2, save=True, log=None):
    """Time every candidate tile on an n x m problem; store and return the best."""
    from . import _solve
    dev = torch.device("cuda", 0)
    tdt = torch.float32 if dtype == "f32" else torch.float64
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    M = torch.rand((n, m), generator=g, device=dev, dtype=tdt)
    a = torch.full((1, n), 1.0 / n, dtype=torch.float64, device=dev)
    b = torch.full((1, m), 1.0 / m, dtype=torch.float64, device=dev)
    z = torch.zeros((1, n), dtype=torch.float64, device=dev)
    zv = torch.zeros((1, m), dtype=torch.float64, device=dev)
    results = []
    for (T, cpt, pf) in candidates(dtype, m):
        for rpb in RPB_CHOICES:
            if rpb > n:
                continue
            tile = (T, cpt, rpb, pf)
            best = float("inf")
            try:
                for _ in range(reps):
                    torch.cuda.synchronize()
                    t0 = time.perf_counter()
                    _solve.dense_solve(M, a, b, z, zv, 0.1, domain=_lib.LOG,
                                       order=_lib.COL_FIRST, max_iter=iters, stop_thr=0.0,
                                       tile=tile)
                    torch.cuda.synchronize()
                    best = min(best, (time.perf_counter() - t0) / iters)
            except _lib.LibraryError:
                continue
            results.append((best, tile))
            if log:
                log(f"  tile {tile}: {best * 1e6:.1f} us/iteration")
    results.sort()
    best_t, best_tile = results[0]
    if save:
        with _lock:
            mem = _load()
            mem[_key(dtype, n, m)] = {"tile": list(best_tile), "us_per_iter": best_t * 1e6,
                                      "measured": [[t * 1e6, list(c)] for t, c in results]}
            with open(CACHE, "w") as fh:
                json.dump(mem, fh, indent=1, sort_keys=True)
    return best_tile, best_t


# Ragged batch configurations compiled in csrc/batch.hip (waves, rows
# prefetched, register cap in waves per SIMD; 255 = no cap) and
# csrc/batch_points.hip (waves).
BATCH_CONFIGS = {
    "dense": [(8, 1, 4), (8, 1, 255), (4, 1, 0), (8, 2, 0), (4, 2, 0)],
    "points": [(8, 0, 0), (16, 0, 0)],
}


def _pow2(x):
    """Nearest power of two (in the log) of a positive number."""
    return 1 << max(0, int(round(np.log2(max(float(x), 1.0)))))


def _batch_key(cost, ns, ms, d):
    """Largest sizes (rounded up) and mean sizes (rounded) of the batch."""
    mean = f"{_pow2(np.mean(ns))}x{_pow2(np.mean(ms))}"
    return f"{_arch()}|batch-{cost}|{_class(int(np.max(ns)), int(np.max(ms)))}|mean{mean}|d{int(d)}"


def lookup_batch(cost, ns, ms, d=0):
    """Cached (waves, prefetch, occupancy) of a ragged batch class, or None."""
    if len(ns) == 0:
        return None
    ent = _load().get(_batch_key(cost, ns, ms, d if cost == "points" else 0))
    return tuple(ent["config"]) if ent else None


def tune_batch(ns, ms, d=16, cost="points", reg=1e-2, iters=20, reps=2, save=True, log=None,
               seed=0):
    """Time every batch configuration on ``iters`` fixed iterations of
    synthetic Gaussian clouds with these sizes; store and return the best."""
    from . import ragged
    if cost not in BATCH_CONFIGS:
        raise ValueError("cost must be 'dense' or 'points'")
    rng = np.random.default_rng(seed)
    ns = [int(n) for n in ns]
    ms = [int(m) for m in ms]
    xs = [rng.standard_normal((n, d)).astype(np.float32) for n in ns]
    ys = [(rng.standard_normal((m, d)) + 0.5).astype(np.float32) for m in ms]
    regs = np.full(len(ns), float(reg))
    if cost == "points":
        P = ragged.points_from_clouds(xs, ys, regs)

        def run(cfg):
            ragged.sinkhorn_log_ragged_points(P, numItermax=iters, stopThr=0.0, waves=cfg[0],
                                              unsolved="nan")
    else:
        R = ragged.from_clouds(xs, ys, regs)

        def run(cfg):
            ragged.sinkhorn_log_ragged(R, numItermax=iters, stopThr=0.0, waves=cfg[0],
                                       prefetch=cfg[1], occupancy=cfg[2])
    results = []
    for cfg in BATCH_CONFIGS[cost]:
        best = float("inf")
        try:
            run(cfg)  # warm-up
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                run(cfg)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
        except _lib.LibraryError:
            continue
        results.append((best, cfg))
        if log:
            log(f"  batch config {cfg}: {best * 1e3:.2f} ms for {iters} iterations")
    results.sort()
    best_t, best_cfg = results[0]
    if save:
        with _lock:
            mem = _load()
            mem[_batch_key(cost, ns, ms, d if cost == "points" else 0)] = {
                "config": list(best_cfg), "ms": best_t * 1e3, "iters": iters,
                "batch": len(ns), "measured": [[t * 1e3, list(c)] for t, c in results]}
            with open(CACHE, "w") as fh:
                json.dump(mem, fh, indent=1, sort_keys=True)
    return best_cfg, best_t


def _main_batch(argv):
    """python -m pot_amd.autotune batch COST B NMIN NMAX D: sizes uniform in
    [NMIN, NMAX] (seed 0), as in the ragged and warm-start workloads."""
    cost = argv[0] if argv else "points"
    B = int(argv[1]) if len(argv) > 1 else 4096
    lo = int(argv[2]) if len(argv) > 2 else 1024
    hi = int(argv[3]) if len(argv) > 3 else lo
    d = int(argv[4]) if len(argv) > 4 else 16
    rng = np.random.default_rng(0)
    ns = rng.integers(lo, hi + 1, size=B)
    ms = rng.integers(lo, hi + 1, size=B)
    cfg, t = tune_batch(ns, ms, d=d, cost=cost, log=lambda s: print(s, flush=True))
    print(json.dumps({"cost": cost, "batch": B, "sizes": [lo, hi], "d": d, "config": list(cfg),
                      "ms": t * 1e3}))


if __name__ == "__main__":
    import sys
    if len(sys.argv) > 1 and sys.argv[1] == "batch":
        _main_batch(sys.argv[2:])
        sys.exit(0)
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 16384
    m = int(sys.argv[2]) if len(sys.argv) > 2 else n
    dt = sys.argv[3] if len(sys.argv) > 3 else "f32"
    tile, t = tune(n, m, dt, log=lambda s: print(s, flush=True))
    print(json.dumps({"n": n, "m": m, "dtype": dt, "tile": tile, "us_per_iter": t * 1e6}))
