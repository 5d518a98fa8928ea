"""Tile-shape autotuner of the dense solver.

For a (dtype, n, m) shape class it times every compiled row-sweep
configuration that covers m -- threads per workgroup, columns per thread,
rows prefetched, rows per workgroup -- on a few fixed iterations of a
synthetic problem of that class, and keeps the fastest.  Results are cached
in a JSON file keyed by GPU architecture and shape class; the solver looks
the cache up and falls back to the library's default shape when it has no
entry.
"""

import json
import os
import threading
import time

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


if __name__ == "__main__":
    import sys
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 16384
    m = int(sys.argv[2]) if len(sys.argv) > 2 else n
    dt = sys.argv[3] if len(sys.argv) > 3 else "f32"
    tile, t = tune(n, m, dt, log=lambda s: print(s, flush=True))
    print(json.dumps({"n": n, "m": m, "dtype": dt, "tile": tile, "us_per_iter": t * 1e6}))
