"""ctypes binding of ``libpot_amd.so`` (the C ABI declared in include/pot_amd.h).

The library is loaded from the package directory only.  There is no fallback:
if it is missing or lacks a symbol, every compute entry point raises.
"""

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpot_amd.so")

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_size = ctypes.c_size_t
c_vp = ctypes.c_void_p
c_ip = ctypes.POINTER(ctypes.c_int)

# name -> (restype, argtypes); must match include/pot_amd.h
SIGNATURES = {
    "potamd_version": (c_int, []),
    "potamd_status_string": (ctypes.c_char_p, [c_int]),
    "potamd_dense_workspace_size": (c_size, [c_int, c_int, c_int, c_int, c_int]),
    "potamd_sinkhorn_dense": (c_int, [
        c_int, c_int, c_int, c_int,          # dtype, domain, order, kernel_given
        c_vp, c_i64, c_i64, c_int, c_int, c_int,  # M, ldm, bstride, batch, n, m
        c_dbl, c_vp, c_vp, c_vp, c_vp,       # reg, a, b, u, v
        c_int, c_dbl, c_int, c_vp, c_int,    # max_iter, stop_thr, err_norm, err_hist, err_cap
        c_int, c_int, c_int, c_int,          # threads, cols_per_thread, rows_per_block, prefetch
        c_vp, c_size, c_vp, c_vp, c_ip]),    # workspace, ws_bytes, host_ctrl, stream, info
    "potamd_sinkhorn_dense_each": (c_int, [
        c_int, c_int, c_int, c_int,          # dtype, domain, order, kernel_given
        c_vp, c_i64, c_i64, c_int, c_int, c_int,  # M, ldm, bstride, batch, n, m
        c_dbl, c_vp, c_vp, c_vp, c_vp,       # reg, a, b, u, v
        c_int, c_dbl,                        # max_iter, stop_thr
        c_int, c_int, c_int, c_int,          # threads, cols_per_thread, rows_per_block, prefetch
        c_vp, c_size, c_vp, c_vp,            # workspace, ws_bytes, host_ctrl, stream
        c_vp, c_vp, c_ip]),                  # niters, statuses, info
    "potamd_plan_workspace_size": (c_size, [c_int, c_int]),
    "potamd_plan_dense": (c_int, [
        c_int, c_int, c_int, c_vp, c_i64, c_i64, c_int, c_int, c_int, c_dbl,
        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_size, c_vp]),
    "potamd_sinkhorn_log_batch": (c_int, [
        c_int, c_int, c_int, c_int, c_vp,    # dtype, batch, max_n, max_m, M
        c_vp, c_vp, c_vp, c_vp,              # moff, ns, ms, ldms
        c_vp, c_vp, c_vp,                    # aoff, boff, regs
        c_vp, c_vp, c_vp, c_vp,              # a, b, u, v
        c_int, c_int, c_dbl, c_dbl, c_vp, c_int,  # max_iter, check_every, stop, fallback, order, vec
        c_vp, c_vp, c_vp, c_vp,              # niter, status, err_last, err_first
        c_int, c_vp]),                       # waves, stream
    "potamd_sinkhorn_log_batch_points": (c_int, [
        c_int, c_int, c_int, c_int, c_vp, c_vp,  # batch, max_n, max_m, d, X, Y
        c_vp, c_vp, c_vp, c_vp, c_vp, c_int,     # xoff, yoff, ns, ms, regs, normalize
        c_vp, c_vp, c_vp, c_vp,                  # a, b, u, v
        c_int, c_int, c_dbl, c_dbl, c_vp,        # max_iter, check_every, stop, fallback, order
        c_vp, c_vp, c_vp, c_vp, c_vp,            # niter, status, err_last, err_first, cmax
        c_int, c_vp, c_vp]),                     # waves, queue, stream
    "potamd_sinkhorn_log_batch_points_split": (c_int, [
        c_int, c_int, c_int, c_int, c_vp, c_vp,  # batch, max_n, max_m, d, X, Y
        c_vp, c_vp, c_vp, c_vp, c_vp, c_int,     # xoff, yoff, ns, ms, regs, normalize
        c_vp, c_vp, c_vp, c_vp,                  # a, b, u, v
        c_int, c_int, c_dbl, c_dbl,              # max_iter, check_every, stop, fallback
        c_int, c_vp, c_vp, c_int, c_vp, c_size,  # items, order, split, nsplit, ws, ws bytes
        c_vp, c_vp, c_vp, c_vp, c_vp,            # niter, status, err_last, err_first, cmax
        c_vp, c_vp]),                            # queue, stream
    "potamd_points_split_workspace_size": (c_size, [c_int]),
    "potamd_sinkhorn_log_batch_points_split_noalloc": (c_int, [
        c_int, c_int, c_int, c_int, c_vp, c_vp,  # batch, max_n, max_m, d, X, Y
        c_vp, c_vp, c_vp, c_vp, c_vp, c_int,     # xoff, yoff, ns, ms, regs, normalize
        c_vp, c_vp, c_vp, c_vp,                  # a, b, u, v
        c_int, c_int, c_dbl, c_dbl,              # max_iter, check_every, stop, fallback
        c_int, c_vp, c_vp, c_int,                # items, order, split, nsplit
        c_vp, c_vp, c_vp, c_vp, c_vp,            # niter, status, err_last, err_first, cmax
        c_vp]),                                  # stream
    "potamd_noalloc_scratch_bytes": (c_size, []),
    "potamd_ragged_points_cmax": (c_int, [
        c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "potamd_sqeuclidean_ragged": (c_int, [
        c_int, c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
        c_vp, c_vp, c_vp, c_vp, c_vp]),
    "potamd_gemm_bf16": (c_int, [
        c_int, c_int, c_int, c_vp, c_i64, c_int, c_vp, c_i64, c_int, c_vp, c_i64,
        c_int, ctypes.c_float, c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_int, c_vp]),
    "potamd_mlp_fused": (c_int, [
        c_int, c_int, c_int, c_vp, c_vp, c_vp, c_vp, c_vp,   # mode, rows, seg, inp, pts, wts, regs, cmax
        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "potamd_mlp_op": (c_int, [
        c_int, c_int, c_int, c_int, c_int, c_int, c_i64, c_int, c_dbl, c_vp,
        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "potamd_online_workspace_size": (c_size, [c_int, c_int, c_int, c_int]),
    "potamd_sinkhorn_online": (c_int, [
        c_int, c_int, c_vp, c_i64, c_vp, c_i64,  # precision, dtype, X, ldx, Y, ldy
        c_int, c_int, c_int, c_dbl, c_dbl,   # n, m, d, reg, cost_scale
        c_vp, c_vp, c_vp, c_vp,              # a, b, u, v
        c_int, c_dbl, c_int,                 # max_iter, stop_thr, semantics
        c_vp, c_int, c_int,                  # err_hist, err_cap, splits
        c_vp, c_size, c_vp, c_vp, c_ip,      # workspace, ws_bytes, host_ctrl, stream, info
        ctypes.POINTER(c_dbl)]),             # cost_scale_out
    "potamd_online_plan_workspace_size": (c_size, [c_int, c_int]),
    "potamd_online_plan": (c_int, [
        c_int, c_vp, c_i64, c_vp, c_i64, c_int, c_int, c_int, c_dbl, c_dbl,
        c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_size, c_vp]),
    "potamd_greenkhorn_workspace_size": (c_size, [c_int, c_int]),
    "potamd_greenkhorn": (c_int, [
        c_int, c_vp, c_i64, c_int, c_int, c_dbl,   # dtype, M, ldm, n, m, reg
        c_vp, c_vp, c_vp, c_vp, c_int, c_dbl,      # a, b, u, v, max_iter, stop_thr
        c_vp, c_vp, c_size, c_vp, c_ip]),          # P, workspace, ws_bytes, stream, info
    "potamd_stabilized_workspace_size": (c_size, [c_int, c_int, c_int]),
    "potamd_sinkhorn_stabilized": (c_int, [
        c_int, c_vp, c_i64, c_int, c_int, c_dbl,   # dtype, M, ldm, n, m, reg
        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,        # a, b, alpha, beta, u, v
        c_dbl, c_int, c_dbl, c_int,                # tau, max_iter, stop_thr, print_period
        c_vp, c_int, c_vp,                         # err_hist, err_cap, marg_sq
        c_vp, c_size, c_vp, c_vp, c_ip]),          # workspace, ws_bytes, host_ctrl, stream, info
    "potamd_stabilized_multi_workspace_size": (c_size, [c_int, c_int, c_int, c_int]),
    "potamd_sinkhorn_stabilized_multi": (c_int, [
        c_int, c_vp, c_i64, c_int, c_int, c_int, c_dbl,  # dtype, M, ldm, n, m, H, reg
        c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,        # a, b (H x m), alpha, beta, u, v
        c_dbl, c_int, c_dbl, c_int,                # tau, max_iter, stop_thr, print_period
        c_vp, c_int,                               # err_hist, err_cap
        c_vp, c_size, c_vp, c_vp, c_ip]),          # workspace, ws_bytes, host_ctrl, stream, info
    "potamd_pairwise_batch": (c_int, [
        c_int, c_int, c_vp, c_vp, c_int, c_int, c_int, c_int,  # dtype, metric, X, Y, B, n, m, d
        c_dbl, c_dbl, c_dbl, c_vp, c_vp]),                     # p, q, eps, M, stream
    "potamd_gemm_batched": (c_int, [
        c_int, c_int, c_int, c_int, c_int, c_dbl,              # dtype, B, M, N, K, alpha
        c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_i64, c_i64,  # A, sA, rsA, csA, B, sB, rsB, csB
        c_dbl, c_vp, c_i64, c_i64, c_vp]),                     # beta, C, sC, ldc, stream
    "potamd_sqeuclidean": (c_int, [
        c_int, c_vp, c_vp, c_int, c_int, c_int, c_int, c_i64, c_i64, c_vp, c_i64,
        c_i64, c_int, c_int, c_int, c_vp]),
}

CONVERGED, MAX_ITER, NUMERICAL, NEED_EXACT = 0, 1, 2, 3
F32, F64 = 0, 1
LOG, SCALING = 0, 1
COL_FIRST, ROW_FIRST = 0, 1
ONLINE_PRECISIONS = {"f32": 0, "bf16": 1, "bf16x3": 2, "fp8": 3, "f64": 4}

_lock = threading.Lock()
_lib = None


class LibraryError(RuntimeError):
    """The HIP library is missing, incomplete or returned an error."""


def load():
    """Load and bind the library once; raises LibraryError if unusable."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise LibraryError(
                f"{LIB_PATH} not found: build it with `python -m pot_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            try:
                fn = getattr(lib, name)
            except AttributeError as e:
                raise LibraryError(f"{LIB_PATH} lacks symbol {name}") from e
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code, what):
    """Raise on negative return codes of the C ABI."""
    if code < 0:
        msg = load().potamd_status_string(code).decode()
        raise LibraryError(f"{what} failed: {msg} (code {code})")
    return code
