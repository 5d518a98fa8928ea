"""Thin device-tensor wrappers over the C ABI (no argument semantics here)."""

import ctypes
from dataclasses import dataclass

import torch

from . import _device, _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


@dataclass
class DenseResult:
    u: torch.Tensor        # (B, n) float64: log potentials or scalings
    v: torch.Tensor        # (B, m) float64
    niter: int             # the reference's `ii` at the stop
    errs: list             # marginal errors of the check iterations
    status: int            # _lib.CONVERGED / MAX_ITER / NUMERICAL
    bad_iter: int
    niters: torch.Tensor = None    # (B,) int32 per-problem stops (each=True only)
    statuses: torch.Tensor = None  # (B,) int32 per-problem statuses (each=True only)


def dense_solve(M, a, b, u0, v0, reg, *, domain, order, kernel_given=False,
                max_iter=1000, stop_thr=1e-9, err_norm=0, batch=None,
                bstride=None, tile=None, each=False):
    """Run potamd_sinkhorn_dense (or, with ``each``, potamd_sinkhorn_dense_each:
    every problem stops on its own check, and the result's ``niters`` /
    ``statuses`` hold the per-problem stops).

    M: (n, m) or (B, n, m) device tensor (float32/float64), rows contiguous.
    a: (B, n), b: (B, m), u0: (B, n), v0: (B, m) float64 device tensors.
    ``batch``/``bstride`` override the batch layout (bstride = 0 shares M).
    """
    lib = _lib.load()
    dev = M.device
    code = _lib.F32 if M.dtype == torch.float32 else _lib.F64
    if M.dim() == 2:
        n, m = M.shape
        ldm = M.stride(0)
        B = batch or 1
        bs = 0 if bstride is None else bstride
    else:
        B, n, m = M.shape
        ldm = M.stride(1)
        bs = M.stride(0) if bstride is None else bstride
    if M.stride(-1) != 1:
        raise ValueError("cost rows must be contiguous")
    if tile is None:
        from . import autotune
        tile = autotune.lookup("f32" if code == _lib.F32 else "f64", n, m)
    T, cpt, rpb, pf = (tuple(tile) + (0,) * 4)[:4] if tile else (0, 0, 0, 0)
    ws_bytes = lib.potamd_dense_workspace_size(code, B, n, m, rpb)
    if ws_bytes == 0:
        raise _lib.LibraryError("invalid shape for the dense solver")
    ws = _device.workspace(ws_bytes, dev)
    u = u0.to(torch.float64).contiguous().clone()
    v = v0.to(torch.float64).contiguous().clone()
    err_cap = max(1, max_iter // 10 + 2)
    err_hist = torch.zeros(err_cap, dtype=torch.float64, device=dev)
    info = (ctypes.c_int * 4)()
    hc = _device.host_ctrl()
    if a.numel() != B * n or b.numel() != B * m or u.numel() != B * n or v.numel() != B * m:
        raise ValueError("weights and potentials must hold B x n and B x m values")
    niters = statuses = None
    if each:
        niters = torch.full((B,), -1, dtype=torch.int32, device=dev)
        statuses = torch.full((B,), -1, dtype=torch.int32, device=dev)
        rc = lib.potamd_sinkhorn_dense_each(
            code, domain, order, int(bool(kernel_given)),
            _ptr(M), ldm, bs, B, n, m, float(reg) if reg is not None else 1.0,
            _ptr(a), _ptr(b), _ptr(u), _ptr(v), int(max_iter), float(stop_thr),
            T, cpt, rpb, pf, _ptr(ws), ws_bytes, ctypes.c_void_p(hc.data_ptr()),
            ctypes.c_void_p(_device.stream_ptr(dev)), _ptr(niters), _ptr(statuses), info)
        _lib.check(rc, "potamd_sinkhorn_dense_each")
    else:
        rc = lib.potamd_sinkhorn_dense(
            code, domain, order, int(bool(kernel_given)),
            _ptr(M), ldm, bs, B, n, m, float(reg) if reg is not None else 1.0,
            _ptr(a), _ptr(b), _ptr(u), _ptr(v),
            int(max_iter), float(stop_thr), int(err_norm), _ptr(err_hist), err_cap,
            T, cpt, rpb, pf, _ptr(ws), ws_bytes, ctypes.c_void_p(hc.data_ptr()),
            ctypes.c_void_p(_device.stream_ptr(dev)), info)
        _lib.check(rc, "potamd_sinkhorn_dense")
    if rc == _lib.NEED_EXACT:
        raise _lib.LibraryError(
            "a column mass fell below the fp32 range at iteration %d; "
            "use float64 inputs" % info[0])
    n_err = 0 if each else min(info[1], err_cap)
    errs = err_hist[:n_err].cpu().tolist() if n_err else []
    return DenseResult(u=u.view(B, n), v=v.view(B, m), niter=info[0], errs=errs,
                       status=info[2], bad_iter=info[3], niters=niters, statuses=statuses)


def plan(M, u, v, reg, *, domain, kernel_given=False, a=None, b=None,
         want_plan=True, want_log=False, want_sums=True, batch=None, bstride=None,
         out_dtype=None):
    """Run potamd_plan_dense; returns (P, logP, sums) (None where not asked)."""
    lib = _lib.load()
    dev = M.device
    code = _lib.F32 if M.dtype == torch.float32 else _lib.F64
    if M.dim() == 2:
        n, m = M.shape
        ldm = M.stride(0)
        B = batch or 1
        bs = 0 if bstride is None else bstride
    else:
        B, n, m = M.shape
        ldm = M.stride(1)
        bs = M.stride(0) if bstride is None else bstride
    P = torch.empty((B, n, m), dtype=M.dtype, device=dev) if want_plan else None
    L = torch.empty((B, n, m), dtype=M.dtype, device=dev) if want_log else None
    sums = torch.empty((B, 4), dtype=torch.float64, device=dev) if want_sums else None
    ws_bytes = lib.potamd_plan_workspace_size(B, n)
    ws = _device.workspace(ws_bytes, dev)
    rc = lib.potamd_plan_dense(
        code, domain, int(bool(kernel_given)), _ptr(M), ldm, bs, B, n, m,
        float(reg) if reg is not None else 1.0,
        _ptr(u.contiguous()), _ptr(v.contiguous()),
        _ptr(a) if a is not None else None, _ptr(b) if b is not None else None,
        _ptr(P), _ptr(L), n * m, _ptr(sums), _ptr(ws), ws_bytes,
        ctypes.c_void_p(_device.stream_ptr(dev)))
    _lib.check(rc, "potamd_plan_dense")
    return P, L, sums


def sqeuclidean(X, Y, *, squared=True, clamp=True, zero_diag=False):
    """(B, n, d), (B, m, d) -> (B, n, m) via potamd_sqeuclidean."""
    lib = _lib.load()
    X = X.contiguous()
    Y = Y.contiguous()
    B, n, d = X.shape
    m = Y.shape[1]
    M = torch.empty((B, n, m), dtype=X.dtype, device=X.device)
    code = _lib.F32 if X.dtype == torch.float32 else _lib.F64
    rc = lib.potamd_sqeuclidean(
        code, _ptr(X), _ptr(Y), B, n, m, d, n * d, m * d, _ptr(M), m, n * m,
        int(squared), int(clamp), int(zero_diag),
        ctypes.c_void_p(_device.stream_ptr(X.device)))
    _lib.check(rc, "potamd_sqeuclidean")
    return M


def gemm(A, Bm, C, *, alpha=1.0, beta=0.0, transA=False, transB=False):
    """``C = alpha op(A) op(B) + beta C`` per problem on the library's batched
    GEMM; A (P, M, K) or (P, K, M), B (P, K, N) or (P, N, K), C (P, M, N)
    contiguous, same float dtype."""
    lib = _lib.load()
    P, Mr, N = C.shape
    K = A.shape[1] if transA else A.shape[2]
    a_rows = A.shape[2] if transA else A.shape[1]
    b_k, b_cols = (Bm.shape[2], Bm.shape[1]) if transB else (Bm.shape[1], Bm.shape[2])
    if (A.dim() != 3 or Bm.dim() != 3 or A.shape[0] != P or Bm.shape[0] != P
            or a_rows != Mr or b_k != K or b_cols != N or C.stride(2) != 1
            or not (A.dtype == Bm.dtype == C.dtype)
            or C.dtype not in (torch.float32, torch.float64)
            or not (A.device == Bm.device == C.device)):
        raise ValueError(f"gemm: shapes {tuple(A.shape)} x {tuple(Bm.shape)} -> "
                         f"{tuple(C.shape)} (transA={transA}, transB={transB}) do not agree")
    sa = (A.stride(2), A.stride(1)) if transA else (A.stride(1), A.stride(2))
    sb = (Bm.stride(2), Bm.stride(1)) if transB else (Bm.stride(1), Bm.stride(2))
    code = _lib.F32 if C.dtype == torch.float32 else _lib.F64
    rc = lib.potamd_gemm_batched(
        code, P, Mr, N, K, float(alpha), ctypes.c_void_p(A.data_ptr()), A.stride(0),
        sa[0], sa[1], ctypes.c_void_p(Bm.data_ptr()), Bm.stride(0), sb[0], sb[1],
        float(beta), ctypes.c_void_p(C.data_ptr()), C.stride(0), C.stride(1),
        ctypes.c_void_p(_device.stream_ptr(C.device)))
    _lib.check(rc, "potamd_gemm_batched")
    return C


def bmm(A, Bm, *, transA=False, transB=False, alpha=1.0):
    """New ``alpha op(A) op(B)`` per problem on the library's batched GEMM."""
    rows = A.shape[2] if transA else A.shape[1]
    cols = Bm.shape[1] if transB else Bm.shape[2]
    C = torch.empty((A.shape[0], rows, cols), dtype=A.dtype, device=A.device)
    return gemm(A, Bm, C, alpha=alpha, transA=transA, transB=transB)
