"""Many uneven ``sinkhorn_log`` problems in one launch.

The reference batch (ot/batch/_utils.py:165-290) needs one shape, one ``reg``
and stops all problems together; this module generalises it: every problem
keeps its own n, m, reg, initial ``u`` and stop, and each one follows
``ot.bregman.sinkhorn_log`` exactly (ot/bregman/_sinkhorn.py:880-906), so a
problem's ``niter`` and final potentials are those of a lone
``sinkhorn_log`` call on it.

Costs are packed row-major into one float32 buffer (row stride rounded up to
4 for 16-byte loads).  A cost normalised by its maximum, ``M / max M``, is
kept unnormalised in memory and the maximum is folded into ``reg``
(``exp(-(M/max)/reg) = exp(-M/(reg*max))``).
"""

import ctypes
import warnings
from dataclasses import dataclass

import numpy as np
import torch

from . import _device, _lib

MAX_M = 1024
MAX_N = 1024


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _tuned(cost, ns, ms, d, waves, prefetch=0, occupancy=0):
    """The caller's batch configuration, or the autotuner's for this batch
    class when the caller leaves all of it at 0 (then the library default if
    the class was never tuned)."""
    if waves or prefetch or occupancy:
        return int(waves), int(prefetch), int(occupancy)
    from . import autotune
    cfg = autotune.lookup_batch(cost, ns, ms, d)
    return tuple(int(c) for c in cfg) if cfg else (0, 0, 0)


def _round4(x):
    return (x + 3) // 4 * 4


@dataclass
class RaggedProblems:
    """Packed device representation of B problems."""
    device: torch.device
    ns: np.ndarray          # (B,) int
    ms: np.ndarray
    M: torch.Tensor         # packed float32 costs
    moff: torch.Tensor      # (B,) int64
    ns_d: torch.Tensor      # (B,) int32
    ms_d: torch.Tensor
    ldms_d: torch.Tensor
    aoff: torch.Tensor      # (B,) int64 offsets into row arrays
    boff: torch.Tensor      # (B,) int64 offsets into column arrays
    a: torch.Tensor         # (sum n,) float64
    b: torch.Tensor         # (sum m,) float64
    regs: torch.Tensor      # (B,) float64 effective reg (max folded in)
    cmax: torch.Tensor      # (B,) float32 cost maxima (1 when not normalised)

    @property
    def batch(self):
        return len(self.ns)

    def cost(self, p):
        """Problem p's cost as stored (unnormalised), an (n, m) view."""
        n, m = int(self.ns[p]), int(self.ms[p])
        ld = _round4(m)
        o = int(self.moff[p].item())
        return self.M[o:o + n * ld].view(n, ld)[:, :m]


def _layout(ns, ms, device):
    ns = np.asarray(ns, dtype=np.int64)
    ms = np.asarray(ms, dtype=np.int64)
    if ms.max() > MAX_M or ns.max() > MAX_N:
        raise ValueError(f"ragged solver supports n <= {MAX_N}, m <= {MAX_M}")
    lds = (ms + 3) // 4 * 4
    sizes = ns * lds
    moff = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    aoff = np.concatenate([[0], np.cumsum(ns)[:-1]])
    boff = np.concatenate([[0], np.cumsum(ms)[:-1]])
    t = lambda x, dt: torch.tensor(x, dtype=dt, device=device)
    return (ns, ms, int(sizes.sum()), t(moff, torch.int64), t(ns, torch.int32),
            t(ms, torch.int32), t(lds, torch.int32), t(aoff, torch.int64),
            t(boff, torch.int64))


def _weights(ws, counts, device):
    if ws is None:
        counts = np.asarray(counts, dtype=np.int64)
        if np.all(counts == counts[0]):
            # one size: a single fill with the same value 1 / n
            return torch.full((int(counts.sum()),), 1.0 / float(counts[0]), dtype=torch.float64,
                              device=device)
        inv = torch.from_numpy(1.0 / counts.astype(np.float64)).to(device)
        return torch.repeat_interleave(inv, torch.from_numpy(counts).to(device),
                                       output_size=int(counts.sum()))
    return torch.cat([torch.as_tensor(np.asarray(w), dtype=torch.float64).reshape(-1)
                      for w in ws]).to(device)


def _pack_points(xs, device):
    """(sum n, d) float32 device tensor from a list of clouds or a (B, n, d) tensor."""
    if isinstance(xs, torch.Tensor):
        return xs.to(device=device, dtype=torch.float32).reshape(-1, xs.shape[-1]).contiguous()
    return torch.cat([x.to(device=device, dtype=torch.float32) if isinstance(x, torch.Tensor)
                      else torch.as_tensor(np.asarray(x, dtype=np.float32)).to(device)
                      for x in xs]).contiguous()


def from_clouds(xs, ys, regs, a=None, b=None, normalize=True, device=None):
    """Costs ``|x - y|^2`` (clamped at 0) of point-cloud pairs, built on the GPU.

    xs[p]: (n_p, d), ys[p]: (m_p, d) float arrays or tensors (or (B, n, d)
    tensors for equal sizes); regs: (B,).
    With ``normalize`` each cost is taken relative to its maximum.
    """
    device = device or _device.require_gpu()
    lib = _lib.load()
    if isinstance(xs, torch.Tensor) and isinstance(ys, torch.Tensor):
        ns = [int(xs.shape[1])] * int(xs.shape[0])
        ms = [int(ys.shape[1])] * int(ys.shape[0])
    else:
        ns = [len(x) for x in xs]
        ms = [len(y) for y in ys]
    d = int(xs[0].shape[1])
    ns, ms, total, moff, ns_d, ms_d, ld_d, aoff, boff = _layout(ns, ms, device)
    X = _pack_points(xs, device)
    Y = _pack_points(ys, device)
    M = torch.empty(total, dtype=torch.float32, device=device)
    cmax = torch.empty(len(ns), dtype=torch.float32, device=device)
    rc = lib.potamd_sqeuclidean_ragged(
        len(ns), int(ns.max()), int(ms.max()), d, _ptr(X), _ptr(Y), _ptr(aoff), _ptr(boff),
        _ptr(ns_d), _ptr(ms_d), _ptr(M), _ptr(moff), _ptr(ld_d), _ptr(cmax),
        ctypes.c_void_p(_device.stream_ptr(device)))
    _lib.check(rc, "potamd_sqeuclidean_ragged")
    regs_t = torch.as_tensor(np.asarray(regs, dtype=np.float64)).to(device)
    if normalize:
        regs_eff = regs_t * cmax.to(torch.float64)
    else:
        cmax.fill_(1.0)
        regs_eff = regs_t
    return RaggedProblems(device=device, ns=ns, ms=ms, M=M, moff=moff, ns_d=ns_d, ms_d=ms_d,
                          ldms_d=ld_d, aoff=aoff, boff=boff, a=_weights(a, ns, device),
                          b=_weights(b, ms, device), regs=regs_eff, cmax=cmax)


def from_costs(Ms, regs, a=None, b=None, device=None):
    """Pack given cost matrices (float32) without normalisation."""
    device = device or _device.require_gpu()
    ns = [int(np.shape(M)[0]) for M in Ms]
    ms = [int(np.shape(M)[1]) for M in Ms]
    ns, ms, total, moff, ns_d, ms_d, ld_d, aoff, boff = _layout(ns, ms, device)
    host = np.zeros(total, dtype=np.float32)
    off = 0
    for M, n, m in zip(Ms, ns, ms):
        ld = _round4(int(m))
        blk = np.zeros((n, ld), dtype=np.float32)
        blk[:, :m] = np.asarray(M, dtype=np.float32)
        host[off:off + n * ld] = blk.ravel()
        off += n * ld
    Mt = torch.from_numpy(host).to(device)
    regs_t = torch.as_tensor(np.asarray(regs, dtype=np.float64)).to(device)
    return RaggedProblems(device=device, ns=ns, ms=ms, M=Mt, moff=moff, ns_d=ns_d, ms_d=ms_d,
                          ldms_d=ld_d, aoff=aoff, boff=boff, a=_weights(a, ns, device),
                          b=_weights(b, ms, device), regs=regs_t,
                          cmax=torch.ones(len(ns), dtype=torch.float32, device=device))


@dataclass
class RaggedResult:
    u: torch.Tensor          # (sum n,) float64 log potentials
    v: torch.Tensor          # (sum m,) float64
    niter: torch.Tensor      # (B,) int32, the reference's `ii` at the stop
    status: torch.Tensor     # (B,) int32: 0 converged, 1 max-iter
    err_first: torch.Tensor  # (B,) float64 marginal error at ii = 0
    err_last: torch.Tensor   # (B,) float64 last checked marginal error
    cmax: torch.Tensor = None  # (B,) float64 cost maxima (point-cloud solver)

    def split(self, R, p):
        ao, bo = int(R.aoff[p].item()), int(R.boff[p].item())
        return self.u[ao:ao + int(R.ns[p])], self.v[bo:bo + int(R.ms[p])]


def launch_order(R, work=None):
    """Problems by decreasing expected work (default n*m/reg), longest first."""
    if work is None:
        regs = R.regs.double().cpu().numpy() / R.cmax.double().cpu().numpy()
        work = R.ns.astype(np.float64) * R.ms / regs
    order = np.argsort(-np.asarray(work), kind="stable").astype(np.int32)
    return torch.tensor(order, device=R.device)


def sinkhorn_log_ragged(R, u0=None, numItermax=1000, stopThr=1e-9, order=None,
                        fallback_thr=0.0, waves=0, prefetch=0, check_every=10, occupancy=0,
                        rotate=0):
    """Solve every problem of ``R`` with sinkhorn_log semantics in one launch.

    u0: (sum n,) float64 initial log potentials (None = zeros, the cold start).
    fallback_thr > 0: a problem whose first checked error exceeds it restarts
    cold (its status gets bit 4).  check_every: iterations between checks (10
    is the reference's cadence and keeps ``niter`` identical to a lone
    sinkhorn_log call; 1 stops at the first iteration meeting stopThr).  Returns a RaggedResult; the call is
    asynchronous on the current stream.
    """
    lib = _lib.load()
    dev = R.device
    B = R.batch
    total_n = int(R.ns.sum())
    u = torch.zeros(total_n, dtype=torch.float64, device=dev) if u0 is None \
        else u0.to(device=dev, dtype=torch.float64).contiguous().clone()
    v = torch.empty(int(R.ms.sum()), dtype=torch.float64, device=dev)
    niter = torch.empty(B, dtype=torch.int32, device=dev)
    status = torch.empty(B, dtype=torch.int32, device=dev)
    e_first = torch.empty(B, dtype=torch.float64, device=dev)
    e_last = torch.empty(B, dtype=torch.float64, device=dev)
    if order is None:
        order = launch_order(R)
    waves, prefetch, occupancy = _tuned("dense", R.ns, R.ms, 0, waves, prefetch, occupancy)
    rc = lib.potamd_sinkhorn_log_batch(
        _lib.F32, B, int(R.ns.max()), int(R.ms.max()), _ptr(R.M), _ptr(R.moff), _ptr(R.ns_d),
        _ptr(R.ms_d), _ptr(R.ldms_d), _ptr(R.aoff), _ptr(R.boff), _ptr(R.regs), _ptr(R.a),
        _ptr(R.b), _ptr(u), _ptr(v), int(numItermax), int(check_every), float(stopThr),
        float(fallback_thr),
        _ptr(order), 1, _ptr(niter), _ptr(status), _ptr(e_last), _ptr(e_first),
        int(waves) | (int(prefetch) << 8) | (int(occupancy) << 16) | (int(bool(rotate)) << 24),
        ctypes.c_void_p(_device.stream_ptr(dev)))
    _lib.check(rc, "potamd_sinkhorn_log_batch")
    return RaggedResult(u=u, v=v, niter=niter, status=status, err_first=e_first,
                        err_last=e_last)


MAX_D_POINTS = 16
SPLIT_MAX = 8  # workgroups per split problem (the kernel's exchange slots hold 8 parts)


@dataclass
class RaggedPoints:
    """Packed point clouds of B problems for the on-the-fly cost solver."""
    device: torch.device
    ns: np.ndarray          # (B,) int
    ms: np.ndarray
    d: int
    X: torch.Tensor         # (sum n, d) float32
    Y: torch.Tensor         # (sum m, d) float32
    aoff: torch.Tensor      # (B,) int64 row offsets (points, a and u)
    boff: torch.Tensor      # (B,) int64 column offsets (points, b and v)
    ns_d: torch.Tensor      # (B,) int32
    ms_d: torch.Tensor
    a: torch.Tensor         # (sum n,) float64
    b: torch.Tensor         # (sum m,) float64
    regs: torch.Tensor      # (B,) float64 reg (on the normalised cost when normalize)
    normalize: bool
    regs_host: np.ndarray = None  # (B,) float64 host copy of regs (launch order)
    cmax: torch.Tensor = None  # (B,) float64 cost maxima once found (points_cmax)
    order: torch.Tensor = None  # (B,) int32 longest-first launch order, on the device

    @property
    def batch(self):
        return len(self.ns)


def points_cmax(P):
    """Per-problem maximum of |x - y|^2 (float64, on the device), found on the
    GPU without solving; also stored in ``P.cmax``."""
    lib = _lib.load()
    cmax = torch.empty(P.batch, dtype=torch.float64, device=P.device)
    rc = lib.potamd_ragged_points_cmax(
        P.batch, int(P.ns.max()), int(P.ms.max()), P.d, _ptr(P.X), _ptr(P.Y), _ptr(P.aoff),
        _ptr(P.boff), _ptr(P.ns_d), _ptr(P.ms_d), _ptr(cmax),
        ctypes.c_void_p(_device.stream_ptr(P.device)))
    _lib.check(rc, "potamd_ragged_points_cmax")
    P.cmax = cmax
    return cmax


def points_from_clouds(xs, ys, regs, a=None, b=None, normalize=True, device=None):
    """Pack point-cloud pairs for ``sinkhorn_log_ragged_points``.

    xs[p]: (n_p, d), ys[p]: (m_p, d) arrays or tensors (or (B, n, d) tensors
    for equal sizes), n_p, m_p <= 1024, d <= 16; regs: (B,), or one value
    for every problem (any other length raises ValueError).  With
    ``normalize`` each problem's cost is |x - y|^2 / max |x - y|^2.
    """
    device = device or _device.require_gpu()
    if isinstance(xs, torch.Tensor) and isinstance(ys, torch.Tensor):
        ns = np.full(int(xs.shape[0]), int(xs.shape[1]), dtype=np.int64)
        ms = np.full(int(ys.shape[0]), int(ys.shape[1]), dtype=np.int64)
    else:
        ns = [len(x) for x in xs]
        ms = [len(y) for y in ys]
    d = int(xs[0].shape[-1])
    if d > MAX_D_POINTS:
        raise ValueError(f"the point-cloud solver supports d <= {MAX_D_POINTS}")
    ns = np.asarray(ns, dtype=np.int64)
    ms = np.asarray(ms, dtype=np.int64)
    if ns.min() < 1 or ms.min() < 1:
        raise ValueError("every problem needs at least one point on each side")
    if ms.max() > MAX_M or ns.max() > MAX_N:
        raise ValueError(f"ragged solver supports n <= {MAX_N}, m <= {MAX_M}")
    B = len(ns)
    aoff = np.concatenate([[0], np.cumsum(ns)[:-1]])
    boff = np.concatenate([[0], np.cumsum(ms)[:-1]])
    regs_h = np.asarray(regs, dtype=np.float64).reshape(-1)
    if regs_h.size == 1:
        regs_h = np.full(B, regs_h[0])
    if regs_h.size != B:
        raise ValueError(f"regs has {regs_h.size} entries for {B} problems")
    # The host-built tables (offsets, regs, sizes, and the launch order, which
    # an upload at solve time would make wait for the work already queued on
    # the stream, such as a warm start's network) go up in one copy of one
    # byte buffer, viewed on the device in their own types (no conversion
    # kernels), from pinned memory without a host wait.
    raw_t = torch.empty(3 * B * 8 + 3 * B * 4, dtype=torch.uint8, pin_memory=True)
    raw = raw_t.numpy()
    h64 = raw[:24 * B].view(np.int64)
    h64[0:B] = aoff
    h64[B:2 * B] = boff
    h64[2 * B:3 * B] = regs_h.view(np.int64)
    h32 = raw[24 * B:].view(np.int32)
    h32[0:B] = ns
    h32[B:2 * B] = ms
    h32[2 * B:3 * B] = _points_order(ns, ms, regs_h)
    tab = raw_t.to(device, non_blocking=True)
    t64 = tab[:24 * B].view(torch.int64)
    t32 = tab[24 * B:].view(torch.int32)
    return RaggedPoints(device=device, ns=ns, ms=ms, d=d, X=_pack_points(xs, device),
                        Y=_pack_points(ys, device), aoff=t64[0:B], boff=t64[B:2 * B],
                        ns_d=t32[0:B], ms_d=t32[B:2 * B], a=_weights(a, ns, device),
                        b=_weights(b, ms, device), regs=t64[2 * B:3 * B].view(torch.float64),
                        normalize=bool(normalize), regs_host=regs_h, order=t32[2 * B:3 * B])


def _check_order(order, B):
    """A launch order must be a permutation of range(B) (it indexes problems)."""
    o = order.cpu().numpy() if isinstance(order, torch.Tensor) else np.asarray(order)
    o = o.astype(np.int64).reshape(-1)
    if o.shape != (B,) or not np.array_equal(np.sort(o), np.arange(B)):
        raise ValueError("order must be a permutation of the problem indices")
    return o.astype(np.int32)


def _points_order(ns, ms, regs):
    work = ns.astype(np.float64) * ms / np.asarray(regs, dtype=np.float64)
    return np.argsort(-work, kind="stable").astype(np.int32)


def split_plan(ns, ms, cus=None, regs=None):
    """Workgroups per problem for a batch too small to fill the GPU with one
    workgroup per problem (at most 2 problems per CU): 8 for problems of at
    least 3 * 2^18 elements (n >= 256), 4 from 3 * 2^17 (n >= 128), 2 from
    2^17 (n >= 64), else 1.  Larger batches are not split (their work already
    fills the GPU).

    The parts of all split problems together never exceed the number of CUs
    (the persistent grid).  Over that budget, 8-part problems fall back to 4
    first (the eighth-size parts spend most of each iteration exchanging
    columns: about 3x the speed of one workgroup on 8 CUs, against 2.5x on
    4), smallest first, then the smallest split problems give up parts.
    ``regs`` (optional): sizes are then weighed as n m / reg, a proxy for the
    iterations a problem needs (the automatic split leaves it out: a graph
    is planned before it sees the regs, and an eager solve of the same batch
    must split alike).  With the split problems queued first
    (``_split_items``), every part is taken by the first wave of workgroups,
    so no part waits for a partner that a busy workgroup has not dequeued
    yet."""
    ns = np.asarray(ns, dtype=np.int64)
    ms = np.asarray(ms, dtype=np.int64)
    parts = np.ones(len(ns), dtype=np.int32)
    if cus is None:
        cus = torch.cuda.get_device_properties(_device.require_gpu()).multi_processor_count
    if len(ns) > 2 * cus:
        return parts
    nm = ns * ms
    parts[(nm >= 1 << 17) & (ns >= 64)] = 2
    parts[(nm >= 3 << 17) & (ns >= 128)] = 4
    parts[(nm >= 3 << 18) & (ns >= 256)] = SPLIT_MAX
    weight = nm.astype(np.float64)
    if regs is not None:
        weight = weight / np.broadcast_to(np.asarray(regs, dtype=np.float64), weight.shape)
    by_size = np.argsort(weight, kind="stable")
    while parts[parts > 1].sum() > cus:
        eight = by_size[parts[by_size] == SPLIT_MAX]
        k = eight[0] if len(eight) else by_size[np.nonzero(parts[by_size] > 1)[0][0]]
        parts[k] = parts[k] // 2
    return parts


def _split_tables(parts):
    """split[p] = parts | slot << 4 for the split problems (0 otherwise) and
    the number of slots."""
    split = np.zeros(len(parts), dtype=np.int32)
    big = np.nonzero(parts > 1)[0]
    split[big] = parts[big] | (np.arange(len(big), dtype=np.int32) << 4)
    return split, len(big)


def _split_items(order, parts):
    """Queue items: the split problems first, then the others, each group in
    launch order, each problem as its parts in a row."""
    order = np.asarray(order, dtype=np.int64)
    big = parts[order] > 1
    order = np.concatenate([order[big], order[~big]])
    reps = parts[order]
    probs = np.repeat(order, reps)
    first = np.repeat(np.cumsum(reps) - reps, reps)
    part = np.arange(len(probs)) - first
    return (probs | (part << 24)).astype(np.int32)


def _warn_unsolved(status):
    """A problem left with status 2 has NaN potentials; say so instead of
    returning NaNs quietly."""
    bad = int(((status & 3) == 2).sum().item())
    if bad:
        warnings.warn(f"{bad} problem(s) of the batch were not solved (status 2: a cost "
                      "maximum of 0, whose normalised cost is NaN as in the reference); "
                      "their potentials are NaN")


def _resolve_unsolved(P, res, u0, numItermax, stopThr):
    """Solve again, one at a time on the float64 dense GPU solver, the problems
    the point-cloud kernel reported as not solved (status 2): scaled
    coordinates outside the f16 operand range (small reg on an unnormalised
    cost), or split parts that gave up waiting for each other.  Each runs the
    reference's sinkhorn_log on ot.dist(x, y) (divided by its maximum when
    P.normalize; ot/bregman/_sinkhorn.py:799-918, the marginal checked every
    10 iterations) from the same start; its potentials, niter, errors and
    status replace the entry.  A problem whose cost maximum is 0 stays NaN
    (its normalised cost is 0/0, NaN in the reference too).  Synchronises the
    host with the stream once, to read the statuses."""
    st = res.status.cpu().numpy()
    bad = np.nonzero((st & 3) == 2)[0]
    if len(bad) == 0:
        return
    from . import _solve, bregman
    d = P.d  # P.X, P.Y: (sum n, d), (sum m, d)
    aoff = P.aoff.cpu().numpy()
    boff = P.boff.cpu().numpy()
    regs = P.regs_host if P.regs_host is not None else P.regs.cpu().numpy()
    for p in bad:
        n, m = int(P.ns[p]), int(P.ms[p])
        a0, b0 = int(aoff[p]), int(boff[p])
        x = P.X[a0:a0 + n].reshape(1, n, d).to(torch.float64)
        y = P.Y[b0:b0 + m].reshape(1, m, d).to(torch.float64)
        M = _solve.sqeuclidean(x, y)[0]
        mx = float(M.max().item())
        if P.normalize:
            if not mx > 0.0:
                continue
            M = M / mx
        ws = None
        if u0 is not None:
            ws = (u0[a0:a0 + n].to(torch.float64),
                  torch.zeros(m, dtype=torch.float64, device=M.device))
        _, lg = bregman.sinkhorn_log(P.a[a0:a0 + n], P.b[b0:b0 + m], M, float(regs[p]),
                                     numItermax=numItermax, stopThr=stopThr, log=True,
                                     warn=False, warmstart=ws)
        res.u[a0:a0 + n] = torch.as_tensor(lg["log_u"], device=M.device)
        res.v[b0:b0 + m] = torch.as_tensor(lg["log_v"], device=M.device)
        errs = [float(e) for e in lg["err"]]
        res.niter[p] = int(lg["niter"])
        res.status[p] = 0 if (errs and errs[-1] < stopThr) else 1
        res.err_first[p] = errs[0] if errs else float("nan")
        res.err_last[p] = errs[-1] if errs else float("nan")
        res.cmax[p] = mx if P.normalize else 1.0


def _check_parts(parts, ns):
    parts = np.asarray(parts, dtype=np.int32).reshape(-1)
    if len(parts) != len(ns) or parts.min() < 1 or parts.max() > SPLIT_MAX:
        raise ValueError(f"split: one part count in 1..{SPLIT_MAX} per problem")
    if np.any(np.asarray(ns) < 32 * parts):
        raise ValueError("split: each part needs at least 32 rows (n >= 32 * parts)")
    return parts


def _launch_points(lib, P, X, Y, regs, u, v, niter, status, e_last, e_first, cmax, order,
                   numItermax, check_every, stopThr, fallback_thr, stream, cmax_given=False,
                   waves=0, queue=None):
    mode = (2 if cmax_given else 1) if P.normalize else 0
    rc = lib.potamd_sinkhorn_log_batch_points(
        P.batch, int(P.ns.max()), int(P.ms.max()), P.d, _ptr(X), _ptr(Y), _ptr(P.aoff),
        _ptr(P.boff), _ptr(P.ns_d), _ptr(P.ms_d), _ptr(regs), mode, _ptr(P.a),
        _ptr(P.b), _ptr(u), _ptr(v), int(numItermax), int(check_every), float(stopThr),
        float(fallback_thr), _ptr(order), _ptr(niter), _ptr(status), _ptr(e_last),
        _ptr(e_first), _ptr(cmax), int(waves), _ptr(queue), stream)
    _lib.check(rc, "potamd_sinkhorn_log_batch_points")


def _launch_points_split(lib, P, X, Y, regs, u, v, niter, status, e_last, e_first, cmax, items,
                         split, nsplit, sws, numItermax, check_every, stopThr, fallback_thr,
                         stream, queue, cmax_given=False):
    mode = (2 if cmax_given else 1) if P.normalize else 0
    rc = lib.potamd_sinkhorn_log_batch_points_split(
        P.batch, int(P.ns.max()), int(P.ms.max()), P.d, _ptr(X), _ptr(Y), _ptr(P.aoff),
        _ptr(P.boff), _ptr(P.ns_d), _ptr(P.ms_d), _ptr(regs), mode, _ptr(P.a),
        _ptr(P.b), _ptr(u), _ptr(v), int(numItermax), int(check_every), float(stopThr),
        float(fallback_thr), int(items.numel()), _ptr(items), _ptr(split), int(nsplit),
        _ptr(sws), sws.numel(), _ptr(niter), _ptr(status), _ptr(e_last), _ptr(e_first),
        _ptr(cmax), _ptr(queue), stream)
    _lib.check(rc, "potamd_sinkhorn_log_batch_points_split")


def _split_workspace(lib, nsplit, dev):
    """Zeroed exchange slots, 256-byte aligned (the kernel leaves them zero)."""
    nbytes = int(lib.potamd_points_split_workspace_size(int(nsplit)))
    buf = torch.zeros(nbytes + 256, dtype=torch.uint8, device=dev)
    off = (-buf.data_ptr()) % 256
    return buf[off:off + nbytes]


def sinkhorn_log_ragged_points(P, u0=None, numItermax=1000, stopThr=1e-9, order=None,
                               fallback_thr=0.0, check_every=10, cmax=None, waves=0,
                               split="auto", unsolved="dense"):
    """``sinkhorn_log_ragged`` on point clouds, the cost computed on the fly.

    P: a RaggedPoints.  Same loop, stop rule, warm start and fallback as
    ``sinkhorn_log_ragged``; the cost of problem p is |x - y|^2 (divided by
    its maximum when P.normalize), evaluated on MFMA from the points in every
    sweep and never stored.  ``cmax``: the maxima from ``points_cmax`` (the
    solve then skips its own search).  ``waves``: wavefronts per problem (8
    or 16; 0 = the autotuned value of this batch class, else the library
    default).  ``split``: "auto" (``split_plan``: only batches too small to
    fill the GPU are split), None (one workgroup per problem) or workgroups
    per problem (1..8): a split problem's rows are shared by several
    workgroups (8 waves each).  ``unsolved``: "dense" (default) solves the
    problems the kernel cannot take (status 2: scaled coordinates outside its
    f16 operand range, or split parts that gave up waiting) again on the
    float64 dense solver (``_resolve_unsolved``; reads the statuses back, one
    host synchronisation); "nan" leaves them NaN with status 2 and returns
    without synchronising.  Returns a RaggedResult whose ``cmax`` holds the
    cost maxima.
    """
    if unsolved not in ("dense", "nan"):
        raise ValueError(f"unsolved must be 'dense' or 'nan', not {unsolved!r}")
    lib = _lib.load()
    dev = P.device
    B = P.batch
    u = torch.zeros(int(P.ns.sum()), dtype=torch.float64, device=dev) if u0 is None \
        else u0.to(device=dev, dtype=torch.float64).contiguous().clone()
    v = torch.empty(int(P.ms.sum()), dtype=torch.float64, device=dev)
    niter = torch.empty(B, dtype=torch.int32, device=dev)
    status = torch.empty(B, dtype=torch.int32, device=dev)
    e_first = torch.empty(B, dtype=torch.float64, device=dev)
    e_last = torch.empty(B, dtype=torch.float64, device=dev)
    given = cmax is not None and P.normalize
    cmax = cmax.to(device=dev, dtype=torch.float64).contiguous() if given \
        else torch.empty(B, dtype=torch.float64, device=dev)
    if order is None and P.order is not None:
        order = P.order
    elif order is None:
        regs_h = P.regs_host if P.regs_host is not None else P.regs.cpu().numpy()
        order = torch.from_numpy(_points_order(P.ns, P.ms, regs_h)).to(dev)
    else:
        order = torch.as_tensor(_check_order(order, B), device=dev)
    waves = _tuned("points", P.ns, P.ms, P.d, waves)[0]
    # work-queue counters of the persistent launch (the kernel leaves them zero)
    queue = torch.zeros(2, dtype=torch.int32, device=dev)
    parts = None
    if split is not None:
        # (sizes only, as RaggedGraph plans before it sees the regs, so the
        # eager solve and a graph replay of one batch split alike and agree
        # bit for bit)
        parts = split_plan(P.ns, P.ms) if (isinstance(split, str) and split == "auto") \
            else _check_parts(split, P.ns)
    if parts is not None and (parts > 1).any():
        table, nsplit = _split_tables(parts)
        items = torch.from_numpy(_split_items(order.cpu().numpy(), parts)).to(dev)
        sws = _split_workspace(lib, nsplit, dev)
        _launch_points_split(lib, P, P.X, P.Y, P.regs, u, v, niter, status, e_last, e_first,
                             cmax, items, torch.from_numpy(table).to(dev), nsplit, sws,
                             numItermax, check_every, stopThr, fallback_thr,
                             ctypes.c_void_p(_device.stream_ptr(dev)), queue, cmax_given=given)
    else:
        _launch_points(lib, P, P.X, P.Y, P.regs, u, v, niter, status, e_last, e_first, cmax,
                       order, numItermax, check_every, stopThr, fallback_thr,
                       ctypes.c_void_p(_device.stream_ptr(dev)), cmax_given=given, waves=waves,
                       queue=queue)
    res = RaggedResult(u=u, v=v, niter=niter, status=status, err_first=e_first,
                       err_last=e_last, cmax=cmax)
    if unsolved == "dense":
        _resolve_unsolved(P, res, u0, numItermax, stopThr)
        _warn_unsolved(status)
    return res


class RaggedGraph:
    """Cost construction and solve of a fixed ragged layout, captured once as a
    hipGraph (through ``torch.cuda.CUDAGraph``) and replayed per batch.

    The layout (sizes, d, iteration settings) is fixed at construction; each
    ``solve`` copies new points, regs and (optionally) initial potentials into
    the graph's static buffers, sets the launch order, and replays the
    kernels: with ``cost="dense"`` the ragged squared distances with
    per-problem maxima, then the batched sinkhorn_log over them; with
    ``cost="points"`` the one point-cloud kernel that evaluates the cost on
    the fly.  The captured work contains no memset nodes and no
    host synchronisation.  Results are views of static buffers, overwritten
    by the next ``solve``.
    """

    def __init__(self, ns, ms, d, numItermax=1000, stopThr=1e-9, check_every=10,
                 fallback_thr=0.0, waves=0, prefetch=0, occupancy=0, device=None,
                 cost="dense", split="auto"):
        if cost not in ("dense", "points"):
            raise ValueError("cost must be 'dense' or 'points'")
        self.cost = cost
        self.normalize = True
        self.device = device or _device.require_gpu()
        dev = self.device
        self.lib = _lib.load()
        (self.ns, self.ms, total, self.moff, self.ns_d, self.ms_d, self.ld_d, self.aoff,
         self.boff) = _layout(ns, ms, dev)
        B = len(self.ns)
        self.d = int(d)
        tn, tm = int(self.ns.sum()), int(self.ms.sum())
        self.X = torch.zeros((tn, self.d), dtype=torch.float32, device=dev)
        self.Y = torch.zeros((tm, self.d), dtype=torch.float32, device=dev)
        if cost == "points" and self.d > MAX_D_POINTS:
            raise ValueError(f"the point-cloud solver supports d <= {MAX_D_POINTS}")
        self.M = torch.empty(total if cost == "dense" else 1, dtype=torch.float32, device=dev)
        self.cmax = torch.empty(B, dtype=torch.float32 if cost == "dense" else torch.float64,
                                device=dev)
        self.reg_in = torch.ones(B, dtype=torch.float64, device=dev)
        self.reg_eff = torch.ones(B, dtype=torch.float64, device=dev)
        self.a = _weights(None, self.ns, dev)
        self.b = _weights(None, self.ms, dev)
        self.u0 = torch.zeros(tn, dtype=torch.float64, device=dev)
        self.u = torch.zeros(tn, dtype=torch.float64, device=dev)
        self.v = torch.zeros(tm, dtype=torch.float64, device=dev)
        self.niter = torch.zeros(B, dtype=torch.int32, device=dev)
        self.status = torch.zeros(B, dtype=torch.int32, device=dev)
        self.e_first = torch.zeros(B, dtype=torch.float64, device=dev)
        self.e_last = torch.zeros(B, dtype=torch.float64, device=dev)
        self.order = torch.arange(B, dtype=torch.int32, device=dev)
        self.queue = torch.zeros(2, dtype=torch.int32, device=dev)  # left zero by each launch
        # Split problems (cost="points"): parts per problem fixed by the sizes,
        # the queue items re-ordered on each solve, the exchange slots zero.
        self.parts = None
        if split is not None and cost == "points":
            parts = split_plan(self.ns, self.ms) if (isinstance(split, str) and split == "auto") \
                else _check_parts(split, self.ns)
            if (parts > 1).any():
                self.parts = parts
                table, self.nsplit = _split_tables(parts)
                self.split = torch.from_numpy(table).to(dev)
                self.sws = _split_workspace(self.lib, self.nsplit, dev)
                self.items = torch.from_numpy(
                    _split_items(np.arange(B), parts)).to(dev)
        waves, prefetch, occupancy = _tuned(cost, self.ns, self.ms,
                                            self.d if cost == "points" else 0,
                                            waves, prefetch, occupancy)
        self.settings = (int(numItermax), int(check_every), float(stopThr), float(fallback_thr),
                         int(waves) | (int(prefetch) << 8) | (int(occupancy) << 16))
        # eager warm-up on a side stream, then capture
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream(dev))
        with torch.cuda.stream(side):
            self._enqueue()
        torch.cuda.current_stream(dev).wait_stream(side)
        torch.cuda.synchronize(dev)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._enqueue()
        torch.cuda.synchronize(dev)

    @property
    def batch(self):
        return len(self.ns)

    def _enqueue(self):
        dev = self.device
        stream = ctypes.c_void_p(_device.stream_ptr(dev))
        if self.cost == "points":
            self.u.copy_(self.u0)
            max_iter, check_every, stop, fb, wv = self.settings
            if self.parts is not None:
                _launch_points_split(self.lib, self, self.X, self.Y, self.reg_in, self.u,
                                     self.v, self.niter, self.status, self.e_last,
                                     self.e_first, self.cmax, self.items, self.split,
                                     self.nsplit, self.sws, max_iter, check_every, stop, fb,
                                     stream, self.queue)
                return
            _launch_points(self.lib, self, self.X, self.Y, self.reg_in, self.u, self.v,
                           self.niter, self.status, self.e_last, self.e_first, self.cmax,
                           self.order, max_iter, check_every, stop, fb, stream,
                           waves=wv & 0xff, queue=self.queue)
            return
        rc = self.lib.potamd_sqeuclidean_ragged(
            len(self.ns), int(self.ns.max()), int(self.ms.max()), self.d, _ptr(self.X),
            _ptr(self.Y), _ptr(self.aoff), _ptr(self.boff), _ptr(self.ns_d), _ptr(self.ms_d),
            _ptr(self.M), _ptr(self.moff), _ptr(self.ld_d), _ptr(self.cmax), stream)
        _lib.check(rc, "potamd_sqeuclidean_ragged")
        torch.mul(self.reg_in, self.cmax.to(torch.float64), out=self.reg_eff)
        self.u.copy_(self.u0)
        max_iter, check_every, stop, fb, wv = self.settings
        rc = self.lib.potamd_sinkhorn_log_batch(
            _lib.F32, len(self.ns), int(self.ns.max()), int(self.ms.max()), _ptr(self.M),
            _ptr(self.moff), _ptr(self.ns_d), _ptr(self.ms_d), _ptr(self.ld_d), _ptr(self.aoff),
            _ptr(self.boff), _ptr(self.reg_eff), _ptr(self.a), _ptr(self.b), _ptr(self.u),
            _ptr(self.v), max_iter, check_every, stop, fb, _ptr(self.order), 1,
            _ptr(self.niter), _ptr(self.status), _ptr(self.e_last), _ptr(self.e_first), wv,
            stream)
        _lib.check(rc, "potamd_sinkhorn_log_batch")

    def solve(self, xs, ys, regs, u0=None, order=None):
        """Replay on new data: xs, ys packed (sum n, d) / (sum m, d) or (B, n, d)
        tensors; regs (B,) host values.  Returns a RaggedResult of views."""
        self.X.copy_(_pack_points(xs, self.device).view_as(self.X))
        self.Y.copy_(_pack_points(ys, self.device).view_as(self.Y))
        regs = np.asarray(regs, dtype=np.float64)
        self.reg_in.copy_(torch.from_numpy(regs))
        if u0 is None:
            self.u0.zero_()
        else:
            self.u0.copy_(u0)
        if order is None:
            order = _points_order(self.ns, self.ms, regs)
        else:
            order = _check_order(order, len(self.ns))
        self.order.copy_(torch.as_tensor(order))
        if self.parts is not None:
            self.items.copy_(torch.from_numpy(_split_items(order, self.parts)))
        self.graph.replay()
        if self.parts is not None:
            _warn_unsolved(self.status)
        return RaggedResult(u=self.u, v=self.v, niter=self.niter, status=self.status,
                            err_first=self.e_first, err_last=self.e_last,
                            cmax=self.cmax if self.cost == "points" else None)
