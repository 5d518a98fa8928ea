"""The no-allocation entry point (potamd_sinkhorn_log_batch_points_split_noalloc)
against the split point-cloud entry point it mirrors, through the C-ABI: two
calls in a row on one GPU, the second with a smaller batch, each equal bit
for bit to potamd_sinkhorn_log_batch_points_split on the same inputs."""

import ctypes

import numpy as np
import pytest
import torch

from pot_amd import _device, _lib, ragged

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _clouds(seed, sizes, d=16):
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    for n, m in sizes:
        xs.append(rng.standard_normal((n, d)).astype(np.float32))
        ys.append((rng.standard_normal((m, d)) * 1.3 + 0.4).astype(np.float32))
    regs = np.exp(rng.uniform(np.log(5e-3), np.log(5e-2), len(sizes)))
    return xs, ys, regs


def _outputs(P, dev):
    B = P.batch
    u = torch.zeros(int(P.ns.sum()), dtype=torch.float64, device=dev)
    v = torch.full((int(P.ms.sum()),), -7.0, dtype=torch.float64, device=dev)
    niter = torch.full((B,), -1, dtype=torch.int32, device=dev)
    status = torch.full((B,), -1, dtype=torch.int32, device=dev)
    e_last = torch.zeros(B, dtype=torch.float64, device=dev)
    e_first = torch.zeros(B, dtype=torch.float64, device=dev)
    cmax = torch.zeros(B, dtype=torch.float64, device=dev)
    return u, v, niter, status, e_last, e_first, cmax


def _solve_both(lib, P, dev, parts, max_iter=2000, stop=1e-6, check_every=10):
    order = P.order.cpu().numpy()
    table, nsplit = ragged._split_tables(parts)
    items = torch.from_numpy(ragged._split_items(order, parts)).to(dev)
    split = torch.from_numpy(table).to(dev)
    stream = ctypes.c_void_p(_device.stream_ptr(dev))
    head = (P.batch, int(P.ns.max()), int(P.ms.max()), P.d, _ptr(P.X), _ptr(P.Y),
            _ptr(P.aoff), _ptr(P.boff), _ptr(P.ns_d), _ptr(P.ms_d), _ptr(P.regs), 1,
            _ptr(P.a), _ptr(P.b))
    ref = _outputs(P, dev)
    u, v, niter, status, e_last, e_first, cmax = ref
    sws = ragged._split_workspace(lib, max(nsplit, 1), dev)
    queue = torch.zeros(2, dtype=torch.int32, device=dev)
    rc = lib.potamd_sinkhorn_log_batch_points_split(
        *head, _ptr(u), _ptr(v), max_iter, check_every, stop, 0.0, int(items.numel()),
        _ptr(items), _ptr(split), max(nsplit, 1), _ptr(sws), sws.numel(), _ptr(niter),
        _ptr(status), _ptr(e_last), _ptr(e_first), _ptr(cmax), _ptr(queue), stream)
    assert rc == 0, rc
    got = _outputs(P, dev)
    u, v, niter, status, e_last, e_first, cmax = got
    rc = lib.potamd_sinkhorn_log_batch_points_split_noalloc(
        *head, _ptr(u), _ptr(v), max_iter, check_every, stop, 0.0, int(items.numel()),
        _ptr(items), _ptr(split), nsplit, _ptr(niter), _ptr(status), _ptr(e_last),
        _ptr(e_first), _ptr(cmax), stream)
    assert rc == 0, rc
    torch.cuda.synchronize()
    return ref, got, nsplit


def test_noalloc_matches_split_entry_twice():
    lib = _lib.load()
    dev = _device.require_gpu()
    # first call: 24 problems, the large ones split over 2-4 workgroups
    rng = np.random.default_rng(0)
    sizes = [(int(n), int(m)) for n, m in rng.integers(200, 1025, size=(24, 2))]
    sizes[0] = (1024, 1024)
    sizes[1] = (1000, 700)
    # second call: a smaller batch of smaller problems
    sizes2 = [(int(n), int(m)) for n, m in rng.integers(64, 520, size=(9, 2))]
    sizes2[0] = (512, 512)
    for call, (sz, seed) in enumerate(((sizes, 1), (sizes2, 2))):
        xs, ys, regs = _clouds(seed, sz)
        P = ragged.points_from_clouds(xs, ys, regs)
        parts = ragged.split_plan(P.ns, P.ms)
        ref, got, nsplit = _solve_both(lib, P, dev, parts)
        print(f"no-alloc call {call}: {P.batch} problems, {nsplit} split, parts "
              f"{int(parts.sum())}", flush=True)
        if call == 0:
            assert nsplit > 0
        st = ref[3].cpu().numpy()
        assert ((st & 3) == 0).all(), st
        names = ("u", "v", "niter", "status", "err_last", "err_first", "cmax")
        for name, r, g in zip(names, ref, got):
            assert torch.equal(r, g), (call, name)

