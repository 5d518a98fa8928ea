"""GPU checks of the warm-start network kernels.

* GEMM: every transpose / epilogue mode against a float32 torch matmul of the
  same bf16-representable inputs (fp32 accumulation: 1e-4 relative for fp32
  outputs; bf16 outputs within one bf16 ulp, 1e-2 relative).
* Forward and backward of the whole network against the float64 oracle
  (oracle/warmnet.py) on a tiny batch: bf16 activations bound the agreement
  to 3e-2 relative (norm-wise per tensor).
* A short training run lowers the loss.
"""

import ctypes

import numpy as np
import pytest
import torch

import oracle.warmnet as own
from pot_amd import _lib
from pot_amd.warmstart import net as wnet

pytestmark = pytest.mark.gpu


def _bf16_exact(shape, gen, scale=1.0):
    return (torch.randn(shape, generator=gen) * scale).to(torch.bfloat16)


@pytest.mark.parametrize("transA,transB", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (100, 70, 45), (513, 256, 256)])
def test_gemm_modes_match_torch(transA, transB, M, N, K):
    gen = torch.Generator().manual_seed(M + N + K)
    dev = torch.device("cuda")
    A = _bf16_exact((K, M) if transA else (M, K), gen)
    B = _bf16_exact((N, K) if transB else (K, N), gen)
    opA = A.float().t() if transA else A.float()
    opB = B.float().t() if transB else B.float()
    ref = opA @ opB
    Ad, Bd = A.to(dev).contiguous(), B.to(dev).contiguous()
    C = torch.empty((M, N), dtype=torch.float32, device=dev)
    lib = _lib.load()
    rc = lib.potamd_gemm_bf16(M, N, K, ctypes.c_void_p(Ad.data_ptr()), Ad.shape[1], transA,
                              ctypes.c_void_p(Bd.data_ptr()), Bd.shape[1], transB,
                              ctypes.c_void_p(C.data_ptr()), N, 1, 1.0, None, None, 1, 0, None,
                              N, 1, None)
    assert rc == 0
    torch.cuda.synchronize()
    err = (C.cpu() - ref).abs().max() / ref.abs().max()
    assert err < 1e-4, err


def test_gemm_epilogues_and_split_k():
    gen = torch.Generator().manual_seed(0)
    dev = torch.device("cuda")
    M, N, K, seg = 256, 128, 64, 64
    A = _bf16_exact((M, K), gen).to(dev)
    B = _bf16_exact((K, N), gen).to(dev)
    bias = torch.randn(N, generator=gen).to(dev)
    rowvec = torch.randn((M // seg, N), generator=gen).to(dev)
    mask = _bf16_exact((M, N), gen).to(dev)
    net = wnet.WarmStartNet(wnet.WarmStartNet.init_params(0), dev)
    out = torch.empty((M, N), dtype=torch.bfloat16, device=dev)
    net._gemm(A, B, out, M, N, K, bias=bias, rowvec=rowvec, seg_len=seg, relu=1, mask=mask)
    ref = A.float() @ B.float() + bias + rowvec.repeat_interleave(seg, 0)
    # leaky ReLU (slope 0.01), then the backward mask: keep where mask > 0,
    # scale by the slope elsewhere
    ref = torch.where(ref > 0, ref, 0.01 * ref)
    ref = torch.where(mask.float() > 0, ref, 0.01 * ref)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.float().cpu(), ref.cpu(), rtol=1e-2, atol=1e-2)
    # split-K accumulation into fp32 (weight-gradient form A^T B over many rows)
    R = 4096
    X = _bf16_exact((R, 32), gen).to(dev)
    Y = _bf16_exact((R, N), gen).to(dev)
    G = torch.zeros((32, N), dtype=torch.float32, device=dev)
    net._gemm(X, Y, G, 32, N, R, transA=1, out_kind=2, splits=8)
    torch.cuda.synchronize()
    ref = X.float().t() @ Y.float()
    assert ((G - ref).abs().max() / ref.abs().max()).item() < 1e-4


def _bf16(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _net_inputs(net, P, n, m, seed):
    rng = np.random.default_rng(seed)
    dev = net.device
    X = torch.tensor(rng.standard_normal((P * n, 16)).astype(np.float32), device=dev)
    Y = torch.tensor(rng.standard_normal((P * m, 16)).astype(np.float32) + 0.5, device=dev)
    a = torch.full((P * n,), 1.0 / n, dtype=torch.float64, device=dev)
    b = torch.full((P * m,), 1.0 / m, dtype=torch.float64, device=dev)
    regs = torch.tensor([0.01, 0.03][:P], dtype=torch.float64, device=dev)
    cmax = torch.full((P,), 4.0, dtype=torch.float32, device=dev)
    t = torch.tensor(rng.standard_normal(P * n) * 0.1, dtype=torch.float32, device=dev)
    return X, Y, a, b, regs, cmax, t


def test_forward_backward_match_oracle():
    dev = torch.device("cuda")
    P, n, m = 2, 64, 96
    params = wnet.WarmStartNet.init_params(3)
    rng = np.random.default_rng(5)
    for k in ("b_e1", "b_e2", "b_s1", "b_s2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.1
    params["w_m3"] = rng.standard_normal(params["w_m3"].shape) * 0.05
    net = wnet.WarmStartNet(params, dev)
    X, Y, a, b, regs, cmax, target = _net_inputs(net, P, n, m, 1)
    Sin, Ein, qn = net.inputs(X, Y, a, b, regs, cmax, P, n, m)
    f = net.forward(Sin, Ein, qn, P, n, m, keep=True)
    lp, df = net.loss_grad(f, target, regs, P, n)
    g = net.backward(df)
    torch.cuda.synchronize()

    # oracle weights = the GPU's (bf16 copies for the GEMM layers)
    w = net.params_numpy()
    for k in wnet.GEMM_WEIGHTS:
        w[k] = net.wb[k].float().cpu().double().numpy()
    op = {k: w[k] for k in own.PARAMS if k not in ("w_m3",)}
    for k in ("W_e1", "W_s1", "W_m1x"):
        op[k] = w[k][:18]
    op["w_m3"] = w["w_m3"][:, None]
    S = Sin.float().cpu().double().numpy()
    E = Ein.float().cpu().double().numpy()
    regs_l = [float(r) for r in regs.cpu().numpy()]
    xs = [S[p * n:(p + 1) * n, :16] for p in range(P)]
    ys = [E[p * m:(p + 1) * m, :16] for p in range(P)]
    as_ = [(S[p * n:(p + 1) * n, 16] + 1.0) / n for p in range(P)]
    bs = [(E[p * m:(p + 1) * m, 16] + 1.0) / m for p in range(P)]
    oregs = [float(np.exp(E[p * m, 17] * own.LR_S + own.LR_C)) for p in range(P)]
    fo, cache = own.forward(op, xs, ys, as_, bs, oregs)
    ts = list(target.cpu().double().numpy().reshape(P, n))
    errs = {"f": _rel(f.cpu().numpy(), np.concatenate(fo))}
    lo = own.loss(fo, ts, regs_l)
    errs["loss"] = abs(lp.sum().item() - lo) / lo
    fg = list(f.cpu().double().numpy().reshape(P, n))
    dfo = own.loss_grad_f(fg, ts, regs_l)
    errs["df"] = _rel(df.cpu().numpy(), np.concatenate(dfo))
    assert errs["f"] < 3e-2 and errs["loss"] < 3e-2 and errs["df"] < 1e-4, errs

    # backward from the GPU's own activations and df, with the GPU's bf16
    # storage of the backward signals emulated: only accumulation order differs
    Sin_, Ein_, _, _, _, H1, H2, zt, S1, S2, zs, G1, G2 = net._cache
    t64 = lambda t: t.float().cpu().double().numpy()
    gc = []
    for p in range(P):
        rn, rm = slice(p * n, (p + 1) * n), slice(p * m, (p + 1) * m)
        te = (t64(H1)[rm], t64(H1)[rm], t64(H2)[rm], t64(H2)[rm], t64(zt)[p])
        se = (t64(S1)[rn], t64(S1)[rn], t64(S2)[rn], t64(S2)[rn], t64(zs)[p])
        s_rows = S[rn, :18]
        gc.append((s_rows[:, :16], E[rm, :18], s_rows, te, se, t64(G1)[rn], t64(G1)[rn],
                   t64(G2)[rn], t64(G2)[rn]))
    dfs = [d.astype(np.float64) for d in df.cpu().double().numpy().reshape(P, n)]
    go = own.backward_from_df(op, gc, dfs, rnd=_bf16)
    for k, ref in go.items():
        mine = g[k].cpu().double().numpy()
        if k in ("W_e1", "W_s1", "W_m1x"):
            mine = mine[:18]
        if k == "w_m3":
            ref = ref[:, 0]
        if k == "b_m3":  # zero by the gauge invariance of the centred loss
            assert abs(float(mine.sum())) < 1e-6
            continue
        errs[k] = _rel(mine.reshape(ref.shape), ref)
    print("relative errors:", errs, flush=True)
    assert all(v < 1e-2 for v in errs.values()), errs


def test_training_lowers_loss():
    # overfit a handful of problems: the kernels' gradients must descend
    dev = torch.device("cuda")
    from pot_amd.warmstart import train as wtrain
    net = wnet.WarmStartNet(wnet.WarmStartNet.init_params(0), dev)
    data = wtrain.make_dataset(net, 8, seed=77, chunk=8, log=lambda s: None)
    l0 = wtrain.evaluate(net, data, batch=8)
    wtrain.train(net, data, epochs=150, batch=8, lr=2e-3, lr_min=2e-3, log=lambda s: None)
    l1 = wtrain.evaluate(net, data, batch=8)
    print("loss before / after", l0, l1, flush=True)
    assert l1 < 0.6 * l0, (l0, l1)


def test_fused_forward_matches_gemm_path_and_oracle():
    # The inference kernel (two layers per launch, activations on chip) must
    # give the per-layer GEMM path's result up to fp32 rounding order (and the
    # head's bf16 hi + lo split of w3), and the oracle's within the bf16 bar
    # of the GEMM path.
    dev = torch.device("cuda")
    P, n, m = 2, 128, 192
    params = wnet.WarmStartNet.init_params(4)
    rng = np.random.default_rng(6)
    for k in ("b_e1", "b_e2", "b_s1", "b_s2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.1
    params["w_m3"] = rng.standard_normal(params["w_m3"].shape) * 0.05
    params["b_m3"] = np.array([0.3])
    net = wnet.WarmStartNet(params, dev)
    X, Y, a, b, regs, cmax, _ = _net_inputs(net, P, n, m, 2)
    Sin, Ein, qn = net.inputs(X, Y, a, b, regs, cmax, P, n, m)
    f_gemm = net.forward(Sin, Ein, qn, P, n, m, keep=True)
    f_fused = net.forward(Sin, Ein, qn, P, n, m, keep=False)
    # inputs built inside the fused kernels from the points
    f_pts = net.forward_points(X, Y, a, b, regs, cmax, P, n, m)
    torch.cuda.synchronize()
    err = _rel(f_fused.cpu().numpy(), f_gemm.cpu().numpy())
    err_p = _rel(f_pts.cpu().numpy(), f_gemm.cpu().numpy())
    print("fused vs GEMM path:", err, "from points:", err_p, flush=True)
    assert err < 2e-3, err
    assert err_p < 2e-3, err_p

    w = net.params_numpy()
    for k in wnet.GEMM_WEIGHTS:
        w[k] = net.wb[k].float().cpu().double().numpy()
    op = {k: w[k] for k in own.PARAMS if k not in ("w_m3",)}
    for k in ("W_e1", "W_s1", "W_m1x"):
        op[k] = w[k][:18]
    op["w_m3"] = w["w_m3"][:, None]
    S = Sin.float().cpu().double().numpy()
    E = Ein.float().cpu().double().numpy()
    xs = [S[p * n:(p + 1) * n, :16] for p in range(P)]
    ys = [E[p * m:(p + 1) * m, :16] for p in range(P)]
    as_ = [(S[p * n:(p + 1) * n, 16] + 1.0) / n for p in range(P)]
    bs = [(E[p * m:(p + 1) * m, 16] + 1.0) / m for p in range(P)]
    oregs = [float(np.exp(E[p * m, 17] * own.LR_S + own.LR_C)) for p in range(P)]
    fo, _ = own.forward(op, xs, ys, as_, bs, oregs)
    err_o = _rel(f_fused.cpu().numpy(), np.concatenate(fo))
    assert err_o < 3e-2, err_o


def test_fused_forward_many_segments_per_workgroup():
    # More problems than workgroups: each workgroup walks several segments,
    # so the per-segment bias, scale and log reg change inside one launch and
    # the segment sums restart; one and two blocks per segment.
    dev = torch.device("cuda")
    params = wnet.WarmStartNet.init_params(9)
    rng = np.random.default_rng(12)
    for k in ("b_e1", "b_e2", "b_s1", "b_s2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.1
    params["w_m3"] = rng.standard_normal(params["w_m3"].shape) * 0.05
    params["b_m3"] = np.array([-0.2])
    net = wnet.WarmStartNet(params, dev)
    P, n, m = 600, 64, 128
    X = torch.tensor(rng.standard_normal((P * n, 16)).astype(np.float32), device=dev)
    Y = torch.tensor(rng.standard_normal((P * m, 16)).astype(np.float32) + 0.5, device=dev)
    a = torch.tensor(rng.dirichlet(np.ones(n), size=P).reshape(-1), device=dev)
    b = torch.tensor(rng.dirichlet(np.ones(m), size=P).reshape(-1), device=dev)
    regs = torch.tensor(np.exp(rng.uniform(np.log(5e-3), np.log(5e-2), P)), device=dev)
    cmax = torch.tensor(rng.uniform(2.0, 8.0, P).astype(np.float32), device=dev)
    Sin, Ein, qn = net.inputs(X, Y, a, b, regs, cmax, P, n, m)
    f_gemm = net.forward(Sin, Ein, qn, P, n, m, keep=True)
    f_fused = net.forward(Sin, Ein, qn, P, n, m, keep=False)
    f_pts = net.forward_points(X, Y, a, b, regs, cmax, P, n, m)
    torch.cuda.synchronize()
    g = f_gemm.cpu().numpy().reshape(P, n)
    for name, f in (("fused", f_fused), ("points", f_pts)):
        f = f.cpu().numpy().reshape(P, n)
        per = np.linalg.norm(f - g, axis=1) / np.linalg.norm(g, axis=1)
        print(name, "worst per-problem relative difference", per.max(), flush=True)
        assert per.max() < 2e-3, (name, int(per.argmax()), per.max())


def test_fused_forward_encoder_subsample():
    # Clouds of 1024 points made of 256 distinct points, each repeated four
    # times in a row (weights equal within a group): the encoders' means over
    # the stride-4 sample equal those over every point, so the prediction with
    # ENC_POINTS = 256 must equal the one from every point (fp32 rounding
    # order aside), and that one the per-layer GEMM path.
    dev = torch.device("cuda")
    params = wnet.WarmStartNet.init_params(5)
    rng = np.random.default_rng(21)
    for k in ("b_e1", "b_e2", "b_s1", "b_s2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.1
    params["w_m3"] = rng.standard_normal(params["w_m3"].shape) * 0.05
    net = wnet.WarmStartNet(params, dev)
    P, n, m = 6, 1024, 1024
    base_x = rng.standard_normal((P, n // 4, 16)).astype(np.float32)
    base_y = (rng.standard_normal((P, m // 4, 16)) + 0.5).astype(np.float32)
    X = torch.tensor(np.repeat(base_x, 4, axis=1).reshape(-1, 16), device=dev)
    Y = torch.tensor(np.repeat(base_y, 4, axis=1).reshape(-1, 16), device=dev)
    wa = np.repeat(rng.dirichlet(np.ones(n // 4), size=P) / 4, 4, axis=1)
    wb = np.repeat(rng.dirichlet(np.ones(m // 4), size=P) / 4, 4, axis=1)
    a = torch.tensor(wa.reshape(-1), device=dev)
    b = torch.tensor(wb.reshape(-1), device=dev)
    regs = torch.tensor(np.exp(rng.uniform(np.log(5e-3), np.log(5e-2), P)), device=dev)
    cmax = torch.tensor(rng.uniform(2.0, 8.0, P).astype(np.float32), device=dev)
    assert net.ENC_POINTS == 256
    f_sub = net.forward_points(X, Y, a, b, regs, cmax, P, n, m)
    net.ENC_POINTS = 0
    f_all = net.forward_points(X, Y, a, b, regs, cmax, P, n, m)
    Sin, Ein, qn = net.inputs(X, Y, a, b, regs, cmax, P, n, m)
    f_gemm = net.forward(Sin, Ein, qn, P, n, m, keep=True)
    torch.cuda.synchronize()
    e_sub = _rel(f_sub.cpu().numpy(), f_all.cpu().numpy())
    e_all = _rel(f_all.cpu().numpy(), f_gemm.cpu().numpy())
    print("stride-4 sample vs every point:", e_sub, "every point vs GEMM path:", e_all, flush=True)
    assert e_sub < 1e-4, e_sub
    assert e_all < 2e-3, e_all
