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
    ref = torch.relu(ref) * (mask.float() > 0)
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


def _tiny_problem(seed, P=2, n=64, m=64):
    rng = np.random.default_rng(seed)
    xs = [rng.standard_normal((n, 16)).astype(np.float32) for _ in range(P)]
    ys = [rng.standard_normal((m, 16)).astype(np.float32) + 0.5 for _ in range(P)]
    bs = [np.full(m, 1.0 / m) for _ in range(P)]
    regs = np.array([0.01, 0.03])[:P]
    ts = [rng.standard_normal(n) * 0.1 for _ in range(P)]
    return xs, ys, bs, regs, ts


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_backward_match_oracle():
    dev = torch.device("cuda")
    P, n, m = 2, 64, 64
    xs, ys, bs, regs, ts = _tiny_problem(1, P, n, m)
    params = wnet.WarmStartNet.init_params(3)
    rng = np.random.default_rng(5)
    for k in ("b_e1", "b_e2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.1
    net = wnet.WarmStartNet(params, dev)
    X = torch.tensor(np.concatenate(xs), device=dev)
    Y = torch.tensor(np.concatenate(ys), device=dev)
    b = torch.tensor(np.concatenate(bs), device=dev)
    regs_t = torch.tensor(regs, device=dev)
    cmax = torch.ones(P, dtype=torch.float32, device=dev)
    Xin = net.build_inputs(X, P, n, cmax=cmax)
    Ein = net.build_inputs(Y, P, m, b=b, regs=regs_t, cmax=cmax, extra=True)
    f = net.forward(Xin, Ein, P, n, m, keep=True)
    target = torch.tensor(np.concatenate(ts), dtype=torch.float32, device=dev)
    lp, df = net.loss_grad(f, target, regs_t, P, n)
    g = net.backward(df)
    torch.cuda.synchronize()

    # oracle on the same bf16-rounded inputs and weights
    p64 = net.params_numpy()
    for k in wnet.GEMM_WEIGHTS:
        p64[k] = net.wb[k].float().cpu().double().numpy()
    ox = [Xin.float().cpu().numpy()[p * n:(p + 1) * n, :16].astype(np.float64) for p in range(P)]
    eo = Ein.float().cpu().numpy().astype(np.float64)
    oy = [eo[p * m:(p + 1) * m, :16] for p in range(P)]
    ob = [eo[p * m:(p + 1) * m, 16] / m for p in range(P)]
    oregs = [float(np.exp(eo[p * m, 17])) for p in range(P)]
    op = {"W_e1": p64["W_e1"][:18], "b_e1": p64["b_e1"], "W_e2": p64["W_e2"],
          "b_e2": p64["b_e2"], "W_m1x": p64["W_m1x"][:16], "W_m1z": p64["W_m1z"],
          "b_m1": p64["b_m1"], "W_m2": p64["W_m2"], "b_m2": p64["b_m2"],
          "w_m3": p64["w_m3"][:, None], "b_m3": p64["b_m3"]}
    fo, cache = own.forward(op, ox, oy, ob, oregs)
    go = own.backward(op, cache, fo, ts, list(regs))
    lo = own.loss(fo, ts, list(regs))
    errs = {"f": _rel(f.cpu().numpy(), np.concatenate(fo)),
            "loss": abs(lp.sum().item() - lo) / lo}
    for k, ref in go.items():
        mine = g[k].cpu().double().numpy()
        if k == "W_e1":
            mine = mine[:18]
        if k == "W_m1x":
            mine = mine[:16]
        if k == "w_m3":
            ref = ref[:, 0]
        errs[k] = _rel(mine.reshape(ref.shape), ref)
    print("relative errors:", errs, flush=True)
    assert all(v < 3e-2 for v in errs.values()), errs


def test_training_lowers_loss():
    dev = torch.device("cuda")
    from pot_amd.warmstart import train as wtrain
    net = wnet.WarmStartNet(wnet.WarmStartNet.init_params(0), dev)
    data = wtrain.make_dataset(net, 64, seed=77, chunk=64, log=lambda s: None)
    l0 = wtrain.evaluate(net, data, batch=16)
    wtrain.train(net, data, epochs=15, batch=16, lr=1e-3, log=lambda s: print(s, flush=True))
    l1 = wtrain.evaluate(net, data, batch=16)
    print("loss before / after", l0, l1, flush=True)
    assert l1 < 0.8 * l0, (l0, l1)
