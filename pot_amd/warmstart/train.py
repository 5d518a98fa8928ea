"""Train the warm-start network on (problem, converged potential) pairs.

1. Problems of the family (n = m = 1024, d = 16, GMM clouds of 1-8
   components, reg ~ logU[5e-3, 5e-2], uniform weights) are drawn on the GPU
   from a seeded torch generator, in chunks.
2. Each chunk is solved by the ragged GPU solver (sinkhorn_log semantics,
   stopThr ``label_tol``); the label of a problem is its converged dual
   potential f = reg * log_u in units of the normalised cost.
3. The network (pot_amd.warmstart.net) trains with Adam on the centred,
   reg-weighted squared error, using only the library's forward/backward
   kernels, and is saved to ``weights.npz``.

Usage: ``python -m pot_amd.warmstart.train --problems 200000 --epochs 4``.
"""

import argparse
import json
import math
import os
import time

import numpy as np
import torch

from .. import ragged
from . import WEIGHTS
from .net import DX, WarmStartNet


def gmm_clouds(gen, P, n, d, kmin, kmax, device):
    """P Gaussian-mixture clouds of n points: means ~ U[-3,3]^d, std ~ U[0.3,1.5]."""
    K = torch.randint(kmin, kmax + 1, (P,), generator=gen, device=device)
    means = torch.rand((P, kmax, d), generator=gen, device=device) * 6.0 - 3.0
    stds = torch.rand((P, kmax), generator=gen, device=device) * 1.2 + 0.3
    comp = (torch.rand((P, n), generator=gen, device=device) * K[:, None]).long()
    comp = torch.minimum(comp, (K[:, None] - 1))
    idx = torch.arange(P, device=device)[:, None]
    noise = torch.randn((P, n, d), generator=gen, device=device)
    return means[idx, comp] + stds[idx, comp][..., None] * noise


def draw_family(gen, P, n=1024, d=DX, kmin=1, kmax=8, reg_range=(5e-3, 5e-2), device=None):
    x = gmm_clouds(gen, P, n, d, kmin, kmax, device)
    y = gmm_clouds(gen, P, n, d, kmin, kmax, device)
    lo, hi = math.log(reg_range[0]), math.log(reg_range[1])
    regs = torch.exp(torch.rand(P, generator=gen, device=device, dtype=torch.float64)
                     * (hi - lo) + lo)
    return x.float().contiguous(), y.float().contiguous(), regs


def make_dataset(net, count, seed, chunk=2048, label_tol=1e-6, max_iter=5000, log=print):
    """Inputs (bf16 rows) and labels for `count` problems, all resident on the GPU."""
    dev = net.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    Xs, Es, Qs, Ts, Rs = [], [], [], [], []
    done, t0, iters = 0, time.time(), []
    while done < count:
        P = min(chunk, count - done)
        x, y, regs = draw_family(gen, P, device=dev)
        R = ragged.from_clouds(x, y, regs.cpu().numpy())
        res = ragged.sinkhorn_log_ragged(R, numItermax=max_iter, stopThr=label_tol)
        n = x.shape[1]
        f = (res.u.view(P, n) * regs[:, None]).float()
        Sin, Ein, qn = net.inputs(x.view(-1, DX), y.view(-1, DX), R.a, R.b, regs, R.cmax,
                                  P, n, n)
        Xs.append(Sin.view(P, n, -1))
        Es.append(Ein.view(P, n, -1))
        Qs.append(qn.view(P, n))
        Ts.append(f)
        Rs.append(regs)
        iters.append(res.niter.float().mean().item())
        done += P
        log(f"  labels {done}/{count}  mean iters {iters[-1]:.1f}  "
            f"{time.time() - t0:.1f}s")
    return torch.cat(Xs), torch.cat(Es), torch.cat(Qs), torch.cat(Ts), torch.cat(Rs)


def train(net, data, epochs, batch=32, lr=1e-3, lr_min=1e-5, val=None, log=print):
    X, E, Q, T, Rg = data
    N, n = T.shape
    m = E.shape[1]
    steps_per_epoch = N // batch
    total = epochs * steps_per_epoch
    gen = torch.Generator(device=X.device)
    gen.manual_seed(1234)
    history = []
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        perm = torch.randperm(N, generator=gen, device=X.device)
        run = 0.0
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            xb = X[idx].reshape(batch * n, -1).contiguous()
            eb = E[idx].reshape(batch * m, -1).contiguous()
            qb = Q[idx].reshape(-1).contiguous()
            tb = T[idx].reshape(-1).contiguous()
            rb = Rg[idx].contiguous()
            f = net.forward(xb, eb, qb, batch, n, m, keep=True)
            lp, df = net.loss_grad(f, tb, rb, batch, n)
            net.backward(df)
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, total)))
            net.adam_step(cur)
            step += 1
            if s % 200 == 0 or (steps_per_epoch < 20 and s == 0):
                run = lp.sum().item()
                log(f"  epoch {ep} step {s}/{steps_per_epoch} loss {run:.4e} lr {cur:.2e} "
                    f"{time.time() - t0:.1f}s")
        entry = {"epoch": ep, "train_loss_last": run}
        if val is not None:
            entry["val_loss"] = evaluate(net, val, batch)
            log(f"  epoch {ep} val loss {entry['val_loss']:.4e}")
        history.append(entry)
    return history


def evaluate(net, data, batch=32):
    X, E, Q, T, Rg = data
    N, n = T.shape
    m = E.shape[1]
    tot = 0.0
    cnt = 0
    for s in range(0, N - batch + 1, batch):
        f = net.forward(X[s:s + batch].reshape(batch * n, -1).contiguous(),
                        E[s:s + batch].reshape(batch * m, -1).contiguous(),
                        Q[s:s + batch].reshape(-1).contiguous(), batch, n, m)
        lp, _ = net.loss_grad(f, T[s:s + batch].reshape(-1).contiguous(),
                              Rg[s:s + batch].contiguous(), batch, n)
        tot += lp.sum().item()
        cnt += 1
    return tot / max(cnt, 1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=200_000)
    ap.add_argument("--val", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=10_000)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--log", default=None)
    args = ap.parse_args()
    dev = torch.device("cuda", 0)
    net = WarmStartNet(WarmStartNet.init_params(args.seed), dev)
    t0 = time.time()
    print(f"generating {args.problems} training problems", flush=True)
    data = make_dataset(net, args.problems, args.seed, log=lambda s: print(s, flush=True))
    val = make_dataset(net, args.val, args.seed + 1, log=lambda s: print(s, flush=True))
    t_data = time.time() - t0
    hist = train(net, data, args.epochs, batch=args.batch, lr=args.lr, val=val,
                 log=lambda s: print(s, flush=True))
    net.save(args.out)
    info = {"problems": args.problems, "val": args.val, "epochs": args.epochs,
            "batch": args.batch, "lr": args.lr, "seed": args.seed, "data_s": t_data,
            "train_s": time.time() - t0 - t_data, "history": hist}
    if args.log:
        with open(args.log, "w") as fh:
            json.dump(info, fh, indent=1)
    print(json.dumps(info), flush=True)


if __name__ == "__main__":
    main()
