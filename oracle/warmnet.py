"""Float64 NumPy forward/backward of the warm-start network.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The network has no
counterpart in the reference; this restatement is pinned by central finite
differences (tests/test_warmnet_oracle.py) and is the checker of the GPU
forward/backward kernels.

Per problem (sources x: n x dx, targets y: m x dx, target weights b, reg):

    e_j  = [y_j, m * b_j, log reg]                    (m x (dx + 2))
    H1   = relu(e W_e1 + b_e1),  H2 = relu(H1 W_e2 + b_e2)   (m x h)
    z    = mean_j H2                                   (h)
    c    = z W_m1z + b_m1                              (h)
    G1   = relu(x W_m1x + c),    G2 = relu(G1 W_m2 + b_m2)   (n x h)
    f    = G2 w_m3 + b_m3                              (n)

Loss of a batch: mean over problems of mean_i (fc_i - tc_i)^2 * (r0 / reg)^2
where fc, tc are the centred prediction and target potentials (the gauge
f + const is free) and r0 = 1e-2 fixes the scale.
"""

import numpy as np

PARAMS = ("W_e1", "b_e1", "W_e2", "b_e2", "W_m1x", "W_m1z", "b_m1", "W_m2", "b_m2",
          "w_m3", "b_m3")
R0 = 1e-2


def init_params(rng, dx=16, h=256):
    def lin(fan_in, fan_out):
        return rng.standard_normal((fan_in, fan_out)) * np.sqrt(2.0 / fan_in)
    return {
        "W_e1": lin(dx + 2, h), "b_e1": np.zeros(h),
        "W_e2": lin(h, h), "b_e2": np.zeros(h),
        "W_m1x": lin(dx + h, h)[:dx], "W_m1z": lin(dx + h, h)[dx:], "b_m1": np.zeros(h),
        "W_m2": lin(h, h), "b_m2": np.zeros(h),
        "w_m3": lin(h, 1) * 0.1, "b_m3": np.zeros(1),
    }


def forward(params, xs, ys, bs, regs):
    """Lists of per-problem arrays -> (list of f, cache)."""
    P = params
    out, cache = [], []
    for x, y, b, reg in zip(xs, ys, bs, regs):
        m = y.shape[0]
        e = np.concatenate([y, (m * b)[:, None], np.full((m, 1), np.log(reg))], axis=1)
        A1 = e @ P["W_e1"] + P["b_e1"]
        H1 = np.maximum(A1, 0)
        A2 = H1 @ P["W_e2"] + P["b_e2"]
        H2 = np.maximum(A2, 0)
        z = H2.mean(0)
        c = z @ P["W_m1z"] + P["b_m1"]
        B1 = x @ P["W_m1x"] + c
        G1 = np.maximum(B1, 0)
        B2 = G1 @ P["W_m2"] + P["b_m2"]
        G2 = np.maximum(B2, 0)
        f = G2 @ P["w_m3"][:, 0] + P["b_m3"][0]
        out.append(f)
        cache.append((x, e, A1, H1, A2, H2, z, B1, G1, B2, G2))
    return out, cache


def loss(fs, targets, regs):
    vals = []
    for f, t, reg in zip(fs, targets, regs):
        d = (f - f.mean()) - (t - t.mean())
        vals.append(np.mean(d * d) * (R0 / reg) ** 2)
    return float(np.mean(vals))


def backward(params, cache, fs, targets, regs):
    """Gradients of ``loss`` with respect to every parameter."""
    P = params
    g = {k: np.zeros_like(v) for k, v in P.items()}
    nb = len(fs)
    for (x, e, A1, H1, A2, H2, z, B1, G1, B2, G2), f, t, reg in zip(cache, fs, targets, regs):
        n = f.shape[0]
        d = (f - f.mean()) - (t - t.mean())
        df = 2.0 * d / n * (R0 / reg) ** 2 / nb
        df = df - df.mean()  # through the centring
        g["w_m3"][:, 0] += G2.T @ df
        g["b_m3"][0] += df.sum()
        dG2 = np.outer(df, P["w_m3"][:, 0])
        dB2 = dG2 * (B2 > 0)
        g["W_m2"] += G1.T @ dB2
        g["b_m2"] += dB2.sum(0)
        dG1 = dB2 @ P["W_m2"].T
        dB1 = dG1 * (B1 > 0)
        g["W_m1x"] += x.T @ dB1
        dc = dB1.sum(0)
        g["b_m1"] += dc
        g["W_m1z"] += np.outer(z, dc)
        dz = P["W_m1z"] @ dc
        m = H2.shape[0]
        dH2 = np.broadcast_to(dz / m, H2.shape)
        dA2 = dH2 * (A2 > 0)
        g["W_e2"] += H1.T @ dA2
        g["b_e2"] += dA2.sum(0)
        dH1 = dA2 @ P["W_e2"].T
        dA1 = dH1 * (A1 > 0)
        g["W_e1"] += e.T @ dA1
        g["b_e1"] += dA1.sum(0)
    return g
