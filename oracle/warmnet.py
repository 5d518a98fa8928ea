"""Float64 NumPy forward/backward of the warm-start network.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The network has no
counterpart in the reference; this restatement is pinned by central finite
differences (tests/test_warmnet_oracle.py) and is the checker of the GPU
forward/backward kernels.

Per problem (sources x: n x dx with weights a, targets y: m x dx with
weights b, reg; points already scaled by 1/sqrt(max cost)):

    e_j  = [y_j, m b_j - 1, lr]       H1 = act(e W_e1 + b_e1)
    H2   = act(H1 W_e2 + b_e2)        zt = mean_j H2        (target encoder)
    s_i  = [x_i, n a_i - 1, lr]       S1 = act(s W_s1 + b_s1)
    S2   = act(S1 W_s2 + b_s2)        zs = mean_i S2        (source encoder)
    c    = zt W_m1z + zs W_m1s + b_m1
    G1   = act(s W_m1x + c)           G2 = act(G1 W_m2 + b_m2)
    f    = |x_i|^2 + G2 w_m3 + b_m3

with act the leaky ReLU of slope 0.01 and lr = (log reg - LR_C) / LR_S
(about [-1, 1] for reg in [5e-3, 5e-2]).

The |x_i|^2 term is exact: f - |x|^2 is a soft-max of functions affine in x
(the c-transform of the squared cost), the kind of function a ReLU network
represents well.

Loss of a batch: mean over problems of mean_i (fc_i - tc_i)^2 * (r0 / reg)^2
where fc, tc are the centred prediction and target potentials (the gauge
f + const is free) and r0 = 1e-2 fixes the scale.
"""

import numpy as np

PARAMS = ("W_e1", "b_e1", "W_e2", "b_e2", "W_s1", "b_s1", "W_s2", "b_s2",
          "W_m1x", "W_m1z", "W_m1s", "b_m1", "W_m2", "b_m2", "w_m3", "b_m3")
R0 = 1e-2
LEAK = 0.01
LR_C, LR_S = -4.147366, 1.151293


def act(a):
    return np.where(a > 0, a, LEAK * a)


def dact(a):
    return np.where(a > 0, 1.0, LEAK)


def features(pts, w, reg):
    """Input rows [point, len * weight - 1, normalised log reg]."""
    k = pts.shape[0]
    lr = (np.log(reg) - LR_C) / LR_S
    return np.concatenate([pts, (k * w - 1.0)[:, None], np.full((k, 1), lr)], axis=1)


def init_params(rng, dx=16, h=256):
    def lin(fan_in, fan_out):
        return rng.standard_normal((fan_in, fan_out)) * np.sqrt(2.0 / fan_in)
    return {
        "W_e1": lin(dx + 2, h), "b_e1": np.zeros(h),
        "W_e2": lin(h, h), "b_e2": np.zeros(h),
        "W_s1": lin(dx + 2, h), "b_s1": np.zeros(h),
        "W_s2": lin(h, h), "b_s2": np.zeros(h),
        "W_m1x": lin(dx + 2 * h, h)[:dx + 2], "W_m1z": lin(dx + 2 * h, h)[:h],
        "W_m1s": lin(dx + 2 * h, h)[:h], "b_m1": np.zeros(h),
        "W_m2": lin(h, h), "b_m2": np.zeros(h),
        "w_m3": lin(h, 1) * 0.1, "b_m3": np.zeros(1),
    }


def _encode(P, e, W1, b1, W2, b2):
    A1 = e @ P[W1] + P[b1]
    H1 = act(A1)
    A2 = H1 @ P[W2] + P[b2]
    H2 = act(A2)
    return A1, H1, A2, H2, H2.mean(0)


def forward(params, xs, ys, as_, bs, regs):
    """Lists of per-problem arrays -> (list of f, cache)."""
    P = params
    out, cache = [], []
    for x, y, a, b, reg in zip(xs, ys, as_, bs, regs):
        e = features(y, b, reg)
        s = features(x, a, reg)
        te = _encode(P, e, "W_e1", "b_e1", "W_e2", "b_e2")
        se = _encode(P, s, "W_s1", "b_s1", "W_s2", "b_s2")
        c = te[4] @ P["W_m1z"] + se[4] @ P["W_m1s"] + P["b_m1"]
        B1 = s @ P["W_m1x"] + c
        G1 = act(B1)
        B2 = G1 @ P["W_m2"] + P["b_m2"]
        G2 = act(B2)
        f = np.sum(x * x, axis=1) + G2 @ P["w_m3"][:, 0] + P["b_m3"][0]
        out.append(f)
        cache.append((x, e, s, te, se, B1, G1, B2, G2))
    return out, cache


def loss(fs, targets, regs):
    vals = []
    for f, t, reg in zip(fs, targets, regs):
        d = (f - f.mean()) - (t - t.mean())
        vals.append(np.mean(d * d) * (R0 / reg) ** 2)
    return float(np.mean(vals))


def loss_grad_f(fs, targets, regs):
    """d loss / d f per problem (through the centring)."""
    nb = len(fs)
    out = []
    for f, t, reg in zip(fs, targets, regs):
        n = f.shape[0]
        d = (f - f.mean()) - (t - t.mean())
        df = 2.0 * d / n * (R0 / reg) ** 2 / nb
        out.append(df - df.mean())
    return out


def backward(params, cache, fs, targets, regs):
    """Gradients of ``loss`` with respect to every parameter."""
    return backward_from_df(params, cache, loss_grad_f(fs, targets, regs))


def _encode_back(P, g, inp, enc, dz, W1, b1, W2, b2, rnd):
    A1, H1, A2, H2, _ = enc
    rows = H2.shape[0]
    dA2 = rnd(np.broadcast_to(dz / rows, H2.shape) * dact(A2))
    g[W2] += H1.T @ dA2
    g[b2] += dA2.sum(0)
    dA1 = rnd((dA2 @ P[W2].T) * dact(A1))
    g[W1] += inp.T @ dA1
    g[b1] += dA1.sum(0)


def backward_from_df(params, cache, dfs, rnd=None):
    """Gradients of every parameter given d loss / d f of each problem.

    ``rnd`` (optional) rounds the backward signals dB2, dB1, dA2, dA1 the way
    a lower-precision implementation stores them."""
    rnd = rnd or (lambda t: t)
    P = params
    g = {k: np.zeros_like(v) for k, v in P.items()}
    for (x, e, s, te, se, B1, G1, B2, G2), df in zip(cache, dfs):
        g["w_m3"][:, 0] += G2.T @ df
        g["b_m3"][0] += df.sum()
        dB2 = rnd(np.outer(df, P["w_m3"][:, 0]) * dact(B2))
        g["W_m2"] += G1.T @ dB2
        g["b_m2"] += dB2.sum(0)
        dB1 = rnd((dB2 @ P["W_m2"].T) * dact(B1))
        g["W_m1x"] += s.T @ dB1
        dc = dB1.sum(0)
        g["b_m1"] += dc
        g["W_m1z"] += np.outer(te[4], dc)
        g["W_m1s"] += np.outer(se[4], dc)
        _encode_back(P, g, e, te, P["W_m1z"] @ dc, "W_e1", "b_e1", "W_e2", "b_e2", rnd)
        _encode_back(P, g, s, se, P["W_m1s"] @ dc, "W_s1", "b_s1", "W_s2", "b_s2", rnd)
    return g
