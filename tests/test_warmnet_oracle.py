"""Pins the float64 warm-start network oracle by central finite differences."""

import numpy as np

import oracle.warmnet as wn


def _tiny_batch(rng, dx=3, h=8):
    params = wn.init_params(rng, dx=dx, h=h)
    # random biases so every ReLU sees both signs
    for k in ("b_e1", "b_e2", "b_s1", "b_s2", "b_m1", "b_m2"):
        params[k] = rng.standard_normal(params[k].shape) * 0.3
    xs = [rng.standard_normal((5, dx)), rng.standard_normal((4, dx))]
    ys = [rng.standard_normal((6, dx)), rng.standard_normal((3, dx))]
    as_ = [rng.dirichlet(np.ones(5)), rng.dirichlet(np.ones(4))]
    bs = [rng.dirichlet(np.ones(6)), rng.dirichlet(np.ones(3))]
    regs = [0.02, 0.05]
    ts = [rng.standard_normal(5), rng.standard_normal(4)]
    return params, xs, ys, as_, bs, regs, ts


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    params, xs, ys, as_, bs, regs, ts = _tiny_batch(rng)
    fs, cache = wn.forward(params, xs, ys, as_, bs, regs)
    grads = wn.backward(params, cache, fs, ts, regs)
    eps = 1e-6
    for name in wn.PARAMS:
        p = params[name]
        flat = p.reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for k in idx:
            old = flat[k]
            flat[k] = old + eps
            lp = wn.loss(wn.forward(params, xs, ys, as_, bs, regs)[0], ts, regs)
            flat[k] = old - eps
            lm = wn.loss(wn.forward(params, xs, ys, as_, bs, regs)[0], ts, regs)
            flat[k] = old
            fd = (lp - lm) / (2 * eps)
            an = grads[name].reshape(-1)[k]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, k, fd, an)


def test_loss_is_gauge_invariant():
    rng = np.random.default_rng(1)
    params, xs, ys, as_, bs, regs, ts = _tiny_batch(rng)
    fs, _ = wn.forward(params, xs, ys, as_, bs, regs)
    l0 = wn.loss(fs, ts, regs)
    l1 = wn.loss(fs, [t + 5.0 for t in ts], regs)
    assert abs(l0 - l1) < 1e-12
    assert wn.loss(ts, ts, regs) == 0.0
