"""Seeded synthetic Sinkhorn workloads.

All inputs are drawn with ``numpy.random.default_rng`` in float64 on the host
and cast once, so a CPU check and the GPU see bit-identical data.

* ``gauss_pair``      two Gaussian clouds (single-problem workloads)
* ``gmm_cloud``       isotropic Gaussian mixture in d dimensions
* ``ragged_family``   many uneven problems: n_i, m_i ~ U{lo..hi}, GMM clouds with
                      K ~ U{kmin..kmax} components (means ~ U[-3, 3]^d, std ~
                      U[0.3, 1.5]), uniform weights, reg_i ~ logU[reg_lo, reg_hi];
                      the cost of each problem is its squared Euclidean
                      distance divided by its own maximum
"""

from dataclasses import dataclass

import numpy as np


def gauss_pair(n, m, d, seed, shift=0.5, scale_y=np.sqrt(2.0)):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    y = shift + scale_y * rng.standard_normal((m, d))
    return rng, x, y


def gmm_cloud(rng, n, d, kmin, kmax):
    k = int(rng.integers(kmin, kmax + 1))
    means = rng.uniform(-3.0, 3.0, size=(k, d))
    stds = rng.uniform(0.3, 1.5, size=k)
    comp = rng.integers(0, k, size=n)
    return means[comp] + stds[comp, None] * rng.standard_normal((n, d))


@dataclass
class Family:
    xs: list          # float32 arrays (n_i, d)
    ys: list          # float32 arrays (m_i, d)
    regs: np.ndarray  # float64 (B,)
    seeds: list


def ragged_family(count, seed, n_range=(64, 1024), m_range=None, d=16, kmin=1, kmax=4,
                  reg_range=(5e-3, 5e-2), first_index=0):
    """Problems ``first_index .. first_index+count-1`` of the family ``seed``.

    Problem i uses the sub-stream ``SeedSequence(seed).spawn(first_index+count)[i]``,
    so any subset can be regenerated alone.
    """
    m_range = m_range or n_range
    children = np.random.SeedSequence(seed).spawn(first_index + count)[first_index:]
    xs, ys, regs = [], [], []
    for ss in children:
        rng = np.random.default_rng(ss)
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        m = int(rng.integers(m_range[0], m_range[1] + 1))
        xs.append(gmm_cloud(rng, n, d, kmin, kmax).astype(np.float32))
        ys.append(gmm_cloud(rng, m, d, kmin, kmax).astype(np.float32))
        regs.append(float(np.exp(rng.uniform(np.log(reg_range[0]), np.log(reg_range[1])))))
    return Family(xs=xs, ys=ys, regs=np.asarray(regs), seeds=list(range(first_index,
                                                                        first_index + count)))


def config1(seed=0, n=2048):
    """x ~ N(0, I2), y ~ N((2, 2), [[1, .5], [.5, 1]]), uniform weights, M / max."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 2))
    y = rng.multivariate_normal([2.0, 2.0], [[1.0, 0.5], [0.5, 1.0]], n)
    M = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    M = M / M.max()
    a = np.full(n, 1.0 / n)
    return a, a.copy(), M


def config2_clouds(seed=1, n=16384, d=64):
    """x ~ N(0, I), y ~ N(0.5 * 1, 2 I) in fp32; Dirichlet(1) weights."""
    rng, x, y = gauss_pair(n, n, d, seed)
    a = rng.dirichlet(np.ones(n))
    b = rng.dirichlet(np.ones(n))
    return x.astype(np.float32), y.astype(np.float32), a, b


def config4_clouds(seed=3, n=65536, d=128):
    """x ~ N(0, I), y ~ N(0, I) + 0.25 * U[-1, 1]^d (one shift vector), fp32."""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    shift = 0.25 * rng.uniform(-1.0, 1.0, size=d)
    y = rng.standard_normal((n, d)) + shift
    return x.astype(np.float32), y.astype(np.float32)


def config3(seed=2, B=512):
    return ragged_family(B, seed, n_range=(64, 1024), d=16, kmin=1, kmax=4)


def config5(seed, count, first_index=0):
    """n = m = 1024, d = 16, GMMs of 1-8 components, reg ~ logU[5e-3, 5e-2]."""
    return ragged_family(count, seed, n_range=(1024, 1024), d=16, kmin=1, kmax=8,
                         first_index=first_index)


def cost_host(x, y):
    """float64 squared Euclidean cost divided by its max (the oracle's input)."""
    x = x.astype(np.float64)
    y = y.astype(np.float64)
    M = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    return M / M.max()
