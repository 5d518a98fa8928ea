"""``pot_amd.solve`` / ``solve_sample`` (ot/solvers.py entropic route) against
the float64 oracle.

Tolerances: float64 inputs -- plan 1e-10 relative on entries >= 1e-30,
values 1e-12 relative, same iteration count; float32 inputs -- plan 1e-5
relative, values 1e-6 relative.
"""

import numpy as np
import pytest
import torch

import oracle.solvers as osol
import pot_amd

pytestmark = pytest.mark.gpu


def _problem(n, m, seed, d=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, d))
    y = rng.standard_normal((m, d)) + 0.7
    M = ((x[:, None] - y[None]) ** 2).sum(-1)
    M /= M.max()
    return x, y, rng.dirichlet(np.ones(n)), rng.dirichlet(np.ones(m)), M


def _rel(P, Pr, floor=1e-30):
    mask = Pr >= floor
    return float(np.max(np.abs(P[mask] - Pr[mask]) / Pr[mask]))


@pytest.mark.parametrize("reg_type", ["KL", "entropy"])
def test_solve_f64_matches_oracle(reg_type):
    _, _, a, b, M = _problem(90, 70, seed=1)
    res = pot_amd.solve(M, a, b, reg=0.05, reg_type=reg_type)
    P, (u, v), val, vl, _ = osol.solve_entropic(M, a, b, 0.05, reg_type=reg_type)
    assert _rel(res.plan, P) < 1e-10
    np.testing.assert_allclose(res.value, val, rtol=1e-12)
    np.testing.assert_allclose(res.value_linear, vl, rtol=1e-12)
    np.testing.assert_allclose(res.potentials[0] + res.potentials[1][0], u + v[0], atol=1e-9)
    assert res.status is None


def test_solve_f32_torch_matches_oracle():
    _, _, a, b, M = _problem(200, 150, seed=2)
    Mt = torch.tensor(M, dtype=torch.float32, device="cuda")
    res = pot_amd.solve(Mt, torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda"),
                        reg=0.02)
    assert res.plan.dtype == torch.float32 and res.plan.is_cuda
    P, _, val, vl, _ = osol.solve_entropic(Mt.double().cpu().numpy(), a, b, 0.02)
    assert _rel(res.plan.double().cpu().numpy(), P) < 1e-5
    np.testing.assert_allclose(float(res.value), val, rtol=1e-6)
    np.testing.assert_allclose(float(res.value_linear), vl, rtol=1e-6)


def test_solve_envelope_gradients():
    _, _, a, b, M = _problem(40, 30, seed=3)
    Mt = torch.tensor(M, dtype=torch.float64, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda", requires_grad=True)
    bt = torch.tensor(b, device="cuda", requires_grad=True)
    res = pot_amd.solve(Mt, at, bt, reg=0.1, grad="envelope")
    res.value.backward()
    np.testing.assert_allclose(Mt.grad.cpu().numpy(), res.plan.detach().cpu().numpy(),
                               rtol=1e-12)
    u, v = (p.detach().cpu().numpy() for p in res.potentials)
    np.testing.assert_allclose(at.grad.cpu().numpy(), 0.1 * (u - u.mean()), rtol=1e-12)
    np.testing.assert_allclose(bt.grad.cpu().numpy(), 0.1 * (v - v.mean()), rtol=1e-12)
    with pytest.raises(NotImplementedError):
        pot_amd.solve(Mt, reg=0.1, grad="autodiff")


def _last_step_value(Mp, M, a, b, u, v, reg):
    """The reference's last-step value as a function of the differentiated
    cost Mp, with the potentials of the first max_iter - 1 iterations and the
    first factor of value_linear fixed (ot/solvers.py:463-481)."""
    from scipy.special import logsumexp
    vn = np.log(b) - logsumexp(-Mp / reg + u[:, None], 0)
    un = np.log(a) - logsumexp(-Mp / reg + v[None, :], 1)
    P = np.exp(-Mp / reg + un[:, None] + vn[None, :])
    return np.sum(M * P) + reg * np.sum(P * np.log(P + 1e-16))


def test_solve_last_step_not_converged_matches_oracle_and_fd():
    # small reg, Dirichlet weights, few iterations: the step order matters
    _, _, a, b, M = _problem(40, 30, seed=3)
    reg, it = 0.01, 12
    P, (u1, v1), val, vl, _ = osol.solve_entropic(M, a, b, reg, max_iter=it, tol=0.0,
                                                  grad="last_step", reg_type="entropy")
    # without gradients: the device path
    res0 = pot_amd.solve(M, a, b, reg=reg, grad="last_step", max_iter=it, tol=0.0,
                         reg_type="entropy")
    assert _rel(res0.plan, P) < 1e-10
    np.testing.assert_allclose(res0.value, val, rtol=1e-11)
    np.testing.assert_allclose(res0.potentials[0], u1, atol=1e-9)
    np.testing.assert_allclose(res0.potentials[1], v1, atol=1e-9)
    # with gradients: one differentiable update on the caller's tensors
    Mt = torch.tensor(M, dtype=torch.float64, device="cuda", requires_grad=True)
    at = torch.tensor(a, device="cuda")
    bt = torch.tensor(b, device="cuda")
    res = pot_amd.solve(Mt, at, bt, reg=reg, grad="last_step", max_iter=it, tol=0.0,
                        reg_type="entropy")
    assert _rel(res.plan.detach().cpu().numpy(), P) < 1e-10
    np.testing.assert_allclose(float(res.value), val, rtol=1e-11)
    np.testing.assert_allclose(float(res.value_linear), vl, rtol=1e-11)
    res.value.backward()
    g = Mt.grad.cpu().numpy()
    # central differences of the reference's definition, from the potentials
    # of the first it - 1 iterations
    from oracle.sinkhorn import sinkhorn_log as osl
    _, lg = osl(a, b, M, reg, numItermax=it - 1, stopThr=0.0, log=True, warn=False)
    rng = np.random.default_rng(0)
    h = 1e-7
    for _ in range(12):
        i, j = int(rng.integers(40)), int(rng.integers(30))
        Mp, Mm = M.copy(), M.copy()
        Mp[i, j] += h
        Mm[i, j] -= h
        fd = (_last_step_value(Mp, M, a, b, lg["log_u"], lg["log_v"], reg)
              - _last_step_value(Mm, M, a, b, lg["log_u"], lg["log_v"], reg)) / (2 * h)
        assert abs(g[i, j] - fd) <= 1e-6 * max(1.0, abs(fd)), (i, j, g[i, j], fd)


def test_solve_sample_and_unsupported_routes():
    x, y, a, b, _ = _problem(60, 50, seed=4)
    res = pot_amd.solve_sample(x, y, a, b, reg=0.5)
    M = ((x[:, None] - y[None]) ** 2).sum(-1)
    P, _, val, _, _ = osol.solve_entropic(M, a, b, 0.5)
    assert _rel(res.plan, P) < 1e-9
    np.testing.assert_allclose(res.value, val, rtol=1e-10)
    with pytest.raises(NotImplementedError):
        pot_amd.solve(M)
    with pytest.raises(NotImplementedError):
        pot_amd.solve(M, reg=0.1, unbalanced=1.0)
    with pytest.raises(NotImplementedError):
        pot_amd.solve(M, reg=0.1, reg_type="L2")
    with pytest.raises(NotImplementedError):
        pot_amd.solve_sample(x, y, reg=0.1, lazy=True)
