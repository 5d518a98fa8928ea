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


def test_solve_last_step_and_envelope_gradients():
    _, _, a, b, M = _problem(40, 30, seed=3)
    Mt = torch.tensor(M, dtype=torch.float64, device="cuda", requires_grad=True)
    res = pot_amd.solve(Mt, reg=0.1, grad="envelope")
    res.value.backward()
    np.testing.assert_allclose(Mt.grad.cpu().numpy(), res.plan.detach().cpu().numpy(),
                               rtol=1e-12)
    M2 = torch.tensor(M, dtype=torch.float64, device="cuda", requires_grad=True)
    res2 = pot_amd.solve(M2, reg=0.1, grad="last_step", max_iter=30, tol=0.0)
    P, _, val, _, _ = osol.solve_entropic(M, reg=0.1, max_iter=30, tol=0.0, grad="last_step")
    assert _rel(res2.plan.detach().cpu().numpy(), P) < 1e-10
    res2.value.backward()
    assert torch.isfinite(M2.grad).all()
    res3 = pot_amd.solve(M, reg=0.1, grad="last_step", max_iter=30, tol=0.0)
    assert _rel(res3.plan, P) < 1e-10
    with pytest.raises(NotImplementedError):
        pot_amd.solve(Mt, reg=0.1, grad="autodiff")


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
