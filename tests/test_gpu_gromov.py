"""Batched (fused) Gromov-Wasserstein (ot/batch/_quadratic.py) against the
float64 oracle.  The inner projections run on the GPU dense solver; with
float64 inputs they reproduce the oracle's iteration sequence, so plans and
values agree to 1e-8 relative.
"""

import numpy as np
import pytest
import torch

import oracle.quadratic as oq
from pot_amd import batch as pb

pytestmark = pytest.mark.gpu


def _graphs(B, n, m, seed):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((B, n, 2))
    Y = rng.standard_normal((B, m, 2))
    C1 = np.sqrt(((X[:, :, None] - X[:, None]) ** 2).sum(-1))
    C2 = np.sqrt(((Y[:, :, None] - Y[:, None]) ** 2).sum(-1))
    return C1, C2


def test_gromov_batch_matches_oracle():
    C1, C2 = _graphs(3, 12, 10, seed=1)
    res = pb.solve_gromov_batch(C1, C2, reg=0.05, max_iter=20, tol=1e-7)
    ref = oq.solve_gromov_batch(C1, C2, reg=0.05, max_iter=20, tol=1e-7)
    assert res.log["n_iter"] == ref["n_iter"]
    np.testing.assert_allclose(res.plan, ref["T"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(res.value, ref["value"], rtol=1e-8)
    np.testing.assert_allclose(res.value_quad, ref["value_quad"], rtol=1e-8)


def test_fused_gromov_batch_and_losses():
    C1, C2 = _graphs(2, 9, 11, seed=2)
    rng = np.random.default_rng(3)
    M = rng.random((2, 9, 11))
    res = pb.solve_gromov_batch(C1, C2, reg=0.1, M=M, alpha=0.4, max_iter=15)
    ref = oq.solve_gromov_batch(C1, C2, reg=0.1, M=M, alpha=0.4, max_iter=15)
    np.testing.assert_allclose(res.plan, ref["T"], rtol=1e-8, atol=1e-14)
    np.testing.assert_allclose(res.value_linear, ref["value_linear"], rtol=1e-8)
    a = np.full((2, 9), 1 / 9)
    b = np.full((2, 11), 1 / 11)
    T = ref["T"]
    got = pb.loss_quadratic_samples_batch(a, b, C1, C2, T, recompute_const=True)
    want = oq.loss_quadratic_batch(oq.tensor_batch(a, b, C1, C2), T, recompute_const=True)
    np.testing.assert_allclose(got, want, rtol=1e-12)
    with pytest.raises(ValueError):
        pb.solve_gromov_batch(C1, C2, M=M)


def test_gromov_batch_torch_envelope_gradient():
    C1, C2 = _graphs(2, 8, 8, seed=4)
    C1t = torch.tensor(C1, dtype=torch.float32, device="cuda", requires_grad=True)
    C2t = torch.tensor(C2, dtype=torch.float32, device="cuda")
    res = pb.solve_gromov_batch(C1t, C2t, reg=0.1, max_iter=10)
    assert res.plan.dtype == torch.float32 and res.plan.is_cuda
    res.value.sum().backward()
    assert C1t.grad is not None and torch.isfinite(C1t.grad).all()
    with pytest.raises(NotImplementedError):
        pb.solve_gromov_batch(C1t, C2t, reg=0.1, grad="autodiff")
