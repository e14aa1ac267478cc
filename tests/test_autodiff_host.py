"""autograd plumbing of trench01_amd.autodiff with a stand-in batch whose
Jacobian is known (the device batch is covered by test_gpu_autodiff.py)."""
import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

from trench01_amd import autodiff


class LinearBatch:
    """increment = A theta + B rho (in place of the traced batch)."""

    def __init__(self, m=5, seed=0):
        rng = np.random.default_rng(seed)
        self.A = rng.standard_normal((m, 2))
        self.B = rng.standard_normal((m, m))

    def ion_jacobian(self, rho, sigma, shift):
        rho = np.zeros(self.B.shape[0]) if rho is None else rho
        return self.A @ np.array([sigma, shift]) + self.B @ rho, self.A.copy()

    def increment(self, rho):
        return self.B @ rho

    def charge_jvp(self, rho, drho):
        return self.B @ rho, self.B @ drho

    def vjp(self, rho, g):
        return self.B @ rho, self.B.T @ g, self.A.T @ g


def test_ion_increment_backward_and_jvp():
    b = LinearBatch()
    theta = torch.tensor([3.0, 0.5], dtype=torch.float64, requires_grad=True)
    out = autodiff.IonIncrement.apply(theta, b)
    np.testing.assert_allclose(out.detach().numpy(), b.A @ [3.0, 0.5])
    w = torch.arange(5, dtype=torch.float64)
    (out @ w).backward()
    np.testing.assert_allclose(theta.grad.numpy(), b.A.T @ w.numpy())
    v = torch.tensor([1.0, -2.0], dtype=torch.float64)
    with fwAD.dual_level():
        t = fwAD.make_dual(torch.tensor([3.0, 0.5], dtype=torch.float64), v)
        tangent = fwAD.unpack_dual(autodiff.IonIncrement.apply(t, b)).tangent
    np.testing.assert_allclose(tangent.numpy(), b.A @ v.numpy())


def test_charge_increment_backward_is_the_transpose_of_its_jvp():
    b = LinearBatch(m=6, seed=1)
    rho = torch.tensor(np.random.default_rng(2).standard_normal(6), requires_grad=True)
    out = autodiff.ChargeIncrement.apply(rho, b)
    g = torch.tensor(np.random.default_rng(3).standard_normal(6))
    (out @ g).backward()
    np.testing.assert_allclose(rho.grad.numpy(), b.B.T @ g.numpy(), rtol=1e-12)
    v = torch.tensor(np.random.default_rng(4).standard_normal(6))
    with fwAD.dual_level():
        d = fwAD.make_dual(rho.detach(), v)
        tangent = fwAD.unpack_dual(autodiff.ChargeIncrement.apply(d, b)).tangent
    np.testing.assert_allclose(tangent.numpy(), b.B @ v.numpy(), rtol=1e-12)


class ScoreBatch:
    """increment = a sigma, gradient by the (here exact) likelihood ratio a."""

    def __init__(self, m=4):
        self.a = np.arange(1.0, m + 1.0)

    def sigma_score_gradient(self, rho, sigma, shift):
        return self.a * sigma, self.a.copy()


def test_sigma_increment_backward_and_jvp():
    b = ScoreBatch()
    sigma = torch.tensor(3.0, dtype=torch.float64, requires_grad=True)
    out = autodiff.SigmaIncrement.apply(sigma, b)
    np.testing.assert_allclose(out.detach().numpy(), 3.0 * b.a)
    w = torch.tensor([1.0, -1.0, 2.0, 0.5], dtype=torch.float64)
    (out @ w).backward()
    assert float(sigma.grad) == pytest.approx(float(b.a @ w.numpy()))
    with fwAD.dual_level():
        d = fwAD.make_dual(torch.tensor(3.0, dtype=torch.float64), torch.tensor(0.25, dtype=torch.float64))
        tan = fwAD.unpack_dual(autodiff.SigmaIncrement.apply(d, b)).tangent
    np.testing.assert_allclose(tan.numpy(), 0.25 * b.a)
