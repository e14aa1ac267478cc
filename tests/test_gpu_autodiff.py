"""torch.autograd over the device's pathwise batch derivative
(trench01_amd.autodiff over tangent.PathwiseBatch): the forward value is the
batch's face-smoothed charge increment, backward and forward-mode products
use the tangent push, and the surface-charge backward (one tangent push per
surface cell) is the transpose of the surface-charge JVP."""
from dataclasses import replace

import numpy as np
import pytest
import torch
import torch.autograd.forward_ad as fwAD

from trench01_amd import autodiff, model, tangent

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def small_params():
    return replace(model.Params(), nx=30, trench_w=10, trench_d=20, floor_t=5, mask_t=6, n_vac=12, bias="dc",
                   sigma_s=0.0, n_per_batch=3000, max_steps=1000, wall_model="barrier", mask_type="conductor",
                   field_mode="floating", ion_distribution="mono", ion_energy_eV=100.0, ion_angle_deg=3.0)


@pytest.fixture(scope="module")
def batch():
    b = tangent.PathwiseBatch(small_params(), seed=3)
    rng = np.random.default_rng(8)
    rho = np.where(b.cond_s, 0.0, rng.uniform(-2e7, 4e7, b.surface.size))
    return b, rho


def test_ion_increment_through_autograd(batch):
    _case("autodiff ion increment")
    b, rho = batch
    inc, jac = b.ion_jacobian(rho, 3.0, 0.0)
    assert np.abs(jac).max() > 0 and np.abs(inc).max() > 0
    np.testing.assert_array_equal(b.increment(rho, 3.0, 0.0), inc)   # same random stream, same batch
    theta = torch.tensor([3.0, 0.0], dtype=torch.float64, requires_grad=True)
    out = autodiff.IonIncrement.apply(theta, b, rho)
    np.testing.assert_array_equal(out.detach().numpy(), inc)
    w = torch.tensor(np.random.default_rng(1).standard_normal(inc.size))
    (out @ w).backward()
    np.testing.assert_allclose(theta.grad.numpy(), jac.T @ w.numpy(), rtol=1e-12)
    v = torch.tensor([0.5, -1.0], dtype=torch.float64)
    with fwAD.dual_level():
        t = fwAD.make_dual(torch.tensor([3.0, 0.0], dtype=torch.float64), v)
        tan = fwAD.unpack_dual(autodiff.IonIncrement.apply(t, b, rho)).tangent
    np.testing.assert_allclose(tan.numpy(), jac @ v.numpy(), rtol=1e-12)


def test_charge_backward_is_the_transpose_of_the_charge_jvp(batch):
    _case("autodiff charge adjoint identity")
    b, rho = batch
    rng = np.random.default_rng(5)
    g = torch.tensor(rng.standard_normal(rho.size))
    v = torch.tensor(np.where(b.cond_s, 0.0, rng.standard_normal(rho.size)) * 1e7)
    r = torch.tensor(rho, requires_grad=True)
    out = autodiff.ChargeIncrement.apply(r, b)
    (out @ g).backward()
    with fwAD.dual_level():
        d = fwAD.make_dual(torch.tensor(rho), v)
        jv = fwAD.unpack_dual(autodiff.ChargeIncrement.apply(d, b)).tangent
    lhs = float(r.grad @ v)
    rhs = float(g @ jv)
    print(f"  <J^T g, v> = {lhs:.6e}, <g, J v> = {rhs:.6e}", flush=True)
    assert abs(rhs) > 0
    assert lhs == pytest.approx(rhs, rel=1e-9)


@pytest.fixture(scope="module")
def charged_ar8():
    """A charged AR 8 trench: 300 batches of the charging loop (same charge per
    batch at 2.5e4 particles per species), and a 2e5-per-species batch on it."""
    from trench01_amd import charging, configs
    p, _ = configs.product_params("AR8", n_per_batch=200_000)
    pre = charging.run(replace(p, n_per_batch=25_000, n_batches=300, until_steady=False), log=None,
                       injection="device")
    b = tangent.PathwiseBatch(p, seed=7)
    rho = np.asarray(pre["rho_raw"]).ravel()[b.surface]
    yield p, b, rho
    b.close()


def test_sigma_score_gradient_follows_the_expected_increment(charged_ar8):
    """The likelihood-ratio gradient with respect to the ion angular sigma
    estimates the derivative of the expected increment: on a charged AR 8
    trench it points along the central difference of the increment at a
    20 % step with common random numbers (reproducible between streams), and
    SigmaIncrement's backward is that gradient."""
    _case("autodiff sigma score gradient")
    p, b, rho = charged_ar8
    sig = p.ion_angle_deg
    inc, grad = b.sigma_score_gradient(rho, sig, 0.0)
    np.testing.assert_array_equal(inc, b.increment(rho, sig))
    fd = (b.increment(rho, sig * 1.2) - b.increment(rho, sig * 0.8)) / (0.4 * sig)
    cos = float(grad @ fd / (np.linalg.norm(grad) * np.linalg.norm(fd)))
    ratio = float(np.linalg.norm(grad) / np.linalg.norm(fd))
    print(f"  cos(score gradient, difference) {cos:.3f}, norm ratio {ratio:.3f}", flush=True)
    assert cos > 0.8 and 0.6 < ratio < 1.6
    s = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    out = autodiff.SigmaIncrement.apply(s, b, rho)
    w = torch.tensor(np.random.default_rng(2).standard_normal(inc.size))
    (out @ w).backward()
    assert float(s.grad) == pytest.approx(float(grad @ w.numpy()), rel=1e-12)


def test_energy_score_gradient_follows_the_expected_increment(charged_ar8):
    """The likelihood-ratio gradient with respect to the mean ion energy along
    the exponentially tilted family (arcsine 40-260 eV) points along the
    central difference of the increment over that family (tilts moving the
    mean by +-10 %, common random numbers), and EnergyIncrement's backward is
    that gradient."""
    _case("autodiff energy score gradient")
    p, b, rho = charged_ar8
    sampler = tangent.ion_sampler(p)
    mean, var = sampler.energy_moments()
    inc, grad = b.energy_score_gradient(rho)
    np.testing.assert_array_equal(inc, b.increment(rho))
    t = 0.1 * mean / var
    a = np.linspace(0.0, 0.5 * np.pi, 200_001)
    e = p.ion_energy_lo_eV + (p.ion_energy_hi_eV - p.ion_energy_lo_eV) * np.sin(a) ** 2

    def tilted_mean(tt):
        w = np.exp(tt * (e - mean))
        return float(np.trapz(w * e, a) / np.trapz(w, a))
    fd = (b.increment_tilted(rho, t) - b.increment_tilted(rho, -t)) / (tilted_mean(t) - tilted_mean(-t))
    cos = float(grad @ fd / (np.linalg.norm(grad) * np.linalg.norm(fd)))
    ratio = float(np.linalg.norm(grad) / np.linalg.norm(fd))
    print(f"  cos(energy score gradient, difference) {cos:.3f}, norm ratio {ratio:.3f}", flush=True)
    assert cos > 0.8 and 0.6 < ratio < 1.6
    m = torch.tensor(mean, dtype=torch.float64, requires_grad=True)
    out = autodiff.EnergyIncrement.apply(m, b, rho)
    w = torch.tensor(np.random.default_rng(3).standard_normal(inc.size))
    (out @ w).backward()
    assert float(m.grad) == pytest.approx(float(grad @ w.numpy()), rel=1e-12)
