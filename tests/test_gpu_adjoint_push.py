"""Reverse-mode pathwise derivative of the push (th_trace_adjoint) on the device.

* It is the transpose of the forward-mode tangent push (th_trace_tangent):
  for random cotangents of the recorded end points and final velocities,
  <cotangents, tangent of the outputs> equals <adjoint, tangent of the
  inputs>, per particle for initial-state directions and summed over the
  batch for a field direction (the field of a surface-charge direction), to
  1e-9 relative.
* It agrees with central finite differences with common random numbers at a
  relative step of 1e-4 of the oracle's C push: along the ion angular sigma
  and the energy shift per particle, and along a surface-charge direction
  summed over the particles that keep their discrete events (status, hit
  cell, step count).
* The batch's vector-Jacobian product (PathwiseBatch.vjp: one primal trace
  and one adjoint push) is the transpose of its JVPs (charge_jvp,
  ion_jacobian) on the AR 8 trench, to 1e-9.
* Path lengths that do not match the particles are refused with an error.
"""
import time

import numpy as np
import pytest

from tests.test_gpu_tangent import H, _directions, draw, oracle_record, trench  # noqa: F401 (fixture)
from trench01_amd import HipTracer, configs, tangent

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def _adjoint_and_tangent(case, rng, dstate=None, dfield=None):
    parts, _ = draw(case)
    n = parts[0].size
    t = HipTracer(case["vac"])
    p = case["p"]
    barrier = (case["phi"], case["wface"])
    r = t.tangent(*parts, case["Ex"], case["Ey"], p, barrier, dstate=dstate, dfield=dfield)
    bar_end = rng.standard_normal((n, 2)) / p.dx
    bar_vel = rng.standard_normal((n, 2)) / np.abs(r["final_velocity"]).max()
    bar_state, bar_field = t.adjoint(*parts, case["Ex"], case["Ey"], p, barrier, r["steps"], bar_end, bar_vel)
    return parts, r, bar_end, bar_vel, bar_state, bar_field


def test_adjoint_is_the_transpose_of_the_tangent_per_particle(trench):
    _case("adjoint vs tangent, initial-state direction")
    rng = np.random.default_rng(41)
    n = 2 * trench["n"]
    p = trench["p"]
    vs = 1e5
    dstate = (rng.standard_normal(n) * p.dx, rng.standard_normal(n) * p.dx, rng.standard_normal(n) * vs,
              rng.standard_normal(n) * vs)
    parts, r, be, bv, bs, _ = _adjoint_and_tangent(trench, rng, dstate=dstate)
    lhs = (be * r["d_endpoints"]).sum(axis=1) + (bv * r["d_final_velocity"]).sum(axis=1)
    rhs = (bs * np.stack(dstate, axis=1)).sum(axis=1)
    scale = np.abs(be * r["d_endpoints"]).sum(axis=1) + np.abs(bv * r["d_final_velocity"]).sum(axis=1) + \
        np.abs(bs * np.stack(dstate, axis=1)).sum(axis=1)
    err = np.abs(lhs - rhs) / np.maximum(scale, 1e-300)
    print(f"  relative transpose error: median {np.median(err):.2e}, max {err.max():.2e}", flush=True)
    assert err.max() < 1e-9
    assert np.abs(bs).max() > 0


def test_adjoint_field_cotangent_is_the_transpose_of_the_field_tangent(trench):
    _case("adjoint vs tangent, field direction")
    rng = np.random.default_rng(42)
    vac = trench["vac"]
    parts, r, be, bv, bs, bf = _adjoint_and_tangent(trench, rng, dfield=(trench["dEx"], trench["dEy"]))
    lhs = float((be * r["d_endpoints"]).sum() + (bv * r["d_final_velocity"]).sum())
    rhs = float((bf[..., 0] * trench["dEx"] * vac).sum() + (bf[..., 1] * trench["dEy"] * vac).sum())
    scale = float(np.abs(be * r["d_endpoints"]).sum() + np.abs(bv * r["d_final_velocity"]).sum())
    print(f"  <bar, J v> = {lhs:.12e}, <J^T bar, v> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-9 * scale


@pytest.mark.parametrize("direction", ["sigma", "energy", "charge"])
def test_adjoint_matches_central_differences(trench, direction):
    """d/dtheta of sum_p <bar_end_p, end_p> from the adjoint against central
    differences of the oracle's record traces (common random numbers), over
    the particles whose events stay the same."""
    _case(f"adjoint vs differences: {direction}")
    states, flds, step, kw = _directions(trench)[direction]
    p = trench["p"]
    t = HipTracer(trench["vac"])
    barrier = (trench["phi"], trench["wface"])
    mid = t.tangent(*states[1], trench["Ex"], trench["Ey"], p, barrier)
    lo, hi = (oracle_record(trench, s, *f) for s, f in ((states[0], flds[0]), (states[2], flds[2])))
    same = mid["status"] != 3
    for ref in (lo, hi):
        same &= (ref["status"] == mid["status"]) & (ref["hit"] == mid["hit"]) & (ref["steps"] == mid["steps"])
    assert same.mean() > 0.95
    rng = np.random.default_rng(43)
    n = same.size
    bar_end = rng.standard_normal((n, 2)) * same[:, None] / p.dx
    bs, bf = t.adjoint(*states[1], trench["Ex"], trench["Ey"], p, barrier, mid["steps"], bar_end)
    fd = ((hi["endpoints"] - lo["endpoints"]) / step * bar_end).sum(axis=1)
    if direction == "charge":
        # Summed over particles (the field cotangent is a sum).  The long,
        # chaotic electron paths limit the difference quotient itself to about
        # 1e-4 per particle; the ions' short paths resolve it to 1e-5 or better.
        vac = trench["vac"]
        ions = states[1][5] == 0
        for label, keep, tol in (("ions", same & ions, 1e-5), ("all", same, 1e-3)):
            be = bar_end * keep[:, None]
            _, bf = t.adjoint(*states[1], trench["Ex"], trench["Ey"], p, barrier, mid["steps"], be)
            ad = float((bf[..., 0] * trench["dEx"] * vac).sum() + (bf[..., 1] * trench["dEy"] * vac).sum())
            f = fd * keep
            print(f"  {label}: adjoint {ad:.9e}, difference {f.sum():.9e} (sum of |terms| {np.abs(f).sum():.3e})",
                  flush=True)
            assert abs(ad - f.sum()) <= tol * np.abs(f).sum()
    else:
        dv = kw["dstate"]
        ad = bs[:, 2] * dv[2] + bs[:, 3] * dv[3]
        moving = same & (np.abs(ad) > 0)
        assert moving.mean() > 0.3
        err = np.abs(ad - fd)[moving] / (np.abs(ad)[moving] + 1e-6 * np.median(np.abs(ad[moving])))
        q = np.quantile(err, [0.5, 0.99])
        print(f"  per-particle relative error: median {q[0]:.2e}, 99% {q[1]:.2e}", flush=True)
        assert q[0] < 1e-5 and q[1] < 1e-3
        np.testing.assert_array_equal(fd[same & (np.abs(ad) == 0)], 0.0)


def test_batch_vjp_is_the_transpose_of_its_jvps():
    """PathwiseBatch.vjp on the tapered AR 8 trench (arcsine ions, conductive
    floating mask), charged: <vjp, v> = <g, charge_jvp(v)> for random
    surface-charge directions, and the ion gradient = g . ion_jacobian."""
    _case("batch VJP vs JVPs, AR8")
    p, _ = configs.product_params("AR8", n_per_batch=20_000)
    b = tangent.PathwiseBatch(p, seed=5)
    rng = np.random.default_rng(44)
    m = b.surface.size
    rho = np.where(b.cond_s, 0.0, rng.normal(0, 2e6, m))
    rho[b.setup.c_idx] = -3e8
    g = rng.standard_normal(m)
    t0 = time.perf_counter()
    inc, grad_rho, grad_ion = b.vjp(rho, g)
    t_vjp = time.perf_counter() - t0
    inc2, jac = b.ion_jacobian(rho)
    np.testing.assert_array_equal(inc, inc2)
    np.testing.assert_allclose(grad_ion, jac.T @ g, rtol=1e-9)
    for k in range(3):
        v = rng.standard_normal(m) * 1e6
        t0 = time.perf_counter()
        _, dinc = b.charge_jvp(rho, v)
        t_jvp = time.perf_counter() - t0
        lhs, rhs = float(g @ dinc), float(grad_rho @ v)
        print(f"  <g, J v> = {lhs:.12e}, <J^T g, v> = {rhs:.12e}", flush=True)
        assert lhs == pytest.approx(rhs, rel=1e-9, abs=1e-9 * float(np.abs(g) @ np.abs(dinc)))
    print(f"  VJP {t_vjp * 1e3:.0f} ms (all {m} surface cells and 2 ion parameters) vs one JVP {t_jvp * 1e3:.0f} ms",
          flush=True)
    b.close()


def test_adjoint_refuses_path_lengths_of_other_particles(trench):
    _case("adjoint step-count check")
    parts, _ = draw(trench)
    t = HipTracer(trench["vac"])
    p = trench["p"]
    barrier = (trench["phi"], trench["wface"])
    r = t.tangent(*parts, trench["Ex"], trench["Ey"], p, barrier)
    for delta in (1, -1):
        steps = r["steps"].copy()
        k = int(np.flatnonzero((steps > 1) & (steps < p.max_steps))[0])
        steps[k] += delta
        with pytest.raises(ValueError, match="different number of steps"):
            t.adjoint(*parts, trench["Ex"], trench["Ey"], p, barrier, steps, np.zeros((steps.size, 2)))
