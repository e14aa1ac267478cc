"""The relaxed-barrier push (th_trace_soft) and the Jacobian of the batch map
built on it (trench01_amd.relaxed): the device's face-smoothed deposits
against the host's over its deposit list, the sharp limit against the
tangent push, the relaxed decision against its closed form on straight
flights, the derivative against central differences of the same relaxed
model with common random numbers, and the host charge update against the
device's."""
from dataclasses import replace

import numpy as np
import pytest

from trench01_amd import charging, configs, model, relaxed, tangent
from trench01_amd.tracer import HipTracer

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.fixture(scope="module")
def charged():
    """A charged AR 8 trench (150 batches of 2.5e4 particles per species at
    the full batch's charge per batch) and a 2e4-per-species batch on it."""
    p, _ = configs.product_params("AR8", n_per_batch=25_000)
    pre = charging.run(replace(p, n_batches=150, until_steady=False), log=None, injection="device")
    b = tangent.PathwiseBatch(replace(p, n_per_batch=20_000), seed=7)
    rho = np.asarray(pre["rho_raw"]).ravel()[b.surface]
    yield b, rho
    b.close()


def test_device_deposit_is_the_face_deposit_of_its_events(charged):
    """The deposits accumulated on the device equal the host's face-smoothed
    deposit of the listed events; an absorbed particle's shares sum to one."""
    _case("relaxed push: device vs host deposit")
    b, rho = charged
    inc, dinc, r = b.relaxed(rho, 1.0, drho_s=rho, events=400_000)
    ev = r["events"]
    soft = int(np.sum(ev["kind"] == 0))
    print(f"  {ev.size} deposits, {soft} relaxed decisions", flush=True)
    assert soft > 1000
    dep, ddep = tangent.event_deposit(b.geo.vac, b.surface, ev, b.p.dx)
    np.testing.assert_allclose(r["deposit"], dep, rtol=0, atol=1e-12 * np.abs(dep).max())
    np.testing.assert_allclose(r["d_deposit"], ddep, rtol=0, atol=1e-12 * np.abs(ddep).max())
    n = r["status"].size
    w = np.bincount(ev["pid"], ev["a"], minlength=n)
    absorbed = r["status"] == 1
    np.testing.assert_allclose(w[absorbed], 1.0, rtol=0, atol=1e-13)
    assert np.all(w[~absorbed] < 1.0 + 1e-13)
    assert np.all(ev["kind"][np.isin(ev["pid"], np.flatnonzero(~absorbed))] == 0)


def test_sharp_limit_is_the_tangent_push(charged):
    """With a width far below any margin the relaxed push takes no relaxed
    decision: its deposit is the face-smoothed deposit of the plain record
    push and its derivative the pathwise derivative of the tangent push."""
    _case("relaxed push: sharp limit")
    b, rho = charged
    inc, dinc, r = b.relaxed(rho, 1e-12, drho_s=rho, events=200_000)
    assert np.sum(r["events"]["kind"] == 0) == 0
    inc_t, dinc_t = b.charge_jvp(rho, rho)
    # the device sums deposits with FP64 atomics, the host by bincount: equal up to the order of the sums
    np.testing.assert_allclose(inc, inc_t, rtol=0, atol=1e-13 * np.abs(inc_t).max())
    np.testing.assert_allclose(dinc, dinc_t, rtol=0, atol=1e-12 * np.abs(dinc_t).max())


def test_relaxed_decision_closed_form():
    """Field-free vertical flights onto a floor held at V0 above the vacuum:
    each ion deposits 1 - p, p = 1 / (1 + exp((E - wface V0) / w)), and the
    reflected share escapes; along a unit potential tangent on the floor the
    deposit moves by -p (1 - p) wface / w per ion."""
    _case("relaxed push: closed form")
    p = replace(model.Params(), nx=30, trench_w=10, trench_d=20, floor_t=5, mask_t=6, n_vac=12, bias="dc",
                sigma_s=0.0, wall_model="barrier", mask_type="dielectric", field_mode="floating", max_steps=2000,
                trench_taper_deg=0.0, mask_taper_deg=0.0)
    geo = model.Geometry(p)
    fields = model.Fields(p, geo)
    vac = geo.vac
    nx, ny = vac.shape
    floor = np.zeros(vac.shape, bool)
    floor[:, p.floor_t - 1] = True
    V0, w = 40.0, 0.7
    phi = np.where(floor, V0, 0.0)
    wf = float(fields.wface[floor][0])
    t = HipTracer(vac)
    surf = t.surface_cells()
    x0, x1 = geo.opening_span(p.floor_t)
    n = 4000
    rng = np.random.default_rng(3)
    E = wf * V0 + rng.uniform(-6 * w, 6 * w, n)
    qm = model.E_CHARGE / (p.ion_mass_amu * model.AMU)
    x = rng.uniform(x0 + 0.5, x1 - 0.5, n) * p.dx
    y = np.full(n, (ny - 0.5) * p.dx)
    vx = np.zeros(n)
    vy = -np.sqrt(2.0 * E * qm)
    t.reserve(n)
    t.upload(0, x, y, vx, vy, np.full(n, qm), np.zeros(n, np.int8))
    zero = np.zeros(vac.shape)
    t.set_fields(zero, zero, (phi, fields.wface))
    r = t.soft(0, n, p, w, 30.0, dphi=floor.astype(float), events=4 * n, per_particle=True)
    pr = 1.0 / (1.0 + np.exp((0.5 * vy ** 2 / qm - wf * V0) / w))
    print(f"  deposit {r['deposit'][0].sum():.6f} vs {np.sum(1 - pr):.6f}", flush=True)
    assert np.all(np.isin(r["status"], (1, 2)))
    np.testing.assert_allclose(r["deposit"][0].sum(), np.sum(1.0 - pr), rtol=1e-12)
    np.testing.assert_allclose(r["d_deposit"][0].sum(), np.sum(-pr * (1.0 - pr) * wf / w), rtol=1e-11)
    on_floor = np.isin(surf, np.flatnonzero(floor.ravel()))
    assert np.all(r["deposit"][0][~on_floor] == 0.0)
    t.close()


def test_relaxed_derivative_matches_central_differences(charged):
    """Per particle, the derivative of every deposit (weight and crossing
    point) along the charge's own direction equals the central difference of
    the relaxed push at rho (1 +- 1e-5) with the same particles, for the
    particles whose events (cells, faces, kinds, step count) stay the same
    (the largest errors are those of long electron paths, whose sensitivity
    makes the difference nonlinear at any step that changes no event)."""
    _case("relaxed push: derivative vs differences")
    b, rho = charged
    h, cap = 1e-5, 400_000
    _, _, r0 = b.relaxed(rho, 1.0, drho_s=rho, events=cap)
    _, _, rp = b.relaxed(rho * (1 + h), 1.0, events=cap)
    _, _, rm = b.relaxed(rho * (1 - h), 1.0, events=cap)
    n = r0["status"].size
    order = lambda e: e[np.argsort(e["pid"], kind="stable")]   # a particle's deposits in the order made
    e0, ep, em = order(r0["events"]), order(rp["events"]), order(rm["events"])
    c0, cp, cm = (np.bincount(e["pid"], minlength=n) for e in (e0, ep, em))
    same = (c0 == cp) & (c0 == cm) & (r0["steps"] == rp["steps"]) & (r0["steps"] == rm["steps"])
    a0, ap, am = e0[same[e0["pid"]]], ep[same[ep["pid"]]], em[same[em["pid"]]]
    ok = np.ones(a0.size, bool)
    for f in ("cell", "face", "kind"):
        ok &= (a0[f] == ap[f]) & (a0[f] == am[f])
    bad = np.zeros(n, bool)
    bad[a0["pid"][~ok]] = True
    keep = ~bad[a0["pid"]]
    print(f"  particles keeping their events: {np.mean(same & ~bad):.4f}", flush=True)
    assert np.mean(same & ~bad) > 0.95
    soft = a0["kind"][keep] == 0
    assert soft.sum() > 1000
    xf = a0["face"][keep] == 1
    # the weight, and the crossing point's coordinate along the crossed face (across it the point is the face)
    quantities = {"weight": (ap["a"][keep] - am["a"][keep], a0["da"][keep]),
                  "along the face": (np.where(xf, ap["y"][keep] - am["y"][keep], ap["x"][keep] - am["x"][keep]),
                                     np.where(xf, a0["dy"][keep], a0["dx"][keep]))}
    across = np.where(xf, a0["dx"][keep], a0["dy"][keep])
    assert np.abs(across).max() <= 1e-9 * np.abs(quantities["along the face"][1]).max()
    for name, (diff, d) in quantities.items():
        fd = diff / (2 * h)
        moving = np.abs(d) > 0
        floor = 1e-6 * np.median(np.abs(d[moving]))
        err = np.abs(fd - d)[moving] / (np.abs(d[moving]) + floor)
        q = np.quantile(err, [0.5, 0.9, 0.95, 0.99])
        qs = np.quantile(np.abs(fd - d)[moving & soft] / (np.abs(d[moving & soft]) + floor), [0.5, 0.95])
        print(f"  {name}: relative error median {q[0]:.2e}, 90% {q[1]:.2e}, 95% {q[2]:.2e}, 99% {q[3]:.2e}; relaxed "
              f"decisions median {qs[0]:.2e}, 95% {qs[1]:.2e}", flush=True)
        assert q[0] < 1e-6 and q[2] < 1e-3 and q[3] < 5e-2
        assert qs[0] < 1e-6 and qs[1] < 1e-3
        assert np.abs(fd[~moving]).max(initial=0.0) <= 1e-6 * np.abs(d).max()


def test_charge_update_restates_the_device_update():
    """relaxed.ChargeUpdate applied to a device batch's counts reproduces the
    device's charge after that batch (deposit, leak, floating conductor)."""
    _case("relaxed: host charge update vs device")
    p, _ = configs.product_params("AR8", n_per_batch=20_000)
    setup = charging.ChargingSetup(p)
    dev = charging.DeviceCharging(setup)
    try:
        rng = np.random.default_rng(4)
        rho0 = np.where(setup.cond_s.astype(bool), 0.0, rng.uniform(-1e7, 3e7, setup.geo.surface.size))
        dev.reserve(2 * p.n_per_batch)
        dev.set_rho(rho0)
        dev.inject_device(0, p.n_per_batch, 0, 5)
        dev.step(0, 2 * p.n_per_batch)
        rho1, _, counts = dev.read(phi=False)
        m = setup.geo.surface.size
        up = relaxed.ChargeUpdate(setup)
        host, _ = up.apply(rho0, counts[:m].astype(float), counts[m:].astype(float), setup.q_macro)
        np.testing.assert_allclose(host, rho1, rtol=0, atol=1e-12 * np.abs(rho1).max())
    finally:
        dev.close()


def test_map_jacobian_directions_and_linearity():
    """MapJacobian's potential tangent (Green's matrix) equals the LU solve of
    the same charge direction on the rows the push reads, and its JVPs are
    linear in the direction; with a lower step limit its deposit is the
    relaxed push's with that limit."""
    _case("relaxed: map Jacobian")
    p, _ = configs.product_params("AR8", n_per_batch=20_000)
    setup = charging.ChargingSetup(p)
    dev = charging.DeviceCharging(setup)
    try:
        n = 2 * p.n_per_batch
        dev.reserve(n)
        rng = np.random.default_rng(6)
        rho = np.where(setup.cond_s.astype(bool), 0.0, rng.uniform(0, 2e7, setup.geo.surface.size))
        dev.inject_device(0, p.n_per_batch, 0, 0)
        jac = relaxed.MapJacobian(setup, dev, 1.0)
        jac.at(rho, n, setup.q_macro)
        k = setup.green.shape[1]
        du, dv = rng.standard_normal(k), rng.standard_normal(k)
        dQ, drho, dphi, _, _ = jac.directional(du)
        full = np.zeros(setup.geo.solid.size)
        full[setup.geo.surface] = drho
        dphi_lu, _, _ = tangent.charge_direction(setup.fields, setup.geo.vac, p, full.reshape(setup.geo.solid.shape))
        rows = setup.rows
        np.testing.assert_allclose(dphi.ravel()[rows], dphi_lu.ravel()[rows], rtol=0,
                                   atol=1e-10 * np.abs(dphi_lu).max())
        np.testing.assert_allclose(jac.P @ jac.totals(drho), du, rtol=0, atol=1e-10 * np.abs(du).max())
        ju, jv, jw = jac(du), jac(dv), jac(2.0 * du - 0.5 * dv)
        np.testing.assert_allclose(jw, 2.0 * ju - 0.5 * jv, rtol=0, atol=1e-9 * np.abs(jw).max())
        assert np.abs(ju - du).max() > 0
        # paths ended early: the relaxed push of the same particles with that step limit
        capped = relaxed.MapJacobian(setup, dev, 1.0, max_steps=300)
        assert capped.p_push.max_steps == 300 and jac.p_push.max_steps == p.max_steps
        dep_c = capped.at(rho, n, setup.q_macro)
        ref = dev.tracer.soft(0, n, replace(p, max_steps=300), 1.0)
        np.testing.assert_allclose(dep_c, ref["deposit"], rtol=1e-12, atol=0)
        assert dep_c.sum() < jac.deposit.sum()
        jc, jd = capped(du), capped(2.0 * du)
        np.testing.assert_allclose(jd, 2.0 * jc, rtol=0, atol=1e-9 * np.abs(jd).max())
    finally:
        dev.close()


def test_relaxed_adjoint_is_the_transpose_of_the_relaxed_tangent(charged):
    """th_trace_soft_adjoint against th_trace_soft on the same particles and
    fields: for random deposit cotangents g and random directions of the
    initial states, the field and the potential, <g, d deposit> equals
    <bar_state, d state> + <bar_field, d field> + <bar_phi, d phi>."""
    _case("relaxed push: adjoint identity")
    b, rho = charged
    t, p = b.tracer, b.p
    sampler = tangent.ion_sampler(p)
    parts, _ = tangent.draw_batch(p, b.geo.ny, b.seed, sampler, b.n)
    phi, ex, ey = b.field(rho)
    n = parts[0].size
    t.reserve(n)
    t.upload(0, *parts)
    t.set_fields(ex, ey, (phi, b.fields.wface))
    rng = np.random.default_rng(12)
    vac = b.geo.vac
    ds = [rng.standard_normal(n) * s for s in (1e-10, 1e-10, 1e2, 1e2)]
    dex, dey = (rng.standard_normal(vac.shape) * 1e5 * vac for _ in range(2))
    dphi = rng.standard_normal(vac.shape) * 0.1
    g = rng.standard_normal((2, b.surface.size))
    for width in (1.0, 0.2):
        r = t.soft(0, n, p, width, 30.0, ds, (dex, dey), dphi, per_particle=True)
        bs, bf, bp = t.soft_adjoint(0, n, p, width, r["steps"], g)
        lhs = float(np.sum(g * r["d_deposit"]))
        parts_rhs = [sum(float(bs[:, c] @ ds[c]) for c in range(4)),
                     float(np.sum(bf[..., 0] * dex + bf[..., 1] * dey)), float(np.sum(bp * dphi))]
        rhs = sum(parts_rhs)
        print(f"  width {width}: <g, J v> = {lhs:.9e}, <J^T g, v> = {rhs:.9e} (state, field, potential "
              f"{parts_rhs[0]:.3e}, {parts_rhs[1]:.3e}, {parts_rhs[2]:.3e}); adjoint {t.last_adjoint_ms:.1f} ms",
              flush=True)
        assert all(abs(x) > 0 for x in parts_rhs)
        assert rhs == pytest.approx(lhs, rel=1e-9)


def test_relaxed_batch_gradient_is_the_transpose_of_its_jvps(charged):
    """PathwiseBatch.relaxed_vjp (one relaxed push and its adjoint, the field
    stencil's and G's transposes) against the relaxed JVPs along a surface-
    charge direction and along the two ion parameters."""
    _case("relaxed push: batch gradient")
    b, rho = charged
    rng = np.random.default_rng(13)
    g = rng.standard_normal(b.surface.size)
    v = np.where(b.cond_s, 0.0, rng.standard_normal(b.surface.size)) * 1e7
    inc, grad, ion = b.relaxed_vjp(rho, g, 1.0)
    inc1, jv, _ = b.relaxed(rho, 1.0, drho_s=v)
    np.testing.assert_allclose(inc, inc1, rtol=0, atol=1e-13 * np.abs(inc1).max())   # FP64 atomics' order
    assert float(grad @ v) == pytest.approx(float(g @ jv), rel=1e-9)
    for k, key in enumerate(("sigma", "energy")):
        _, jk, _ = b.relaxed(rho, 1.0, key=key)
        assert ion[k] == pytest.approx(float(g @ jk), rel=1e-9)
