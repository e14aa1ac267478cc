"""Pathwise derivative of the push (th_trace_tangent) on the device.

* Its primal is the record-mode FP64 trace: status, hit cell, step count, end
  point and final velocity equal the oracle's bit for bit.
* Its derivative equals central finite differences with common random
  numbers at a relative step of 1e-4, per particle, for every particle whose
  discrete events (status, hit cell, step count) are the same at theta - h,
  theta and theta + h; three directions: the ion angular sigma, a shift of
  the ion energy distribution, and the surface charge (the field and wall
  potential of a charge direction).  The differences come from the oracle's
  C push.
* In a field-free trench the crossing point on the floor is a closed form of
  the launch angle: d x_floor / d sigma = sec^2(theta) (theta / sigma) (y0 -
  y_floor) and d x_floor / dE = 0.
* The face deposit's derivative equals the central difference of the face
  deposit over the particles that keep their events.
"""
from dataclasses import replace

import numpy as np
import pytest

from oracle import charging_ref as R
from oracle import trace_ref
from tests import cases
from trench01_amd import HipTracer, configs, tangent
from trench01_amd.tracer import tapered_of

pytestmark = pytest.mark.gpu

H = 1e-4   # relative step of the central differences


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.fixture(scope="module", params=["ar5", "ar8"])
def trench(request):
    """AR 5 (mono 100 eV ions) and AR 8 (tapered walls, arcsine 40-260 eV ions)
    trenches with a charged barrier-model field, 4000 particles per species."""
    if request.param == "ar5":
        p, ion = cases.ar5_params(), configs.AR5_ION
    else:
        p, ion = cases.ar8_params(), configs.AR8_ION
    c = cases.charged_case(p, n=10)
    solid, mask, i0, i1 = R.build_geometry(p)
    fields = R.setup_fields(p, solid, mask)
    rho = np.zeros(solid.shape)
    rho[i0 - 1, p.floor_t:p.floor_t + p.trench_d] = 2e7
    rho[i1, p.floor_t:p.floor_t + p.trench_d] = -1e7
    dphi = fields["solve"]((-rho / R.EPS0).ravel()).reshape(solid.shape)
    dex, dey = tangent.linear_field(dphi, ~solid, p)
    return dict(p=p, vac=c["vac"], Ex=c["Ex"], Ey=c["Ey"], phi=c["barrier"][0], wface=c["barrier"][1],
                dphi=dphi, dEx=dex, dEy=dey, ion=ion, n=4000, seed=31)


def draw(case, sigma_scale=1.0, shift=0.0):
    ion = case["ion"]
    s = tangent.IonSampler(case["p"].ion_mass_amu, ion["kind"], ion["sigma_deg"] * sigma_scale,
                           energy_eV=ion.get("energy_eV"), e_lo=ion.get("e_lo"), e_hi=ion.get("e_hi"),
                           energy_shift=shift)
    parts = R.inject(case["p"], case["n"], case["vac"].shape[1], np.random.default_rng(case["seed"]),
                     ion_velocity=s)
    n = case["n"]
    tan = {k: tuple(np.r_[v, np.zeros(n)] for v in dv) for k, dv in s.tangents.items()}
    return parts, tan


def oracle_record(case, parts, Ex, Ey, phi):
    p = case["p"]
    return trace_ref.trace_arrays(*parts, Ex, Ey, case["vac"], p.dx, p.cfl, p.max_steps, tapered_of(p),
                                  (phi, case["wface"]), record=True, probe_trajectories=1)


def test_tangent_primal_is_the_record_trace(trench):
    _case("tangent primal")
    parts, tan = draw(trench)
    t = HipTracer(trench["vac"])
    r = t.tangent(*parts, trench["Ex"], trench["Ey"], trench["p"], (trench["phi"], trench["wface"]), dstate=(None, None, *tan["sigma"]))
    ref = oracle_record(trench, parts, trench["Ex"], trench["Ey"], trench["phi"])
    np.testing.assert_array_equal(r["status"], ref["status"])
    np.testing.assert_array_equal(r["hit"], ref["hit"])
    np.testing.assert_array_equal(r["steps"], ref["steps"])
    np.testing.assert_array_equal(r["endpoints"], ref["endpoints"])
    np.testing.assert_array_equal(r["final_velocity"], ref["final_velocity"])
    assert r["total_steps"] == ref["total_steps"] and r["lost"] == ref["lost"]
    np.testing.assert_array_equal(r["escaped"], ref["escaped"])
    done = r["status"] != 3
    assert np.all((r["face"] >= 0) == done)
    assert np.all(r["face"][r["status"] == 2] == 0)   # escapes cross the top, a y-face


def _directions(case):
    """name -> (states and fields at theta - h, theta, theta + h; tangent inputs)."""
    base, tan = draw(case)
    sig = case["ion"]["sigma_deg"]
    e_ref = case["ion"].get("energy_eV") or 0.5 * (case["ion"]["e_lo"] + case["ion"]["e_hi"])
    f0 = (case["Ex"], case["Ey"], case["phi"])
    out = {
        "sigma": ([draw(case, 1 - H)[0], base, draw(case, 1 + H)[0]], [f0] * 3, 2 * H * sig,
                  dict(dstate=(None, None, *tan["sigma"]))),
        "energy": ([draw(case, shift=-H * e_ref)[0], base, draw(case, shift=H * e_ref)[0]], [f0] * 3, 2 * H * e_ref,
                   dict(dstate=(None, None, *tan["energy"]))),
    }
    fm = tuple(a - H * d for a, d in zip(f0, (case["dEx"], case["dEy"], case["dphi"])))
    fp = tuple(a + H * d for a, d in zip(f0, (case["dEx"], case["dEy"], case["dphi"])))
    out["charge"] = ([base] * 3, [fm, f0, fp], 2 * H, dict(dfield=(case["dEx"], case["dEy"])))
    return out


@pytest.mark.parametrize("direction", ["sigma", "energy", "charge"])
def test_tangent_matches_central_differences(trench, direction):
    _case(f"tangent vs differences: {direction}")
    states, flds, step, kw = _directions(trench)[direction]
    t = HipTracer(trench["vac"])
    p = trench["p"]
    r = t.tangent(*states[1], trench["Ex"], trench["Ey"], p, (trench["phi"], trench["wface"]), **kw)
    lo, hi = (oracle_record(trench, s, *f) for s, f in ((states[0], flds[0]), (states[2], flds[2])))
    same = np.ones(r["status"].size, bool)
    for ref in (lo, hi):
        same &= (ref["status"] == r["status"]) & (ref["hit"] == r["hit"]) & (ref["steps"] == r["steps"])
    same &= r["status"] != 3
    print(f"  particles keeping their events: {same.mean():.4f}", flush=True)
    assert same.mean() > 0.95
    fd_end = (hi["endpoints"] - lo["endpoints"]) / step
    fd_vel = (hi["final_velocity"] - lo["final_velocity"]) / step
    for name, fd, tg in (("end point", fd_end, r["d_endpoints"]), ("final velocity", fd_vel, r["d_final_velocity"])):
        fd, tg = fd[same], tg[same]
        scale = np.abs(tg).max(axis=1)
        moving = scale > 0
        assert moving.mean() > 0.3
        floor = 1e-6 * np.median(scale[moving])
        np.testing.assert_array_equal(fd[~moving], 0.0)   # electrons under an ion-only direction
        err = np.abs(fd - tg)[moving].max(axis=1) / (scale[moving] + floor)
        q = np.quantile(err, [0.5, 0.99, 1.0])
        print(f"  {name}: relative error median {q[0]:.2e}, 99% {q[1]:.2e}, max {q[2]:.2e}", flush=True)
        assert q[0] < 1e-5 and q[1] < 1e-3
    assert np.abs(r["d_endpoints"][same]).max() > 0


def test_field_free_floor_crossing_closed_form():
    """Straight flights (no field, absorbing walls): the floor crossing point
    x0 + tan(theta) (y0 - y_floor) and its derivatives."""
    _case("tangent closed form")
    p = replace(cases.small_params(wall_model="absorb"), mask_t=0)
    solid = R.build_geometry(p)[0]
    vac = ~solid
    ion = dict(kind="mono", sigma_deg=2.0, energy_eV=80.0)
    s = tangent.IonSampler(p.ion_mass_amu, **ion)
    n = 3000
    x, y, vx, vy, qm, spc = R.inject(p, n, solid.shape[1], np.random.default_rng(4), ion_velocity=s)
    parts = tuple(a[:n] for a in (x, y, vx, vy, qm, spc))   # ions only
    zero = np.zeros(solid.shape)
    t = HipTracer(vac)
    rs = t.tangent(*parts, zero, zero, p, None, dstate=(None, None, *s.tangents["sigma"]))
    re_ = t.tangent(*parts, zero, zero, p, None, dstate=(None, None, *s.tangents["energy"]))
    floor = (rs["status"] == 1) & (rs["face"] == 0) & (rs["hit"] % solid.shape[1] == p.floor_t - 1)
    assert floor.sum() > 0.15 * n
    theta = np.arctan2(parts[2], -parts[3])[floor]
    depth = parts[1][floor] - rs["endpoints"][floor, 1]
    expect = depth * (theta / ion["sigma_deg"]) / np.cos(theta) ** 2
    np.testing.assert_allclose(rs["d_endpoints"][floor, 0], expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max())
    assert np.abs(rs["d_endpoints"][floor, 1]).max() <= 1e-12 * np.abs(expect).max()   # on the face
    assert np.abs(re_["d_endpoints"][floor]).max() <= 1e-12 * np.abs(expect).max()


def test_face_deposit_derivative_matches_its_differences(trench):
    """d(face deposit)/d(theta) of the tangent push against the central
    difference of the face deposit over the particles that keep their events
    (device primal at theta +- h)."""
    _case("face deposit derivative")
    states, flds, step, kw = _directions(trench)["sigma"]
    t = HipTracer(trench["vac"])
    p = trench["p"]
    bd = tangent.BatchDerivative(t, p, trench["Ex"], trench["Ey"], (trench["phi"], trench["wface"]))
    mid = bd(states[1], **kw)
    side = [t.tangent(*s, *f[:2], p, (f[2], trench["wface"])) for s, f in ((states[0], flds[0]), (states[2], flds[2]))]
    same = mid["status"] != 3
    for r in side:
        same &= (r["status"] == mid["status"]) & (r["hit"] == mid["hit"]) & (r["steps"] == mid["steps"]) & \
                (r["face"] == mid["face"])
    spc = states[1][5]
    surf = bd.surface

    def dep(r, keep):
        st = np.where(keep, r["status"], 0).astype(np.int8)
        return tangent.face_deposit(trench["vac"], surf, spc, st, r["hit"], r["face"], r["endpoints"],
                                    r["d_endpoints"], p.dx)
    fd = (dep(side[1], same)[0] - dep(side[0], same)[0]) / step
    tg = dep(mid, same)[1]
    np.testing.assert_allclose(tg, fd, rtol=1e-4, atol=1e-6 * np.abs(tg).max())
    # the whole batch's face deposit conserves the absorbed count per species
    absorbed = [np.sum((mid["status"] == 1) & (spc == s)) for s in (0, 1)]
    np.testing.assert_allclose(mid["deposit"].sum(axis=1), absorbed, rtol=0, atol=1e-9)
    np.testing.assert_allclose(mid["d_deposit"].sum(axis=1), 0.0, atol=1e-9 * np.abs(mid["d_deposit"]).max())
