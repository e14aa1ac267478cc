"""The oracle against golden vectors written by the reference itself
(tests/golden/make_golden.py imports the reference in the build container).

These pin what the reference's own fixtures leave open: the barrier wall
model, conductive masks, tapered walls, record mode, the conductor-collapsed
LU, the leak operator, the floating-conductor step, and a whole AR 5 run."""
import json
from dataclasses import fields as dc_fields, replace
from pathlib import Path

import numpy as np
import pytest

from oracle import charging_ref as R
from oracle import trace_ref

GOLDEN = Path(__file__).resolve().parent / "golden"
TRACE_CASES = sorted(p.stem[6:] for p in GOLDEN.glob("trace_*.npz"))
SETUP_CASES = sorted(p.stem[6:] for p in GOLDEN.glob("setup_*.npz"))


def load(name):
    with np.load(GOLDEN / f"{name}.npz") as d:
        return {k: d[k] for k in d.files}


def params_of(g):
    """Oracle Params from the reference Params stored with a vector."""
    stored = json.loads(str(g["params"]))
    names = {f.name for f in dc_fields(R.Params)}
    return replace(R.Params(), **{k: v for k, v in stored.items() if k in names})


@pytest.fixture(scope="module", params=TRACE_CASES)
def trace_case(request):
    return request.param, load(f"trace_{request.param}")


def test_geometry_and_fields(trace_case):
    name, g = trace_case
    p = params_of(g)
    solid, mask, _, _ = R.build_geometry(p)
    np.testing.assert_array_equal(solid, g["solid"])
    np.testing.assert_array_equal(mask, g["mask"])
    fields = R.setup_fields(p, solid, mask)
    phi = fields["solve"]((-g["rho"] / R.EPS0).ravel()).reshape(solid.shape) + fields["external"]
    np.testing.assert_allclose(phi, g["phi"], rtol=0, atol=1e-12 * np.abs(g["phi"]).max())
    ex, ey = R.compute_field(g["phi"], ~solid, p)
    np.testing.assert_array_equal(ex, g["Ex"])
    np.testing.assert_array_equal(ey, g["Ey"])
    np.testing.assert_array_equal(np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"])), g["wface"])


def trace_golden(g, p, threads=0):
    barrier = (g["phi"], g["wface"]) if p.wall_model == "barrier" else None
    return trace_ref.trace_arrays(g["x"], g["y"], g["vx"], g["vy"], g["qm"], g["spc"], g["Ex"], g["Ey"],
                                  ~g["solid"], p.dx, p.cfl, p.max_steps, trace_ref.is_tapered(p), barrier,
                                  threads=threads, record=True, probe_trajectories=p.probe_trajectories)


def test_record_mode_trace_is_bit_identical(trace_case):
    name, g = trace_case
    p = params_of(g)
    r = trace_golden(g, p)
    np.testing.assert_array_equal(r["cnt_i"], g["cnt_i"])
    np.testing.assert_array_equal(r["cnt_e"], g["cnt_e"])
    np.testing.assert_array_equal(r["escaped"], g["escaped"])
    assert r["lost"] == int(g["lost"])
    np.testing.assert_array_equal(r["status"], g["status"])
    ny = g["solid"].shape[1]
    cells = np.where(r["status"][:, None] == 1, np.column_stack((r["hit"] // ny, r["hit"] % ny)), -1)
    np.testing.assert_array_equal(cells, g["hit_cells"])
    np.testing.assert_array_equal(r["endpoints"], g["endpoints"])
    np.testing.assert_array_equal(r["final_velocity"], g["final_velocity"])
    np.testing.assert_array_equal(r["selected"], g["selected"])
    np.testing.assert_array_equal([len(q) for q in r["paths"]], g["path_lengths"])
    np.testing.assert_array_equal(np.concatenate(r["paths"]), g["paths"])


@pytest.fixture(scope="module", params=SETUP_CASES)
def setup_case(request):
    g = load(f"setup_{request.param}")
    p = params_of(g)
    solid, mask, _, _ = R.build_geometry(p)
    return p, g, solid, mask, R.setup_fields(p, solid, mask)


def test_conductor_collapsed_green_columns(setup_case):
    p, g, solid, mask, fields = setup_case
    flat = g["flat"]
    B = np.zeros((solid.size, flat.size))
    B[flat, np.arange(flat.size)] = -1.0 / (p.dx ** 2 * R.EPS0)
    green = fields["solve"](B)
    np.testing.assert_allclose(green, g["green"], rtol=1e-12, atol=1e-12 * np.abs(g["green"]).max())
    np.testing.assert_allclose(fields["external"], g["external"], rtol=0, atol=1e-12)


def test_fixed_voltage_external_field(setup_case):
    p, g, solid, mask, _ = setup_case
    pf = replace(p, field_mode="fixed", substrate_voltage=-20.0, mask_voltage=-10.0, top_voltage=5.0)
    ext = R.setup_fields(pf, solid, mask)["external"]
    np.testing.assert_allclose(ext, g["external_fixed"], rtol=0, atol=1e-11)


def test_leak_operator(setup_case):
    p, g, solid, mask, fields = setup_case
    leak = R.build_leakage(solid, fields["solve"], p)
    np.testing.assert_array_equal(leak["flat"], g["flat"])
    np.testing.assert_array_equal(leak["G"], g["G"])
    np.testing.assert_array_equal(leak["region"], g["region"])
    np.testing.assert_allclose(leak["M"], g["M"], rtol=1e-12, atol=1e-12)


def test_conductor_response_and_step(setup_case):
    p, g, solid, mask, fields = setup_case
    g_cond = R.conductor_response(fields["solve"], fields["cond"], p.dx)
    np.testing.assert_allclose(g_cond, g["g_cond"], rtol=1e-12, atol=0)
    C = 1.0 / g_cond[fields["cond"]][0]
    assert C == pytest.approx(float(g["C_cond"]), rel=1e-12)
    for n_i, n_e, drift, dv in g["conductor_steps"]:
        got = R.conductor_step(int(n_i), int(n_e), 1e-15, float(g["C_cond"]), p.electron_temp_eV, drift)
        assert got == pytest.approx(dv, rel=1e-14, abs=1e-300)


def test_ar5_whole_run_matches_reference():
    """Time-marching oracle vs the reference's own run() on the plain AR 5
    workload: 50 batches x 1e4 particles per species."""
    path = GOLDEN / "run_ar5.npz"
    if not path.exists():
        pytest.skip("run_ar5.npz not generated")
    g = load("run_ar5")
    p = params_of(g)
    from trench01_amd import configs
    sampler = configs.ion_velocity_sampler(p.ion_mass_amu, **configs.AR5_ION)
    res = R.run(p, ion_velocity=sampler, tracer=lambda *a: _threaded_tracer(*a))
    for k in ("ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost", "t"):
        np.testing.assert_array_equal(res["hist"][k], g[f"hist_{k}"], err_msg=k)
    for k in res["probes"]:
        np.testing.assert_allclose(res["hist"][k], g[f"hist_{k}"], rtol=1e-11, atol=1e-11, err_msg=k)
    np.testing.assert_allclose(res["rho_raw"], g["rho_raw"], rtol=1e-12, atol=1e-12 * np.abs(g["rho_raw"]).max())
    np.testing.assert_allclose(res["phi"], g["phi"], rtol=0, atol=1e-11 * np.abs(g["phi"]).max())


def _threaded_tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None, stop=None, rf=None):
    r = trace_ref.trace_arrays(x, y, vx, vy, qm, spc, Ex, Ey, vac, p.dx, p.cfl, p.max_steps,
                               trace_ref.is_tapered(p), barrier, threads=0)
    return r["cnt_i"], r["cnt_e"], r["escaped"], r["lost"]
