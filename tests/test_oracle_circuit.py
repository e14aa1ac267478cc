"""The circuit field modes on the CPU: the oracle's set-up and whole runs
against runs of the reference itself (tests/golden/runrf_*.npz, rftrace_*.npz),
the linear RF basis against independent Poisson solves at every phase, and
the host-side set-up of the HIP path (trench01_amd.model) against the oracle."""
import json
from dataclasses import replace

import numpy as np
import pytest

from oracle import charging_ref as R
from trench01_amd import model
from tests.test_golden import GOLDEN, load, params_of

RUN_CASES = ("circuit_rf", "circuit")
RF_TRACE_CASES = sorted(p.stem[8:] for p in GOLDEN.glob("rftrace_*.npz"))


def circuit_of(g):
    return {k: g[f"circuit_{k}"] for k in ("t", "uw", "um", "up", "V1")}


def sine_circuit(period=4e-10, n=256, amplitude=35.0):
    """One period of a sinusoidal bias (the layout of the reference's circuit
    solver output: t, uw, um, up, V1)."""
    t = np.arange(n) * period / n
    w = 2 * np.pi * t / period
    up = 10 + 2 * np.cos(w)
    return dict(t=t, up=up, uw=-80 - amplitude * np.sin(w) + up, um=-40 - amplitude / 2 * np.sin(w) + up,
                V1=40 + amplitude / 2 * np.sin(w))


def small(**changes):
    base = replace(R.Params(), nx=20, trench_w=6, trench_d=12, floor_t=4, mask_t=4, n_vac=8,
                   field_mode="circuit_rf", side_bc="pillar", bias="rf", sigma_s=0.0, n_batches=3,
                   n_per_batch=128, max_steps=1000, probe_trajectories=8)
    return replace(base, **changes)


@pytest.mark.parametrize("name", RUN_CASES)
def test_whole_circuit_run_matches_reference(name):
    """Oracle run() vs the reference's run() in the circuit field modes (same
    seed, same circuit waveform, single-threaded BLAS on both sides)."""
    g = load(f"runrf_{name}")
    p = params_of(g)
    res = R.run(p, circuit=circuit_of(g), blas_tail=True)
    for k in ("ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost", "t"):
        np.testing.assert_array_equal(res["hist"][k], g[f"hist_{k}"], err_msg=k)
    for k in res["probes"]:
        np.testing.assert_allclose(res["hist"][k], g[f"hist_{k}"], rtol=1e-12, atol=1e-12, err_msg=k)
    scale = np.abs(g["rho_raw"]).max()
    np.testing.assert_allclose(res["rho_raw"], g["rho_raw"], rtol=1e-12, atol=1e-12 * scale)
    for key in ("phi", "mean_phi", "external_phi"):
        np.testing.assert_allclose(res[key], g[key], rtol=0, atol=1e-11 * np.abs(g[key]).max(), err_msg=key)
    for key in ("Ex", "Ey"):
        np.testing.assert_allclose(res[key], g[key], rtol=0, atol=1e-10 * np.abs(g[key]).max(), err_msg=key)
    cond = g["rho"] != g["rho_raw"]
    np.testing.assert_allclose(res["rho"][cond], g["rho"][cond], rtol=1e-8, atol=1e-8 * np.abs(g["rho"]).max())
    volts = json.loads(str(g["voltages"]))
    for k, v in volts.items():
        assert res["voltages"][k] == pytest.approx(v, rel=1e-14, abs=1e-12), k
    assert np.isfinite(res["hist"]["bottom center"]).all()


def test_display_phase_changes_only_what_is_shown():
    """circuit_rf: rf_phase selects the phase of the shown potential; the
    charge and the histories (period means) do not depend on it
    (reference test tests/test_rf_cycle.py:127-146)."""
    g = load("runrf_circuit_rf")
    p = replace(params_of(g), n_batches=3)
    c = circuit_of(g)
    a = R.run(p, circuit=c)
    b = R.run(replace(p, rf_phase=p.rf_phase + 90), circuit=c)
    np.testing.assert_array_equal(a["rho_raw"], b["rho_raw"])
    for key in a["hist"]:
        np.testing.assert_array_equal(a["hist"][key], b["hist"][key])
    np.testing.assert_array_equal(a["mean_phi"], b["mean_phi"])
    assert np.max(np.abs(a["phi"] - b["phi"])) > 10


@pytest.mark.parametrize("side_bc", ["periodic", "pillar", "fixed"])
def test_linear_basis_matches_independent_poisson_at_every_phase(side_bc):
    """The period-mean external potential plus the wave's deviation times the
    unit responses equals an independent solve with the instantaneous
    voltages (field_mode circuit at that phase): the superposition that the
    RF push relies on."""
    p = small(side_bc=side_bc, left_voltage=-20.0, right_voltage=5.0)
    c = sine_circuit()
    solid, mask, _, _ = R.build_geometry(p)
    f = R.setup_fields(p, solid, mask, c)
    for phase in (0, 45, 90, 180, 270, 359.9, 405):
        shown, ext, rhs, volts = R.field_snapshot(f, f["external"], replace(p, rf_phase=phase))
        direct = R.setup_fields(replace(p, field_mode="circuit", rf_phase=phase), solid, mask, c)
        np.testing.assert_allclose(shown, direct["external"], atol=2e-12, rtol=2e-13)
        np.testing.assert_allclose(ext, direct["external"], atol=2e-12, rtol=2e-13)
        np.testing.assert_allclose(rhs, direct["boundary_rhs"], atol=1e3, rtol=2e-13)
        assert volts["substrate"] == pytest.approx(direct["volts"]["substrate"])
        assert volts["mask"] == pytest.approx(direct["volts"]["mask"])


def test_pillar_sides_ramp_from_substrate_to_mask():
    """side_bc pillar: the side faces of the dielectric carry a linear ramp
    from the substrate to the mask voltage (midpoint of each cell row)."""
    p = small(field_mode="circuit", side_bc="pillar")
    solid, mask, _, _ = R.build_geometry(p)
    eps = R.permittivity(solid, mask, p)
    volts = dict(substrate=-50.0, mask=-10.0, top=0.0)
    rhs = R.boundary_rhs(p, eps, volts)
    top_ox = p.floor_t + p.trench_d
    side = -rhs[0, :top_ox] * p.dx ** 2 / (2 * eps[0, :top_ox])
    bottom = np.zeros(top_ox)
    bottom[0] = 2 * eps[0, 0] / p.dx ** 2 * volts["substrate"] * p.dx ** 2 / (2 * eps[0, 0])
    y = (np.arange(top_ox) + 0.5) / top_ox
    np.testing.assert_allclose(side - bottom, -50.0 + 40.0 * y, rtol=1e-13)
    assert np.all(rhs[1:-1, 1:top_ox] == 0)


def test_circuit_voltages_interpolate_periodically():
    c = sine_circuit()
    period = c["t"][1] * c["t"].size
    for phase in (0.0, 10.0, 359.0, 720.0 + 33.0):
        v = R.circuit_voltages(c, phase)
        at = (phase % 360) / 360 * period
        w = 2 * np.pi * at / period
        # linear interpolation of smooth curves over 256 nodes: within 1e-3 V
        assert v["substrate"] == pytest.approx(-80 - 35 * np.sin(w), abs=1e-2)
        assert v["mask"] == pytest.approx(-40 - 17.5 * np.sin(w), abs=1e-2)
        assert v["top"] == 0.0
    # at a node the value is the node's exactly
    v = R.circuit_voltages(c, 360.0 * 17 / 256)
    assert v["substrate"] == c["uw"][17] - c["up"][17]


def test_launch_times_cover_the_cycle_in_shuffled_order():
    times = R.launch_times(np.random.default_rng(123), 128, 1e-8)
    np.testing.assert_array_equal(np.sort((times / 1e-8 * 128).astype(int)), np.arange(128))
    assert abs(np.corrcoef(times, np.arange(128))[0, 1]) < 0.2
    again = model.launch_times(np.random.default_rng(123), 128, 1e-8)
    np.testing.assert_array_equal(times, again)


@pytest.mark.parametrize("change", [dict(bias="dc"), dict(top_bc="neumann"), dict(rf_time_steps=0),
                                    dict(dt_batch=1e-12)])
def test_invalid_cycle_settings_are_rejected(change):
    p = small(**change)
    solid, mask, _, _ = R.build_geometry(p)
    with pytest.raises(ValueError):
        R.setup_fields(p, solid, mask, sine_circuit())
    mp = model.Params(**{k: getattr(p, k) for k in vars(p) if hasattr(model.Params, k)})
    with pytest.raises(ValueError):
        model.Fields(mp, model.Geometry(mp), sine_circuit())


@pytest.mark.parametrize("name", RF_TRACE_CASES)
def test_rf_basis_matches_reference_setup(name):
    """Oracle set-up of circuit_rf (unit responses, their fields, wave, dt cap)
    vs the arrays the reference's setup_fields produced for the RF traces."""
    g = load(f"rftrace_{name}")
    p = params_of(g)
    solid, mask = g["solid"], g["mask"]
    from tests.golden.make_golden import rf_circuit
    f = R.setup_fields(p, solid, mask, rf_circuit())
    rf = f["rf"]
    assert rf["period"] == g["rf_period"] and rf["dt_max"] == g["rf_dt_max"]
    np.testing.assert_array_equal(rf["wave"], g["rf_wave"])
    for key in ("phi", "Ex", "Ey"):
        np.testing.assert_allclose(rf[key], g[f"rf_{key}"], rtol=0, atol=1e-11 * np.abs(g[f"rf_{key}"]).max(),
                                   err_msg=key)


@pytest.mark.parametrize("name", RUN_CASES)
def test_host_setup_matches_oracle(name):
    """trench01_amd.model.Fields (the HIP path's host set-up) vs the oracle
    in the circuit modes: external potential, voltages, RF basis."""
    g = load(f"runrf_{name}")
    rp = params_of(g)
    stored = json.loads(str(g["params"]))
    p = model.Params(**{k: v for k, v in stored.items() if hasattr(model.Params, k)})
    c = circuit_of(g)
    geo = model.Geometry(p)
    mine = model.Fields(p, geo, c)
    solid, mask, _, _ = R.build_geometry(rp)
    ref = R.setup_fields(rp, solid, mask, c)
    np.testing.assert_allclose(mine.external, ref["external"], rtol=0, atol=1e-12 * np.abs(ref["external"]).max())
    np.testing.assert_allclose(mine.boundary_rhs, ref["boundary_rhs"], rtol=1e-14)
    for k in ("substrate", "mask", "top"):
        assert mine.volts[k] == pytest.approx(ref["volts"][k], rel=1e-14, abs=1e-13)
    if ref["rf"] is None:
        assert mine.rf is None
        return
    for key in ("phi", "Ex", "Ey", "boundary_rhs"):
        np.testing.assert_allclose(mine.rf[key], ref["rf"][key], rtol=0,
                                   atol=1e-12 * max(np.abs(ref["rf"][key]).max(), 1.0), err_msg=key)
    np.testing.assert_array_equal(mine.rf["wave"], ref["rf"]["wave"])
    assert mine.rf["period"] == ref["rf"]["period"] and mine.rf["dt_max"] == ref["rf"]["dt_max"]
    shown, ext, rhs, volts = mine.snapshot(ref["external"], p)
    o_shown, o_ext, o_rhs, o_volts = R.field_snapshot(ref, ref["external"], rp)
    np.testing.assert_allclose(shown, o_shown, rtol=0, atol=1e-12 * np.abs(o_shown).max())
    assert volts["substrate"] == pytest.approx(o_volts["substrate"], rel=1e-14)
