"""The circuit field modes on the device: one batch from a charged state
against the oracle (push counts bit for bit, update to rounding), and whole
runs of trench01_amd.charging.run against the oracle's run()."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import charging_ref as R
from oracle import trace_ref
from tests.test_gpu_charging import product_params_from
from tests.test_oracle_circuit import RUN_CASES, circuit_of, load, params_of
from trench01_amd import charging, model

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def bohm(p):
    return lambda rng, n: np.full(n, p.electron_temp_eV / 2)


@pytest.mark.parametrize("name", RUN_CASES)
def test_device_circuit_batch_matches_oracle(name):
    """One device batch from a charged state in a circuit field mode (circuit_rf:
    the push follows the voltage wave from stratified launch phases): counts,
    escapes, losses and steps equal the oracle push on the device's field; the
    update (deposit, leak with the external-voltage current, driven mask,
    potential) agrees with the oracle's to 1e-12 / 1e-10 of the largest values."""
    _case(f"device circuit batch {name}")
    g = load(f"runrf_{name}")
    c = circuit_of(g)
    rp = replace(params_of(g), n_per_batch=3000)
    p = product_params_from(rp)
    setup = charging.ChargingSetup(p, c)
    dev = charging.DeviceCharging(setup)
    geo = setup.geo
    surf = geo.surface
    rng = np.random.default_rng(7)
    rho0 = rng.normal(0, 2e6, surf.size) * (1 - setup.cond_s)
    dev.set_rho(rho0)
    ctx = R.batch_context(rp, c)
    rho_full = np.zeros(geo.solid.shape)
    rho_full.flat[surf] = rho0
    _, phi_dev, _ = dev.read()
    phi_grid = phi_dev.reshape(geo.solid.shape)
    ex, ey = R.compute_field(phi_grid, geo.vac, rp, ctx["fields"]["volts"]["top"])
    ex_dev, ey_dev = dev.tracer.read_field()
    np.testing.assert_array_equal(ex_dev[geo.vac], ex[geo.vac])
    np.testing.assert_array_equal(ey_dev[geo.vac], ey[geo.vac])
    parts = model.inject(p, p.n_per_batch, geo.ny, np.random.default_rng(9), bohm(p))
    dev.reserve(2 * p.n_per_batch)
    dev.tracer.upload(0, *parts)
    rf = ctx["fields"]["rf"]
    barrier = (phi_grid, setup.fields.wface) if p.wall_model == "barrier" else None
    if rf is not None:
        times = np.concatenate([model.launch_times(np.random.default_rng(s), p.n_per_batch, rf["period"])
                                for s in (1, 2)])
        dev.tracer.upload_times(0, times)
        ref = trace_ref.trace_rf_arrays(*parts, ex, ey, geo.vac, p.dx, p.cfl, p.max_steps, model.is_tapered(p),
                                        {**rf, "initial_time": times}, barrier)
    else:
        ref = trace_ref.trace_arrays(*parts, ex, ey, geo.vac, p.dx, p.cfl, p.max_steps, model.is_tapered(p),
                                     barrier)
    d = dev.step(0, 2 * p.n_per_batch)
    rho1, phi1, counts = dev.read()
    m = surf.size
    np.testing.assert_array_equal(counts[:m], ref["cnt_i"].ravel()[surf])
    np.testing.assert_array_equal(counts[m:], ref["cnt_e"].ravel()[surf])
    assert d.tally.escaped[0] == ref["escaped"][0] and d.tally.escaped[1] == ref["escaped"][1]
    assert d.tally.lost == ref["lost"] and d.tally.steps == ref["total_steps"]
    assert counts.sum() > 0
    rho_ref, phi_ref1, _, _, _ = R.batch_update(ctx, rho_full, ref["cnt_i"], ref["cnt_e"])
    np.testing.assert_allclose(rho1, rho_ref.ravel()[surf], rtol=0, atol=1e-12 * np.abs(rho_ref).max())
    rows = setup.rows
    np.testing.assert_allclose(phi1[rows], phi_ref1.ravel()[rows], rtol=0, atol=1e-10 * np.abs(phi_ref1).max())


@pytest.mark.parametrize("name", RUN_CASES)
def test_circuit_run_on_device_matches_oracle_run(name):
    """Whole runs with host injection (the reference's PCG64 streams for the
    particles and the launch phases): batch 1 is identical to the oracle's;
    later batches follow the same charging history within its Monte Carlo
    noise; the shown fields are the oracle's snapshot of the device's mean
    potential (period mean + wave at rf_phase)."""
    _case(f"circuit run {name}")
    g = load(f"runrf_{name}")
    c = circuit_of(g)
    rp = params_of(g)
    p = product_params_from(rp)
    res = charging.run(p, log=None, circuit=c)
    ref = R.run(rp, circuit=c)
    for k in ("ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost"):
        assert res["hist"][k][0] == ref["hist"][k][0], k
    for k in ("bottom center", "sidewall middle", "mask top"):
        a, b = res["hist"][k], ref["hist"][k]
        print(k, a[-4:], b[-4:])
        assert abs(a[-4:].mean() - b[-4:].mean()) <= 0.05 * max(1.0, abs(b[-4:].mean())), k
    ctx = R.batch_context(rp, c)
    shown, ext, _, volts = R.field_snapshot(ctx["fields"], res["mean_phi"], rp)
    np.testing.assert_allclose(res["phi"], shown, rtol=0, atol=1e-12 * np.abs(shown).max())
    np.testing.assert_allclose(res["external_phi"], ext, rtol=0, atol=1e-12 * np.abs(ext).max())
    assert res["voltages"]["substrate"] == pytest.approx(volts["substrate"], rel=1e-14)
    np.testing.assert_allclose(res["mean_phi"][res["mask"]], ref["mean_phi"][ref["mask"]], rtol=1e-13)
    assert res["circuit"] is c
    if ctx["fields"]["rf"] is not None:
        assert res["checkpoint"]["phase_rng_state"] is not None


def test_circuit_rf_resume_equals_continuous_run(tmp_path):
    """A circuit_rf run stopped, saved and resumed equals the continuous run
    bit for bit: the stored checkpoint carries the launch-phase stream and the
    result file the circuit waveform (the reference's tests/test_resume.py:35-49
    check, in the RF-cycle mode); the display phase may change on resume."""
    _case("circuit_rf resume")
    from tests.test_gpu_results import same_result
    from trench01_amd import results
    g = load("runrf_circuit_rf")
    c = circuit_of(g)
    p = replace(product_params_from(params_of(g)), n_batches=8, n_per_batch=600)
    full = charging.run(p, log=None, circuit=c)
    part = charging.run(replace(p, n_batches=3), log=None, circuit=c)
    path = results.save_results(part, str(tmp_path / "part"))
    loaded, _ = results.load_results(path)
    resumed = charging.run(replace(p, n_batches=5), log=None, resume=loaded)
    assert resumed["start_batch"] == 3 and resumed["checkpoint"]["batches_completed"] == 8
    same_result(resumed, full)
    shifted = charging.run(replace(p, n_batches=5, rf_phase=p.rf_phase + 90), log=None, resume=loaded)
    np.testing.assert_array_equal(shifted["rho_raw"], full["rho_raw"])
    assert np.max(np.abs(shifted["phi"] - full["phi"])) > 1


@pytest.mark.parametrize("name", RUN_CASES)
def test_circuit_run_on_device_against_the_reference_run(name):
    """The device run against the reference's own run() in the circuit field
    modes (tests/golden/runrf_*.npz, same seed and waveform): every
    count-derived history identical, probe potentials and the final charge
    within 1e-12 of their range.  (circuit_rf: the RF bound's rounding that
    the reference leaves to its BLAS does not change any particle's cell in
    these runs.)"""
    _case(f"circuit run vs reference {name}")
    g = load(f"runrf_{name}")
    res = charging.run(product_params_from(params_of(g)), log=None, circuit=circuit_of(g))
    for k in ("ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost", "t"):
        np.testing.assert_array_equal(res["hist"][k], g[f"hist_{k}"], err_msg=k)
    for k in res["probes"]:
        b = g[f"hist_{k}"]
        np.testing.assert_allclose(res["hist"][k], b, rtol=0, atol=1e-12 * max(1.0, np.abs(b).max()), err_msg=k)
    np.testing.assert_allclose(res["rho_raw"], g["rho_raw"], rtol=0, atol=1e-12 * np.abs(g["rho_raw"]).max())
