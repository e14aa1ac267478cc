"""The RF full-cycle push on the GPU (th_trace_rf / th_trace_rf_record).

The device keeps one rounding for the acceleration bound's two-term product
for every particle (the block form of NumPy's matrix product), so it is
checked bit for bit against the oracle run the same way, and against the
reference's own RF traces (tests/golden/rftrace_*.npz) to the 1e-12 the
reference compares its own CPU and CUDA RF paths at
(tests/test_rf_cycle.py:196-206).  The closed-form sinusoidal flight and the
time-dependent wall barrier (tests/test_rf_cycle.py:59-113) run on the device
too, as do the non-record path, the multi-pass schedule and cancellation."""
import threading
from dataclasses import replace

import numpy as np
import pytest

from oracle import trace_ref
from tests import rf_cases
from trench01_amd import HipTracer, model

pytestmark = pytest.mark.gpu

RECORD_KEYS = ("status", "hit_cells", "endpoints", "final_velocity", "flight_time", "final_phase", "launch_phase")


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def hip_rf(args, p, rf, barrier=None, tracer=None):
    t = tracer or HipTracer(args[-1])
    record = {}
    out = t(*args, p, barrier, record, rf=rf)
    return out, record


def oracle_rf(args, p, rf, barrier=None):
    record = {}
    out = trace_ref.oracle_tracer(*args, p, barrier, record, rf=rf)
    return out, record


def same_record(a, b):
    for key in RECORD_KEYS:
        np.testing.assert_array_equal(a[key], b[key], err_msg=key)
    np.testing.assert_array_equal(a["selected"], b["selected"])
    assert len(a["paths"]) == len(b["paths"])
    for pa, pb in zip(a["paths"], b["paths"]):
        np.testing.assert_array_equal(pa, pb)


@pytest.mark.parametrize("name", rf_cases.CASES)
def test_rf_record_matches_oracle_bit_for_bit_and_reference(name):
    _case(f"rf record {name}")
    g, p, rf, barrier, args = rf_cases.golden(name)
    (ci, ce, esc, lost), rec = hip_rf(args, p, rf, barrier)
    (oci, oce, oesc, olost), orec = oracle_rf(args, p, rf, barrier)
    np.testing.assert_array_equal(ci, oci)
    np.testing.assert_array_equal(ce, oce)
    np.testing.assert_array_equal(esc, oesc)
    assert lost == olost
    same_record(rec, orec)
    # the reference's own trace, to its CPU-vs-CUDA tolerance
    np.testing.assert_array_equal(ci, g["cnt_i"])
    np.testing.assert_array_equal(ce, g["cnt_e"])
    np.testing.assert_array_equal(rec["status"], g["status"])
    np.testing.assert_array_equal(rec["launch_phase"], g["launch_phase"])
    for key in ("endpoints", "final_velocity", "flight_time", "final_phase"):
        np.testing.assert_allclose(rec[key], g[key], rtol=1e-12, atol=1e-12 * np.abs(g[key]).max(), err_msg=key)


def test_rf_flight_matches_sinusoidal_closed_form_on_device():
    _case("rf analytic")
    errors = {}
    for steps in (96, 192):
        args, p, rf, t0 = rf_cases.analytic(steps)
        _, rec = hip_rf(args, p, rf)
        _, orec = oracle_rf(args, p, rf)
        same_record(rec, orec)
        assert np.all(rec["status"] == 3)
        np.testing.assert_allclose(rec["flight_time"], 2 / 3, atol=1e-13)
        np.testing.assert_allclose(rec["final_phase"], (rec["launch_phase"] + 240) % 360, atol=1e-10)
        errors[steps] = rf_cases.analytic_error(rec, t0)
    assert errors[192] < 2e-6
    assert errors[96] > 3.8 * errors[192]


def test_rf_wall_barrier_uses_instantaneous_voltage_on_device():
    _case("rf barrier")
    args, p, barrier, rf = rf_cases.barrier_case()
    t = HipTracer(args[-1])
    (ci, ce, esc, lost), rec = hip_rf(args, p, rf, barrier, tracer=t)
    assert ci.sum() == 0 and lost == 1 and rec["final_velocity"][0, 1] > 0
    ci, ce, esc, lost = t(*args, p, barrier)
    assert ci.sum() == 1 and lost == 0


def big_batch(n_per_species=20000, seed=5):
    """A larger RF batch on the tapered golden geometry: host injection of the
    reference's draw order, random launch times over the period."""
    g, p, rf, barrier, _ = rf_cases.golden("taper_barrier")
    p = replace(p, n_per_batch=n_per_species, max_steps=1000)
    rng = np.random.default_rng(seed)
    x, y, vx, vy, qm, spc = model.inject(p, n_per_species, g["solid"].shape[1], rng)
    rf = {**rf, "initial_time": rng.uniform(0.0, rf["period"], 2 * n_per_species)}
    return (x, y, vx, vy, qm, spc, g["Ex"], g["Ey"], ~g["solid"]), p, rf, barrier


def test_rf_tallies_single_and_multi_pass_match_oracle():
    _case("rf multi-pass tallies")
    args, p, rf, barrier = big_batch()
    oci, oce, oesc, olost = trace_ref.oracle_tracer(*args, p, barrier, rf=rf)
    t = HipTracer(args[-1])
    for pass_steps in (0, 64, 250):
        t.pass_steps = pass_steps
        ci, ce, esc, lost = t(*args, p, barrier, rf=rf, keep_status=True)
        np.testing.assert_array_equal(ci, oci, err_msg=f"pass_steps={pass_steps}")
        np.testing.assert_array_equal(ce, oce, err_msg=f"pass_steps={pass_steps}")
        np.testing.assert_array_equal(esc, oesc)
        assert lost == olost
    # the record path (one pass) gives the same per-particle outcome
    rec = {}
    p1 = replace(p, probe_trajectories=4)
    t(*args, p1, barrier, rec, rf=rf)
    np.testing.assert_array_equal(rec["status"], t.last_status)


def test_rf_resident_pool_matches_host_call():
    """th_particles_upload_times + th_trace_resident with params.rf = 1."""
    _case("rf resident")
    args, p, rf, barrier = big_batch(4000, seed=9)
    x, y, vx, vy, qm, spc, ex, ey, vac = args
    t = HipTracer(vac)
    ci, ce, esc, lost = t(*args, p, barrier, rf=rf)
    n = x.size
    t.set_fields(ex, ey, barrier)
    t.set_rf(rf)
    t.reserve(n)
    t.upload(0, x, y, vx, vy, qm, spc)
    t.upload_times(0, rf["initial_time"])
    tally, counts, _ = t.trace_resident(0, n, t.params_for(p, True, rf=True), with_counts=True)
    np.testing.assert_array_equal(counts[0], ci)
    np.testing.assert_array_equal(counts[1], ce)
    assert tally["escaped"] == tuple(int(v) for v in esc) and tally["lost"] == lost


def test_rf_cancellation_keeps_launch_state():
    _case("rf cancellation")
    g, p, rf, barrier, args = rf_cases.golden("pillar_barrier")
    stop = threading.Event()
    stop.set()
    record = {}
    t = HipTracer(args[-1])
    ci, ce, esc, lost = t(*args, p, barrier, record, stop=stop, rf=rf)
    n = args[0].size
    assert lost == n and ci.sum() == 0 and record["cancelled"]
    np.testing.assert_array_equal(record["flight_time"], np.zeros(n))
    np.testing.assert_array_equal(record["final_phase"], record["launch_phase"])
    np.testing.assert_array_equal(record["endpoints"], np.column_stack((args[0], args[1])))


def test_rf_argument_errors():
    _case("rf errors")
    g, p, rf, barrier, args = rf_cases.golden("pillar_barrier")
    t = HipTracer(args[-1])
    with pytest.raises(ValueError):
        t(*args, p, barrier, rf={**rf, "initial_time": rf["initial_time"][:-1]})
    with pytest.raises(ValueError):
        t(*args, p, barrier, rf={**rf, "period": 0.0})
    with pytest.raises(ValueError):
        t(*args, p, barrier, rf={**rf, "wave": rf["wave"][:1]})
    bad = rf["Ex"].copy()
    bad[0, 0, 0] = np.nan
    with pytest.raises(ValueError):
        t(*args, p, barrier, rf={**rf, "Ex": bad})
    t.precision = 1
    with pytest.raises(ValueError):
        t(*args, p, barrier, rf=rf)
