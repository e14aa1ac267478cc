"""The RF full-cycle push of the oracle (oracle/trace_ref.c, oracle_trace_rf)
pinned three ways:

* bit for bit against the reference's own RF traces (trace_cpu with rf=,
  tests/golden/rftrace_*.npz), with the oracle reproducing the per-position
  rounding of NumPy's matrix product (blas_tail=True);
* against the closed-form flight in a uniform sinusoidal field, with the
  second-order convergence of the trapezoidal push
  (tests/test_rf_cycle.py:59-88 of the reference);
* against the known answers of the time-dependent wall barrier
  (tests/test_rf_cycle.py:92-113) and of a constant wave, which must reduce to
  the frozen-field push (tests/test_rf_cycle.py:116-127).
"""
from dataclasses import replace

import numpy as np
import pytest

from oracle import trace_ref
from tests import rf_cases

RECORD_KEYS = ("status", "endpoints", "final_velocity", "flight_time", "final_phase", "launch_phase")


def oracle_rf(args, p, rf, barrier=None, blas_tail=False):
    record = {}
    out = trace_ref.oracle_tracer(*args, p, barrier, record, rf=rf, blas_tail=blas_tail)
    return out, record


@pytest.mark.parametrize("name", rf_cases.CASES)
def test_rf_trace_matches_reference_bit_for_bit(name):
    g, p, rf, barrier, args = rf_cases.golden(name)
    (ci, ce, esc, lost), rec = oracle_rf(args, p, rf, barrier, blas_tail=True)
    np.testing.assert_array_equal(ci, g["cnt_i"])
    np.testing.assert_array_equal(ce, g["cnt_e"])
    np.testing.assert_array_equal(esc, g["escaped"])
    assert lost == int(g["lost"])
    for key in RECORD_KEYS:
        np.testing.assert_array_equal(rec[key], g[key], err_msg=key)
    np.testing.assert_array_equal(rec["hit_cells"], g["hit_cells"])
    np.testing.assert_array_equal(rec["selected"], g["selected"])
    np.testing.assert_array_equal([len(q) for q in rec["paths"]], g["path_lengths"])
    np.testing.assert_array_equal(np.concatenate(rec["paths"]), g["paths"])


@pytest.mark.parametrize("name", rf_cases.CASES)
def test_rf_block_rounding_stays_within_reference_tolerance(name):
    """The per-particle order (block rounding for every particle, the order the
    device keeps) agrees with the reference to the 1e-12 its own CPU and CUDA
    RF paths are compared at (tests/test_rf_cycle.py:196-206)."""
    g, p, rf, barrier, args = rf_cases.golden(name)
    (ci, ce, esc, lost), rec = oracle_rf(args, p, rf, barrier, blas_tail=False)
    np.testing.assert_array_equal(ci, g["cnt_i"])
    np.testing.assert_array_equal(ce, g["cnt_e"])
    np.testing.assert_array_equal(rec["status"], g["status"])
    for key in ("endpoints", "final_velocity", "flight_time", "final_phase"):
        np.testing.assert_allclose(rec[key], g[key], rtol=1e-12, atol=1e-12 * np.abs(g[key]).max(), err_msg=key)


def test_rf_flight_matches_sinusoidal_closed_form_and_converges():
    errors = {}
    for steps in (96, 192):
        args, p, rf, t0 = rf_cases.analytic(steps)
        _, rec = oracle_rf(args, p, rf)
        assert np.all(rec["status"] == 3)
        np.testing.assert_allclose(rec["flight_time"], 2 / 3, atol=1e-13)
        np.testing.assert_allclose(rec["final_phase"], (rec["launch_phase"] + 240) % 360, atol=1e-10)
        errors[steps] = rf_cases.analytic_error(rec, t0)
    assert errors[192] < 2e-6
    assert errors[96] > 3.8 * errors[192]     # second order: halving dt cuts the error ~4x


def test_rf_wall_barrier_uses_instantaneous_voltage():
    args, p, barrier, rf = rf_cases.barrier_case()
    (ci, ce, esc, lost), rec = oracle_rf(args, p, rf, barrier)
    assert ci.sum() == 0 and lost == 1 and rec["final_velocity"][0, 1] > 0
    ci, ce, esc, lost = trace_ref.oracle_tracer(*args, p, barrier)
    assert ci.sum() == 1 and lost == 0


def test_constant_wave_reduces_to_frozen_field():
    g, p, rf, barrier, args = rf_cases.golden("taper_barrier")
    flat = {**rf, "wave": np.zeros_like(rf["wave"]), "initial_time": np.zeros_like(rf["initial_time"])}
    (ci, ce, esc, lost), dyn = oracle_rf(args, p, flat, barrier)
    frozen = {}
    fci, fce, fesc, flost = trace_ref.oracle_tracer(*args, p, barrier, frozen)
    np.testing.assert_array_equal(dyn["status"], frozen["status"])
    np.testing.assert_array_equal(ci, fci)
    np.testing.assert_allclose(dyn["final_velocity"], frozen["final_velocity"], atol=1e-7, rtol=1e-12)


def test_rf_coefficients_hit_the_wave_nodes():
    """At t = k period / n the interpolated voltage is the k-th wave row; half
    way it is the mean of the neighbours (trench_rf.py:5-11).  Traced through
    one step of a particle in a uniform basis field, where the step's
    acceleration is the interpolated coefficient times qm."""
    n_wave = 8
    wave = np.column_stack((np.arange(n_wave) * 1.0, -np.arange(n_wave) * 2.0))
    shape = (3, 50)
    vac = np.ones(shape, bool)
    z = np.zeros(shape)
    p = replace(rf_cases.model.Params(), dx=0.1, max_steps=1, probe_trajectories=1)
    dt = 1e-3
    # the step samples the wave at clock + dt/2: launch half a step before each
    # node and each midpoint between nodes (the last one wraps to row 0)
    at = np.concatenate([np.arange(n_wave), np.arange(n_wave) + 0.5]) / n_wave
    expect = np.concatenate([wave, 0.5 * (wave + np.roll(wave, -1, axis=0))])
    rf = dict(period=1.0, dt_max=dt, initial_time=at - 0.5 * dt, wave=wave, Ex=np.array([np.ones(shape), z]),
              Ey=np.array([z, np.ones(shape)]), phi=np.zeros((2, *shape)))
    n = at.size
    args = (np.full(n, 0.15), np.full(n, 2.5), np.zeros(n), np.zeros(n), np.ones(n), np.zeros(n, np.int8), z, z, vac)
    _, rec = oracle_rf(args, p, rf)
    np.testing.assert_allclose(rec["final_velocity"] / dt, expect, rtol=0, atol=1e-9)
    np.testing.assert_allclose(rec["flight_time"], dt, rtol=1e-15)
