"""Steady-state criterion and charge monitors (host side of the batch loop):
known answers on synthetic series, as the reference's tests/test_steady_charge.py
checks its own criterion, applied to this package's monitors and the oracle."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import charging_ref as R
from tests import cases
from trench01_amd import model, monitors


def params(**kw):
    return replace(model.Params(), steady_window=30, **kw)


@pytest.mark.parametrize("sign", [-1, 1])
@pytest.mark.parametrize("age", [10_000, 1_000_000])
def test_constant_charging_is_rejected(sign, age):
    p = params()
    q = sign * (age + np.arange(2 * p.steady_window)) * 1e-12
    cur, prev = q[-p.steady_window:].mean(), q[:p.steady_window].mean()
    assert abs(cur - prev) < p.steady_charge_rtol * max(abs(cur), abs(prev))
    check = monitors.check_charge_steady({"right": q}, p)
    assert not check["ok"] and check["kind"] == "drift"
    assert check["d"] == pytest.approx(sign * p.steady_window * 1e-12, rel=1e-8)
    assert check["rate"] == pytest.approx(sign * 1e-12 / p.dt_batch, rel=1e-8)
    assert not R.check_charge_steady({"right": q}, p)["ok"]


@pytest.mark.parametrize("seed", [3, 17])
def test_stationary_noise_is_accepted(seed):
    p = params()
    noise = np.random.default_rng(seed).normal(0, 0.01e-12, 2 * p.steady_window)
    assert monitors.check_charge_steady({"left": 20e-12 + noise, "right": -50e-12 - noise}, p)["ok"]


def test_recent_step_fails_the_block_mean():
    p = params()
    q = np.r_[np.full(p.steady_window, 20e-12), np.full(p.steady_window, 21e-12)]
    check = monitors.check_charge_steady({"bottom": q}, p)
    assert not check["ok"] and check["kind"] == "mean" and check["d"] == pytest.approx(1e-12)


def test_negligible_drift_uses_absolute_tolerance():
    p = params()
    q = np.arange(2 * p.steady_window) * (p.steady_charge_atol / (2 * p.steady_window))
    assert monitors.check_charge_steady({"left": q}, p)["ok"]
    assert monitors.check_charge_steady({"left": np.zeros_like(q)}, replace(p, steady_charge_atol=0))["ok"]


@pytest.mark.parametrize("changes", [dict(steady_charge_rtol=0), dict(steady_charge_rtol=np.nan),
                                     dict(steady_charge_atol=-1), dict(steady_window=4), dict(dt_batch=0)])
def test_invalid_criterion_is_rejected(changes):
    with pytest.raises(ValueError):
        monitors.check_charge_steady({"left": np.zeros(60)}, replace(params(), **changes))


def test_voltage_criterion_known_answers():
    p = params()
    W = p.steady_window
    flat = {"a": np.r_[np.full(W, 10.0), np.full(W, 10.0)]}
    ok, name, d, tol = monitors.check_steady(flat, ["a"], p)
    assert ok and d == 0 and tol == pytest.approx(p.steady_rtol * 10.0)
    step = {"a": np.r_[np.full(W, 10.0), np.full(W, 11.0)]}
    ok, name, d, tol = monitors.check_steady(step, ["a"], p)
    assert not ok and name == "a" and d == pytest.approx(1.0)


def test_mean_and_se_matches_independent_formula():
    # white noise: SE of the mean ~ sigma/sqrt(n) with the detrended variance
    x = np.random.default_rng(0).normal(5.0, 2.0, 400)
    mean, se = monitors.mean_and_se(x)
    t = np.arange(x.size)
    resid = x - np.polyval(np.polyfit(t, x, 1), t)
    r1 = max(0.0, min(0.95, np.corrcoef(resid[:-1], resid[1:])[0, 1]))
    n_eff = x.size * (1 - r1) / (1 + r1)
    assert mean == pytest.approx(x.mean())
    assert se == pytest.approx(np.sqrt(resid.var(ddof=2) / n_eff))


@pytest.mark.parametrize("make", [cases.small_params, cases.taper_params, cases.ar8_params])
def test_probes_and_regions_match_oracle(make):
    rp = make()
    p = replace(model.Params(), **{k: getattr(rp, k) for k in model.Params.__dataclass_fields__ if hasattr(rp, k)})
    geo = model.Geometry(p)
    fields = model.Fields(p, geo)
    solid, mask, i0, i1 = R.build_geometry(rp)
    assert monitors.probes(geo, p, fields.cond) == R.probes_for(solid, rp, R.setup_fields(rp, solid, mask)["cond"])
    mine = monitors.charge_regions(geo, p)
    ref = R.charge_monitors(solid, mask, i0, i1, rp)
    assert mine.keys() == ref.keys()
    for k in ref:
        np.testing.assert_array_equal(mine[k], ref[k])
    rho = np.random.default_rng(1).normal(size=solid.shape)
    assert monitors.sample_charges(rho, mine, p.dx) == R.sample_charge_monitors(rho, ref, rp.dx)


def test_band_regions_partition_the_sidewalls():
    p = replace(model.Params(), **{k: getattr(cases.taper_params(), k) for k in model.Params.__dataclass_fields__
                                   if hasattr(cases.taper_params(), k)})
    regions = monitors.charge_regions(model.Geometry(p), p)
    for side in ("left", "right"):
        bands = np.sort(np.concatenate([regions[f"{side} band {b}"] for b in (1, 2, 3)]))
        np.testing.assert_array_equal(bands, regions[side])
