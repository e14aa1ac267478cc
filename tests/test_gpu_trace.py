"""HIP push parity: libtrench_hip.so against the oracle's C push on identical
host-generated particle states.  Integer results (per-cell counts, escaped,
lost, per-particle status and hit cell, total steps) must be bit-identical."""
import threading

import numpy as np
import pytest

from oracle import charging_ref as R
from oracle import trace_ref
from tests import cases
from trench01_amd import HipTracer, configs
from trench01_amd.tracer import tapered_of

pytestmark = pytest.mark.gpu


def _print_case(name):
    print(f"[gpu case] {name}", flush=True)


def test_device_arithmetic_matches_host_c_library():
    _print_case("device arithmetic")
    rng = np.random.default_rng(7)
    n = 1 << 20
    parts = []
    for ea, eb in ((0, 0), (20, 17), (60, 65), (-30, 5), (300, 290), (-470, -470), (-1060, -1060)):
        parts.append((rng.standard_normal(n // 8) * 2.0 ** ea, rng.standard_normal(n // 8) * 2.0 ** eb))
    e = rng.integers(-1070, 1020, n // 8)
    parts.append((np.ldexp(rng.random(n // 8) + .5, e), np.ldexp(rng.random(n // 8) + .5, e + rng.integers(-40, 40, n // 8))))
    # hypot operands at the edges of its range-free core: widely separated pairs around the 2^-54 ratio and
    # the smaller operand around 2^-383, over the exponent range
    x = np.ldexp(rng.random(n // 8) + .5, rng.integers(-383, 511, n // 8))
    parts.append((x, x * 2.0 ** -54 * (1 + rng.uniform(-1e-3, 1e-3, x.size))))
    y = np.ldexp(rng.random(n // 8) + .5, rng.integers(-386, -380, n // 8))
    parts.append((y * 2.0 ** rng.uniform(0, 60, y.size), y))
    # quotients near the ranges where the division's scaling steps act
    parts.append((np.ldexp(rng.random(n // 16) + .5, rng.integers(-975, -960, n // 16)),
                  np.ldexp(rng.random(n // 16) + .5, rng.integers(-10, 10, n // 16))))
    parts.append((np.ldexp(rng.random(n // 16) + .5, rng.integers(-60, 60, n // 16)),
                  np.ldexp(rng.random(n // 16) + .5, rng.integers(1015, 1024, n // 16))))
    parts.append((np.ldexp(rng.random(n // 16) + .5, rng.integers(700, 1023, n // 16)),
                  np.ldexp(rng.random(n // 16) + .5, rng.integers(-80, 10, n // 16))))
    a = np.concatenate([p[0] for p in parts])
    b = np.concatenate([p[1] for p in parts])
    t = HipTracer(np.ones((4, 4), bool))
    out = t.selftest_math(a, b)
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(out[:, 0], trace_ref.glibc_hypot(a, b))
        np.testing.assert_array_equal(out[:, 7], out[:, 0])   # the paired hypot, both pairs
        np.testing.assert_array_equal(out[:, 8], trace_ref.glibc_hypot(b * 2.0 ** -500, a))
        np.testing.assert_array_equal(out[:, 1], np.sqrt(a))
        np.testing.assert_array_equal(out[:, 2], a / b)
        np.testing.assert_array_equal(out[:, 3], np.fmod(a, b))
        # the range-free forms equal the library operations wherever the IEEE sequences do not rescale
        ok = (np.abs(a) >= 2.0 ** -767) & np.isfinite(a)
        assert ok.sum() > n // 2
        np.testing.assert_array_equal(out[ok, 4], np.sqrt(np.abs(a[ok])))
        q = a / b
        ea, eb = np.frexp(a)[1], np.frexp(b)[1]
        ok = (np.abs(b) >= 2.0 ** -1022) & (np.abs(b) < 2.0 ** 1021) & np.isfinite(a) & \
             ((a == 0) | (np.abs(a) >= 2.0 ** -968)) & ((a == 0) | (np.abs(q) >= 2.0 ** -1021)) & \
             np.isfinite(q) & (ea - eb < 767)
        assert ok.sum() > n // 2
        np.testing.assert_array_equal(out[ok, 5], q[ok])
        np.testing.assert_array_equal(out[:, 6], 1.75e-9 / (np.abs(a) + np.sqrt(1e-8 * np.abs(b)) + 1.0))


CASES = {
    "regression-absorb": lambda: cases.charged_case(cases.regression_params(), n=3000, seed=3),
    "small-absorb": lambda: cases.charged_case(cases.small_params(wall_model="absorb")),
    "small-barrier": lambda: cases.charged_case(cases.small_params(wall_model="barrier")),
    "taper-absorb": lambda: cases.charged_case(cases.taper_params(wall_model="absorb"), n=2000),
    "taper-barrier": lambda: cases.charged_case(cases.taper_params(wall_model="barrier"), n=2000),
    "ar5-barrier": lambda: cases.charged_case(cases.ar5_params(), n=5000, ion=configs.AR5_ION),
    "ar8-barrier": lambda: cases.charged_case(cases.ar8_params(), n=20000, ion=configs.AR8_ION,
                                              q_left=4e7, q_right=-2e7),
    "ar5-marched": lambda: cases.marched_case(cases.ar5_params(), 30, 2000, 5000, ion=configs.AR5_ION),
    "ar8-marched": lambda: cases.marched_case(cases.ar8_params(), 30, 5000, 20000, ion=configs.AR8_ION),
    "ar8-marched-1e6": lambda: cases.marched_case(cases.ar8_params(), 30, 5000, 500_000, ion=configs.AR8_ION),
    "ar8-marched-short": lambda: cases.marched_case(cases.ar8_params(max_steps=450), 30, 5000, 20000,
                                                    ion=configs.AR8_ION),
    "regression-marched-absorb": lambda: cases.marched_case(cases.regression_params(), 40, 3000, 3000),
}


@pytest.mark.parametrize("name", list(CASES))
def test_trace_is_bit_identical_to_oracle(name):
    _print_case(f"trace {name}")
    c = CASES[name]()
    p = c["p"]
    x, y, vx, vy, qm, spc = c["particles"]
    ref = trace_ref.trace_arrays(x, y, vx, vy, qm, spc, c["Ex"], c["Ey"], c["vac"], p.dx, p.cfl,
                                 p.max_steps, tapered_of(p), c["barrier"])
    t = HipTracer(c["vac"])
    ci, ce, esc, lost = t(x.copy(), y.copy(), vx.copy(), vy.copy(), qm.copy(), spc.copy(),
                          np.asfortranarray(c["Ex"]), np.asfortranarray(c["Ey"]), c["vac"], p, c["barrier"],
                          keep_status=True)
    np.testing.assert_array_equal(t.last_status, ref["status"])
    np.testing.assert_array_equal(t.last_hit, ref["hit"])
    np.testing.assert_array_equal(ci, ref["cnt_i"])
    np.testing.assert_array_equal(ce, ref["cnt_e"])
    np.testing.assert_array_equal(esc, ref["escaped"])
    assert lost == ref["lost"]
    assert t.last_tally["steps"] == ref["total_steps"]
    assert t.last_tally["off_surface"] == 0
    assert ci.sum() + ce.sum() + esc.sum() + lost == x.size
    np.testing.assert_array_equal(t.surface_cells(), np.flatnonzero(R.surface_cells(c["solid"])))


def test_whole_run_reproduces_reference_regression_fixture():
    """The oracle's batch loop with the HIP push in place of the C push
    reproduces the reference's base_noleak fixture: charge bit for bit."""
    _print_case("whole run base_noleak")
    p = cases.regression_params()
    solid, _, _, _ = R.build_geometry(p)
    tracer = HipTracer(~solid)
    res = R.run(p, tracer=tracer)
    with np.load(cases.GOLDEN / "base_noleak.npz") as d:
        np.testing.assert_array_equal(res["rho"], d["rho"])
        for k in ("ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost"):
            np.testing.assert_array_equal(res["hist"][k], d[f"hist_{k}"])


def test_empty_batch_returns_zeros():
    _print_case("empty batch")
    p = cases.small_params()
    solid, _, _, _ = R.build_geometry(p)
    z = np.zeros(solid.shape)
    t = HipTracer(~solid)
    e = np.zeros(0)
    ci, ce, esc, lost = t(e, e, e, e, e, np.zeros(0, np.int8), z, z, ~solid, p)
    assert ci.sum() == ce.sum() == esc.sum() == lost == 0


def test_argument_errors_match_reference_contract():
    _print_case("argument errors")
    p = cases.small_params()
    solid, _, _, _ = R.build_geometry(p)
    z = np.zeros(solid.shape)
    t = HipTracer(~solid)
    x, y, vx, vy, qm, spc = R.inject(p, 4, solid.shape[1], np.random.default_rng(0))
    with pytest.raises(ValueError):
        t(x, y, vx, vy, qm, np.full(8, 2, np.int8), z, z, ~solid, p)
    with pytest.raises(ValueError):
        t(x, y, vx, vy, qm, spc, z, z, solid, p)
    from dataclasses import replace
    with pytest.raises(ValueError):
        t(x, y, vx, vy, qm, spc, z, z, ~solid, replace(p, cfl=1.5))
    with pytest.raises(ValueError):
        t(x[:3], y, vx, vy, qm, spc, z, z, ~solid, p)
    with pytest.raises(RuntimeError):
        HipTracer(~solid, device=999)


def test_stop_before_start_reports_every_particle_lost():
    _print_case("cancellation")
    p = cases.small_params()
    solid, _, _, _ = R.build_geometry(p)
    z = np.zeros(solid.shape)
    t = HipTracer(~solid)
    parts = R.inject(p, 20, solid.shape[1], np.random.default_rng(1))
    stop = threading.Event()
    stop.set()
    ci, ce, esc, lost = t(*parts, z, z, ~solid, p, stop=stop)
    assert ci.sum() == ce.sum() == esc.sum() == 0 and lost == 40
    assert t.last_tally["cancelled"]


@pytest.mark.parametrize("precision", [0, 1])
def test_multi_pass_schedule_does_not_change_results(precision):
    """Stopping particles at a cap and continuing them in later passes (with
    their full state) gives the single-pass results bit for bit: per-particle
    status and hit cell, tallies and the total step count."""
    _print_case(f"multi-pass precision {precision}")
    c = CASES["ar8-marched"]()
    p = c["p"]
    parts = c["particles"]
    runs = []
    for pass_steps in (0, 37, 100, 333, 1000):
        t = HipTracer(c["vac"])
        t.precision = precision
        t.pass_steps = pass_steps
        r = t(*parts, c["Ex"], c["Ey"], c["vac"], p, c["barrier"], keep_status=True)
        runs.append((r, t.last_status.copy(), t.last_hit.copy(), t.last_tally["steps"]))
    (r0, s0, h0, n0) = runs[0]
    for r, st, h, nsteps in runs[1:]:
        for a, b in zip(r, r0):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(st, s0)
        np.testing.assert_array_equal(h, h0)
        assert nsteps == n0


@pytest.mark.parametrize("name", ["ar8-marched", "ar5-marched", "taper-barrier"])
def test_tile_push_does_not_change_results(name):
    """The late passes with the trench's field staged in LDS (tile_pass) give
    the plain multi-pass push's results bit for bit: per-particle status and
    hit cell, tallies and total steps, whichever pass the tile starts at."""
    _print_case(f"tile push {name}")
    c = CASES[name]()
    p = c["p"]
    parts = c["particles"]
    runs = []
    for pass_steps, tile_pass in ((100, 0), (100, 1), (100, 3), (37, 2), (250, 4)):
        t = HipTracer(c["vac"])
        t.pass_steps = pass_steps
        t.tile_pass = tile_pass
        r = t(*parts, c["Ex"], c["Ey"], c["vac"], p, c["barrier"], keep_status=True)
        runs.append((r, t.last_status.copy(), t.last_hit.copy(), t.last_tally["steps"]))
    (r0, s0, h0, n0) = runs[0]
    for r, st, h, nsteps in runs[1:]:
        for a, b in zip(r, r0):
            np.testing.assert_array_equal(a, b)
        np.testing.assert_array_equal(st, s0)
        np.testing.assert_array_equal(h, h0)
        assert nsteps == n0
