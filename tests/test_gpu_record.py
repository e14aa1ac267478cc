"""Record mode and the FP32 push on the GPU.

Record mode (trench_trace.cu:149-183, trench_cuda.py:81-137) is checked bit
for bit against the reference's own record-mode outputs (tests/golden/
trace_*.npz): status, hit cells, end points clipped to the crossed face,
final velocities and probe paths.  The FP32 push is a precision variant: its
tallies agree with the FP64 push statistically, not bit for bit."""
import threading

import numpy as np
import pytest

from oracle import charging_ref as R
from tests import cases
from tests.test_golden import TRACE_CASES, load, params_of
from trench01_amd import HipTracer, configs

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.mark.parametrize("name", TRACE_CASES)
def test_record_mode_matches_reference_golden(name):
    _case(f"record {name}")
    g = load(f"trace_{name}")
    p = params_of(g)
    vac = ~g["solid"]
    barrier = (g["phi"], g["wface"]) if p.wall_model == "barrier" else None
    t = HipTracer(vac)
    record = {}
    ci, ce, esc, lost = t(g["x"], g["y"], g["vx"], g["vy"], g["qm"], g["spc"], g["Ex"], g["Ey"], vac, p, barrier,
                          record)
    np.testing.assert_array_equal(ci, g["cnt_i"])
    np.testing.assert_array_equal(ce, g["cnt_e"])
    np.testing.assert_array_equal(esc, g["escaped"])
    assert lost == int(g["lost"])
    np.testing.assert_array_equal(record["status"], g["status"])
    np.testing.assert_array_equal(record["hit_cells"], g["hit_cells"])
    np.testing.assert_array_equal(record["endpoints"], g["endpoints"])
    np.testing.assert_array_equal(record["final_velocity"], g["final_velocity"])
    np.testing.assert_array_equal(record["selected"], g["selected"])
    np.testing.assert_array_equal([len(q) for q in record["paths"]], g["path_lengths"])
    np.testing.assert_array_equal(np.concatenate(record["paths"]), g["paths"])
    np.testing.assert_array_equal(record["initial_x"], g["x"])
    assert record["cancelled"] is False


def test_record_cancellation_keeps_initial_paths():
    """trench01's tests/test_cuda.py:82-92: a stop before the first launch."""
    _case("record cancellation")
    p = cases.small_params(probe_trajectories=12)
    solid, _, _, _ = R.build_geometry(p)
    parts = R.inject(p, 20, solid.shape[1], np.random.default_rng(1))
    stop = threading.Event()
    stop.set()
    record = {}
    z = np.zeros(solid.shape)
    ci, ce, esc, lost = HipTracer(~solid)(*parts, z, z, ~solid, p, record=record, stop=stop)
    assert ci.sum() == ce.sum() == esc.sum() == 0 and lost == 40
    assert record["cancelled"] and all(v.shape == (1, 2) for v in record["paths"])
    np.testing.assert_array_equal(record["endpoints"], np.column_stack(parts[:2]))


def test_record_empty_batch():
    """trench01's tests/test_cuda.py:95-103."""
    _case("record empty")
    p = cases.small_params(probe_trajectories=12)
    solid, _, _, _ = R.build_geometry(p)
    z = np.zeros(solid.shape)
    e = np.zeros(0)
    record = {}
    ci, ce, esc, lost = HipTracer(~solid)(e, e, e, e, e, np.zeros(0, np.int8), z, z, ~solid, p, record=record)
    assert ci.sum() == ce.sum() == esc.sum() == lost == 0
    assert record["paths"] == []


def test_fp32_push_agrees_statistically_with_fp64():
    """Same AR 8 batch through a time-marched (charged) field: the FP32 push's
    totals per species and outcome, and its charge on the bottom and the two
    sidewalls, lie within 5 standard deviations (Poisson) of the FP64 push's."""
    _case("fp32 vs fp64")
    c = cases.marched_case(cases.ar8_params(), 30, 5000, 100_000, ion=configs.AR8_ION)
    p = c["p"]
    x, y, vx, vy, qm, spc = c["particles"]
    t64 = HipTracer(c["vac"])
    t32 = HipTracer(c["vac"])
    t32.precision = 1
    r64 = t64(x, y, vx, vy, qm, spc, c["Ex"], c["Ey"], c["vac"], p, c["barrier"])
    r32 = t32(x, y, vx, vy, qm, spc, c["Ex"], c["Ey"], c["vac"], p, c["barrier"])
    solid = c["solid"]
    regions = R.charge_monitors(solid, R.build_geometry(p)[1], *R.build_geometry(p)[2:], p)

    def within(a, b, what):
        assert abs(a - b) <= 5 * np.sqrt(max(a + b, 1)) + 2, (what, a, b)

    for k in (0, 1):
        within(r64[k].sum(), r32[k].sum(), f"absorbed species {k}")
        within(r64[2][k], r32[2][k], f"escaped species {k}")
        for name in ("bottom", "left", "right"):
            within(r64[k].ravel()[regions[name]].sum(), r32[k].ravel()[regions[name]].sum(), f"{name} {k}")
    within(r64[3], r32[3], "lost")
    print("fp64", [int(r64[k].sum()) for k in (0, 1)], list(r64[2]), r64[3],
          "fp32", [int(r32[k].sum()) for k in (0, 1)], list(r32[2]), r32[3], flush=True)
