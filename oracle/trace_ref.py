"""ORACLE -- test infrastructure only (tests, smoke() and the bench's cpu_baseline
leg).  ctypes binding of ``trace_ref.c``, the C restatement of trench01's FP64
particle push (trench_trace.cu:17-193, trench_charging_2d.py:542-702).

``oracle_tracer`` has the reference tracer signature
``tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None,
stop=None, rf=None) -> (cnt_i, cnt_e, escaped, lost)`` (trench_charging_2d.py:542,
:702), with the record dict filled as trace_cpu fills it.  ``trace_rf_arrays``
runs the RF full-cycle push (trench_charging_2d.py:556-602); its ``blas_tail``
switch is explained in trace_ref.c.
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle_trace.so"
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_i8p = ctypes.POINTER(ctypes.c_int8)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build(force=False):
    """Compile trace_ref.c with contraction off (the reference's --fmad=false)."""
    if LIB.exists() and not force:
        return LIB
    tmp = LIB.with_suffix(".so.tmp%d" % os.getpid())
    subprocess.run(["gcc", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-shared",
                    "-fPIC", str(HERE / "trace_ref.c"), "-o", str(tmp), "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        L.oracle_trace.restype = ctypes.c_int
        L.oracle_trace.argtypes = [_dp, _dp, _dp, _dp, _dp, _i8p, ctypes.c_int64,
                                   _dp, _dp, _u8p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                   ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                   ctypes.c_int, _i8p, _i32p, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p,
                                   _dp, _dp, _i32p, _dp, _i32p]
        L.oracle_trace_rf.restype = ctypes.c_int
        L.oracle_trace_rf.argtypes = [_dp, _dp, _dp, _dp, _dp, _i8p, ctypes.c_int64,
                                      _dp, _dp, _u8p, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      _dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                      _dp, _dp, ctypes.c_int,
                                      _i8p, _i32p, _i64p, _i64p, _i64p, _i64p, _i64p, _i64p,
                                      _dp, _dp, _i32p, _dp, _i32p, _dp, _dp]
        L.oracle_hypot.restype = None
        L.oracle_hypot.argtypes = [_dp, _dp, _dp, ctypes.c_int64]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def is_tapered(p):
    """trench_charging_2d.py:554 / trench_cuda.py:113."""
    return bool(p.trench_taper_deg or (p.mask_t > 0 and p.mask_taper_deg))


def probe_selection(n, probe_trajectories):
    """Probe particles of record mode (trench_cuda.py:85-87, trench_charging_2d.py:566-569)."""
    return np.unique(np.linspace(0, n - 1, min(probe_trajectories, n), dtype=int))


def trace_arrays(x, y, vx, vy, qm, spc, Ex, Ey, vac, dx, cfl, max_steps, tapered,
                 barrier=None, threads=0, record=False, probe_trajectories=0):
    """Run the C push.  Returns dict(status, hit, steps, cnt_i, cnt_e, escaped, lost, total_steps)
    plus, in record mode, endpoints, final_velocity, selected and paths (list of (k, 2))."""
    vac = np.ascontiguousarray(vac, dtype=np.uint8)
    nx, ny = vac.shape
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    x, y, vx, vy, qm = (f64(a) for a in (x, y, vx, vy, qm))
    spc = np.ascontiguousarray(spc, dtype=np.int8)
    n = x.size
    Ex, Ey = f64(Ex), f64(Ey)
    phi = wface = None
    if barrier is not None:
        phi, wface = (f64(a) for a in barrier)
    out = dict(status=np.empty(n, np.int8), hit=np.empty(n, np.int32), steps=np.empty(n, np.int64),
               cnt_i=np.empty((nx, ny), np.int64), cnt_e=np.empty((nx, ny), np.int64),
               escaped=np.zeros(2, np.int64))
    lost = ctypes.c_int64(0)
    total = ctypes.c_int64(0)
    ends = vel = slots = paths = plen = None
    if record:
        ends = np.empty((n, 2))
        vel = np.empty((n, 2))
        selected = probe_selection(n, probe_trajectories) if n else np.zeros(0, int)
        slots = np.full(n, -1, np.int32)
        slots[selected] = np.arange(selected.size)
        paths = np.zeros((max(selected.size, 1), max_steps + 1, 2))
        plen = np.zeros(max(selected.size, 1), np.int32)
    rc = lib().oracle_trace(_ptr(x, _dp), _ptr(y, _dp), _ptr(vx, _dp), _ptr(vy, _dp), _ptr(qm, _dp),
                            _ptr(spc, _i8p), n, _ptr(Ex, _dp), _ptr(Ey, _dp), _ptr(vac, _u8p),
                            nx, ny, _ptr(phi, _dp), _ptr(wface, _dp), float(dx), float(cfl),
                            int(max_steps), int(bool(tapered)), int(threads),
                            _ptr(out["status"], _i8p), _ptr(out["hit"], _i32p), _ptr(out["steps"], _i64p),
                            _ptr(out["cnt_i"], _i64p), _ptr(out["cnt_e"], _i64p),
                            _ptr(out["escaped"], _i64p), ctypes.byref(lost), ctypes.byref(total),
                            _ptr(ends, _dp), _ptr(vel, _dp), _ptr(slots, _i32p), _ptr(paths, _dp),
                            _ptr(plen, _i32p))
    if rc != 0:
        raise ValueError("species must be 0 (ion) or 1 (electron)")
    if record:
        out.update(endpoints=ends, final_velocity=vel, selected=selected,
                   paths=[paths[k, :plen[k]].copy() for k in range(selected.size)])
    out["lost"] = int(lost.value)
    out["total_steps"] = int(total.value)
    return out


def rf_arrays(rf, n, shape):
    """The reference's rf dict (trench_charging_2d.py:410-412, :1232-1233) as
    contiguous arrays, checked as trench_rf.validate (trench_rf.py:21-32) does."""
    period, dt_max = float(rf["period"]), float(rf["dt_max"])
    if not (np.isfinite(period) and period > 0 and np.isfinite(dt_max) and dt_max > 0):
        raise ValueError("RF period and dt_max must be positive and finite")
    out = {}
    for key in ("Ex", "Ey", "phi"):
        a = np.ascontiguousarray(rf[key], dtype=np.float64)
        if a.shape != (2, *shape) or not np.isfinite(a).all():
            raise ValueError(f"RF basis {key} must be finite with shape (2, nx, ny)")
        out[key] = a
    wave = np.ascontiguousarray(rf["wave"], dtype=np.float64)
    if wave.ndim != 2 or wave.shape[1] != 2 or len(wave) < 2 or not np.isfinite(wave).all():
        raise ValueError("RF wave must be a finite (n_phase >= 2, 2) array")
    t0 = np.ascontiguousarray(rf["initial_time"], dtype=np.float64)
    if t0.shape != (n,) or not np.isfinite(t0).all():
        raise ValueError("RF initial_time must be finite with one entry per particle")
    out.update(wave=wave, initial_time=t0, period=period, dt_max=dt_max,
               bound=np.ascontiguousarray(np.abs(wave).max(axis=0)))
    return out


def trace_rf_arrays(x, y, vx, vy, qm, spc, Ex, Ey, vac, dx, cfl, max_steps, tapered, rf,
                    barrier=None, record=False, probe_trajectories=0, blas_tail=False):
    """Run the C RF push.  Returns the dict of ``trace_arrays`` plus, in record
    mode, flight_time, launch_phase and final_phase [deg]."""
    vac = np.ascontiguousarray(vac, dtype=np.uint8)
    nx, ny = vac.shape
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    x, y, vx, vy, qm = (f64(a) for a in (x, y, vx, vy, qm))
    spc = np.ascontiguousarray(spc, dtype=np.int8)
    n = x.size
    r = rf_arrays(rf, n, (nx, ny))
    Ex, Ey = f64(Ex), f64(Ey)
    phi = wface = None
    if barrier is not None:
        phi, wface = (f64(a) for a in barrier)
    out = dict(status=np.empty(n, np.int8), hit=np.empty(n, np.int32), steps=np.empty(n, np.int64),
               cnt_i=np.empty((nx, ny), np.int64), cnt_e=np.empty((nx, ny), np.int64),
               escaped=np.zeros(2, np.int64))
    lost = ctypes.c_int64(0)
    total = ctypes.c_int64(0)
    ends = vel = slots = paths = plen = flight = clock = None
    if record:
        ends, vel = np.empty((n, 2)), np.empty((n, 2))
        flight, clock = np.empty(n), np.empty(n)
        selected = probe_selection(n, probe_trajectories) if n else np.zeros(0, int)
        slots = np.full(n, -1, np.int32)
        slots[selected] = np.arange(selected.size)
        paths = np.zeros((max(selected.size, 1), max_steps + 1, 2))
        plen = np.zeros(max(selected.size, 1), np.int32)
    rc = lib().oracle_trace_rf(
        _ptr(x, _dp), _ptr(y, _dp), _ptr(vx, _dp), _ptr(vy, _dp), _ptr(qm, _dp), _ptr(spc, _i8p), n,
        _ptr(Ex, _dp), _ptr(Ey, _dp), _ptr(vac, _u8p), nx, ny, _ptr(phi, _dp), _ptr(wface, _dp),
        float(dx), float(cfl), int(max_steps), int(bool(tapered)),
        _ptr(r["Ex"], _dp), _ptr(r["Ey"], _dp), _ptr(r["phi"], _dp), _ptr(r["wave"], _dp), len(r["wave"]),
        r["period"], r["dt_max"], _ptr(r["bound"], _dp), _ptr(r["initial_time"], _dp), int(bool(blas_tail)),
        _ptr(out["status"], _i8p), _ptr(out["hit"], _i32p), _ptr(out["steps"], _i64p),
        _ptr(out["cnt_i"], _i64p), _ptr(out["cnt_e"], _i64p), _ptr(out["escaped"], _i64p),
        ctypes.byref(lost), ctypes.byref(total),
        _ptr(ends, _dp), _ptr(vel, _dp), _ptr(slots, _i32p), _ptr(paths, _dp), _ptr(plen, _i32p),
        _ptr(flight, _dp), _ptr(clock, _dp))
    if rc != 0:
        raise ValueError("species must be 0 (ion) or 1 (electron)")
    if record:
        period = r["period"]
        out.update(endpoints=ends, final_velocity=vel, selected=selected,
                   paths=[paths[k, :plen[k]].copy() for k in range(selected.size)],
                   flight_time=flight, final_phase=clock / period * 360,
                   launch_phase=(r["initial_time"] % period) / period * 360)
    out["lost"] = int(lost.value)
    out["total_steps"] = int(total.value)
    return out


def fill_record(record, r, n):
    """The record dict of trace_cpu (trench_charging_2d.py:565-577, :652-686, :699-701)."""
    hit_cells = np.full((n, 2), -1, dtype=int)
    absorbed = r["status"] == 1
    hit_cells[absorbed, 0] = r["hit"][absorbed] // r["cnt_i"].shape[1]
    hit_cells[absorbed, 1] = r["hit"][absorbed] % r["cnt_i"].shape[1]
    record.update(selected=r["selected"], status=r["status"], hit_cells=hit_cells,
                  endpoints=r["endpoints"], final_velocity=r["final_velocity"], paths=r["paths"],
                  cancelled=False)
    for key in ("launch_phase", "final_phase", "flight_time"):
        if key in r:
            record[key] = r[key]


def oracle_tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None, stop=None, rf=None,
                  blas_tail=False):
    """Reference tracer API (trench_charging_2d.py:542) over the C push."""
    args = (x, y, vx, vy, qm, spc, Ex, Ey, vac, p.dx, p.cfl, p.max_steps, is_tapered(p))
    rec = record is not None
    if rf is not None:
        r = trace_rf_arrays(*args, rf, barrier, rec, p.probe_trajectories, blas_tail)
    else:
        r = trace_arrays(*args, barrier, record=rec, probe_trajectories=p.probe_trajectories)
    if rec:
        record["initial_x"] = np.array(x, dtype=np.float64)
        fill_record(record, r, np.size(x))
    return r["cnt_i"], r["cnt_e"], r["escaped"], r["lost"]


def glibc_hypot(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty_like(a)
    lib().oracle_hypot(_ptr(a, _dp), _ptr(b, _dp), _ptr(out, _dp), a.size)
    return out
