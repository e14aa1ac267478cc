"""HIP particle tracer with the reference tracer API.

``HipTracer`` is a drop-in for ``CudaTracer`` (trench_cuda.py:21-138):
constructed once per geometry with the vacuum map, then called once per batch
as ``tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None,
stop=None, rf=None) -> (cnt_i, cnt_e, escaped, lost)`` (trench_cuda.py:41,
trench_charging_2d.py:542, :702).  Argument checks and their exception types
follow trench_cuda.py:42-59.  ``rf=`` runs the RF full-cycle push
(trench_charging_2d.py:556-602; trench_cuda.py:65-71, :134-137): the rf dict
holds ``period``, ``dt_max``, ``wave`` (n_phase, 2), the basis ``Ex``, ``Ey``,
``phi`` (2, nx, ny) and ``initial_time`` (n,).
"""
import ctypes
import threading

import numpy as np

from . import _lib

BACKENDS = ("hip", "auto")


def tapered_of(p):
    """Staircase corner crossings are tested when either wall is tapered
    (trench_charging_2d.py:554, trench_cuda.py:113)."""
    return bool(p.trench_taper_deg or (p.mask_t > 0 and p.mask_taper_deg))


class HipTracer:
    """Reuses the device geometry; uploads the frozen field and particles per call
    and traces them on the GPU (trench_cuda.py:21-39)."""

    #: 0 = FP64 push, bit-identical with the reference; 1 = FP32 push.
    precision = 0
    #: push schedule (see trench01_amd.autotune); the defaults are the fastest
    #: of the schedule sweep on the AR 8 workload (MI355X)
    refill = 16
    order = 1
    prio_steps = 0
    check_steps = 8
    pass_steps = 250
    blocks_per_cu = 0
    tile_pass = 0

    def __init__(self, vac, device=0):
        L = _lib.load()
        self.device = int(device)
        n_dev = _lib.device_count()
        if self.device < 0 or self.device >= n_dev:
            raise RuntimeError(f"HIP device {self.device} is not available ({n_dev} visible)")
        self.vac = np.asarray(vac, dtype=bool).copy()
        if self.vac.ndim != 2:
            raise ValueError("vac must be a 2-D (nx, ny) array")
        nx, ny = self.vac.shape
        handle = ctypes.c_void_p()
        v8 = np.ascontiguousarray(self.vac, dtype=np.uint8)
        _lib.check(L.th_tracer_create(_lib.ptr(v8, _lib.u8p), nx, ny, self.device, ctypes.byref(handle)))
        self._h = handle
        buf = ctypes.create_string_buffer(256)
        _lib.check(L.th_tracer_device_name(self._h, buf, 256))
        self.name = buf.value.decode(errors="replace")
        self.last_tally = None
        self.last_status = None
        self.last_hit = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().th_tracer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def surface_cells(self):
        """Flat indices (ix*ny+iy) of the surface cells the tallies are kept on."""
        L = _lib.load()
        m = ctypes.c_int32(0)
        _lib.check(L.th_tracer_surface(self._h, ctypes.byref(m), None))
        flat = np.empty(m.value, dtype=np.int32)
        _lib.check(L.th_tracer_surface(self._h, ctypes.byref(m), _lib.ptr(flat, _lib.i32p)))
        return flat

    def __call__(self, x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None, stop=None, rf=None,
                 keep_status=False):
        if np.shape(vac) != self.vac.shape or not np.array_equal(vac, self.vac):
            raise ValueError("the HIP tracer was built for another geometry; create a new tracer")
        if not (0 < p.cfl <= 1) or p.max_steps < 1:
            raise ValueError("HIP tracing needs 0 < cfl <= 1 and max_steps >= 1")
        return self._trace(x, y, vx, vy, qm, spc, Ex, Ey, p, barrier, record, stop, keep_status, rf)

    def set_rf(self, rf, n=None):
        """Upload the RF basis and wave (checked as trench_rf.validate,
        trench_rf.py:21-32); returns the launch times as a float64 array."""
        nx, ny = self.vac.shape
        period, dt_max = float(rf["period"]), float(rf["dt_max"])
        if not (np.isfinite(period) and period > 0 and np.isfinite(dt_max) and dt_max > 0):
            raise ValueError("RF period and dt_max must be positive and finite")
        basis = {}
        for key in ("phi", "Ex", "Ey"):
            a = np.ascontiguousarray(rf[key], dtype=np.float64)
            if a.shape != (2, nx, ny) or not np.isfinite(a).all():
                raise ValueError(f"RF basis {key} must be finite with shape (2, nx, ny)")
            basis[key] = a
        wave = np.ascontiguousarray(rf["wave"], dtype=np.float64)
        if wave.ndim != 2 or wave.shape[1] != 2 or len(wave) < 2 or not np.isfinite(wave).all():
            raise ValueError("RF wave must be a finite (n_phase >= 2, 2) array")
        t0 = None
        if n is not None:
            t0 = np.ascontiguousarray(rf["initial_time"], dtype=np.float64)
            if t0.shape != (n,) or not np.isfinite(t0).all():
                raise ValueError("RF initial_time must be finite with one entry per particle")
        _lib.check(_lib.load().th_set_rf(self._h, _lib.ptr(basis["Ex"], _lib.dp), _lib.ptr(basis["Ey"], _lib.dp),
                                         _lib.ptr(basis["phi"], _lib.dp), _lib.ptr(wave, _lib.dp), len(wave),
                                         period, dt_max))
        return t0

    def _trace(self, x, y, vx, vy, qm, spc, Ex, Ey, p, barrier, record, stop, keep_status, rf=None):
        L = _lib.load()
        n = len(x)
        nx, ny = self.vac.shape
        if any(np.shape(v) != (n,) for v in (x, y, vx, vy, qm, spc)):
            raise ValueError("particle arrays differ in length")
        if any(np.shape(v) != (nx, ny) for v in (Ex, Ey)):
            raise ValueError("field shape does not match the geometry")
        spc = np.ascontiguousarray(spc, dtype=np.int8)
        if np.any((spc != 0) & (spc != 1)):
            raise ValueError("species must be 0 (ion) or 1 (electron)")
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        ex, ey = f64(Ex), f64(Ey)
        if barrier is None:
            phi = wface = None
        else:
            if any(np.shape(v) != (nx, ny) for v in barrier):
                raise ValueError("wall potential shape does not match the geometry")
            phi, wface = (f64(v) for v in barrier)
        _lib.check(L.th_set_fields(self._h, _lib.ptr(ex, _lib.dp), _lib.ptr(ey, _lib.dp),
                                   _lib.ptr(phi, _lib.dp), _lib.ptr(wface, _lib.dp)))
        xs = [f64(v) for v in (x, y, vx, vy, qm)]
        prm = self.params_for(p, barrier is not None, rf is not None)
        t0 = None
        if rf is not None:
            if self.precision != 0:
                raise ValueError("the RF push runs in FP64 only")
            t0 = self.set_rf(rf, n)
        cnt_i = np.empty((nx, ny), dtype=np.int64)
        cnt_e = np.empty((nx, ny), dtype=np.int64)
        tally = _lib.Tally()
        recording = record is not None
        status = np.empty(n, dtype=np.int8) if (keep_status or recording) else None
        hit = np.empty(n, dtype=np.int32) if (keep_status or recording) else None
        if recording:
            if p.probe_trajectories < 1:
                raise ValueError("record mode needs probe_trajectories >= 1")
            if prm.precision != 0:
                raise ValueError("record mode runs the FP64 push only")
            # probe particles: trench_cuda.py:85-87
            selected = np.unique(np.linspace(0, n - 1, min(p.probe_trajectories, n), dtype=int))
            slots = np.full(n, -1, dtype=np.int32)
            slots[selected] = np.arange(selected.size, dtype=np.int32)
            stride = p.max_steps + 1
            ends = np.empty((n, 2))
            vel = np.empty((n, 2))
            paths = np.empty((max(selected.size, 1), stride, 2))
            lengths = np.ones(max(selected.size, 1), dtype=np.int32)
            flight = np.empty(n) if rf is not None else None
            clock = np.empty(n) if rf is not None else None
        flag = None
        watcher = None
        done = threading.Event()
        if stop is not None:
            # The C side polls this flag between launches; a watcher mirrors stop into it.
            flag = ctypes.c_int32(1 if stop.is_set() else 0)

            def watch():
                while not done.is_set():
                    if stop.is_set():
                        flag.value = 1
                        return
                    done.wait(0.002)

            watcher = threading.Thread(target=watch, daemon=True)
            watcher.start()
        try:
            head = (self._h, *(_lib.ptr(v, _lib.dp) for v in xs), _lib.ptr(spc, _lib.i8p))
            if rf is not None:
                head = head + (_lib.ptr(t0, _lib.dp),)
            args = (*head, n, ctypes.byref(prm),
                    ctypes.byref(flag) if flag is not None else None, _lib.ptr(cnt_i, _lib.i64p),
                    _lib.ptr(cnt_e, _lib.i64p), ctypes.byref(tally), _lib.ptr(status, _lib.i8p),
                    _lib.ptr(hit, _lib.i32p))
            if recording:
                rec_args = (_lib.ptr(ends, _lib.dp), _lib.ptr(vel, _lib.dp), _lib.ptr(slots, _lib.i32p),
                            int(selected.size), _lib.ptr(paths, _lib.dp), _lib.ptr(lengths, _lib.i32p))
                if rf is not None:
                    _lib.check(L.th_trace_rf_record(*args, *rec_args, _lib.ptr(flight, _lib.dp),
                                                    _lib.ptr(clock, _lib.dp)))
                else:
                    _lib.check(L.th_trace_record(*args, *rec_args))
            elif rf is not None:
                _lib.check(L.th_trace_rf(*args))
            else:
                _lib.check(L.th_trace(*args))
        finally:
            done.set()
            if watcher is not None:
                watcher.join()
        escaped = np.array([tally.escaped[0], tally.escaped[1]], dtype=np.int64)
        self.last_tally = dict(steps=int(tally.steps), off_surface=int(tally.off_surface),
                               cancelled=bool(tally.cancelled))
        self.last_status, self.last_hit = status, hit
        if recording:
            # the record dictionary of trench_cuda.py:123-133
            cells = np.full((n, 2), -1, dtype=int)
            absorbed = status == 1
            cells[absorbed] = np.column_stack((hit[absorbed] // ny, hit[absorbed] % ny))
            record.update(selected=selected, status=status, hit_cells=cells, initial_x=np.asarray(x).copy(),
                          endpoints=ends, final_velocity=vel,
                          paths=[paths[k, :lengths[k]].copy() for k in range(selected.size)],
                          cancelled=bool(tally.cancelled) or (stop is not None and stop.is_set()))
            if rf is not None:
                # trench_charging_2d.py:575-577, :680-682; trench_cuda.py:134-137
                period = float(rf["period"])
                record.update(launch_phase=t0 % period / period * 360, final_phase=clock / period * 360,
                              flight_time=flight)
        return cnt_i, cnt_e, escaped, int(tally.lost)

    # ---- device-resident batches (no reference counterpart; for batch loops and benchmarks)
    def params_for(self, p, barrier, rf=False):
        return _lib.TraceParams(dx=float(p.dx), cfl=float(p.cfl), max_steps=int(p.max_steps),
                                tapered=int(tapered_of(p)), barrier=int(bool(barrier)),
                                precision=0 if rf else int(self.precision), refill=int(self.refill), chunk=0,
                                order=int(self.order), prio_steps=int(self.prio_steps),
                                check_steps=int(self.check_steps), pass_steps=int(self.pass_steps), rf=int(bool(rf)),
                                blocks_per_cu=int(self.blocks_per_cu), tile_pass=int(self.tile_pass))

    def upload_times(self, offset, initial_time):
        """Launch times of pool[offset, offset+n) for resident RF traces."""
        t0 = np.ascontiguousarray(initial_time, dtype=np.float64)
        _lib.check(_lib.load().th_particles_upload_times(self._h, int(offset), _lib.ptr(t0, _lib.dp), t0.size))

    def set_fields(self, Ex, Ey, barrier=None):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        phi, wface = (None, None) if barrier is None else (f64(v) for v in barrier)
        _lib.check(_lib.load().th_set_fields(self._h, _lib.ptr(f64(Ex), _lib.dp), _lib.ptr(f64(Ey), _lib.dp),
                                             _lib.ptr(phi, _lib.dp), _lib.ptr(wface, _lib.dp)))

    def reserve(self, capacity):
        _lib.check(_lib.load().th_particles_reserve(self._h, int(capacity)))

    def upload(self, offset, x, y, vx, vy, qm, spc):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        xs = [f64(v) for v in (x, y, vx, vy, qm)]
        spc = np.ascontiguousarray(spc, dtype=np.int8)
        _lib.check(_lib.load().th_particles_upload(self._h, int(offset), *(_lib.ptr(v, _lib.dp) for v in xs),
                                                   _lib.ptr(spc, _lib.i8p), len(xs[0])))

    def trace_resident(self, offset, n, prm, with_counts=False):
        """Trace pool[offset, offset+n); returns (tally dict, counts or None, push_ms)."""
        L = _lib.load()
        tally = _lib.Tally()
        nx, ny = self.vac.shape
        ci = np.empty((nx, ny), np.int64) if with_counts else None
        ce = np.empty((nx, ny), np.int64) if with_counts else None
        _lib.check(L.th_trace_resident(self._h, int(offset), int(n), ctypes.byref(prm), _lib.ptr(ci, _lib.i64p),
                                       _lib.ptr(ce, _lib.i64p), ctypes.byref(tally), None, None))
        ms = ctypes.c_double(0.0)
        _lib.check(L.th_last_push_ms(self._h, ctypes.byref(ms)))
        t = dict(escaped=(int(tally.escaped[0]), int(tally.escaped[1])), lost=int(tally.lost),
                 steps=int(tally.steps), off_surface=int(tally.off_surface))
        return t, ((ci, ce) if with_counts else None), float(ms.value)

    def read_particles(self, offset, n):
        """Copy pool[offset, offset+n) back: (x, y, vx, vy, qm, spc)."""
        out = [np.empty(n) for _ in range(5)] + [np.empty(n, np.int8)]
        _lib.check(_lib.load().th_particles_read(self._h, int(offset), int(n),
                                                 *(_lib.ptr(a, _lib.dp) for a in out[:5]),
                                                 _lib.ptr(out[5], _lib.i8p)))
        return tuple(out)

    def read_field(self):
        """The push's current field (Ex, Ey) on the grid; zero on solid cells."""
        nx, ny = self.vac.shape
        ex, ey = np.empty((nx, ny)), np.empty((nx, ny))
        _lib.check(_lib.load().th_read_field(self._h, _lib.ptr(ex, _lib.dp), _lib.ptr(ey, _lib.dp)))
        return ex, ey

    def tangent(self, x, y, vx, vy, qm, spc, Ex, Ey, p, barrier=None, dstate=None, dfield=None):
        """Record-mode FP64 trace with the pathwise derivative of every path
        along one direction: ``dstate`` = (dx, dy, dvx, dvy) initial tangents
        (entries may be None = 0), ``dfield`` = (dEx, dEy) field tangent or
        None.  Returns dict(status, hit, steps, face, endpoints, d_endpoints,
        final_velocity, d_final_velocity, escaped, lost, total_steps)."""
        L = _lib.load()
        n = len(x)
        nx, ny = self.vac.shape
        if not (0 < p.cfl <= 1) or p.max_steps < 1:
            raise ValueError("HIP tracing needs 0 < cfl <= 1 and max_steps >= 1")
        if any(np.shape(v) != (n,) for v in (x, y, vx, vy, qm, spc)):
            raise ValueError("particle arrays differ in length")
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.set_fields(Ex, Ey, barrier)
        xs = [f64(v) for v in (x, y, vx, vy, qm)]
        spc = np.ascontiguousarray(spc, dtype=np.int8)
        ds = [None] * 4 if dstate is None else [f64(v) for v in dstate]
        if any(v is not None and v.shape != (n,) for v in ds):
            raise ValueError("initial tangents must have one entry per particle")
        dex, dey = (None, None) if dfield is None else (f64(v) for v in dfield)
        if dex is not None and (dex.shape != (nx, ny) or dey.shape != (nx, ny)):
            raise ValueError("field tangent shape does not match the geometry")
        prm = self.params_for(p, barrier is not None)
        prm.precision = 0
        out = dict(status=np.empty(n, np.int8), hit=np.empty(n, np.int32), steps=np.empty(n, np.int32),
                   face=np.empty(n, np.int8), endpoints=np.empty((n, 2)), d_endpoints=np.empty((n, 2)),
                   final_velocity=np.empty((n, 2)), d_final_velocity=np.empty((n, 2)))
        tally = _lib.Tally()
        _lib.check(L.th_trace_tangent(self._h, *(_lib.ptr(v, _lib.dp) for v in xs), _lib.ptr(spc, _lib.i8p), n,
                                      ctypes.byref(prm), *(_lib.ptr(v, _lib.dp) for v in ds),
                                      _lib.ptr(dex, _lib.dp), _lib.ptr(dey, _lib.dp), ctypes.byref(tally),
                                      _lib.ptr(out["status"], _lib.i8p), _lib.ptr(out["hit"], _lib.i32p),
                                      _lib.ptr(out["steps"], _lib.i32p), _lib.ptr(out["face"], _lib.i8p),
                                      *(_lib.ptr(out[k], _lib.dp) for k in ("endpoints", "d_endpoints",
                                                                           "final_velocity", "d_final_velocity"))))
        out.update(escaped=np.array([tally.escaped[0], tally.escaped[1]], dtype=np.int64), lost=int(tally.lost),
                   total_steps=int(tally.steps))
        return out

    def soft(self, offset, n, p, width, cut=30.0, dstate=None, dfield=None, dphi=None, events=0, per_particle=False):
        """Relaxed-barrier push of the resident pool[offset, offset+n) through
        the fields last set (``set_fields`` with a barrier), with the forward
        derivative along one direction (th_trace_soft): a barrier test
        reflects with probability 1 / (1 + exp(t)), t = (vn^2/2 - q/m dphi) /
        (|q/m| ``width`` [V]), sharp for |t| >= ``cut``; the absorbed shares
        are face-smoothed deposits.  ``dstate`` = (dx, dy, dvx, dvy) initial
        tangents (n each, entries may be None), ``dfield`` = (dEx, dEy) and
        ``dphi`` [nx, ny] the field and potential tangents (None = 0).
        Returns dict(deposit [2, m], d_deposit [2, m], escaped, lost,
        total_steps, push_ms) plus status / hit / steps with
        ``per_particle`` and the deposit list (numpy structured array of
        th_soft_event) when ``events`` > 0 gives its capacity."""
        L = _lib.load()
        nx, ny = self.vac.shape
        m = self.surface_cells().size
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        ds = [None] * 4 if dstate is None else [f64(v) for v in dstate]
        if any(v is not None and v.shape != (n,) for v in ds):
            raise ValueError("initial tangents must have one entry per particle")
        dex, dey = (None, None) if dfield is None else (f64(v) for v in dfield)
        dph = f64(dphi)
        for a in (dex, dey, dph):
            if a is not None and a.shape != (nx, ny):
                raise ValueError("field and potential tangents must match the geometry")
        prm = self.params_for(p, True)
        prm.precision = 0
        dep, ddep = np.empty((2, m)), np.empty((2, m))
        pp = dict(status=np.empty(n, np.int8), hit=np.empty(n, np.int32),
                  steps=np.empty(n, np.int32)) if per_particle else {}
        ev = np.zeros(int(events), dtype=SOFT_EVENT) if events else None
        n_ev = ctypes.c_int64(0)
        tally = _lib.Tally()
        _lib.check(L.th_trace_soft(self._h, int(offset), int(n), ctypes.byref(prm), float(width), float(cut),
                                   *(_lib.ptr(v, _lib.dp) for v in ds), _lib.ptr(dex, _lib.dp),
                                   _lib.ptr(dey, _lib.dp), _lib.ptr(dph, _lib.dp), ctypes.byref(tally),
                                   _lib.ptr(dep, _lib.dp), _lib.ptr(ddep, _lib.dp),
                                   _lib.ptr(pp.get("status"), _lib.i8p), _lib.ptr(pp.get("hit"), _lib.i32p),
                                   _lib.ptr(pp.get("steps"), _lib.i32p), _lib.ptr(ev, ctypes.POINTER(_lib.SoftEvent)),
                                   int(events), ctypes.byref(n_ev)))
        ms = ctypes.c_double(0.0)
        _lib.check(L.th_last_push_ms(self._h, ctypes.byref(ms)))
        out = dict(deposit=dep, d_deposit=ddep, escaped=np.array([tally.escaped[0], tally.escaped[1]], np.int64),
                   lost=int(tally.lost), total_steps=int(tally.steps), push_ms=float(ms.value),
                   n_events=int(n_ev.value), **pp)
        if events:
            if n_ev.value > events:
                raise RuntimeError(f"{n_ev.value} relaxed-push deposits exceed the event capacity {events}")
            out["events"] = ev[:n_ev.value]
        return out

    def soft_adjoint(self, offset, n, p, width, steps, gbar, cut=30.0, field=True, potential=True):
        """Reverse mode of ``soft`` (th_trace_soft_adjoint): for the cotangent
        ``gbar`` [2, m] of the relaxed deposits of pool[offset, offset+n)
        (``steps`` from the ``soft`` call with ``per_particle``), returns
        (bar_state [n, 4], bar_field [nx, ny, 2] or None, bar_phi [nx, ny] or
        None): the cotangents of the initial states, the vacuum-masked field
        and the potential."""
        L = _lib.load()
        nx, ny = self.vac.shape
        steps = np.ascontiguousarray(steps, dtype=np.int32)
        g = np.ascontiguousarray(gbar, dtype=np.float64)
        if steps.shape != (n,) or g.shape != (2, self.surface_cells().size):
            raise ValueError("steps must be [n] and gbar [2, m]")
        prm = self.params_for(p, True)
        prm.precision = 0
        bar_state = np.empty((n, 4))
        bar_field = np.empty((nx, ny, 2)) if field else None
        bar_phi = np.empty((nx, ny)) if potential else None
        ms = ctypes.c_double(0.0)
        _lib.check(L.th_trace_soft_adjoint(self._h, int(offset), int(n), ctypes.byref(prm), float(width), float(cut),
                                           _lib.ptr(steps, _lib.i32p), _lib.ptr(g, _lib.dp),
                                           _lib.ptr(bar_state, _lib.dp), _lib.ptr(bar_field, _lib.dp),
                                           _lib.ptr(bar_phi, _lib.dp), ctypes.byref(ms)))
        self.last_adjoint_ms = ms.value
        return bar_state, bar_field, bar_phi

    def adjoint(self, x, y, vx, vy, qm, spc, Ex, Ey, p, barrier, steps, bar_end=None, bar_vel=None, field=True):
        """Reverse-mode pathwise derivative (th_trace_adjoint): the transpose
        of ``tangent``.  ``steps`` are the particles' path lengths from a
        ``tangent`` (or record) trace of the same particles and field;
        ``bar_end`` / ``bar_vel`` ([n, 2] each, None = 0) the cotangents of the
        recorded end points and final velocities.  Returns (bar_state [n, 4]
        = cotangent of (x, y, vx, vy) at launch, bar_field [nx, ny, 2] =
        cotangent of the vacuum-masked (Ex, Ey) per cell, or None)."""
        L = _lib.load()
        n = len(x)
        nx, ny = self.vac.shape
        if any(np.shape(v) != (n,) for v in (x, y, vx, vy, qm, spc, steps)):
            raise ValueError("particle arrays differ in length")
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self.set_fields(Ex, Ey, barrier)
        xs = [f64(v) for v in (x, y, vx, vy, qm)]
        spc = np.ascontiguousarray(spc, dtype=np.int8)
        steps = np.ascontiguousarray(steps, dtype=np.int32)
        be, bv = f64(bar_end), f64(bar_vel)
        for b in (be, bv):
            if b is not None and b.shape != (n, 2):
                raise ValueError("cotangents must be [n, 2]")
        prm = self.params_for(p, barrier is not None)
        prm.precision = 0
        bar_state = np.empty((n, 4))
        bar_field = np.empty((nx, ny, 2)) if field else None
        ms = ctypes.c_double(0.0)
        _lib.check(L.th_trace_adjoint(self._h, *(_lib.ptr(v, _lib.dp) for v in xs), _lib.ptr(spc, _lib.i8p), n,
                                      ctypes.byref(prm), _lib.ptr(steps, _lib.i32p), _lib.ptr(be, _lib.dp),
                                      _lib.ptr(bv, _lib.dp), _lib.ptr(bar_state, _lib.dp),
                                      _lib.ptr(bar_field, _lib.dp), ctypes.byref(ms)))
        self.last_adjoint_ms = ms.value
        return bar_state, bar_field

    def selftest_math(self, a, b):
        """Device arithmetic on arrays (for arithmetic pinning): columns hypot(a, b), sqrt(a), a / b,
        fmod(a, b), the range-free sqrt(|a|) and a / b, the CFL step 1.75e-9 / (|a| + sqrt(1e-8 |b|) + 1), and
        hypot(a, b), hypot(b 2^-500, a) from the paired evaluation."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.empty((a.size, 9), dtype=np.float64)
        _lib.check(_lib.load().th_selftest_math(self._h, _lib.ptr(a, _lib.dp), _lib.ptr(b, _lib.dp), a.size,
                                                _lib.ptr(out, _lib.dp)))
        return out


# numpy layout of th_soft_event (64 bytes)
SOFT_EVENT = np.dtype([("x", "f8"), ("y", "f8"), ("dx", "f8"), ("dy", "f8"), ("a", "f8"), ("da", "f8"),
                       ("pid", "i4"), ("cell", "i4"), ("face", "i1"), ("kind", "i1"), ("species", "i1"),
                       ("pad0", "i1"), ("pad1", "i4")])


def make_particle_tracer(p, vac, log=None):
    """Backend factory with the reference's contract (trench_charging_2d.py:705-724):
    returns ``(tracer, backend_used)``.  ``"hip"`` raises RuntimeError with the
    reason when the GPU path cannot start.  ``"auto"`` also raises: this package
    has no CPU path to fall back to (the reference's own factory keeps its
    ``trace_cpu`` for that)."""
    backend = getattr(p, "backend", "hip")
    if backend not in BACKENDS:
        raise ValueError(f"backend must be one of {' / '.join(BACKENDS)}")
    device = int(getattr(p, "hip_device", getattr(p, "cuda_device", 0)))
    try:
        tracer = HipTracer(vac, device)
    except Exception as exc:
        raise RuntimeError(f"HIP tracer could not start: {exc}") from exc
    if log is not None:
        log(f"particle tracing: HIP ({tracer.name}, device {device}, float64)")
    return tracer, "hip"


def trace(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier=None, record=None, stop=None, rf=None):
    """One-shot tracer call (trench_charging_2d.py:727-730)."""
    tracer, _ = make_particle_tracer(p, vac)
    try:
        return tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p, barrier, record, stop, rf=rf)
    finally:
        tracer.close()

