"""Many trench instances in one persistent push launch (th_ensemble).

``Ensemble`` binds the C-ABI: instances are added with their vacuum map,
frozen field (and wall potential for the barrier model), taper flag and batch
size; a batch is traced by one launch per pass over all instances, by a
captured hipGraph of that batch, or -- for comparison -- by one launch per
instance and pass, the schedule of a tracer that takes one geometry per call
(trench_cuda.py:41-138).  Every instance gets bit-identical results in all
three modes and equal to a single-instance trace of the same particles.

``batching_workload`` draws the heterogeneous workload: trench width, aspect
ratio, mask thickness and kind, tapers, ion energy distribution and particle
count per instance; ``instance_field`` sets up each instance's charged field
on the host (Poisson LU of its own geometry).
"""
import ctypes
from dataclasses import dataclass, replace

import numpy as np

from . import _lib, model
from .charging import ion_spec
from .model import EPS0
from .tracer import tapered_of

MODES = {"one": 0, "graph": 1, "per_instance": 2}


class Ensemble:
    """A set of trench instances traced together on one device."""

    def __init__(self, device=0):
        n_dev = _lib.device_count()
        if not 0 <= int(device) < n_dev:
            raise RuntimeError(f"HIP device {device} is not available ({n_dev} visible)")
        h = ctypes.c_void_p()
        _lib.check(_lib.load().th_ensemble_create(int(device), ctypes.byref(h)))
        self._h = h
        self.shapes = []
        self.sizes = []

    def close(self):
        if self._h:
            _lib.load().th_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, vac, Ex, Ey, barrier=None, tapered=False, n_particles=0):
        """Append an instance; returns its index.  ``barrier`` = (phi, wface) or None."""
        vac8 = np.ascontiguousarray(vac, dtype=np.uint8)
        nx, ny = vac8.shape
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        ex, ey = f64(Ex), f64(Ey)
        phi, wf = (None, None) if barrier is None else (f64(barrier[0]), f64(barrier[1]))
        for a in (ex, ey) + (() if phi is None else (phi, wf)):
            if a.shape != (nx, ny):
                raise ValueError("fields must have the vacuum map's shape")
        idx = ctypes.c_int32()
        _lib.check(_lib.load().th_ensemble_add(self._h, _lib.ptr(vac8, _lib.u8p), nx, ny, _lib.ptr(ex, _lib.dp),
                                               _lib.ptr(ey, _lib.dp), _lib.ptr(phi, _lib.dp), _lib.ptr(wf, _lib.dp),
                                               int(bool(tapered)), int(n_particles), ctypes.byref(idx)))
        self.shapes.append((nx, ny))
        self.sizes.append(int(n_particles))
        return idx.value

    def set_fields(self, index, Ex, Ey, barrier=None):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        phi, wf = (None, None) if barrier is None else (f64(barrier[0]), f64(barrier[1]))
        _lib.check(_lib.load().th_ensemble_set_fields(self._h, int(index), _lib.ptr(f64(Ex), _lib.dp),
                                                      _lib.ptr(f64(Ey), _lib.dp), _lib.ptr(phi, _lib.dp),
                                                      _lib.ptr(wf, _lib.dp)))

    def surface(self, index):
        m = ctypes.c_int32()
        _lib.check(_lib.load().th_ensemble_surface(self._h, int(index), ctypes.byref(m), None))
        flat = np.empty(m.value, dtype=np.int32)
        _lib.check(_lib.load().th_ensemble_surface(self._h, int(index), ctypes.byref(m), _lib.ptr(flat, _lib.i32p)))
        return flat

    def set_inject(self, index, p):
        """Device injection of the instance's batch with the distributions of
        Params ``p`` (its ion_distribution fields), half ions, half electrons."""
        nx, ny = self.shapes[index]
        spec = ion_spec(p)
        spec.n = self.sizes[index] // 2
        spec.lx = nx * p.dx
        spec.y0 = ny * p.dx - 1e-3 * p.dx
        _lib.check(_lib.load().th_ensemble_set_inject(self._h, int(index), ctypes.byref(spec)))

    def upload(self, index, x, y, vx, vy, qm, spc):
        f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        xs = [f64(a) for a in (x, y, vx, vy, qm)]
        sp = np.ascontiguousarray(spc, dtype=np.int8)
        _lib.check(_lib.load().th_ensemble_upload(self._h, int(index), *(_lib.ptr(a, _lib.dp) for a in xs),
                                                  _lib.ptr(sp, _lib.i8p), sp.size))

    def run(self, prm, seed=0, first_batch=0, batches=1, mode="one", inject=False, status=False):
        """Trace ``batches`` batches (tallies accumulate); returns the elapsed ms."""
        flags = (1 if inject else 0) | (2 if status else 0)
        _lib.check(_lib.load().th_ensemble_run(self._h, ctypes.byref(prm), int(seed) & (2**64 - 1),
                                               int(first_batch), int(batches), MODES[mode], flags))
        ms = ctypes.c_double()
        _lib.check(_lib.load().th_ensemble_last_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def read(self, index, status=False):
        """Tallies of one instance: cnt_i, cnt_e (nx, ny), escaped [2], lost,
        steps, off_surface, and (status=True) per-particle status and hit cell."""
        nx, ny = self.shapes[index]
        ci = np.zeros((nx, ny), dtype=np.int64)
        ce = np.zeros_like(ci)
        t = _lib.Tally()
        st = np.empty(self.sizes[index], dtype=np.int8) if status else None
        hit = np.empty(self.sizes[index], dtype=np.int32) if status else None
        _lib.check(_lib.load().th_ensemble_read(self._h, int(index), _lib.ptr(ci, _lib.i64p),
                                                _lib.ptr(ce, _lib.i64p), ctypes.byref(t), _lib.ptr(st, _lib.i8p),
                                                _lib.ptr(hit, _lib.i32p)))
        out = dict(cnt_i=ci, cnt_e=ce, escaped=np.array(t.escaped[:], dtype=np.int64), lost=int(t.lost),
                   steps=int(t.steps), off_surface=int(t.off_surface))
        if status:
            out.update(status=st, hit=hit)
        return out

    def totals(self):
        """All instances' tallies: counts [2, slots] (instance-major surface
        slots) and scalars [n_instances, 5] (escaped ions, escaped electrons,
        lost, steps, off-surface absorptions)."""
        slots = sum(self.surface(i).size for i in range(len(self.shapes)))
        counts = np.empty((2, slots), dtype=np.int64)
        scal = np.empty((len(self.shapes), 5), dtype=np.int64)
        _lib.check(_lib.load().th_ensemble_totals(self._h, _lib.ptr(counts, _lib.i64p), _lib.ptr(scal, _lib.i64p)))
        return counts, scal

    def read_particles(self, index):
        n = self.sizes[index]
        arrs = [np.empty(n) for _ in range(5)]
        spc = np.empty(n, dtype=np.int8)
        _lib.check(_lib.load().th_ensemble_read_particles(self._h, int(index), *(_lib.ptr(a, _lib.dp) for a in arrs),
                                                          _lib.ptr(spc, _lib.i8p)))
        return (*arrs, spc)


# ---------------------------------------------------------------- workload
@dataclass
class InstanceSpec:
    p: model.Params          # geometry, mask, ion distribution, n_per_batch (per species)
    charge: tuple            # (floor, left wall, right wall) charge density [C/m^3]


def batching_workload(count=256, seed=0, dx=5e-9, particles=(1e4, 1e6), angle_deg=3.0):
    """The heterogeneous batching workload, ``count`` instances: trench width
    uniform in 10-40 cells, aspect ratio uniform in 2-20, mask 0-40 cells
    (conductive, or dielectric with eps_r 3), tapers uniform in -3..+5 degrees
    (redrawn while the trench bottom would be narrower than a third of its
    width), ions DC (one peak) or bimodal (arcsine between two peaks) with
    peaks in 50-300 eV and a polar-angle sigma of ``angle_deg``; traced
    particles per batch (ions + electrons) log-uniform in ``particles``.
    Floor / sidewall charge densities set a charged field."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        w = int(rng.integers(10, 41))
        d = int(round(rng.uniform(2.0, 20.0) * w))
        mask_t = int(rng.integers(0, 41))
        while True:
            t_tr, t_mk = rng.uniform(-3.0, 5.0, 2)
            bottom = w - 2 * d * np.tan(np.radians(t_tr))
            if bottom >= w / 3:
                break
        top = w + 2 * mask_t * np.tan(np.radians(t_mk))
        widest = max(bottom, float(w), top)
        margin = max(12, w)
        nx = int(np.ceil(widest)) + 2 * margin
        nx += nx % 2
        n = int(round(np.exp(rng.uniform(np.log(particles[0]), np.log(particles[1])))))
        n_per = max(n // 2, 1)
        peaks = np.sort(rng.uniform(50.0, 300.0, 2))
        dc = rng.random() < 0.5
        ion = (dict(ion_distribution="mono", ion_energy_eV=float(peaks[1])) if dc else
               dict(ion_distribution="arcsine", ion_energy_lo_eV=float(peaks[0]), ion_energy_hi_eV=float(peaks[1])))
        conductive = rng.random() < 0.5
        p = replace(model.Params(), nx=nx, trench_w=w, trench_d=d, floor_t=10, mask_t=mask_t, n_vac=40 + w, dx=dx,
                    trench_taper_deg=float(t_tr), mask_taper_deg=float(t_mk),
                    mask_type="conductor" if conductive else "dielectric", mask_eps_r=3.0,
                    field_mode="floating", sigma_s=0.0, wall_model="barrier", bias="dc", electron_temp_eV=4.0,
                    n_per_batch=n_per, seed=int(rng.integers(2**31)), ion_angle_deg=float(angle_deg), **ion)
        scale = rng.uniform(0.5, 1.5)
        out.append(InstanceSpec(p=p, charge=(3e7 * scale, 2e7 * scale, -1e7 * scale)))
    return out


def instance_field(spec):
    """Charged frozen field of one instance: floor cells of the trench at
    charge[0], left / right trench-wall surface cells at charge[1] / charge[2]
    [C/m^3]; potential by the instance's Poisson LU (field_mode floating, so a
    conductive mask floats).  Returns (vac, Ex, Ey, phi, wface)."""
    p = spec.p
    geo = model.Geometry(p)
    fields = model.Fields(p, geo)
    nx, ny = geo.solid.shape
    rho = np.zeros(geo.solid.shape)
    surf = geo.surface
    ix, iy = surf // ny, surf % ny
    top_ox = p.floor_t + p.trench_d
    centre = nx / 2
    floor = iy == p.floor_t - 1
    wall = (iy >= p.floor_t) & (iy < top_ox) & ~fields.cond.ravel()[surf]
    rho.flat[surf[floor]] = spec.charge[0]
    rho.flat[surf[wall & (ix + 0.5 < centre)]] = spec.charge[1]
    rho.flat[surf[wall & (ix + 0.5 >= centre)]] = spec.charge[2]
    phi = fields.solver.solve((-rho / EPS0).ravel()).reshape(nx, ny) + fields.external
    ex, ey = model.compute_field(phi, geo.vac, p, fields.volts["top"])
    return geo.vac, ex, ey, phi, fields.wface


def build_ensemble(specs, fields, device=0, barrier=True):
    """An Ensemble of the given instances (fields from ``instance_field``),
    each with device injection of its own distributions."""
    ens = Ensemble(device)
    for spec, (vac, ex, ey, phi, wf) in zip(specs, fields):
        ens.add(vac, ex, ey, (phi, wf) if barrier else None, tapered_of(spec.p), 2 * spec.p.n_per_batch)
    for i, spec in enumerate(specs):
        ens.set_inject(i, spec.p)
    return ens
