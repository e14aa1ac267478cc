"""Charging batch loop with the per-batch hot path on the GPU.

``run(p, log, progress, stop)`` keeps the contract of the reference's
``run()`` (trench_charging_2d.py:1045-1369): same batch order (inject, trace,
deposit, leak, floating conductor, Poisson, field, record, steady test), same
histories, probes, charge monitors, steady criterion and result keys.  The
difference is where the batch runs: after a one-off set-up (LU of the
Poisson operator on the host, the Green's matrix columns by FP64 triangular
solves on the device, the leak operator), every batch
stays on the device -- push, deposit tallies, leak, conductor step,
potential (phi = G Q + phi_ext) and field -- and only a few kilobytes of
diagnostics come back per batch.

Injection: ``"host"`` draws every batch with the reference's PCG64 stream
and draw order (trench_charging_2d.py:458-489) and uploads it; ``"device"``
draws it on the GPU (Philox, same distributions, different stream).
"""
import ctypes
import time
from types import SimpleNamespace
from dataclasses import asdict

import numpy as np

from . import _lib, model, monitors
from .model import E_CHARGE, EPS0, M_E, AMU
from .tracer import HipTracer

STEADY_CRITERION_VERSION = 2


def lu_factors(lu):
    """SuperLU's factors (Pr A Pc = L U) in the form th_lu_solve takes: the
    strictly lower part of the unit-diagonal L and the strictly upper part of
    U in CSR, and U's diagonal."""
    n = lu.shape[0]
    Lc, Uc = lu.L.tocsr(), lu.U.tocsr()
    Lc.sort_indices()
    Uc.sort_indices()
    last, first = Lc.indptr[1:] - 1, Uc.indptr[:-1]
    rows = np.arange(n)
    if (np.any(np.diff(Lc.indptr) < 1) or np.any(np.diff(Uc.indptr) < 1)
            or not np.array_equal(Lc.indices[last], rows) or not np.array_equal(Uc.indices[first], rows)):
        raise RuntimeError("LU factors without a stored diagonal")
    keep_l = np.ones(Lc.nnz, dtype=bool)
    keep_l[last] = False
    keep_u = np.ones(Uc.nnz, dtype=bool)
    keep_u[first] = False
    shift = np.arange(n + 1)
    i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)
    f64 = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    return dict(l_ptr=i32(Lc.indptr - shift), l_idx=i32(Lc.indices[keep_l]), l_val=f64(Lc.data[keep_l]),
                u_ptr=i32(Uc.indptr - shift), u_idx=i32(Uc.indices[keep_u]), u_val=f64(Uc.data[keep_u]),
                u_diag=f64(Uc.data[first]))


def _lu_solve(lu, z, device):
    """Solves L U w = z on the device (th_lu_solve) for a C-ordered block z."""
    n, k = z.shape
    w = np.empty((n, k))
    f = lu_factors(lu)
    ms = ctypes.c_double()
    _lib.check(_lib.load().th_lu_solve(int(device), n, *(_lib.ptr(f[key], _lib.i32p if key.endswith(("ptr", "idx"))
                                                                 else _lib.dp)
                                                         for key in ("l_ptr", "l_idx", "l_val", "u_ptr", "u_idx",
                                                                     "u_val", "u_diag")),
                                       k, _lib.ptr(np.ascontiguousarray(z), _lib.dp), _lib.ptr(w, _lib.dp),
                                       ctypes.byref(ms)))
    return w, ms.value


def device_solve(solver, B, device=0):
    """``solver.solve(B)`` (a model.PoissonSolver) for a block of right-hand
    sides, with the triangular solves in FP64 on the device (th_lu_solve).
    The LU factorisation and the conductor mapping stay on the host.  Returns
    (solution, kernel ms)."""
    lu = solver.lu
    r = np.asarray(solver.reduce(B), dtype=np.float64)
    if r.ndim != 2:
        raise ValueError("device_solve takes a block of right-hand sides")
    z = np.empty(r.shape)
    z[lu.perm_r] = r                      # Pr r  (SuperLU: Pr A Pc = L U)
    w, ms = _lu_solve(lu, z, device)
    return solver.expand(w[lu.perm_c]), ms   # Pc w


def unit_source_solve(solver, cells, value, device=0):
    """The potentials of unit sources, one column per source: column k solves
    the Poisson system whose right-hand side is ``value`` at cell ``cells[k]``
    and zero elsewhere (solver.solve of that block), with the triangular
    solves on the device.  The right-hand side is built already permuted and
    padded to whole waves, and the potentials are read only where asked:
    returns (potential, kernel ms), where potential(cells_out) is the
    [len(cells_out), k] array and potential(cells_out, j) its column j."""
    lu = solver.lu
    unknown = solver.unknown_of()
    n, k = lu.shape[0], len(cells)
    src = unknown[np.asarray(cells)]
    if np.any(src < 0):
        raise ValueError("a unit source on a driven conductor's cell")
    z = np.zeros((n, -(-k // 64) * 64))
    z[lu.perm_r[src], np.arange(k)] = value
    w, ms = _lu_solve(lu, z, device)

    def potential(cells_out, column=None):
        u = unknown[np.asarray(cells_out)]
        at = lu.perm_c[np.maximum(u, 0)]
        out = w[at, :k] if column is None else w[at, column]
        out[u < 0] = 0.0
        return out

    return potential, ms


class ChargingSetup:
    """Once per geometry: geometry, LU (host), Green's matrix on the rows the
    push reads (columns solved on ``device``), leak operator,
    floating-conductor response, monitors."""

    def __init__(self, p, circuit=None, device=0):
        if p.wall_model not in model.WALL_MODELS:
            raise ValueError(f"wall_model must be one of {model.WALL_MODELS}: {p.wall_model!r}")
        self.p = p
        self.geo = geo = model.Geometry(p)
        self.fields = fields = model.Fields(p, geo, circuit)
        N = geo.solid.size
        surf = geo.surface
        m = surf.size
        cond_flat = fields.cond.ravel()
        self.cond_s = cond_flat[surf].astype(np.uint8)
        # One Green's column per distinct surface response: a floating
        # conductor's cells share the conductor's column; a driven conductor's
        # cells carry no charge (:1257-1258) and get none.
        col_of = np.full(m, -1, dtype=np.int32)
        reps = []
        cond_col = -1
        for k in range(m):
            if self.cond_s[k]:
                if fields.floating:
                    if cond_col < 0:
                        cond_col = len(reps)
                        reps.append(k)
                    col_of[k] = cond_col
            else:
                col_of[k] = len(reps)
                reps.append(k)
        self.col_of = col_of
        self.cond_col = cond_col
        # 1 C/m per column (:795-797)
        green_at, self.green_solve_ms = unit_source_solve(fields.solver, surf[reps], -1.0 / (p.dx ** 2 * EPS0),
                                                          device)
        # rows the push reads: vacuum cells (field) and surface cells (barrier)
        keep = geo.vac.ravel().copy()
        keep[surf] = True
        self.rows = np.flatnonzero(keep).astype(np.int32)
        self.green = np.ascontiguousarray(green_at(self.rows))
        # surface response P (m x m) for the leak operator (:795-798)
        P = np.zeros((m, m))
        has = col_of >= 0
        P[:, has] = green_at(surf)[:, col_of[has]]
        self.surface_response = P      # surface potentials per unit surface charge [V per C/m]
        self.leak = self._leakage(P) if p.sigma_s > 0 else None
        self.g_cond = green_at(np.arange(N), cond_col).reshape(geo.solid.shape) if fields.floating else None
        if fields.floating:
            self.C_cond = 1.0 / self.g_cond[fields.cond][0]                      # :1089
            c_cell = np.flatnonzero(geo.surface_mask() & fields.cond)[0]        # :1090
            self.c_idx = int(geo.smap[c_cell])
        else:
            self.C_cond, self.c_idx = 0.0, -1
        self.q_macro = E_CHARGE * (p.flux * (geo.nx * p.dx) * p.dt_batch / p.n_per_batch)   # :1093-1094
        self.probes = monitors.probes(geo, p, fields.cond)
        self.regions = monitors.charge_regions(geo, p)
        self.b0, self.b1 = geo.opening_span(p.floor_t)
        self.e0, self.e1 = geo.opening_span(p.floor_t + p.trench_d + p.mask_t - 1)
        bottom = np.arange(self.b0, self.b1) * geo.ny + (p.floor_t - 1)
        self.bottom_s = geo.smap[bottom]
        if (self.bottom_s < 0).any():
            raise RuntimeError("trench floor cells are expected to be surface cells")

    def _leakage(self, P):
        """Surface leak (:768-803): conductance graph of 4-connected surface cells
        (g = sigma_s/dx, x periodic), backward-Euler operator M = inv(I - dt G P)."""
        p, geo = self.p, self.geo
        m = geo.surface.size
        smap = geo.smap.reshape(geo.solid.shape)
        g = p.sigma_s / p.dx
        G = np.zeros((m, m))
        for a, b in ((smap, np.roll(smap, -1, axis=0)), (smap[:, :-1], smap[:, 1:])):
            sel = (a >= 0) & (b >= 0)
            a, b = a[sel], b[sel]
            np.add.at(G, (a, b), g)
            np.add.at(G, (b, a), g)
            np.add.at(G, (a, a), -g)
            np.add.at(G, (b, b), -g)
        M = np.linalg.inv(np.eye(m) - p.dt_batch * (G @ P))
        iy = geo.surface % geo.ny
        cut = p.floor_t + p.trench_d // 2
        region = geo.inner_surface_mask().ravel()[geo.surface] & (iy < cut)
        ext = self.fields.external.ravel()
        c = p.dt_batch * (G @ ext[geo.surface]) if np.any(ext) else None
        return dict(G=G, M=M, region=region, c=c)


def ion_spec(p):
    """Ion energy / angle distribution of a parameter set for device injection:
    the reference's dc injection (energy ion_energy_eV plus a lateral thermal
    normal of ion_temp_eV), or a polar-angle Gaussian with a monoenergetic or
    arcsine-shaped IED (``ion_distribution`` "mono" / "arcsine")."""
    kind = getattr(p, "ion_distribution", "thermal")
    e_ion = p.ion_energy_eV
    if p.field_mode in model.CIRCUIT_FIELDS:
        # the circuit fields accelerate the ions inside the domain: they enter at
        # the Bohm energy Te/2 (trench_charging_2d.py:1131-1136)
        kind, e_ion = "thermal", p.electron_temp_eV / 2
    m_i = p.ion_mass_amu * AMU
    spec = _lib.InjectParams(qm_i=E_CHARGE / m_i, qm_e=-E_CHARGE / M_E, e_ion=float(e_ion),
                        e_lo=float(getattr(p, "ion_energy_lo_eV", 0.0)),
                        e_hi=float(getattr(p, "ion_energy_hi_eV", 0.0)),
                        sigma_theta=float(np.radians(getattr(p, "ion_angle_deg", 0.0))),
                        vth_i=float(np.sqrt(p.ion_temp_eV * E_CHARGE / m_i)),
                        v_per_sqrt_ev=float(np.sqrt(2.0 * E_CHARGE / m_i)),
                        s_e=float(np.sqrt(p.electron_temp_eV * E_CHARGE / M_E)))
    spec.ion_kind = {"thermal": 0, "mono": 1, "arcsine": 2}[kind]
    return spec


class DeviceCharging:
    """The device-resident charging state of one geometry (th_charging)."""

    def __init__(self, setup, tracer=None, device=0):
        self.setup = s = setup
        p = s.p
        self.tracer = tracer or HipTracer(s.geo.vac, device)
        flat = self.tracer.surface_cells()
        if not np.array_equal(flat, s.geo.surface):
            raise RuntimeError("tracer and host disagree on the surface cells")
        f64 = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        self._keep = dict(col_of=np.ascontiguousarray(s.col_of, np.int32), rows=s.rows, green=s.green,
                          external=f64(s.fields.external.ravel()), wface=f64(s.fields.wface.ravel()),
                          leak_M=f64(s.leak["M"]) if s.leak else None,
                          leak_c=f64(s.leak["c"]) if s.leak and s.leak["c"] is not None else None,
                          region=np.ascontiguousarray(s.leak["region"], np.uint8) if s.leak else None,
                          cond_s=np.ascontiguousarray(s.cond_s, np.uint8),
                          g_cond_s=f64(s.g_cond.ravel()[s.geo.surface]) if s.g_cond is not None else None)
        k = self._keep
        su = _lib.ChargingSetupC(n_cols=int(s.green.shape[1]), n_rows=int(s.rows.size),
                    col_of=_lib.ptr(k["col_of"], _lib.i32p), rows=_lib.ptr(k["rows"], _lib.i32p),
                    green=_lib.ptr(k["green"], _lib.dp), external=_lib.ptr(k["external"], _lib.dp),
                    wface=_lib.ptr(k["wface"], _lib.dp), leak_M=_lib.ptr(k["leak_M"], _lib.dp),
                    leak_c=_lib.ptr(k["leak_c"], _lib.dp), region=_lib.ptr(k["region"], _lib.u8p),
                    cond_s=_lib.ptr(k["cond_s"], _lib.u8p), g_cond_s=_lib.ptr(k["g_cond_s"], _lib.dp),
                    q_macro=s.q_macro, dx=p.dx, C_cond=s.C_cond, Te=p.electron_temp_eV, dt_batch=p.dt_batch,
                    top_voltage=s.fields.volts["top"], substrate_voltage=p.substrate_voltage, c_idx=s.c_idx,
                    floating=int(s.fields.floating), driven=int(s.fields.driven and s.fields.cond.any()),
                    top_dirichlet=int(p.top_bc == "dirichlet"))
        L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(L.th_charging_create(self.tracer._h, ctypes.byref(su), ctypes.byref(h)))
        self._h = h
        self.m = s.geo.surface.size
        self.N = s.geo.solid.size
        self.prm = self.tracer.params_for(p, p.wall_model == "barrier", rf=s.fields.rf is not None)
        if s.fields.rf is not None:
            self.tracer.set_rf(s.fields.rf)
        self.capacity = 0

    def set_schedule(self, **kw):
        """Push options of the device batches: precision (0 FP64, 1 FP32) and the
        schedule fields of trench01_amd.autotune (order, refill, check_steps,
        prio_steps)."""
        for k, v in kw.items():
            if not hasattr(self.prm, k):
                raise KeyError(k)
            setattr(self.prm, k, int(v))

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().th_charging_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_green_precision(self, mode):
        """Potential update in FP64 (0) or with the Green's matrix split into
        BF16 (1) or scaled FP16 (2) halves, FP32 products and sums; 3 / 4:
        the same splits on matrix cores (th_charging_set_green)."""
        _lib.check(_lib.load().th_charging_set_green(self._h, int(mode)))
        self.green_precision = int(mode)

    def potentials(self, Q):
        """G Q for 1 to 16 charge vectors Q [R][n_cols] (C/m per Green's
        column) in one pass over G, on matrix cores; needs split precision 3 or
        4 (th_charging_potentials).  Returns ([R][n_rows] potentials without the
        external part, kernel ms)."""
        Q = np.ascontiguousarray(np.atleast_2d(Q), dtype=np.float64)
        rows, cols = self.setup.green.shape
        if Q.ndim != 2 or Q.shape[1] != cols or not 1 <= Q.shape[0] <= 16:
            raise ValueError(f"charge vectors must be [1..16][{cols}]")
        out = np.empty((Q.shape[0], rows))
        ms = ctypes.c_double()
        _lib.check(_lib.load().th_charging_potentials(self._h, Q.shape[0], _lib.ptr(Q, _lib.dp),
                                                      _lib.ptr(out, _lib.dp), ctypes.byref(ms)))
        return out, ms.value

    def set_rho(self, rho_s):
        rho_s = np.ascontiguousarray(rho_s, dtype=np.float64)
        _lib.check(_lib.load().th_charging_set_rho(self._h, _lib.ptr(rho_s, _lib.dp)))

    def reserve(self, n):
        if n > self.capacity:
            self.tracer.reserve(n)
            self.capacity = n

    def potential_adjoint(self, w):
        """G^T w: the gradient of sum(w * phi) over the grid with respect to the
        charge of each Green's column [V per C/m] (th_charging_green_adjoint)."""
        w = np.ascontiguousarray(w, dtype=np.float64).ravel()
        if w.size != self.setup.geo.solid.size:
            raise ValueError("weights must cover the grid")
        out = np.empty(self.setup.green.shape[1])
        _lib.check(_lib.load().th_charging_green_adjoint(self._h, _lib.ptr(w, _lib.dp), _lib.ptr(out, _lib.dp)))
        return out

    def inject_device(self, offset, n_per_species, seed, batch):
        p = self.setup.p
        spec = ion_spec(p)
        spec.n = int(n_per_species)
        spec.lx = p.nx * p.dx
        spec.y0 = self.setup.geo.ny * p.dx - 1e-3 * p.dx
        _lib.check(_lib.load().th_inject(self.tracer._h, int(offset), ctypes.byref(spec), int(seed) & (2**64 - 1),
                                         int(batch)))

    def run_batches(self, offset, n_per_species, seed, first_batch, batches):
        """``batches`` device-injected batches replayed from one captured
        hipGraph (th_charging_run).  Returns a list of per-batch records
        (probe potentials dict, region charges dict incl. 'surface absolute',
        floor ion count, floor electron count, tally dict, diag dict) and the
        GPU time [ms]."""
        p = self.setup.p
        spec = ion_spec(p)
        spec.n = int(n_per_species)
        spec.lx = p.nx * p.dx
        spec.y0 = self.setup.geo.ny * p.dx - 1e-3 * p.dx
        L = _lib.load()
        size = ctypes.c_int32()
        _lib.check(L.th_charging_row_size(self._h, ctypes.byref(size)))
        out = []
        total_ms = 0.0
        done = 0
        while done < batches:
            k = min(1024, batches - done)
            rows = np.empty((k, size.value))
            ms = ctypes.c_double()
            _lib.check(L.th_charging_run(self._h, int(offset), ctypes.byref(spec), int(seed) & (2**64 - 1),
                                         int(first_batch + done), k, ctypes.byref(self.prm), _lib.ptr(rows, _lib.dp),
                                         ctypes.byref(ms)))
            total_ms += ms.value
            npb = len(getattr(self, "monitor_probes", []))
            nr = len(getattr(self, "monitor_regions", []))
            nm = npb + nr + 3 if npb or nr else 0
            for r in rows:
                probes = dict(zip(self.monitor_probes, r[:npb].tolist())) if nm else {}
                charges = dict(zip(self.monitor_regions, r[npb:npb + nr].tolist())) if nm else {}
                if nm:
                    charges["surface absolute"] = float(r[npb + nr])
                t = r[nm:nm + 5]
                tally = dict(escaped=(int(t[0]), int(t[1])), lost=int(t[2]), steps=int(t[3]), off_surface=int(t[4]))
                dg = r[nm + 5:nm + 10]
                diag = dict(leak_dq=float(dg[0]), drift=float(dg[1]), dv=float(dg[2]), n_i_cond=float(dg[3]),
                            n_e_cond=float(dg[4]))
                out.append((probes, charges, float(r[npb + nr + 1]) if nm else 0.0,
                            float(r[npb + nr + 2]) if nm else 0.0, tally, diag))
            done += k
        return out, total_ms

    def step(self, offset, n):
        d = _lib.BatchDiag()
        _lib.check(_lib.load().th_charging_step(self._h, int(offset), int(n), ctypes.byref(self.prm),
                                                ctypes.byref(d)))
        return d

    def set_q_macro(self, q_macro):
        """Charge per macro-particle of the following batches [C/m]."""
        _lib.check(_lib.load().th_charging_set_q_macro(self._h, float(q_macro)))

    def step_replicas(self, offset, n, replicas):
        """``replicas`` batches from the current charge in one push (replica r's
        particles at pool[offset + r n, offset + (r+1) n)), each updating its
        own copy of the charge (th_charging_step_replicas).  Returns (rho
        [replicas][m], per-replica {escaped ions, escaped electrons, lost,
        off-surface} [replicas][4], push totals); the device state is left
        at the current charge."""
        rho = np.empty((replicas, self.m))
        scal = np.empty((replicas, 4), np.int64)
        d = _lib.BatchDiag()
        _lib.check(_lib.load().th_charging_step_replicas(self._h, int(offset), int(n), int(replicas),
                                                         ctypes.byref(self.prm), _lib.ptr(rho, _lib.dp),
                                                         _lib.ptr(scal, _lib.i64p), ctypes.byref(d)))
        return rho, scal, d

    def set_monitors(self, probes, regions, abs_key="surface"):
        """Per-batch device monitors: potential at the probe cells, charge per
        region (regions: name -> flat cells), |charge| of ``abs_key``, and the
        counts on the trench floor."""
        geo = self.setup.geo
        self.monitor_probes = list(probes)
        self.monitor_regions = list(regions)
        cells = np.array([ix * geo.ny + iy for ix, iy in probes.values()], dtype=np.int32)
        idx, ptr = [], [0]
        for key in self.monitor_regions:
            s_idx = geo.smap[np.asarray(regions[key], dtype=np.int64)]
            s_idx = s_idx[s_idx >= 0]          # charge only ever sits on surface cells
            idx.append(s_idx.astype(np.int32))
            ptr.append(ptr[-1] + s_idx.size)
        idx = np.concatenate(idx) if idx else np.zeros(0, np.int32)
        ptr = np.asarray(ptr, dtype=np.int32)
        bottom = np.ascontiguousarray(self.setup.bottom_s, dtype=np.int32)
        self._mon_keep = (cells, ptr, idx, bottom)
        _lib.check(_lib.load().th_charging_set_monitors(
            self._h, cells.size, _lib.ptr(cells, _lib.i32p), len(self.monitor_regions), _lib.ptr(ptr, _lib.i32p),
            _lib.ptr(idx, _lib.i32p), self.monitor_regions.index(abs_key), bottom.size, _lib.ptr(bottom, _lib.i32p)))
        self._mon_out = np.empty(cells.size + len(self.monitor_regions) + 3)

    def monitors(self):
        """The last batch's monitors: (probe potentials dict, region charges dict
        incl. 'surface absolute', floor ion count, floor electron count)."""
        out = self._mon_out
        _lib.check(_lib.load().th_charging_monitors(self._h, _lib.ptr(out, _lib.dp)))
        npb, nr = len(self.monitor_probes), len(self.monitor_regions)
        probes = dict(zip(self.monitor_probes, out[:npb].tolist()))
        charges = dict(zip(self.monitor_regions, out[npb:npb + nr].tolist()))
        charges["surface absolute"] = float(out[npb + nr])
        return probes, charges, float(out[npb + nr + 1]), float(out[npb + nr + 2])

    def read(self, phi=True, counts=True):
        rho = np.empty(self.m)
        ph = np.empty(self.N) if phi else None
        cn = np.empty(2 * self.m, np.int32) if counts else None
        _lib.check(_lib.load().th_charging_read(self._h, _lib.ptr(rho, _lib.dp), _lib.ptr(ph, _lib.dp),
                                                _lib.ptr(cn, _lib.i32p)))
        return rho, ph, cn


# Settings a resumed run may change (trench_charging_2d.py:952-954).
RESUME_EDITABLE = {"until_steady", "n_batches", "max_batches", "steady_window", "steady_rtol",
                   "steady_charge_rtol", "steady_charge_atol", "steady_hold", "backend", "cuda_device",
                   "probe_ions", "probe_energy_eV", "probe_trajectories"}
# Settings of the steady criterion: a resumed run keeps its count of passed
# checks only if none of them changed (trench_charging_2d.py:955-956, :1203-1207).
STEADY_SETTINGS = ("until_steady", "steady_window", "steady_rtol", "steady_charge_rtol", "steady_charge_atol",
                   "steady_hold")


def resume_params(res, p):
    """The stored run's parameters with the editable settings taken from ``p``;
    any other difference is an error (trench_charging_2d.py:959-973)."""
    from dataclasses import fields as dc_fields, replace
    names = {f.name for f in dc_fields(model.Params)}
    source = model.Params(**{k: v for k, v in res["params"].items() if k in names})
    # the display phase of an RF-cycle run changes only what is shown (:963)
    editable = RESUME_EDITABLE | ({"rf_phase"} if source.field_mode == "circuit_rf" else set())
    changed = []
    for k, v in asdict(source).items():
        entered = getattr(p, k, v)
        equal = np.isclose(entered, v, rtol=1e-9, atol=0) if isinstance(v, float) else entered == v
        if k not in editable and not equal and not (k == "backend"):
            changed.append(k)
    if changed:
        raise ValueError("a resumed run must keep the geometry and physics of the stored one: " + ", ".join(changed))
    return replace(source, **{k: getattr(p, k) for k in editable if hasattr(p, k)})


def resumed_steady_count(resume, p, start_batch, history_start, hist, probes, charge_monitor_hist,
                         steady_reference=None):
    """Consecutive passed steady checks a resumed run starts from
    (trench_charging_2d.py:1200-1222): the stored count if the stored run used
    the same criterion version and steady settings and its convergence was not
    revoked on load; else, for a result without a checkpoint, the passes
    re-counted backwards over its histories; otherwise 0."""
    source = resume["params"]
    same = (resume.get("steady_criterion_version") == STEADY_CRITERION_VERSION
            and not resume.get("convergence_note")
            and all(source.get(k) == getattr(p, k) for k in STEADY_SETTINGS))
    if not same:
        return 0
    checkpoint = resume.get("checkpoint")
    if checkpoint:
        return int(checkpoint["steady_n_ok"])
    n_ok, W = 0, p.steady_window
    for size in range(start_batch - start_batch % W, history_start + 2 * W - 1, -W):
        h = {k: np.asarray(v[:size], dtype=float) for k, v in hist.items()}
        q = {k: np.asarray(v[:size], dtype=float) for k, v in charge_monitor_hist.items()}
        if not (monitors.check_steady(h, probes, p, steady_reference)[0] and monitors.check_charge_steady(q, p)["ok"]):
            break
        n_ok += 1
        if n_ok >= p.steady_hold:
            break
    return n_ok


def run(p, log=print, progress=None, stop=None, injection="host", ion_energy=None, ion_velocity=None,
        device=0, resume=None, device_monitors=True, circuit=None, graph=True):
    """Charging run (trench_charging_2d.py:1045-1369) with the batch on the GPU.
    ``resume``: a result of ``run`` or ``results.load_results``; its charge,
    histories and random state are continued (fixed length: n_batches more;
    until_steady: max_batches in total).  ``circuit``: the RF circuit waveform
    for field_mode "circuit" / "circuit_rf" (the reference's
    PC.solve_circuit_cached output: t, uw, um, up, V1 over one period); with
    "circuit_rf" every batch draws stratified launch phases and the push
    follows the voltage wave in flight.  ``graph``: device-injected runs
    without a progress callback replay a captured batch up to each steady
    check (same results; ``stop`` is then polled between such blocks and
    timing.push_ms holds the whole batches' GPU time).  Returns the reference's result dict
    (plus ``timing``), or None if stopped before the first batch."""
    log = log or (lambda s: None)
    start_batch = 0
    if resume is not None:
        p = resume_params(resume, p)
        start_batch = len(resume["hist"]["t"])
        if p.until_steady and p.max_batches <= start_batch:
            raise ValueError(f"max_batches counts all batches; raise it above the {start_batch} already run")
        if not p.until_steady and p.n_batches < 1:
            raise ValueError("a resumed fixed-length run needs n_batches >= 1")
    if injection not in ("host", "device"):
        raise ValueError("injection must be 'host' or 'device'")
    if p.until_steady and (p.steady_window < 5 or p.steady_hold < 1 or p.steady_rtol <= 0
                           or not np.isfinite(p.steady_charge_rtol) or p.steady_charge_rtol <= 0
                           or not np.isfinite(p.steady_charge_atol) or p.steady_charge_atol < 0):
        raise ValueError("steady_window >= 5, steady_hold >= 1, positive tolerances are required")
    if p.field_mode in model.CIRCUIT_FIELDS:
        if p.bias != "rf":
            raise ValueError("circuit field modes need bias=rf")
        if circuit is None and resume is not None:
            circuit = resume.get("circuit")
        if circuit is None:
            raise ValueError("circuit field modes need the circuit waveform (circuit=...)")
        # ions enter at the Bohm energy and are accelerated inside the domain (:1131-1136)
        ion_velocity = None
        ion_energy = lambda rng_, n_: np.full(n_, p.electron_temp_eV / 2)
    t_setup = time.perf_counter()
    s = ChargingSetup(p, circuit, device)
    geo, fields = s.geo, s.fields
    nx, ny = geo.solid.shape
    dev = DeviceCharging(s, device=device)
    log(f"particle tracing: HIP ({dev.tracer.name}, device {device}, float64); grid {nx} x {ny}, "
        f"{geo.surface.size} surface cells, {s.green.shape[1]} Green's columns on {s.rows.size} rows")
    m = geo.surface.size
    rng = np.random.default_rng(p.seed)
    # stratified RF launch phases (trench_charging_2d.py:1079, :1231-1233)
    phase_rng = np.random.default_rng(np.random.SeedSequence([p.seed, 0x5246])) if fields.rf is not None else None
    rho_full = np.zeros(geo.solid.size)
    n_ok = 0
    if resume is not None:
        if not np.array_equal(resume["solid"], geo.solid) or not np.array_equal(resume["mask"], geo.mask):
            raise ValueError("the stored geometry does not match the parameters")
        rho_full[:] = np.asarray(resume["rho_raw"], dtype=float).ravel()
        off = np.setdiff1d(np.flatnonzero(rho_full), geo.surface)
        if off.size:
            raise ValueError("stored charge lies outside the surface cells")
        checkpoint = resume.get("checkpoint") or {}
        if injection == "host":
            if not checkpoint.get("rng_state"):
                raise ValueError("host injection needs the stored random state to resume")
            rng.bit_generator.state = checkpoint["rng_state"]
        if phase_rng is not None:
            if not checkpoint.get("phase_rng_state"):
                raise ValueError("an RF-cycle run needs the stored phase random state to resume")
            phase_rng.bit_generator.state = checkpoint["phase_rng_state"]
    dev.set_rho(rho_full[geo.surface])
    n = 2 * p.n_per_batch
    dev.reserve(n)
    t_setup = time.perf_counter() - t_setup

    probes = s.probes
    # with the RF cycle the criterion scales by the potential relative to the
    # period-mean external potential (:1167-1168)
    steady_reference = ({k: float(fields.external[ix, iy]) for k, (ix, iy) in probes.items()}
                        if fields.rf is not None else None)
    hist = {k: [] for k in ("t", "ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost", "leak", *probes)}
    charge_hist = {"left": [], "right": []}
    charge_monitor_hist = {k: [] for k in (*s.regions, "surface absolute")}
    if resume is not None:
        # Histories a stored result lacks are NaN, except their last sample,
        # which is recomputed from the stored charge (trench_charging_2d.py:1175-1197).
        # A stored charge-monitor history keeps its own last sample: it is the
        # same integral of the same charge, summed on the device in the run
        # that wrote it, so the resumed run continues it bit for bit.
        phi0 = fields.solve((-rho_full / EPS0)) + fields.external.ravel()
        initial_charge = monitors.sample_charges(rho_full, s.regions, p.dx)
        for key in hist:
            hist[key] = list(np.asarray(resume["hist"].get(key, np.full(start_batch, np.nan)), dtype=float))
            if key in probes and key not in resume["hist"]:
                ix, iy = probes[key]
                hist[key][-1] = float(phi0[ix * ny + iy])
        stored_charge = resume.get("charge_hist") or {}
        for key in charge_hist:
            charge_hist[key] = list(np.asarray(stored_charge.get(key, np.full(start_batch, np.nan)), dtype=float))
            if key not in stored_charge:
                charge_hist[key][-1] = initial_charge[key]
        stored_monitors = resume.get("charge_monitor_hist") or {}
        for key in charge_monitor_hist:
            saved = stored_monitors.get(key)
            charge_monitor_hist[key] = list(np.asarray(
                saved if saved is not None else charge_hist[key] if key in charge_hist else np.full(start_batch, np.nan),
                dtype=float))
            if saved is None:
                charge_monitor_hist[key][-1] = initial_charge[key]
    frac_open = (s.b1 - s.b0) / nx
    if device_monitors:
        dev.set_monitors(probes, s.regions)
    n_max = p.max_batches if p.until_steady else start_batch + p.n_batches
    converged_at, stopped = None, False
    history_start = monitors.history_start(hist, probes, charge_monitor_hist)
    if resume is not None and p.until_steady:
        n_ok = resumed_steady_count(resume, p, start_batch, history_start, hist, probes, charge_monitor_hist,
                                    steady_reference)
    push_ms = update_ms = 0.0
    phi = fields.external.ravel().copy()
    t0 = time.time()
    # Device-injected runs without per-batch host work replay a captured batch
    # (th_charging_run) up to the next steady check, then read the records.
    replay = graph and injection == "device" and device_monitors and phase_rng is None and progress is None
    pending = []
    for b in range(start_batch, n_max):
        if stop is not None and stop.is_set():
            stopped = True
            break
        if replay:
            if not pending:
                k = n_max - b
                if p.until_steady:
                    k = min(k, p.steady_window - b % p.steady_window)
                records, ms = dev.run_batches(0, p.n_per_batch, p.seed, b, k)
                push_ms += ms
                pending = records
            values, charges, n_i_floor, n_e_floor, tally, dg = pending.pop(0)
            d = SimpleNamespace(tally=SimpleNamespace(escaped=tally["escaped"], lost=tally["lost"],
                                                      off_surface=tally["off_surface"]), leak_dq=dg["leak_dq"])
        else:
            if injection == "host":
                dev.tracer.upload(0, *model.inject(p, p.n_per_batch, ny, rng, ion_energy, ion_velocity))
            else:
                dev.inject_device(0, p.n_per_batch, p.seed, b)
            if phase_rng is not None:
                dev.tracer.upload_times(0, np.concatenate(
                    [model.launch_times(phase_rng, p.n_per_batch, fields.rf["period"]) for _ in range(2)]))
            d = dev.step(0, n)
            push_ms += d.push_ms
            update_ms += d.update_ms
        if d.tally.off_surface:
            raise RuntimeError(f"{d.tally.off_surface} particles were absorbed off the surface cells")
        hist["t"].append((b + 1) * p.dt_batch)
        norm = p.n_per_batch * frac_open
        if replay:
            for k in probes:
                hist[k].append(values[k])
        elif device_monitors:
            values, charges, n_i_floor, n_e_floor = dev.monitors()
            for k in probes:
                hist[k].append(values[k])
        else:
            rho_s, phi, counts = dev.read()
            rho_full[:] = 0.0
            rho_full[geo.surface] = rho_s
            for k, (ix, iy) in probes.items():
                hist[k].append(phi[ix * ny + iy])
            n_i_floor, n_e_floor = counts[s.bottom_s].sum(), counts[m + s.bottom_s].sum()
            charges = monitors.sample_charges(rho_full, s.regions, p.dx)
        hist["ratio_i_bottom"].append(n_i_floor / norm)
        hist["ratio_e_bottom"].append(n_e_floor / norm)
        hist["esc_e"].append(d.tally.escaped[1] / p.n_per_batch)
        hist["lost"].append(d.tally.lost / (2 * p.n_per_batch))
        hist["leak"].append(d.leak_dq / p.dt_batch / (E_CHARGE * p.flux * (s.e1 - s.e0) * p.dx)
                            if s.leak is not None else 0.0)
        for side in ("left", "right"):
            charge_hist[side].append(charges[side])
        for key, q in charges.items():
            charge_monitor_hist[key].append(q)
        if (b + 1) % 10 == 0 or b == 0:
            log(f"[{b + 1:4d}/{n_max}] t={hist['t'][-1] * 1e6:7.1f} us  "
                f"phi(bottom)={hist['bottom center'][-1]:7.2f} V  "
                f"phi(side mid)={hist['sidewall middle'][-1]:7.2f} V  lost={hist['lost'][-1] * 100:.2f}%  "
                f"({time.time() - t0:.1f}s)")
        W = p.steady_window
        check_info = None
        if p.until_steady and (b + 1) % W == 0 and (b + 1) - history_start >= 2 * W:
            v_ok, name, dvolt, tol = monitors.check_steady(hist, probes, p, steady_reference)
            charge = monitors.check_charge_steady(charge_monitor_hist, p)
            ok = bool(v_ok and charge["ok"])
            n_ok = min(n_ok + 1, p.steady_hold) if ok else 0
            voltage = dict(ok=bool(v_ok), name=name, d=float(dvolt), tol=float(tol), unit="V", kind="voltage")
            check_info = {**(charge if not charge["ok"] else voltage), "ok": ok, "n_ok": n_ok,
                          "voltage": voltage, "charge": charge}
            log(f"  [steady] voltage {'OK' if v_ok else 'NG'} ({name} {dvolt:+.3g} V, tol {tol:.3g}); "
                f"charge {'OK' if charge['ok'] else 'NG'} -> {n_ok}/{p.steady_hold}")
            if n_ok >= p.steady_hold:
                converged_at = b + 1
        if progress is not None:
            if device_monitors:
                rho_s, phi, _ = dev.read(counts=False)
                rho_full[:] = 0.0
                rho_full[geo.surface] = rho_s
            progress(dict(batch=b + 1, start_batch=start_batch, n_max=n_max, elapsed=time.time() - t0,
                          phi=phi.reshape(nx, ny).copy(), rho=rho_full.reshape(nx, ny).copy(),
                          hist={k: np.array(v) for k, v in hist.items()},
                          charge_hist={k: np.array(v) for k, v in charge_hist.items()},
                          check=check_info, converged=converged_at is not None, probes=list(probes)))
        if converged_at is not None:
            break
    if len(hist["t"]) == start_batch:
        return None
    wall = time.time() - t0
    rho_s, phi, _ = dev.read(counts=False)
    rho_full[:] = 0.0
    rho_full[geo.surface] = rho_s
    hist = {k: np.array(v) for k, v in hist.items()}
    Wl = min(p.steady_window, len(hist["t"]))
    steady = {k: float(hist[k][-Wl:].mean()) for k in probes}
    rho = rho_full.reshape(nx, ny).copy()
    # Final potential and field on the whole grid with the LU (the device keeps
    # only the rows the push reads); with the RF cycle the fields are shown at
    # p.rf_phase around the period-mean potential (:1353-1358).
    phi_full = fields.solve((-rho / EPS0).ravel()).reshape(nx, ny) + fields.external
    shown_phi, shown_ext, shown_rhs, shown_volts = fields.snapshot(phi_full, p)
    Ex, Ey = model.compute_field(shown_phi, geo.vac, p, shown_volts["top"])
    shown = rho.copy()
    if fields.cond.any():
        induced = (fields.A @ shown_phi.ravel()).reshape(rho.shape) - shown_rhs
        shown[fields.cond] = (-EPS0 * induced)[fields.cond]
    t_conv = converged_at * p.dt_batch if converged_at else np.nan
    nb = len(hist["t"]) - start_batch
    return dict(phi=shown_phi, mean_phi=phi_full, rho=shown, Ex=Ex, Ey=Ey, rho_raw=rho,
                external_phi=shown_ext, voltages=shown_volts,
                charge_hist={k: np.array(v) for k, v in charge_hist.items()},
                charge_monitor_hist={k: np.array(v) for k, v in charge_monitor_hist.items()},
                steady_criterion_version=STEADY_CRITERION_VERSION,
                checkpoint=dict(version=1, batches_completed=len(hist["t"]),
                                rng_state=rng.bit_generator.state if injection == "host" else None,
                                phase_rng_state=phase_rng.bit_generator.state if phase_rng is not None else None,
                                steady_n_ok=n_ok),
                start_batch=start_batch, params=asdict(p), converged=converged_at is not None, backend_used="hip",
                solid=geo.solid, mask=geo.mask, i0=geo.i0, i1=geo.i1, hist=hist, probes=probes, steady=steady,
                t_conv=t_conv, stopped=stopped, circuit=circuit,
                timing=dict(setup_s=t_setup, loop_s=wall, batches=nb, push_ms=push_ms, update_ms=update_ms,
                            device_phi=phi.reshape(nx, ny)))
