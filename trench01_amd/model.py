"""Host-side model of the masked trench: parameters, staircase geometry,
materials, the Poisson operator and its LU, boundary potentials, the field
stencil, and host injection.  These are the set-up pieces the reference runs
once per run around its batch loop (trench_charging_2d.py:103-454, :458-489);
the per-batch work runs on the GPU (``charging``).

Names, arguments and defaults follow the reference so that a reference
``Params`` can be passed straight in.
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sps
from scipy.sparse.linalg import splu

# trench_charging_2d.py:90-93
E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
M_E = 9.1093837015e-31
AMU = 1.66053906660e-27

MASK_TYPES = ("conductor", "dielectric")
WALL_MODELS = ("barrier", "absorb")
FIELD_MODES = ("floating", "fixed", "circuit", "circuit_rf")      # trench_charging_2d.py:97-98
CIRCUIT_FIELDS = ("circuit", "circuit_rf")
SIDE_BCS = ("periodic", "pillar", "fixed")                         # :99


@dataclass
class Params:
    """The charging-path fields of trench_charging_2d.py:103-174, same names and
    defaults; ``backend`` picks this package's HIP tracer."""
    dx: float = 5e-9
    nx: int = 60
    trench_w: int = 20
    trench_offset: int = 0
    trench_taper_deg: float = 0.0
    trench_d: int = 80
    floor_t: int = 10
    n_vac: int = 50
    eps_r: float = 3.9
    top_bc: str = "dirichlet"
    mask_t: int = 60
    mask_taper_deg: float = 0.0
    mask_type: str = "conductor"
    mask_eps_r: float = 3.0
    field_mode: str = "floating"
    rf_phase: float = 0.0
    rf_time_steps: int = 256
    substrate_voltage: float = 0.0
    mask_voltage: float = 0.0
    top_voltage: float = 0.0
    side_bc: str = "periodic"
    left_voltage: float = 0.0
    right_voltage: float = 0.0
    probe_ions: int = 1000
    probe_energy_eV: float = 100.0
    probe_trajectories: int = 40
    flux: float = 1e20
    ion_mass_amu: float = 40.0
    ion_energy_eV: float = 100.0
    ion_temp_eV: float = 0.5
    electron_temp_eV: float = 3.0
    # Ion distribution for device-side injection (no reference field: the
    # reference takes such distributions through its ion_velocity hook):
    # "thermal" (ion_energy_eV + lateral ion_temp_eV, the reference's dc
    # injection), "mono" (ion_energy_eV) or "arcsine" (between the lo / hi
    # energies), the last two with a Gaussian polar angle of ion_angle_deg.
    ion_distribution: str = "thermal"
    ion_angle_deg: float = 0.0
    ion_energy_lo_eV: float = 0.0
    ion_energy_hi_eV: float = 0.0
    bias: str = "dc"
    # RF bias source (the equivalent-circuit inputs of the reference; the
    # circuit itself is solved outside this package and passed as ``circuit``)
    rf_freq: float = 13.56e6
    rf_volt: float = 100.0
    rf_wave: str = "sine"
    rf_duty: float = 0.1
    rf_rise: float = 1e-9
    c_block: float = 4900e-12
    wafer_d: float = 0.3
    wall_ratio: float = 5.0
    ied_model: str = "sheath"
    gas_pressure: float = 1.0
    gas_temp: float = 300.0
    sigma_s: float = 1e-15
    until_steady: bool = False
    max_batches: int = 3000
    steady_window: int = 30
    steady_rtol: float = 0.005
    steady_charge_rtol: float = 0.005
    steady_charge_atol: float = 1e-15
    steady_hold: int = 2
    backend: str = "hip"
    cuda_device: int = 0
    n_batches: int = 300
    n_per_batch: int = 3000
    dt_batch: float = 1e-4
    cfl: float = 0.35
    max_steps: int = 3000
    wall_model: str = "barrier"
    seed: int = 1


def is_tapered(p):
    """trench_charging_2d.py:554."""
    return bool(p.trench_taper_deg or (p.mask_t > 0 and p.mask_taper_deg))


class Geometry:
    """Staircase trench geometry (trench_charging_2d.py:178-248).

    Attributes: ``solid``, ``mask`` (nx, ny) bool; ``i0``, ``i1`` opening at
    the dielectric top; ``vac``; ``surface`` (flat indices of the surface
    cells, ascending) and ``smap`` (cell -> surface index or -1)."""

    def __init__(self, p):
        self.p = p
        self.solid, self.mask, self.i0, self.i1 = self._build(p)
        self.vac = ~self.solid
        self.nx, self.ny = self.solid.shape
        self.surface = np.flatnonzero(self.surface_mask())
        self.smap = np.full(self.solid.size, -1, dtype=np.int64)
        self.smap[self.surface] = np.arange(self.surface.size)

    @staticmethod
    def opening_widths(p):
        """trench_charging_2d.py:178-182."""
        return dict(bottom=p.trench_w - 2 * p.trench_d * np.tan(np.radians(p.trench_taper_deg)),
                    interface=float(p.trench_w),
                    top=p.trench_w + 2 * p.mask_t * np.tan(np.radians(p.mask_taper_deg)))

    @classmethod
    def _build(cls, p):
        """trench_charging_2d.py:185-226: a cell is vacuum when its centre lies in the
        (tapered) opening; the mask spans the top mask_t solid rows."""
        ints = ("nx", "trench_w", "trench_offset", "trench_d", "floor_t", "n_vac", "mask_t")
        if any(not isinstance(getattr(p, k), (int, np.integer)) for k in ints) or p.nx < 3 \
                or not 1 <= p.trench_w <= p.nx - 2 or min(p.floor_t, p.trench_d, p.n_vac) < 1:
            raise ValueError("cell counts must be positive integers and trench_w <= nx-2")
        i0 = (p.nx - p.trench_w) // 2 + p.trench_offset
        i1 = i0 + p.trench_w
        if i0 < 1 or i1 > p.nx - 1:
            raise ValueError("at least one solid column must remain on each side of the opening")
        if p.mask_t < 0:
            raise ValueError(f"mask_t must be >= 0: {p.mask_t}")
        if p.mask_type not in MASK_TYPES:
            raise ValueError(f"mask_type must be one of {MASK_TYPES}: {p.mask_type!r}")
        for key in ("trench_taper_deg", "mask_taper_deg"):
            angle = getattr(p, key)
            if not np.isfinite(angle) or not -90 < angle < 90:
                raise ValueError(f"{key}: taper angle must be finite and inside (-90, 90) degrees")
        centre = (i0 + i1) / 2
        for level, width in cls.opening_widths(p).items():
            if width < 1:
                raise ValueError(f"the opening at the {level} is narrower than one cell")
            if centre - width / 2 < 1 or centre + width / 2 > p.nx - 1:
                raise ValueError(f"the taper at the {level} leaves the domain")
        top_ox = p.floor_t + p.trench_d
        ny = top_ox + p.mask_t + p.n_vac
        solid = np.zeros((p.nx, ny), dtype=bool)
        solid[:, :top_ox + p.mask_t] = True
        rows = np.arange(p.floor_t, top_ox + p.mask_t) + 0.5
        slope = np.tan(np.radians(np.where(rows < top_ox, p.trench_taper_deg, p.mask_taper_deg)))
        half = p.trench_w / 2 + (rows - top_ox) * slope
        xc = (np.arange(p.nx) + 0.5)[:, None]
        opening = (xc >= centre - half) & (xc < centre + half)
        solid[:, p.floor_t:top_ox + p.mask_t] = ~opening
        mask = solid.copy()
        mask[:, :top_ox] = False
        return solid, mask, i0, i1

    def surface_mask(self, vacuum=None):
        """trench_charging_2d.py:734-744: solid cells with a vacuum cell among the
        8 neighbours (x periodic, beyond the y range counts as solid)."""
        vac = self.vac if vacuum is None else vacuum
        padded = np.pad(vac, ((0, 0), (1, 1)), constant_values=False)
        near = np.zeros_like(self.solid)
        for sx in (-1, 0, 1):
            shifted = np.roll(padded, -sx, axis=0)
            for sy in (-1, 0, 1):
                if sx or sy:
                    near |= shifted[:, 1 + sy:1 + sy + self.ny]
        return self.solid & near

    def inner_surface_mask(self):
        """trench_charging_2d.py:747-752."""
        p = self.p
        inner = self.vac.copy()
        inner[:, :p.floor_t] = False
        inner[:, p.floor_t + p.trench_d + p.mask_t:] = False
        return self.surface_mask(inner)

    def opening_span(self, row):
        """trench_charging_2d.py:229-232."""
        cells = np.flatnonzero(self.vac[:, row])
        return int(cells[0]), int(cells[-1] + 1)

    def wall_bounds(self):
        """trench_charging_2d.py:235-241."""
        p = self.p
        iy = np.arange(p.floor_t, p.floor_t + p.trench_d + p.mask_t)
        v = self.vac[:, iy]
        return np.argmax(v, axis=0) - 1, self.nx - np.argmax(v[::-1], axis=0), iy

    def permittivity(self):
        """trench_charging_2d.py:244-248."""
        p = self.p
        eps = np.where(self.solid, p.eps_r, 1.0)
        return np.where(self.mask, p.mask_eps_r, eps) if p.mask_type == "dielectric" else eps


def poisson_matrix(eps, cond, p):
    """trench_charging_2d.py:252-294: div(eps grad phi) on the cell-centred grid.
    Faces take the harmonic mean of the two cell permittivities; a face touching
    a conductor cell couples at twice the other side's permittivity, a face
    between two conductor cells not at all; the substrate row (and the top row
    when top_bc is dirichlet) sees its boundary half a cell away; x is periodic
    except inside the dielectric when side_bc is not periodic."""
    nx, ny = eps.shape
    dx2 = p.dx ** 2

    def coupling(a, b, ca, cb):
        face = np.where(ca, 2.0 * b, np.where(cb, 2.0 * a, 2.0 * a * b / (a + b)))
        return np.where(ca & cb, 0.0, face)

    top_ox = p.floor_t + p.trench_d
    c_east = coupling(eps, np.roll(eps, -1, axis=0), cond, np.roll(cond, -1, axis=0)) / dx2
    if p.side_bc != "periodic":
        c_east[-1, :top_ox] = 0.0
    c_west = np.roll(c_east, 1, axis=0)
    c_vert = coupling(eps[:, :-1], eps[:, 1:], cond[:, :-1], cond[:, 1:]) / dx2
    c_north = np.zeros((nx, ny))
    c_north[:, :-1] = c_vert
    c_south = np.zeros((nx, ny))
    c_south[:, 1:] = c_vert
    d_north, d_south = c_north.copy(), c_south.copy()
    d_south[:, 0] += 2.0 * eps[:, 0] / dx2
    if p.top_bc == "dirichlet":
        d_north[:, -1] += 2.0 * eps[:, -1] / dx2
    diag = -(c_east + c_west + d_north + d_south)
    if p.side_bc != "periodic":
        diag[0, :top_ox] -= 2.0 * eps[0, :top_ox] / dx2
        diag[-1, :top_ox] -= 2.0 * eps[-1, :top_ox] / dx2
    idx = np.arange(nx * ny).reshape(nx, ny)
    rows = [idx.ravel(), idx.ravel(), idx.ravel(), idx[:, :-1].ravel(), idx[:, 1:].ravel()]
    cols = [idx.ravel(), np.roll(idx, -1, axis=0).ravel(), np.roll(idx, 1, axis=0).ravel(),
            idx[:, 1:].ravel(), idx[:, :-1].ravel()]
    vals = [diag.ravel(), c_east.ravel(), c_west.ravel(), c_north[:, :-1].ravel(), c_south[:, 1:].ravel()]
    return sps.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(nx * ny, nx * ny)).tocsc()


class PoissonSolver:
    """LU of the Poisson operator, factored once (trench_charging_2d.py:297-309,
    :376-392).  Floating conductors: all conductor cells share one unknown
    (E^T A E u = E^T b).  Driven conductors: conductor cells are eliminated with
    their potential fixed.  ``solve(b)`` accepts one right-hand side or a block
    of columns."""

    def __init__(self, A, cond, driven):
        self.A = A
        self.cond = cond
        flat = cond.ravel()
        if driven and flat.any():
            self.free = ~flat
            self.lu = splu(A[self.free][:, self.free].tocsc())
            self.E = None
        elif flat.any():
            n_free = int((~flat).sum())
            col = np.where(flat, n_free, np.cumsum(~flat) - 1)
            self.E = sps.csr_matrix((np.ones(flat.size), (np.arange(flat.size), col)),
                                    shape=(flat.size, n_free + 1))
            self.lu = splu((self.E.T @ A @ self.E).tocsc())
            self.free = None
        else:
            self.lu = splu(A)
            self.E = None
            self.free = None

    def solve(self, b):
        return self.expand(self.lu.solve(self.reduce(b)))

    def reduce(self, b):
        """The factored system's right-hand side for ``b`` (one column or a block)."""
        if self.free is not None:
            return b[self.free]
        if self.E is not None:
            return self.E.T @ b
        return b

    def unknown_of(self):
        """Per cell, the unknown of the factored system that holds its potential
        (the shared conductor unknown for a floating conductor's cells), or -1
        for a driven conductor's cells (fixed potential, zero in ``expand``)."""
        n_cells = self.A.shape[0]
        if self.free is not None:
            out = np.full(n_cells, -1, dtype=np.int64)
            out[self.free] = np.arange(int(self.free.sum()))
            return out
        if self.E is not None:
            return self.E.indices.astype(np.int64)      # one entry per row of E
        return np.arange(n_cells, dtype=np.int64)

    def expand(self, y):
        """The potential on every cell from a solution of the factored system."""
        if self.free is not None:
            out = np.zeros((self.free.size,) + y.shape[1:])
            out[self.free] = y
            return out
        if self.E is not None:
            return self.E @ y
        return y


def circuit_voltages(c, phase):
    """trench_charging_2d.py:312-320: electrode potentials of the circuit
    waveform at a phase [deg], periodically interpolated, with the plasma
    potential as 0 V."""
    t = np.asarray(c["t"])
    period = (t[1] - t[0]) * t.size
    at = (phase % 360.0) / 360.0 * period
    values = {k: float(np.interp(at, np.append(t, period), np.append(c[k], c[k][0]))) for k in ("uw", "um", "up", "V1")}
    return dict(substrate=values["uw"] - values["up"], mask=values["um"] - values["up"], top=0.0,
                sheath=values["V1"])


def rf_coefficients(rf, at):
    """trench_rf.py:5-11: the voltage wave (deviation from the period mean,
    [substrate, mask]) at times ``at``, periodic linear interpolation."""
    wave = np.asarray(rf["wave"])
    n_phase = len(wave)
    pos = np.mod(np.asarray(at, dtype=float), rf["period"]) / rf["period"] * n_phase
    node = np.floor(pos)
    k = node.astype(np.int64) % n_phase
    frac = (pos - node)[..., None]
    return wave[k] * (1 - frac) + wave[(k + 1) % n_phase] * frac


def launch_times(rng, n, period):
    """trench_rf.py:14-18: one launch time per stratum of the period, shuffled."""
    times = (np.arange(n) + rng.random(n)) * period / n
    rng.shuffle(times)
    return times


def boundary_rhs(p, eps, volts, fixed_sides=True):
    """trench_charging_2d.py:323-337: right-hand side of the Dirichlet faces
    (substrate, top, and the dielectric's side faces: fixed voltages, or with
    side_bc pillar a linear ramp from the substrate to the mask potential)."""
    rhs = np.zeros_like(eps)
    rhs[:, 0] -= 2.0 * eps[:, 0] / p.dx ** 2 * volts["substrate"]
    if p.top_bc == "dirichlet":
        rhs[:, -1] -= 2.0 * eps[:, -1] / p.dx ** 2 * volts["top"]
    top_ox = p.floor_t + p.trench_d
    if p.side_bc == "pillar":
        yfrac = (np.arange(top_ox) + 0.5) / top_ox
        vl = vr = volts["substrate"] + (volts["mask"] - volts["substrate"]) * yfrac
    else:
        vl, vr = (p.left_voltage, p.right_voltage) if fixed_sides else (0.0, 0.0)
    if p.side_bc != "periodic":
        rhs[0, :top_ox] -= 2.0 * eps[0, :top_ox] / p.dx ** 2 * vl
        rhs[-1, :top_ox] -= 2.0 * eps[-1, :top_ox] / p.dx ** 2 * vr
    return rhs


class Fields:
    """Poisson set-up of one geometry (trench_charging_2d.py:340-413): the LU,
    the potential of the applied boundary voltages (``external``), the wall
    barrier face weights, the conductor bookkeeping, and for field_mode
    "circuit_rf" the RF basis (``rf``: unit substrate / mask responses, the
    voltage wave around its period mean, the step cap dt_max).  The circuit
    modes take the circuit waveform (dict of t, uw, um, up, V1 over one
    period) as ``circuit``."""

    def __init__(self, p, geo, circuit=None):
        if p.field_mode not in FIELD_MODES or p.side_bc not in SIDE_BCS:
            raise ValueError(f"field_mode must be one of {FIELD_MODES} and side_bc one of {SIDE_BCS}")
        if p.top_bc not in ("dirichlet", "neumann"):
            raise ValueError("top_bc must be dirichlet or neumann")
        numeric = (p.rf_phase, p.substrate_voltage, p.mask_voltage, p.top_voltage, p.left_voltage, p.right_voltage)
        if not np.isfinite(numeric).all():
            raise ValueError("potentials and the RF phase must be finite")
        driven = p.field_mode != "floating"
        if driven and geo.mask.any() and p.mask_type != "conductor":
            raise ValueError("a mask voltage needs mask_type=conductor")
        if p.side_bc == "pillar" and not driven:
            raise ValueError("side_bc pillar needs field_mode fixed / circuit / circuit_rf")
        if p.field_mode in CIRCUIT_FIELDS:
            if p.bias != "rf" or circuit is None or p.top_bc != "dirichlet":
                raise ValueError("circuit field modes need bias=rf, the circuit waveform and top_bc=dirichlet")
            volts = circuit_voltages(circuit, p.rf_phase)
        else:
            volts = dict(substrate=p.substrate_voltage, mask=p.mask_voltage, top=p.top_voltage)
        if p.field_mode == "circuit_rf":
            if not isinstance(p.rf_time_steps, (int, np.integer)) or p.rf_time_steps < 16:
                raise ValueError("rf_time_steps must be an integer >= 16")
            t = np.asarray(circuit["t"])
            period = (t[1] - t[0]) * t.size
            if (not np.isfinite(period) or period <= 0 or not np.isclose(t[0], 0, atol=period * 1e-12)
                    or not np.allclose(np.diff(t), period / t.size, rtol=1e-10, atol=period * 1e-12)):
                raise ValueError("the circuit waveform must be one equally spaced period starting at 0")
            if not np.isfinite(p.dt_batch) or p.dt_batch < period * (1 - 1e-12):
                raise ValueError("circuit_rf needs dt_batch of at least one RF period")
            volt_wave = np.column_stack((circuit["uw"] - circuit["up"], circuit["um"] - circuit["up"]))
            mean = volt_wave.mean(axis=0)
            volts = dict(substrate=float(mean[0]), mask=float(mean[1]), top=0.0, sheath=float(np.mean(circuit["V1"])))
        self.volts = volts
        self.cond = geo.mask if p.mask_type == "conductor" else np.zeros_like(geo.mask)
        self.eps = geo.permittivity()
        self.A = poisson_matrix(self.eps, self.cond, p)
        self.solver = PoissonSolver(self.A, self.cond, driven)
        if driven and self.cond.any():
            phi_cond = self.cond.ravel().astype(float) * volts["mask"]
            mask_rhs = -(self.A @ phi_cond)
        else:
            phi_cond = np.zeros(self.cond.size)
            mask_rhs = np.zeros(self.cond.size)
        self.boundary_rhs = boundary_rhs(p, self.eps, volts)
        self.external = (self.solver.solve(self.boundary_rhs.ravel() + mask_rhs) + phi_cond).reshape(self.eps.shape)
        self.floating = bool(self.cond.any() and not driven)
        self.driven = driven
        self.wface = np.where(self.cond, 1.0, self.eps / (1.0 + self.eps))   # trench_charging_2d.py:1085
        self.rf = None
        if p.field_mode == "circuit_rf":
            # the instantaneous field is linear in the electrode voltages (:397-412)
            basis, rhs_basis = [], []
            for vs, vm in ((1.0, 0.0), (0.0, 1.0)):
                brhs = boundary_rhs(p, self.eps, dict(substrate=vs, mask=vm, top=0.0), fixed_sides=False)
                bc = self.cond.ravel().astype(float) * vm
                basis.append((self.solver.solve(brhs.ravel() - self.A @ bc) + bc).reshape(self.eps.shape))
                rhs_basis.append(brhs)
            basis = np.array(basis)
            electric = [compute_field(v, geo.vac, p, 0.0) for v in basis]
            dt_max = period / p.rf_time_steps
            if p.rf_wave == "pulse":
                dt_max = min(dt_max, p.rf_rise / 8)
            self.rf = dict(period=period, wave=volt_wave - mean, phi=basis, Ex=np.array([v[0] for v in electric]),
                           Ey=np.array([v[1] for v in electric]), dt_max=dt_max, boundary_rhs=np.array(rhs_basis))

    def solve(self, b):
        return self.solver.solve(b)

    def snapshot(self, mean_phi, p):
        """trench_charging_2d.py:416-426: the potential, external potential,
        boundary right-hand side and electrode voltages shown at p.rf_phase
        (circuit_rf); the period means otherwise."""
        if self.rf is None:
            return mean_phi, self.external, self.boundary_rhs, self.volts
        delta = rf_coefficients(self.rf, (p.rf_phase % 360) / 360 * self.rf["period"])
        shift = np.einsum("i,ijk->jk", delta, self.rf["phi"])
        rhs = self.boundary_rhs + np.einsum("i,ijk->jk", delta, self.rf["boundary_rhs"])
        volts = {**self.volts, "substrate": self.volts["substrate"] + float(delta[0]),
                 "mask": self.volts["mask"] + float(delta[1])}
        return mean_phi + shift, self.external + shift, rhs, volts


def compute_field(phi, vac, p, top_voltage=None):
    """E = -grad(phi) at cell centres from vacuum cells only
    (trench_charging_2d.py:429-454): central differences where both
    neighbours are vacuum, one-sided where one is, zero where none is; ghost
    rows at the substrate and at the top (Dirichlet or Neumann)."""
    dx = p.dx

    def neg_grad(c, lo, hi, v_lo, v_hi):
        g = np.zeros_like(c)
        g = np.where(v_lo & v_hi, (hi - lo) / (2 * dx), g)
        g = np.where(v_hi & ~v_lo, (hi - c) / dx, g)
        g = np.where(v_lo & ~v_hi, (c - lo) / dx, g)
        return -g

    ex = neg_grad(phi, np.roll(phi, 1, 0), np.roll(phi, -1, 0), np.roll(vac, 1, 0), np.roll(vac, -1, 0))
    vtop = p.top_voltage if top_voltage is None else top_voltage
    ghost_top = 2 * vtop - phi[:, -1] if p.top_bc == "dirichlet" else phi[:, -1]
    ghost_bot = 2 * p.substrate_voltage - phi[:, 0]
    lo = np.concatenate([ghost_bot[:, None], phi[:, :-1]], axis=1)
    hi = np.concatenate([phi[:, 1:], ghost_top[:, None]], axis=1)
    edge = np.zeros((phi.shape[0], 1), dtype=bool)
    v_lo = np.concatenate([edge, vac[:, :-1]], axis=1)
    v_hi = np.concatenate([vac[:, 1:], ~edge], axis=1)
    return ex, neg_grad(phi, lo, hi, v_lo, v_hi)


def inject(p, n, ny, rng, ion_energy=None, ion_velocity=None):
    """Host injection at the top boundary (trench_charging_2d.py:458-489): n ions
    then n electrons at y = Ly - 1e-3 dx.  Random draws in the reference's
    order: ion x; ion velocity (``ion_velocity`` hook, or energies and a lateral
    thermal normal); electron x; electron vx; electron uniform for the
    flux-weighted (Rayleigh) normal speed."""
    dx = p.dx
    Lx, Ly = p.nx * dx, ny * dx
    m_i = p.ion_mass_amu * AMU
    xi = rng.uniform(0, Lx, n)
    if ion_velocity is not None:
        vxi, vyi = ion_velocity(rng, n)
    else:
        Ei = p.ion_energy_eV * np.ones(n) if ion_energy is None else ion_energy(rng, n)
        vyi = -np.sqrt(2 * Ei * E_CHARGE / m_i)
        vxi = rng.normal(0, np.sqrt(p.ion_temp_eV * E_CHARGE / m_i), n)
    s = np.sqrt(p.electron_temp_eV * E_CHARGE / M_E)
    xe = rng.uniform(0, Lx, n)
    vxe = rng.normal(0, s, n)
    vye = -s * np.sqrt(-2.0 * np.log(1.0 - rng.uniform(size=n)))
    return (np.concatenate([xi, xe]), np.full(2 * n, Ly - 1e-3 * dx), np.concatenate([vxi, vxe]),
            np.concatenate([vyi, vye]),
            np.concatenate([np.full(n, E_CHARGE / m_i), np.full(n, -E_CHARGE / M_E)]),
            np.concatenate([np.zeros(n, dtype=np.int8), np.ones(n, dtype=np.int8)]))
