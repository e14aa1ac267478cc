"""ORACLE -- test infrastructure only (tests, smoke() and the bench's cpu_baseline
leg).  NumPy/SciPy restatement of trench01's per-batch charging loop around the
push: geometry, Poisson assembly and LU, field, injection, surface leak,
floating conductor, steady-state criterion and the time-marching batch loop.
The push itself is the C restatement in ``trace_ref.c``.

Every function cites the reference lines it restates.  Floating-point
expressions keep the reference's evaluation order so that a whole run can be
compared bit for bit with the reference's own regression fixtures
(tests/data/base_*.npz, made by tests/test_regression.py:25-50).

Scope: field_mode floating / fixed (bias dc, optionally with ``ion_energy`` /
``ion_velocity`` samplers) and the circuit modes "circuit" (the circuit's
voltages frozen at one phase) and "circuit_rf" (the push follows one RF period
of voltages, launch phases drawn per batch); side boundaries periodic, fixed
and pillar; no resume.
"""
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sps
from scipy.sparse.linalg import splu

from . import trace_ref

# trench_charging_2d.py:90-93
E_CHARGE = 1.602176634e-19
EPS0 = 8.8541878128e-12
M_E = 9.1093837015e-31
AMU = 1.66053906660e-27


@dataclass
class Params:
    """Subset of trench_charging_2d.py:103-174 used by the charging path (same names and defaults)."""
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
    bias: str = "dc"
    rf_phase: float = 0.0
    rf_time_steps: int = 256
    rf_wave: str = "sine"
    rf_rise: float = 1e-9
    sigma_s: float = 1e-15
    until_steady: bool = False
    max_batches: int = 3000
    steady_window: int = 30
    steady_rtol: float = 0.005
    steady_charge_rtol: float = 0.005
    steady_charge_atol: float = 1e-15
    steady_hold: int = 2
    n_batches: int = 300
    n_per_batch: int = 3000
    dt_batch: float = 1e-4
    cfl: float = 0.35
    max_steps: int = 3000
    wall_model: str = "barrier"
    seed: int = 1


# ------------------------------------------------------------------ geometry
def opening_widths(p):
    """trench_charging_2d.py:178-182."""
    return dict(bottom=p.trench_w - 2 * p.trench_d * np.tan(np.radians(p.trench_taper_deg)),
                interface=float(p.trench_w),
                top=p.trench_w + 2 * p.mask_t * np.tan(np.radians(p.mask_taper_deg)))


def build_geometry(p):
    """trench_charging_2d.py:185-226: staircase masked trench, cell-centre opening test."""
    ints = ("nx", "trench_w", "trench_offset", "trench_d", "floor_t", "n_vac", "mask_t")
    if any(not isinstance(getattr(p, k), (int, np.integer)) for k in ints) or p.nx < 3 \
            or not 1 <= p.trench_w <= p.nx - 2 or min(p.floor_t, p.trench_d, p.n_vac) < 1:
        raise ValueError("cell counts must be positive integers and trench_w <= nx-2")
    i0 = (p.nx - p.trench_w) // 2 + p.trench_offset
    i1 = i0 + p.trench_w
    if i0 < 1 or i1 > p.nx - 1:
        raise ValueError("at least one solid column must remain on each side")
    if p.mask_t < 0 or p.mask_type not in ("conductor", "dielectric"):
        raise ValueError("bad mask")
    for key in ("trench_taper_deg", "mask_taper_deg"):
        a = getattr(p, key)
        if not np.isfinite(a) or not -90 < a < 90:
            raise ValueError(f"{key} must be finite and inside (-90, 90)")
    centre = (i0 + i1) / 2
    for width in opening_widths(p).values():
        if width < 1 or centre - width / 2 < 1 or centre + width / 2 > p.nx - 1:
            raise ValueError("opening closes or leaves the domain")
    top_ox = p.floor_t + p.trench_d
    ny = top_ox + p.mask_t + p.n_vac
    solid = np.zeros((p.nx, ny), dtype=bool)
    solid[:, :top_ox + p.mask_t] = True
    yc = np.arange(p.floor_t, top_ox + p.mask_t) + 0.5
    slope = np.tan(np.radians(np.where(yc < top_ox, p.trench_taper_deg, p.mask_taper_deg)))
    half = p.trench_w / 2 + (yc - top_ox) * slope
    xc = (np.arange(p.nx) + 0.5)[:, None]
    solid[:, p.floor_t:top_ox + p.mask_t] = ~((xc >= centre - half) & (xc < centre + half))
    mask = solid.copy()
    mask[:, :top_ox] = False
    return solid, mask, i0, i1


def opening_span(solid, row):
    """trench_charging_2d.py:229-232."""
    cells = np.flatnonzero(~solid[:, row])
    return int(cells[0]), int(cells[-1] + 1)


def wall_bounds(solid, p):
    """trench_charging_2d.py:235-241."""
    iy = np.arange(p.floor_t, p.floor_t + p.trench_d + p.mask_t)
    vac = ~solid[:, iy]
    return np.argmax(vac, axis=0) - 1, solid.shape[0] - np.argmax(vac[::-1], axis=0), iy


def permittivity(solid, mask, p):
    """trench_charging_2d.py:244-248."""
    eps = np.where(solid, p.eps_r, 1.0)
    return np.where(mask, p.mask_eps_r, eps) if p.mask_type == "dielectric" else eps


# ------------------------------------------------------------------ Poisson
def build_poisson(eps, cond, p):
    """trench_charging_2d.py:252-294: 5-point div(eps grad) with harmonic face
    permittivity, conductor faces at 2*eps, x periodic, half-cell Dirichlet
    faces at the substrate (and at the top when top_bc is dirichlet)."""
    nx, ny = eps.shape
    dx2 = p.dx ** 2

    def face(a, b, ca, cb):
        h = np.where(ca, 2.0 * b, np.where(cb, 2.0 * a, 2.0 * a * b / (a + b)))
        return np.where(ca & cb, 0.0, h)

    east = face(eps, np.roll(eps, -1, axis=0), cond, np.roll(cond, -1, axis=0)) / dx2
    top_ox = p.floor_t + p.trench_d
    if p.side_bc != "periodic":
        east[-1, :top_ox] = 0.0
    west = np.roll(east, 1, axis=0)
    vert = face(eps[:, :-1], eps[:, 1:], cond[:, :-1], cond[:, 1:]) / dx2
    north = np.zeros((nx, ny))
    north[:, :-1] = vert
    south = np.zeros((nx, ny))
    south[:, 1:] = vert
    dn, ds = north.copy(), south.copy()
    ds[:, 0] += 2.0 * eps[:, 0] / dx2
    if p.top_bc == "dirichlet":
        dn[:, -1] += 2.0 * eps[:, -1] / dx2
    diag = -(east + west + dn + ds)
    if p.side_bc != "periodic":
        diag[0, :top_ox] -= 2.0 * eps[0, :top_ox] / dx2
        diag[-1, :top_ox] -= 2.0 * eps[-1, :top_ox] / dx2
    idx = np.arange(nx * ny).reshape(nx, ny)
    rows = np.concatenate([idx.ravel(), idx.ravel(), idx.ravel(), idx[:, :-1].ravel(), idx[:, 1:].ravel()])
    cols = np.concatenate([idx.ravel(), np.roll(idx, -1, axis=0).ravel(), np.roll(idx, 1, axis=0).ravel(),
                           idx[:, 1:].ravel(), idx[:, :-1].ravel()])
    vals = np.concatenate([diag.ravel(), east.ravel(), west.ravel(), north[:, :-1].ravel(),
                           south[:, 1:].ravel()])
    return sps.coo_matrix((vals, (rows, cols)), shape=(nx * ny, nx * ny)).tocsc()


def make_solver(A, cond):
    """trench_charging_2d.py:297-309: LU once; conductor cells share one unknown (E^T A E)."""
    if not cond.any():
        return splu(A).solve
    flat = cond.ravel()
    n_free = int((~flat).sum())
    col = np.where(flat, n_free, np.cumsum(~flat) - 1)
    E = sps.csr_matrix((np.ones(flat.size), (np.arange(flat.size), col)), shape=(flat.size, n_free + 1))
    lu = splu((E.T @ A @ E).tocsc())
    return lambda b: E @ lu.solve(E.T @ b)


def boundary_rhs(p, eps, volts, fixed_sides=True):
    """trench_charging_2d.py:323-337: Dirichlet faces of the substrate, the top
    and (side_bc fixed / pillar) the dielectric's side faces; pillar sides
    ramp linearly from the substrate to the mask voltage over the dielectric's
    height.  ``fixed_sides`` False: fixed sides at 0 V (the RF basis)."""
    rhs = np.zeros_like(eps)
    rhs[:, 0] -= 2.0 * eps[:, 0] / p.dx ** 2 * volts["substrate"]
    if p.top_bc == "dirichlet":
        rhs[:, -1] -= 2.0 * eps[:, -1] / p.dx ** 2 * volts["top"]
    top_ox = p.floor_t + p.trench_d
    if p.side_bc == "pillar":
        frac = (np.arange(top_ox) + 0.5) / top_ox
        left = right = volts["substrate"] + (volts["mask"] - volts["substrate"]) * frac
    elif fixed_sides:
        left, right = p.left_voltage, p.right_voltage
    else:
        left = right = 0.0
    if p.side_bc != "periodic":
        rhs[0, :top_ox] -= 2.0 * eps[0, :top_ox] / p.dx ** 2 * left
        rhs[-1, :top_ox] -= 2.0 * eps[-1, :top_ox] / p.dx ** 2 * right
    return rhs


def circuit_voltages(c, phase):
    """trench_charging_2d.py:312-320: the circuit's electrode voltages at
    ``phase`` [deg] by periodic linear interpolation, relative to the plasma
    potential (top at 0 V)."""
    t = np.asarray(c["t"])
    period = (t[1] - t[0]) * t.size
    at = (phase % 360.0) / 360.0 * period
    v = {k: float(np.interp(at, np.append(t, period), np.append(c[k], c[k][0]))) for k in ("uw", "um", "up", "V1")}
    return dict(substrate=v["uw"] - v["up"], mask=v["um"] - v["up"], top=0.0, sheath=v["V1"])


def launch_times(rng, n, period):
    """trench_rf.py:14-18: one time per n-th of the period, uniform inside its
    stratum, in shuffled order."""
    t = (np.arange(n) + rng.random(n)) * period / n
    rng.shuffle(t)
    return t


def rf_coefficients(rf, at):
    """trench_rf.py:35-44: the wave's deviation from its mean at times ``at``,
    linear between the n_phase nodes of one period (periodic)."""
    wave = np.asarray(rf["wave"])
    n = len(wave)
    u = np.mod(np.asarray(at, dtype=float), rf["period"]) / rf["period"] * n
    k = np.floor(u).astype(np.int64) % n
    f = (u - np.floor(u))[..., None]
    return wave[k] * (1 - f) + wave[(k + 1) % n] * f


def setup_fields(p, solid, mask, circuit=None):
    """trench_charging_2d.py:340-413: charge-free potential of the electrode
    voltages and the charge response (floating / fixed / circuit /
    circuit_rf).  For circuit_rf the external potential is the period mean and
    ``rf`` holds the unit-voltage responses of substrate and mask (potential,
    field, boundary terms) with the wave's deviation from its mean."""
    driven = p.field_mode != "floating"
    if p.field_mode in ("circuit", "circuit_rf"):
        if p.bias != "rf" or circuit is None or p.top_bc != "dirichlet":
            raise ValueError("circuit field modes need bias rf, a circuit waveform and a Dirichlet top")
        volts = circuit_voltages(circuit, p.rf_phase)
    elif p.field_mode in ("floating", "fixed"):
        volts = dict(substrate=p.substrate_voltage, mask=p.mask_voltage, top=p.top_voltage)
    else:
        raise ValueError(f"unknown field_mode {p.field_mode!r}")
    if p.side_bc == "pillar" and not driven:
        raise ValueError("pillar sides need driven electrodes")
    if p.field_mode == "circuit_rf":
        t = np.asarray(circuit["t"])
        period = (t[1] - t[0]) * t.size
        if p.rf_time_steps < 16 or p.dt_batch < period * (1 - 1e-12):
            raise ValueError("circuit_rf needs rf_time_steps >= 16 and dt_batch of at least one period")
        wave = np.column_stack((circuit["uw"] - circuit["up"], circuit["um"] - circuit["up"]))
        mean = wave.mean(axis=0)
        volts = dict(substrate=float(mean[0]), mask=float(mean[1]), top=0.0, sheath=float(np.mean(circuit["V1"])))
    cond = mask if p.mask_type == "conductor" else np.zeros_like(mask)
    eps = permittivity(solid, mask, p)
    A = build_poisson(eps, cond, p)
    if driven and cond.any():
        free = ~cond.ravel()
        lu = splu(A[free][:, free].tocsc())

        def solve(b):
            out = np.zeros_like(b, dtype=float)
            out[free] = lu.solve(b[free])
            return out

        phi_cond = cond.ravel().astype(float) * volts["mask"]
        mask_rhs = -(A @ phi_cond)
    else:
        solve = make_solver(A, cond)
        phi_cond = np.zeros(cond.size)
        mask_rhs = np.zeros(cond.size)
    rhs = boundary_rhs(p, eps, volts)
    external = (solve(rhs.ravel() + mask_rhs) + phi_cond).reshape(eps.shape)
    fields = dict(A=A, solve=solve, cond=cond, eps=eps, external=external, boundary_rhs=rhs,
                  volts=volts, floating=bool(cond.any() and not driven), rf=None)
    if p.field_mode == "circuit_rf":
        # unit responses: substrate at 1 V, then mask at 1 V (:397-406); the
        # field of each with a 0 V top ghost (:407)
        phis, rhss = [], []
        for vs, vm in ((1.0, 0.0), (0.0, 1.0)):
            b = boundary_rhs(p, eps, dict(substrate=vs, mask=vm, top=0.0), fixed_sides=False)
            on_mask = cond.ravel().astype(float) * vm
            phis.append((solve(b.ravel() - A @ on_mask) + on_mask).reshape(eps.shape))
            rhss.append(b)
        e = [compute_field(v, ~solid, p, 0.0) for v in phis]
        dt_max = period / p.rf_time_steps
        if p.rf_wave == "pulse":
            dt_max = min(dt_max, p.rf_rise / 8)
        fields["rf"] = dict(period=period, wave=wave - mean, phi=np.array(phis), Ex=np.array([a for a, _ in e]),
                            Ey=np.array([b for _, b in e]), dt_max=dt_max, boundary_rhs=np.array(rhss))
    return fields


def field_snapshot(fields, mean_phi, p):
    """trench_charging_2d.py:416-426: potential, external potential, boundary
    terms and voltages shown at p.rf_phase (circuit_rf); the period means
    otherwise."""
    rf = fields["rf"]
    if rf is None:
        return mean_phi, fields["external"], fields["boundary_rhs"], fields["volts"]
    d = rf_coefficients(rf, (p.rf_phase % 360) / 360 * rf["period"])
    shift = d[0] * rf["phi"][0] + d[1] * rf["phi"][1]
    rhs = fields["boundary_rhs"] + (d[0] * rf["boundary_rhs"][0] + d[1] * rf["boundary_rhs"][1])
    volts = {**fields["volts"], "substrate": fields["volts"]["substrate"] + float(d[0]),
             "mask": fields["volts"]["mask"] + float(d[1])}
    return mean_phi + shift, fields["external"] + shift, rhs, volts


def _neg_grad(c, m, pl, vm, vp, dx):
    """trench_charging_2d.py:429-436: vacuum-only central / one-sided difference."""
    g = np.zeros_like(c)
    g = np.where(vm & vp, (pl - m) / (2 * dx), g)
    g = np.where(vp & ~vm, (pl - c) / dx, g)
    g = np.where(vm & ~vp, (c - m) / dx, g)
    return -g


def compute_field(phi, vac, p, top_voltage=None):
    """trench_charging_2d.py:439-454: E = -grad(phi) at cell centres, ghost rows top and bottom."""
    dx = p.dx
    Ex = _neg_grad(phi, np.roll(phi, 1, 0), np.roll(phi, -1, 0), np.roll(vac, 1, 0), np.roll(vac, -1, 0), dx)
    vtop = p.top_voltage if top_voltage is None else top_voltage
    ghost_top = 2 * vtop - phi[:, -1] if p.top_bc == "dirichlet" else phi[:, -1]
    ghost_bot = 2 * p.substrate_voltage - phi[:, 0]
    below = np.concatenate([ghost_bot[:, None], phi[:, :-1]], axis=1)
    above = np.concatenate([phi[:, 1:], ghost_top[:, None]], axis=1)
    edge = np.zeros((phi.shape[0], 1), dtype=bool)
    vb = np.concatenate([edge, vac[:, :-1]], axis=1)
    va = np.concatenate([vac[:, 1:], ~edge], axis=1)
    return Ex, _neg_grad(phi, below, above, vb, va, dx)


# ------------------------------------------------------------------ injection
def inject(p, n, ny, rng, ion_energy=None, ion_velocity=None):
    """trench_charging_2d.py:458-489.  Draw order is the contract: ion x, ion
    velocity (sampler, or lateral normal), electron x, electron vx, electron
    uniform for the flux-weighted normal speed."""
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
    x = np.concatenate([xi, xe])
    y = np.full(2 * n, Ly - 1e-3 * dx)
    vx = np.concatenate([vxi, vxe])
    vy = np.concatenate([vyi, vye])
    qm = np.concatenate([np.full(n, E_CHARGE / m_i), np.full(n, -E_CHARGE / M_E)])
    spc = np.concatenate([np.zeros(n, dtype=np.int8), np.ones(n, dtype=np.int8)])
    return x, y, vx, vy, qm, spc


# ------------------------------------------------------------------ surface leak
def surface_cells(solid, vacuum=None):
    """trench_charging_2d.py:734-744: solid cells with a vacuum cell among their 8 neighbours."""
    nx, ny = solid.shape
    vp = np.pad(~solid if vacuum is None else vacuum, ((0, 0), (1, 1)), constant_values=False)
    near = np.zeros_like(solid)
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            if sx or sy:
                near |= np.roll(vp, -sx, axis=0)[:, 1 + sy:1 + sy + ny]
    return solid & near


def inner_surface_cells(solid, p):
    """trench_charging_2d.py:747-752."""
    inner = ~solid
    inner[:, :p.floor_t] = False
    inner[:, p.floor_t + p.trench_d + p.mask_t:] = False
    return surface_cells(solid, inner)


def build_leakage(solid, solve, p):
    """trench_charging_2d.py:768-803: G (graph Laplacian of 4-connected surface
    cells, g = sigma_s/dx), P = surface rows of solve(B), M = inv(I - dt G P)."""
    nx, ny = solid.shape
    dx = p.dx
    ix, iy = np.nonzero(surface_cells(solid))
    m = ix.size
    smap = -np.ones((nx, ny), dtype=np.int64)
    smap[ix, iy] = np.arange(m)
    g = p.sigma_s / dx
    G = np.zeros((m, m))
    for a, b in ((smap, np.roll(smap, -1, axis=0)), (smap[:, :-1], smap[:, 1:])):
        sel = (a >= 0) & (b >= 0)
        a, b = a[sel], b[sel]
        np.add.at(G, (a, b), g)
        np.add.at(G, (b, a), g)
        np.add.at(G, (a, a), -g)
        np.add.at(G, (b, b), -g)
    flat = ix * ny + iy
    B = np.zeros((nx * ny, m))
    B[flat, np.arange(m)] = -1.0 / (dx ** 2 * EPS0)
    P = solve(B)[flat, :]
    M = np.linalg.inv(np.eye(m) - p.dt_batch * (G @ P))
    cut = p.floor_t + p.trench_d // 2
    region = inner_surface_cells(solid, p)[ix, iy] & (iy < cut)
    return dict(flat=flat, M=M, G=G, region=region)


# ------------------------------------------------------------------ floating conductor
def conductor_response(solve, cond, dx):
    """trench_charging_2d.py:807-812: potential everywhere from 1 C/m on the conductor."""
    b = np.zeros(cond.size)
    b[np.flatnonzero(cond)[0]] = -1.0 / (dx ** 2 * EPS0)
    return solve(b).reshape(cond.shape)


def conductor_step(n_i, n_e, q, C, Te, drift=0.0):
    """trench_charging_2d.py:815-830: exact solution over one batch of
    dv/dt = (q n_i/C + drift - q n_e/C exp(v/Te)) / dt_batch."""
    x_i, x_e = q * n_i / C + drift, q * n_e / C
    if n_e == 0:
        return x_i
    k = x_i / Te
    frac = -np.expm1(-k) / k if k != 0 else 1.0
    return -Te * np.logaddexp(-k, np.log(x_e / Te * frac))


# ------------------------------------------------------------------ steady criterion
def charge_monitors(solid, mask, i0, i1, p):
    """trench_charging_2d.py:850-868."""
    exposed = surface_cells(solid)
    dielectric = exposed & ~(mask if p.mask_type == "conductor" else np.zeros_like(mask))
    ix, iy = np.indices(solid.shape)
    wall = (iy >= p.floor_t) & (iy < p.floor_t + p.trench_d)
    inner = inner_surface_cells(solid, p)
    centre = (i0 + i1) / 2
    b0, b1 = opening_span(solid, p.floor_t)
    regions = dict(left=(ix + 0.5 < centre) & wall & inner, right=(ix + 0.5 >= centre) & wall & inner,
                   bottom=(ix >= b0) & (ix < b1) & (iy == p.floor_t - 1), surface=dielectric)
    for side in ("left", "right"):
        for band, cells in enumerate(np.array_split(np.arange(p.floor_t, p.floor_t + p.trench_d), 3)):
            if len(cells):
                regions[f"{side} band {band + 1}"] = regions[side] & np.isin(iy, cells)
    if mask.any():
        regions["mask"] = mask
    return {k: np.flatnonzero(r) for k, r in regions.items()}


def sample_charge_monitors(rho, regions, dx):
    """trench_charging_2d.py:871-874."""
    out = {k: float(rho.ravel()[c].sum() * dx ** 2) for k, c in regions.items()}
    out["surface absolute"] = float(np.abs(rho.ravel()[regions["surface"]]).sum() * dx ** 2)
    return out


def mean_and_se(x):
    """trench_charging_2d.py:877-890: detrended, AR(1)-corrected standard error."""
    n = x.size
    t = np.arange(n)
    resid = x - np.polyval(np.polyfit(t, x, 1), t)
    var = resid.var(ddof=2)
    r1 = 0.0
    if var > 0:
        r1 = np.corrcoef(resid[:-1], resid[1:])[0, 1]
        r1 = float(np.clip(r1, 0.0, 0.95)) if np.isfinite(r1) else 0.0
    return x.mean(), np.sqrt(var / max(n * (1 - r1) / (1 + r1), 1.0))


def check_steady(hist, probes, p, reference=None):
    """trench_charging_2d.py:893-910."""
    W = p.steady_window
    stats = {k: (mean_and_se(np.asarray(hist[k][-W:])), mean_and_se(np.asarray(hist[k][-2 * W:-W])))
             for k in probes}
    scale = max(1.0, max(abs(c[0] - (reference[k] if reference is not None else 0.0))
                         for k, (c, _) in stats.items()))
    ok, worst = True, (None, 0.0, 1.0, -1.0)
    for k, ((mc, sc), (mp, sp)) in stats.items():
        d = mc - mp
        tol = max(p.steady_rtol * scale, 3.0 * np.hypot(sc, sp))
        ok &= abs(d) <= tol
        if abs(d) / tol > worst[3]:
            worst = (k, d, tol, abs(d) / tol)
    return ok, worst[0], worst[1], worst[2]


def check_charge_steady(charge_hist, p):
    """trench_charging_2d.py:913-947: block-mean difference and net drift per region."""
    W = p.steady_window
    ok, worst = True, None
    for key, values in charge_hist.items():
        q = np.asarray(values[-2 * W:], dtype=float)
        cur, se_cur = mean_and_se(q[-W:])
        prev, se_prev = mean_and_se(q[:W])
        diff = float(cur - prev)
        mean_tol = max(p.steady_charge_atol, p.steady_charge_rtol * max(abs(cur), abs(prev)),
                       3 * np.hypot(se_cur, se_prev))
        rate, se_rate = mean_and_se(np.diff(q))
        drift = float(rate * W)
        drift_tol = max(p.steady_charge_atol, float(3 * se_rate * W))
        for kind, d, tol in (("mean", diff, mean_tol), ("drift", drift, drift_tol)):
            ratio = abs(d) / tol if tol > 0 else (0.0 if d == 0 else np.inf)
            ok &= abs(d) <= tol
            if worst is None or ratio > worst["ratio"]:
                worst = dict(name=key, kind=kind, d=d, tol=float(tol), ratio=float(ratio))
    return {**worst, "ok": bool(ok)}


def probes_for(solid, p, cond):
    """trench_charging_2d.py:1152-1166: probe cells (ix, iy)."""
    top_ox = p.floor_t + p.trench_d
    left, right, _ = wall_bounds(solid, p)
    b0, b1 = opening_span(solid, p.floor_t)
    e0, _ = opening_span(solid, top_ox + p.mask_t - 1)
    probes = {"mask top": (max(0, e0 - 10), top_ox + p.mask_t - 1)}
    if p.mask_t > 0 and not cond.any():
        ym = top_ox + p.mask_t // 2
        probes["mask sidewall"] = (int(left[ym - p.floor_t]), ym)
    for level, yw in (("upper", max(p.floor_t, top_ox - 4)), ("middle", p.floor_t + p.trench_d // 2),
                      ("lower", min(top_ox - 1, p.floor_t + 3))):
        probes[f"sidewall {level}"] = (int(left[yw - p.floor_t]), yw)
    probes["bottom center"] = ((b0 + b1) // 2, p.floor_t - 1)
    for level in ("upper", "middle", "lower"):
        yw = probes[f"sidewall {level}"][1]
        probes[f"right sidewall {level}"] = (int(right[yw - p.floor_t]), yw)
    return probes


# ------------------------------------------------------------------ batch loop
def batch_context(p, circuit=None):
    """Set-up of the per-batch update for a geometry (trench_charging_2d.py:1063-1094)."""
    solid, mask, i0, i1 = build_geometry(p)
    fields = setup_fields(p, solid, mask, circuit)
    cond, solve = fields["cond"], fields["solve"]
    ctx = dict(p=p, vac=~solid, fields=fields, cond=cond, conductor=fields["floating"],
               driven=p.field_mode != "floating", solve=solve,
               leak=build_leakage(solid, solve, p) if p.sigma_s > 0 else None,
               q_macro=E_CHARGE * (p.flux * (solid.shape[0] * p.dx) * p.dt_batch / p.n_per_batch))
    ctx["e0"], ctx["e1"] = opening_span(solid, p.floor_t + p.trench_d + p.mask_t - 1)
    if ctx["conductor"]:
        g = conductor_response(solve, cond, p.dx)
        ctx.update(g_cond=g, C_cond=1.0 / g[cond][0],
                   c_cell=np.flatnonzero(surface_cells(solid) & cond)[0])
    return ctx


def batch_update(ctx, rho, cnt_i, cnt_e):
    """One batch after the push (trench_charging_2d.py:1239-1269): deposit,
    surface leak, driven / floating conductor, Poisson solve and field.
    Updates ``rho`` in place; returns (rho, phi, Ex, Ey, leak_ratio)."""
    p, fields, cond = ctx["p"], ctx["fields"], ctx["cond"]
    dx = p.dx
    nx, ny = rho.shape
    leak = ctx["leak"]
    rho_prev = rho.copy() if ctx["conductor"] else None
    drho = (cnt_i - cnt_e) * ctx["q_macro"] / dx ** 2
    if cond.any():
        drho[cond] = 0.0
    rho += drho
    leak_ratio = 0.0
    if leak is not None:
        Qs = rho.flat[leak["flat"]] * dx ** 2
        Qs_old = Qs
        if np.any(fields["external"]):
            Qs = Qs + p.dt_batch * (leak["G"] @ fields["external"].ravel()[leak["flat"]])
        Qs_new = leak["M"] @ Qs
        rho.flat[leak["flat"]] = Qs_new / dx ** 2
        dQ = Qs_old[leak["region"]].sum() - Qs_new[leak["region"]].sum()
        leak_ratio = dQ / p.dt_batch / (E_CHARGE * p.flux * (ctx["e1"] - ctx["e0"]) * dx)
    if ctx["driven"]:
        rho[cond] = 0.0
    if ctx["conductor"]:
        drift = ((rho - rho_prev) * ctx["g_cond"]).sum() * dx ** 2
        dV = conductor_step(cnt_i[cond].sum(), cnt_e[cond].sum(), ctx["q_macro"], ctx["C_cond"],
                            p.electron_temp_eV, drift)
        rho.flat[ctx["c_cell"]] += ctx["C_cond"] * (dV - drift) / dx ** 2
    phi = ctx["solve"]((-rho / EPS0).ravel()).reshape(nx, ny) + fields["external"]
    Ex, Ey = compute_field(phi, ctx["vac"], p, fields["volts"]["top"])
    return rho, phi, Ex, Ey, leak_ratio


def run(p, tracer=None, ion_energy=None, ion_velocity=None, batches=None, on_batch=None, circuit=None,
        blas_tail=False):
    """trench_charging_2d.py:1045-1369 (setup :1063-1223, loop :1225-1335,
    result :1337-1369), restated without resume.

    ``tracer`` defaults to the C push.  ``on_batch(b, state)`` is called after
    every batch with the loop's arrays (for per-batch comparisons).
    ``circuit``: the circuit waveform (t, uw, um, up, V1) of the circuit field
    modes.  ``blas_tail``: the C push rounds the RF field bound as NumPy's
    single-threaded matrix product does (for comparisons with the reference)."""
    if tracer is None:
        tracer = (lambda *a, **k: trace_ref.oracle_tracer(*a, **k, blas_tail=blas_tail)) if blas_tail \
            else trace_ref.oracle_tracer
    rng = np.random.default_rng(p.seed)
    solid, mask, i0, i1 = build_geometry(p)
    vac = ~solid
    nx, ny = solid.shape
    dx = p.dx
    fields = setup_fields(p, solid, mask, circuit)
    # stratified launch phases on a stream of their own (:1079)
    phase_rng = np.random.default_rng(np.random.SeedSequence([p.seed, 0x5246])) if fields["rf"] else None
    if p.field_mode in ("circuit", "circuit_rf"):
        # ions enter at the Bohm energy Te/2 and are accelerated in the domain (:1131-1136)
        ion_velocity = None
        ion_energy = lambda rng_, n_: np.full(n_, p.electron_temp_eV / 2)
    cond, eps, solve = fields["cond"], fields["eps"], fields["solve"]
    conductor = fields["floating"]
    driven = p.field_mode != "floating"
    wface = np.where(cond, 1.0, eps / (1.0 + eps))
    leak = build_leakage(solid, solve, p) if p.sigma_s > 0 else None
    if conductor:
        g_cond = conductor_response(solve, cond, dx)
        C_cond = 1.0 / g_cond[cond][0]
        c_cell = np.flatnonzero(surface_cells(solid) & cond)[0]
    q_macro = E_CHARGE * (p.flux * (nx * dx) * p.dt_batch / p.n_per_batch)
    b0, b1 = opening_span(solid, p.floor_t)
    e0, e1 = opening_span(solid, p.floor_t + p.trench_d + p.mask_t - 1)
    ctx = dict(p=p, vac=vac, fields=fields, cond=cond, conductor=conductor, driven=driven, leak=leak,
               q_macro=q_macro, e0=e0, e1=e1, solve=solve)
    if conductor:
        ctx.update(g_cond=g_cond, C_cond=C_cond, c_cell=c_cell)
    probes = probes_for(solid, p, cond)
    reference = ({k: float(fields["external"][ix, iy]) for k, (ix, iy) in probes.items()}
                 if fields["rf"] is not None else None)                                   # :1167-1168
    hist = {k: [] for k in ("t", "ratio_i_bottom", "ratio_e_bottom", "esc_e", "lost", "leak", *probes)}
    charge_hist = {"left": [], "right": []}
    q_regions = charge_monitors(solid, mask, i0, i1, p)
    charge_monitor_hist = {k: [] for k in (*q_regions, "surface absolute")}
    rho = np.zeros((nx, ny))
    phi = fields["external"].copy()
    Ex, Ey = compute_field(phi, vac, p, fields["volts"]["top"])
    frac_open = (b1 - b0) / nx
    n_max = batches if batches is not None else (p.max_batches if p.until_steady else p.n_batches)
    n_ok, converged_at = 0, None
    for b in range(n_max):
        x, y, vx, vy, qm, spc = inject(p, p.n_per_batch, ny, rng, ion_energy, ion_velocity)
        options = {}
        if fields["rf"] is not None:                                                      # :1231-1234
            options["rf"] = {**fields["rf"], "initial_time": np.concatenate(
                [launch_times(phase_rng, p.n_per_batch, fields["rf"]["period"]) for _ in range(2)])}
        cnt_i, cnt_e, esc, n_lost = tracer(x, y, vx, vy, qm, spc, Ex, Ey, vac, p,
                                           (phi, wface) if p.wall_model == "barrier" else None, **options)
        rho, phi, Ex, Ey, leak_ratio = batch_update(ctx, rho, cnt_i, cnt_e)
        hist["t"].append((b + 1) * p.dt_batch)
        for k, (ix, iy) in probes.items():
            hist[k].append(phi[ix, iy])
        norm = p.n_per_batch * frac_open
        hist["ratio_i_bottom"].append(cnt_i[b0:b1, p.floor_t - 1].sum() / norm)
        hist["ratio_e_bottom"].append(cnt_e[b0:b1, p.floor_t - 1].sum() / norm)
        hist["esc_e"].append(esc[1] / p.n_per_batch)
        hist["lost"].append(n_lost / (2 * p.n_per_batch))
        hist["leak"].append(leak_ratio)
        charges = sample_charge_monitors(rho, q_regions, dx)
        for side in ("left", "right"):
            charge_hist[side].append(charges[side])
        for key, q in charges.items():
            charge_monitor_hist[key].append(q)
        if on_batch is not None:
            on_batch(b, dict(cnt_i=cnt_i, cnt_e=cnt_e, escaped=esc, lost=n_lost, rho=rho, phi=phi,
                             Ex=Ex, Ey=Ey))
        W = p.steady_window
        if p.until_steady and (b + 1) % W == 0 and b + 1 >= 2 * W:
            v_ok = check_steady(hist, probes, p, reference)[0]
            ok = bool(v_ok and check_charge_steady(charge_monitor_hist, p)["ok"])
            n_ok = min(n_ok + 1, p.steady_hold) if ok else 0
            if n_ok >= p.steady_hold:
                converged_at = b + 1
                break
    hist = {k: np.array(v) for k, v in hist.items()}
    Wl = min(p.steady_window, len(hist["t"]))
    steady = {k: float(hist[k][-Wl:].mean()) for k in probes}
    shown_phi, shown_ext, shown_rhs, volts = field_snapshot(fields, phi, p)          # :1351-1354
    if fields["rf"] is not None:
        Ex, Ey = compute_field(shown_phi, vac, p, volts["top"])
    rho_shown = rho.copy()
    if cond.any():
        induced = (fields["A"] @ shown_phi.ravel()).reshape(rho.shape) - shown_rhs
        rho_shown[cond] = (-EPS0 * induced)[cond]
    return dict(phi=shown_phi, mean_phi=phi, external_phi=shown_ext, voltages=volts, rho=rho_shown, rho_raw=rho,
                Ex=Ex, Ey=Ey, solid=solid, mask=mask, i0=i0, i1=i1,
                hist=hist, probes=probes, steady=steady, charge_hist=charge_hist,
                charge_monitor_hist={k: np.array(v) for k, v in charge_monitor_hist.items()},
                t_conv=converged_at * p.dt_batch if converged_at else np.nan)
