"""Pathwise (forward-mode) derivative of one charging batch on the device.

The deposit of a batch (trench_charging_2d.py:1237-1243) is a sum of integer
counts per surface cell, whose pathwise derivative is zero almost everywhere.
The quantity differentiated here is the face-smoothed deposit: each absorbed
particle's charge is shared between its hit cell and the next surface cell
along the face it crossed, linearly in the crossing point (cloud-in-cell on
the wall: weight |u - 1/2| to the neighbour on the side of u, the rest to the
hit cell, u the crossing point's coordinate along the face in cell units).  It
conserves the absorbed count and is continuous in the crossing point between
the discrete events of a path (cell sequence, wall hits, barrier decisions,
step count) and at wall corners without a neighbour along the face, where the
hit cell keeps the whole charge.

``HipTracer.tangent`` (th_trace_tangent) traces a batch and carries the
derivative of every path along one direction theta; this module supplies the
directions the charging loop is differentiated in:

* the ion distribution: angular sigma [deg] and a shift of the ion energy
  distribution [eV], through the initial velocities (``IonSampler``, the
  draws of configs.ion_velocity_sampler with their derivatives);
* the surface charge, through the field: a charge direction drho gives
  dphi = LU^-1(-drho / eps0) and dE = -grad(dphi) by the field stencil with
  zero ghost voltages (``charge_direction``).

With common random numbers (the same seed at every theta), the derivative
equals the central difference of the traced paths for every particle whose
discrete events are the same at theta - h and theta + h; the derivative of
the face deposit is the sum of those particles' contributions.  Particles that
change events (a barrier decision, the hit cell) contribute jumps that the
pathwise derivative leaves out by construction.
"""
from dataclasses import replace

import numpy as np

from . import model
from .model import AMU, E_CHARGE, EPS0


class IonSampler:
    """The ion velocity draws of configs.ion_velocity_sampler (energy E,
    monoenergetic or E = lo + (hi - lo) sin^2(pi U / 2); polar angle theta ~
    N(0, sigma); vx = v sin(theta), vy = -v cos(theta)), with the energy
    distribution shifted by ``energy_shift`` eV, and the derivatives of the
    last draw with respect to sigma [per degree] and to the shift [per eV] in
    ``tangents`` = {"sigma": (dvx, dvy), "energy": (dvx, dvy)}."""

    def __init__(self, ion_mass_amu, kind, sigma_deg, energy_eV=None, e_lo=None, e_hi=None, energy_shift=0.0,
                 tilt=0.0):
        if kind not in ("mono", "arcsine"):
            raise ValueError(f"unknown ion energy distribution {kind!r}")
        if tilt and kind != "arcsine":
            raise ValueError("only the arcsine energy distribution can be tilted")
        self.m_i = ion_mass_amu * AMU
        self.kind, self.sigma_deg = kind, float(sigma_deg)
        self.energy_eV, self.e_lo, self.e_hi = energy_eV, e_lo, e_hi
        self.energy_shift = float(energy_shift)
        self.tilt = float(tilt)
        self.tangents = None

    def energy_moments(self):
        """Mean and variance of the (untilted, unshifted) energy distribution
        [eV, eV^2]: arcsine on [lo, hi]: (lo + hi) / 2, (hi - lo)^2 / 8."""
        if self.kind == "mono":
            return float(self.energy_eV), 0.0
        return 0.5 * (self.e_lo + self.e_hi), (self.e_hi - self.e_lo) ** 2 / 8.0

    def _tilted_energy(self, u):
        """Energies of the exponentially tilted arcsine distribution
        p(E) exp(tilt (E - mean)) / Z from the uniforms u, by the inverse of
        its distribution function.  The arcsine law is uniform in the angle
        a = pi u / 2 of E = lo + (hi - lo) sin^2(a); the tilted one has density
        proportional to exp(tilt (E(a) - mean)) in a, tabulated on 2^16
        intervals and inverted by linear interpolation."""
        mean, _ = self.energy_moments()
        a = np.linspace(0.0, 0.5 * np.pi, (1 << 16) + 1)
        e = self.e_lo + (self.e_hi - self.e_lo) * np.sin(a) ** 2
        w = np.exp(self.tilt * (e - mean))
        cdf = np.concatenate([[0.0], np.cumsum(0.5 * (w[1:] + w[:-1]) * np.diff(a))])
        return np.interp(u, cdf / cdf[-1], e)

    def __call__(self, rng, n):
        sig = np.radians(self.sigma_deg)
        if self.kind == "mono":
            E = np.full(n, float(self.energy_eV))
        elif self.tilt:
            E = self._tilted_energy(rng.uniform(size=n))
        else:
            E = self.e_lo + (self.e_hi - self.e_lo) * np.sin(0.5 * np.pi * rng.uniform(size=n)) ** 2
        if self.energy_shift:
            E = E + self.energy_shift
        theta = rng.normal(0.0, sig, n)
        v = np.sqrt(2.0 * E * E_CHARGE / self.m_i)
        s, c = np.sin(theta), np.cos(theta)
        dtheta = theta / self.sigma_deg if self.sigma_deg else np.zeros(n)   # theta = sigma * z
        dv = E_CHARGE / (self.m_i * v)                                      # dv/dE
        self.tangents = {"sigma": (v * c * dtheta, v * s * dtheta), "energy": (s * dv, -c * dv)}
        self.theta = theta
        self.energy = E
        return v * s, -v * c

    def energy_score(self):
        """Score of the last draw's energies with respect to the mean energy
        along the exponentially tilted family p(E) exp(t (E - mean)) / Z at
        t = 0: d log p_t / dt = E - mean, and d mean / dt = Var(E), so the
        score per eV of mean energy is (E - mean) / Var(E) (mean and variance
        of the untilted distribution, the shift included in the mean)."""
        mean, var = self.energy_moments()
        if var <= 0.0:
            raise ValueError("a monoenergetic distribution has no tilted family")
        return (self.energy - (mean + self.energy_shift)) / var

    def sigma_score(self):
        """Score of the last draw's polar angles with respect to sigma [per
        degree]: d/dsigma log N(theta; 0, sigma) = (theta^2 / sigma^2 - 1) / sigma
        (sigma in radians, times the radians per degree)."""
        s = np.radians(self.sigma_deg)
        return (self.theta ** 2 / s ** 2 - 1.0) / s * np.radians(1.0)


def ion_sampler(p, energy_shift=0.0, sigma_deg=None, tilt=0.0):
    """IonSampler of Params ``p``'s ion distribution fields."""
    sig = p.ion_angle_deg if sigma_deg is None else sigma_deg
    if p.ion_distribution == "mono":
        return IonSampler(p.ion_mass_amu, "mono", sig, energy_eV=p.ion_energy_eV, energy_shift=energy_shift, tilt=tilt)
    return IonSampler(p.ion_mass_amu, "arcsine", sig, e_lo=p.ion_energy_lo_eV, e_hi=p.ion_energy_hi_eV,
                      energy_shift=energy_shift, tilt=tilt)


def draw_batch(p, ny, seed, sampler, n=None):
    """One host-injected batch (model.inject, the reference's draw order) with
    ``sampler`` as the ion velocity hook; returns the particles and the ion
    velocity tangents (zero for electrons) as {"sigma": ..., "energy": ...}
    of (dvx, dvy) over the whole batch."""
    n = p.n_per_batch if n is None else n
    parts = model.inject(p, n, ny, np.random.default_rng(seed), ion_velocity=sampler)
    total = parts[0].size
    tan = {}
    for key, (dvx, dvy) in sampler.tangents.items():
        full_x, full_y = np.zeros(total), np.zeros(total)
        full_x[:n], full_y[:n] = dvx, dvy
        tan[key] = (full_x, full_y)
    return parts, tan


def linear_field(dphi, vac, p):
    """dE = -grad(dphi) by the field stencil (model.compute_field) with zero
    ghost voltages: the linear part of the field map."""
    return model.compute_field(dphi, vac, replace(p, substrate_voltage=0.0), 0.0)


def charge_direction(fields, vac, p, drho):
    """Potential and field tangents of a charge-density direction drho
    [C/m^3] on the grid: dphi = LU^-1(-drho / eps0), dE = linear_field(dphi)."""
    dphi = fields.solve((-np.asarray(drho, dtype=float) / EPS0).ravel()).reshape(vac.shape)
    dex, dey = linear_field(dphi, vac, p)
    return dphi, dex, dey


def _face_geometry(vac, surface, spc, status, hit, face, endpoints, dx, weights=None):
    """Per absorbed particle on a surface cell: its index, species, weight,
    hit and neighbour surface indices, the neighbour's share |u - 1/2| and
    the slope of that share with respect to the crossing point's coordinate
    along the face (d share / d coordinate [1/m], zero where the crossing
    point is clipped to the face's end or no neighbour takes a share), and
    which end-point coordinate runs along the face (1: y, for an x-face)."""
    vac = np.asarray(vac, dtype=bool)
    nx, ny = vac.shape
    m = surface.size
    smap = np.full(vac.size, -1, dtype=np.int64)
    smap[surface] = np.arange(m)
    a = np.flatnonzero(status == 1)
    a = a[smap[hit[a]] >= 0]   # hit cells off the surface carry no surface charge
    c = hit[a].astype(np.int64)
    cx, cy = c // ny, c % ny
    xf = face[a] == 1
    e = endpoints[a]
    sp = np.asarray(spc)[a].astype(np.int64)
    wt = np.ones(a.size) if weights is None else np.asarray(weights, dtype=float)[a]
    # coordinate along the face and the side of the face the particle came from
    along = np.where(xf, e[:, 1] / dx - cy, np.mod(e[:, 0] / dx - cx, nx))
    across = np.where(xf, np.mod(e[:, 0] / dx - cx, nx), e[:, 1] / dx - cy)
    moving = (along > 0.0) & (along < 1.0)   # a clipped crossing point stays put
    along = np.clip(along, 0.0, 1.0)
    side = np.where(np.rint(across) % nx == 1, 1, -1)   # +1: from the cell after, -1: from the cell before
    step = np.where(along >= 0.5, 1, -1)
    # neighbour along the face and the vacuum cell in front of it
    nbx = np.where(xf, cx, np.mod(cx + step, nx))
    nby = np.where(xf, cy + step, cy)
    fx = np.where(xf, np.mod(nbx + side, nx), nbx)
    fy = np.where(xf, nby, nby + side)
    inside = (nby >= 0) & (nby < ny) & (fy >= 0) & (fy < ny)
    nbx_c, nby_c = np.clip(nbx, 0, nx - 1), np.clip(nby, 0, ny - 1)
    fx_c, fy_c = np.clip(fx, 0, nx - 1), np.clip(fy, 0, ny - 1)
    nb = nbx_c * ny + nby_c
    valid = inside & ~vac[nbx_c, nby_c] & vac[fx_c, fy_c] & (smap[nb] >= 0)
    w_nb = np.where(valid, np.abs(along - 0.5), 0.0)
    slope = np.where(valid & moving, np.sign(along - 0.5), 0.0) / dx
    return a, sp, wt, smap[c], smap[np.where(valid, nb, c)], w_nb, slope, xf


def face_deposit(vac, surface, spc, status, hit, face, endpoints, d_endpoints, dx, weights=None, d_weights=None):
    """Face-smoothed deposit per species on the surface cells and its
    derivative: (dep, ddep), each [2, m] in particles (``surface`` = flat
    indices of the surface cells, ascending).  ``weights`` (per particle)
    scale each particle's contribution to both; ``d_weights`` (per particle)
    are their derivatives, which the derivative then includes."""
    m = surface.size
    dep, ddep = np.zeros((2, m)), np.zeros((2, m))
    a, sp, wt, k_hit, k_nb, w_nb, slope, xf = _face_geometry(vac, surface, spc, status, hit, face, endpoints, dx,
                                                             weights)
    if a.size == 0:
        return dep, ddep
    de = d_endpoints[a]
    dw_nb = slope * np.where(xf, de[:, 1], de[:, 0])
    dwt = np.zeros(a.size) if d_weights is None else np.asarray(d_weights, dtype=float)[a]
    for s in (0, 1):
        sel = sp == s
        dep[s] = np.bincount(k_hit[sel], wt[sel] * (1.0 - w_nb[sel]), m) + np.bincount(k_nb[sel], wt[sel] * w_nb[sel], m)
        ddep[s] = (np.bincount(k_hit[sel], dwt[sel] * (1.0 - w_nb[sel]) - wt[sel] * dw_nb[sel], m) +
                   np.bincount(k_nb[sel], dwt[sel] * w_nb[sel] + wt[sel] * dw_nb[sel], m))
    return dep, ddep


def event_deposit(vac, surface, events, dx):
    """face_deposit over the deposit list of the relaxed-barrier push
    (HipTracer.soft events): each entry is one weighted absorption with its
    weight tangent."""
    ev = events
    n = ev.size
    return face_deposit(vac, surface, ev["species"], np.ones(n, np.int8), ev["cell"], ev["face"],
                        np.column_stack((ev["x"], ev["y"])), np.column_stack((ev["dx"], ev["dy"])), dx,
                        weights=ev["a"], d_weights=ev["da"])


def face_deposit_vjp(vac, surface, spc, status, hit, face, endpoints, dx, gbar, weights=None):
    """Transpose of face_deposit's derivative: for cotangents ``gbar`` [2, m]
    of the deposit, the cotangent of every particle's recorded end point
    [n, 2], so that sum(gbar * ddep) = sum(bar_end * d_endpoints) for any
    d_endpoints."""
    bar = np.zeros((len(status), 2))
    a, sp, wt, k_hit, k_nb, _, slope, xf = _face_geometry(vac, surface, spc, status, hit, face, endpoints, dx,
                                                          weights)
    if a.size:
        coef = wt * slope * (gbar[sp, k_nb] - gbar[sp, k_hit])
        bar[a[xf], 1] = coef[xf]
        bar[a[~xf], 0] = coef[~xf]
    return bar


def linear_field_transpose(bex, bey, vac, p):
    """Transpose of ``linear_field`` (E = -grad(phi) by the field stencil with
    zero ghost voltages): for cotangents (bex, bey) of the field on the grid,
    the cotangent of phi, so that sum(bex Ex + bey Ey) = sum(bphi phi)."""
    vac = np.asarray(vac, dtype=bool)
    dx = p.dx
    bex = np.asarray(bex, dtype=float)
    bey = np.asarray(bey, dtype=float)
    out = np.zeros(vac.shape)
    # x: neighbours west (i-1) and east (i+1), periodic
    v_lo, v_hi = np.roll(vac, 1, 0), np.roll(vac, -1, 0)
    both, only_hi, only_lo = v_lo & v_hi, v_hi & ~v_lo, v_lo & ~v_hi
    c_hi = -bex * (both / (2 * dx) + only_hi / dx)
    c_lo = bex * (both / (2 * dx) + only_lo / dx)
    out += bex * (only_hi / dx - only_lo / dx)
    out += np.roll(c_hi, 1, 0)     # coefficient of phi[i+1] lands on cell i+1
    out += np.roll(c_lo, -1, 0)    # coefficient of phi[i-1] lands on cell i-1
    # y: neighbours below (j-1) and above (j+1); below the grid no vacuum, above
    # it the ghost row (Dirichlet: -phi of the top cell, Neumann: +phi)
    edge = np.zeros((vac.shape[0], 1), dtype=bool)
    v_lo = np.concatenate([edge, vac[:, :-1]], axis=1)
    v_hi = np.concatenate([vac[:, 1:], ~edge], axis=1)
    both, only_hi, only_lo = v_lo & v_hi, v_hi & ~v_lo, v_lo & ~v_hi
    c_hi = -bey * (both / (2 * dx) + only_hi / dx)
    c_lo = bey * (both / (2 * dx) + only_lo / dx)
    out += bey * (only_hi / dx - only_lo / dx)
    out[:, 1:] += c_hi[:, :-1]
    out[:, -1] += c_hi[:, -1] * (-1.0 if p.top_bc == "dirichlet" else 1.0)
    out[:, :-1] += c_lo[:, 1:]
    return out


class BatchDerivative:
    """Traces batches on one geometry and frozen field with the derivative
    along one direction and returns the face deposit and its derivative.

    ``tracer`` is a HipTracer of the geometry; ``barrier`` = (phi, wface) for
    the barrier wall model, else None."""

    def __init__(self, tracer, p, Ex, Ey, barrier=None):
        self.tracer, self.p = tracer, p
        self.Ex, self.Ey, self.barrier = Ex, Ey, barrier
        self.surface = tracer.surface_cells()

    def __call__(self, particles, dstate=None, dfield=None):
        x, y, vx, vy, qm, spc = particles
        r = self.tracer.tangent(x, y, vx, vy, qm, spc, self.Ex, self.Ey, self.p, self.barrier, dstate, dfield)
        dep, ddep = face_deposit(self.tracer.vac, self.surface, spc, r["status"], r["hit"], r["face"],
                                 r["endpoints"], r["d_endpoints"], self.p.dx)
        r.update(deposit=dep, d_deposit=ddep)
        return r


class PathwiseBatch:
    """One batch of the charging loop at a given surface charge with its
    random stream fixed (host injection, seed ``seed``): the face-smoothed
    surface-charge increment [C/m^3 on the surface cells, conductor cells
    zero] (the deposit of trench_charging_2d.py:1237-1243 with the face
    deposit in place of the counts) and its derivatives along the ion
    distribution (sigma, energy shift) and along the surface charge.

    The state is the charge density on the surface cells, ``rho_s`` [m]; the
    field follows the batch loop: phi = LU^-1(-rho / eps0) + phi_ext, E =
    -grad(phi) (trench_charging_2d.py:1268-1269)."""

    def __init__(self, p, seed=0, n=None, device=0):
        from .charging import ChargingSetup
        from .tracer import HipTracer
        self.p = p
        self.setup = setup = ChargingSetup(p, device=device)
        self.geo = setup.geo
        self.fields = setup.fields
        self.tracer = HipTracer(self.geo.vac, device)
        self.surface = self.tracer.surface_cells()
        self.n = p.n_per_batch if n is None else int(n)
        self.seed = int(seed)
        # charge per traced particle: the batch's flux spread over the n traced particles per species
        self.scale = setup.q_macro * p.n_per_batch / self.n / p.dx ** 2
        self.cond_s = setup.cond_s.astype(bool)

    def close(self):
        """Releases the tracer's device memory."""
        self.tracer.close()

    def field(self, rho_s):
        rho = np.zeros(self.geo.solid.size)
        rho[self.surface] = rho_s
        phi = self.fields.solve(-rho / EPS0).reshape(self.geo.solid.shape) + self.fields.external
        ex, ey = model.compute_field(phi, self.geo.vac, self.p, self.fields.volts["top"])
        return phi, ex, ey

    def _trace(self, rho_s, sigma_deg, energy_shift, key=None, dfield=None):
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        parts, tan = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        barrier = (phi, self.fields.wface) if self.p.wall_model == "barrier" else None
        dstate = None if key is None else (None, None, *tan[key])
        bd = BatchDerivative(self.tracer, self.p, ex, ey, barrier)
        r = bd(parts, dstate, dfield)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        dinc = (r["d_deposit"][0] - r["d_deposit"][1]) * self.scale
        inc[self.cond_s] = 0.0
        dinc[self.cond_s] = 0.0
        return inc, dinc, r

    def increment(self, rho_s=None, sigma_deg=None, energy_shift=0.0):
        rho_s = np.zeros(self.surface.size) if rho_s is None else rho_s
        return self._trace(rho_s, sigma_deg, energy_shift)[0]

    def ion_jacobian(self, rho_s=None, sigma_deg=None, energy_shift=0.0):
        """(increment [m], d increment / d (sigma [deg], energy shift [eV]) [m, 2])."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else rho_s
        inc, d_sig, _ = self._trace(rho_s, sigma_deg, energy_shift, "sigma")
        _, d_en, _ = self._trace(rho_s, sigma_deg, energy_shift, "energy")
        return inc, np.column_stack((d_sig, d_en))

    def increment_tilted(self, rho_s, tilt):
        """The increment with the ion energies drawn from the exponentially
        tilted distribution (same random stream)."""
        return self._trace_sampler(rho_s, ion_sampler(self.p, tilt=tilt))[0]

    def _trace_sampler(self, rho_s, sampler, weights_from=None):
        parts, _ = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        barrier = (phi, self.fields.wface) if self.p.wall_model == "barrier" else None
        r = BatchDerivative(self.tracer, self.p, ex, ey, barrier)(parts)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        inc[self.cond_s] = 0.0
        grad = None
        if weights_from is not None:
            score = np.zeros(parts[0].size)
            score[:self.n] = weights_from(sampler)   # ions first; electrons do not depend on it
            dep, _ = face_deposit(self.tracer.vac, self.surface, parts[5], r["status"], r["hit"], r["face"],
                                  r["endpoints"], r["d_endpoints"], self.p.dx, weights=score)
            grad = (dep[0] - dep[1]) * self.scale
            grad[self.cond_s] = 0.0
        return inc, grad

    def energy_score_gradient(self, rho_s=None, sigma_deg=None, energy_shift=0.0):
        """Likelihood-ratio estimate of the derivative of the expected
        increment with respect to the mean ion energy [per eV], along the
        exponentially tilted family of the energy distribution (IonSampler.
        energy_score): each ion's face deposit weighted by (E - mean) /
        Var(E), from one trace.  Unbiased for that family, with a bounded
        score (a shift of the arcsine law has a score of infinite variance at
        its edges).  Returns (increment [m], d E[increment] / d mean [m])."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else rho_s
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        return self._trace_sampler(rho_s, sampler, weights_from=lambda s: s.energy_score())

    def sigma_score_gradient(self, rho_s=None, sigma_deg=None, energy_shift=0.0):
        """Likelihood-ratio (score-function) estimate of the derivative of the
        expected increment with respect to the ion angular sigma [per degree],
        from one trace: each ion's face deposit weighted by the score of its
        polar angle (IonSampler.sigma_score), whose mean is zero.  Unbiased:
        unlike the pathwise derivative it includes the particles that change
        events.  Returns (increment [m], d E[increment] / d sigma [m])."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else rho_s
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        parts, _ = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        barrier = (phi, self.fields.wface) if self.p.wall_model == "barrier" else None
        bd = BatchDerivative(self.tracer, self.p, ex, ey, barrier)
        r = bd(parts)
        score = np.zeros(parts[0].size)
        score[:self.n] = sampler.sigma_score()   # ions first; electrons do not depend on sigma
        dep, _ = face_deposit(self.tracer.vac, self.surface, parts[5], r["status"], r["hit"], r["face"],
                              r["endpoints"], r["d_endpoints"], self.p.dx, weights=score)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        grad = (dep[0] - dep[1]) * self.scale
        inc[self.cond_s] = 0.0
        grad[self.cond_s] = 0.0
        return inc, grad

    def vjp(self, rho_s, g, sigma_deg=None, energy_shift=0.0):
        """Reverse-mode derivative of <g, increment> from one primal trace and
        one adjoint push (th_trace_adjoint): returns (increment [m], gradient
        with respect to the surface charge rho_s [m], gradient with respect
        to (ion angular sigma [deg], energy shift [eV]) [2]).  The transpose of
        charge_jvp and ion_jacobian: the same pathwise derivative (events held
        fixed), for all m + 2 directions at once."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else np.asarray(rho_s, dtype=float)
        g = np.asarray(g, dtype=float)
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        parts, tan = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        barrier = (phi, self.fields.wface) if self.p.wall_model == "barrier" else None
        r = BatchDerivative(self.tracer, self.p, ex, ey, barrier)(parts)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        inc[self.cond_s] = 0.0
        gs = np.where(self.cond_s, 0.0, g) * self.scale
        bar_end = face_deposit_vjp(self.geo.vac, self.surface, parts[5], r["status"], r["hit"], r["face"],
                                   r["endpoints"], self.p.dx, np.stack([gs, -gs]))
        bar_state, bar_field = self.tracer.adjoint(*parts, ex, ey, self.p, barrier, r["steps"], bar_end)
        vac = self.geo.vac
        bphi = linear_field_transpose(bar_field[..., 0] * vac, bar_field[..., 1] * vac, vac, self.p)
        s = self.setup
        per_col = s.green.T @ bphi.ravel()[s.rows]            # d<g, inc> / d(column charge) [per C/m]
        grad_rho = np.where(s.col_of >= 0, per_col[np.maximum(s.col_of, 0)], 0.0) * self.p.dx ** 2
        ion = np.array([bar_state[:, 2] @ tan[k][0] + bar_state[:, 3] @ tan[k][1] for k in ("sigma", "energy")])
        return inc, grad_rho, ion

    def relaxed(self, rho_s, width, drho_s=None, key=None, cut=30.0, sigma_deg=None, energy_shift=0.0, events=0,
                trajectory=True):
        """The increment with the relaxed barrier decision (relaxed.py,
        th_trace_soft: reflection with probability 1 / (1 + exp(t)) of the
        normal-energy margin t in units of ``width`` volts), and its derivative
        along a surface-charge direction ``drho_s`` and/or an ion-parameter
        direction ``key`` ("sigma" [per degree] or "energy" [per eV]).
        ``trajectory`` False leaves the field tangent of the charge direction
        out (paths held fixed, only the decisions respond).  Returns
        (increment [m], derivative [m], the push's output dict)."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else rho_s
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        parts, tan = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        total = parts[0].size
        t = self.tracer
        t.reserve(total)
        t.upload(0, *parts)
        t.set_fields(ex, ey, (phi, self.fields.wface))
        dfield = dphi = None
        if drho_s is not None:
            drho = np.zeros(self.geo.solid.size)
            drho[self.surface] = drho_s
            dphi, dex, dey = charge_direction(self.fields, self.geo.vac, self.p, drho.reshape(self.geo.solid.shape))
            dfield = (dex, dey) if trajectory else None
        dstate = None if key is None else (None, None, *tan[key])
        r = t.soft(0, total, self.p, width, cut, dstate, dfield, dphi, events=events, per_particle=True)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        dinc = (r["d_deposit"][0] - r["d_deposit"][1]) * self.scale
        inc[self.cond_s] = 0.0
        dinc[self.cond_s] = 0.0
        return inc, dinc, r

    def relaxed_vjp(self, rho_s, g, width, cut=30.0, sigma_deg=None, energy_shift=0.0):
        """Reverse mode of ``relaxed``: from one relaxed push and its adjoint
        (th_trace_soft_adjoint), the gradient of <g, relaxed increment> with
        respect to every surface cell's charge (through the field and through
        the potential at the walls' relaxed decisions) and to (ion angular
        sigma [deg], energy shift [eV]).  Returns (increment [m], gradient
        [m], ion gradient [2])."""
        rho_s = np.zeros(self.surface.size) if rho_s is None else np.asarray(rho_s, dtype=float)
        sampler = ion_sampler(self.p, energy_shift, sigma_deg)
        parts, tan = draw_batch(self.p, self.geo.ny, self.seed, sampler, self.n)
        phi, ex, ey = self.field(rho_s)
        total = parts[0].size
        t = self.tracer
        t.reserve(total)
        t.upload(0, *parts)
        t.set_fields(ex, ey, (phi, self.fields.wface))
        r = t.soft(0, total, self.p, width, cut, per_particle=True)
        inc = (r["deposit"][0] - r["deposit"][1]) * self.scale
        inc[self.cond_s] = 0.0
        gs = np.where(self.cond_s, 0.0, np.asarray(g, dtype=float)) * self.scale
        bar_state, bar_field, bar_phi = t.soft_adjoint(0, total, self.p, width, r["steps"], np.stack([gs, -gs]), cut)
        vac = self.geo.vac
        bphi = linear_field_transpose(bar_field[..., 0] * vac, bar_field[..., 1] * vac, vac, self.p) + bar_phi
        s = self.setup
        per_col = s.green.T @ bphi.ravel()[s.rows]
        grad_rho = np.where(s.col_of >= 0, per_col[np.maximum(s.col_of, 0)], 0.0) * self.p.dx ** 2
        ion = np.array([bar_state[:, 2] @ tan[k][0] + bar_state[:, 3] @ tan[k][1] for k in ("sigma", "energy")])
        return inc, grad_rho, ion

    def charge_jvp(self, rho_s, drho_s, sigma_deg=None, energy_shift=0.0):
        """(increment [m], d increment along the surface-charge direction drho_s [m])."""
        drho = np.zeros(self.geo.solid.size)
        drho[self.surface] = drho_s
        _, dex, dey = charge_direction(self.fields, self.geo.vac, self.p, drho.reshape(self.geo.solid.shape))
        inc, dinc, _ = self._trace(rho_s, sigma_deg, energy_shift, dfield=(dex, dey))
        return inc, dinc


def _cos(a, b):
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    return float(a @ b / (na * nb)) if na > 0 and nb > 0 else 0.0


def derivative_agreement(p, rho_s, seeds=(7, 8), steps=(0.01, 0.05, 0.2), n=None, device=0, relaxed_width=None):
    """How well the pathwise derivative of the batch increment estimates the
    derivative of its expectation, per direction (ion angular sigma, energy
    shift, the surface charge's own direction rho_s), and the likelihood-ratio
    estimate for sigma ("sigma_score"): the cosine between the estimates of
    two random streams, between the central differences (common random
    numbers, relative steps ``steps``) of the two streams, and between
    estimate and difference, with their norms.  The differences include the
    particles that change events; the pathwise derivative leaves them out,
    the likelihood ratio does not.  With ``relaxed_width`` (volts) also the
    derivative of the relaxed-barrier increment ("*_relaxed"), against the
    differences of the plain increment."""
    batches = [PathwiseBatch(p, seed=s, n=n, device=device) for s in seeds]
    sig = p.ion_angle_deg
    e_mean = p.ion_energy_eV if p.ion_distribution == "mono" else 0.5 * (p.ion_energy_lo_eV + p.ion_energy_hi_eV)
    out = {}
    names = ["sigma", "sigma_score", "energy", "charge"]
    if p.ion_distribution != "mono":
        names.insert(3, "energy_score")
    if relaxed_width:
        names += ["sigma_relaxed", "energy_relaxed", "charge_relaxed"]
    for name in names:
        pw, fds = [], []
        for b in batches:
            if name.endswith("_relaxed"):
                base = name[:-len("_relaxed")]
                if base == "charge":
                    pw.append(b.relaxed(rho_s, relaxed_width, drho_s=rho_s)[1])
                    f = lambda e: b.increment(rho_s * (1.0 + e))
                    scale = 1.0
                elif base == "sigma":
                    pw.append(b.relaxed(rho_s, relaxed_width, key="sigma", sigma_deg=sig)[1])
                    f = lambda e: b.increment(rho_s, sig * (1.0 + e))
                    scale = sig
                else:
                    pw.append(b.relaxed(rho_s, relaxed_width, key="energy", sigma_deg=sig)[1])
                    f = lambda e: b.increment(rho_s, sig, e * e_mean)
                    scale = e_mean
            elif name == "sigma_score":
                pw.append(b.sigma_score_gradient(rho_s, sig, 0.0)[1])
                f = lambda e: b.increment(rho_s, sig * (1.0 + e))
                scale = sig
            elif name == "energy_score":
                # along the tilted family: mean energy changes by tilt * Var(E) to first order
                mean, var = ion_sampler(p).energy_moments()
                pw.append(b.energy_score_gradient(rho_s, sig, 0.0)[1])
                f = lambda e: b.increment_tilted(rho_s, e * mean / var)
                scale = mean
            elif name == "charge":
                pw.append(b.charge_jvp(rho_s, rho_s)[1])
                f = lambda e: b.increment(rho_s * (1.0 + e))
                scale = 1.0
            elif name == "sigma":
                pw.append(b.ion_jacobian(rho_s, sig, 0.0)[1][:, 0])
                f = lambda e: b.increment(rho_s, sig * (1.0 + e))
                scale = sig
            else:
                pw.append(b.ion_jacobian(rho_s, sig, 0.0)[1][:, 1])
                f = lambda e: b.increment(rho_s, sig, e * e_mean)
                scale = e_mean
            fds.append([(f(e) - f(-e)) / (2.0 * e * scale) for e in steps])
        out[name] = dict(
            estimate_norm=float(np.linalg.norm(pw[0])), estimate_seed_cosine=_cos(pw[0], pw[1]),
            steps=list(steps), difference_norm=[float(np.linalg.norm(d)) for d in fds[0]],
            difference_seed_cosine=[_cos(a, b) for a, b in zip(fds[0], fds[1])],
            estimate_difference_cosine=[_cos(pw[0] + pw[1], a + b) for a, b in zip(fds[0], fds[1])])
    for b in batches:
        b.close()
    return out
