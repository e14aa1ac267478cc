"""Relaxed barrier decision: derivatives of the charging batch map that see
the reflect / absorb test, for a Newton-Krylov steady solve.

The batch map's response to the surface charge lives mostly in the barrier
test of trench_charging_2d.py:623-640 (trench_trace.cu:128-143): a particle
whose normal energy vn^2/2 lies below q/m times the wall's potential step is
reflected, else absorbed.  Along a fixed particle path that decision is a
step function of the charge, so the pathwise derivative of the deposit leaves
it out.  The relaxed model replaces it by a split: the particle reflects with
probability p = 1 / (1 + exp(t)), t = (vn^2/2 - q/m dphi) / (|q/m| w), the
share 1 - p of its weight is deposited where it hit and the share p goes on
along the reflected path (``HipTracer.soft``, th_trace_soft).  Its deposit is
a smooth function of the charge between the path's other events, and the
device push carries its derivative: the path tangent of the record push plus
dp = -p (1 - p) dt, dt from the normal velocity's tangent and from the
potential tangent at the wall.  As w -> 0 the model tends to the sharp one;
the width w (volts) sets the trade between the bias of the relaxed
expectation (of order w^2 times the curvature of the normal-energy density at
the barrier) and the variance of its derivative (of order 1 / (n w)).

``MapJacobian`` composes the relaxed deposit's derivative with the rest of
the batch map -- the charge update (deposit, leak, floating conductor;
trench_charging_2d.py:1237-1266), which ``ChargeUpdate`` restates on the host
with its derivative -- in the surface-potential coordinates of the steady
solve, for Jacobian-vector products J v of the map u -> u'.
"""
from dataclasses import replace

import numpy as np
from scipy.linalg import lu_factor, lu_solve

from .tangent import linear_field


def conductor_step(n_i, n_e, q, C, Te, drift):
    """Potential change [V] of the floating conductor over one batch: the
    exact solution of dv/dt = (q n_i / C + drift - q n_e / C exp(v / Te)) /
    dt_batch over the batch (trench_charging_2d.py:815-830; the device's
    conductor_step_d)."""
    x_i, x_e = q * n_i / C + drift, q * n_e / C
    if n_e == 0:
        return x_i
    k = x_i / Te
    frac = -np.expm1(-k) / k if k != 0 else 1.0
    return -Te * np.logaddexp(-k, np.log(x_e / Te * frac))


def _dlog_frac(k):
    """d/dk log((1 - e^-k) / k) = 1 / (e^k - 1) - 1 / k (-1/2 + k/12 near 0)."""
    if abs(k) < 1e-4:
        return -0.5 + k / 12.0
    return 1.0 / np.expm1(k) - 1.0 / k


def conductor_step_grad(n_i, n_e, q, C, Te, drift):
    """Partial derivatives of ``conductor_step`` with respect to (n_i, n_e,
    drift).  With k = x_i / Te and L = logaddexp(-k, log(x_e / Te) + log
    frac(k)), v = -Te L and dL = s_a (-dk) + s_b (dx_e / x_e + dlog frac),
    s_a = exp(-k - L), s_b = 1 - s_a."""
    if n_e == 0:
        return q / C, 0.0, 1.0
    x_i, x_e = q * n_i / C + drift, q * n_e / C
    k = x_i / Te
    frac = -np.expm1(-k) / k if k != 0 else 1.0
    b = np.log(x_e / Te * frac)
    L = np.logaddexp(-k, b)
    s_a = np.exp(-k - L)
    s_b = np.exp(b - L)
    dL_dk = -s_a + s_b * _dlog_frac(k)
    dv_dxi = -Te * dL_dk / Te
    dv_dxe = -Te * s_b / x_e
    return dv_dxi * q / C, dv_dxe * q / C, dv_dxi


class ChargeUpdate:
    """The batch's charge update on the surface cells (the device's
    charge_deposit / leak_rows / charge_finish; trench_charging_2d.py:
    1237-1266) with real-valued deposits, and its derivative:

        r = rho + (n_i - n_e) q / dx^2  (conductor cells: + 0)
        rho' = M (r dx^2 + c) / dx^2    (surface leak; identity without)
        driven conductor: rho'[cond] = 0
        floating: drift = sum (rho' - rho) g_cond dx^2,
                  rho'[c_idx] += C (conductor_step(N_i, N_e, q, C, Te, drift) - drift) / dx^2
                  with N_i, N_e the deposits on the conductor's cells.
    """

    def __init__(self, setup):
        s = setup
        self.m = s.geo.surface.size
        self.dx2 = s.p.dx ** 2
        self.cond = s.cond_s.astype(bool)
        self.M = s.leak["M"] if s.leak else None
        self.c = s.leak["c"] if s.leak else None
        self.floating = bool(s.fields.floating)
        self.driven = not self.floating
        self.g = s.g_cond.ravel()[s.geo.surface] if self.floating else None
        self.C, self.c_idx, self.Te = s.C_cond, s.c_idx, s.p.electron_temp_eV

    def _linear(self, rho, n_i, n_e, q, affine):
        r = rho + np.where(self.cond, 0.0, (n_i - n_e) * q / self.dx2)
        if self.M is not None:
            qn = r * self.dx2
            if affine and self.c is not None:
                qn = qn + self.c
            r = self.M @ qn / self.dx2
        if self.driven:
            r = np.where(self.cond, 0.0, r)
        return r

    def apply(self, rho, n_i, n_e, q):
        """The updated charge and (N_i, N_e, drift) of the conductor step."""
        r = self._linear(rho, n_i, n_e, q, True)
        point = None
        if self.floating:
            N_i, N_e = float(n_i[self.cond].sum()), float(n_e[self.cond].sum())
            drift = float(((r - rho) * self.g).sum() * self.dx2)
            dv = conductor_step(N_i, N_e, q, self.C, self.Te, drift)
            r = r.copy()
            r[self.c_idx] += self.C * (dv - drift) / self.dx2
            point = (N_i, N_e, drift)
        return r, point

    def jvp(self, point, q, drho, dn_i, dn_e):
        """Derivative of ``apply`` at the point it returned, along (drho,
        dn_i, dn_e)."""
        dr = self._linear(drho, dn_i, dn_e, q, False)
        if self.floating:
            N_i, N_e, drift = point
            ddrift = float(((dr - drho) * self.g).sum() * self.dx2)
            a_i, a_e, a_d = conductor_step_grad(N_i, N_e, q, self.C, self.Te, drift)
            ddv = a_i * float(dn_i[self.cond].sum()) + a_e * float(dn_e[self.cond].sum()) + a_d * ddrift
            dr = dr.copy()
            dr[self.c_idx] += self.C * (ddv - ddrift) / self.dx2
        return dr


class MapJacobian:
    """Jacobian-vector products of the batch map in surface-potential
    coordinates (one per Green's column: u = P Q, Q the column charges), with
    the relaxed barrier decision: for a direction du,

        dQ = P^-1 du,  drho = dQ / dx^2 on each column's representative cell,
        dphi = G dQ on the rows the push reads,  dE = -grad(dphi),
        (dn_i, dn_e) = derivative of the relaxed face-smoothed deposit of the
                       particles in the device pool (th_trace_soft),
        drho' = derivative of the charge update,  du' = P dQ'(drho').

    ``dev`` is the DeviceCharging whose tracer holds the particles
    (pool[offset, offset + n)) and whose charge sets the field; ``replicas``
    batches of n / replicas particles each are averaged.  ``q`` is the charge
    per particle of those batches.  ``trajectory`` False leaves the field tangent out: the
    paths are held fixed and only the relaxed decisions respond (through the
    potential step at the wall), which drops the path tangents' variance.
    ``max_steps`` (below the map's own) ends the derivative's paths earlier:
    the few particles still in flight then add nothing, and the push's
    latency-bound tail -- the longest paths, traced one step after another --
    shortens."""

    def __init__(self, setup, dev, width, cut=30.0, trajectory=True, max_steps=None):
        s = setup
        self.s, self.dev, self.p = s, dev, s.p
        self.p_push = s.p if max_steps is None else replace(s.p, max_steps=min(int(max_steps), s.p.max_steps))
        self.width, self.cut = float(width), float(cut)
        self.trajectory = bool(trajectory)
        m = s.geo.surface.size
        cols = np.full(s.green.shape[1], -1)
        for k in range(m):
            if s.col_of[k] >= 0 and cols[s.col_of[k]] < 0:
                cols[s.col_of[k]] = k
        self.cols = cols
        self.P = s.surface_response[np.ix_(cols, cols)]
        self.lu = lu_factor(self.P)
        self.has = s.col_of >= 0
        self.idx = s.col_of[self.has]
        self.update = ChargeUpdate(s)
        self.N = s.geo.solid.size
        self.point = None
        self.jvps = 0
        self.push_ms = 0.0

    def totals(self, rho):
        q = np.zeros(self.cols.size)
        np.add.at(q, self.idx, rho[self.has] * self.p.dx ** 2)
        return q

    def at(self, rho, n, q, replicas=1, offset=0):
        """Linearise at charge ``rho`` over pool[offset, offset + n) (``replicas`` batches),
        particles of charge ``q``: sets the device field and traces the
        relaxed deposit (its primal part is the same for every direction)."""
        self.dev.set_rho(rho)
        self.rho = np.asarray(rho, dtype=float).copy()
        self.n, self.q, self.R, self.offset = int(n), float(q), int(replicas), int(offset)
        r = self.dev.tracer.soft(self.offset, self.n, self.p_push, self.width, self.cut)
        self.push_ms += r["push_ms"]
        dep = r["deposit"] / self.R
        _, self.point = self.update.apply(self.rho, dep[0], dep[1], self.q)
        self.deposit = dep
        return dep

    def directional(self, du):
        """(dQ, drho, dphi grid, dEx, dEy) of a surface-potential direction."""
        s, p = self.s, self.p
        dQ = lu_solve(self.lu, du)
        drho = np.zeros(self.rho.size)
        drho[self.cols] = dQ / p.dx ** 2
        dphi = np.zeros(self.N)
        dphi[s.rows] = s.green @ dQ
        dphi = dphi.reshape(s.geo.solid.shape)
        vac = s.geo.vac
        dex, dey = linear_field(dphi, vac, p)
        return dQ, drho, dphi, dex * vac, dey * vac

    def __call__(self, du):
        """J du: the derivative of the map's surface potentials along du."""
        _, drho, dphi, dex, dey = self.directional(np.asarray(du, dtype=float))
        r = self.dev.tracer.soft(self.offset, self.n, self.p_push, self.width, self.cut,
                                 dfield=(dex, dey) if self.trajectory else None, dphi=dphi)
        self.jvps += 1
        self.push_ms += r["push_ms"]
        dd = r["d_deposit"] / self.R
        dr = self.update.jvp(self.point, self.q, drho, dd[0], dd[1])
        return self.P @ self.totals(dr)
