"""Steady-state surface charge: a Krylov (Anderson) solve of the batch map's
fixed point, then the reference's own criterion on plain batches.

Plain time-marching (``charging.run`` with ``until_steady``) applies the batch
map -- inject, push, deposit, leak, floating conductor, potential, field
(trench_charging_2d.py:1225-1269) -- once per batch and stops when the
reference criterion passes (:893-947, :1299-1316).  At the AR 8 workload it
needs about 1600 batches: the trench floor and sidewall charge relax with a
time constant of a few hundred batches.

The steady state is the fixed point of that map, including every detail of
its per-batch arithmetic: at this workload the long-run potentials move with
dt_batch (about +1 V at the trench floor per doubling, independent of the
particle count), so the fixed point of a different discretisation would not
reproduce the time-marched answer.  ``anderson`` solves for the fixed point
of the map itself:

* one evaluation g(x) runs the plain batch from the charge x, R times with
  different particles, and averages the resulting surface potentials; at
  every evaluation of a stage the particles are the same (device injection
  with fixed seed and batch indices: common random numbers), so g is a
  deterministic map;
* the unknowns are the surface potentials (one per Green's column, volts);
* Anderson acceleration (type II, Walker & Ni, SIAM J. Numer. Anal. 49(4),
  2011) extrapolates from the last ``memory`` secant pairs; for a linear map
  it is GMRES on (I - A) x = b (their sec. 4): a Krylov solve that needs
  nothing beyond evaluations of the map;
* stages raise R (same K batches per replica, so the same Jacobian up to
  sampling error) and carry the secant pairs over: the first stage finds the
  fixed point of a small particle sample, later ones refine it for the
  sampling error of R*K batches in a few evaluations;
* a stage may trace a fraction f of the particles per batch at 1/f the charge
  per particle: the expected deposit, and so the expected map and its
  Jacobian, are those of the full batch (the conductor's one-batch step has
  the same root), at a fraction of the cost where a batch is short enough
  for the push's tail of long paths to dominate;
* the R replicas of a one-batch evaluation start from the same charge, so
  they share its field and are traced by a single push
  (th_charging_step_replicas), which pays the tail once;
* then plain batches with fresh particles run until the reference criterion
  (window W, hold) passes; the reported steady values are their last-window
  means, as the plain method reports its own (:1342-1343).

No pathwise adjoint of the push enters the Anderson solve: deposit counts are
integers, so their derivative with respect to the charge along a fixed
particle path is zero almost everywhere; secant information over finite steps
is what carries the response.  ``newton_krylov`` solves for the same fixed
point with Newton steps whose GMRES Jacobian-vector products come from the
push itself, with the barrier decision relaxed to a logistic
(relaxed.MapJacobian), or from differences of the map.
"""
import time

import numpy as np

from . import monitors
from .charging import ChargingSetup, DeviceCharging

# (K batches per replica, R replicas, tolerance [V], particle fraction): a
# coarse approach from zero charge with a quarter of the particles per batch
# (at four times the charge each: the same expected batch map), a refinement
# at half, and the final solve at full size over 24 replicas.
DEFAULT_STAGES = ((1, 2, 0.01, 0.25), (1, 4, 0.003, 0.5), (1, 24, 0.001, 1.0))
# Batch size (particles per species) the default tolerances are set for; the
# sampling noise of a batch map scales as 1 / sqrt(n_per_batch), and so do
# the default tolerances for other batch sizes.
DEFAULT_STAGES_N = 500_000


def default_stages(p):
    """DEFAULT_STAGES with the tolerances scaled to ``p.n_per_batch``'s
    sampling noise (never below the defaults)."""
    scale = max(1.0, float(np.sqrt(DEFAULT_STAGES_N / max(1, p.n_per_batch))))
    return tuple((K, R, tol * scale, f) for K, R, tol, f in DEFAULT_STAGES)


def anderson_update(pairs, f, gx, reg=1e-12):
    """Type-II Anderson step: with secant pairs (dF, dG) (newest last), the
    residual f = g(x) - x and the image g(x), returns the next iterate
    g(x) - dG gamma, gamma = argmin |f - dF gamma| (Tikhonov ``reg`` relative
    to the trace).  No pairs: the plain step g(x)."""
    if not pairs:
        return gx.copy()
    dF = np.array([a for a, _ in pairs]).T
    dG = np.array([b for _, b in pairs]).T
    A = dF.T @ dF
    gamma = np.linalg.solve(A + reg * np.trace(A) * np.eye(A.shape[0]), dF.T @ f)
    return gx - dG @ gamma


def anderson_iterate(g, x, pairs, tol, memory=10, max_evals=40, patience=None, log=None):
    """Iterate x <- Anderson(g) until max |g(x) - x| < tol, ``max_evals``
    evaluations, or ``patience`` evaluations without a new smallest residual
    (the sampling noise of the map then dominates the residual).  ``g(x)``
    returns (state, image); ``pairs`` (secant pairs, newest last) is updated in
    place and may come from an earlier map with the same Jacobian.  Returns
    (state at the last evaluation, its image, evaluations)."""
    last, evals, best, since_best = None, 0, np.inf, 0
    while True:
        state, gx = g(x)
        evals += 1
        f = gx - x
        res = float(np.max(np.abs(f)))
        if log:
            log(evals, res)
        since_best = 0 if res < best else since_best + 1
        if res < tol or evals >= max_evals or (patience is not None and since_best >= patience):
            return state, gx, evals
        if res > 4 * best and pairs:        # safeguard: drop the history, take a plain step
            pairs.clear()
            last = None
        best = min(best, res)
        if last is not None:
            pairs.append((f - last[0], gx - last[1]))
            if len(pairs) > memory:
                pairs.pop(0)
        last = (f, gx)
        x = anderson_update(pairs, f, gx)


def potential_map(setup):
    """Surface potentials from charge, one unknown per Green's column (each
    dielectric surface cell, and a floating conductor's total charge):
    u = P Q [V] and back by LU.  A conductor's charge may sit on any of its
    cells (they share one potential); only its total enters."""
    from scipy.linalg import lu_factor, lu_solve
    s = setup
    m = s.geo.surface.size
    cols = np.full(s.green.shape[1], -1)
    for k in range(m):                      # one representative cell per column
        if s.col_of[k] >= 0 and cols[s.col_of[k]] < 0:
            cols[s.col_of[k]] = k
    P = s.surface_response[np.ix_(cols, cols)]
    lu = lu_factor(P)
    dx2 = s.p.dx ** 2
    has = s.col_of >= 0
    idx = s.col_of[has]

    def totals(rho_s):
        q = np.zeros(cols.size)
        np.add.at(q, idx, rho_s[has] * dx2)
        return q

    def to_u(rho_s):
        return P @ totals(rho_s)

    def to_rho(u, like):
        """Charge with surface potentials u; within a column (a conductor) the
        distribution of ``like`` is kept and its representative cell takes
        the change of the total."""
        rho = like.copy()
        rho[cols] += (lu_solve(lu, u) - totals(like)) / dx2
        return rho

    return to_u, to_rho


class _Session:
    """One device-resident charging state for a steady solve: the batch map
    with common random numbers, the potential coordinates, and the plain
    phase with the reference criterion."""

    def __init__(self, p, device, setup, push_precision, green_precision, replicas=1, merge=True):
        self.t_start = time.perf_counter()
        self.p = p
        self.s = s = setup or ChargingSetup(p, device=device)
        self.dev = dev = DeviceCharging(s, device=device)
        dev.set_schedule(precision=push_precision)
        if green_precision:
            dev.set_green_precision(green_precision)
        dev.set_monitors(s.probes, s.regions)
        self.n = 2 * p.n_per_batch
        self.merge = merge
        dev.reserve(self.n * (max(1, replicas) if merge else 1))
        self.to_u, self.to_rho = potential_map(s)
        self.setup_s = time.perf_counter() - self.t_start
        self.traced = 0        # batches traced, in units of full batches
        self.frac = 1.0

    def set_fraction(self, frac):
        """Batches of ``frac`` times the particles at 1/frac the charge per
        particle (the same expected deposit per batch)."""
        self.frac = frac
        self.dev.set_q_macro(self.s.q_macro / frac)

    def evaluate(self, rho, first, K, R):
        """The batch map from charge ``rho``: R replicas of K batches (device
        injection, batch indices first .. first + R*K - 1), surface potentials
        averaged over the replicas.  Returns (charge of the last replica,
        mean surface potentials).  With K = 1 the replicas share the field of
        ``rho`` and are traced by one push (th_charging_step_replicas): the
        same results as R separate batches, with the push's tail of long
        paths paid once."""
        p, dev = self.p, self.dev
        n_per = max(1, int(round(p.n_per_batch * self.frac)))
        n = 2 * n_per
        if K == 1 and self.merge:
            dev.set_rho(rho)
            for r in range(R):
                dev.inject_device(r * n, n_per, p.seed, first + r)
            rhos, _, _ = dev.step_replicas(0, n, R)
            self.traced += R * self.frac
            return rhos[-1], np.mean([self.to_u(x) for x in rhos], axis=0)
        us = []
        for r in range(R):
            dev.set_rho(rho)
            for b in range(K):
                dev.inject_device(0, n_per, p.seed, first + r * K + b)
                dev.step(0, n)
            self.traced += K * self.frac
            rho_g = dev.read(phi=False, counts=False)[0]
            us.append(self.to_u(rho_g))
        return rho_g, np.mean(us, axis=0)

    def plain_phase(self, rho, log):
        """Plain batches with fresh particles from ``rho`` until the reference
        criterion passes (trench_charging_2d.py:1299-1316); the steady values
        are the last-window means (:1342-1343)."""
        p, dev, s = self.p, self.dev, self.s
        probes = s.probes
        self.set_fraction(1.0)
        dev.set_rho(rho)
        hist = {k: [] for k in probes}
        qhist = {k: [] for k in (*s.regions, "surface absolute")}
        W = p.steady_window
        n_ok, done, b2 = 0, False, 0
        first = 1 << 20                              # batch indices away from the solver's
        pending = []
        while self.traced < p.max_batches:
            if not pending:   # replay the captured batch up to the next check
                k = int(min(W - b2 % W, np.ceil(p.max_batches - self.traced)))
                pending, _ = dev.run_batches(0, p.n_per_batch, p.seed, first + b2, max(1, k))
            values, charges, _, _, tally, _ = pending.pop(0)
            if tally["off_surface"]:
                raise RuntimeError(f"{tally['off_surface']} particles were absorbed off the surface cells")
            b2 += 1
            self.traced += 1
            for k in probes:
                hist[k].append(values[k])
            for k, q in charges.items():
                qhist[k].append(q)
            if b2 % W == 0 and b2 >= 2 * W:
                v_ok, name, dv, tol = monitors.check_steady(hist, probes, p)
                q_ok = monitors.check_charge_steady(qhist, p)["ok"]
                n_ok = min(n_ok + 1, p.steady_hold) if (v_ok and q_ok) else 0
                log(f"  [plain] batch {b2}: voltage {'OK' if v_ok else 'NG'} ({name} {dv:+.3g} V, tol {tol:.3g}); "
                    f"charge {'OK' if q_ok else 'NG'} -> {n_ok}/{p.steady_hold}")
                if n_ok >= p.steady_hold:
                    done = True
                    break
        rho_s, _, _ = dev.read(phi=False, counts=False)
        Wl = min(W, len(hist[next(iter(probes))]))
        steady = {k: float(np.mean(hist[k][-Wl:])) for k in probes}
        return done, b2, steady, {k: np.array(v) for k, v in hist.items()}, rho_s

    def close(self):
        self.dev.close()


def _secant_stages(ss, stages, memory, max_evals, patience, log):
    """The Anderson stages of ``anderson`` on session ``ss`` from zero charge:
    returns (charge, surface potentials, evaluations per stage, next free
    batch index)."""
    pairs = []
    rho = np.zeros(ss.dev.m)
    x = ss.to_u(rho)
    evals_per_stage = []
    first = 0
    for stage in stages:
        K, R, tol = stage[:3]
        ss.set_fraction(stage[3] if len(stage) > 3 else 1.0)
        box = {"like": rho}

        def g(x, first=first, K=K, R=R):
            rho_in = ss.to_rho(x, box["like"])
            rho_g, gx = ss.evaluate(rho_in, first, K, R)
            box["like"] = rho_g
            return rho_g, gx

        rho_g, x, e = anderson_iterate(
            g, x, pairs, tol, memory, max_evals, patience,
            log=lambda k, res, K=K, R=R: log(f"  [anderson K={K} R={R}] eval {k}: max |g(x) - x| = {res:.4g} V"))
        rho = ss.to_rho(x, rho_g)
        evals_per_stage.append(e)
        first += K * R
    ss.set_fraction(1.0)
    return rho, x, evals_per_stage, first


def anderson(p, stages=None, memory=10, max_evals=40, patience=10, log=None, device=0, setup=None,
             push_precision=0, green_precision=0, merge=True, secant_push_precision=None):
    """Steady state by multilevel Anderson acceleration of the batch map with
    common random numbers, then plain batches to the reference criterion.
    ``stages`` lists (K, R, tol[, f]): K batches per replica, R replicas
    averaged per evaluation, stop when no surface potential moves by tol
    volts (or after ``patience`` evaluations without progress); f (default 1)
    the fraction of the batch's particles traced, at 1/f the charge each.
    Default: ``default_stages(p)``.  ``push_precision`` /
    ``green_precision`` select the precision variants of the batch (FP32
    push; BF16 / FP16 split Green's matrix).  ``merge``: trace the replicas
    of a K = 1 evaluation in one push (same results).
    ``secant_push_precision`` (if given) is the push precision of the secant
    stages only; the plain phase then runs with ``push_precision``.
    Returns a dict: converged, batches (all traced), phases, steady (probe ->
    last-window mean), hist (probe histories of the plain phase), rho_s,
    wall_s and setup_s."""
    log = log or (lambda s: None)
    stages = default_stages(p) if stages is None else stages
    ss = _Session(p, device, setup, push_precision, green_precision, max(st[1] for st in stages), merge)
    if secant_push_precision is not None:
        ss.dev.set_schedule(precision=secant_push_precision)
    rho, x, evals_per_stage, _ = _secant_stages(ss, stages, memory, max_evals, patience, log)
    secant_batches = ss.traced
    if secant_push_precision is not None:
        ss.dev.set_schedule(precision=push_precision)
    done, b2, steady, hist, rho_s = ss.plain_phase(rho, log)
    wall = time.perf_counter() - ss.t_start
    ss.close()
    return dict(method="multilevel Anderson acceleration of the batch map (common random numbers), "
                       "then plain batches to the reference criterion",
                converged=done, batches=ss.traced,
                phases=[dict(kind="secant", stages=[list(st) for st in stages], memory=memory,
                             evaluations_per_stage=evals_per_stage, batches=secant_batches),
                        dict(kind="plain", batches=b2, converged=done)],
                steady=steady, hist=hist, rho_s=rho_s, wall_s=wall, setup_s=ss.setup_s)


def gmres(matvec, b, krylov, rtol=None):
    """``krylov`` steps of GMRES from x0 = 0 (Arnoldi with modified
    Gram-Schmidt, least squares on the Hessenberg matrix; Saad & Schultz,
    SIAM J. Sci. Stat. Comput. 7(3), 1986).  Stops early when the Krylov space
    is exhausted or, with ``rtol``, once the least-squares residual |b - A x|
    falls to rtol |b| (the forcing term of an inexact Newton step).  Returns
    (x, matvecs used)."""
    n = b.size
    beta = float(np.linalg.norm(b))
    if beta == 0.0:
        return np.zeros(n), 0
    V = np.zeros((krylov + 1, n))
    H = np.zeros((krylov + 1, krylov))
    V[0] = b / beta
    k = 0
    for j in range(krylov):
        w = matvec(V[j])
        k += 1
        for i in range(j + 1):
            H[i, j] = V[i] @ w
            w = w - H[i, j] * V[i]
        H[j + 1, j] = float(np.linalg.norm(w))
        if H[j + 1, j] <= 1e-14 * beta:
            j += 1
            break
        V[j + 1] = w / H[j + 1, j]
        if rtol is not None:
            e1 = np.zeros(j + 2)
            e1[0] = beta
            y = np.linalg.lstsq(H[:j + 2, :j + 1], e1, rcond=None)[0]
            if np.linalg.norm(H[:j + 2, :j + 1] @ y - e1) <= rtol * beta:
                j += 1
                break
    else:
        j = krylov
    e1 = np.zeros(j + 1)
    e1[0] = beta
    y = np.linalg.lstsq(H[:j + 1, :j], e1, rcond=None)[0]
    return V[:j].T @ y, k


def newton_krylov_iterate(g, x, tol, h, krylov=30, max_newton=6, backtrack=4, log=None, jvp=None, forcing=None):
    """Jacobian-free Newton-Krylov on F(x) = g(x) - x: every Newton step
    solves (J_g - I) dx = -F(x) by ``krylov`` GMRES steps whose Jacobian-vector
    products are forward differences of g with step ``h`` along the unit
    direction (J_g v ~ |v| (g(x + h v/|v|) - g(x)) / h; g sees the same
    particles at every evaluation: common random numbers), or ``jvp(x, v)``
    (J_g v at the last evaluated point x) when given; ``forcing`` stops
    GMRES once its residual is that fraction of |F(x)|.  A step that does
    not lower max |F| is halved up to ``backtrack`` times (then the plain step
    x <- g(x) is taken).  Stops when max |F| < tol or after ``max_newton``
    steps.  ``g(x)`` returns (state, image).  Returns (state at the last
    evaluation, its image, evaluations)."""
    state, gx = g(x)
    evals = 1
    res = float(np.max(np.abs(gx - x)))
    for it in range(max_newton):
        if log:
            log(it, res, evals)
        if res < tol:
            break
        f = gx - x

        def matvec(v, x=x, gx=gx):
            if jvp is not None:
                return jvp(x, v) - v
            nrm = float(np.linalg.norm(v))
            if nrm == 0.0:
                return np.zeros_like(v)
            _, gp = g(x + (h / nrm) * v)
            return (gp - gx) * (nrm / h) - v

        dx, k = gmres(matvec, -f, krylov, forcing)
        if jvp is None:   # each difference JVP is one evaluation of g
            evals += k
        lam = 1.0
        for _ in range(backtrack + 1):
            xt = x + lam * dx
            st, gt = g(xt)
            evals += 1
            rt = float(np.max(np.abs(gt - xt)))
            if rt < res:
                break
            lam *= 0.5
        else:                                   # no descent along dx: plain step
            xt, st, gt = gx, *g(gx)
            evals += 1
            rt = float(np.max(np.abs(gt - xt)))
        x, state, gx, res = xt, st, gt, rt
    return state, gx, evals


def newton_krylov(p, K=1, R=16, h_volts=2.0, krylov=30, max_newton=1, tol=None, start=None, log=None, device=0,
                  setup=None, jvp="relaxed", width=1.0, jvp_fraction=0.25, trajectory=True, forcing=0.03,
                  jvp_max_steps=1000, push_precision=0, green_precision=0):
    """Steady state by Newton-Krylov on the batch map (GMRES, ``krylov``
    steps per Newton step, R replicas of K batches per map evaluation with
    common random numbers, traced in one push for K = 1), started from the
    Anderson stages ``start`` (default: the first two default stages), then
    plain batches to the reference criterion.  Jacobian-vector products:
    ``jvp`` = "difference", forward differences of step ``h_volts`` on the
    surface potentials; "relaxed", the derivative of the map with the barrier
    decision relaxed to a logistic of ``width`` volts (relaxed.MapJacobian)
    over a batch of ``jvp_fraction`` times the particles (at 1 / jvp_fraction
    the charge each; the same batch for every JVP: common random numbers),
    with the paths' response to the field (``trajectory``)
    or with the paths held fixed.  ``forcing`` ends a Newton step's GMRES
    early once its residual is that fraction of the map's residual (None: all
    ``krylov`` steps).  ``jvp_max_steps`` ends the relaxed JVPs' paths after
    that many steps (None: the map's own limit): the few longer paths -- the
    trapped particles, whose tangents grow along the path -- add variance
    to the JVPs, and tracing them is the push's latency-bound tail.
    ``push_precision`` / ``green_precision`` select the map's push and
    Green's-update precision variants (as in ``anderson``); the relaxed JVP
    push is FP64 either way.
    Same return dict as ``anderson``."""
    if jvp not in ("difference", "relaxed"):
        raise ValueError(f"jvp must be 'difference' or 'relaxed': {jvp!r}")
    if jvp == "relaxed" and not 0.0 < jvp_fraction <= 1.0:
        raise ValueError("jvp_fraction must lie in (0, 1]")
    log = log or (lambda s: None)
    start = default_stages(p)[:2] if start is None else start
    noise = max(1.0, float(np.sqrt(DEFAULT_STAGES_N / max(1, p.n_per_batch))))
    tol = 0.002 * noise if tol is None else tol
    ss = _Session(p, device, setup, push_precision, green_precision, max([R] + [st[1] for st in start]))
    rho, x, evals_start, first = _secant_stages(ss, start, 10, 40, 10, log)
    start_batches = ss.traced
    box = {"like": rho}

    def g(x):
        rho_in = ss.to_rho(x, box["like"])
        rho_g, gx = ss.evaluate(rho_in, first, K, R)
        box["like"] = rho_g
        box["x"], box["rho_in"] = x.copy(), rho_in
        return rho_g, gx

    jac = None
    if jvp == "relaxed":
        from .relaxed import MapJacobian
        jac = MapJacobian(ss.s, ss.dev, width, trajectory=trajectory, max_steps=jvp_max_steps)
        n_jvp = max(1, int(round(p.n_per_batch * jvp_fraction)))   # per species
        jvp_offset = 2 * max(1, int(round(p.n_per_batch * ss.frac))) * R * K   # after the evaluations' particles
        ss.dev.reserve(jvp_offset + 2 * n_jvp)

        def jvp_fn(x, v):
            if not np.array_equal(x, box["x"]):
                raise RuntimeError("relaxed JVP requested away from the last evaluated point")
            if box.get("lin") is not box["rho_in"]:   # linearise once per point (one relaxed push)
                ss.dev.inject_device(jvp_offset, n_jvp, p.seed, (1 << 19) + first)   # the JVPs' own batch
                jac.at(box["rho_in"], 2 * n_jvp, ss.s.q_macro * p.n_per_batch / n_jvp, offset=jvp_offset)
                box["lin"] = box["rho_in"]
            return jac(v)
    else:
        jvp_fn = None

    rho_g, x, evals = newton_krylov_iterate(
        g, x, tol, h_volts, krylov, max_newton,
        log=lambda k, res, e: log(f"  [newton-krylov] step {k}: max |g(x) - x| = {res:.4g} V after {e} evaluations"),
        jvp=jvp_fn, forcing=forcing)
    rho = ss.to_rho(x, rho_g)
    solver_batches = ss.traced
    done, b2, steady, hist, rho_s = ss.plain_phase(rho, log)
    wall = time.perf_counter() - ss.t_start
    ss.close()
    how = ("forward-difference JVPs over merged replicas" if jvp == "difference" else
           f"JVPs of the relaxed-barrier map (width {width} V) over {jvp_fraction:g} of a batch")
    return dict(method=f"Newton-Krylov (GMRES) on the batch map (common random numbers, {how}) from "
                       f"{len(start)} Anderson stage(s), then plain batches to the reference criterion",
                converged=done, batches=ss.traced,
                phases=[dict(kind="secant", stages=[list(st) for st in start], evaluations_per_stage=evals_start,
                             batches=start_batches),
                        dict(kind="newton-krylov", K=K, R=R, h_volts=h_volts, krylov=krylov, forcing=forcing, jvp=jvp,
                             width=width if jac else None, trajectory=trajectory if jac else None,
                             jvp_max_steps=jvp_max_steps if jac else None,
                             relaxed_jvps=jac.jvps if jac else 0,
                             relaxed_push_ms=jac.push_ms if jac else 0.0,
                             evaluations=evals, batches=solver_batches - start_batches),
                        dict(kind="plain", batches=b2, converged=done)],
                steady=steady, hist=hist, rho_s=rho_s, wall_s=wall, setup_s=ss.setup_s)
