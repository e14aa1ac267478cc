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
* then plain batches with fresh particles run until the reference criterion
  (window W, hold) passes; the reported steady values are their last-window
  means, as the plain method reports its own (:1342-1343).

No pathwise adjoint of the push enters: deposit counts are integers, so their
derivative with respect to the charge along a fixed particle path is zero
almost everywhere; secant information over finite steps is what carries the
response.
"""
import time

import numpy as np

from . import monitors
from .charging import ChargingSetup, DeviceCharging

DEFAULT_STAGES = ((1, 1, 0.02), (1, 4, 0.003), (1, 8, 0.002))


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


def anderson(p, stages=DEFAULT_STAGES, memory=10, max_evals=40, patience=10, log=None, device=0, setup=None,
             push_precision=0, green_precision=0):
    """Steady state by multilevel Anderson acceleration of the batch map with
    common random numbers, then plain batches to the reference criterion.
    ``stages`` lists (K, R, tol): K batches per replica, R replicas averaged
    per evaluation, stop when no surface potential moves by tol volts (or
    after ``patience`` evaluations without progress).  ``push_precision`` /
    ``green_precision`` select the precision variants of the batch (FP32
    push; BF16 / FP16 split Green's matrix).
    Returns a dict: converged, batches (all traced), phases, steady (probe ->
    last-window mean), hist (probe histories of the plain phase), rho_s,
    wall_s and setup_s."""
    log = log or (lambda s: None)
    t_start = time.perf_counter()
    s = setup or ChargingSetup(p)
    dev = DeviceCharging(s, device=device)
    dev.set_schedule(precision=push_precision)
    if green_precision:
        dev.set_green_precision(green_precision)
    probes = s.probes
    dev.set_monitors(probes, s.regions)
    n = 2 * p.n_per_batch
    dev.reserve(n)
    to_u, to_rho = potential_map(s)
    setup_s = time.perf_counter() - t_start
    traced = 0

    def evaluate(rho, first, K, R):
        nonlocal traced
        us = []
        for r in range(R):
            dev.set_rho(rho)
            for b in range(K):
                dev.inject_device(0, p.n_per_batch, p.seed, first + r * K + b)
                dev.step(0, n)
            traced += K
            rho_g = dev.read(phi=False, counts=False)[0]
            us.append(to_u(rho_g))
        return rho_g, np.mean(us, axis=0)

    pairs = []
    rho = np.zeros(dev.m)
    x = to_u(rho)
    evals_per_stage = []
    first = 0
    for K, R, tol in stages:
        box = {"like": rho}

        def g(x, first=first, K=K, R=R):
            rho_in = to_rho(x, box["like"])
            rho_g, gx = evaluate(rho_in, first, K, R)
            box["like"] = rho_g
            return rho_g, gx

        rho_g, x, e = anderson_iterate(
            g, x, pairs, tol, memory, max_evals, patience,
            log=lambda k, res, K=K, R=R: log(f"  [anderson K={K} R={R}] eval {k}: max |g(x) - x| = {res:.4g} V"))
        rho = to_rho(x, rho_g)
        evals_per_stage.append(e)
        first += K * R
    secant_batches = traced
    # plain batches with fresh particles until the reference criterion passes
    dev.set_rho(rho)
    hist = {k: [] for k in probes}
    qhist = {k: [] for k in (*s.regions, "surface absolute")}
    W = p.steady_window
    n_ok, done, b2 = 0, False, 0
    first = 1 << 20                              # batch indices away from the secant phase's
    while traced < p.max_batches:
        dev.inject_device(0, p.n_per_batch, p.seed, first + b2)
        d = dev.step(0, n)
        if d.tally.off_surface:
            raise RuntimeError(f"{d.tally.off_surface} particles were absorbed off the surface cells")
        b2 += 1
        traced += 1
        values, charges, _, _ = dev.monitors()
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
    wall = time.perf_counter() - t_start
    Wl = min(W, len(hist[next(iter(probes))]))
    steady = {k: float(np.mean(hist[k][-Wl:])) for k in probes}
    dev.close()
    return dict(method="multilevel Anderson acceleration of the batch map (common random numbers), "
                       "then plain batches to the reference criterion",
                converged=done, batches=traced,
                phases=[dict(kind="secant", stages=[list(st) for st in stages], memory=memory,
                             evaluations_per_stage=evals_per_stage, batches=secant_batches),
                        dict(kind="plain", batches=b2, converged=done)],
                steady=steady, hist={k: np.array(v) for k, v in hist.items()}, rho_s=rho_s, wall_s=wall,
                setup_s=setup_s)
