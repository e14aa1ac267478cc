"""Probes, charge monitors and the steady-state criterion of the batch loop
(trench_charging_2d.py:850-947, probes :1152-1166).  Host-side: they read a
few numbers per batch and decide when the charging run has saturated."""
import numpy as np


def probes(geo, p, cond):
    """Potential probe cells (ix, iy): mask top, dielectric sidewalls at three
    heights on both sides, and the trench bottom centre (:1152-1166)."""
    top_ox = p.floor_t + p.trench_d
    left, right, _ = geo.wall_bounds()
    b0, b1 = geo.opening_span(p.floor_t)
    e0, _ = geo.opening_span(top_ox + p.mask_t - 1)
    out = {"mask top": (max(0, e0 - 10), top_ox + p.mask_t - 1)}
    if p.mask_t > 0 and not cond.any():
        ym = top_ox + p.mask_t // 2
        out["mask sidewall"] = (int(left[ym - p.floor_t]), ym)
    for level, yw in (("upper", max(p.floor_t, top_ox - 4)), ("middle", p.floor_t + p.trench_d // 2),
                      ("lower", min(top_ox - 1, p.floor_t + 3))):
        out[f"sidewall {level}"] = (int(left[yw - p.floor_t]), yw)
    out["bottom center"] = ((b0 + b1) // 2, p.floor_t - 1)
    for level in ("upper", "middle", "lower"):
        yw = out[f"sidewall {level}"][1]
        out[f"right sidewall {level}"] = (int(right[yw - p.floor_t]), yw)
    return out


def charge_regions(geo, p):
    """Cells whose raw charge is monitored (:850-868): dielectric sidewalls
    left / right (and three height bands each), the floor of the opening, all
    dielectric surface cells, and the mask."""
    exposed = geo.surface_mask()
    dielectric = exposed & ~(geo.mask if p.mask_type == "conductor" else np.zeros_like(geo.mask))
    ix, iy = np.indices(geo.solid.shape)
    wall = (iy >= p.floor_t) & (iy < p.floor_t + p.trench_d)
    inner = geo.inner_surface_mask()
    centre = (geo.i0 + geo.i1) / 2
    b0, b1 = geo.opening_span(p.floor_t)
    regions = dict(left=(ix + 0.5 < centre) & wall & inner, right=(ix + 0.5 >= centre) & wall & inner,
                   bottom=(ix >= b0) & (ix < b1) & (iy == p.floor_t - 1), surface=dielectric)
    for side in ("left", "right"):
        for band, rows in enumerate(np.array_split(np.arange(p.floor_t, p.floor_t + p.trench_d), 3)):
            if len(rows):
                regions[f"{side} band {band + 1}"] = regions[side] & np.isin(iy, rows)
    if geo.mask.any():
        regions["mask"] = geo.mask
    return {k: np.flatnonzero(r) for k, r in regions.items()}


def sample_charges(rho, regions, dx):
    """Integrated charge per region [C/m] (:871-874); ``rho`` is the full grid."""
    flat = rho.ravel()
    out = {k: float(flat[c].sum() * dx ** 2) for k, c in regions.items()}
    out["surface absolute"] = float(np.abs(flat[regions["surface"]]).sum() * dx ** 2)
    return out


def history_start(hist, probe_names, charge_monitor_hist):
    """First batch from which every probe and charge-monitor history holds
    finite values: result files of older runs lack some histories, which a
    resume fills with NaN, and the steady checks only look at the part where
    all of them are known (trench_charging_2d.py:1034-1041)."""
    first = 0
    for values in [*(hist[k] for k in probe_names), *charge_monitor_hist.values()]:
        missing = np.flatnonzero(~np.isfinite(np.asarray(values, dtype=float)))
        if missing.size:
            first = max(first, int(missing[-1]) + 1)
    return first


def mean_and_se(x):
    """Block mean and its standard error from linearly detrended residuals,
    with the effective sample size of a lag-1 autocorrelated series (:877-890)."""
    n = x.size
    t = np.arange(n)
    resid = x - np.polyval(np.polyfit(t, x, 1), t)
    var = resid.var(ddof=2)
    r1 = 0.0
    if var > 0:
        r1 = np.corrcoef(resid[:-1], resid[1:])[0, 1]
        r1 = float(np.clip(r1, 0.0, 0.95)) if np.isfinite(r1) else 0.0
    n_eff = max(n * (1 - r1) / (1 + r1), 1.0)
    return x.mean(), np.sqrt(var / n_eff)


def check_steady(hist, probe_names, p, reference=None):
    """Voltage test (:893-910): the last two W-batch block means of every probe
    agree within max(rtol * scale, 3 * SE).  Returns (ok, worst probe, d, tol)."""
    W = p.steady_window
    stats = {k: (mean_and_se(np.asarray(hist[k][-W:])), mean_and_se(np.asarray(hist[k][-2 * W:-W])))
             for k in probe_names}
    scale = max(1.0, max(abs(c[0] - (reference[k] if reference is not None else 0.0))
                         for k, (c, _) in stats.items()))
    ok, worst = True, (None, 0.0, 1.0, -1.0)
    for k, ((m_c, se_c), (m_p, se_p)) in stats.items():
        d = m_c - m_p
        tol = max(p.steady_rtol * scale, 3.0 * np.hypot(se_c, se_p))
        ok &= abs(d) <= tol
        if abs(d) / tol > worst[3]:
            worst = (k, d, tol, abs(d) / tol)
    return ok, worst[0], worst[1], worst[2]


def check_charge_steady(charge_hist, p):
    """Charge test (:913-947): per region, the block-mean difference and the net
    drift over the last 2W batches must both sit inside their tolerances."""
    W = p.steady_window
    if (not np.isfinite(p.steady_charge_rtol) or p.steady_charge_rtol <= 0
            or not np.isfinite(p.steady_charge_atol) or p.steady_charge_atol < 0
            or W < 5 or not np.isfinite(p.dt_batch) or p.dt_batch <= 0):
        raise ValueError("charge tolerances must be finite (rtol > 0, atol >= 0), steady_window >= 5, dt_batch > 0")
    if not charge_hist or any(len(q) < 2 * W for q in charge_hist.values()):
        raise ValueError("the charge criterion needs two blocks of history")
    ok, worst = True, None
    for key, values in charge_hist.items():
        q = np.asarray(values[-2 * W:], dtype=float)
        if q.ndim != 1 or not np.isfinite(q).all():
            raise ValueError("charge history must be finite")
        cur, se_cur = mean_and_se(q[-W:])
        prev, se_prev = mean_and_se(q[:W])
        difference = float(cur - prev)
        mean_tol = max(p.steady_charge_atol, p.steady_charge_rtol * max(abs(cur), abs(prev)),
                       3 * np.hypot(se_cur, se_prev))
        rate, se_rate = mean_and_se(np.diff(q))
        drift = float(rate * W)
        drift_tol = max(p.steady_charge_atol, float(3 * se_rate * W))
        for kind, d, tol in (("mean", difference, mean_tol), ("drift", drift, drift_tol)):
            ratio = abs(d) / tol if tol > 0 else (0.0 if d == 0 else np.inf)
            ok &= abs(d) <= tol
            if worst is None or ratio > worst["ratio"]:
                worst = dict(name=key, kind=kind, d=d, tol=float(tol), ratio=float(ratio),
                             rate=float(rate / p.dt_batch), unit="C/m")
    return {**worst, "ok": bool(ok)}
