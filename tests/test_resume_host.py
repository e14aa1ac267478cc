"""Resume bookkeeping of the batch loop on the host: the count of passed steady
checks a resumed run starts from, and the first batch whose histories are all
known (trench_charging_2d.py:1034-1041, :1200-1222).  A resumed run may only
keep its stored count when its criterion is the stored one; otherwise it must
pass steady_hold fresh checks before it may stop."""
from dataclasses import asdict, replace

import numpy as np
import pytest

from trench01_amd import charging, model, monitors

PROBES = {"bottom center": (0, 0), "mask top": (0, 0)}


def params(**kw):
    return replace(model.Params(), until_steady=True, steady_window=30, steady_hold=2, **kw)


def histories(n, seed=0, drift=0.0):
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    hist = {"t": list((t + 1) * 1e-4),
            "bottom center": list(100.0 + drift * t + rng.normal(0, 0.01, n)),
            "mask top": list(50.0 + rng.normal(0, 0.01, n))}
    charges = {"left": list(2e-11 + rng.normal(0, 1e-15, n) + drift * 1e-12 * t),
               "right": list(-5e-11 + rng.normal(0, 1e-15, n)),
               "surface absolute": list(9e-11 + rng.normal(0, 1e-15, n))}
    return hist, charges


def stored(p, n_ok=2, **kw):
    res = dict(params=asdict(p), steady_criterion_version=charging.STEADY_CRITERION_VERSION,
               checkpoint=dict(version=1, batches_completed=90, steady_n_ok=n_ok))
    res.update(kw)
    return res


def count(res, p, start=90, hist=None, charges=None):
    if hist is None:
        hist, charges = histories(start)
    first = monitors.history_start(hist, PROBES, charges)
    return charging.resumed_steady_count(res, p, start, first, hist, PROBES, charges)


def test_same_criterion_keeps_the_stored_count():
    p = params()
    assert count(stored(p, n_ok=1), p) == 1
    assert count(stored(p, n_ok=2), p) == 2


@pytest.mark.parametrize("change", [dict(steady_rtol=2e-3), dict(steady_window=20), dict(steady_hold=3),
                                    dict(steady_charge_rtol=0.5), dict(steady_charge_atol=1e-13)])
def test_changed_steady_setting_restarts_the_count(change):
    p = params()
    assert count(stored(p), replace(p, **change)) == 0


def test_revoked_convergence_or_other_criterion_version_restarts_the_count():
    p = params()
    assert count(stored(p, convergence_note="stored charge is not saturated under the charge criterion"), p) == 0
    assert count(stored(p, steady_criterion_version=charging.STEADY_CRITERION_VERSION - 1), p) == 0


def test_result_without_checkpoint_recounts_its_passed_checks():
    p = params()
    res = stored(p)
    res["checkpoint"] = None
    assert count(res, p, start=120) == 2          # stationary histories pass at 120 and 90
    hist, charges = histories(120, drift=0.05)     # a drifting potential fails the last check
    assert count(res, p, start=120, hist=hist, charges=charges) == 0


def test_history_start_skips_the_unknown_prefix():
    hist, charges = histories(150)
    assert monitors.history_start(hist, PROBES, charges) == 0
    charges["left"][:40] = [np.nan] * 40
    hist["mask top"][:25] = [np.nan] * 25
    assert monitors.history_start(hist, PROBES, charges) == 40
    hist["t"][:100] = [np.nan] * 100               # only probe and charge histories count
    assert monitors.history_start(hist, PROBES, charges) == 40
    # checks before history_start + 2W are skipped, also when recounting
    res = stored(params())
    res["checkpoint"] = None
    assert count(res, params(), start=90, hist=hist, charges=charges) == 0
    assert count(res, params(), start=120, hist=hist, charges=charges) == 1
