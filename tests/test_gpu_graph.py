"""The device-injected batch captured as a hipGraph and replayed
(th_charging_run) against the same batches run eagerly (th_inject +
th_charging_step + monitors): every replayed batch's monitors, tallies and
update diagnostics, and the charge after each block, equal the eager ones bit
for bit, with eager work and host synchronisation between the replays, after
a change of the potential update's precision (a re-capture), and through the
batch loop charging.run."""
from dataclasses import replace

import numpy as np
import pytest

from trench01_amd import charging, configs

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


def _eager(dev, n_per, seed, first, batches):
    out = []
    for b in range(first, first + batches):
        dev.inject_device(0, n_per, seed, b)
        d = dev.step(0, 2 * n_per)
        probes, charges, ni, ne = dev.monitors()
        out.append((probes, charges, ni, ne,
                    dict(escaped=(d.tally.escaped[0], d.tally.escaped[1]), lost=d.tally.lost, steps=d.tally.steps,
                         off_surface=d.tally.off_surface),
                    dict(leak_dq=d.leak_dq, drift=d.drift, dv=d.dv, n_i_cond=d.n_i_cond, n_e_cond=d.n_e_cond)))
    return out


def _same(a, b):
    assert len(a) == len(b)
    for ra, rb in zip(a, b):
        assert ra[0] == rb[0] and ra[1] == rb[1] and ra[2] == rb[2] and ra[3] == rb[3]
        assert ra[4] == rb[4], (ra[4], rb[4])
        assert ra[5] == rb[5], (ra[5], rb[5])


def test_replayed_batches_equal_eager_batches():
    _case("graph replay vs eager batches, AR8 2e4 per species")
    p, _ = configs.product_params("AR8", n_per_batch=20_000)
    s = charging.ChargingSetup(p)
    devs = [charging.DeviceCharging(s) for _ in range(2)]
    for dev in devs:
        dev.reserve(2 * p.n_per_batch)
        dev.set_rho(np.zeros(dev.m))
        dev.set_monitors(s.probes, s.regions)
    g, e = devs
    n_per, seed = p.n_per_batch, 3
    a, _ = g.run_batches(0, n_per, seed, 0, 7)
    _same(a, _eager(e, n_per, seed, 0, 7))
    np.testing.assert_array_equal(g.read(counts=False)[0], e.read(counts=False)[0])
    # eager work on the replaying state between two replays
    _same(_eager(g, n_per, seed, 7, 2), _eager(e, n_per, seed, 7, 2))
    a, _ = g.run_batches(0, n_per, seed, 9, 5)
    _same(a, _eager(e, n_per, seed, 9, 5))
    # a different precision of the potential update is captured anew
    for dev in devs:
        dev.set_green_precision(1)
    a, _ = g.run_batches(0, n_per, seed, 14, 4)
    _same(a, _eager(e, n_per, seed, 14, 4))
    rg, phg, cg = g.read()
    re_, phe, ce = e.read()
    np.testing.assert_array_equal(rg, re_)
    np.testing.assert_array_equal(phg, phe)
    np.testing.assert_array_equal(cg, ce)
    for dev in devs:
        dev.close()


def test_batch_loop_with_replays_equals_eager_loop():
    _case("charging.run with and without replays, AR8 until steady (capped)")
    p, _ = configs.product_params("AR8", n_per_batch=10_000)
    p = replace(p, until_steady=True, max_batches=95, steady_window=30)
    a = charging.run(p, log=None, injection="device", graph=True)
    b = charging.run(p, log=None, injection="device", graph=False)
    assert len(a["hist"]["t"]) == len(b["hist"]["t"])
    for key, v in b["hist"].items():
        np.testing.assert_array_equal(a["hist"][key], v, err_msg=key)
    for key, v in b["charge_monitor_hist"].items():
        np.testing.assert_array_equal(a["charge_monitor_hist"][key], v, err_msg=key)
    np.testing.assert_array_equal(a["rho_raw"], b["rho_raw"])
    assert a["converged"] == b["converged"]
