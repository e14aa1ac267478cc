"""Ensemble push on the device: every instance of a heterogeneous ensemble
gets the oracle's push results bit for bit, in every launch mode (one launch
per pass over all instances, a replayed hipGraph of the batch, one launch per
instance), with single- and multi-pass schedules and device injection."""
from dataclasses import replace

import numpy as np
import pytest

from oracle import trace_ref
from trench01_amd import configs, ensemble as E, model
from trench01_amd.tracer import HipTracer, tapered_of

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[gpu case] {name}", flush=True)


@pytest.fixture(scope="module")
def workload():
    specs = E.batching_workload(7, seed=21, particles=(2e3, 1.2e4))
    specs[0] = replace(specs[0], p=replace(specs[0].p, trench_taper_deg=0.0, mask_taper_deg=0.0))   # untapered
    fields = [E.instance_field(s) for s in specs]
    parts = []
    for k, s in enumerate(specs):
        p = s.p
        ion = dict(kind="mono", energy_eV=p.ion_energy_eV, sigma_deg=p.ion_angle_deg) \
            if p.ion_distribution == "mono" else \
            dict(kind="arcsine", e_lo=p.ion_energy_lo_eV, e_hi=p.ion_energy_hi_eV, sigma_deg=p.ion_angle_deg)
        sampler = configs.ion_velocity_sampler(p.ion_mass_amu, **ion)
        parts.append(model.inject(p, p.n_per_batch, fields[k][0].shape[1], np.random.default_rng(100 + k),
                                  ion_velocity=sampler))
    return specs, fields, parts


def params(p, pass_steps=0, blocks_per_cu=0, refill=32):
    t = HipTracer.__new__(HipTracer)   # schedule attributes only
    prm = HipTracer.params_for(t, p, True)
    prm.pass_steps = pass_steps
    prm.blocks_per_cu = blocks_per_cu
    prm.refill = refill
    prm.tapered = 1
    return prm


def oracle(specs, fields, parts):
    out = []
    for s, (vac, ex, ey, phi, wf), pt in zip(specs, fields, parts):
        p = s.p
        out.append(trace_ref.trace_arrays(*pt, ex, ey, vac, p.dx, p.cfl, p.max_steps, tapered_of(p), (phi, wf)))
    return out


def same(got, ref, status=True):
    np.testing.assert_array_equal(got["cnt_i"], ref["cnt_i"])
    np.testing.assert_array_equal(got["cnt_e"], ref["cnt_e"])
    np.testing.assert_array_equal(got["escaped"], ref["escaped"])
    assert got["lost"] == ref["lost"] and got["steps"] == ref["total_steps"]
    assert got["off_surface"] == 0
    if status:
        np.testing.assert_array_equal(got["status"], ref["status"])
        np.testing.assert_array_equal(got["hit"], np.where(ref["status"] == 1, ref["hit"], -1))


def build(specs, fields, parts):
    ens = E.Ensemble()
    for s, (vac, ex, ey, phi, wf), pt in zip(specs, fields, parts):
        ens.add(vac, ex, ey, (phi, wf), tapered_of(s.p), pt[0].size)
    for i, pt in enumerate(parts):
        ens.upload(i, *pt)
    return ens


def test_every_instance_matches_the_oracle_in_every_mode(workload):
    _case("ensemble modes")
    specs, fields, parts = workload
    ref = oracle(specs, fields, parts)
    assert any(r["status"].tolist().count(1) for r in ref)
    ens = build(specs, fields, parts)
    p = specs[0].p
    for mode, pass_steps, bpc in (("one", 0, 0), ("one", 120, 0), ("per_instance", 120, 0), ("one", 250, 4),
                                  ("per_instance", 0, 3)):
        print(f"  mode {mode} pass_steps {pass_steps} blocks_per_cu {bpc}", flush=True)
        ens.run(params(p, pass_steps, bpc), mode=mode, status=True)
        for i in range(len(specs)):
            same(ens.read(i, status=True), ref[i])
    ens.close()


def test_graph_replays_equal_eager_batches(workload):
    """A captured batch replayed several times, with eager runs and host reads
    in between, gives the eager tallies every time."""
    _case("ensemble graph")
    specs, fields, parts = workload
    ref = oracle(specs, fields, parts)
    ens = build(specs, fields, parts)
    p = specs[0].p
    prm = params(p, 120)
    for rep in range(3):
        ens.run(prm, mode="graph", batches=2)
        for i in range(len(specs)):
            got = ens.read(i)
            np.testing.assert_array_equal(got["cnt_i"], 2 * ref[i]["cnt_i"])
            np.testing.assert_array_equal(got["cnt_e"], 2 * ref[i]["cnt_e"])
            assert got["steps"] == 2 * ref[i]["total_steps"] and got["lost"] == 2 * ref[i]["lost"]
        ens.run(prm, mode="one", status=True)
        for i in range(len(specs)):
            same(ens.read(i, status=True), ref[i])
    ens.close()


def test_device_injection_batches_match_the_oracle(workload):
    """Device injection per instance (its own distributions and stream): the
    injected batch traced by the ensemble equals the oracle push of the same
    states; eager and graph-replayed batches draw the same particles; the
    draws follow the instance's layout (ions then electrons at the top)."""
    _case("ensemble injection")
    specs, fields, parts = workload
    ens = build(specs, fields, parts)
    for i, s in enumerate(specs):
        ens.set_inject(i, s.p)
    prm = params(specs[0].p, 120)
    ens.run(prm, seed=5, first_batch=7, mode="one", inject=True, status=True)
    states = [ens.read_particles(i) for i in range(len(specs))]
    ref = oracle(specs, fields, states)
    for i, s in enumerate(specs):
        same(ens.read(i, status=True), ref[i])
        x, y, vx, vy, qm, spc = states[i]
        n = s.p.n_per_batch
        nx, ny = fields[i][0].shape
        np.testing.assert_array_equal(spc, np.r_[np.zeros(n), np.ones(n)])
        np.testing.assert_array_equal(y, np.full(2 * n, ny * s.p.dx - 1e-3 * s.p.dx))
        assert x.min() >= 0 and x.max() < nx * s.p.dx
        assert np.all(vy[:n] < 0) and np.all(vy[n:] < 0)
    tallies = [ens.read(i) for i in range(len(specs))]
    ens.run(prm, seed=5, first_batch=7, mode="graph", inject=True)
    for i in range(len(specs)):
        got = ens.read(i)
        np.testing.assert_array_equal(got["cnt_i"], tallies[i]["cnt_i"])
        assert got["steps"] == tallies[i]["steps"]
        np.testing.assert_array_equal(ens.read_particles(i)[0], states[i][0])
    ens.run(prm, seed=5, first_batch=8, mode="one", inject=True)
    assert not np.array_equal(ens.read_particles(0)[0], states[0][0])
    ens.close()


def test_argument_errors(workload):
    _case("ensemble errors")
    specs, fields, parts = workload
    ens = E.Ensemble()
    vac, ex, ey, phi, wf = fields[0]
    with pytest.raises(ValueError):
        ens.add(vac, ex, ey, (phi, None), False, 10)
    ens.add(vac, ex, ey, None, False, 10)
    with pytest.raises(ValueError):
        ens.upload(0, *(a[:4] for a in parts[0]))
    prm = params(specs[0].p)
    with pytest.raises(RuntimeError):          # barrier without phi
        ens.run(prm)
    prm.barrier = 0
    with pytest.raises(RuntimeError):          # injection without distributions
        ens.run(prm, inject=True)
    with pytest.raises(ValueError):
        ens.run(prm, mode="graph", status=True)
    ens.close()


def test_autotuner_and_bulk_totals(workload):
    """The ensemble autotuner times candidate schedules on one injected batch
    and checks that they agree; the bulk readback equals the per-instance one."""
    _case("ensemble autotune")
    from trench01_amd import autotune
    specs, fields, parts = workload
    ens = build(specs, fields, parts)
    for i, s in enumerate(specs):
        ens.set_inject(i, s.p)
    prm = params(specs[0].p)
    best, table = autotune.tune_ensemble(ens, prm, space=dict(pass_steps=(0, 120), refill=(16,), check_steps=(4, 8),
                                                              blocks_per_cu=(0, 4)))
    assert len(table) == 8 and best in [t[0] for t in table]
    counts, scal = ens.totals()
    off = 0
    for i in range(len(specs)):
        r = ens.read(i)
        flat = ens.surface(i)
        m = flat.size
        np.testing.assert_array_equal(counts[0, off:off + m], r["cnt_i"].ravel()[flat])
        np.testing.assert_array_equal(counts[1, off:off + m], r["cnt_e"].ravel()[flat])
        np.testing.assert_array_equal(scal[i], [r["escaped"][0], r["escaped"][1], r["lost"], r["steps"],
                                                r["off_surface"]])
        off += m
    assert off == counts.shape[1]
    ens.close()
