#!/usr/bin/env python3
"""Throughput of the per-batch charging hot path on one MI355X.

Workload: the AR 8 production case -- 120x320 cells (dx 5 nm), trench 20x160
with 2 / 5 degree tapers, conductive floating mask of 40 cells, barrier wall
model, surface leak sigma_s 1e-15 S; 1e6 particles per batch (5e5 ions from
the arcsine 40-260 eV IED with a 4 degree polar spread, 5e5 electrons at
Te = 4 eV).

One step = one batch, entirely on the device: injection (Philox), FP64 push
with LDS deposit tallies, deposit, leak, floating-conductor step, potential
(Green's matrix GEMV) and field, plus the batch's diagnostics read back for
the host loop (tallies, probe potentials and region charges: what the
reference's loop records per batch, trench_charging_2d.py:1271-1285).  Timed from a pre-charged state (the same loop run first with
2e4 particles per species per batch, which deposits the same charge per
batch), so path lengths are those of a charged trench.

Prints one JSON line: the metric, the roofline of the push kernel (FP64
FLOP at 69 per particle-step over its HIP-event time), and the CPU baseline
(the oracle's C push on this host's cores over one whole batch, same states,
same field).  The second half of the metric, time-to-steady at the same
workload, is measured too (``--no-steady`` skips it): plain time-marching from
zero charge to the reference's steady criterion, and the chosen method
(trench01_amd.steady.newton_krylov: two Anderson stages, then a Newton step
whose GMRES Jacobian-vector products come from the relaxed-barrier push, all
with common random numbers, then plain batches to the same criterion), with
the bottom-centre potential of each; trench01_amd.steady.anderson alone is
measured beside it.
"""
import argparse
import json
import os
import sys
import time
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_STEP = 69          # algorithmic FP64 FLOP per particle-step
BYTES_PER_PARTICLE = 78     # algorithmic HBM bytes per particle (state in/out, species, status, hit)
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector (= FP64 matrix) peak
FP32_PEAK_TFLOPS = 157.3    # MI355X FP32 vector peak
HBM_PEAK_GBS = 8000.0
CPU_SAMPLE_S = 10.0         # minimum CPU work behind the cpu_baseline figure
# HBM bytes of the push per batch at this workload, measured with rocprofv3
# (FETCH_SIZE / WRITE_SIZE per dispatch, summed over the passes of one batch,
# scaled by the copy-kernel calibration); the command that made it is stored
# with it.
TRAFFIC_FILES = {0: os.path.join(ROOT, "profiles", "push64_ar8_1e6_traffic.json"),
                 1: os.path.join(ROOT, "profiles", "push32_ar8_1e6_traffic.json")}
ENSEMBLE_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "ensemble_push_256_traffic.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-per-batch", type=int, default=500_000, help="particles per species per batch")
    ap.add_argument("--precharge", type=int, default=600, help="pre-charge batches (2e4 per species)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: all cores this process may use")
    ap.add_argument("--no-steady", action="store_true", help="skip the time-to-steady measurement")
    ap.add_argument("--max-batches", type=int, default=5000)
    ap.add_argument("--steady-seeds", default="0,1,2,3", help="seeds of the time-to-steady comparison")
    ap.add_argument("--long-run", type=int, default=2000,
                    help="plain batches run on past the first seed's plain convergence; the mean over the last "
                         "three quarters is the long-run time-marched value (0: skip)")
    ap.add_argument("--no-host-array", action="store_true", help="skip the host-array (PCIe-inclusive) tracer rate")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 1), help="0: FP64 push, 1: FP32 push")
    ap.add_argument("--green", type=int, default=0, choices=(0, 1, 2, 3, 4),
                    help="potential update: 0 FP64 Green's matrix, 1 BF16 hi/lo split, 2 FP16 hi/lo split, "
                         "3 / 4 the same splits on MFMA")
    ap.add_argument("--pass-steps", type=int, default=None,
                    help="multi-pass schedule's first cap (0: one pass); default: the library's tuned value")
    ap.add_argument("--tile-pass", type=int, default=None,
                    help="first pass (1-based) of the push that stages the trench's field in LDS (0: none); "
                         "default: the library's tuned value")
    ap.add_argument("--batching", type=int, default=256, metavar="INSTANCES",
                    help="instances of the heterogeneous batching workload (0: skip)")
    ap.add_argument("--batching-batches", type=int, default=5)
    ap.add_argument("--batching-schedule", default=None, metavar="JSON",
                    help="ensemble schedule instead of the autotuner's pick, e.g. '{\"pass_steps\": 0}'")
    ap.add_argument("--differentiation", action="store_true",
                    help="also time the tangent push and compare pathwise derivatives with differences")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (traffic-counter calibration)")
    return ap.parse_args()


def copy_probe(mib):
    import ctypes
    from trench01_amd import HipTracer, _lib
    t = HipTracer(np.ones((4, 4), bool))
    ms = ctypes.c_double(0.0)
    nbytes = mib << 20
    _lib.check(_lib.load().th_probe_copy(t._h, nbytes, 5, ctypes.byref(ms)))
    print(json.dumps({"copy_probe_bytes": nbytes, "avg_kernel_ms": ms.value,
                      "gbs": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)


def precharge(p, batches, log):
    """Charge the trench with small batches (same charge per batch) and return rho on surface cells."""
    from trench01_amd import charging
    pc = replace(p, n_per_batch=20_000)
    dev = charging.DeviceCharging(charging.ChargingSetup(pc))
    dev.set_rho(np.zeros(dev.m))
    dev.reserve(2 * pc.n_per_batch)
    t0 = time.perf_counter()
    dev.run_batches(0, pc.n_per_batch, 1000 + p.seed, 0, batches)     # replays of the captured batch
    rho, phi, _ = dev.read(counts=False)
    log(f"pre-charge: {batches} batches of {2 * pc.n_per_batch} in {time.perf_counter() - t0:.2f}s")
    dev.close()
    dev.tracer.close()
    return rho


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; this package has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    if a.copy_probe:
        copy_probe(a.copy_probe)
        return
    from trench01_amd import charging, configs
    log = lambda s: print(s, file=sys.stderr, flush=True)
    p, _ = configs.product_params("AR8", n_per_batch=a.n_per_batch)
    rho0 = precharge(p, a.precharge, log) if a.precharge > 0 else None
    setup = charging.ChargingSetup(p)
    dev = charging.DeviceCharging(setup)
    dev.set_schedule(precision=a.precision)
    if a.pass_steps is not None:
        dev.set_schedule(pass_steps=a.pass_steps)
    if a.tile_pass is not None:
        dev.set_schedule(tile_pass=a.tile_pass)
    m = dev.m
    dev.set_rho(rho0 if rho0 is not None else np.zeros(m))
    precision = precision_errors(dev, setup, a.green) if (a.green or a.precision) else None
    n = 2 * p.n_per_batch
    dev.reserve(n)
    dev.set_monitors(setup.probes, setup.regions)   # what the batch loop records per batch
    for b in range(a.warmup):
        dev.inject_device(0, p.n_per_batch, p.seed, b)
        dev.step(0, n)
        dev.monitors()
    # The first timed batch, traced once beforehand to keep its inputs and the
    # device's tallies for the parity check and the CPU baseline; the charge
    # is then put back, so the timed loop starts from the same state.
    rho_before = dev.read(phi=False, counts=False)[0]
    dev.inject_device(0, p.n_per_batch, p.seed, a.warmup)
    cpu_inputs = (dev.tracer.read_particles(0, n), dev.tracer.read_field(), dev.read(counts=False)[1])
    d_first = dev.step(0, n)
    gpu_first = dict(counts=dev.read(phi=False)[2], escaped=list(d_first.tally.escaped), lost=d_first.tally.lost,
                     steps=d_first.tally.steps, off_surface=d_first.tally.off_surface)
    dev.set_rho(rho_before)
    host_array = host_array_rate(p, setup, *cpu_inputs) if not a.no_host_array else None
    steps_total = 0
    push_ms = update_ms = 0.0
    t0 = time.perf_counter()
    for b in range(a.warmup, a.warmup + a.steps):
        dev.inject_device(0, p.n_per_batch, p.seed, b)
        d = dev.step(0, n)
        dev.monitors()
        steps_total += d.tally.steps
        push_ms += d.push_ms
        update_ms += d.update_ms
    wall = time.perf_counter() - t0
    value = a.steps * n / wall
    # The same loop with each batch replayed from its captured hipGraph and the
    # monitors read back once (what the batch loop does between steady checks).
    dev.run_batches(0, p.n_per_batch, p.seed, 10_000, 1)
    t1 = time.perf_counter()
    _, graph_ms = dev.run_batches(0, p.n_per_batch, p.seed, 10_001, a.steps)
    graph_wall = time.perf_counter() - t1
    achieved = FLOP_PER_STEP * steps_total / (push_ms * 1e-3) / 1e12
    peak = FP64_PEAK_TFLOPS if a.precision == 0 else FP32_PEAK_TFLOPS
    traffic, traffic_src = measured_traffic(a.precision)
    out = {
        "metric": "traced particles/s/GPU + time-to-steady-charge at AR=8 (plain vs chosen)",
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if a.precision == 0 else "f32",
        "data": "synthetic (device-injected particles, Philox)",
        "config": {"workload": f"AR8 trench 120x320 (dx 5 nm), trench 20x160 with 2/5 deg tapers, conductive "
                               f"floating mask 40, barrier wall, sigma_s 1e-15 S; ions arcsine IED 40-260 eV with "
                               f"4 deg polar sigma, electrons isotropic Te 4 eV; 1e6 particles per batch "
                               f"(5e5 per species), seed {p.seed}; full batch on device (inject, "
                               f"{'FP64' if a.precision == 0 else 'FP32'} push, deposit, leak, conductor, "
                               f"{GREEN_NAMES[a.green]} Green's potential, field); multi-pass push schedule",
                   "particles_per_batch": n, "grid": [setup.geo.nx, setup.geo.ny], "surface_cells": m,
                   "green_rows_cols": list(setup.green.shape), "precharge_batches": a.precharge,
                   "max_steps": p.max_steps, "mean_steps_per_particle": steps_total / (a.steps * n),
                   "push_ms_per_batch": push_ms / a.steps, "update_ms_per_batch": update_ms / a.steps},
        "roofline": {"bound": "fp64-valu" if a.precision == 0 else "fp32-valu", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "push64_kernel" if a.precision == 0 else "push32_kernel",
                     "kernel_ms_per_batch": push_ms / a.steps,
                     "launches_per_batch": len(pass_caps(dev.prm.pass_steps, p.max_steps)),
                     "flop_per_particle_step": FLOP_PER_STEP,
                     "note": "vector-ALU roof of the push's precision (FP64 vector 78.6 TFLOP/s; FP32 vector "
                             "157.3); the push is ALU/latency bound, not HBM: algorithmic HBM traffic is 78 B/particle; "
                             "one batch's push is launches_per_batch stream-ordered launches (multi-pass schedule), "
                             "timed together with HIP events on the library stream; traffic is HBM bytes per batch",
                     "hbm_gbs_algorithmic": BYTES_PER_PARTICLE * n / (push_ms / a.steps * 1e-3) / 1e9},
    }
    out["graph_replay"] = {"value": a.steps * n / graph_wall, "unit": "particles/s",
                           "ms_per_step": graph_wall / a.steps * 1e3, "gpu_ms_per_step": graph_ms / a.steps,
                           "note": "the timed loop's batches replayed from one captured hipGraph "
                                   "(th_charging_run), monitors read back after the last"}
    if host_array is not None:
        out["host_array_tracer"] = host_array
    if precision is not None:
        out["precision"] = precision
    if not a.no_cpu:
        out["cpu_baseline"], out["parity"] = cpu_baseline(p, setup, *cpu_inputs, a.cpu_threads, gpu_first)
    if not a.no_steady:
        out["time_to_steady"] = time_to_steady(p, a.max_batches, log, a.precision, a.green,
                                               [int(x) for x in a.steady_seeds.split(",")], a.long_run)
        if "cpu_baseline" in out:
            # the plain run's batch count at the CPU push rate (push only; extrapolated, not run)
            nb = out["time_to_steady"]["plain"]["batches"]
            out["time_to_steady"]["plain_cpu_extrapolated"] = {
                "wall_s": nb * n / out["cpu_baseline"]["value"], "cores": out["cpu_baseline"]["cores"],
                "note": "extrapolated: plain batches x particles per batch / cpu_baseline rate (push only)"}
    if a.differentiation:
        out["differentiation"] = differentiation(p, dev, rho0 if rho0 is not None else np.zeros(m), push_ms / a.steps)
    if a.batching > 0:
        dev.close()
        dev.tracer.close()
        out["batching"] = batching(a.batching, a.batching_batches, log,
                                   json.loads(a.batching_schedule) if a.batching_schedule else None)
    print(json.dumps(out), flush=True)


def differentiation(p, dev, rho, push_ms):
    """The pathwise derivative at the timed workload's pre-charged state: the
    tangent push over one batch of it (1e6 particles, device-injected, single
    pass, along the surface charge's own direction) next to the plain push,
    and the agreement of pathwise derivatives with central differences of the
    batch increment (tangent.derivative_agreement, 5e5 particles per batch)."""
    import ctypes
    from trench01_amd import _lib, tangent
    n = 2 * p.n_per_batch
    dev.inject_device(0, p.n_per_batch, p.seed, 0)
    parts = dev.tracer.read_particles(0, n)
    b = tangent.PathwiseBatch(p, seed=7)
    phi, ex, ey = b.field(rho)
    drho = np.zeros(b.geo.solid.size)
    drho[b.surface] = rho
    _, dex, dey = tangent.charge_direction(b.fields, b.geo.vac, p, drho.reshape(b.geo.solid.shape))
    times = []
    for _ in range(3):
        b.tracer.tangent(*parts, ex, ey, p, (phi, b.fields.wface), None, (dex, dey))
        ms = ctypes.c_double(0.0)
        _lib.check(_lib.load().th_last_push_ms(b.tracer._h, ctypes.byref(ms)))
        times.append(ms.value)
    # the relaxed-barrier push (width 1 V) with field and potential tangents on the same batch
    dphi, _, _ = tangent.charge_direction(b.fields, b.geo.vac, p, drho.reshape(b.geo.solid.shape))
    b.tracer.reserve(n)
    b.tracer.upload(0, *parts)
    b.tracer.set_fields(ex, ey, (phi, b.fields.wface))
    soft_times = [b.tracer.soft(0, n, p, 1.0, dfield=(dex, dey), dphi=dphi)["push_ms"] for _ in range(3)]
    b.close()
    # reverse mode: one primal trace + one adjoint push give <g, increment>'s
    # gradient for every surface cell and both ion parameters
    vb = tangent.PathwiseBatch(p, seed=7)
    g = np.random.default_rng(1).standard_normal(vb.surface.size)
    vb.vjp(rho, g)                                   # warm-up
    t0 = time.perf_counter()
    vb.vjp(rho, g)
    vjp_s = time.perf_counter() - t0
    adjoint_ms = vb.tracer.last_adjoint_ms
    t0 = time.perf_counter()
    vb.charge_jvp(rho, rho)
    jvp_s = time.perf_counter() - t0
    # the relaxed-barrier push's reverse mode: the same gradient for the relaxed increment
    vb.relaxed_vjp(rho, g, 1.0)                      # warm-up
    t0 = time.perf_counter()
    vb.relaxed_vjp(rho, g, 1.0)
    relaxed_vjp_s = time.perf_counter() - t0
    relaxed_adjoint_ms = vb.tracer.last_adjoint_ms
    m = int(vb.surface.size)
    vb.close()
    agreement = tangent.derivative_agreement(replace(p, n_per_batch=250_000), rho, relaxed_width=1.0)
    return {"tangent_push_ms": min(times), "plain_push_ms": push_ms, "particles": n,
            "kernel": "push64_tangent_kernel", "agreement": agreement,
            "relaxed_push_ms": min(soft_times), "relaxed_kernel": "push64_soft_kernel",
            "reverse_mode": {"vjp_s": vjp_s, "adjoint_kernel_ms": adjoint_ms, "jvp_s": jvp_s, "surface_cells": m,
                             "m_jvp_loop_s_estimate": m * jvp_s, "relaxed_vjp_s": relaxed_vjp_s,
                             "relaxed_adjoint_kernel_ms": relaxed_adjoint_ms,
                             "note": "PathwiseBatch.vjp (primal trace, face-deposit VJP, th_trace_adjoint, field "
                                     "stencil and G^T on the host) against one forward-mode JVP; the backward that "
                                     "the forward mode alone gives takes one JVP per surface cell"}}


def batching(count, batches, log, schedule=None):
    """The heterogeneous batching workload (trench01_amd.ensemble.batching_workload,
    seed 0): ``count`` instances in one persistent launch per pass, schedule
    picked by the autotuner, device injection per instance; timed as eager
    launches, as a replayed hipGraph of the batch, and as one launch per
    instance and pass (the one-geometry-per-call schedule).  ``schedule``
    (a dict of schedule fields) replaces the autotuner's pick."""
    from concurrent.futures import ThreadPoolExecutor
    from trench01_amd import autotune, ensemble as E
    from trench01_amd.tracer import HipTracer
    t0 = time.perf_counter()
    specs = E.batching_workload(count, seed=0)
    with ThreadPoolExecutor(max(1, min(16, len(os.sched_getaffinity(0))))) as ex:
        fields = list(ex.map(E.instance_field, specs))
    ens = E.build_ensemble(specs, fields)
    setup_s = time.perf_counter() - t0
    n = sum(2 * s.p.n_per_batch for s in specs)
    cells = sum(f[0].size for f in fields)
    prm = HipTracer.params_for(HipTracer.__new__(HipTracer), specs[0].p, True)
    log(f"batching: {count} instances, {n} particles per batch, {cells} cells; set-up {setup_s:.1f}s; tuning")
    if schedule is None:
        best, table = autotune.tune_ensemble(ens, prm, log=log)
    else:
        best, table = dict(schedule), []
    autotune.apply(prm, best)
    ens.run(prm, seed=1, first_batch=0, batches=1, inject=True)          # warm-up
    ms = {}
    for mode, nb in (("one", batches), ("graph", batches), ("per_instance", max(1, batches // 2))):
        ms[mode] = ens.run(prm, seed=1, first_batch=100, batches=nb, inject=True, mode=mode) / nb
        log(f"  {mode}: {ms[mode]:.1f} ms per batch")
    _, scal = ens.totals()
    steps = int(scal[:, 3].sum())
    nb_last = max(1, batches // 2)
    parity = ensemble_parity(ens, specs, fields, prm, log)
    ens.close()
    steps_per_batch = steps / nb_last
    achieved = FLOP_PER_STEP * steps_per_batch / (ms["one"] * 1e-3) / 1e12
    try:
        with open(ENSEMBLE_TRAFFIC_FILE) as f:
            t = json.load(f)
        traffic = t["fetch_bytes_per_batch"] + t["write_bytes_per_batch"] if count == 256 else None
    except (OSError, ValueError, KeyError):
        traffic = None
    return {"workload": f"{count} heterogeneous trench instances (width 10-40 cells, AR 2-20, mask 0-40 "
                        "conductive or dielectric eps 3, tapers -3..+5 deg, DC or bimodal IED 50-300 eV; "
                        "particles per batch log-uniform 1e4-1e6 per instance), charged floating fields, FP64",
            "instances": count, "particles_per_batch": n, "cells": cells,
            "mean_steps_per_particle": steps / (nb_last * n),
            "schedule": best, "tuned_candidates": len(table),
            "ms_per_batch": ms, "particles_per_s": {k: n / (v * 1e-3) for k, v in ms.items()},
            "speedup_vs_per_instance": ms["per_instance"] / ms["one"],
            "graph_vs_eager": ms["one"] / ms["graph"], "setup_s": setup_s, "parity": parity,
            "roofline": {"bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                         "traffic_source": os.path.relpath(ENSEMBLE_TRAFFIC_FILE, ROOT) if traffic else None,
                         "kernel": "ensemble_push_kernel",
                         "note": "FP64 FLOP at 69 per particle-step over the eager batch time (one push launch plus "
                                 "injection, counter clears and the end-record reduction, about 1% of it)"}}


def ensemble_parity(ens, specs, fields, prm, log, sample=4, max_particles=400_000):
    """One more device-injected batch of the whole ensemble at full size: every
    instance's accounting (absorbed + escaped + lost = its particle count, none
    absorbed off the surface cells), and a sample of instances -- the largest
    aspect ratio, the smallest and largest particle counts up to
    ``max_particles``, and one drawn at random -- traced again by the oracle's
    C push from the batch's own initial states and compared tally by tally."""
    from oracle import trace_ref
    from trench01_amd.tracer import tapered_of
    ens.run(prm, seed=7, first_batch=0, batches=1, inject=True, mode="one")
    counts, scal = ens.totals()
    sizes = np.array([2 * s.p.n_per_batch for s in specs])
    offs = np.cumsum([0] + [ens.surface(i).size for i in range(len(specs))])
    absorbed = np.array([counts[:, offs[i]:offs[i + 1]].sum() for i in range(len(specs))])
    accounted = absorbed + scal[:, 0] + scal[:, 1] + scal[:, 2] + scal[:, 4]
    accounting_ok = bool(np.array_equal(accounted, sizes) and not scal[:, 4].any())
    small = [i for i in range(len(specs)) if sizes[i] <= max_particles]
    rng = np.random.default_rng(5)
    pick = {int(np.argmax([s.p.trench_d / s.p.trench_w if sizes[i] <= max_particles else 0
                           for i, s in enumerate(specs)])),
            int(min(small, key=lambda i: sizes[i])), int(max(small, key=lambda i: sizes[i]))}
    pick.add(int(rng.choice([i for i in small if i not in pick])))
    cores = len(os.sched_getaffinity(0))
    checked = []
    for i in sorted(pick)[:sample]:
        p = specs[i].p
        vac, ex, ey, phi, wf = fields[i]
        got = ens.read(i)
        ref = trace_ref.trace_arrays(*ens.read_particles(i), ex, ey, vac, p.dx, p.cfl, p.max_steps, tapered_of(p),
                                     (phi, wf), threads=cores)
        ok = (np.array_equal(got["cnt_i"], ref["cnt_i"]) and np.array_equal(got["cnt_e"], ref["cnt_e"])
              and np.array_equal(got["escaped"], ref["escaped"]) and got["lost"] == ref["lost"]
              and got["steps"] == ref["total_steps"])
        checked.append({"instance": i, "particles": int(sizes[i]), "aspect_ratio": p.trench_d / p.trench_w,
                        "bit_exact": bool(ok)})
    log(f"batching parity: accounting {'ok' if accounting_ok else 'FAILED'} over {len(specs)} instances; "
        f"oracle sample {[c['instance'] for c in checked]} bit-exact {all(c['bit_exact'] for c in checked)}")
    return {"accounting_ok": accounting_ok, "instances": len(specs), "particles": int(sizes.sum()),
            "oracle_sample": checked, "bit_exact": accounting_ok and all(c["bit_exact"] for c in checked)}


GREEN_NAMES = {0: "FP64", 1: "BF16 hi+lo, FP32 sums", 2: "FP16 hi+lo (scaled), FP32 sums",
               3: "BF16 hi+lo on MFMA (charges split too), FP32 accumulation",
               4: "FP16 hi+lo (scaled) on MFMA (charges split too), FP32 accumulation"}


def precision_errors(dev, setup, green):
    """Potential and field of the pre-charged state with the FP64 Green's
    matrix against the selected split (left selected)."""
    dev.set_green_precision(0)
    _, phi64, _ = dev.read(counts=False)
    ex64, ey64 = dev.tracer.read_field()
    dev.set_green_precision(green)
    _, phi, _ = dev.read(counts=False)
    ex, ey = dev.tracer.read_field()
    rows = setup.rows
    emax = max(np.abs(ex64).max(), np.abs(ey64).max())
    return {"green": GREEN_NAMES[green], "potential_max_abs_err_V": float(np.abs(phi[rows] - phi64[rows]).max()),
            "potential_max_V": float(np.abs(phi64[rows]).max()),
            "field_max_rel_err": float(max(np.abs(ex - ex64).max(), np.abs(ey - ey64).max()) / emax)}


def pass_caps(pass_steps, max_steps):
    """Step caps of the multi-pass schedule (one launch each)."""
    if pass_steps <= 0 or pass_steps >= max_steps:
        return [max_steps]
    caps, cap = [], pass_steps
    while cap < max_steps and len(caps) < 31:
        caps.append(cap)
        cap = min(2 * cap, max_steps)
    return caps + [max_steps]


def measured_traffic(precision):
    """HBM bytes per batch of the push (FP64 or FP32) from the committed
    rocprofv3 counter summary, with its path (None, None without one)."""
    path = TRAFFIC_FILES.get(precision)
    if path is None or not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    return t["fetch_bytes_per_batch"] + t["write_bytes_per_batch"], os.path.relpath(path, ROOT)


def host_array_rate(p, setup, parts, field, phi, calls=5):
    """The tracer call as the reference defines it: host arrays in (particle
    states and fields uploaded over PCIe), push, per-cell counts back on the
    host (HipTracer.__call__, the CudaTracer signature), on the first timed
    batch's particles and field; the best of ``calls`` calls after a warm-up."""
    from trench01_amd import HipTracer
    geo = setup.geo
    ex, ey = field
    barrier = (phi.reshape(geo.solid.shape), setup.fields.wface)
    t = HipTracer(geo.vac)
    t(*parts, ex, ey, geo.vac, p, barrier)
    best = np.inf
    for _ in range(calls):
        t0 = time.perf_counter()
        ci, ce, esc, lost = t(*parts, ex, ey, geo.vac, p, barrier)
        best = min(best, time.perf_counter() - t0)
    t.close()
    n = parts[0].size
    return {"value": n / best, "unit": "particles/s", "ms_per_call": best * 1e3, "particles": n,
            "note": "HipTracer called with host arrays (x, y, vx, vy, qm, spc, Ex, Ey, phi, wface uploaded; counts "
                    "read back), the reference tracer call's definition of particles/s; value above is the "
                    "device-resident batch"}


def cpu_baseline(p, setup, parts, field, phi, threads, gpu=None):
    """The oracle's C push (OpenMP) over whole batches of the timed workload
    (same particle states and field as the GPU's first timed batch), repeated
    until at least CPU_SAMPLE_S seconds of work.  With ``gpu`` (the device's
    tallies of that batch), also returns the parity of the two: per-cell ion
    and electron counts on every surface cell, escapes per species, losses and
    total particle-steps."""
    from oracle import trace_ref
    geo = setup.geo
    cores = threads or len(os.sched_getaffinity(0))
    trace_ref.lib()
    ex, ey = field
    x, y, vx, vy, qm, spc = parts
    reps, steps = 0, 0
    t0 = time.perf_counter()
    first = None
    while reps == 0 or time.perf_counter() - t0 < CPU_SAMPLE_S:
        r = trace_ref.trace_arrays(x, y, vx, vy, qm, spc, ex, ey, geo.vac, p.dx, p.cfl, p.max_steps,
                                   True, (phi.reshape(geo.solid.shape), setup.fields.wface), threads=cores)
        first = first or r
        reps += 1
        steps += r["total_steps"]
    wall = time.perf_counter() - t0
    # one thread, on every 64th particle of the same batch (same species mix and path lengths)
    sub = [a[::64] for a in (x, y, vx, vy, qm, spc)]
    reps1, steps1 = 0, 0
    t1 = time.perf_counter()
    while reps1 == 0 or time.perf_counter() - t1 < CPU_SAMPLE_S / 2:
        r1 = trace_ref.trace_arrays(*sub, ex, ey, geo.vac, p.dx, p.cfl, p.max_steps, True,
                                    (phi.reshape(geo.solid.shape), setup.fields.wface), threads=1)
        reps1 += 1
        steps1 += r1["total_steps"]
    wall1 = time.perf_counter() - t1
    parity = None
    if gpu is not None:
        r0 = first
        m = geo.surface.size
        ci = r0["cnt_i"].ravel()
        ce = r0["cnt_e"].ravel()
        off = int(ci.sum() + ce.sum() - ci[geo.surface].sum() - ce[geo.surface].sum())
        checks = {"counts_ions": bool(np.array_equal(gpu["counts"][:m], ci[geo.surface])),
                  "counts_electrons": bool(np.array_equal(gpu["counts"][m:], ce[geo.surface])),
                  "escaped": [int(v) for v in r0["escaped"]] == [int(v) for v in gpu["escaped"]],
                  "lost": int(r0["lost"]) == int(gpu["lost"]),
                  "steps": int(r0["total_steps"]) == int(gpu["steps"]),
                  "off_surface": off == 0 and int(gpu["off_surface"]) == 0}
        parity = {"bit_exact": all(checks.values()), "checks": checks, "particles": int(x.size),
                  "absorbed": int(ci.sum() + ce.sum()), "escaped": [int(v) for v in r0["escaped"]],
                  "lost": int(r0["lost"]), "particle_steps": int(r0["total_steps"]),
                  "cells_differing": int(np.count_nonzero(gpu["counts"][:m] != ci[geo.surface])
                                         + np.count_nonzero(gpu["counts"][m:] != ce[geo.surface])),
                  "against": "oracle C push (FP64, contraction off, glibc hypot) on the same initial states and "
                             "field as the device's first timed batch"}
    base = {"value": reps * x.size / wall, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"{reps} pass(es) over one whole batch ({x.size} particles, {r['total_steps']} "
                      f"particle-steps each) of the timed workload, push only, {wall:.1f} s",
            "single_thread": {"value": reps1 * sub[0].size / wall1, "unit": "particles/s", "cores": 1,
                              "sample": f"{reps1} pass(es) over every 64th particle of the same batch "
                                        f"({sub[0].size} particles, {r1['total_steps']} particle-steps each), "
                                        f"{wall1:.1f} s"}}
    return base, parity


def time_to_steady(p, max_batches, log, push_precision=0, green_precision=0, seeds=(0, 1, 2, 3), long_run=300):
    """Time-to-steady from zero charge at this workload (device injection,
    1e6 particles per batch), per seed: plain time-marching until the
    reference's criterion passes (W = 30, hold = 2), the chosen method
    (Newton-Krylov on the batch map with Jacobian-vector products of the
    relaxed-barrier push, from two Anderson stages, then plain batches to the
    same criterion) and the Anderson solve alone, with the bottom-centre
    potential of each.  The first seed's plain run is then continued
    ``long_run`` batches: the mean over the last three quarters is the
    long-run time-marched value (the plain run stops while still drifting up
    towards it).  Newton-Krylov with difference JVPs, and the chosen method
    with the precision variants (FP32 push, BF16 split Green's update, unless
    others are given), run on the first seed.  Wall times include each
    method's host set-up."""
    from trench01_amd import charging, steady
    per_seed, tails = [], []
    for seed in seeds:
        ps = replace(p, seed=seed, until_steady=True, max_batches=max_batches)
        t0 = time.perf_counter()
        res = charging.run(ps, log=None, injection="device")
        wall = time.perf_counter() - t0
        nb = res["timing"]["batches"]
        plain = {"method": "plain time-marching", "converged": bool(res["converged"]), "batches": nb, "wall_s": wall,
                 "setup_s": res["timing"]["setup_s"], "bottom_center_V": res["steady"]["bottom center"],
                 "t_conv_s": res["t_conv"]}
        if long_run > 0 and not tails:
            # one long time-marched chain: the first seed's plain run continued
            # long_run batches; the first quarter is left to relax
            more = charging.run(replace(ps, until_steady=False, n_batches=long_run), log=None, injection="device",
                                resume=res)
            tails.append(np.asarray(more["hist"]["bottom center"][nb + long_run // 4:]))
        log(f"time-to-steady seed {seed}, plain: converged={res['converged']} after {nb} batches in {wall:.1f}s, "
            f"bottom centre {plain['bottom_center_V']:.3f} V")
        r = steady.newton_krylov(ps)
        chosen = {"method": r["method"], "converged": bool(r["converged"]), "batches": r["batches"],
                  "wall_s": r["wall_s"], "setup_s": r["setup_s"], "phases": r["phases"],
                  "bottom_center_V": r["steady"]["bottom center"]}
        log(f"time-to-steady seed {seed}, chosen (newton-krylov): converged={r['converged']} after {r['batches']} "
            f"batches in {r['wall_s']:.1f}s, bottom centre {chosen['bottom_center_V']:.3f} V")
        a = steady.anderson(ps)
        anderson = {"method": a["method"], "converged": bool(a["converged"]), "batches": a["batches"],
                    "wall_s": a["wall_s"], "setup_s": a["setup_s"], "phases": a["phases"],
                    "bottom_center_V": a["steady"]["bottom center"], "speedup_vs_plain": wall / a["wall_s"],
                    "bottom_center_diff_V": a["steady"]["bottom center"] - plain["bottom_center_V"]}
        log(f"time-to-steady seed {seed}, anderson: converged={a['converged']} after {a['batches']} batches in "
            f"{a['wall_s']:.1f}s, bottom centre {anderson['bottom_center_V']:.3f} V")
        per_seed.append({"seed": seed, "plain": plain, "chosen": chosen, "speedup": wall / r["wall_s"],
                         "bottom_center_diff_V": chosen["bottom_center_V"] - plain["bottom_center_V"],
                         "anderson": anderson})
    out = {"seeds": per_seed, "plain": per_seed[0]["plain"], "chosen": per_seed[0]["chosen"],
           "speedup": per_seed[0]["speedup"], "bottom_center_diff_V": per_seed[0]["bottom_center_diff_V"],
           "speedup_min": min(e["speedup"] for e in per_seed),
           "speedup_mean": float(np.mean([e["speedup"] for e in per_seed])),
           "max_abs_diff_vs_plain_V": max(abs(e["bottom_center_diff_V"]) for e in per_seed),
           "all_converged": all(e["plain"]["converged"] and e["chosen"]["converged"] for e in per_seed)}
    if tails:
        tail = tails[0]
        # standard error of the mean of an autocorrelated series by batch means (10 blocks)
        blocks = np.array([b.mean() for b in np.array_split(tail, 10)])
        long = {"value_V": float(tail.mean()), "batches_averaged": int(tail.size), "seed": seeds[0],
                "batches_after_plain_stop": [int(long_run // 4), int(long_run)],
                "se_V": float(blocks.std(ddof=1) / np.sqrt(blocks.size))}
        out["long_run_time_marched"] = long
        for e in per_seed:
            e["chosen_minus_long_run_V"] = e["chosen"]["bottom_center_V"] - long["value_V"]
            e["plain_minus_long_run_V"] = e["plain"]["bottom_center_V"] - long["value_V"]
            e["anderson"]["minus_long_run_V"] = e["anderson"]["bottom_center_V"] - long["value_V"]
        out["max_abs_diff_vs_long_run_V"] = max(abs(e["chosen_minus_long_run_V"]) for e in per_seed)
    an = [e["anderson"] for e in per_seed]
    out["anderson"] = {
        "method": an[0]["method"], "worst_speedup": min(r["speedup_vs_plain"] for r in an),
        "mean_speedup": float(np.mean([r["speedup_vs_plain"] for r in an])),
        "max_abs_diff_vs_plain_V": max(abs(r["bottom_center_diff_V"]) for r in an),
        "max_abs_diff_vs_long_run_V": (max(abs(r["minus_long_run_V"]) for r in an)
                                       if "long_run_time_marched" in out else None),
        "all_converged": all(r["converged"] for r in an)}
    ps = replace(p, seed=seeds[0], until_steady=True, max_batches=max_batches)
    nk = steady.newton_krylov(ps, R=8, jvp="difference")
    out["newton_krylov_difference"] = {
        "seed": seeds[0], "method": nk["method"], "converged": bool(nk["converged"]), "batches": nk["batches"],
        "wall_s": nk["wall_s"], "phases": nk["phases"], "bottom_center_V": nk["steady"]["bottom center"],
        "speedup_vs_plain": per_seed[0]["plain"]["wall_s"] / nk["wall_s"]}
    log(f"time-to-steady, newton-krylov (difference JVPs): {nk['wall_s']:.1f}s, bottom centre "
        f"{nk['steady']['bottom center']:.3f} V")
    # the precision pattern in the steady solve: FP32 push + BF16 split Green's unless asked otherwise (the
    # one-column update is faster as the scalar split GEMV than on MFMA, 7.0 vs 10.8 us)
    if not (push_precision or green_precision):
        push_precision, green_precision = 1, 1
    ps = replace(p, seed=seeds[0], until_steady=True, max_batches=max_batches)
    rv = steady.newton_krylov(ps, push_precision=push_precision, green_precision=green_precision)
    plain0, chosen0 = per_seed[0]["plain"], per_seed[0]["chosen"]
    out["chosen_precision_variant"] = {
        "seed": seeds[0], "push": "FP32" if push_precision else "FP64", "green": GREEN_NAMES[green_precision],
        "converged": bool(rv["converged"]), "batches": rv["batches"], "wall_s": rv["wall_s"],
        "speedup_vs_plain": plain0["wall_s"] / rv["wall_s"], "bottom_center_V": rv["steady"]["bottom center"],
        "drift_vs_plain_fp64_V": rv["steady"]["bottom center"] - plain0["bottom_center_V"],
        "drift_vs_chosen_fp64_V": rv["steady"]["bottom center"] - chosen0["bottom_center_V"]}
    log(f"time-to-steady, chosen with precision variants: {rv['wall_s']:.1f}s, bottom centre "
        f"{rv['steady']['bottom center']:.3f} V")
    return out


if __name__ == "__main__":
    main()
