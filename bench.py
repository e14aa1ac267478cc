#!/usr/bin/env python3
"""Throughput of the per-batch charging hot path on one MI355X.

One step = one batch of the AR 8 production workload (120x320 cells, trench
20x160 with 2 / 5 degree tapers, conductive mask 40 cells, barrier wall model):
1e6 particles (5e5 ions from the arcsine 40-260 eV IED with 4 degree angular
spread, 5e5 electrons at Te = 4 eV) traced through the frozen field with
LDS-privatised deposit tallies.  Particle states are generated on the host
(the reference's injection) and uploaded before the timed region, one
distinct batch per step, so the timed work starts from HBM.

Prints one JSON line (metric, value, roofline of the push kernel, and the CPU
baseline: the oracle's C restatement of the push on this host's cores, on a
bounded sample of the same workload).
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
FP64_PEAK_TFLOPS = 78.6     # MI355X FP64 vector / matrix peak
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-per-batch", type=int, default=500_000, help="particles per species per batch")
    ap.add_argument("--cpu-sample", type=int, default=20_000, help="particles per species in the CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--refill", type=int, default=0)
    return ap.parse_args()


def workload(n_per_batch):
    from trench01_amd import configs, model
    p = replace(model.Params(), **{**configs.AR8, "n_per_batch": n_per_batch})
    geo = model.Geometry(p)
    fields = model.Fields(p, geo)
    # Frozen field of a charged trench: positive charge on the left dielectric
    # wall, negative on the right (the reference's CUDA parity field,
    # tests/test_cuda.py:52-57), on top of the applied boundary potentials.
    rho = np.zeros(geo.solid.shape)
    rho[geo.i0 - 1, p.floor_t:p.floor_t + p.trench_d] = 4e7
    rho[geo.i1, p.floor_t:p.floor_t + p.trench_d] = -2e7
    phi = fields.solve((-rho / model.EPS0).ravel()).reshape(rho.shape) + fields.external
    ex, ey = model.compute_field(phi, geo.vac, p)
    sampler = configs.ion_velocity_sampler(p.ion_mass_amu, **configs.AR8_ION)
    return p, geo, ex, ey, (phi, fields.wface), sampler


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    from trench01_amd import HipTracer, model
    p, geo, ex, ey, barrier, sampler = workload(a.n_per_batch)
    tracer = HipTracer(geo.vac)
    tracer.refill = a.refill
    tracer.set_fields(ex, ey, barrier)
    prm = tracer.params_for(p, barrier)
    nb = a.warmup + a.steps
    n = 2 * p.n_per_batch
    rng = np.random.default_rng(p.seed)
    tracer.reserve(nb * n)
    first = None
    for b in range(nb):
        parts = model.inject(p, p.n_per_batch, geo.ny, rng, ion_velocity=sampler)
        if first is None:
            first = parts
        tracer.upload(b * n, *parts)
    for b in range(a.warmup):
        tracer.trace_resident(b * n, n, prm)
    steps_total = 0
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for b in range(a.warmup, nb):
        tally, _, ms = tracer.trace_resident(b * n, n, prm)
        steps_total += tally["steps"]
        kernel_ms += ms
    wall = time.perf_counter() - t0
    value = a.steps * n / wall
    achieved = FLOP_PER_STEP * steps_total / (kernel_ms * 1e-3) / 1e12
    out = {
        "metric": "traced particles/s/GPU (AR=8, 1e6 particles per batch)",
        "value": value, "unit": "particles/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (host-injected particle states)",
        "config": {"workload": "AR8 trench 120x320, tapers 2/5 deg, conductive mask 40, barrier wall, "
                               "1e6 particles per batch, FP64 push + LDS deposit tallies",
                   "particles_per_batch": n, "grid": [geo.nx, geo.ny], "surface_cells": int(geo.surface.size),
                   "max_steps": p.max_steps, "mean_steps_per_particle": steps_total / (a.steps * n)},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "kernel": "push64_kernel", "kernel_ms_per_step": kernel_ms / a.steps,
                     "flop_per_particle_step": FLOP_PER_STEP,
                     "hbm_gbs_algorithmic": BYTES_PER_PARTICLE * n * a.steps / (kernel_ms * 1e-3) / 1e9},
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, geo, ex, ey, barrier, first, a.cpu_sample)
    print(json.dumps(out), flush=True)


def cpu_baseline(p, geo, ex, ey, barrier, parts, sample):
    """The oracle's C push (OpenMP) on a bounded sample of batch 0: ``sample``
    ions and ``sample`` electrons, on all host cores this process may use."""
    from oracle import trace_ref
    from trench01_amd.model import is_tapered
    n = p.n_per_batch
    idx = np.concatenate([np.arange(sample), n + np.arange(sample)])
    x, y, vx, vy, qm, spc = (v[idx] for v in parts)
    cores = len(os.sched_getaffinity(0))
    trace_ref.lib()
    t0 = time.perf_counter()
    r = trace_ref.trace_arrays(x, y, vx, vy, qm, spc, ex, ey, geo.vac, p.dx, p.cfl, p.max_steps,
                               is_tapered(p), barrier, threads=cores)
    wall = time.perf_counter() - t0
    return {"value": idx.size / wall, "unit": "particles/s", "cores": cores, "kind": "port",
            "sample": f"first {sample} ions + {sample} electrons of batch 0 (same field), "
                      f"{r['total_steps']} particle-steps in {wall:.2f} s"}


if __name__ == "__main__":
    main()
