"""Shared inputs for parity tests: geometries at the reference's test sizes and
at the workload sizes, charged fields, and host-generated particle states."""
from dataclasses import replace

import numpy as np

from oracle import charging_ref as R
from trench01_amd import configs
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def regression_params(sigma_s=0.0):
    """tests/test_regression.py:28-29 of the reference."""
    return replace(R.Params(), mask_t=0, n_batches=40, bias="dc", sigma_s=sigma_s, wall_model="absorb")


def small_params(**kw):
    """tests/test_cuda.py:15-19 of the reference."""
    return replace(R.Params(), nx=30, trench_w=10, trench_d=20, floor_t=5, mask_t=6, n_vac=12, bias="dc",
                   sigma_s=0, n_batches=8, n_per_batch=400, max_steps=1000, **kw)


def taper_params(**kw):
    """tests/test_taper.py:19-24 of the reference (tan 0.25 and tan 0.5 walls)."""
    return replace(R.Params(), nx=40, trench_w=16, trench_d=20, floor_t=4, mask_t=12, n_vac=10,
                   trench_taper_deg=float(np.degrees(np.arctan(.25))),
                   mask_taper_deg=float(np.degrees(np.arctan(.5))), bias="dc", sigma_s=0, n_batches=3,
                   n_per_batch=256, max_steps=800, **kw)


def ar5_params(**kw):
    """Plain method at AR 5: 80x200, trench 20x100, conductive mask 20, sigma_s 0, Te 4 eV."""
    return replace(R.Params(), **{**configs.AR5, **kw})


def ar8_params(**kw):
    """Production size at AR 8: 120x320, trench 20x160, mask 40, tapers 2/5 deg, sigma_s 1e-15."""
    return replace(R.Params(), **{**configs.AR8, **kw})


def charged_case(p, q_left=2e7, q_right=-1e7, seed=20, n=None, ion=None):
    """Geometry, a charged field (uniform sidewall charge, as tests/test_cuda.py:52-57)
    and injected particles.  Returns dict(p, solid, vac, Ex, Ey, barrier, particles)."""
    solid, mask, i0, i1 = R.build_geometry(p)
    fields = R.setup_fields(p, solid, mask)
    rho = np.zeros(solid.shape)
    rho[i0 - 1, p.floor_t:p.floor_t + p.trench_d] = q_left
    rho[i1, p.floor_t:p.floor_t + p.trench_d] = q_right
    phi = fields["solve"]((-rho / R.EPS0).ravel()).reshape(solid.shape) + fields["external"]
    ex, ey = R.compute_field(phi, ~solid, p)
    wface = np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"]))
    barrier = (phi, wface) if p.wall_model == "barrier" else None
    rng = np.random.default_rng(seed)
    nb = p.n_per_batch if n is None else n
    sampler = configs.ion_velocity_sampler(p.ion_mass_amu, **ion) if ion else None
    particles = R.inject(p, nb, solid.shape[1], rng, ion_velocity=sampler)
    return dict(p=p, solid=solid, vac=~solid, Ex=ex, Ey=ey, barrier=barrier, particles=particles)


def marched_case(p, batches, n_per_batch, n_trace, seed=11, ion=None):
    """Field after ``batches`` batches of the oracle's time-marching loop (a
    realistically charged trench: reflections, trapped electrons), then a fresh
    injected batch of ``n_trace`` particles per species."""
    sampler = configs.ion_velocity_sampler(p.ion_mass_amu, **ion) if ion else None
    pr = replace(p, n_per_batch=n_per_batch, n_batches=batches, until_steady=False)
    res = R.run(pr, ion_velocity=sampler)
    solid = res["solid"]
    fields = R.setup_fields(p, solid, res["mask"])
    wface = np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"]))
    barrier = (res["phi"], wface) if p.wall_model == "barrier" else None
    particles = R.inject(p, n_trace, solid.shape[1], np.random.default_rng(seed), ion_velocity=sampler)
    return dict(p=p, solid=solid, vac=~solid, Ex=res["Ex"], Ey=res["Ey"], barrier=barrier, particles=particles)
