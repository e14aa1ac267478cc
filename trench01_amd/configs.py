"""Workload definitions: geometry / plasma settings and ion samplers.

The ion samplers follow the ``ion_velocity(rng, n) -> (vx, vy)`` hook of the
reference's injection (trench_charging_2d.py:458-474): they are drawn right
after the ion x positions, so the host random stream keeps the reference's
draw order.
"""
import numpy as np

E_CHARGE = 1.602176634e-19
AMU = 1.66053906660e-27

# Plain method, CPU-runnable: 80x200 cells, AR 5, conductive mask 20 cells.
AR5 = dict(nx=80, trench_w=20, trench_d=100, floor_t=10, mask_t=20, n_vac=70, dx=5e-9,
           mask_type="conductor", field_mode="floating", sigma_s=0.0, wall_model="barrier", bias="dc",
           electron_temp_eV=4.0, n_per_batch=10_000, n_batches=50, seed=0,
           ion_energy_eV=100.0)
AR5_ION = dict(kind="mono", energy_eV=100.0, sigma_deg=3.0)

# Production size: 120x320 cells, AR 8, mask 40, tapers 2 / 5 degrees.
AR8 = dict(nx=120, trench_w=20, trench_d=160, floor_t=10, mask_t=40, n_vac=110, dx=5e-9,
           trench_taper_deg=2.0, mask_taper_deg=5.0, mask_type="conductor", field_mode="floating",
           sigma_s=1e-15, wall_model="barrier", bias="dc", electron_temp_eV=4.0, n_per_batch=500_000,
           until_steady=True, max_batches=5000, seed=0)
AR8_ION = dict(kind="arcsine", e_lo=40.0, e_hi=260.0, sigma_deg=4.0)


def ion_velocity_sampler(ion_mass_amu, kind, sigma_deg, energy_eV=None, e_lo=None, e_hi=None):
    """Ion velocity sampler: energy E (monoenergetic, or arcsine-shaped
    bimodal E = lo + (hi-lo) sin^2(pi U / 2)), polar angle theta ~ N(0, sigma)
    about the downward vertical; vx = v sin(theta), vy = -v cos(theta)."""
    m_i = ion_mass_amu * AMU
    sig = np.radians(sigma_deg)

    def sample(rng, n):
        if kind == "mono":
            E = np.full(n, float(energy_eV))
        elif kind == "arcsine":
            E = e_lo + (e_hi - e_lo) * np.sin(0.5 * np.pi * rng.uniform(size=n)) ** 2
        else:
            raise ValueError(f"unknown ion energy distribution {kind!r}")
        theta = rng.normal(0.0, sig, n)
        v = np.sqrt(2.0 * E * E_CHARGE / m_i)
        return v * np.sin(theta), -v * np.cos(theta)

    return sample


def product_params(name, **changes):
    """Parameters of a named workload ("AR5" or "AR8") for this package's
    ``model.Params``, with the ion distribution as Params fields (used by the
    device-side injection) and the matching host sampler."""
    from dataclasses import replace
    from .model import Params
    geo, ion = {"AR5": (AR5, AR5_ION), "AR8": (AR8, AR8_ION)}[name]
    extra = dict(ion_distribution=ion["kind"], ion_angle_deg=ion["sigma_deg"])
    if ion["kind"] == "mono":
        extra["ion_energy_eV"] = ion["energy_eV"]
    else:
        extra.update(ion_energy_lo_eV=ion["e_lo"], ion_energy_hi_eV=ion["e_hi"])
    p = replace(Params(), **{**geo, **extra, **changes})
    return p, ion_velocity_sampler(p.ion_mass_amu, **ion)
