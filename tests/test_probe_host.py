"""The vertical-ion diagnostic's argument checks (trench_charging_2d.py:1375-1387;
reference tests/test_features.py:165-171, tests/test_taper.py:190-195), which
run before any device work."""
from dataclasses import asdict, replace

import numpy as np
import pytest

from trench01_amd import model, probe


def small_params(**changes):
    """The reference's tests/test_features.py:16-20."""
    return replace(model.Params(), nx=20, trench_w=6, trench_d=12, floor_t=4, mask_t=4, n_vac=8, bias="dc",
                   sigma_s=0, n_batches=3, n_per_batch=200, max_steps=800, probe_ions=120, probe_trajectories=8,
                   **changes)


def steady_result(p):
    geo = model.Geometry(p)
    z = np.zeros(geo.solid.shape)
    return dict(params=asdict(p), solid=geo.solid, mask=geo.mask, i0=geo.i0, i1=geo.i1, rho_raw=z.copy(),
                t_conv=0.001, stopped=False, circuit=None)


def test_probe_requires_a_saturated_result():
    p = small_params()
    res = steady_result(p)
    res["t_conv"] = np.nan
    with pytest.raises(ValueError, match="steady"):
        probe.probe_vertical_ions(res, log=lambda _: None)
    res["t_conv"] = 0.001
    res["stopped"] = True
    with pytest.raises(ValueError, match="steady"):
        probe.probe_vertical_ions(res, log=lambda _: None)


@pytest.mark.parametrize("change", [dict(trench_w=8), dict(mask_t=6), dict(wall_model="absorb"),
                                    dict(ion_mass_amu=20.0)])
def test_probe_keeps_the_geometry_and_boundaries_of_the_run(change):
    p = small_params()
    with pytest.raises(ValueError, match="geometry"):
        probe.probe_vertical_ions(steady_result(p), replace(p, **change), log=lambda _: None)


@pytest.mark.parametrize("change", [dict(probe_ions=0), dict(probe_trajectories=0), dict(probe_trajectories=121),
                                    dict(probe_energy_eV=0.0), dict(probe_energy_eV=np.nan)])
def test_probe_settings_are_checked(change):
    p = small_params()
    with pytest.raises(ValueError, match="probe"):
        probe.probe_vertical_ions(steady_result(p), replace(p, **change), log=lambda _: None)
