"""Vertical-ion diagnostic on the device (trench_charging_2d.py:1373-1439).

``probe_vertical_ions(res, p=None, phase=None, log=print, stop=None)`` fires an
equally spaced, exactly vertical beam of monoenergetic ions across the trench
entrance, once through the saturated charge of a charging result and once
through the uncharged field, with the tracer in record mode. It returns the
reference's dictionary: per beam the hit counts, bottom fraction and flux,
escapes, losses, probe paths, impact angles and energies, and the field shown
at the display phase. With the RF cycle (field_mode "circuit_rf") the launch
phases are spread over the period from the stream
SeedSequence([seed, 0x50524F42]), as the reference spreads them.
"""
from dataclasses import replace

import numpy as np

from . import model
from .model import AMU, E_CHARGE, EPS0
from .tracer import make_particle_tracer

# Parameters the diagnostic must share with the charging run (trench_charging_2d.py:1379-1384).
SAME_AS_RUN = ("dx", "nx", "trench_w", "trench_offset", "trench_taper_deg", "trench_d", "floor_t", "n_vac",
               "mask_t", "mask_taper_deg", "mask_type", "eps_r", "mask_eps_r", "field_mode", "side_bc",
               "substrate_voltage", "mask_voltage", "top_voltage", "left_voltage", "right_voltage", "top_bc",
               "ion_mass_amu", "wall_model", "bias")


def probe_vertical_ions(res, p=None, phase=None, log=print, stop=None):
    """The vertical-ion diagnostic of a saturated charging result ``res`` (a
    result of charging.run or a loaded result file). ``p`` may change the
    probe settings (probe_ions, probe_energy_eV, probe_trajectories, seed);
    ``phase`` sets the display phase of the RF cycle."""
    if not np.isfinite(res.get("t_conv", np.nan)) or res.get("stopped", False):
        raise ValueError("the diagnostic needs a result that passed the steady criterion: "
                         "run the charging until steady")
    source = model.Params(**res["params"])
    p = source if p is None else p
    for key in SAME_AS_RUN:
        if getattr(p, key) != getattr(source, key):
            raise ValueError("the diagnostic must use the geometry and boundary conditions of the charging run")
    if (p.probe_ions < 1 or p.probe_trajectories < 1 or p.probe_trajectories > p.probe_ions
            or not np.isfinite(p.probe_energy_eV) or p.probe_energy_eV <= 0):
        raise ValueError("probe ions and energy must be positive, and the probe trajectories at most the ions")
    if phase is not None:
        p = replace(p, rf_phase=phase)
    solid = np.asarray(res["solid"], dtype=bool)
    vac = ~solid
    geo = model.Geometry(p)
    if geo.solid.shape != solid.shape or not np.array_equal(geo.solid, solid):
        raise ValueError("the result's geometry does not match its parameters")
    fields = model.Fields(p, geo, res.get("circuit"))
    # the diagnostic runs on the device whichever backend computed the stored charge
    tracer, backend_used = make_particle_tracer(replace(p, backend="hip"), vac, log)
    ny = solid.shape[1]
    i0, i1 = geo.opening_span(p.floor_t + p.trench_d + p.mask_t - 1)
    b0, b1 = geo.opening_span(p.floor_t)
    # an equally spaced beam removes angle and position sampling from the comparison (vx is exactly 0)
    x = (i0 + (np.arange(p.probe_ions) + 0.5) * (i1 - i0) / p.probe_ions) * p.dx
    y = np.full(x.size, (ny - 1e-3) * p.dx)
    vx = np.zeros(x.size)
    vy = np.full(x.size, -np.sqrt(2 * p.probe_energy_eV * E_CHARGE / (p.ion_mass_amu * AMU)))
    qm = np.full(x.size, E_CHARGE / (p.ion_mass_amu * AMU))
    spc = np.zeros(x.size, dtype=np.int8)
    rho = np.asarray(res["rho_raw"], dtype=float)
    charged_phi = fields.solve((-rho / EPS0).ravel()).reshape(solid.shape) + fields.external
    result = dict(energy_eV=p.probe_energy_eV, n_ions=x.size, phase=p.rf_phase,
                  entrance_i0=i0, entrance_i1=i1, bottom_i0=b0, bottom_i1=b1,
                  source_t_conv=res["t_conv"], source_converged=True,
                  cancelled=False, backend_used=backend_used,
                  rf_cycle=fields.rf is not None)
    options = {}
    if fields.rf is not None:
        rng = np.random.default_rng(np.random.SeedSequence([p.seed, 0x50524F42]))
        options["rf"] = {**fields.rf, "initial_time": model.launch_times(rng, x.size, fields.rf["period"])}
        result["rf_period"] = fields.rf["period"]
        result["rf_time_steps"] = p.rf_time_steps
        log("vertical diagnostic over the whole RF cycle: launch phases spread over 0-360 deg, "
            "circuit field updated during flight")
    try:
        for label, phi in (("charged", charged_phi), ("uncharged", fields.external)):
            log(f"vertical ion diagnostic: {label}, {x.size} ions, {p.probe_energy_eV:g} eV")
            Ex, Ey = model.compute_field(phi, vac, p, fields.volts["top"])
            record = {}
            ci, _, esc, lost = tracer(x.copy(), y.copy(), vx.copy(), vy.copy(), qm.copy(), spc.copy(),
                                      Ex, Ey, vac, p, (phi, fields.wface) if p.wall_model == "barrier" else None,
                                      record=record, stop=stop, **options)
            shown_phi, _, _, volts = fields.snapshot(phi, p)
            shown_ex, shown_ey = model.compute_field(shown_phi, vac, p, volts["top"])
            record.update(phi=shown_phi, Ex=shown_ex, Ey=shown_ey, hits=ci, escaped=int(esc[0]), lost=lost,
                          bottom_fraction=float(ci[b0:b1, p.floor_t - 1].sum() / x.size),
                          bottom_flux=ci[b0:b1, p.floor_t - 1] * p.flux * (i1 - i0) / x.size)
            record["angle_deg"] = np.degrees(np.arctan2(record["final_velocity"][:, 0],
                                                       -record["final_velocity"][:, 1]))
            record["impact_energy_eV"] = (p.ion_mass_amu * AMU / (2 * E_CHARGE)
                                          * np.sum(record["final_velocity"] ** 2, axis=1))
            result[label] = record
            if record["cancelled"]:
                result["cancelled"] = True
                return result
            log(f"  bottom {record['bottom_fraction']:.1%}, escaped at the top {record['escaped']}, "
                f"not arrived {lost}")
    finally:
        tracer.close()
    return result
