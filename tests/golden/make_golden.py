#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ by importing the reference
(Reasonia-TK/trench01, read-only checkout).

Only run where that checkout exists; the reference does not travel with the tests.  Each
vector pins a reference function that the reference's own fixtures
(tests/data/base_*.npz: absorb wall, no mask, no taper) leave open:

* ``trace_*.npz``   trace_cpu (trench_charging_2d.py:542-702) in record mode on
  barrier / absorb walls, conductive masks and tapered walls, with charged
  fields from the reference's own Poisson set-up (setup_fields :340-413,
  compute_field :439-454): inputs, fields, per-particle status / hit cell /
  end point / final velocity, per-cell counts.
* ``setup_*.npz``   the conductor-collapsed LU (make_solver :297-309),
  conductor_response / C_cond (:807-812, :1089), leak operator
  (build_leakage :768-803) and conductor_step (:815-830) on small geometries.
* ``rftrace_*.npz`` trace_cpu with the RF full-cycle push (rf=, :556-602,
  :629-632, :654-682): circuit_rf fields (setup_fields :361-371, :397-412) of
  a sinusoidal circuit waveform, launch times from trench_rf.launch_times,
  charged sidewalls; record-mode outputs incl. flight time and phases.  Run
  with single-threaded BLAS (the bound's matrix product rounds per position;
  see oracle/trace_ref.c).
* ``runrf_*.npz``   whole run()s in the circuit field modes, with the circuit
  solver (plasma_circuit.solve_circuit_cached) replaced by the sinusoidal
  waveform above as the reference's tests/test_rf_cycle.py:127-133 does:
  ``circuit_rf`` (full-cycle push, pillar sides, surface leak) and
  ``circuit`` (voltages frozen at rf_phase 120, fixed sides); few batches,
  single-threaded BLAS; ``result_circuit_rf.npz``: the reference's result file
  of the circuit_rf run (save_results), for the circuit waveform on disk.
* ``run_ar5.npz``   a whole run() (:1045-1369) of the plain AR 5 workload
  (80x200, conductive mask, sigma_s 0, barrier wall, 100 eV ions with 3 degree
  polar spread, Te 4 eV), 50 batches of 1e4 particles per species, seed 0;
  ``trace_ar5_marched.npz``: a fresh record-mode batch through its final field;
  ``result_ar5.npz``: the reference's own result file of that run (save_results
  :1687-1734), for the on-disk format.

Usage:  python tests/golden/make_golden.py [--reference /root/reference] [--only NAME]
"""
import argparse
import json
import sys
import time
from dataclasses import asdict, replace
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parents[1]


def load_reference(path):
    sys.path.insert(0, str(path))
    import matplotlib
    matplotlib.use("Agg")
    import trench_charging_2d as T
    return T


def ion_sampler(T, kind, sigma_deg, energy_eV=None, e_lo=None, e_hi=None, mass_amu=40.0):
    """Polar-angle Gaussian ion velocities (monoenergetic or arcsine IED),
    the ion_velocity(rng, n) hook of the reference's inject (:458-474)."""
    m_i = mass_amu * T.AMU
    sig = np.radians(sigma_deg)

    def sample(rng, n):
        if kind == "mono":
            E = np.full(n, float(energy_eV))
        else:
            E = e_lo + (e_hi - e_lo) * np.sin(0.5 * np.pi * rng.uniform(size=n)) ** 2
        theta = rng.normal(0.0, sig, n)
        v = np.sqrt(2.0 * E * T.E_CHARGE / m_i)
        return v * np.sin(theta), -v * np.cos(theta)

    return sample


def case_params(T, name):
    base = T.Params()
    if name.startswith("small"):
        return replace(base, nx=30, trench_w=10, trench_d=20, floor_t=5, mask_t=6, n_vac=12, bias="dc",
                       sigma_s=0, n_per_batch=400, max_steps=1000, probe_trajectories=12,
                       wall_model="absorb" if name.endswith("absorb") else "barrier")
    if name.startswith("taper"):
        return replace(base, nx=40, trench_w=16, trench_d=20, floor_t=4, mask_t=12, n_vac=10,
                       trench_taper_deg=float(np.degrees(np.arctan(.25))),
                       mask_taper_deg=float(np.degrees(np.arctan(.5))), bias="dc", sigma_s=0,
                       n_per_batch=1000, max_steps=800, probe_trajectories=12,
                       wall_model="absorb" if name.endswith("absorb") else "barrier")
    if name.startswith("ar5"):
        return replace(base, nx=80, trench_w=20, trench_d=100, floor_t=10, mask_t=20, n_vac=70,
                       mask_type="conductor", field_mode="floating", sigma_s=0.0, wall_model="barrier",
                       bias="dc", electron_temp_eV=4.0, n_per_batch=2000, seed=0, probe_trajectories=12)
    if name.startswith("ar8"):
        return replace(base, nx=120, trench_w=20, trench_d=160, floor_t=10, mask_t=40, n_vac=110,
                       trench_taper_deg=2.0, mask_taper_deg=5.0, mask_type="conductor",
                       field_mode="floating", sigma_s=1e-15, wall_model="barrier", bias="dc",
                       electron_temp_eV=4.0, n_per_batch=2000, seed=0, probe_trajectories=12)
    raise KeyError(name)


TRACE_CASES = {
    # name: (sidewall charge left/right [C/m^3], ion sampler)
    "small_barrier": ((2e7, -1e7), None),
    "small_absorb": ((2e7, -1e7), None),
    "taper_barrier": ((2e7, -1e7), None),
    "taper_absorb": ((2e7, -1e7), None),
    "ar5_barrier": ((3e7, -1.5e7), dict(kind="mono", energy_eV=100.0, sigma_deg=3.0)),
    "ar8_barrier": ((4e7, -2e7), dict(kind="arcsine", e_lo=40.0, e_hi=260.0, sigma_deg=4.0)),
}


RF_CASES = {
    # name: (params changes, sidewall charge left/right [C/m^3])
    "pillar_barrier": (dict(nx=20, trench_w=6, trench_d=12, floor_t=4, mask_t=4, n_vac=8, side_bc="pillar",
                               wall_model="barrier"), (2e7, -1e7)),
    "taper_absorb": (dict(nx=40, trench_w=16, trench_d=20, floor_t=4, mask_t=12, n_vac=10,
                             trench_taper_deg=float(np.degrees(np.arctan(.25))),
                             mask_taper_deg=float(np.degrees(np.arctan(.5))), wall_model="absorb"), (2e7, -1e7)),
    "taper_barrier": (dict(nx=40, trench_w=16, trench_d=20, floor_t=4, mask_t=12, n_vac=10,
                              trench_taper_deg=float(np.degrees(np.arctan(.25))),
                              mask_taper_deg=float(np.degrees(np.arctan(.5))), wall_model="barrier"), (3e7, -1e7)),
}


def rf_circuit(period=4e-10, n=256, amplitude=45.0):
    """One period of a sinusoidal bias in the circuit-waveform layout that
    setup_fields reads (t, uw, um, up, V1; :313-321, :363-371)."""
    t = np.arange(n) * period / n
    phase = 2 * np.pi * t / period
    up = 12 + 3 * np.cos(phase)
    uw = -70 - amplitude * np.sin(phase) + up
    um = -30 - 0.5 * amplitude * np.sin(phase + 0.3) + up
    return dict(t=t, uw=uw, um=um, up=up, V1=up - um)


RUN_RF_CASES = {
    "circuit_rf": dict(nx=20, trench_w=6, trench_d=12, floor_t=4, mask_t=4, n_vac=8, field_mode="circuit_rf",
                       side_bc="pillar", sigma_s=1e-14, n_batches=12, n_per_batch=1000, rf_phase=75.0),
    "circuit": dict(nx=40, trench_w=16, trench_d=20, floor_t=4, mask_t=12, n_vac=10, field_mode="circuit",
                    trench_taper_deg=float(np.degrees(np.arctan(.25))), side_bc="fixed", left_voltage=-15.0,
                    right_voltage=4.0, sigma_s=1e-14, n_batches=8, n_per_batch=1000, rf_phase=120.0),
}


def params_json(p):
    """The run's Params as JSON (stored with every vector)."""
    return np.array(json.dumps(asdict(p)))


def write_trace_case(T, name):
    (ql, qr), ion = TRACE_CASES[name]
    p = case_params(T, name)
    solid, mask, i0, i1 = T.build_geometry(p)
    fields = T.setup_fields(p, solid, mask)
    rho = np.zeros(solid.shape)
    rho[i0 - 1, p.floor_t:p.floor_t + p.trench_d] = ql
    rho[i1, p.floor_t:p.floor_t + p.trench_d] = qr
    phi = fields["solve"]((-rho / T.EPS0).ravel()).reshape(solid.shape) + fields["external"]
    ex, ey = T.compute_field(phi, ~solid, p)
    wface = np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"]))
    barrier = (phi, wface) if p.wall_model == "barrier" else None
    sampler = ion_sampler(T, **ion) if ion else None
    x, y, vx, vy, qm, spc = T.inject(p, p.n_per_batch, solid.shape[1], np.random.default_rng(20),
                                     ion_velocity=sampler)
    record = {}
    ci, ce, esc, lost = T.trace_cpu(x.copy(), y.copy(), vx.copy(), vy.copy(), qm.copy(), spc.copy(), ex, ey,
                                    ~solid, p, barrier, record)
    np.savez_compressed(HERE / f"trace_{name}.npz", params=params_json(p), solid=solid, mask=mask, rho=rho,
                        phi=phi, Ex=ex, Ey=ey,
                        wface=wface, x=x, y=y, vx=vx, vy=vy, qm=qm, spc=spc, cnt_i=ci, cnt_e=ce,
                        escaped=esc, lost=np.int64(lost), status=record["status"], hit_cells=record["hit_cells"],
                        endpoints=record["endpoints"], final_velocity=record["final_velocity"],
                        selected=record["selected"],
                        path_lengths=np.array([len(v) for v in record["paths"]]),
                        paths=np.concatenate(record["paths"]) if record["paths"] else np.zeros((0, 2)))
    print(f"trace_{name}: n={x.size} absorbed={int((record['status'] == 1).sum())} escaped={esc} lost={lost}")


def write_rf_case(T, name):
    from threadpoolctl import threadpool_limits
    import trench_rf as RF
    changes, (ql, qr) = RF_CASES[name]
    p = replace(T.Params(), field_mode="circuit_rf", rf_freq=2.5e9, bias="rf", mask_type="conductor",
                sigma_s=0.0, n_per_batch=300, max_steps=1000, probe_trajectories=12, **changes)
    solid, mask, i0, i1 = T.build_geometry(p)
    fields = T.setup_fields(p, solid, mask, rf_circuit())
    rho = np.zeros(solid.shape)
    rho[i0 - 1, p.floor_t:p.floor_t + p.trench_d] = ql
    rho[i1, p.floor_t:p.floor_t + p.trench_d] = qr
    phi = fields["solve"]((-rho / T.EPS0).ravel()).reshape(solid.shape) + fields["external"]
    ex, ey = T.compute_field(phi, ~solid, p)
    wface = np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"]))
    barrier = (phi, wface) if p.wall_model == "barrier" else None
    sampler = ion_sampler(T, kind="mono", energy_eV=60.0, sigma_deg=3.0)
    rng = np.random.default_rng(22)
    x, y, vx, vy, qm, spc = T.inject(p, p.n_per_batch, solid.shape[1], rng, ion_velocity=sampler)
    rf = {**fields["rf"], "initial_time": np.concatenate(
        [RF.launch_times(rng, p.n_per_batch, fields["rf"]["period"]) for _ in range(2)])}
    record = {}
    with threadpool_limits(1):
        ci, ce, esc, lost = T.trace_cpu(x.copy(), y.copy(), vx.copy(), vy.copy(), qm.copy(), spc.copy(), ex,
                                        ey, ~solid, p, barrier, record, rf=rf)
    np.savez_compressed(HERE / f"rftrace_{name}.npz", params=params_json(p), solid=solid, mask=mask, rho=rho,
                        phi=phi, Ex=ex, Ey=ey, wface=wface, x=x, y=y, vx=vx, vy=vy, qm=qm, spc=spc,
                        rf_period=np.float64(rf["period"]), rf_dt_max=np.float64(rf["dt_max"]),
                        rf_wave=rf["wave"], rf_Ex=rf["Ex"], rf_Ey=rf["Ey"], rf_phi=rf["phi"],
                        rf_initial_time=rf["initial_time"], cnt_i=ci, cnt_e=ce, escaped=esc,
                        lost=np.int64(lost), status=record["status"], hit_cells=record["hit_cells"],
                        endpoints=record["endpoints"], final_velocity=record["final_velocity"],
                        flight_time=record["flight_time"], final_phase=record["final_phase"],
                        launch_phase=record["launch_phase"], selected=record["selected"],
                        path_lengths=np.array([len(v) for v in record["paths"]]),
                        paths=np.concatenate(record["paths"]) if record["paths"] else np.zeros((0, 2)))
    print(f"rftrace_{name}: n={x.size} absorbed={int((record['status'] == 1).sum())} escaped={esc} lost={lost}")


def write_run_rf(T, name):
    from threadpoolctl import threadpool_limits
    p = replace(T.Params(), bias="rf", rf_freq=2.5e9, mask_type="conductor", wall_model="barrier",
                max_steps=1000, probe_trajectories=8, seed=3, **RUN_RF_CASES[name])
    circuit = dict(rf_circuit(), us=None, i=None, E=np.zeros(0), cached=False, ied_computed=False,
                   sheath=None, ied_model="sheath", n_periods=1, elements={})
    circuit["us"] = circuit["uw"].copy()
    circuit["i"] = circuit["t"] * 0
    original = T.PC.solve_circuit_cached
    T.PC.solve_circuit_cached = lambda *a, **k: circuit
    try:
        with threadpool_limits(1):
            res = T.run(p, log=lambda s: None)
    finally:
        T.PC.solve_circuit_cached = original
    arrays = dict(params=params_json(p), rho_raw=res["rho_raw"], rho=res["rho"], phi=res["phi"],
                  mean_phi=res["mean_phi"], external_phi=res["external_phi"], Ex=res["Ex"], Ey=res["Ey"],
                  solid=res["solid"], voltages=np.array(json.dumps(res["voltages"])),
                  **{f"circuit_{k}": circuit[k] for k in ("t", "uw", "um", "up", "V1")})
    arrays.update({f"hist_{k}": v for k, v in res["hist"].items()})
    arrays.update({f"charge_{k}": v for k, v in res["charge_monitor_hist"].items()})
    np.savez_compressed(HERE / f"runrf_{name}.npz", **arrays)
    print(f"runrf_{name}: bottom centre {res['hist']['bottom center'][-1]:.3f} V, "
          f"lost {res['hist']['lost'][-1]:.4f}")
    if name == "circuit_rf":
        # the reference's result file of this run (NPZ only): circuit waveform and phase-stream checkpoint
        import shutil
        import tempfile
        with tempfile.TemporaryDirectory() as tmp:
            T.save_results(res, p, str(Path(tmp) / "out"))
            shutil.copy(Path(tmp) / "out.npz", HERE / "result_circuit_rf.npz")


def write_setup_case(T, name):
    """Leak / conductor operators on the small geometries (conductor mask,
    sigma_s 1e-14), floating field mode, plus a fixed-voltage external field."""
    p = replace(case_params(T, name + "_barrier"), sigma_s=1e-14)
    solid, mask, i0, i1 = T.build_geometry(p)
    fields = T.setup_fields(p, solid, mask)
    leak = T.build_leakage(solid, fields["solve"], i0, i1, p)
    ix, iy = np.nonzero(T.surface_cells(solid))
    flat = ix * solid.shape[1] + iy
    B = np.zeros((solid.size, flat.size))
    B[flat, np.arange(flat.size)] = -1.0 / (p.dx ** 2 * T.EPS0)
    green = fields["solve"](B)
    cond = fields["cond"]
    g_cond = T.conductor_response(fields["solve"], cond, p.dx)
    C_cond = 1.0 / g_cond[cond][0]
    pf = replace(p, field_mode="fixed", substrate_voltage=-20.0, mask_voltage=-10.0, top_voltage=5.0)
    ext_fixed = T.setup_fields(pf, solid, mask)["external"]
    steps = []
    for n_i, n_e, drift in ((10, 0, 0.0), (10, 12, 0.0), (0, 7, 0.3), (25, 25, -0.2), (3, 40, 1.5)):
        steps.append((n_i, n_e, drift, T.conductor_step(n_i, n_e, 1e-15, C_cond, p.electron_temp_eV, drift)))
    np.savez_compressed(HERE / f"setup_{name}.npz", params=params_json(p), solid=solid, mask=mask, flat=flat,
                        M=leak["M"], G=leak["G"],
                        region=leak["region"], green=green, g_cond=g_cond, C_cond=np.float64(C_cond),
                        external=fields["external"], external_fixed=ext_fixed, conductor_steps=np.array(steps))
    print(f"setup_{name}: m={flat.size} C_cond={C_cond:.4g}")


def write_run_ar5(T):
    p = replace(case_params(T, "ar5"), n_per_batch=10_000, n_batches=50)
    sampler = ion_sampler(T, kind="mono", energy_eV=100.0, sigma_deg=3.0)
    original = T.inject
    T.inject = lambda p_, n, ny, rng, ion_energy=None, ion_velocity=None: original(p_, n, ny, rng, ion_energy,
                                                                                    sampler)
    try:
        t0 = time.time()
        res = T.run(p, log=lambda s: None)
    finally:
        T.inject = original
    arrays = dict(params=params_json(p), rho_raw=res["rho_raw"], phi=res["phi"], Ex=res["Ex"], Ey=res["Ey"],
                  solid=res["solid"])
    arrays.update({f"hist_{k}": v for k, v in res["hist"].items()})
    arrays.update({f"charge_{k}": v for k, v in res["charge_monitor_hist"].items()})
    np.savez_compressed(HERE / "run_ar5.npz", **arrays)
    print(f"run_ar5: {time.time() - t0:.0f}s, bottom centre {res['hist']['bottom center'][-1]:.3f} V")
    # The reference's own result file of this run (save_results writes plots too;
    # only the NPZ is kept, as result_ar5.npz).
    import shutil
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        T.save_results(res, p, str(Path(tmp) / "out"))
        shutil.copy(Path(tmp) / "out.npz", HERE / "result_ar5.npz")
    # A fresh batch traced in record mode through the charged field the run ends with.
    solid = res["solid"]
    fields = T.setup_fields(p, solid, res["mask"])
    wface = np.where(fields["cond"], 1.0, fields["eps"] / (1 + fields["eps"]))
    x, y, vx, vy, qm, spc = T.inject(replace(p, n_per_batch=2000), 2000, solid.shape[1],
                                     np.random.default_rng(21), ion_velocity=sampler)
    record = {}
    ci, ce, esc, lost = T.trace_cpu(x.copy(), y.copy(), vx.copy(), vy.copy(), qm.copy(), spc.copy(),
                                    res["Ex"], res["Ey"], ~solid, p, (res["phi"], wface), record)
    np.savez_compressed(HERE / "trace_ar5_marched.npz", params=params_json(p), solid=solid, mask=res["mask"],
                        rho=res["rho_raw"],
                        phi=res["phi"], Ex=res["Ex"], Ey=res["Ey"], wface=wface, x=x, y=y, vx=vx, vy=vy, qm=qm,
                        spc=spc, cnt_i=ci, cnt_e=ce, escaped=esc, lost=np.int64(lost), status=record["status"],
                        hit_cells=record["hit_cells"], endpoints=record["endpoints"],
                        final_velocity=record["final_velocity"], selected=record["selected"],
                        path_lengths=np.array([len(v) for v in record["paths"]]),
                        paths=np.concatenate(record["paths"]))
    print(f"trace_ar5_marched: absorbed={int((record['status'] == 1).sum())} escaped={esc} lost={lost}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", default=None)
    a = ap.parse_args()
    T = load_reference(a.reference)
    jobs = [*(f"trace_{k}" for k in TRACE_CASES), *(f"rftrace_{k}" for k in RF_CASES),
            *(f"runrf_{k}" for k in RUN_RF_CASES), "setup_small", "setup_taper", "run_ar5"]
    for job in jobs:
        if a.only and a.only != job:
            continue
        if job.startswith("rftrace_"):
            write_rf_case(T, job[8:])
        elif job.startswith("runrf_"):
            write_run_rf(T, job[6:])
        elif job.startswith("trace_"):
            write_trace_case(T, job[6:])
        elif job.startswith("setup_"):
            write_setup_case(T, job[6:])
        else:
            write_run_ar5(T)


if __name__ == "__main__":
    main()
