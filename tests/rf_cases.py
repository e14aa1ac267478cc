"""Inputs of the RF full-cycle push tests, shared by the oracle and GPU tests.

* ``golden(name)``: the reference's own RF traces (tests/golden/rftrace_*.npz,
  written by trace_cpu with rf=, trench_charging_2d.py:556-602).
* ``analytic(steps)``: a spatially uniform field E_y = A sin(2 pi t) acting on
  12 particles launched at different phases, traced for 2/3 of a period
  without reaching a wall; the closed-form velocity and position after the
  flight are known (the known answer of tests/test_rf_cycle.py:59-88).
* ``barrier_case()``: one ion moving into a wall whose potential is raised
  only by the RF wave (tests/test_rf_cycle.py:92-113): the instantaneous
  voltage must reflect it, the frozen field alone absorbs it.
"""
import json
from dataclasses import fields as dc_fields, replace
from pathlib import Path

import numpy as np

from trench01_amd import model

GOLDEN = Path(__file__).resolve().parent / "golden"
CASES = sorted(p.stem[8:] for p in GOLDEN.glob("rftrace_*.npz"))


def golden(name):
    with np.load(GOLDEN / f"rftrace_{name}.npz", allow_pickle=False) as d:
        g = {k: d[k] for k in d.files}
    stored = json.loads(str(g["params"]))
    names = {f.name for f in dc_fields(model.Params)}
    p = replace(model.Params(), **{k: v for k, v in stored.items() if k in names})
    rf = dict(period=float(g["rf_period"]), dt_max=float(g["rf_dt_max"]), wave=g["rf_wave"], Ex=g["rf_Ex"],
              Ey=g["rf_Ey"], phi=g["rf_phi"], initial_time=g["rf_initial_time"])
    barrier = (g["phi"], g["wface"]) if p.wall_model == "barrier" else None
    args = (g["x"], g["y"], g["vx"], g["vy"], g["qm"], g["spc"], g["Ex"], g["Ey"], ~g["solid"])
    return g, p, rf, barrier, args


AMPLITUDE = 0.05
N_ANALYTIC = 12


def analytic(steps):
    """(args, p, rf, t0) of the uniform sinusoidal field case; max_steps = 2/3 period."""
    p = replace(model.Params(), dx=0.1, max_steps=steps * 2 // 3, probe_trajectories=4)
    shape = (3, 100)
    vac = np.ones(shape, bool)
    z = np.zeros(shape)
    n = N_ANALYTIC
    t0 = np.linspace(0.0, 0.9, n)
    k = np.arange(4096)
    wave = np.column_stack((AMPLITUDE * np.sin(k * 2 * np.pi / 4096), np.zeros(4096)))
    rf = dict(period=1.0, dt_max=1.0 / steps, initial_time=t0, wave=wave, Ex=np.zeros((2, *shape)),
              Ey=np.array([np.ones(shape), z]), phi=np.zeros((2, *shape)))
    args = (np.full(n, 0.15), np.full(n, 8.0), np.zeros(n), np.full(n, -0.1), np.ones(n), np.zeros(n, np.int8),
            z, z, vac)
    return args, p, rf, t0


def analytic_error(record, t0):
    """Largest deviation of the final vertical velocity and height from the
    closed form of dv/dt = A sin(omega (t0 + t))."""
    t, omega = record["flight_time"], 2 * np.pi
    v = -0.1 + AMPLITUDE / omega * (np.cos(omega * t0) - np.cos(omega * (t0 + t)))
    y = (8.0 - 0.1 * t + AMPLITUDE / omega * np.cos(omega * t0) * t
         - AMPLITUDE / omega ** 2 * (np.sin(omega * (t0 + t)) - np.sin(omega * t0)))
    return max(np.max(np.abs(record["final_velocity"][:, 1] - v)), np.max(np.abs(record["endpoints"][:, 1] - y)))


def barrier_case():
    """(args, p, barrier, rf) of the time-dependent wall barrier."""
    p = replace(model.Params(), dx=0.001, max_steps=2, probe_trajectories=1)
    vac = np.ones((3, 4), bool)
    vac[:, :2] = False
    z = np.zeros(vac.shape)
    basis = np.zeros((2, *vac.shape))
    basis[0, :, :2] = 1.0
    rf = dict(period=1.0, dt_max=1.0, initial_time=np.zeros(1), wave=np.array([[1.0, 0.0], [1.0, 0.0]]),
              Ex=np.zeros_like(basis), Ey=np.zeros_like(basis), phi=basis)
    args = (np.array([0.0015]), np.array([0.00201]), np.zeros(1), np.array([-0.1]), np.ones(1),
            np.zeros(1, np.int8), z, z, vac)
    return args, p, (z, np.ones_like(z)), rf
