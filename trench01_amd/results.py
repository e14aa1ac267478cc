"""Result files: the reference's NPZ format (trench_charging_2d.py:1687-1801),
without plots.

``save_results(res, out)`` writes ``out.npz`` with the reference's keys (fields,
raw charge, histories, charge monitors, steady values, the run's parameters
as JSON, the checkpoint with the host random state), so the reference's
``load_results`` reads it.  ``load_results(path)`` reads files written by
either side without unpickling anything (``allow_pickle=False``).  Parameters
that the reference's ``Params`` does not have (this package's device-injection
fields) are kept under ``params_extra_json``.
"""
import json
from dataclasses import asdict, fields as dc_fields

import numpy as np

from . import model, monitors

# Params fields of the reference (trench_charging_2d.py:103-174) that this
# package's Params also has; others are written to params_extra_json.
EXTRA_FIELDS = ("ion_distribution", "ion_angle_deg", "ion_energy_lo_eV", "ion_energy_hi_eV")


def _reference_params(params):
    """The run's parameters in the reference's Params vocabulary."""
    out = {k: v for k, v in params.items() if k not in EXTRA_FIELDS}
    if out.get("backend") == "hip":
        out["backend"] = "auto"      # the reference's Params has no "hip" backend
    return out


def _circuit_arrays(c):
    """The circuit waveform as ``rf_*`` arrays plus ``circuit_meta_json`` (its
    scalar entries; a sheath record's scalars nested), :1692-1703."""
    if c is None:
        return {}
    out = {f"rf_{k}": v for k, v in c.items() if isinstance(v, np.ndarray)}
    meta = {k: v for k, v in c.items() if not isinstance(v, np.ndarray) and k != "sheath"}
    sheath = c.get("sheath")
    meta["sheath"] = None if sheath is None else {k: v for k, v in sheath.items() if not isinstance(v, np.ndarray)}
    if sheath is not None:
        out.update({f"rf_sheath_{k}": v for k, v in sheath.items() if isinstance(v, np.ndarray)})
    out["circuit_meta_json"] = np.array(json.dumps(meta))
    return out


def _circuit_from(d):
    """Inverse of _circuit_arrays (:1769-1774)."""
    if "circuit_meta_json" not in d:
        return None
    c = json.loads(str(d["circuit_meta_json"]))
    c.update({k[3:]: d[k].copy() for k in d.files if k.startswith("rf_") and not k.startswith("rf_sheath_")})
    if c.get("sheath") is not None:
        c["sheath"].update({k[10:]: d[k].copy() for k in d.files if k.startswith("rf_sheath_")})
    return c


def save_results(res, out):
    """Write ``out + '.npz'`` (trench_charging_2d.py:1687-1734, the arrays only;
    with an RF circuit its waveform as well)."""
    params = res["params"]
    extra = {k: params[k] for k in EXTRA_FIELDS if k in params}
    path = f"{out}.npz"
    np.savez(path, phi=res["phi"], rho=res["rho"], Ex=res["Ex"], Ey=res["Ey"],
             mean_phi=res.get("mean_phi", res["phi"]), solid=res["solid"], mask=res["mask"], t_conv=res["t_conv"],
             rho_raw=res["rho_raw"], external_phi=res["external_phi"],
             params_json=np.array(json.dumps(_reference_params(params))),
             params_extra_json=np.array(json.dumps(extra)), stopped=res["stopped"],
             backend_used=np.array(res.get("backend_used", "hip")),
             checkpoint_json=np.array(json.dumps(res.get("checkpoint"))),
             convergence_meta_json=np.array(json.dumps(dict(
                 criterion_version=res.get("steady_criterion_version", 0),
                 note=res.get("convergence_note", ""), previous_t_conv=res.get("previous_t_conv")))),
             **{f"qmonitor_{k}": v for k, v in res.get("charge_monitor_hist", {}).items()},
             **{f"charge_{k}": v for k, v in res["charge_hist"].items()},
             **{f"hist_{k}": v for k, v in res["hist"].items()},
             **{f"steady_{k}": v for k, v in res["steady"].items()}, **_circuit_arrays(res.get("circuit")))
    return path


def load_results(filename):
    """Read a result NPZ (trench_charging_2d.py:1737-1801) written by the
    reference or by ``save_results``.  Returns (res, Params)."""
    with np.load(filename, allow_pickle=False) as d:
        if "params_json" not in d or "rho_raw" not in d:
            raise ValueError("result file lacks params_json / rho_raw (older format)")
        stored = json.loads(str(d["params_json"]))
        if "params_extra_json" in d:
            stored.update(json.loads(str(d["params_extra_json"])))
        names = {f.name for f in dc_fields(model.Params)}
        p = model.Params(**{k: v for k, v in stored.items() if k in names})
        geo = model.Geometry(p)
        res = {k: d[k].copy() for k in ("phi", "rho", "rho_raw", "Ex", "Ey", "external_phi", "solid", "mask")}
        res["mean_phi"] = d["mean_phi"].copy() if "mean_phi" in d else res["phi"].copy()
        if not np.array_equal(res["solid"], geo.solid) or not np.array_equal(res["mask"], geo.mask):
            raise ValueError("stored geometry does not match its parameters")
        for key in ("phi", "mean_phi", "rho", "rho_raw", "Ex", "Ey", "external_phi"):
            if res[key].shape != geo.solid.shape or not np.isfinite(res[key]).all():
                raise ValueError(f"stored {key} has a wrong shape or non-finite values")
        res.update(i0=geo.i0, i1=geo.i1, params=asdict(p), t_conv=float(d["t_conv"]), stopped=bool(d["stopped"]),
                   hist={k[5:]: d[k].copy() for k in d.files if k.startswith("hist_")},
                   steady={k[7:]: float(d[k]) for k in d.files if k.startswith("steady_")},
                   charge_hist={k[7:]: d[k].copy() for k in d.files if k.startswith("charge_")}, circuit=None)
        res["probes"] = list(res["steady"])
        res["converged"] = bool(np.isfinite(res["t_conv"]) and not res["stopped"])
        res["backend_used"] = str(d["backend_used"]) if "backend_used" in d else "cpu"
        res["charge_monitor_hist"] = {k[9:]: d[k].copy() for k in d.files if k.startswith("qmonitor_")}
        conv = json.loads(str(d["convergence_meta_json"])) if "convergence_meta_json" in d else {}
        res["steady_criterion_version"] = conv.get("criterion_version", 0)
        if "checkpoint_json" in d:
            cp = json.loads(str(d["checkpoint_json"]))
            if cp is not None:
                res["checkpoint"] = cp
        if conv.get("note"):
            res["convergence_note"] = conv["note"]
        if conv.get("previous_t_conv") is not None:
            res["previous_t_conv"] = conv["previous_t_conv"]
        res["circuit"] = _circuit_from(d)
    # A stored "converged" flag is re-checked against the charge histories
    # (trench_charging_2d.py:1780-1800).
    if res["converged"]:
        qhist = res["charge_monitor_hist"] or res["charge_hist"]
        valid = bool(qhist)
        for block in range(p.steady_hold) if valid else ():
            size = len(res["hist"]["t"]) - block * p.steady_window
            part = {k: q[:size] for k, q in qhist.items()}
            if size < 2 * p.steady_window or not monitors.check_charge_steady(part, p)["ok"]:
                valid = False
                break
        if not valid:
            res["previous_t_conv"] = res["t_conv"]
            res["t_conv"], res["converged"] = np.nan, False
            res["convergence_note"] = "stored charge is not saturated under the charge criterion"
    return res, p
