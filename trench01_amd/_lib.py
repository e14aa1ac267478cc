"""ctypes binding of libtrench_hip.so (C-ABI declared in include/trench_hip.h).

Loading fails loudly: there is no CPU fallback in this package.
"""
import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtrench_hip.so"

dp = ctypes.POINTER(ctypes.c_double)
i8p = ctypes.POINTER(ctypes.c_int8)
u8p = ctypes.POINTER(ctypes.c_uint8)
i32p = ctypes.POINTER(ctypes.c_int32)
i64p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p


class TraceParams(ctypes.Structure):
    """th_trace_params."""
    _fields_ = [("dx", ctypes.c_double), ("cfl", ctypes.c_double), ("max_steps", ctypes.c_int32),
                ("tapered", ctypes.c_int32), ("barrier", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("refill", ctypes.c_int32), ("chunk", ctypes.c_int32), ("order", ctypes.c_int32),
                ("prio_steps", ctypes.c_int32), ("check_steps", ctypes.c_int32), ("pass_steps", ctypes.c_int32),
                ("rf", ctypes.c_int32), ("blocks_per_cu", ctypes.c_int32), ("tile_pass", ctypes.c_int32)]


class Tally(ctypes.Structure):
    """th_tally."""
    _fields_ = [("escaped", ctypes.c_int64 * 2), ("lost", ctypes.c_int64), ("steps", ctypes.c_int64),
                ("off_surface", ctypes.c_int64), ("cancelled", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class ChargingSetupC(ctypes.Structure):
    """th_charging_setup."""
    _fields_ = [("n_cols", ctypes.c_int32), ("n_rows", ctypes.c_int32), ("col_of", i32p),
                ("rows", i32p), ("green", dp), ("external", dp), ("wface", dp),
                ("leak_M", dp), ("leak_c", dp), ("region", u8p), ("cond_s", u8p),
                ("g_cond_s", dp), ("q_macro", ctypes.c_double), ("dx", ctypes.c_double),
                ("C_cond", ctypes.c_double), ("Te", ctypes.c_double), ("dt_batch", ctypes.c_double),
                ("top_voltage", ctypes.c_double), ("substrate_voltage", ctypes.c_double),
                ("c_idx", ctypes.c_int32), ("floating", ctypes.c_int32), ("driven", ctypes.c_int32),
                ("top_dirichlet", ctypes.c_int32)]


class BatchDiag(ctypes.Structure):
    _fields_ = [("tally", Tally), ("leak_dq", ctypes.c_double), ("drift", ctypes.c_double),
                ("dv", ctypes.c_double), ("n_i_cond", ctypes.c_double), ("n_e_cond", ctypes.c_double),
                ("push_ms", ctypes.c_double), ("update_ms", ctypes.c_double)]


class InjectParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("lx", ctypes.c_double), ("y0", ctypes.c_double),
                ("qm_i", ctypes.c_double), ("qm_e", ctypes.c_double), ("ion_kind", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("e_ion", ctypes.c_double), ("e_lo", ctypes.c_double),
                ("e_hi", ctypes.c_double), ("sigma_theta", ctypes.c_double), ("vth_i", ctypes.c_double),
                ("v_per_sqrt_ev", ctypes.c_double), ("s_e", ctypes.c_double)]


class SoftEvent(ctypes.Structure):
    """th_soft_event."""
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("a", ctypes.c_double), ("da", ctypes.c_double), ("pid", ctypes.c_int32), ("cell", ctypes.c_int32),
                ("face", ctypes.c_int8), ("kind", ctypes.c_int8), ("species", ctypes.c_int8), ("pad0", ctypes.c_int8),
                ("pad1", ctypes.c_int32)]


# name -> (restype, argtypes); every symbol include/trench_hip.h declares.
SIGNATURES = {
    "th_last_error": (ctypes.c_char_p, []),
    "th_device_count": (ctypes.c_int, [i32p]),
    "th_tracer_create": (ctypes.c_int, [u8p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.POINTER(vp)]),
    "th_tracer_destroy": (None, [vp]),
    "th_tracer_device_name": (ctypes.c_int, [vp, ctypes.c_char_p, ctypes.c_int32]),
    "th_tracer_surface": (ctypes.c_int, [vp, i32p, i32p]),
    "th_set_fields": (ctypes.c_int, [vp, dp, dp, dp, dp]),
    "th_trace": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                i32p, i64p, i64p, ctypes.POINTER(Tally), i8p, i32p]),
    "th_trace_record": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                       i32p, i64p, i64p, ctypes.POINTER(Tally), i8p, i32p, dp, dp, i32p,
                                       ctypes.c_int32, dp, i32p]),
    "th_set_rf": (ctypes.c_int, [vp, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_double, ctypes.c_double]),
    "th_trace_rf": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, dp, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                   i32p, i64p, i64p, ctypes.POINTER(Tally), i8p, i32p]),
    "th_trace_rf_record": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, dp, ctypes.c_int64,
                                          ctypes.POINTER(TraceParams), i32p, i64p, i64p, ctypes.POINTER(Tally),
                                          i8p, i32p, dp, dp, i32p, ctypes.c_int32, dp, i32p, dp, dp]),
    "th_trace_tangent": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                        dp, dp, dp, dp, dp, dp, ctypes.POINTER(Tally), i8p, i32p, i32p, i8p, dp, dp,
                                        dp, dp]),
    "th_trace_soft": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(TraceParams), ctypes.c_double,
                                     ctypes.c_double, dp, dp, dp, dp, dp, dp, dp, ctypes.POINTER(Tally), dp, dp, i8p,
                                     i32p, i32p, ctypes.POINTER(SoftEvent), ctypes.c_int64, i64p]),
    "th_trace_soft_adjoint": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                             ctypes.c_double, ctypes.c_double, i32p, dp, dp, dp, dp, dp]),
    "th_trace_adjoint": (ctypes.c_int, [vp, dp, dp, dp, dp, dp, i8p, ctypes.c_int64, ctypes.POINTER(TraceParams),
                                        i32p, dp, dp, dp, dp, dp]),
    "th_particles_upload_times": (ctypes.c_int, [vp, ctypes.c_int64, dp, ctypes.c_int64]),
    "th_particles_reserve": (ctypes.c_int, [vp, ctypes.c_int64]),
    "th_particles_upload": (ctypes.c_int, [vp, ctypes.c_int64, dp, dp, dp, dp, dp, i8p, ctypes.c_int64]),
    "th_trace_resident": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(TraceParams), i64p,
                                         i64p, ctypes.POINTER(Tally), i8p, i32p]),
    "th_particles_read": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, dp, dp, dp, dp, dp, i8p]),
    "th_read_field": (ctypes.c_int, [vp, dp, dp]),
    "th_read_counts": (ctypes.c_int, [vp, i32p]),
    "th_last_push_ms": (ctypes.c_int, [vp, dp]),
    "th_selftest_math": (ctypes.c_int, [vp, dp, dp, ctypes.c_int64, dp]),
    "th_probe_copy": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int32, dp]),
    "th_lu_solve": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, i32p, i32p, dp, i32p, i32p, dp, dp, ctypes.c_int32,
                                   dp, dp, dp]),
    "th_inject": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.POINTER(InjectParams), ctypes.c_uint64,
                                 ctypes.c_uint64]),
    "th_charging_create": (ctypes.c_int, [vp, ctypes.POINTER(ChargingSetupC), ctypes.POINTER(vp)]),
    "th_charging_destroy": (None, [vp]),
    "th_charging_set_rho": (ctypes.c_int, [vp, dp]),
    "th_charging_set_q_macro": (ctypes.c_int, [vp, ctypes.c_double]),
    "th_charging_row_size": (ctypes.c_int, [vp, i32p]),
    "th_charging_potentials": (ctypes.c_int, [vp, ctypes.c_int32, dp, dp, dp]),
    "th_charging_run": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.POINTER(InjectParams), ctypes.c_uint64,
                                       ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(TraceParams), dp, dp]),
    "th_charging_set_green": (ctypes.c_int, [vp, ctypes.c_int32]),
    "th_charging_step": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64,
                                        ctypes.POINTER(TraceParams), ctypes.POINTER(BatchDiag)]),
    "th_charging_read": (ctypes.c_int, [vp, dp, dp, i32p]),
    "th_charging_step_replicas": (ctypes.c_int, [vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32,
                                                 ctypes.POINTER(TraceParams), dp, i64p, ctypes.POINTER(BatchDiag)]),
    "th_charging_set_monitors": (ctypes.c_int, [vp, ctypes.c_int32, i32p, ctypes.c_int32, i32p, i32p,
                                                ctypes.c_int32, ctypes.c_int32, i32p]),
    "th_charging_monitors": (ctypes.c_int, [vp, dp]),
    "th_charging_green_adjoint": (ctypes.c_int, [vp, dp, dp]),
    "th_charging_green_adjoint_noalloc": (ctypes.c_int, [vp, dp, dp]),
    "th_ensemble_create": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(vp)]),
    "th_ensemble_destroy": (None, [vp]),
    "th_ensemble_add": (ctypes.c_int, [vp, u8p, ctypes.c_int32, ctypes.c_int32, dp, dp, dp, dp, ctypes.c_int32,
                                       ctypes.c_int64, i32p]),
    "th_ensemble_set_fields": (ctypes.c_int, [vp, ctypes.c_int32, dp, dp, dp, dp]),
    "th_ensemble_surface": (ctypes.c_int, [vp, ctypes.c_int32, i32p, i32p]),
    "th_ensemble_set_inject": (ctypes.c_int, [vp, ctypes.c_int32, ctypes.POINTER(InjectParams)]),
    "th_ensemble_upload": (ctypes.c_int, [vp, ctypes.c_int32, dp, dp, dp, dp, dp, i8p, ctypes.c_int64]),
    "th_ensemble_run": (ctypes.c_int, [vp, ctypes.POINTER(TraceParams), ctypes.c_uint64, ctypes.c_uint64,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "th_ensemble_read": (ctypes.c_int, [vp, ctypes.c_int32, i64p, i64p, ctypes.POINTER(Tally), i8p, i32p]),
    "th_ensemble_totals": (ctypes.c_int, [vp, i64p, i64p]),
    "th_ensemble_read_particles": (ctypes.c_int, [vp, ctypes.c_int32, dp, dp, dp, dp, dp, i8p]),
    "th_ensemble_last_ms": (ctypes.c_int, [vp, dp]),
}

_lib = None
_lock = threading.Lock()


class HipError(RuntimeError):
    pass


def load():
    """Load the library (once).  Raises RuntimeError if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(f"{LIB_PATH.name} is not built; run __graft_entry__.build() "
                                   "or python -m trench01_amd.build")
            L = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = load().th_last_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise HipError(f"libtrench_hip error {rc}: {msg}")


def ptr(a, t):
    return None if a is None else a.ctypes.data_as(t)


def device_count():
    c = ctypes.c_int32(0)
    check(load().th_device_count(ctypes.byref(c)))
    return int(c.value)
