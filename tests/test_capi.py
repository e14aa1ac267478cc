"""The C-ABI library loads and exports every symbol include/trench_hip.h declares
(CPU only: no compute call is made)."""
import ctypes
import re
from pathlib import Path

from trench01_amd import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "trench_hip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(th_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "th_trace" in names and "th_tracer_create" in names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared()) == set(_lib.SIGNATURES)


def test_device_count_without_gpu_is_not_an_error():
    assert _lib.device_count() >= 0


def test_struct_layouts_match_the_header(tmp_path):
    """Every ctypes struct of the binding has the size and field offsets the C
    compiler gives the header's struct (a short Python struct would let the
    library read past it)."""
    import subprocess
    structs = {"th_trace_params": _lib.TraceParams, "th_tally": _lib.Tally,
               "th_charging_setup": _lib.ChargingSetupC, "th_batch_diag": _lib.BatchDiag,
               "th_inject_params": _lib.InjectParams, "th_soft_event": _lib.SoftEvent}
    lines = ["#include <stdio.h>", "#include <stddef.h>", f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_every_entry_point_rejects_null_arguments():
    """Called with NULL handles and pointers and zero sizes, every entry point
    returns an error code with a message (the destroy functions return
    quietly) -- before any GPU work, so this runs without a GPU."""
    L = _lib.load()
    for name, (res, args) in _lib.SIGNATURES.items():
        if name == "th_last_error":
            continue
        call = []
        for a in args:
            if a in (ctypes.c_int, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64):
                call.append(0)
            elif a is ctypes.c_double:
                call.append(0.0)
            else:
                call.append(None)
        rc = getattr(L, name)(*call)
        if res is None:
            continue
        assert rc != 0, name
        assert L.th_last_error(), name
