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
