"""The C-ABI library loads and exports every entry point include/ampsgpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "amps_amd", "libampsgpu.so")


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\s*\*)\s+(amps_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = declared("ampsgpu.h")
    for n in ["amps_gpu_create", "amps_gpu_move_particles", "amps_gpu_update_j_mass_matrix",
              "amps_gpu_compact", "amps_gpu_set_mesh", "amps_gpu_last_error"]:
        assert n in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libampsgpu.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared("ampsgpu.h") if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libampsgpu.so not built")
def test_error_path_without_context():
    lib = ctypes.CDLL(LIB)
    lib.amps_gpu_synchronize.argtypes = [ctypes.c_void_p]
    lib.amps_gpu_last_error.restype = ctypes.c_char_p
    assert lib.amps_gpu_synchronize(None) != 0
    assert b"null context" in lib.amps_gpu_last_error()
