"""The C-ABI library loads and exports every entry point include/ampsgpu.h
declares, and the Python binding declares each with the header's signature
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
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


def test_uniform_sphere_dirs_closed_forms():
    """BuildUniformSphereDirs: unit vectors with z = 1 - 2(k + 1/2)/n and the
    azimuth advancing by the golden angle 2 pi (1 - 1/phi); near-uniform
    (every octant within 3 % of n/8, mean vector ~0)."""
    import math

    from amps_amd import _lib as L
    from amps_amd import cutoff as CU

    n = 4000
    d = CU.uniform_sphere_dirs(n)
    k = np.arange(n)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, rtol=0, atol=1e-15)
    assert np.allclose(d[:, 2], 1.0 - 2.0 * (k + 0.5) / n, rtol=0, atol=1e-15)
    ga = 2 * math.pi * (1 - 2 / (1 + math.sqrt(5)))
    phi = np.arctan2(d[:, 1], d[:, 0])
    step = np.mod(np.diff(phi) - ga + math.pi, 2 * math.pi) - math.pi
    assert np.abs(step).max() < 1e-9
    octant = (d[:, 0] > 0) * 4 + (d[:, 1] > 0) * 2 + (d[:, 2] > 0)
    assert np.all(np.abs(np.bincount(octant, minlength=8) - n / 8) < 0.03 * n / 8)
    assert np.linalg.norm(d.mean(0)) < 2e-3
    assert np.array_equal(CU.uniform_sphere_dirs(1), [[1.0, 0.0, 0.0]])
    with pytest.raises(L.AmpsGpuError):
        CU.uniform_sphere_dirs(0)


def test_local_enu_frame_and_sky_directions():
    """BuildLocalENU_GSM / LocalLonLatToDir_GSM: a right-handed orthonormal
    frame with U radial and E horizontal (perpendicular to +Z), the +X
    fallback at the pole; lon 0 / lat 0 is +N, lon 90 is +E, lon 180 is -N,
    lat 90 is +U, and a general (lon, lat) has the stated components."""
    from amps_amd import cutoff as CU

    rng = np.random.default_rng(3)
    x0 = rng.normal(size=(200, 3)) * 7e6
    x0[0] = [0.0, 0.0, 6.4e6]
    lon = np.array([0.0, 90.0, 180.0, 0.0, 37.0])
    lat = np.array([0.0, 0.0, 0.0, 90.0, -21.0])
    d, f = CU.local_dirs(x0, lon, lat)
    E, N, U = f[:, 0], f[:, 1], f[:, 2]
    for a in (E, N, U):
        assert np.allclose(np.linalg.norm(a, axis=1), 1.0, atol=1e-15)
    assert np.allclose((E * N).sum(1), 0, atol=1e-15) and np.allclose((E * U).sum(1), 0, atol=1e-15)
    assert np.allclose(np.cross(E, N), U, atol=1e-15)
    assert np.allclose(U, x0 / np.linalg.norm(x0, axis=1)[:, None], atol=1e-15)
    assert np.allclose(E[1:, 2], 0.0, atol=1e-15)
    assert np.allclose(E[0], [0.0, -1.0, 0.0], atol=1e-15)  # unit(X x U) at the pole
    assert np.allclose(d[:, 0], N, atol=1e-15) and np.allclose(d[:, 1], E, atol=1e-15)
    assert np.allclose(d[:, 2], -N, atol=1e-15) and np.allclose(d[:, 3], U, atol=1e-15)
    cl, sl = np.cos(np.radians(-21.0)), np.sin(np.radians(-21.0))
    want = cl * np.sin(np.radians(37.0)) * E + cl * np.cos(np.radians(37.0)) * N + sl * U
    assert np.allclose(d[:, 4], want, atol=1e-15)


def test_pointer_helpers_reject_strided_arrays():
    """A strided view handed to the C ABI would be read as if packed: the
    binding and the oracle refuse it instead."""
    import numpy as np

    from amps_amd import _lib as L
    from oracle import oracle as O

    a = np.zeros((3, 8))[:, ::2]
    with pytest.raises(ValueError):
        L._ptr(a, L.C.c_double)
    with pytest.raises(ValueError):
        O._p(a, L.C.c_double)
    b = np.zeros((3, 8))[:, np.arange(4)]  # advanced indexing can return a Fortran-ordered copy
    if not b.flags["C_CONTIGUOUS"]:
        with pytest.raises(ValueError):
            L._ptr(b, L.C.c_double)
    assert L._ptr(np.ascontiguousarray(b), L.C.c_double) is not None


def _prototypes(header):
    """name -> (return type, [parameter declarations]) of every entry point."""
    txt = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", header)).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(int|int64_t|const char\s*\*)\s+(amps_\w+)\s*\(([^;]*?)\)\s*;", txt, re.M | re.S):
        params = [p.strip() for p in m.group(3).replace("\n", " ").split(",")]
        out[m.group(2)] = (m.group(1).replace(" ", ""), [] if params in (["void"], [""]) else params)
    return out


def _kind_of_decl(p):
    if "*" in p or "[" in p:
        return "ptr"
    t = p.rsplit(" ", 1)[0].replace("const ", "").strip()
    return {"double": "f64", "float": "f32", "int64_t": "i64", "uint64_t": "i64", "int32_t": "i32", "int": "i32",
            "uint32_t": "i32"}.get(t, "other " + t)


def _kind_of_ctype(a):
    if a in (ctypes.c_void_p, ctypes.c_char_p) or (isinstance(a, type) and issubclass(a, ctypes._Pointer)):
        return "ptr"
    return {ctypes.c_double: "f64", ctypes.c_float: "f32", ctypes.c_int64: "i64", ctypes.c_uint64: "i64",
            ctypes.c_int32: "i32", ctypes.c_uint32: "i32"}.get(a, f"other {a}")


@pytest.mark.skipif(not os.path.exists(LIB), reason="libampsgpu.so not built")
def test_binding_signatures_match_the_header():
    """The Python binding declares every entry point with the header's
    parameter list (pointer, 32- or 64-bit integer, double, in order) and
    return type: a mismatch would pass arguments in the wrong registers."""
    from amps_amd import _lib as L

    lib = L.load()
    protos = _prototypes("ampsgpu.h")
    assert len(protos) == len(declared("ampsgpu.h"))
    ret = {"int": ctypes.c_int, "int64_t": ctypes.c_int64, "constchar*": ctypes.c_char_p}
    for name, (rtype, params) in sorted(protos.items()):
        f = getattr(lib, name)
        assert f.argtypes is not None, f"{name} has no binding"
        assert [_kind_of_decl(p) for p in params] == [_kind_of_ctype(a) for a in f.argtypes], name
        assert f.restype == ret[rtype], name


def test_binding_structs_match_the_header(tmp_path):
    """Every parameter and result struct of the header has the same size and
    field offsets in the Python binding (ctypes structures and the numpy
    record types of the trace results and exits), as the C compiler lays
    the header's structs out."""
    import shutil
    import subprocess

    from amps_amd import _lib as L

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = [("amps_mesh_desc", L.MeshDesc), ("amps_push_opts", L.PushOpts), ("amps_aos_layout", L.AosLayout),
             ("amps_ecsim_params", L.EcsimParams), ("amps_trace_params", L.TraceParams),
             ("amps_trace_result", L.TraceResult), ("amps_trace_result", L.TRACE_RESULT_DTYPE),
             ("amps_trace_exit", L.TRACE_EXIT_DTYPE)]
    want, lines = [], []
    for cname, py in pairs:
        if isinstance(py, np.dtype):
            fields = [(f, py.fields[f][1]) for f in py.names]
            size = py.itemsize
        else:
            fields = [(f, getattr(py, f).offset) for f, *_ in py._fields_]
            size = ctypes.sizeof(py)
        want.append(f"{cname} sizeof {size}")
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f, off in fields:
            want.append(f"{cname} {f} {off}")
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"ampsgpu.h\"\nint main(void) {\n"
                   + "\n".join(lines) + "\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE).stdout.decode().split("\n")[:-1]
    assert got == want
