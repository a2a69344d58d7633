"""ctypes binding of libampsgpu.so (the C ABI declared in include/ampsgpu.h).

The library is loaded from this package directory.  If it is missing, or the
machine has no GPU, every call fails loudly: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libampsgpu.so")

AMPS_OK = 0
CODE_REFLECTED = -1
CODE_LEFT_DOMAIN = 2
CODE_FINISHED = 3
CODE_SPHERE = 5
CODE_EXIT_FACE0 = 8  # BC_USER: left through face (code - 8), queued for ProcessOutsideDomainParticles
MOVER_LAPENTA = 0
MOVER_BORIS = 1
MOVER_RELATIVISTIC = 2
BC_DELETE = 0
BC_REFLECT = 1
BC_USER = 2
FP64 = 0
FP32 = 1
K_PUSH, K_COMPACT, K_DEPOSIT_CELL, K_DEPOSIT_NODE, K_TRACE, K_MLP, K_ECSIM = range(7)
K_COUNT = 7


class AmpsGpuError(RuntimeError):
    pass


class MeshDesc(C.Structure):
    _fields_ = [("xmin", C.c_double * 3), ("xmax", C.c_double * 3), ("lmax", C.c_int32),
                ("nblocks", C.c_int32)]


class PushOpts(C.Structure):
    _fields_ = [("mover", C.c_int32), ("bc_mode", C.c_int32), ("compact", C.c_int32),
                ("tile", C.c_int32), ("r_sphere", C.c_double)]


class AosLayout(C.Structure):
    """ParticleBuffer record layout (amps_aos_layout)."""
    _fields_ = [("record_bytes", C.c_int32), ("species_offset", C.c_int32), ("species_bytes", C.c_int32),
                ("velocity_offset", C.c_int32), ("position_offset", C.c_int32), ("weight_offset", C.c_int32)]


# the mandatory-field offsets of picParticleDataMacro.h in its two state-vector modes
def aos_layout_aligned(weight=False, extra_bytes=0):
    """_STATE_VECTOR_MODE_ALIGNED_: next, prev, species (long), v[3], [weight], x[3]."""
    w = 48 if weight else -1
    x = 56 if weight else 48
    return AosLayout(x + 24 + extra_bytes, 16, 8, 24, x, w)


def aos_layout_packed(weight=False, extra_bytes=0):
    """Unaligned state vector: next, prev, species (unsigned char), v[3], x[3], [weight]."""
    return AosLayout(17 + 48 + (8 if weight else 0) + extra_bytes, 16, 1, 17, 41, 65 if weight else -1)


class EcsimParams(C.Structure):
    """amps_ecsim_params: dtTotal, theta, c dt, E_conv, B_conv, length_conv."""
    _fields_ = [("dt", C.c_double), ("theta", C.c_double), ("cDt", C.c_double), ("E_conv", C.c_double),
                ("B_conv", C.c_double), ("length_conv", C.c_double)]


class TraceParams(C.Structure):
    _fields_ = [
        ("q_C", C.c_double), ("qabs_C", C.c_double), ("m0_kg", C.c_double),
        ("dt_trace_s", C.c_double), ("max_time_s", C.c_double), ("max_steps", C.c_int64),
        ("max_distance_m", C.c_double), ("dphi_max", C.c_double), ("adaptive", C.c_int32),
        ("dist_limiter", C.c_int32), ("dt_floor", C.c_int32), ("outer_sphere", C.c_int32),
        ("retry", C.c_int32), ("mover", C.c_int32), ("box_min_Re", C.c_double * 3),
        ("box_max_Re", C.c_double * 3), ("r_outer_Re", C.c_double), ("r_inner_Re", C.c_double),
        ("earth_radius_m", C.c_double), ("moment_Am2", C.c_double), ("m_hat", C.c_double * 3),
        ("B_ext_T", C.c_double * 3), ("field_model", C.c_int32), ("dt_selector", C.c_int32),
        ("policy", C.c_int32), ("pad2_", C.c_int32), ("event_tol_m", C.c_double),
        ("trap", C.c_int32), ("trap_min_mirror", C.c_int32), ("trap_min_bounce", C.c_int32), ("pad3_", C.c_int32),
        ("trap_outer_margin_Re", C.c_double), ("trap_radial_tol_Re", C.c_double), ("trap_energy_tol", C.c_double),
        ("trap_deadband", C.c_double),
    ]


class TraceResult(C.Structure):
    _fields_ = [("termination", C.c_int32), ("retry_count", C.c_int32), ("steps", C.c_int64),
                ("time_s", C.c_double), ("distance_m", C.c_double), ("x", C.c_double * 3),
                ("p", C.c_double * 3)]


TRACE_RESULT_DTYPE = np.dtype([("termination", np.int32), ("retry_count", np.int32),
                               ("steps", np.int64), ("time_s", np.float64),
                               ("distance_m", np.float64), ("x", np.float64, 3),
                               ("p", np.float64, 3)])
TRACE_EXIT_DTYPE = np.dtype([("x_exit_m", np.float64, 3), ("v_exit_unit", np.float64, 3),
                             ("cos_alpha", np.float64), ("valid", np.int32), ("pad_", np.int32)])

_lib = None


def _ptr(a, ctype):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays passed to the library must be C-contiguous (a strided view would be read "
                         "as if packed)")
    return a.ctypes.data_as(C.POINTER(ctype))


def load():
    """Load the HIP library (raises AmpsGpuError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AmpsGpuError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "amps_gpu_create": (C.c_int, [C.c_int, P(vp)]),
        "amps_gpu_destroy": (C.c_int, [vp]),
        "amps_gpu_last_error": (C.c_char_p, []),
        "amps_gpu_synchronize": (C.c_int, [vp]),
        "amps_gpu_device_info": (C.c_int, [vp, C.c_char_p, P(C.c_int), P(i64)]),
        "amps_gpu_set_param": (C.c_int, [vp, C.c_char_p, i64]),
        "amps_gpu_set_mesh": (C.c_int, [vp, P(MeshDesc), P(i32), P(i64)]),
        "amps_gpu_node_count": (i64, [vp]),
        "amps_gpu_get_node_map": (C.c_int, [vp, P(i32), P(i64)]),
        "amps_gpu_set_fields": (C.c_int, [vp, P(dbl), P(dbl)]),
        "amps_gpu_set_species": (C.c_int, [vp, i32, P(dbl), P(dbl), P(dbl), P(dbl)]),
        "amps_gpu_set_particles": (C.c_int, [vp, i64, P(dbl), P(dbl), P(i32), P(i32), P(i32),
                                             P(i64), i32]),
        "amps_gpu_particle_count": (i64, [vp]),
        "amps_gpu_get_particles_restart": (C.c_int, [vp, P(AosLayout), P(i32), vp, i64, P(i64)]),
        "amps_gpu_set_particles_restart": (C.c_int, [vp, P(AosLayout), P(i32), i64, vp, i32]),
        "amps_gpu_get_particles": (C.c_int, [vp, P(dbl), P(dbl), P(i32), P(i32), P(i32),
                                             P(i64), P(i32)]),
        "amps_gpu_move_particles": (C.c_int, [vp, P(PushOpts)]),
        "amps_gpu_compact": (C.c_int, [vp]),
        "amps_gpu_get_exits": (C.c_int, [vp, i64, P(i64), P(i64), P(dbl), P(dbl), P(i32), P(i32), P(i32), i32]),
        "amps_gpu_update_j_mass_matrix": (C.c_int, [vp, dbl, P(dbl), P(dbl)]),
        "amps_gpu_ecsim_mass_apply": (C.c_int, [vp, P(dbl), dbl, dbl, P(dbl), i32, i32]),
        "amps_gpu_particle_energy_cfl": (C.c_int, [vp, dbl, P(dbl), P(dbl)]),
        "amps_gpu_ecsim_operator": (C.c_int, [vp, P(EcsimParams), P(dbl), P(dbl), i32]),
        "amps_gpu_ecsim_rhs": (C.c_int, [vp, P(EcsimParams), P(dbl), P(dbl), P(dbl)]),
        "amps_gpu_ecsim_solve": (C.c_int, [vp, P(EcsimParams), P(dbl), dbl, i32, i32, P(dbl), P(i32), P(dbl)]),
        "amps_gpu_ecsim_update_b": (C.c_int, [vp, P(EcsimParams), P(dbl), P(dbl), P(dbl)]),
        "amps_gpu_ecsim_time_step": (C.c_int, [vp, P(EcsimParams), P(dbl), P(dbl), P(dbl), dbl, i32, i32, P(i32),
                                               P(dbl)]),
        "amps_gpu_particle_energy_cfl_noalloc": (C.c_int, [vp, dbl, C.c_void_p]),
        "amps_gpu_generate_particles": (C.c_int, [vp, i64, i32, P(dbl), P(dbl), C.c_uint64, i32]),
        "amps_gpu_run_steps": (C.c_int, [vp, P(PushOpts), dbl, i64, i32]),
        "amps_gpu_copy_probe": (C.c_int, [vp, i64, i32, P(dbl)]),
        "amps_gpu_set_center_fields": (C.c_int, [vp, P(dbl)]),
        "amps_gpu_interp_center": (C.c_int, [vp, i64, P(dbl), P(i32), P(dbl)]),
        "amps_gpu_set_mode3d_field": (C.c_int, [vp, P(dbl)]),
        "amps_gpu_mode3d_field": (C.c_int, [vp, i64, P(dbl), P(dbl)]),
        "amps_gpu_trace_batch": (C.c_int, [vp, P(TraceParams), i32, i64, P(dbl), P(dbl), P(dbl),
                                           vp, P(i64)]),
        "amps_gpu_trace_batch_ex": (C.c_int, [vp, P(TraceParams), i32, i64, P(dbl), P(dbl), P(dbl),
                                              vp, vp, P(i64)]),
        "amps_gpu_transmission": (C.c_int, [vp, P(TraceParams), i32, i64, P(dbl), i32, P(dbl), i32, P(dbl),
                                            P(dbl), P(i32), P(i32), P(i32), P(i32), P(i64)]),
        "amps_gpu_cutoff_scan": (C.c_int, [vp, P(TraceParams), i32, i32, i64, P(dbl), P(dbl), i32,
                                           P(dbl), P(i32), P(i32), P(C.c_int8), P(i32), P(i64)]),
        "amps_gpu_refine_cutoff": (C.c_int, [vp, P(TraceParams), i32, i64, P(dbl), P(dbl), P(dbl), P(dbl),
                                             dbl, dbl, i32, P(dbl), P(i32), P(i32), P(i64)]),
        "amps_gpu_effective_cutoff": (C.c_int, [vp, P(TraceParams), i32, i64, P(dbl), P(dbl), i32, P(dbl),
                                                P(C.c_int8), P(dbl), P(dbl), dbl, dbl, i32, P(dbl), P(dbl),
                                                P(i32), P(i32), P(i64)]),
        "amps_gpu_band_topology": (C.c_int, [i64, i32, P(C.c_int8), P(i32)]),
        "amps_gpu_trace_set_param": (C.c_int, [vp, C.c_char_p, i64]),
        "amps_gpu_uniform_sphere_dirs": (C.c_int, [i32, P(dbl)]),
        "amps_gpu_local_dirs": (C.c_int, [i64, P(dbl), i32, P(dbl), P(dbl), P(dbl), P(dbl)]),
        "amps_gpu_local_dirs_paired": (C.c_int, [i64, P(dbl), P(dbl), P(dbl), P(dbl)]),
        "amps_gpu_mlp_init": (C.c_int, [vp, P(C.c_float)]),
        "amps_gpu_mlp_get_params": (C.c_int, [vp, P(C.c_float), P(i64)]),
        "amps_gpu_mlp_set_data": (C.c_int, [vp, i64, P(C.c_float), P(C.c_float)]),
        "amps_gpu_mlp_grad": (C.c_int, [vp, i64, P(C.c_float), P(C.c_float)]),
        "amps_gpu_mlp_train": (C.c_int, [vp, i64, i32, dbl, P(C.c_float)]),
        "amps_gpu_mlp_predict": (C.c_int, [vp, i64, P(C.c_float), P(C.c_float)]),
        "amps_gpu_set_timing": (C.c_int, [vp, i32]),
        "amps_gpu_get_timing": (C.c_int, [vp, P(dbl), P(i64)]),
        "amps_gpu_reset_timing": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    """Names every entry point of the C ABI must export."""
    return sorted(_declared_names())


def _declared_names():
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "ampsgpu.h")
    txt = open(hdr).read()
    return set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(amps_gpu_\w+)\s*\(", txt, re.M))


def check(rc, what=""):
    if rc != AMPS_OK:
        msg = load().amps_gpu_last_error().decode(errors="replace")
        raise AmpsGpuError(f"{what}: status {rc}: {msg}")


class Context:
    """One library context on one device (owns all device buffers)."""

    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        check(lib.amps_gpu_create(device, C.byref(h)), "amps_gpu_create")
        self.h = h
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.amps_gpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def sync(self):
        check(self.lib.amps_gpu_synchronize(self.h), "synchronize")

    def device_info(self):
        name = C.create_string_buffer(64)
        ncu = C.c_int()
        hbm = C.c_int64()
        check(self.lib.amps_gpu_device_info(self.h, name, C.byref(ncu), C.byref(hbm)), "device_info")
        return name.value.decode(), ncu.value, hbm.value

    def copy_probe(self, nbytes, reps=10):
        ms = C.c_double()
        check(self.lib.amps_gpu_copy_probe(self.h, int(nbytes), int(reps), C.byref(ms)), "copy_probe")
        return ms.value

    def set_param(self, name, value):
        check(self.lib.amps_gpu_set_param(self.h, name.encode(), int(value)), f"set_param {name}")

    def set_timing(self, on):
        check(self.lib.amps_gpu_set_timing(self.h, 1 if on else 0), "set_timing")

    def reset_timing(self):
        check(self.lib.amps_gpu_reset_timing(self.h), "reset_timing")

    def timing(self):
        ms = np.zeros(K_COUNT, np.float64)
        n = np.zeros(K_COUNT, np.int64)
        check(self.lib.amps_gpu_get_timing(self.h, _ptr(ms, C.c_double), _ptr(n, C.c_int64)), "get_timing")
        return ms, n
