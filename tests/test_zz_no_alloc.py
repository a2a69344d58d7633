"""The allocation-free energy/CFL entry point against the allocating one.

amps_gpu_particle_energy_cfl_noalloc keeps its block partials in the
library's static scratch array and writes the result into a device array the
caller owns.  Called through the C ABI on two populations in a row, the
second smaller (fewer blocks and particles), it must give the allocating entry
point's energy and CFLs bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import mesh as M
from amps_amd.pic import PicState, Species

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _case(mesh, n_per_species, seed):
    E, B = M.corner_fields(mesh)
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    return E, B, sp, dt


def test_noalloc_energy_cfl_matches_allocating(gpu_ctx):
    hip = _hip()
    lib = L.load()
    dout = C.c_void_p()
    assert hip.hipMalloc(C.byref(dout), 8 * 9) == 0
    try:
        # first the larger population (2-level mesh, 120 blocks), then a
        # smaller one (one level, 64 blocks): the second call reads only the
        # partials it writes, whatever the first left in the scratch array
        for mesh, nps, seed in ((M.central_octant_mesh(), 120 * 512 * 4, 21), (M.uniform_mesh(2), 5000, 22)):
            E, B, sp, dt = _case(mesh, nps, seed)
            st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
            st.generate_particles(nps, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=seed)
            st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=True)
            print(f"noalloc energy/CFL: {mesh.nblocks} blocks, {st.particle_count()} particles", flush=True)
            e, cfl = st.ParticleEnergyCfl(dt)
            L.check(lib.amps_gpu_particle_energy_cfl_noalloc(gpu_ctx.h, float(dt), dout), "noalloc")
            gpu_ctx.sync()
            h = np.zeros(9)
            assert hip.hipMemcpy(h.ctypes.data, dout, 8 * 9, 2) == 0
            print(f"  energy {h[0]!r} vs {e!r}, cfl {h[1:3]} vs {cfl}", flush=True)
            assert h[0] == e and np.array_equal(h[1:1 + sp.n], cfl)
            assert e > 0 and np.all(cfl > 0)
    finally:
        hip.hipFree(dout)


def test_noalloc_refuses_unsorted_particles(gpu_ctx):
    """Without a compaction after the push the entry point would need one
    (which may allocate): it refuses instead."""
    mesh = M.uniform_mesh(2)
    E, B, sp, dt = _case(mesh, 2000, 3)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.generate_particles(2000, [0.05, 0.005], seed=3)
    st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=False)
    hip = _hip()
    dout = C.c_void_p()
    assert hip.hipMalloc(C.byref(dout), 8 * 9) == 0
    try:
        rc = L.load().amps_gpu_particle_energy_cfl_noalloc(gpu_ctx.h, float(dt), dout)
        assert rc != L.AMPS_OK
        assert b"compact" in L.load().amps_gpu_last_error()
    finally:
        hip.hipFree(dout)
