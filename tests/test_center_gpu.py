"""GPU parity of the cell-centred linear interpolation
(CellCentered::Linear::InitStencil, pic_interpolation_routines.cpp:224-1070)
against the CPU oracle: bit-exact B on uniform and 2-level AMR meshes, at
random points, at points crowded around the coarse/fine interfaces, block
faces and the domain boundary, with an arbitrary (random) ghost fill."""
import numpy as np
import pytest

from amps_amd import mesh as M
from amps_amd.pic import PicState, Species
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _state(gpu_ctx, mesh, Bc):
    E, B = M.corner_fields(mesh)
    st = PicState(mesh, E, B, Species([1.0], [1.0], [0.01]), ctx=gpu_ctx)
    st.set_center_fields(Bc)
    return st


def _points(mesh, rng, n):
    x = rng.uniform(mesh.xmin, mesh.xmax, size=(n, 3))
    # a third of them within a fine cell of a plane of block faces
    k = n // 3
    a = rng.integers(0, 3, k)
    planes = np.array([-1.0, -0.5, -0.25, 0.0, 0.25, 0.5, 1.0])
    x[np.arange(k), a] = rng.choice(planes, k) + rng.uniform(-0.03, 0.03, k)
    # the upper domain face is open: (1 - 2**-53 - xmin) / size rounds to 1
    x = np.clip(x, -1.0, 1.0 - 1e-12)
    blk = mesh.locate(x).astype(np.int32)
    assert np.all(blk >= 0)
    return x.T.copy(), blk


@pytest.mark.parametrize("amr", [False, True])
def test_center_interp_bitexact(gpu_ctx, amr):
    mesh = M.central_octant_mesh() if amr else M.uniform_mesh(2)
    rng = np.random.default_rng(31 if amr else 32)
    Bc = rng.normal(size=(mesh.nblocks, 1000, 3))
    print(f"center interp case amr={amr} blocks={mesh.nblocks}", flush=True)
    st = _state(gpu_ctx, mesh, Bc)
    x, blk = _points(mesh, rng, 60000)
    g = st.interp_center(x, blk)
    o = O.Mesh(mesh).interp_center(Bc, x, blk)
    bad = np.flatnonzero(np.any(g != o, axis=0))
    assert len(bad) == 0, (len(bad), x[:, bad[:3]].T, g[:, bad[:3]].T, o[:, bad[:3]].T)


def test_center_interp_linear_field(gpu_ctx):
    """The device stencil reproduces a linear field on the AMR mesh."""
    mesh = M.central_octant_mesh()
    A0 = np.array([0.3, -1.2, 2.5])
    G = np.array([[1.5, -0.7, 0.2], [0.4, 2.0, -1.1], [-0.9, 0.3, 0.8]])
    Bc = M.center_fields(mesh, lambda P: A0[None, :] + P @ G.T)
    st = _state(gpu_ctx, mesh, Bc)
    rng = np.random.default_rng(33)
    x = rng.uniform(-1 + 2.0 / 32, 1 - 2.0 / 32, size=(50000, 3))
    g = st.interp_center(x.T.copy(), mesh.locate(x).astype(np.int32))
    assert np.max(np.abs(g - (A0[None, :] + x @ G.T).T)) < 1e-12


@pytest.mark.parametrize("mover", [0, 1])
def test_push_center_B_bitexact(gpu_ctx, mover):
    """Lapenta/Boris push with the cell-centred B (centre-based B mode)."""
    from amps_amd import _lib as L
    mesh = M.central_octant_mesh()
    E, Bk = M.corner_fields(mesh)
    rng = np.random.default_rng(41)
    Bc = M.center_fields(mesh, lambda P: M.dipole_field(P))
    parts = M.uniform_particles(mesh, 30000, 0.5, seed_x=42, seed_v=43)
    sp = Species([1.0], [1.0], [0.05])
    st = PicState(mesh, E, Bk, sp, ctx=gpu_ctx)
    st.set_center_fields(Bc)
    gpu_ctx.set_param("b_center", 1)
    try:
        st.set_particles(parts)
        st.MoveParticles(mover=mover, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=False)
        g, gcode = st.get_particles()
    finally:
        gpu_ctx.set_param("b_center", 0)
    om = O.Mesh(mesh)
    om.set_center_field(Bc)
    o, ocode = O.push(om, E, Bk, sp, parts, mover=mover, bc_mode=1, r_sphere=0.2)
    gi, oi = np.argsort(g.id), np.argsort(o.id)
    assert np.array_equal(gcode[gi], ocode[oi])
    assert np.array_equal(g.x[:, gi], o.x[:, oi]) and np.array_equal(g.v[:, gi], o.v[:, oi])


@pytest.mark.parametrize("exact", [1, 0])
def test_deposit_center_B(gpu_ctx, exact):
    """ECSIM deposit with the cell-centred B: bit-exact (exact mode) or
    within 1e-12 per node."""
    mesh = M.central_octant_mesh()
    E, Bk = M.corner_fields(mesh)
    Bc = M.center_fields(mesh, lambda P: M.dipole_field(P))
    parts = M.per_cell_particles(mesh, [4, 4], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    dt = 0.1 / (np.max(np.linalg.norm(Bk, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    st = PicState(mesh, E, Bk, sp, ctx=gpu_ctx)
    st.set_center_fields(Bc)
    st.set_particles(parts)
    gpu_ctx.set_param("b_center", 1)
    gpu_ctx.set_param("deposit_exact", exact)
    try:
        J, Mm = st.UpdateJMassMatrix(dt)
        g, _ = st.get_particles()
    finally:
        gpu_ctx.set_param("b_center", 0)
        gpu_ctx.set_param("deposit_exact", 0)
    om = O.Mesh(mesh)
    om.set_center_field(Bc)
    noc, keys = om.build_nodes()
    Jo, Mo = O.deposit(om, Bk, sp, dt, g, noc, len(keys))
    if exact:
        assert np.array_equal(J, Jo) and np.array_equal(Mm, Mo)
    else:
        nj = np.linalg.norm(J - Jo, axis=1) / np.maximum(np.linalg.norm(Jo, axis=1), 1e-300)
        nm = np.linalg.norm(Mm - Mo, axis=1) / np.maximum(np.linalg.norm(Mo, axis=1), 1e-300)
        assert nj.max() <= 1e-12 and nm.max() <= 1e-12, (nj.max(), nm.max())
