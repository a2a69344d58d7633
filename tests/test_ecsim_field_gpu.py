"""The ECSIM field equation on the device against the oracle's restatement.

The J and M come from a device deposit of a Maxwellian population and are fed
to both sides, so the operator (GetStencil_Legacy + UpdateMatrixElement), the
right-hand side (UpdateRhs_Legacy) and UpdateB are compared bit for bit; the
GMRES solve is checked by its residual, against the oracle's own GMRES and
against a manufactured solution; the time step against its parts.
"""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import mesh as M
from amps_amd.pic import PicState, Species
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _case(gpu_ctx, level=1, ppc=4):
    mesh = M.uniform_mesh(level)
    E, B = M.corner_fields(mesh)
    parts = M.per_cell_particles(mesh, [ppc, ppc], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    J, Mm = st.UpdateJMassMatrix(dt)
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    nbr = O.ecsim_neighbours(mesh, noc, keys)
    bnd = O.ecsim_boundary(mesh, keys)
    cells8 = O.ecsim_cells_around(mesh, noc, nbr)
    return dict(mesh=mesh, st=st, J=J, M=Mm, noc=noc, keys=keys, nbr=nbr, bnd=bnd, cells8=cells8, dt=dt)


def _oracle_tables(c, p):
    dx = c["mesh"].cell_size()[0]
    return O.ecsim_coefficients(dx, p.cDt, p.theta, p.length_conv)


def test_operator_rhs_update_b_bitexact(gpu_ctx):
    c = _case(gpu_ctx)
    st = c["st"]
    nn = len(c["keys"])
    print(f"ECSIM field equation: {c['mesh'].nblocks} blocks, {nn} nodes", flush=True)
    rng = np.random.default_rng(4)
    for p in (st.ecsim_params(c["dt"], theta=0.5, cDt=40 * c["dt"]),
              st.ecsim_params(c["dt"], theta=0.6, cDt=25 * c["dt"], E_conv=3.0e-3, B_conv=1.0e4, length_conv=1e2)):
        K, R, coeff = _oracle_tables(c, p)
        fpdt = 4 * np.pi * p.dt * p.theta
        x = rng.normal(size=(nn, 3))
        y = st.EcsimOperator(p, x)
        yo = O.ecsim_operator(c["nbr"], c["bnd"], c["M"], K, x, fpdt)
        assert np.array_equal(y, yo)
        En = rng.normal(size=(nn, 3))
        Bc = rng.normal(size=(c["mesh"].nblocks * 512, 3))
        r = st.EcsimRhs(p, En, Bc)
        ro = O.ecsim_rhs(c["nbr"], c["bnd"], c["cells8"], c["M"], c["J"], R, coeff, En, Bc, p.dt, p.theta,
                         p.E_conv, p.B_conv)
        assert np.array_equal(r, ro)
        B1 = st.EcsimUpdateB(p, En, Bc)
        B1o = O.ecsim_update_b(c["noc"], En, Bc, c["mesh"].cell_size()[0], p.cDt, p.E_conv, p.B_conv, p.length_conv)
        assert np.array_equal(B1, B1o)


def test_gmres_solution(gpu_ctx):
    c = _case(gpu_ctx, level=2)
    st = c["st"]
    nn = len(c["keys"])
    # c dt theta / dx = 2: the curl-curl part dominates the identity, as in a
    # run whose light crossing of a cell takes a fraction of a step
    dx = c["mesh"].cell_size()[0][0]
    p = st.ecsim_params(c["dt"], theta=0.5, cDt=4 * dx)
    K, R, coeff = _oracle_tables(c, p)
    fpdt = 4 * np.pi * p.dt * p.theta
    rng = np.random.default_rng(8)
    xs = np.where(c["bnd"][:, None], 0.0, rng.normal(size=(nn, 3)))
    b = O.ecsim_operator(c["nbr"], c["bnd"], c["M"], K, xs, fpdt)
    dE, it, rr = st.EcsimSolve(p, b, tol=1e-11, max_iter=600, restart=40)
    print(f"GMRES on {nn} nodes: {it} iterations, relative residual {rr:.2e}", flush=True)
    assert rr <= 1e-11 and 10 < it < 600
    res = O.ecsim_operator(c["nbr"], c["bnd"], c["M"], K, dE, fpdt) - b
    assert np.linalg.norm(res) <= 2e-11 * np.linalg.norm(b)
    assert np.abs(dE - xs).max() <= 1e-8 * np.abs(xs).max()
    # the physical right-hand side: the oracle's GMRES reaches the same solution
    En = M.corner_fields(c["mesh"])[0]
    Enode = np.zeros((nn, 3))
    Enode[c["noc"].reshape(-1, 729).ravel()] = En.reshape(-1, 3)
    Bc = rng.normal(scale=1e-3, size=(c["mesh"].nblocks * 512, 3))
    rhs = st.EcsimRhs(p, Enode, Bc)
    dE, it, rr = st.EcsimSolve(p, rhs, tol=1e-10, max_iter=600, restart=40)

    def apply(v):
        return O.ecsim_operator(c["nbr"], c["bnd"], c["M"], K, v.reshape(nn, 3), fpdt).ravel()

    xo, ito, rro = O.ecsim_gmres(apply, rhs.ravel(), 1e-10, 600, 40)
    print(f"GMRES physical rhs: device {it} iterations ({rr:.2e}), oracle {ito} ({rro:.2e})", flush=True)
    assert rr <= 1e-10 and rro <= 1e-10
    assert np.abs(dE.ravel() - xo).max() <= 1e-8 * max(np.abs(xo).max(), 1e-300)


def test_time_step_composition(gpu_ctx):
    c = _case(gpu_ctx)
    st = c["st"]
    nn = len(c["keys"])
    dx = c["mesh"].cell_size()[0][0]
    p = st.ecsim_params(c["dt"], theta=0.5, cDt=3 * dx)
    rng = np.random.default_rng(12)
    En = rng.normal(size=(nn, 3))
    Bc = rng.normal(size=(c["mesh"].nblocks * 512, 3))
    E1, Eth, B1, it, rr = st.EcsimTimeStep(p, En, Bc, tol=1e-12, max_iter=400, restart=40)
    print(f"ECSIM time step: {it} GMRES iterations, relative residual {rr:.2e}", flush=True)
    assert rr <= 1e-12 and it > 10
    # the same right-hand side and the same deterministic GMRES through the
    # separate entry points give the same increment, bit for bit
    rhs = st.EcsimRhs(p, En, Bc)
    dE, it2, rr2 = st.EcsimSolve(p, rhs, tol=1e-12, max_iter=400, restart=40)
    assert it2 == it and rr2 == rr
    Eth_o, E1_o = O.ecsim_update_e(En, dE, p.theta, p.E_conv)
    assert np.array_equal(Eth, Eth_o) and np.array_equal(E1, E1_o)
    assert np.array_equal(B1, O.ecsim_update_b(c["noc"], Eth, Bc, c["mesh"].cell_size()[0], p.cDt))
    # boundary corners keep E^n in E^{n+theta} (a zero increment)
    assert np.array_equal(Eth[c["bnd"]], En[c["bnd"]])


def test_field_equation_needs_one_level(gpu_ctx):
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    parts = M.per_cell_particles(mesh, [1, 1], [0.05, 0.005])
    sp = Species([-1.0, 1.0], [0.01, 1.0], [1e-3, 1e-3])
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    st.UpdateJMassMatrix(1e-3, copy_out=False)
    with pytest.raises(L.AmpsGpuError, match="single-level"):
        st.EcsimOperator(st.ecsim_params(1e-3), np.zeros((st.nnodes, 3)))
