"""The oracle's ECSIM field equation pinned by what the mathematics fixes.

No reference vectors exist for the field solver (pic.h needs MPI and the
generated configuration), so the restatement of GetStencil_Legacy,
UpdateRhs_Legacy, UpdateMatrixElement, UpdateE and UpdateB
(src/pic/pic_field_solver_ecsim.cpp) is checked against closed forms:

* every Laplacian / grad-div stencil is the second derivative d_a d_b that
  it stands for, exactly, on every polynomial of degree <= 2, and even under
  reflection;
* without particles the operator is I + (c dt theta)^2 curl curl: the
  identity on linear fields, the analytic curl curl on quadratic ones, and
  symmetric;
* the mass-matrix rows are the mass term already pinned by the one-particle
  closed form (ecsim_mass_apply), and the current enters as -4 pi dt theta J;
* the curl-B part of the right-hand side is c dt theta curl B for a linear B,
  and UpdateB advances B by -c dt curl E exactly for a linear E;
* GMRES recovers a manufactured solution.
A wrong coefficient, sign, slot, component or axis swap fails one of them.
"""
import itertools

import numpy as np
import pytest

from amps_amd import mesh as M
from oracle import oracle as O


def _apply_stencil(st, f, pts):
    return sum(a * f(pts + np.array(o)) for o, a in st.items())


def test_stencils_are_second_derivatives_on_quadratics():
    lap, gd = O.ecsim_stencils()
    pts = np.random.default_rng(0).integers(-5, 5, (20, 3)).astype(float)
    mons = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for a, b in itertools.product(range(3), range(3)):
        st = gd[a][b]
        assert all(abs(o) <= 1 for off in st for o in off)
        # even under reflection
        for off, v in st.items():
            assert st.get(tuple(-o for o in off), 0.0) == v
        # degree <= 1 gives zero
        for f in (lambda p: np.ones(len(p)), lambda p: p[:, 0], lambda p: p[:, 1], lambda p: p[:, 2]):
            assert np.allclose(_apply_stencil(st, f, pts), 0.0, atol=1e-13)
        # d_a d_b (x_c x_d) = delta_ac delta_bd + delta_ad delta_bc
        for c, d in mons:
            want = float((a == c and b == d) + (a == d and b == c))
            got = _apply_stencil(st, lambda p: p[:, c] * p[:, d], pts)
            assert np.allclose(got, want, atol=1e-12), (a, b, c, d)
    for d in range(3):
        assert lap[d] is gd[d][d] or lap[d] == gd[d][d]


def _grid(level=1):
    mesh = M.uniform_mesh(level)
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    nbr = O.ecsim_neighbours(mesh, noc, keys)
    bnd = O.ecsim_boundary(mesh, keys)
    S = (1 << mesh.lmax) + 1
    g = np.stack([keys % S, (keys // S) % S, keys // (S * S)], 1)
    h = (1 << (mesh.lmax - level)) // 8
    xyz = mesh.xmin[None, :] + g * mesh.dx_maxref
    idx = g // h  # node index on the uniform node lattice
    return mesh, noc, keys, nbr, bnd, xyz, idx


def _deep(idx, level, margin=2):
    n = 8 << level
    return ((idx >= margin) & (idx <= n - margin)).all(1)


def test_operator_without_particles_is_identity_plus_curl_curl():
    mesh, noc, keys, nbr, bnd, xyz, idx = _grid(1)
    dx = mesh.cell_size()[0]
    cDt, theta = 0.07, 0.5
    K, R, coeff = O.ecsim_coefficients(dx, cDt, theta)
    M0 = np.zeros((len(keys), 243))
    deep = _deep(idx, 1)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    # linear field: A x = x
    E = np.stack([1 + 2 * x - y, 3 * z + 0.5, x + y + z], 1)
    Ez = np.where(bnd[:, None], 0.0, E)
    AE = O.ecsim_operator(nbr, bnd, M0, K, Ez, 4 * np.pi * 0.01 * theta)
    assert np.allclose(AE[deep], E[deep], rtol=0, atol=1e-12)
    # quadratic field: curl curl E = grad div E - lap E
    E = np.stack([y * y + x * z, x * y, x * z + z * z - y * y], 1)  # div = 2x + 3z, lap = (2, 0, 0)
    grad_div = np.stack([2 * np.ones_like(x), np.zeros_like(x), 3 * np.ones_like(x)], 1)
    lap = np.stack([2 * np.ones_like(x), np.zeros_like(x), np.zeros_like(x)], 1)
    want = E + (cDt * theta) ** 2 * (grad_div - lap)
    Ez = np.where(bnd[:, None], 0.0, E)
    AE = O.ecsim_operator(nbr, bnd, M0, K, Ez, 0.0)
    assert np.allclose(AE[deep], want[deep], rtol=0, atol=1e-11)
    # boundary rows are the identity
    assert np.array_equal(AE[bnd], Ez[bnd])
    # symmetric without particles
    rng = np.random.default_rng(3)
    u = np.where(bnd[:, None], 0.0, rng.normal(size=E.shape))
    v = np.where(bnd[:, None], 0.0, rng.normal(size=E.shape))
    au, av = O.ecsim_operator(nbr, bnd, M0, K, u, 0.0), O.ecsim_operator(nbr, bnd, M0, K, v, 0.0)
    assert abs(np.sum(v * au) - np.sum(u * av)) < 1e-11 * np.sum(np.abs(v * au))
    # and a wrong sign of the grad-div part breaks the quadratic closed form
    Kbad = K.copy()
    Kbad[0, 2] *= -1
    assert not np.allclose(O.ecsim_operator(nbr, bnd, M0, Kbad, Ez, 0.0)[deep], want[deep], atol=1e-6)


def test_operator_mass_rows_and_rhs_terms():
    mesh, noc, keys, nbr, bnd, xyz, idx = _grid(1)
    dx = mesh.cell_size()[0]
    cDt, theta, dt = 0.05, 0.5, 0.02
    K, R, coeff = O.ecsim_coefficients(dx, cDt, theta)
    rng = np.random.default_rng(5)
    nn = len(keys)
    Mm = rng.normal(size=(nn, 243))
    x = np.where(bnd[:, None], 0.0, rng.normal(size=(nn, 3)))
    fpdt = 4 * np.pi * dt * theta
    # the mass part of the operator is the pinned mass term over the kept neighbours
    nbr_kept = np.where((nbr >= 0) & ~bnd[np.maximum(nbr, 0)], nbr, -1)
    A = O.ecsim_operator(nbr, bnd, Mm, K, x, fpdt)
    A0 = O.ecsim_operator(nbr, bnd, np.zeros_like(Mm), K, x, fpdt)
    mass = O.ecsim_mass_apply(nbr_kept, Mm, x, fpdt, 1.0)
    inner = ~bnd
    assert np.allclose((A - A0)[inner], mass[inner], rtol=1e-12, atol=1e-12 * np.abs(mass).max())
    cells8 = O.ecsim_cells_around(mesh, noc, nbr)
    assert (cells8[inner] >= 0).all()
    Z = np.zeros((nn, 3))
    Bz0 = np.zeros((mesh.nblocks * 512, 3))
    # current term alone
    J = rng.normal(size=(nn, 3))
    rhs = O.ecsim_rhs(nbr, bnd, cells8, np.zeros_like(Mm), J, R, coeff, Z, Bz0, dt, theta)
    assert np.allclose(rhs[inner], -4 * np.pi * dt * theta * J[inner], rtol=1e-15, atol=0)
    assert np.all(rhs[bnd] == 0.0)
    # E^n alone: minus (A - I) E^n on deep rows, and the mass part with -4 pi dt theta
    E = np.stack([xyz[:, 1] ** 2, xyz[:, 0] * xyz[:, 2], xyz[:, 2] ** 2], 1)
    rhs = O.ecsim_rhs(nbr, bnd, cells8, np.zeros_like(Mm), Z, R, coeff, E, Bz0, dt, theta)
    deep = _deep(idx, 1)
    Af = O.ecsim_operator(nbr, np.zeros_like(bnd), np.zeros_like(Mm), K, E, 0.0)
    assert np.allclose(rhs[deep], -(Af - E)[deep], rtol=0, atol=1e-11)
    rhs_m = O.ecsim_rhs(nbr, bnd, cells8, Mm, Z, R, coeff, E, Bz0, dt, theta)
    mass_e = O.ecsim_mass_apply(nbr, Mm, E, -fpdt, 1.0)
    assert np.allclose((rhs_m - rhs)[inner], mass_e[inner], rtol=1e-12, atol=1e-12 * np.abs(mass_e).max())
    # curl B of a linear cell-centred B: c dt theta curl B
    G = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.6, 0.9, 1.5]])  # B_i = G[i] . x
    ctr = _cell_centres(mesh)
    Bc = ctr @ G.T
    rhs = O.ecsim_rhs(nbr, bnd, cells8, np.zeros_like(Mm), Z, R, coeff, Z, Bc, dt, theta)
    curl = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    assert np.allclose(rhs[inner], cDt * theta * curl[None, :], rtol=0, atol=1e-12)


def _cell_centres(mesh):
    out = np.zeros((mesh.nblocks * 512, 3))
    dx = mesh.cell_size()
    ii, jj, kk = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    loc = (ii + 8 * jj + 64 * kk).ravel()
    for b in range(mesh.nblocks):
        c = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1) + 0.5
        out[b * 512 + loc] = mesh.blk_xmin[b] + c * dx[b]
    return out


def test_update_e_and_b_closed_forms():
    mesh, noc, keys, nbr, bnd, xyz, idx = _grid(1)
    dx = mesh.cell_size()[0]
    cDt, theta = 0.04, 0.6
    rng = np.random.default_rng(9)
    nn = len(keys)
    En, dE = rng.normal(size=(nn, 3)), rng.normal(size=(nn, 3))
    Eth, E1 = O.ecsim_update_e(En, dE, theta)
    assert np.allclose(Eth, En + dE) and np.allclose(theta * E1 + (1 - theta) * En, Eth)
    G = np.array([[0.2, 1.1, -0.5], [0.4, -0.3, 2.0], [1.5, 0.6, 0.1]])  # E_i = G[i] . x
    Elin = xyz @ G.T
    Bc = rng.normal(size=(mesh.nblocks * 512, 3))
    B1 = O.ecsim_update_b(noc, Elin, Bc, dx, cDt)
    curl = np.array([G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]])
    assert np.allclose(B1 - Bc, -cDt * curl[None, :], rtol=0, atol=1e-12)


def test_gmres_recovers_a_manufactured_solution():
    mesh, noc, keys, nbr, bnd, xyz, idx = _grid(1)
    dx = mesh.cell_size()[0]
    K, R, coeff = O.ecsim_coefficients(dx, 0.06, 0.5)
    rng = np.random.default_rng(11)
    nn = len(keys)
    # a symmetric positive mass term of the kind a deposit gives: a small
    # multiple of the identity on every node's own slot
    Mm = np.zeros((nn, 243))
    Mm[:, [0, 4, 8]] = 0.3
    xs = np.where(bnd[:, None], 0.0, rng.normal(size=(nn, 3)))
    fpdt = 4 * np.pi * 0.01 * 0.5

    def apply(v):
        return O.ecsim_operator(nbr, bnd, Mm, K, v.reshape(nn, 3), fpdt).ravel()

    b = apply(xs.ravel())
    x, it, rel = O.ecsim_gmres(apply, b, 1e-10, 400, 40)
    assert rel <= 1e-10 and it < 400
    assert np.allclose(x.reshape(nn, 3), xs, atol=1e-8)
