"""Pins of the oracle's Mode3D compact-array field: the row stencil of
CellCentered::Linear::InitStencil (pic_interpolation_routines.cpp:224-1070
with a cRowStencil, pic.h:7295-7420) summed over interior-cell B as
InterpolateField_ (GlobalMagneticField.cpp:292-330).

* On a single-level mesh every ghost support cell is the neighbour's interior
  cell under its centre, so the row equals the ghost-layered stencil with the
  ghosts filled from the neighbours: bit for bit against the (pinned)
  cell-centred oracle.
* Constant fields are exact on AMR meshes (every branch's weights, merges and
  renormalisations sum to one).
* Linear fields are exact wherever each row entry's cell centre is the
  support point the stencil names; the reference's canonicalisation of a
  ghost cell into a cell of another level is the only departure, and it is
  rare and confined to rows that mix levels.
* The coarse/fine blend keeps a smooth field continuous.
"""
import numpy as np
import pytest

from amps_amd import mesh as M
from oracle import oracle as O

A0 = np.array([0.3, -1.2, 2.5])
G = np.array([[1.5, -0.7, 0.2], [0.4, 2.0, -1.1], [-0.9, 0.3, 0.8]])


def linear(P):
    return A0[None, :] + P @ G.T


def _ghosted_from_interior(mesh, Bm):
    """[nb][1000][3] ghost-layered array whose ghost cells hold the interior
    value of the cell (of any block) under their centre; NaN outside the
    domain."""
    g = np.arange(-1, M.NC + 1)
    k, j, i = np.meshgrid(g, g, g, indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    out = np.empty((mesh.nblocks, 1000, 3))
    dx = mesh.cell_size()
    for b in range(mesh.nblocks):
        X = mesh.blk_xmin[b] + (ijk + 0.5) * dx[b]
        inside = np.all((X >= mesh.xmin) & (X < mesh.xmax), axis=1)
        Xin = np.where(inside[:, None], X, mesh.xmin)
        owner = mesh.locate(Xin)
        cell = M.cell_of(mesh, Xin.T.copy(), owner.astype(np.int32))
        out[b] = np.where(inside[:, None], Bm[owner, cell], np.nan)
    return out


def test_single_level_row_equals_ghosted_stencil():
    mesh = M.uniform_mesh(2)
    om = O.Mesh(mesh)
    rng = np.random.default_rng(1)
    Bm = rng.normal(size=(mesh.nblocks, 512, 3))
    Bc = _ghosted_from_interior(mesh, Bm)
    x = rng.uniform(-1.0, 1.0 - 1e-12, size=(20000, 3))
    # a third of the points within a cell of a block-face plane
    k = len(x) // 3
    a = rng.integers(0, 3, k)
    x[np.arange(k), a] = rng.choice([-0.5, 0.0, 0.5], k) + rng.uniform(-0.06, 0.06, k)
    blk = mesh.locate(x).astype(np.int32)
    row = om.mode3d_field(Bm, x)
    ghost = om.interp_center(Bc, x.T.copy(), blk).T
    assert np.isfinite(ghost).all()
    assert np.array_equal(row, ghost)


@pytest.mark.parametrize("which", ["octant", "three_level"])
def test_constant_field_exact(which):
    mesh = M.central_octant_mesh() if which == "octant" else M.three_level_mesh()
    om = O.Mesh(mesh)
    Bm = M.interior_center_fields(mesh, lambda P: np.tile([1.25, -0.5, 3.0], (len(P), 1)))
    rng = np.random.default_rng(2)
    x = rng.uniform(-1.0, 1.0 - 1e-12, size=(20000, 3))
    B = om.mode3d_field(Bm, x)
    assert np.max(np.abs(B - [1.25, -0.5, 3.0])) < 1e-14


def test_linear_field_and_level_mixing():
    mesh = M.central_octant_mesh()
    om = O.Mesh(mesh)
    Bm = M.interior_center_fields(mesh, linear)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1 + 2.0 / 32, 1 - 2.0 / 32, size=(20000, 3))
    assert np.max(np.abs(om.mode3d_field(Bm, x) - linear(x))) < 1e-12

    mesh = M.three_level_mesh()
    om = O.Mesh(mesh)
    Bm = M.interior_center_fields(mesh, linear)
    x = rng.uniform(-1 + 2.0 / 32, 1 - 2.0 / 32, size=(20000, 3))
    err = np.max(np.abs(om.mode3d_field(Bm, x) - linear(x)), axis=1)
    bad = np.flatnonzero(err > 1e-12)
    assert len(bad) < 1e-3 * len(x), len(bad)
    dx = mesh.cell_size()
    for p in bad:
        b, c, w = om.row_stencil(x[p])
        assert len(set(mesh.level[b].tolist())) > 1
        # the stencil's centre of mass is off x: the departure is the
        # canonicalised entries' position, not the weights
        cen = mesh.blk_xmin[b] + (np.stack([c % 8, (c // 8) % 8, c // 64], 1) + 0.5) * dx[b]
        assert abs(w.sum() - 1.0) < 1e-14
        off = (w[:, None] * cen).sum(0) - x[p]
        assert np.allclose(linear(x[p][None] + off[None]), om.mode3d_field(Bm, x[p][None]), rtol=0, atol=1e-12)


def test_rows_merge_and_stay_positive():
    mesh = M.three_level_mesh()
    om = O.Mesh(mesh)
    rng = np.random.default_rng(4)
    x = rng.uniform(-0.4, 0.4, size=(3000, 3))
    longest = 0
    for p in range(len(x)):
        b, c, w = om.row_stencil(x[p])
        keys = b.astype(np.int64) * 512 + c
        assert len(np.unique(keys)) == len(keys)
        assert np.all(w > 0.0) and np.all((c >= 0) & (c < 512))
        longest = max(longest, len(w))
    assert longest > 8


def test_blend_continuous():
    """A smooth field along lines through faces of the refined octant."""
    mesh = M.central_octant_mesh()
    om = O.Mesh(mesh)

    def quad(P):
        return np.stack([P[:, 0] ** 2 + P[:, 1] * P[:, 2], np.sin(3 * P[:, 1]), P[:, 2] ** 2 - P[:, 0]], axis=1)

    Bm = M.interior_center_fields(mesh, quad)
    rng = np.random.default_rng(7)
    worst = 0.0
    for line in range(40):
        a = int(rng.integers(0, 3))
        side = float(rng.choice([-0.5, 0.5]))
        p0 = np.zeros(3)
        for d in range(3):
            p0[d] = -0.5 + 0.25 * rng.integers(0, 4) + rng.choice([0.04, 0.165]) + rng.uniform(0.0, 0.045)
        t = np.linspace(-0.06, 0.012, 1501)
        x = np.repeat(p0[None, :], len(t), axis=0)
        x[:, a] = side + t * np.sign(side)
        B = om.mode3d_field(Bm, x)
        dB = np.max(np.abs(np.diff(B, axis=0)), axis=1)
        worst = max(worst, float(np.max(dB) / np.median(dB)))
    assert worst < 4.0, worst
