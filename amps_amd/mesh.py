"""Block-structured AMR meshes and synthetic field/particle workloads.

A mesh is the list of leaf blocks of the AMPS octree (cMeshAMR3d): block b has
refinement level ``level[b]`` and integer position ``ijk[b]`` on that level's
lattice, 8^3 cells per block.  Leaves are ordered along the Morton curve with x
the fastest bit, i.e. the depth-first leaf order of the octree whose children
are numbered i+2(j+2k) (meshAMRgeneric.h:2793-2848, 11587-11620).
"""
import numpy as np

NC = 8
NN = 9
LMAX_DEFAULT = 15  # _MAX_REFINMENT_LEVEL_ default, meshAMRdef.h:108


def _morton(ix, iy, iz, bits):
    key = np.zeros(ix.shape, np.int64)
    for b in range(bits - 1, -1, -1):
        key = (key << 3) | (((iz >> b) & 1) << 2) | (((iy >> b) & 1) << 1) | ((ix >> b) & 1)
    return key


class BlockMesh:
    def __init__(self, level, ijk, xmin=(-1.0, -1.0, -1.0), xmax=(1.0, 1.0, 1.0), lmax=LMAX_DEFAULT):
        level = np.asarray(level, np.int32)
        ijk = np.asarray(ijk, np.int64).reshape(-1, 3)
        g = ijk << (lmax - level)[:, None].astype(np.int64)
        order = np.argsort(_morton(g[:, 0], g[:, 1], g[:, 2], lmax), kind="stable")
        self.level = np.ascontiguousarray(level[order])
        self.ijk = np.ascontiguousarray(ijk[order])
        self.xmin = np.array(xmin, np.float64)
        self.xmax = np.array(xmax, np.float64)
        self.lmax = int(lmax)
        self.nblocks = len(self.level)
        dxr = self.xmax - self.xmin
        self.dx_maxref = dxr / (1 << lmax)
        gl = self.ijk << (lmax - self.level)[:, None].astype(np.int64)
        size = (np.int64(1) << (lmax - self.level)).astype(np.int64)
        self.blk_xmin = self.xmin + gl.astype(np.float64) * self.dx_maxref
        self.blk_xmax = self.xmin + (gl + size[:, None]).astype(np.float64) * self.dx_maxref
        self.lf = int(self.level.max())
        nf = 1 << self.lf
        leaf = -np.ones((nf, nf, nf), np.int32)  # [z][y][x]
        for b in range(self.nblocks):
            s = 1 << (self.lf - self.level[b])
            i, j, k = (self.ijk[b] * s).tolist()
            leaf[k:k + s, j:j + s, i:i + s] = b
        if (leaf < 0).any():
            raise ValueError("blocks do not cover the domain")
        self.leaf = leaf

    @property
    def ncells(self):
        return self.nblocks * NC ** 3

    def cell_size(self):
        return (self.blk_xmax - self.blk_xmin) / NC

    def corner_coords(self):
        """[nb, 729, 3] coordinates of each block's corner nodes (i fastest)."""
        k, j, i = np.meshgrid(np.arange(NN), np.arange(NN), np.arange(NN), indexing="ij")
        ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
        dx = self.cell_size()
        return self.blk_xmin[:, None, :] + ijk[None, :, :] * dx[:, None, :]

    def locate(self, x):
        """Block containing each point (half-open blocks), -1 outside."""
        x = np.asarray(x)
        nf = 1 << self.lf
        g = np.floor((x - self.xmin) / (self.xmax - self.xmin) * nf).astype(np.int64)
        ok = np.all((g >= 0) & (g < nf), axis=1)
        g = np.clip(g, 0, nf - 1)
        b = self.leaf[g[:, 2], g[:, 1], g[:, 0]]
        return np.where(ok, b, -1)

    def desc_arrays(self):
        return self.level.astype(np.int32), self.ijk.astype(np.int64).ravel()


def uniform_mesh(level, lmax=LMAX_DEFAULT, xmin=(-1.0, -1.0, -1.0), xmax=(1.0, 1.0, 1.0)):
    n = 1 << level
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ijk = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    return BlockMesh(np.full(len(ijk), level), ijk, xmin, xmax, lmax)


def refine(level, ijk, predicate):
    """Split every block whose bounds (lo, hi arrays) satisfy predicate."""
    out_l, out_ijk = [], []
    for L, (i, j, k) in zip(level, ijk):
        if predicate(int(L), int(i), int(j), int(k)):
            for c in range(8):
                out_l.append(L + 1)
                out_ijk.append((2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + (c >> 2)))
        else:
            out_l.append(L)
            out_ijk.append((i, j, k))
    return np.array(out_l, np.int32), np.array(out_ijk, np.int64)


def _box_bounds(L, i, j, k, xmin=-1.0, size=2.0):
    h = size / (1 << L)
    lo = np.array([xmin + i * h, xmin + j * h, xmin + k * h])
    return lo, lo + h


def central_octant_mesh(lmax=LMAX_DEFAULT):
    """Level-2 4x4x4 blocks with the central octant [-0.5,0.5]^3 refined to level 3."""
    m = uniform_mesh(2, lmax)

    def pred(L, i, j, k):
        return L == 2 and 1 <= i <= 2 and 1 <= j <= 2 and 1 <= k <= 2

    lv, ijk = refine(m.level, m.ijk, pred)
    return BlockMesh(lv, ijk, lmax=lmax)


def _ball_hits(lo, hi, r):
    c = np.clip(0.0, lo, hi)
    return float(np.sum(c * c)) < r * r


def three_level_mesh(base=4, r1=0.3, slab1=0.0625, r2=0.2, lmax=LMAX_DEFAULT):
    """16^3 level-4 root blocks refined twice around the r<0.3 shell and the
    z=0 current-sheet slab (levels 4, 5, 6)."""
    m = uniform_mesh(base, lmax)

    def p1(L, i, j, k):
        lo, hi = _box_bounds(L, i, j, k)
        slab = lo[2] < slab1 and hi[2] > -slab1 and abs(lo[0] + hi[0]) < 1.0 and abs(lo[1] + hi[1]) < 1.0
        return L == base and (_ball_hits(lo, hi, r1) or slab)

    lv, ijk = refine(m.level, m.ijk, p1)

    def p2(L, i, j, k):
        lo, hi = _box_bounds(L, i, j, k)
        return L == base + 1 and _ball_hits(lo, hi, r2)

    lv, ijk = refine(lv, ijk, p2)
    return BlockMesh(lv, ijk, lmax=lmax)


# ---------------------------------------------------------------------------
# fields (normalized units, c = 1)
# ---------------------------------------------------------------------------
def dipole_field(x, moment=1.0, soft=0.1, bz0=0.05):
    """Centered dipole along +z, softened to |r| = soft inside r < soft, plus uniform Bz."""
    x = np.asarray(x, np.float64)
    r = np.sqrt(np.sum(x * x, axis=-1))
    rs = np.maximum(r, soft)
    with np.errstate(invalid="ignore", divide="ignore"):
        rhat = np.where(r[..., None] > 0, x / r[..., None], 0.0)
    mz = moment
    mdr = rhat[..., 2] * mz
    B = (3.0 * mdr[..., None] * rhat - np.array([0.0, 0.0, mz])) / rs[..., None] ** 3
    B[..., 2] += bz0
    return B


def corner_fields(mesh, E0=(0.01, 0.0, 0.0), moment=1.0, bz0=0.05):
    xc = mesh.corner_coords()
    B = dipole_field(xc, moment=moment, bz0=bz0)
    E = np.broadcast_to(np.array(E0, np.float64), B.shape).copy()
    return np.ascontiguousarray(E), np.ascontiguousarray(B)


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------
class Particles:
    """SoA particle state: x, v as [3, n] planes, species, block, cell, id."""

    def __init__(self, x, v, spec, blk, cell, pid=None):
        self.x = np.ascontiguousarray(x, np.float64)
        self.v = np.ascontiguousarray(v, np.float64)
        self.spec = np.ascontiguousarray(spec, np.int32)
        self.blk = np.ascontiguousarray(blk, np.int32)
        self.cell = np.ascontiguousarray(cell, np.int32)
        n = self.x.shape[1]
        self.id = np.arange(n, dtype=np.int64) if pid is None else np.ascontiguousarray(pid, np.int64)

    @property
    def n(self):
        return self.x.shape[1]

    def copy(self):
        return Particles(self.x.copy(), self.v.copy(), self.spec.copy(), self.blk.copy(),
                         self.cell.copy(), self.id.copy())

    def take(self, idx):
        return Particles(self.x[:, idx], self.v[:, idx], self.spec[idx], self.blk[idx],
                         self.cell[idx], self.id[idx])


def center_fields(mesh, fn):
    """Cell-centred values with one ghost layer per block, [nb][10][10][10][3]
    (cell (i,j,k) in [-1,8]^3 at (i+1) + 10((j+1) + 10(k+1))), from fn(points
    [m,3]) -> [m,3] evaluated at every cell and ghost-cell centre."""
    g = np.arange(-1, NC + 1) + 0.5
    dx = mesh.cell_size()  # [nb, 3]
    X = mesh.blk_xmin[:, None, None, None, 0] + g[None, None, None, :] * dx[:, None, None, None, 0]
    Y = mesh.blk_xmin[:, None, None, None, 1] + g[None, None, :, None] * dx[:, None, None, None, 1]
    Z = mesh.blk_xmin[:, None, None, None, 2] + g[None, :, None, None] * dx[:, None, None, None, 2]
    X, Y, Z = np.broadcast_arrays(X, Y, Z)
    P = np.stack([X, Y, Z], axis=-1).reshape(-1, 3)
    return np.ascontiguousarray(fn(P).reshape(mesh.nblocks, (NC + 2) ** 3, 3))


def interior_center_fields(mesh, fn):
    """Values at interior cell centres, [nb][512][3] with cell i + 8(j + 8k):
    the compact per-leaf layout of the Mode3D global field arrays
    (GlobalMagneticField.h:18-30, block id = Temp_ID)."""
    g = center_fields(mesh, fn).reshape(mesh.nblocks, NC + 2, NC + 2, NC + 2, 3)
    return np.ascontiguousarray(g[:, 1:NC + 1, 1:NC + 1, 1:NC + 1].reshape(mesh.nblocks, NC ** 3, 3))


def cell_of(mesh, x, blk):
    """FindCellIndex for points known to lie in block blk."""
    dxc = (mesh.xmax - mesh.xmin)[None, :] / (1 << mesh.level[blk]).astype(np.float64)[:, None] / float(NC)
    ijk = ((x.T - mesh.blk_xmin[blk]) / dxc).astype(np.int64)
    ijk = np.where(ijk == NC, NC - 1, ijk)
    return (ijk[:, 0] + NC * (ijk[:, 1] + NC * ijk[:, 2])).astype(np.int32)


def uniform_particles(mesh, n, vth, seed_x=1, seed_v=2, r_excl=0.2, mass_spec=None):
    """n points uniform over the domain excluding r < r_excl, Maxwellian v."""
    rx = np.random.default_rng(seed_x)
    rv = np.random.default_rng(seed_v)
    pts = []
    need = n
    while need > 0:
        cand = rx.uniform(mesh.xmin, mesh.xmax, size=(int(need * 1.2) + 16, 3))
        cand = cand[np.sum(cand * cand, axis=1) >= r_excl * r_excl]
        pts.append(cand[:need])
        need -= len(pts[-1])
    x = np.concatenate(pts)[:n].T.copy()
    v = rv.normal(0.0, vth, size=(3, n))
    blk = mesh.locate(x.T).astype(np.int32)
    cell = cell_of(mesh, x, blk)
    spec = np.zeros(n, np.int32)
    return Particles(x, v, spec, blk, cell)


def per_cell_particles(mesh, ppc, vth, drift=(0.0, 0.0, 0.0), seeds=(3, 4), blocks=None):
    """ppc[s] particles of species s uniformly inside every cell (of the
    given blocks, default all), Maxwellian velocities with thermal speed
    vth[s] and a common drift."""
    nsp = len(ppc)
    bl = np.arange(mesh.nblocks, dtype=np.int32) if blocks is None else np.asarray(blocks, np.int32)
    ncell = len(bl) * NC ** 3
    dx = mesh.cell_size()
    blk_all = np.repeat(bl, NC ** 3)
    lc = np.tile(np.arange(NC ** 3, dtype=np.int32), len(bl))
    ci = np.stack([lc % NC, (lc // NC) % NC, lc // (NC * NC)], axis=1)
    lo_all = mesh.blk_xmin[blk_all] + ci * dx[blk_all]
    xs, vs, sps, bs, cs = [], [], [], [], []
    for s in range(nsp):
        rng = np.random.default_rng(seeds[s % len(seeds)] + 17 * s)
        k = ppc[s]
        u = rng.random((ncell, k, 3))
        x = lo_all[:, None, :] + u * dx[blk_all][:, None, :]
        v = rng.normal(0.0, vth[s], size=(ncell, k, 3)) + np.asarray(drift)[None, None, :]
        xs.append(x.reshape(-1, 3))
        vs.append(v.reshape(-1, 3))
        sps.append(np.full(ncell * k, s, np.int32))
        bs.append(np.repeat(blk_all, k))
        cs.append(np.repeat(lc, k))
    x = np.concatenate(xs).T.copy()
    v = np.concatenate(vs).T.copy()
    blk = np.concatenate(bs)
    # positions drawn inside a cell may round onto its upper face: re-locate
    cell = cell_of(mesh, x, blk)
    return Particles(x, v, np.concatenate(sps), blk, cell)
