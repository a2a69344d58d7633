"""ctypes front end of the CPU oracle (liboracle.so) and of the reference's
own compiled gridless sources (_ref/libamps_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF = os.path.join(HERE, "_ref", "libamps_ref.so")

CODE_REFLECTED, CODE_LEFT_DOMAIN, CODE_FINISHED, CODE_SPHERE = -1, 2, 3, 5

# TrajectoryTermination (srcEarth/util/TrajectoryTermination.h:10-21)
TERM_NAMES = ["OUTER_BOUNDARY_ALLOWED", "INNER_BOUNDARY_FORBIDDEN",
              "MAGNETICALLY_TRAPPED_FORBIDDEN", "TIME_LIMIT", "STEP_LIMIT", "DISTANCE_LIMIT",
              "INVALID_TIME_STEP", "INVALID_FIELD", "NUMERICAL_FAILURE"]


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
        ("policy", C.c_int32), ("pad_", C.c_int32), ("event_tol_m", C.c_double),
        ("trap", C.c_int32), ("trap_min_mirror", C.c_int32), ("trap_min_bounce", C.c_int32), ("pad3_", C.c_int32),
        ("trap_outer_margin_Re", C.c_double), ("trap_radial_tol_Re", C.c_double), ("trap_energy_tol", C.c_double),
        ("trap_deadband", C.c_double),
    ]


RESULT_DTYPE = np.dtype([("termination", np.int32), ("retry_count", np.int32),
                         ("steps", np.int64), ("time_s", np.float64), ("distance_m", np.float64),
                         ("x", np.float64, 3), ("p", np.float64, 3)])

_lib = None
_ref = None


def build():
    """Compile the oracle (and the reference pieces when /root/reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _p(a, t):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("oracle arrays must be C-contiguous (a strided view would be read as if packed)")
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        l = C.CDLL(LIB)
        P, vp, i32, i64, dbl = C.POINTER, C.c_void_p, C.c_int, C.c_int64, C.c_double
        l.orc_mesh_create.restype = vp
        l.orc_mesh_create.argtypes = [P(dbl), P(dbl), i32, i32, P(C.c_int), P(i64)]
        l.orc_mesh_destroy.argtypes = [vp]
        l.orc_mesh_eps.restype = dbl
        l.orc_mesh_eps.argtypes = [vp]
        l.orc_find_block.argtypes = [vp, P(dbl), i32]
        l.orc_find_cell.argtypes = [vp, P(dbl), i32]
        l.orc_corner_stencil.argtypes = [vp, P(dbl), i32, P(dbl), P(dbl), P(C.c_int)]
        l.orc_push.argtypes = [vp, P(dbl), P(dbl), P(dbl), P(dbl), P(dbl), i32, i32, dbl, i64,
                               P(dbl), P(dbl), P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), i32]
        l.orc_build_nodes.restype = i64
        l.orc_build_nodes.argtypes = [vp, P(C.c_int), P(i64)]
        l.orc_deposit.argtypes = [vp, P(dbl), P(dbl), P(dbl), P(dbl), dbl, i64, P(dbl), P(dbl),
                                  P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), i64, P(dbl),
                                  P(dbl), i32]
        l.orc_deposit_colored.argtypes = l.orc_deposit.argtypes
        l.orc_energy_cfl.argtypes = [vp, P(dbl), P(dbl), i32, dbl, i64, P(dbl), P(C.c_int), P(C.c_int),
                                     P(C.c_int), P(dbl), P(dbl), P(dbl), P(dbl)]
        l.orc_dipole_B.argtypes = [P(TraceParams), P(dbl), P(dbl)]
        l.orc_select_dt.argtypes = [P(TraceParams), P(dbl), P(dbl), dbl, dbl]
        l.orc_select_dt.restype = dbl
        l.orc_select_dt3d.argtypes = [P(TraceParams), P(dbl), P(dbl), dbl, dbl, P(dbl), P(dbl), dbl]
        l.orc_select_dt3d.restype = dbl
        l.orc_boris_step.argtypes = [P(TraceParams), P(dbl), P(dbl), dbl]
        l.orc_hc4_step.argtypes = [P(TraceParams), P(dbl), P(dbl), dbl]
        l.orc_step_checked.argtypes = [P(TraceParams), P(dbl), P(dbl), dbl, dbl]
        l.orc_step_checked.restype = i32
        l.orc_trace_batch.argtypes = [P(TraceParams), i64, P(dbl), P(dbl), P(dbl), vp, i32]
        l.orc_trace_batch_ex.argtypes = [P(TraceParams), i64, P(dbl), P(dbl), P(dbl), vp, vp, i32]
        l.orc_find_first_event.argtypes = [P(dbl), P(dbl), P(dbl), P(dbl), dbl, dbl, P(i32), P(dbl), P(dbl), P(i32)]
        l.orc_trap_init.argtypes = [vp]
        l.orc_trap_update.argtypes = [P(TraceParams), vp, P(dbl), P(dbl), P(dbl)]
        l.orc_trap_update.restype = i32
        l.orc_finalize_step.argtypes = [dbl, dbl]
        l.orc_finalize_step.restype = dbl
        l.orc_rigidity_grid.argtypes = [dbl, dbl, i32, P(dbl)]
        l.orc_upper_cutoff_bin.argtypes = [P(C.c_int), i32]
        l.orc_band_topology.argtypes = [P(C.c_int), i32, P(C.c_int)]
        l.orc_center_stencil.argtypes = [vp, P(dbl), i32, P(C.c_int), P(C.c_int), P(dbl)]
        l.orc_interp_center.argtypes = [vp, P(dbl), i64, P(dbl), P(C.c_int), P(dbl), i32]
        l.orc_block_neighbours.argtypes = [vp, i32, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int)]
        l.orc_set_center_field.argtypes = [vp, P(dbl)]
        l.orc_row_stencil.argtypes = [vp, P(dbl), P(C.c_int), P(C.c_int), P(dbl)]
        l.orc_row_stencil.restype = i32
        l.orc_mode3d_field.argtypes = [vp, P(dbl), i64, P(dbl), P(dbl)]
        l.orc_mode3d_field.restype = i32
        l.orc_set_trace_mesh.argtypes = [vp, P(dbl)]
        _lib = l
    return _lib


def ref_available():
    return os.path.exists(REF)


REF_CLASSIFY = C.CFUNCTYPE(C.c_int, C.c_double, C.c_void_p)


def ref_refine_cutoff(r_forbid, r_allow, classify, abs_tol=1.0e-3, rel_tol=1.0e-6):
    """The reference's RefineCutoffTransition (CutoffBandSearch.h:212-242)
    with classify(R_GV) -> 0 physical forbidden, 1 allowed, 2 unresolved.
    Returns (cutoff_GV, iterations, unresolved)."""
    cb = REF_CLASSIFY(lambda R, _u: int(classify(R)))
    c, it, un = C.c_double(), C.c_int(), C.c_int()
    ref().ref_refine_cutoff(r_forbid, r_allow, abs_tol, rel_tol, cb, None, C.byref(c), C.byref(it),
                            C.byref(un))
    return c.value, it.value, bool(un.value)


def ref_effective_cutoff(grid, states, lower, upper, classify, abs_tol=1.0e-3, rel_tol=1.0e-6):
    """The reference's IntegrateEffectiveCutoff (CutoffBandSearch.h:284-366)
    on grid/states (0 forbidden, 1 allowed, 2 unresolved) with the same
    classifier convention.  Returns (cutoff_GV, allowed_width_GV, refined,
    unresolved)."""
    cb = REF_CLASSIFY(lambda R, _u: int(classify(R)))
    g = np.ascontiguousarray(grid, np.float64)
    s = np.ascontiguousarray(states, np.int32)
    c, w, it, un = C.c_double(), C.c_double(), C.c_int(), C.c_int()
    ref().ref_effective_cutoff(len(g), g.ctypes.data_as(C.POINTER(C.c_double)),
                               s.ctypes.data_as(C.POINTER(C.c_int)), lower, upper, abs_tol, rel_tol, cb, None,
                               C.byref(c), C.byref(w), C.byref(it), C.byref(un))
    return c.value, w.value, it.value, bool(un.value)


def restart_order(blk, cell, nblocks):
    """The order in which SaveParticleDataBlock (pic_restart.cpp:312-370)
    writes particles given in list order (per cell, head first): blocks in
    id order, cells in i, j, k loop order (k fastest, cell i + 8(j + 8k)),
    each cell's list from its last particle to its first.  Returns the
    per-block ParticleNumberTable [nblocks, 512] in [i][j][k] order and the
    permutation of particle indices in file order."""
    blk = np.asarray(blk)
    cell = np.asarray(cell)
    table = np.zeros((nblocks, 8, 8, 8), np.int32)
    members = {}
    for p in range(len(blk)):
        members.setdefault((int(blk[p]), int(cell[p])), []).append(p)
    order = []
    for b in range(nblocks):
        for i in range(8):
            for j in range(8):
                for k in range(8):
                    lst = members.get((b, i + 8 * (j + 8 * k)), [])
                    table[b, i, j, k] = len(lst)
                    order.extend(reversed(lst))
    return table.reshape(nblocks, 512), np.array(order, np.int64)


def ref():
    """The reference's compiled gridless mover/dipole/band sources."""
    global _ref
    if _ref is None:
        r = C.CDLL(REF)
        P, dbl = C.POINTER, C.c_double
        r.ref_set_dipole.argtypes = [dbl, dbl]
        r.ref_get_dipole.argtypes = [P(dbl), P(dbl)]
        r.ref_field.argtypes = [P(dbl), P(dbl), P(dbl)]
        r.ref_boris_steps.argtypes = [C.c_int, P(dbl), P(dbl), dbl, dbl, dbl, P(dbl), P(dbl)]
        r.ref_step_checked.argtypes = [P(dbl), P(dbl), dbl, dbl, dbl, P(dbl), dbl]
        r.ref_step_checked.restype = C.c_int
        r.ref_mover_steps.argtypes = [C.c_int, C.c_int, P(dbl), P(dbl), dbl, dbl, dbl, P(dbl), P(dbl)]
        r.ref_step_checked_mover.argtypes = [C.c_int, P(dbl), P(dbl), dbl, dbl, dbl, P(dbl), dbl]
        r.ref_step_checked_mover.restype = C.c_int
        r.ref_find_first_event.argtypes = [P(dbl), P(dbl), P(dbl), P(dbl), dbl, dbl, P(C.c_int), P(dbl), P(dbl),
                                           P(C.c_int)]
        r.ref_trap_sequence.argtypes = [C.c_int, P(dbl), P(dbl), P(dbl), C.c_int, C.c_int, dbl, dbl, dbl, dbl,
                                         P(dbl), P(dbl), P(C.c_int), P(C.c_int), P(C.c_int)]
        r.ref_finalize_step.argtypes = [dbl, dbl]
        r.ref_finalize_step.restype = dbl
        r.ref_band_topology.argtypes = [C.c_int, P(C.c_int), P(C.c_int)]
        r.ref_refine_cutoff.argtypes = [dbl, dbl, dbl, dbl, REF_CLASSIFY, C.c_void_p, P(dbl),
                                        P(C.c_int), P(C.c_int)]
        r.ref_effective_cutoff.argtypes = [C.c_int, P(dbl), P(C.c_int), dbl, dbl, dbl, dbl, REF_CLASSIFY,
                                           C.c_void_p, P(dbl), P(dbl), P(C.c_int), P(C.c_int)]
        _ref = r
    return _ref


class Mesh:
    """Oracle octree built from the same leaf list as the GPU library."""

    def __init__(self, m):
        l = lib()
        lv, ijk = m.desc_arrays()
        self._lv, self._ijk = lv.astype(np.int32), ijk.astype(np.int64)
        xmin = np.ascontiguousarray(m.xmin, np.float64)
        xmax = np.ascontiguousarray(m.xmax, np.float64)
        self.h = l.orc_mesh_create(_p(xmin, C.c_double), _p(xmax, C.c_double), m.lmax,
                                   m.nblocks, _p(self._lv, C.c_int), _p(self._ijk, C.c_int64))
        if not self.h:
            raise ValueError("oracle rejected the mesh")
        self.nblocks = m.nblocks

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_destroy(self.h)

    @property
    def eps(self):
        return lib().orc_mesh_eps(self.h)

    def find_block(self, x, start=-1):
        x = np.ascontiguousarray(x, np.float64)
        return lib().orc_find_block(self.h, _p(x, C.c_double), start)

    def find_cell(self, x, b):
        x = np.ascontiguousarray(x, np.float64)
        return lib().orc_find_cell(self.h, _p(x, C.c_double), b)

    def stencil(self, x, b):
        x = np.array(x, np.float64)
        w = np.zeros(8)
        t = np.zeros(8)
        c = np.zeros(8, np.int32)
        lib().orc_corner_stencil(self.h, _p(x, C.c_double), b, _p(w, C.c_double),
                                 _p(t, C.c_double), _p(c, C.c_int))
        return x, w, t, c

    def center_stencil(self, x, b):
        """Cell-centred linear stencil: [(block, ghost-layer cell index, weight)]."""
        x = np.array(x, np.float64)
        bb = np.zeros(80, np.int32)
        cc = np.zeros(80, np.int32)
        ww = np.zeros(80)
        n = lib().orc_center_stencil(self.h, _p(x, C.c_double), int(b), _p(bb, C.c_int), _p(cc, C.c_int),
                                     _p(ww, C.c_double))
        return bb[:n], cc[:n], ww[:n]

    def interp_center(self, Bc, x, blk, nthreads=0):
        """B at points x[3][n] (blocks blk) from ghost-layered cell data Bc[nb][1000][3]."""
        Bc = np.ascontiguousarray(Bc, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        blk = np.ascontiguousarray(blk, np.int32)
        n = x.shape[1]
        B = np.zeros((3, n))
        lib().orc_interp_center(self.h, _p(Bc, C.c_double), n, _p(x, C.c_double), _p(blk, C.c_int),
                                _p(B, C.c_double), nthreads)
        return B

    def set_center_field(self, Bc):
        """Centre-based B (ghost-layered [nb][1000][3]) for push and deposit;
        None restores the corner B."""
        self._Bc = None if Bc is None else np.ascontiguousarray(Bc, np.float64)
        lib().orc_set_center_field(self.h, None if self._Bc is None else _p(self._Bc, C.c_double))

    def row_stencil(self, x):
        """Mode3D row stencil at x: (blocks, interior cells i + 8(j + 8k), weights)."""
        x = np.ascontiguousarray(x, np.float64)
        bb = np.zeros(128, np.int32)
        cc = np.zeros(128, np.int32)
        ww = np.zeros(128)
        n = lib().orc_row_stencil(self.h, _p(x, C.c_double), _p(bb, C.c_int), _p(cc, C.c_int),
                                  _p(ww, C.c_double))
        return bb[:n], cc[:n], ww[:n]

    def mode3d_field(self, Bm, x):
        """Mode3D mesh field at points x[n][3] from interior-cell data Bm[nb][512][3]."""
        Bm = np.ascontiguousarray(Bm, np.float64)
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        B = np.zeros_like(x)
        ok = lib().orc_mode3d_field(self.h, _p(Bm, C.c_double), len(x), _p(x, C.c_double), _p(B, C.c_double))
        if not ok:
            raise ValueError("empty row stencil")
        return B

    def set_trace_field(self, Bm):
        """Mesh field of the tracer's field_model 1 (kept referenced here)."""
        self._Bm = np.ascontiguousarray(Bm, np.float64)
        lib().orc_set_trace_mesh(self.h, _p(self._Bm, C.c_double))

    def neighbours(self, b):
        mm = np.zeros(2, np.int32)
        f = np.zeros(6, np.int32)
        e = np.zeros(12, np.int32)
        c = np.zeros(8, np.int32)
        lib().orc_block_neighbours(self.h, int(b), _p(mm, C.c_int), _p(f, C.c_int), _p(e, C.c_int),
                                   _p(c, C.c_int))
        return mm, f, e, c

    def build_nodes(self):
        noc = np.zeros(self.nblocks * 729, np.int32)
        keys = np.zeros(self.nblocks * 729, np.int64)
        nn = lib().orc_build_nodes(self.h, _p(noc, C.c_int), _p(keys, C.c_int64))
        return noc, keys[:nn]


def push(omesh, E, B, species, parts, mover=0, bc_mode=1, r_sphere=0.0, nthreads=0, copy=True, code=None):
    """One oracle push of a Particles object; returns (new Particles, code).
    copy=False pushes parts itself (and fills code if given)."""
    p = parts.copy() if copy else parts
    code = np.zeros(p.n, np.int32) if code is None else code
    E = np.ascontiguousarray(E, np.float64)
    B = np.ascontiguousarray(B, np.float64)
    lib().orc_push(omesh.h, _p(E, C.c_double), _p(B, C.c_double),
                   _p(species.charge, C.c_double), _p(species.mass, C.c_double),
                   _p(species.dt, C.c_double), mover, bc_mode, r_sphere, p.n,
                   _p(p.x, C.c_double), _p(p.v, C.c_double), _p(p.spec, C.c_int),
                   _p(p.blk, C.c_int), _p(p.cell, C.c_int), _p(code, C.c_int), nthreads)
    return p, code


def energy_cfl(omesh, species, dt, parts, nthreads=0):
    """Total particle energy, per-species max CFL and the per-cell values of
    UpdateJMassMatrix (pic_field_solver_ecsim.cpp:2229-2238, 2355-2358,
    3925-3929, 3972-3987), cells summed in the given particle order."""
    ns = species.n
    e = C.c_double()
    cfl = np.zeros(ns)
    nc = omesh.nblocks * 512
    ce = np.zeros(nc)
    cc = np.zeros((nc, ns))
    lib().orc_energy_cfl(omesh.h, _p(species.mass, C.c_double), _p(species.weight, C.c_double), ns, float(dt),
                         parts.n, _p(parts.v, C.c_double), _p(parts.spec, C.c_int), _p(parts.blk, C.c_int),
                         _p(parts.cell, C.c_int), C.byref(e), _p(cfl, C.c_double), _p(ce, C.c_double),
                         _p(cc, C.c_double))
    return e.value, cfl, ce, cc


def ecsim_neighbours(mesh, node_of_corner, keys):
    """[nnodes][27] neighbour nodes of the ECSIM corner stencil, -1 outside:
    slot ii = a(di) + 3a(dj) + 9a(dk), a(0,-1,+1) = (0,1,2), one cell of the
    node's finest incident cell size away (GetStencil,
    pic_field_solver_ecsim.cpp:1061-1074)."""
    S = (1 << mesh.lmax) + 1
    nn = len(keys)
    h = np.full(nn, np.iinfo(np.int64).max, np.int64)
    cs = (np.int64(1) << (mesh.lmax - mesh.level.astype(np.int64))) // 8
    np.minimum.at(h, np.asarray(node_of_corner).reshape(-1, 729), cs[:, None])
    g = np.stack([keys % S, (keys // S) % S, keys // (S * S)], 1)
    a = np.array([0, -1, 1])
    nbr = -np.ones((nn, 27), np.int64)
    for s in range(27):
        off = np.array([a[s % 3], a[(s // 3) % 3], a[s // 9]])
        o = g + off[None, :] * h[:, None]
        inside = ((o >= 0) & (o < S)).all(1)
        kk = (o[:, 2] * S + o[:, 1]) * S + o[:, 0]
        pos = np.minimum(np.searchsorted(keys, kk), nn - 1)
        found = inside & (keys[pos] == kk)
        nbr[found, s] = pos[found]
    return nbr


def ecsim_mass_apply(nbr, Mm, E, c1, c2, seed=None):
    """The mass-matrix part of UpdateRhs_Legacy (pic_field_solver_ecsim.cpp:
    6116-6124): res_c += (Ex M[9ii+3c] + Ey M[9ii+3c+1] + Ez M[9ii+3c+2])
    * c1 * c2 over the 27 neighbours in slot order (MassMatrixOffsetTable,
    :621-639), absent neighbours skipped; res starts from seed (the row's
    earlier terms) when given."""
    nn = len(Mm)
    out = np.zeros((nn, 3)) if seed is None else np.array(seed, np.float64, copy=True)
    for s in range(27):
        j = nbr[:, s]
        ok = j >= 0
        e = E[np.where(ok, j, 0)]
        for c in range(3):
            m = Mm[:, 9 * s + 3 * c:9 * s + 3 * c + 3]
            term = ((e[:, 0] * m[:, 0] + e[:, 1] * m[:, 1]) + e[:, 2] * m[:, 2]) * c1 * c2
            out[:, c] = np.where(ok, out[:, c] + term, out[:, c])
    return out


def deposit(omesh, B, species, dt, parts, node_of_corner, nnodes, nthreads=0):
    J = np.zeros((nnodes, 3))
    M = np.zeros((nnodes, 243))
    B = np.ascontiguousarray(B, np.float64)
    noc = np.ascontiguousarray(node_of_corner, np.int32)
    lib().orc_deposit(omesh.h, _p(B, C.c_double), _p(species.charge, C.c_double),
                      _p(species.mass, C.c_double), _p(species.weight, C.c_double), dt, parts.n,
                      _p(parts.x, C.c_double), _p(parts.v, C.c_double), _p(parts.spec, C.c_int),
                      _p(parts.blk, C.c_int), _p(parts.cell, C.c_int), _p(noc, C.c_int), nnodes,
                      _p(J, C.c_double), _p(M, C.c_double), nthreads)
    return J, M


def deposit_colored(omesh, B, species, dt, parts, node_of_corner, nnodes, nthreads=0, out=None):
    """The deposit in the reference's parallel arrangement (parity sets of
    cells, UpdateJMassMatrix): equal to deposit() up to rounding.  Rows of
    nodes with no occupied incident cell are left unset (np.empty, or the
    given out=(J, M) arrays), so a sample population on a large mesh touches
    only its own nodes."""
    J, M = (np.empty((nnodes, 3)), np.empty((nnodes, 243))) if out is None else out
    assert J.shape == (nnodes, 3) and M.shape == (nnodes, 243)
    B = np.ascontiguousarray(B, np.float64)
    noc = np.ascontiguousarray(node_of_corner, np.int32)
    lib().orc_deposit_colored(omesh.h, _p(B, C.c_double), _p(species.charge, C.c_double),
                              _p(species.mass, C.c_double), _p(species.weight, C.c_double), dt, parts.n,
                              _p(parts.x, C.c_double), _p(parts.v, C.c_double), _p(parts.spec, C.c_int),
                              _p(parts.blk, C.c_int), _p(parts.cell, C.c_int), _p(noc, C.c_int), nnodes,
                              _p(J, C.c_double), _p(M, C.c_double), nthreads)
    return J, M


EXIT_DTYPE = np.dtype([("x_exit_m", "<f8", 3), ("v_exit_unit", "<f8", 3), ("cos_alpha", "<f8"),
                       ("valid", "<i4"), ("pad_", "<i4")])


def trace_batch_ex(params, x0, dirs, R, nthreads=0):
    """Traces with the exit states of allowed structured traces."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(R, np.float64)
    out = np.zeros(len(R), RESULT_DTYPE)
    ex = np.zeros(len(R), EXIT_DTYPE)
    lib().orc_trace_batch_ex(C.byref(params), len(R), _p(x0, C.c_double), _p(dirs, C.c_double),
                             _p(R, C.c_double), out.ctypes.data_as(C.c_void_p), ex.ctypes.data_as(C.c_void_p),
                             nthreads)
    return out, ex


def find_first_event(a, z, bmin, bmax, r_inner, tol):
    """The first boundary event on the chord a -> z: (type, fraction, position, face)."""
    a, z, bmin, bmax = (np.ascontiguousarray(v, np.float64) for v in (a, z, bmin, bmax))
    t, fr, face, pos = C.c_int(), C.c_double(), C.c_int(), np.zeros(3)
    lib().orc_find_first_event(_p(a, C.c_double), _p(z, C.c_double), _p(bmin, C.c_double), _p(bmax, C.c_double),
                               r_inner, tol, C.byref(t), C.byref(fr), _p(pos, C.c_double), C.byref(face))
    return t.value, fr.value, pos, face.value


TRAP_DTYPE = np.dtype([("initialized", "<i4"), ("trapped", "<i4"), ("last_sign", "<i4"), ("mirror", "<i4"),
                       ("bounce", "<i4"), ("stable", "<i4"), ("have_prev", "<i4"), ("pad_", "<i4"),
                       ("p_ref", "<f8"), ("p_min", "<f8"), ("p_max", "<f8"), ("r_min", "<f8"), ("r_max", "<f8"),
                       ("prev_r_min", "<f8"), ("prev_r_max", "<f8"), ("min_margin", "<f8")])


def trap_sequence(params, x, p, B):
    """The oracle's trap detector over consecutive (x, p, B) samples:
    (trapped, mirror points, bounce cycles) after each sample."""
    st = np.zeros(1, TRAP_DTYPE)
    lib().orc_trap_init(st.ctypes.data_as(C.c_void_p))
    out = np.zeros((len(x), 3), np.int32)
    for i in range(len(x)):
        xi, pi, bi = (np.ascontiguousarray(v[i], np.float64) for v in (x, p, B))
        out[i, 0] = lib().orc_trap_update(C.byref(params), st.ctypes.data_as(C.c_void_p), _p(xi, C.c_double),
                                          _p(pi, C.c_double), _p(bi, C.c_double))
        out[i, 1], out[i, 2] = st["mirror"][0], st["bounce"][0]
    return out


def select_dt(params, x, p, dt_trace, time_remaining):
    """SelectLegacyCutoffTraceDt at one state (x, p in SI)."""
    x, p = (np.ascontiguousarray(a, np.float64) for a in (x, p))
    return lib().orc_select_dt(C.byref(params), _p(x, C.c_double), _p(p, C.c_double), dt_trace, time_remaining)


def select_dt3d(params, x, p, dt_trace, time_remaining, boxmin, boxmax, r_inner):
    """SelectLegacyCutoffTraceDt3D at one state (box and inner radius in metres)."""
    x, p, bmin, bmax = (np.ascontiguousarray(a, np.float64) for a in (x, p, boxmin, boxmax))
    return lib().orc_select_dt3d(C.byref(params), _p(x, C.c_double), _p(p, C.c_double), dt_trace, time_remaining,
                                 _p(bmin, C.c_double), _p(bmax, C.c_double), r_inner)


def finalize_step(dt, time_remaining):
    return lib().orc_finalize_step(dt, time_remaining)


def trace_batch(params, x0, dirs, R, nthreads=0):
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(R, np.float64)
    out = np.zeros(len(R), RESULT_DTYPE)
    lib().orc_trace_batch(C.byref(params), len(R), _p(x0, C.c_double), _p(dirs, C.c_double),
                          _p(R, C.c_double), out.ctypes.data_as(C.c_void_p), nthreads)
    return out


def rigidity_grid(rmin, rmax, n):
    g = np.zeros(n)
    lib().orc_rigidity_grid(rmin, rmax, n, _p(g, C.c_double))
    return g


def upper_cutoff_bin(allowed):
    a = np.ascontiguousarray(allowed, np.int32)
    return lib().orc_upper_cutoff_bin(_p(a, C.c_int), len(a))


def band_topology(states):
    s = np.ascontiguousarray(states, np.int32)
    out = np.zeros(15, np.int32)
    lib().orc_band_topology(_p(s, C.c_int), len(s), _p(out, C.c_int))
    return out


# ----------------------------------------------------------------------------
# ECSIM implicit field equation on a single-level mesh (GetStencil_Legacy,
# UpdateMatrixElement, UpdateRhs_Legacy, ProcessFinalSolution, UpdateE,
# UpdateB of pic_field_solver_ecsim.cpp).  The unknown is dE = E^{n+theta} -
# E^n (times E_conv) at the corner nodes; domain-boundary corners are rows of
# the identity with a zero right-hand side (:782-799) and are dropped from
# the other rows (:1032).
# ----------------------------------------------------------------------------
def _st_add(st, other, di=0, dj=0, dk=0, f=1.0):
    for (i, j, k), a in other.items():
        key = (i + di, j + dj, k + dk)
        st[key] = st.get(key, 0.0) + f * a


def ecsim_stencils():
    """The Laplacian and grad-div stencils of the corner field solver as the
    reference composes them (pic_field_solver_ecsim.cpp:7185-7300): xEdge
    averages of 4 points, d2E/dx2 = the x second difference of xEdge[0]
    averaged over 2x2 (y, z) shifts, the mixed derivatives from differences
    of xEdge[4] and xEdge[8].  Returns (lap[3], gd[3][3]) as dicts
    {(di, dj, dk): coefficient} with offsets in {-1, 0, 1}."""
    def edge(ir, jr, kr):
        return {(i, j, k): 0.25 for i in ir for j in jr for k in kr}

    xe0, xe4, xe8 = edge((0,), (-1, 0), (-1, 0)), edge((-1, 0), (0,), (-1, 0)), edge((-1, 0), (-1, 0), (0,))
    f0 = {}
    _st_add(f0, xe0, 1, 0, 0)
    _st_add(f0, xe0, 0, 0, 0, -2.0)
    _st_add(f0, xe0, -1, 0, 0)
    dxx = {}
    for i in range(2):
        for j in range(2):
            _st_add(dxx, f0, 0, i, j, 0.25)
    dyy = {(j, i, k): a for (i, j, k), a in dxx.items()}
    dzz = {(k, j, i): a for (i, j, k), a in dxx.items()}
    dy0 = dict(xe4)
    _st_add(dy0, xe4, 0, -1, 0, -1.0)
    dz0 = dict(xe8)
    _st_add(dz0, xe8, 0, 0, -1, -1.0)
    dxy, dxz, dyz = {}, {}, {}
    for l in range(2):
        for m in range(2):
            _st_add(dxy, dy0, 1, l, m, 0.25)
            _st_add(dxy, dy0, 0, l, m, -0.25)
            _st_add(dxz, dz0, 1, l, m, 0.25)
            _st_add(dxz, dz0, 0, l, m, -0.25)
            _st_add(dyz, dz0, l, 1, m, 0.25)
            _st_add(dyz, dz0, l, 0, m, -0.25)
    lap = [dxx, dyy, dzz]
    gd = [[dxx, dxy, dxz], [dxy, dyy, dyz], [dxz, dyz, dzz]]
    return lap, gd


def _slot(di, dj, dk):
    a = {0: 0, -1: 1, 1: 2}
    return a[di] + 3 * a[dj] + 9 * a[dk]


def ecsim_coefficients(dx, cDt, theta, length_conv=1.0, corr=0.0):
    """The constant part of the matrix rows (identity, minus the Laplacian
    times coeff^2, plus the grad-div times coeff_iVar coeff_iVarIndex;
    GetStencil_Legacy :873-910) and the E^n coefficients of the right-hand
    side (plus Laplacian, minus grad-div; :1152-1185), as [iVar][iVarIndex]
    [slot] tables, accumulated in the reference's loop order.  dx: the cell
    size per axis."""
    coeff = [cDt / (dx[d] * length_conv) * theta for d in range(3)]
    coeffSqr = [c * c for c in coeff]
    lap, gd = ecsim_stencils()
    K = np.zeros((3, 3, 27))
    R = np.zeros((3, 3, 27))
    for iVar in range(3):
        for idim in range(3):
            for (i, j, k), a in lap[idim].items():
                K[iVar, iVar, _slot(i, j, k)] -= a * coeffSqr[idim]
                R[iVar, iVar, _slot(i, j, k)] += a * coeffSqr[idim]
        K[iVar, iVar, 0] += 1
        for iv in range(3):
            for (i, j, k), a in gd[iVar][iv].items():
                K[iVar, iv, _slot(i, j, k)] += (1 - corr) * a * coeff[iVar] * coeff[iv]
                R[iVar, iv, _slot(i, j, k)] -= (1 - corr) * a * coeff[iVar] * coeff[iv]
    return K, R, coeff


def ecsim_boundary(mesh, keys):
    """Corner nodes on the domain boundary (isBoundaryCorner, :7011-7056, for
    a mesh whose leaves cover the root box)."""
    S = (1 << mesh.lmax) + 1
    g = np.stack([keys % S, (keys // S) % S, keys // (S * S)], 1)
    return ((g == 0) | (g == S - 1)).any(1)


def ecsim_cells_around(mesh, node_of_corner, nbr):
    """[nnodes][8]: the cell whose lower corner is node + (ci, cj, ck),
    ci, cj, ck in {-1, 0}, at (ci+1) + 2(cj+1) + 4(ck+1); -1 outside.  Cell
    ids are block * 512 + i + 8 j + 64 k."""
    nn = len(nbr)
    noc = np.asarray(node_of_corner).reshape(-1, 729)
    lower = -np.ones(nn, np.int64)
    ii, jj, kk = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    corner = (ii + 9 * jj + 81 * kk).ravel()
    cell = (ii + 8 * jj + 64 * kk).ravel()
    for b in range(noc.shape[0]):
        lower[noc[b, corner]] = b * 512 + cell
    out = -np.ones((nn, 8), np.int64)
    for ck in (-1, 0):
        for cj in (-1, 0):
            for ci in (-1, 0):
                m = nbr[:, _slot(ci, cj, ck)]
                out[:, (ci + 1) + 2 * (cj + 1) + 4 * (ck + 1)] = np.where(m >= 0, lower[np.maximum(m, 0)], -1)
    return out


def ecsim_operator(nbr, bnd, Mm, K, x, fourPiDtTheta):
    """y = A x: boundary rows are the identity; row (n, iVar) of an interior
    node sums over iVarIndex (outer) and slot (inner), as the solver's row
    elements are ordered (iVarIndex 27 + slot), the value K + M
    fourPiDtTheta (UpdateMatrixElement, :6004-6016) times x of the neighbour;
    neighbours outside the domain or on its boundary are dropped."""
    nn = len(nbr)
    y = np.zeros((nn, 3))
    for iVar in range(3):
        acc = np.zeros(nn)
        for iv in range(3):
            for s in range(27):
                j = nbr[:, s]
                ok = (j >= 0) & ~bnd[np.maximum(j, 0)]
                val = K[iVar, iv, s] + Mm[:, 9 * s + 3 * iVar + iv] * fourPiDtTheta
                acc = np.where(ok, acc + val * x[np.maximum(j, 0), iv], acc)
        y[:, iVar] = np.where(bnd, x[:, iVar], acc)
    return y


# cells (ci, cj, ck) offsets and signs of the curl-B terms of UpdateRhs_Legacy
# (:1234-1375): per iVar, 16 entries (group of 4 x 4, ii outer, jj inner),
# the first 8 for B component (iVar + 2) % 3, the next 8 for (iVar + 4) % 3;
# each entry (coefficient axis, sign, ci, cj, ck)
def _curlb_table():
    a = (-1, 0)
    t = {0: [], 1: [], 2: []}
    for ii in range(2):
        for jj in range(2):
            t[0].append((1, +1, a[ii], 0, a[jj]))
    for ii in range(2):
        for jj in range(2):
            t[0].append((1, -1, a[ii], -1, a[jj]))
    for ii in range(2):
        for jj in range(2):
            t[0].append((2, -1, a[ii], a[jj], 0))
    for ii in range(2):
        for jj in range(2):
            t[0].append((2, +1, a[ii], a[jj], -1))
    for ii in range(2):
        for jj in range(2):
            t[1].append((2, +1, a[ii], a[jj], 0))
    for ii in range(2):
        for jj in range(2):
            t[1].append((2, -1, a[ii], a[jj], -1))
    for ii in range(2):
        for jj in range(2):
            t[1].append((0, -1, 0, a[jj], a[ii]))
    for ii in range(2):
        for jj in range(2):
            t[1].append((0, +1, -1, a[jj], a[ii]))
    for ii in range(2):
        for jj in range(2):
            t[2].append((0, +1, 0, a[jj], a[ii]))
    for ii in range(2):
        for jj in range(2):
            t[2].append((0, -1, -1, a[jj], a[ii]))
    for ii in range(2):
        for jj in range(2):
            t[2].append((1, -1, a[jj], 0, a[ii]))
    for ii in range(2):
        for jj in range(2):
            t[2].append((1, +1, a[jj], -1, a[ii]))
    return t


CURLB_TABLE = _curlb_table()


def ecsim_rhs(nbr, bnd, cells8, Mm, J, R, coeff, En, Bc, dtTotal, theta, E_conv=1.0, B_conv=1.0):
    """UpdateRhs_Legacy (:6040-6137) for every interior row, in its order:
    the 27 neighbours' E^n times the Laplacian / grad-div coefficients, the
    mass-matrix part times -4 pi dt theta, the current times -4 pi dt theta,
    then the 16 curl-B terms (coefficients +-coeff/4 on the 8 + 8 cell-centred
    B of :1234-1375).  Boundary rows are 0."""
    nn = len(nbr)
    fpdt = 4 * np.pi * dtTotal * theta
    c4 = [0.25 * c for c in coeff]
    out = np.zeros((nn, 3))
    for iVar in range(3):
        res = np.zeros(nn)
        for s in range(27):
            j = np.maximum(nbr[:, s], 0)
            e = En[j]
            res = res + (e[:, 0] * R[iVar, 0, s] + e[:, 1] * R[iVar, 1, s] + e[:, 2] * R[iVar, 2, s]) * E_conv
        for s in range(27):
            j = np.maximum(nbr[:, s], 0)
            e = En[j]
            m = Mm[:, 9 * s + 3 * iVar:9 * s + 3 * iVar + 3]
            res = res + (e[:, 0] * m[:, 0] + e[:, 1] * m[:, 1] + e[:, 2] * m[:, 2]) * (-fpdt) * E_conv
        res = res + J[:, iVar] * (-4 * np.pi * dtTotal * theta)
        for t, (ax, sg, ci, cj, ck) in enumerate(CURLB_TABLE[iVar]):
            comp = (iVar + 2) % 3 if t < 8 else (iVar + 4) % 3
            cl = cells8[:, (ci + 1) + 2 * (cj + 1) + 4 * (ck + 1)]
            res = res + Bc[np.maximum(cl, 0), comp] * (sg * c4[ax]) * B_conv
        out[:, iVar] = np.where(bnd, 0.0, res)
    return out


def ecsim_gmres(apply, b, tol, maxiter, restart):
    """Restarted GMRES with modified Gram-Schmidt from x0 = 0, stopping at
    ||r|| <= tol ||b|| (a CPU check of the device solver)."""
    n = b.size
    x = np.zeros(n)
    bn = np.linalg.norm(b)
    if bn == 0:
        return x, 0, 0.0
    it = 0
    while it < maxiter:
        r = b - apply(x)
        beta = np.linalg.norm(r)
        if beta <= tol * bn:
            break
        V = np.zeros((restart + 1, n))
        H = np.zeros((restart + 1, restart))
        V[0] = r / beta
        k = 0
        for k in range(restart):
            w = apply(V[k])
            for j in range(k + 1):
                H[j, k] = w @ V[j]
                w = w - H[j, k] * V[j]
            H[k + 1, k] = np.linalg.norm(w)
            it += 1
            if H[k + 1, k] > 0:
                V[k + 1] = w / H[k + 1, k]
            e1 = np.zeros(k + 2)
            e1[0] = beta
            yk, *_ = np.linalg.lstsq(H[:k + 2, :k + 1], e1, rcond=None)
            res = np.linalg.norm(e1 - H[:k + 2, :k + 1] @ yk)
            if res <= tol * bn or it >= maxiter or H[k + 1, k] == 0:
                break
        x = x + V[:k + 1].T @ yk
    return x, it, np.linalg.norm(b - apply(x)) / bn


def ecsim_update_e(En, dE, theta, E_conv=1.0):
    """ProcessFinalSolution (:6311-6318) then UpdateE (:5909-5960):
    E^{n+theta} = dE / E_conv + E^n, E^{n+1} = (E^{n+theta} - (1 - theta) E^n) / theta."""
    Eth = dE / E_conv + En
    return Eth, (Eth - (1.0 - theta) * En) / theta


def ecsim_update_b(node_of_corner, Eth, Bc, dx, cDt, E_conv=1.0, B_conv=1.0, length_conv=1.0):
    """UpdateB (:5160-5270) for every cell: B^{n+1} = B^n - c dt curl E^{n+theta}
    from the 2x2 face averages of the cell's 8 corners (coeff4 = c dt / dx / 4)."""
    noc = np.asarray(node_of_corner).reshape(-1, 729)
    nb = noc.shape[0]
    c4 = [0.25 * (cDt / (dx[d] * length_conv)) for d in range(3)]
    out = np.array(Bc, np.float64, copy=True)
    ii, jj, kk = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    for b in range(nb):
        cell = (ii + 8 * jj + 64 * kk).ravel()
        E = {}
        for a in range(2):
            for bb in range(2):
                for cc in range(2):
                    nd = noc[b, ((ii + a) + 9 * (jj + bb) + 81 * (kk + cc)).ravel()]
                    E[(a, bb, cc)] = Eth[nd] * E_conv
        t = np.zeros((len(cell), 3))
        for p in range(2):
            for q in range(2):
                t[:, 0] += -c4[1] * (E[(p, 1, q)][:, 2] - E[(p, 0, q)][:, 2]) + c4[2] * (E[(p, q, 1)][:, 1] - E[(p, q, 0)][:, 1])
        for p in range(2):
            for q in range(2):
                t[:, 1] += -c4[2] * (E[(p, q, 1)][:, 0] - E[(p, q, 0)][:, 0]) + c4[0] * (E[(1, p, q)][:, 2] - E[(0, p, q)][:, 2])
        for p in range(2):
            for q in range(2):
                t[:, 2] += -c4[0] * (E[(1, p, q)][:, 1] - E[(0, p, q)][:, 1]) + c4[1] * (E[(p, 1, q)][:, 0] - E[(p, 0, q)][:, 0])
        out[b * 512 + cell] = Bc[b * 512 + cell] + t / B_conv
    return out
