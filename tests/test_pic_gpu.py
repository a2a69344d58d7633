"""GPU parity of the FP64 push, migration and ECSIM deposit against the CPU
oracle on the same seeded inputs (bit-exact ids, codes, x and v)."""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import mesh as M
from amps_amd.pic import PicState, Species
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _by_id(p, code=None):
    o = np.argsort(p.id)
    q = p.take(o)
    return q, (None if code is None else code[o])


def _case(mesh, n, vth, dt, seed=1):
    E, B = M.corner_fields(mesh)
    parts = M.uniform_particles(mesh, n, vth, seed_x=seed, seed_v=seed + 1)
    return E, B, parts, Species([1.0], [1.0], [dt])


@pytest.mark.parametrize("mover", [L.MOVER_LAPENTA, L.MOVER_BORIS, L.MOVER_RELATIVISTIC])
@pytest.mark.parametrize("bc", [L.BC_REFLECT, L.BC_DELETE])
@pytest.mark.parametrize("amr", [False, True])
def test_push_single_step_bitexact(gpu_ctx, mover, bc, amr):
    mesh = M.central_octant_mesh() if amr else M.uniform_mesh(2)
    E, B, parts, sp = _case(mesh, 40000, 0.5, 0.05)
    print(f"push case mover={mover} bc={bc} amr={amr} n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    st.MoveParticles(mover=mover, bc_mode=bc, r_sphere=0.2, compact=False)
    g, gcode = st.get_particles()
    om = O.Mesh(mesh)
    o, ocode = O.push(om, E, B, sp, parts, mover=mover, bc_mode=bc, r_sphere=0.2)
    g, gcode = _by_id(g, gcode)
    o, ocode = _by_id(o, ocode)
    assert np.array_equal(gcode, ocode)
    assert set(np.unique(ocode)) >= ({-1, 3, 5} if bc == L.BC_REFLECT else {2, 3, 5})
    assert np.array_equal(g.x, o.x) and np.array_equal(g.v, o.v)
    alive = (ocode == 3) | (ocode == -1)
    assert np.array_equal(g.blk[alive], o.blk[alive]) and np.array_equal(g.cell[alive], o.cell[alive])


@pytest.mark.parametrize("mover", [L.MOVER_LAPENTA, L.MOVER_BORIS])
def test_push_waves4_variant(gpu_ctx, mover):
    """The push kernels held to 128 registers: FP64 bit-exact against the
    oracle, FP32 bit-identical to the unconstrained build."""
    mesh = M.central_octant_mesh()
    E, B, parts, sp = _case(mesh, 40000, 0.5, 0.05, seed=5)
    om = O.Mesh(mesh)
    o, ocode = O.push(om, E, B, sp, parts, mover=mover, bc_mode=L.BC_REFLECT, r_sphere=0.2)
    o, ocode = _by_id(o, ocode)
    res = {}
    try:
        for w in (4, 1):
            gpu_ctx.set_param("push_waves", w)
            st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
            st.set_particles(parts)
            st.MoveParticles(mover=mover, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=False)
            g, gcode = _by_id(*st.get_particles())
            if w == 4:
                assert np.array_equal(gcode, ocode)
                assert np.array_equal(g.x, o.x) and np.array_equal(g.v, o.v)
            st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
            st.set_particles(parts, precision=L.FP32)
            st.MoveParticles(mover=mover, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=False)
            res[w] = _by_id(*st.get_particles())
    finally:
        gpu_ctx.set_param("push_waves", 1)
    (a, ca), (b, cb) = res[4], res[1]
    assert np.array_equal(ca, cb) and np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v)


@pytest.mark.parametrize("bc", [L.BC_REFLECT, L.BC_DELETE])
def test_push_relativistic_si_bitexact(gpu_ctx, bc):
    """Relativistic protons (0.1c-0.95c) in SI units on the 2-level mesh
    scaled to metres: gyro-limited sub-steps that cross blocks (fields of
    other blocks read from HBM), exits along the chord, sphere absorption."""
    c = 299792458.0
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh, E0=(1.0e5, 0.0, 0.0), moment=0.1, bz0=0.05)
    parts = M.uniform_particles(mesh, 40000, 1.0, seed_x=21, seed_v=22)
    rng = np.random.default_rng(23)
    speed = rng.uniform(0.1, 0.95, parts.n) * c
    parts.v *= speed / np.linalg.norm(parts.v, axis=0)
    sp = Species([1.602176565e-19], [1.67262192e-27], [1.0e-8])
    print(f"relativistic SI case bc={bc} n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    st.MoveParticles(mover=L.MOVER_RELATIVISTIC, bc_mode=bc, r_sphere=0.2, compact=False)
    g, gcode = st.get_particles()
    om = O.Mesh(mesh)
    o, ocode = O.push(om, E, B, sp, parts, mover=L.MOVER_RELATIVISTIC, bc_mode=bc, r_sphere=0.2)
    g, gcode = _by_id(g, gcode)
    o, ocode = _by_id(o, ocode)
    assert np.array_equal(gcode, ocode)
    assert np.array_equal(g.x, o.x) and np.array_equal(g.v, o.v)
    alive = (ocode == 3) | (ocode == -1)
    assert alive.sum() > 1000
    assert np.array_equal(g.blk[alive], o.blk[alive]) and np.array_equal(g.cell[alive], o.cell[alive])


def _exit_cmp(ex, o, ocode):
    sel = np.flatnonzero(ocode >= L.CODE_EXIT_FACE0)
    assert len(sel) > 100
    oi = sel[np.argsort(o.id[sel])]
    gi = np.argsort(ex["id"])
    assert np.array_equal(ex["id"][gi], o.id[oi])
    assert np.array_equal(ex["x"][gi], o.x[:, oi].T) and np.array_equal(ex["v"][gi], o.v[:, oi].T)
    assert np.array_equal(ex["face"][gi], ocode[oi] - L.CODE_EXIT_FACE0)
    assert np.array_equal(ex["blk"][gi], o.blk[oi]) and np.array_equal(ex["spec"][gi], o.spec[oi])


@pytest.mark.parametrize("mover", [L.MOVER_LAPENTA, L.MOVER_BORIS, L.MOVER_RELATIVISTIC])
def test_user_mode_exit_list_bitexact(gpu_ctx, mover):
    """USER_FUNCTION boundary mode: the device exit list holds exactly the
    oracle's exits with the arguments of ProcessOutsideDomainParticles
    (state moved onto the face, face index, block), bit for bit; the
    survivors match the oracle too."""
    mesh = M.central_octant_mesh()
    if mover == L.MOVER_RELATIVISTIC:
        c = 299792458.0
        E, B = M.corner_fields(mesh, E0=(1.0e5, 0.0, 0.0), moment=0.1, bz0=0.05)
        parts = M.uniform_particles(mesh, 40000, 1.0, seed_x=31, seed_v=32)
        speed = np.random.default_rng(33).uniform(0.1, 0.95, parts.n) * c
        parts.v *= speed / np.linalg.norm(parts.v, axis=0)
        sp = Species([1.602176565e-19], [1.67262192e-27], [1.0e-8])
    else:
        E, B, parts, sp = _case(mesh, 40000, 0.5, 0.05, seed=31)
    print(f"USER exit list mover={mover} n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    st.get_exits()
    st.MoveParticles(mover=mover, bc_mode=L.BC_USER, r_sphere=0.2, compact=True)
    ex = st.get_exits()
    o, ocode = O.push(O.Mesh(mesh), E, B, sp, parts, mover=mover, bc_mode=L.BC_USER, r_sphere=0.2)
    _exit_cmp(ex, o, ocode)
    assert np.all((ex["face"] >= 0) & (ex["face"] < 6)) and np.all(ex["blk"] >= 0)
    g, _ = _by_id(st.get_particles()[0])
    alive = np.flatnonzero((ocode == 3) | (ocode == -1))
    oa = o.take(alive[np.argsort(o.id[alive])])
    assert np.array_equal(g.id, oa.id) and np.array_equal(g.x, oa.x) and np.array_equal(g.v, oa.v)
    assert st.get_exits()["id"].size == 0  # cleared


def test_user_mode_exit_list_graph_steps_and_capacity(gpu_ctx):
    """The exit list accumulates over graph-captured steps exactly as over
    eager ones (every particle lost without a sphere is an exit), and a full
    list reports the dropped exits."""
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    # thermal speeds of 0.3 and steps of 0.05 move a particle ~0.015 per
    # step, so the cells next to the box (1/16 wide) empty within 6 steps
    dt = 0.05
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    res = []
    for graph in (False, True):
        st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
        st.generate_particles(mesh.ncells * 3, [0.3, 0.3], drift=(0.01, 0.0, 0.0), seed=19)
        n0 = st.particle_count()
        st.get_exits()
        st.run_steps(6, dt, bc_mode=L.BC_USER, r_sphere=0.0, use_graph=graph)
        ex = st.get_exits()
        assert n0 - st.particle_count() == len(ex["id"]) > 100
        o = np.argsort(ex["id"])
        res.append({k: v[o] for k, v in ex.items()})
    for k in res[0]:
        assert np.array_equal(res[0][k], res[1][k]), k
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.generate_particles(mesh.ncells * 3, [0.3, 0.3], drift=(0.01, 0.0, 0.0), seed=19)
    gpu_ctx.set_param("exit_capacity", 16)
    try:
        st.run_steps(2, dt, bc_mode=L.BC_USER, r_sphere=0.0, use_graph=False)
        with pytest.raises(L.AmpsGpuError, match="dropped"):
            st.get_exits()
    finally:
        gpu_ctx.set_param("exit_capacity", 1 << 20)


@pytest.mark.parametrize("deterministic", [1, 0])
def test_push_multistep_with_migration(gpu_ctx, deterministic):
    """12 pushes with migration against the oracle, per particle bit for bit;
    deterministic 0 is the scatter in the order its atomics land."""
    mesh = M.central_octant_mesh()
    E, B, parts, sp = _case(mesh, 60000, 0.2, 0.05, seed=7)
    gpu_ctx.set_param("deterministic", deterministic)
    try:
        _multistep(gpu_ctx, mesh, E, B, parts, sp)
    finally:
        gpu_ctx.set_param("deterministic", 1)


def _multistep(gpu_ctx, mesh, E, B, parts, sp):
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    om = O.Mesh(mesh)
    o = parts
    for step in range(12):
        st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=True)
        o, code = O.push(om, E, B, sp, o, mover=0, bc_mode=1, r_sphere=0.2)
        o = o.take(np.flatnonzero((code == 3) | (code == -1)))
        g, _ = st.get_particles()
        key = g.blk.astype(np.int64) * 512 + g.cell
        assert np.all(np.diff(key) >= 0), "not sorted by (block, cell)"
        gs, _ = _by_id(g)
        os_, _ = _by_id(o)
        assert g.n == o.n, step
        assert np.array_equal(gs.id, os_.id)
        assert np.array_equal(gs.x, os_.x) and np.array_equal(gs.v, os_.v)
        assert np.array_equal(gs.blk, os_.blk) and np.array_equal(gs.cell, os_.cell)


@pytest.mark.parametrize("small", [64, 1])
def test_migration_order_is_deterministic(gpu_ctx, small):
    """After every compaction each cell holds the particles that stayed in it,
    in their previous order, then the ones that moved in, in their previous
    array order: a layout fixed by the step itself, not by the order in which
    the scatter's atomics land.  small=1 sends every cell with two or more
    movers through the workgroup merge sort."""
    mesh = M.central_octant_mesh()
    E, B, parts, sp = _case(mesh, 60000, 0.2, 0.05, seed=5)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    gpu_ctx.set_param("migration_sort_small", small)
    try:
        st.set_particles(parts)
        moved = 0
        for step in range(4):
            a, _ = st.get_particles()
            st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_DELETE, r_sphere=0.2, compact=True)
            b, _ = st.get_particles()
            pos = np.full(int(a.id.max()) + 1, -1)
            pos[a.id] = np.arange(a.n)
            old = pos[b.id]
            assert (old >= 0).all()
            ka = a.blk.astype(np.int64) * 512 + a.cell
            kb = b.blk.astype(np.int64) * 512 + b.cell
            mover = kb != ka[old]
            moved += int(mover.sum())
            want = np.lexsort((old, mover, kb))
            assert np.array_equal(want, np.arange(b.n)), step
        per_key = np.bincount(kb[mover], minlength=mesh.ncells)
        print(f"movers over 4 steps: {moved}; last step's cells with >= 2 movers: {(per_key >= 2).sum()}",
              flush=True)
        assert moved > 1000 and (per_key >= 2).sum() > 100
    finally:
        gpu_ctx.set_param("migration_sort_small", 64)


@pytest.mark.parametrize("small,kernel", [(64, 2), (1, 2), (64, 1)])
def test_step_migration_order_with_crowded_arrivals(gpu_ctx, small, kernel):
    """The same layout from the step's own migration, whose push classifies
    the survivors and lists each cell's arrivals (16 per cell, the rest in
    an overflow list): about 200 particles per cell and fast ones, so that
    many cells take more than 16 arrivals and some more than 64 (the
    workgroup merge sort; small=1 sends every cell with two or more there).
    With the matrix-core cell pass the step leaves the records for the next
    push to move; with the record-batch pass (kernel 1) its migration moves
    them through each arrival's destination."""
    from amps_amd import autotune

    mesh = M.uniform_mesh(1)
    E, B = M.corner_fields(mesh)
    dt = 0.05
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.generate_particles(mesh.ncells * 100, [0.8, 0.8], seed=11)
    gpu_ctx.set_param("migration_sort_small", small)
    autotune.set_deposit_variant(gpu_ctx, kernel, 4)
    try:
        st.run_steps(1, dt, r_sphere=0.2, use_graph=False)
        over16 = over64 = 0
        for step in range(3):
            a, _ = st.get_particles()
            st.run_steps(1, dt, r_sphere=0.2, use_graph=False)
            b, _ = st.get_particles()
            pos = np.full(int(a.id.max()) + 1, -1)
            pos[a.id] = np.arange(a.n)
            old = pos[b.id]
            assert (old >= 0).all()
            ka = a.blk.astype(np.int64) * 512 + a.cell
            kb = b.blk.astype(np.int64) * 512 + b.cell
            mover = kb != ka[old]
            want = np.lexsort((old, mover, kb))
            assert np.array_equal(want, np.arange(b.n)), step
            per_key = np.bincount(kb[mover], minlength=mesh.ncells)
            over16 += int((per_key > 16).sum())
            over64 += int((per_key > 64).sum())
            print(f"step {step}: {b.n} particles, {int(mover.sum())} arrivals, at most {per_key.max()} "
                  f"in a cell", flush=True)
        print(f"cells with > 16 arrivals: {over16}, > 64: {over64}", flush=True)
        assert over16 > 100 and over64 > 0
    finally:  # the session context's defaults
        gpu_ctx.set_param("migration_sort_small", 64)
        autotune.set_deposit_variant(gpu_ctx, 2, 12)


def test_config0_workload_steps(gpu_ctx):
    """The plain FP64 workload (2^18 of the 2^20 protons, dt from max
    omega_c dt = 0.1) for 20 steps with reflection and absorption."""
    mesh = M.uniform_mesh(2)
    E, B = M.corner_fields(mesh)
    parts = M.uniform_particles(mesh, 1 << 18, 0.01)
    dt = 0.1 / np.max(np.linalg.norm(B, axis=-1))
    sp = Species([1.0], [1.0], [dt])
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    om = O.Mesh(mesh)
    o = parts
    for _ in range(20):
        st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=True)
        o, code = O.push(om, E, B, sp, o, mover=0, bc_mode=1, r_sphere=0.2)
        o = o.take(np.flatnonzero((code == 3) | (code == -1)))
    g, _ = st.get_particles()
    gs, _ = _by_id(g)
    os_, _ = _by_id(o)
    assert np.array_equal(gs.id, os_.id)
    assert np.array_equal(gs.x, os_.x) and np.array_equal(gs.v, os_.v)
    assert np.array_equal(gs.blk, os_.blk) and np.array_equal(gs.cell, os_.cell)


@pytest.fixture(scope="module")
def dep_case():
    return _dep_case_data()


def _dep_case_data():
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    parts = M.per_cell_particles(mesh, [4, 4], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    return mesh, E, B, parts, sp, dt


DEPOSIT_VARIANTS = [(3, 8), (3, 4), (2, 12), (1, 6), (1, 4), (1, 12), (0, 4), (0, 2), (0, 3), (0, 6)]


def set_deposit_variant(ctx, kernel, nch):
    # a chunk count of 4 is valid for every kernel, so pass through it
    ctx.set_param("deposit_chunks", 4)
    ctx.set_param("deposit_kernel", kernel)
    ctx.set_param("deposit_chunks", nch)


@pytest.mark.parametrize("exact", [1, 0])
@pytest.mark.parametrize("kernel,nch", DEPOSIT_VARIANTS)
def test_deposit_parity(gpu_ctx, dep_case, exact, kernel, nch):
    mesh, E, B, parts, sp, dt = dep_case
    print(f"deposit case kernel={kernel} exact={exact} nch={nch} n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    set_deposit_variant(gpu_ctx, kernel, nch)
    gpu_ctx.set_param("deposit_exact", exact)
    J, Mm = st.UpdateJMassMatrix(dt)
    g, _ = st.get_particles()  # the device order (the oracle sums in it)
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    gnoc, gkeys = st.node_map()
    assert np.array_equal(noc, gnoc) and np.array_equal(keys, gkeys)
    Jo, Mo = O.deposit(om, B, sp, dt, g, noc, len(keys))
    if exact:
        assert np.array_equal(J, Jo) and np.array_equal(Mm, Mo)
    else:
        nj = np.linalg.norm(J - Jo, axis=1) / np.maximum(np.linalg.norm(Jo, axis=1), 1e-300)
        nm = np.linalg.norm(Mm - Mo, axis=1) / np.maximum(np.linalg.norm(Mo, axis=1), 1e-300)
        assert nj.max() <= 1e-12 and nm.max() <= 1e-12, (nj.max(), nm.max())
    gpu_ctx.set_param("deposit_exact", 0)
    set_deposit_variant(gpu_ctx, 1, 6)


@pytest.mark.parametrize("kernel,nch", [(3, 8), (2, 12), (1, 6)])
def test_deposit_fp32_mode_error(gpu_ctx, dep_case, kernel, nch):
    """FP32 mode keeps velocities in FP32 and deposits in FP64: against the
    FP64 oracle on the original FP64 velocities, M (which depends on x and B
    only) stays within 1e-12 and J carries the velocity rounding (stated
    here); against the oracle fed the same FP32-rounded velocities both are
    within 1e-12."""
    mesh, E, B, parts, sp, dt = dep_case
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts, precision=L.FP32)
    set_deposit_variant(gpu_ctx, kernel, nch)
    try:
        J, Mm = st.UpdateJMassMatrix(dt)
    finally:
        set_deposit_variant(gpu_ctx, 1, 6)
    g, _ = st.get_particles()  # device order, velocities widened from FP32
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    by_id = np.argsort(parts.id)
    g64 = g.take(np.arange(g.n))
    g64.v = np.ascontiguousarray(parts.v[:, by_id[np.searchsorted(parts.id[by_id], g.id)]])
    assert np.array_equal(g64.id, g.id) and np.array_equal(g64.x, g.x)
    Jo, Mo = O.deposit(om, B, sp, dt, g64, noc, len(keys))
    Jr, Mr = O.deposit(om, B, sp, dt, g, noc, len(keys))

    def rel(a, b):
        return (np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)).max()

    # J of the two species nearly cancels at many nodes (same drift, opposite
    # charges), so its FP32-velocity error is stated relative to the largest
    # node current; M does not depend on v
    ej = np.abs(J - Jo).max() / np.abs(Jo).max()
    em = rel(Mm, Mo)
    ejr = np.abs(J - Jr).max() / np.abs(Jr).max()
    dv = np.abs(g.v - g64.v).max() / np.abs(g64.v).max()
    print(f"FP32 mode deposit kernel={kernel}: J error {ej:.2e} of max |J|, M rel error {em:.2e} "
          f"against FP64 velocities (J {ejr:.2e} against the FP32-rounded ones; velocity rounding {dv:.2e})",
          flush=True)
    assert em <= 1e-12 and ej <= 1e-6
    assert ejr <= 1e-12 and rel(Mm, Mr) <= 1e-12


@pytest.mark.parametrize("precision", [L.FP64, L.FP32])
def test_particle_energy_cfl_bitexact(gpu_ctx, dep_case, precision):
    """UpdateJMassMatrix's particle energy and per-species max CFL on the
    device, bit-exact with the oracle over the device order (FP32 mode: the
    same sums over the FP32 velocities)."""
    mesh, E, B, parts, sp, dt = dep_case
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts, precision=precision)
    st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=True)
    e, cfl = st.ParticleEnergyCfl(dt)
    g, _ = st.get_particles()
    eo, cflo, _, _ = O.energy_cfl(O.Mesh(mesh), sp, dt, g)
    print(f"energy {e!r} vs {eo!r}, cfl {cfl} vs {cflo}", flush=True)
    assert e == eo and np.array_equal(cfl, cflo)
    assert e > 0 and np.all(cfl > 0)


def test_node_gather_staged_matches_direct(gpu_ctx, dep_case):
    """The staged node gather reads the same operands and sums them in the
    same order as the per-value gather: bit-identical nodes."""
    mesh, E, B, parts, sp, dt = dep_case
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    out = {}
    try:
        for staged in (1, 0):
            gpu_ctx.set_param("node_staged", staged)
            for nb in (0, 1, 8):
                gpu_ctx.set_param("node_batch", nb)
                out[staged, nb] = st.UpdateJMassMatrix(dt)
    finally:
        gpu_ctx.set_param("node_staged", 1)
        gpu_ctx.set_param("node_batch", 8)
    J0, M0 = out[0, 1]
    for k, (J, Mm) in out.items():
        assert np.array_equal(J, J0) and np.array_equal(Mm, M0), k


@pytest.mark.parametrize("kernel,paired,nch", [(3, 0, 8), (3, 0, 4), (2, 1, 12), (2, 0, 12), (1, 0, 12)])
def test_deposit_ragged_cells(gpu_ctx, kernel, paired, nch):
    """Empty cells, cells of 1-3 particles and cells of 65-80 (more than one
    64-particle chunk of the matrix-core pass) side by side."""
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    full = M.per_cell_particles(mesh, [40, 40], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    rng = np.random.default_rng(17)
    key = full.blk.astype(np.int64) * 512 + full.cell
    keep_p = np.array([0.0, 0.015, 0.3, 1.0])[rng.integers(0, 4, mesh.nblocks * 512)]
    parts = full.take(np.flatnonzero(rng.random(full.n) < keep_p[key]))
    counts = np.bincount(parts.blk.astype(np.int64) * 512 + parts.cell, minlength=mesh.nblocks * 512)
    print(f"ragged deposit kernel={kernel} n={parts.n} empty={np.sum(counts == 0)} max={counts.max()}", flush=True)
    assert np.sum(counts == 0) > 1000 and counts.max() > 64 and np.sum((counts > 0) & (counts < 4)) > 100
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    set_deposit_variant(gpu_ctx, kernel, nch)
    gpu_ctx.set_param("deposit_paired", paired)
    try:
        J, Mm = st.UpdateJMassMatrix(dt)
    finally:
        set_deposit_variant(gpu_ctx, 1, 6)
        gpu_ctx.set_param("deposit_paired", 0)
    g, _ = st.get_particles()
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    Jo, Mo = O.deposit(om, B, sp, dt, g, noc, len(keys))
    nj = np.linalg.norm(J - Jo, axis=1) / np.maximum(np.linalg.norm(Jo, axis=1), 1e-300)
    nm = np.linalg.norm(Mm - Mo, axis=1) / np.maximum(np.linalg.norm(Mo, axis=1), 1e-300)
    assert nj.max() <= 1e-12 and nm.max() <= 1e-12, (nj.max(), nm.max())
    # nodes with no particle in any incident cell are exactly zero
    zero = np.all(Mo == 0, axis=1)
    assert np.all(Mm[zero] == 0) and np.all(J[zero] == 0)


def test_deposit_after_push_and_migration(gpu_ctx, dep_case):
    mesh, E, B, parts, sp, dt = dep_case
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    for _ in range(3):
        st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.0, compact=True)
    J, Mm = st.UpdateJMassMatrix(dt)
    g, _ = st.get_particles()
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    Jo, Mo = O.deposit(om, B, sp, dt, g, noc, len(keys))
    nm = np.linalg.norm(Mm - Mo, axis=1) / np.maximum(np.linalg.norm(Mo, axis=1), 1e-300)
    assert nm.max() <= 1e-12
    # J = sum over particles of q W (alpha v) / V: the total is order-independent
    assert np.allclose(J.sum(0), Jo.sum(0), rtol=1e-12, atol=1e-12 * np.abs(Jo).sum())


def test_generator_and_graph_steps_match_eager(gpu_ctx):
    """Device-generated population: per-cell counts and positions inside the
    cells; hipGraph-captured steps reproduce eager steps bit for bit: the
    same particles in the same array order (the migration's layout does not
    depend on atomic scheduling) and so the same deposit."""
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt * 50, dt * 50])
    nps = mesh.ncells * 3 + 1001
    res = []
    for graph in (False, True):
        st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
        st.generate_particles(nps, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=9)
        if not graph:
            g0, _ = st.get_particles()
            assert g0.n == 2 * nps
            key = g0.blk.astype(np.int64) * 512 + g0.cell
            cnt = np.bincount(key, minlength=mesh.ncells)
            assert cnt.min() == 6 and cnt.max() == 8
            lo = mesh.blk_xmin[g0.blk]
            assert np.all(g0.x.T >= lo) and np.all(g0.x.T <= mesh.blk_xmax[g0.blk])
        st.run_steps(5, dt * 50, r_sphere=0.2, use_graph=graph)
        J, Mm = st.UpdateJMassMatrix(dt * 50)
        g, _ = st.get_particles()
        res.append((g, J, Mm))
    (a, Ja, Ma), (b, Jb, Mb) = res
    o1, o2 = np.argsort(a.id), np.argsort(b.id)
    assert np.array_equal(a.id[o1], b.id[o2])
    assert np.array_equal(a.x[:, o1], b.x[:, o2]) and np.array_equal(a.v[:, o1], b.v[:, o2])
    assert np.array_equal(a.blk[o1], b.blk[o2]) and np.array_equal(a.cell[o1], b.cell[o2])
    assert np.array_equal(a.id, b.id)
    assert np.array_equal(Ja, Jb) and np.array_equal(Ma, Mb)


def _record_move_runs(ctx, mesh, E, B, sp, dt, precision, B_center=None):
    """Three steps three ways, one after the other (the states share the
    device context): one run_steps call per step (True), separate
    MoveParticles(compact) and UpdateJMassMatrix calls (False), one
    graph-captured run_steps call ("run"); the arrays and the mass term and
    right-hand side over the resident J and M after each step.  B_center,
    when given, is set as each state's cell-centred B."""
    rng = np.random.default_rng(3)
    res = {}
    for fused in (True, False, "run"):
        st = PicState(mesh, E, B, sp, ctx=ctx)
        if B_center is not None:
            st.set_center_fields(B_center)
        st.generate_particles(mesh.ncells * 3 + 777, [0.3, 0.3], drift=(0.01, 0.0, 0.0), seed=5,
                              precision=precision)
        if fused is True:
            En = rng.normal(size=(st.nnodes, 3))
            Bc = rng.normal(size=(mesh.nblocks * 512, 3))
            prm = st.ecsim_params(dt)
        if fused == "run":
            st.run_steps(3, dt, r_sphere=0.2, use_graph=True)
            res[fused] = (st.get_particles()[0], st.MassMatrixApply(En, -4 * np.pi * dt * 0.5),
                          st.EcsimRhs(prm, En, Bc))
            continue
        out = []
        for _ in range(3):
            if fused:
                st.run_steps(1, dt, r_sphere=0.2, use_graph=False)
            else:
                st.MoveParticles(r_sphere=0.2, compact=True)
                st.UpdateJMassMatrix(dt, copy_out=False)
            y = st.MassMatrixApply(En, -4 * np.pi * dt * 0.5)
            r = st.EcsimRhs(prm, En, Bc)
            out.append((st.get_particles()[0], y, r))
        res[fused] = out
    return res


@pytest.mark.parametrize("kernel", [2, 3])
@pytest.mark.parametrize("precision", [L.FP64, L.FP32])
def test_step_deposit_moves_the_migrated_records(gpu_ctx, precision, kernel):
    """In a step the deterministic migration only works out where each
    particle goes: the cell pass (matrix-core or lane) reads the particles
    where they were, the next push moves them as it updates them, and
    run_steps moves the last step's before it returns.  The arrays (in array
    order) and the resident J and M after each step equal those of separate
    MoveParticles(compact) and UpdateJMassMatrix calls, whose migration moves
    the records itself, and so do those after three steps in one
    graph-captured run_steps call (pushes that move the records); J and M
    are compared through the field equation's mass term and right-hand side,
    which read them on the device."""
    from amps_amd import autotune

    mesh = M.uniform_mesh(1)
    E, B = M.corner_fields(mesh)
    # thermal speeds of 0.3 and steps of 0.05: a particle moves ~0.015 per
    # step, an eighth of a cell, so a good share change cells every step
    dt = 0.05
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    autotune.set_deposit_variant(gpu_ctx, kernel, 4)
    try:
        res = _record_move_runs(gpu_ctx, mesh, E, B, sp, dt, precision)
    finally:  # the session context's default
        autotune.set_deposit_variant(gpu_ctx, 2, 12)
    _check_record_move(res, kernel)


@pytest.mark.parametrize("kernel", [2, 3])
def test_step_deposit_moves_the_migrated_records_center_B(gpu_ctx, kernel):
    """The same with the cell-centred B (FP64 push and cell pass reading the
    centre field): a step's cell pass and the next push read the particles
    through the source indices and match the separate calls bit for bit."""
    from amps_amd import autotune

    mesh = M.uniform_mesh(1)
    E, B = M.corner_fields(mesh)
    Bc = M.center_fields(mesh, lambda P: M.dipole_field(P))
    dt = 0.05
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    autotune.set_deposit_variant(gpu_ctx, kernel, 4)
    gpu_ctx.set_param("b_center", 1)
    try:
        res = _record_move_runs(gpu_ctx, mesh, E, B, sp, dt, L.FP64, B_center=Bc)
    finally:  # the session context's defaults
        gpu_ctx.set_param("b_center", 0)
        autotune.set_deposit_variant(gpu_ctx, 2, 12)
    _check_record_move(res, kernel)


def _check_record_move(res, kernel):
    """Bit-equal arrays and J/M terms after each step; more than a tenth of
    the particles change cells in each step after the first."""
    prev = None
    for step, ((a, ya, ra), (b, yb, rb)) in enumerate(zip(res[True], res[False])):
        key = dict(zip(a.id.tolist(), (a.blk.astype(np.int64) * 512 + a.cell).tolist()))
        moved = 0 if prev is None else sum(prev.get(i, -1) != k for i, k in key.items())
        prev = key
        print(f"kernel {kernel} step {step}: {a.n} particles, {moved} changed cells", flush=True)
        assert a.n == b.n and a.n > 0 and (step == 0 or moved > a.n // 10)
        assert np.array_equal(a.id, b.id)
        assert np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v) and np.array_equal(a.spec, b.spec)
        assert np.array_equal(a.blk, b.blk) and np.array_equal(a.cell, b.cell)
        assert np.array_equal(ya, yb) and np.array_equal(ra, rb)
    (a, ya, ra), (b, yb, rb) = res["run"], res[False][-1]
    assert np.array_equal(a.id, b.id) and np.array_equal(a.spec, b.spec)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.v, b.v)
    assert np.array_equal(a.blk, b.blk) and np.array_equal(a.cell, b.cell)
    assert np.array_equal(ya, yb) and np.array_equal(ra, rb)


@pytest.mark.parametrize("mover", [L.MOVER_LAPENTA, L.MOVER_BORIS])
def test_push_fp32_error_against_fp64_oracle(gpu_ctx, mover):
    """FP32 push (FP32 fields and velocity, FP64 positions): state the error of
    every output against the FP64 oracle after one step and after 20 steps."""
    mesh = M.central_octant_mesh()
    E, B, parts, sp = _case(mesh, 40000, 0.2, 0.05, seed=21)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts, precision=L.FP32)
    om = O.Mesh(mesh)
    o = parts
    for step in range(20):
        st.MoveParticles(mover=mover, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=(step < 19))
        o, ocode = O.push(om, E, B, sp, o, mover=mover, bc_mode=1, r_sphere=0.2)
        if step < 19:
            o = o.take(np.flatnonzero((ocode == 3) | (ocode == -1)))
    g, gcode = st.get_particles()
    g, gcode = _by_id(g, gcode)
    o, ocode = _by_id(o, ocode)
    common = np.intersect1d(g.id, o.id)
    gi = np.searchsorted(g.id, common)
    oi = np.searchsorted(o.id, common)
    v_rel = np.max(np.abs(g.v[:, gi] - o.v[:, oi])) / np.max(np.abs(o.v))
    x_abs = np.max(np.abs(g.x[:, gi] - o.x[:, oi]))
    same_code = np.mean(gcode[gi] == ocode[oi])
    same_cell = np.mean((g.blk[gi] == o.blk[oi]) & (g.cell[gi] == o.cell[oi]))
    print(f"FP32 mover {mover}: 20 steps max|dv|/max|v| {v_rel:.2e}, max|dx| {x_abs:.2e}, "
          f"codes equal {same_code:.6f}, cells equal {same_cell:.6f}, survivors {g.n} vs {o.n}")
    assert v_rel < 1e-4 and x_abs < 1e-5  # FP32 eps ~ 6e-8, 20 steps of rotation
    assert same_code > 0.999 and same_cell > 0.999
    assert abs(g.n - o.n) <= max(2, o.n // 5000)


@pytest.mark.parametrize("layout_name,weight", [("aligned", False), ("aligned", True), ("packed", True)])
def test_restart_records_round_trip(gpu_ctx, layout_name, weight):
    """ParticleBuffer records in restart order: the export follows
    SaveParticleDataBlock's order (oracle restatement) with the fields at the
    picParticleDataMacro.h offsets; the import restores the device state
    (positions, velocities, species, blocks, cells and the order within each
    cell), after migration steps have shuffled the in-cell order."""
    mesh, E, B, parts, sp, dt = _dep_case_data()
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    for _ in range(2):
        st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.0, compact=True)
    g, _ = st.get_particles()
    layout = L.aos_layout_aligned(weight) if layout_name == "aligned" else L.aos_layout_packed(weight)
    counts, rec = st.get_restart(layout)
    table, order = O.restart_order(g.blk, g.cell, mesh.nblocks)
    assert np.array_equal(counts, table)
    assert rec.shape == (g.n, layout.record_bytes)
    def field(off, n, dt_):
        return np.frombuffer(np.ascontiguousarray(rec[:, off:off + 8 * n]).tobytes(), dt_).reshape(-1, n)
    assert np.array_equal(field(layout.velocity_offset, 3, np.float64), g.v[:, order].T)
    assert np.array_equal(field(layout.position_offset, 3, np.float64), g.x[:, order].T)
    assert np.array_equal(rec[:, layout.species_offset], g.spec[order].astype(np.uint8))
    assert np.all(field(0, 2, np.int64) == -1)
    if weight:
        assert np.all(field(layout.weight_offset, 1, np.float64) == 1.0)
    st2 = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st2.set_restart(layout, counts, rec)
    h, _ = st2.get_particles()
    assert np.array_equal(h.x, g.x) and np.array_equal(h.v, g.v) and np.array_equal(h.spec, g.spec)
    assert np.array_equal(h.blk, g.blk) and np.array_equal(h.cell, g.cell)


def test_ecsim_mass_apply_bitexact(gpu_ctx):
    """The ECSIM mass-matrix term over the resident M of the deposit on a
    single-level mesh (the reference solver's), bit-exact with the oracle's
    restatement over the same M (boundary nodes included), also when added
    onto a caller's partial row; linear in E; repeated launches give the same
    result; a multi-level mesh is refused."""
    mesh = M.uniform_mesh(2)
    E, B = M.corner_fields(mesh)
    parts = M.per_cell_particles(mesh, [4, 4], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    print(f"ecsim mass term case n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    J, Mm = st.UpdateJMassMatrix(dt)
    noc, keys = st.node_map()
    nbr = O.ecsim_neighbours(mesh, noc, keys)
    rng = np.random.default_rng(11)
    E1, E2, S = rng.normal(size=(3, len(keys), 3))
    c1, c2 = -4 * np.pi * dt * 0.51, 1.7
    g1 = st.MassMatrixApply(E1, c1, c2)
    assert np.array_equal(g1, O.ecsim_mass_apply(nbr, Mm, E1, c1, c2))
    g2 = st.MassMatrixApply(E2, c1, c2, reps=3)
    assert np.array_equal(g2, O.ecsim_mass_apply(nbr, Mm, E2, c1, c2))
    gs = st.MassMatrixApply(E1, c1, c2, seed=S)
    assert np.array_equal(gs, O.ecsim_mass_apply(nbr, Mm, E1, c1, c2, seed=S))
    g12 = st.MassMatrixApply(E1 + E2, c1, c2)
    assert np.allclose(g12, g1 + g2, rtol=1e-12, atol=1e-12 * np.abs(g12).max())
    assert np.count_nonzero(g1) > 0.9 * g1.size
    amr = M.central_octant_mesh()
    Ea, Ba = M.corner_fields(amr)
    st2 = PicState(amr, Ea, Ba, sp, ctx=gpu_ctx)
    st2.set_particles(M.per_cell_particles(amr, [1, 1], [0.05, 0.005]))
    st2.UpdateJMassMatrix(dt, copy_out=False)
    with pytest.raises(L.AmpsGpuError, match="single-level"):
        st2.MassMatrixApply(np.zeros((st2.nnodes, 3)), c1, c2)


def test_ecsim_mass_apply_single_particle_closed_form(gpu_ctx):
    """Deposit and mass term both on the device for one particle in a dipole
    B: (q h / V) W_i alpha E(x_p) c1 c2 at the 8 corners of its cell, zero
    elsewhere (alpha = (I + h^2 B B^T - h [B]x) / (1 + h^2 |B|^2),
    h = q dt / 2m)."""
    mesh = M.uniform_mesh(1)
    E, B = M.corner_fields(mesh, bz0=0.3)
    x = np.array([[0.123], [-0.456], [0.789]])
    blk = mesh.locate(x.T)
    parts = M.Particles(x, np.array([[0.01], [0.02], [-0.03]]), [0], blk, M.cell_of(mesh, x, blk))
    q, mass, dt = -1.0, 0.01, 1e-3
    sp = Species([q], [mass], [dt], [1.0])
    print("ecsim single-particle closed form", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    st.UpdateJMassMatrix(dt, copy_out=False)
    noc, keys = st.node_map()
    xs, w, t, c = O.Mesh(mesh).stencil(x[:, 0], int(blk[0]))  # trilinear weights of the corners c
    Bi = (w[:, None] * B[blk[0]][c]).sum(0)
    h = q * dt / (2 * mass)
    Bx = np.array([[0, -Bi[2], Bi[1]], [Bi[2], 0, -Bi[0]], [-Bi[1], Bi[0], 0]])
    alpha = (np.eye(3) + h * h * np.outer(Bi, Bi) - h * Bx) / (1 + h * h * Bi @ Bi)
    V = np.prod((mesh.blk_xmax[blk[0]] - mesh.blk_xmin[blk[0]]) / 8)
    En = np.random.default_rng(5).normal(size=(len(keys), 3))
    nodes = noc[int(blk[0]) * 729 + c]
    Ep = (w[:, None] * En[nodes]).sum(0)
    c1, c2 = -4 * np.pi * dt * 0.5, 2.0
    out = st.MassMatrixApply(En, c1, c2)
    want = np.zeros_like(out)
    for wi, ni in zip(w, nodes):
        want[ni] = (q * h / V) * wi * (alpha @ Ep) * c1 * c2
    assert np.allclose(out, want, rtol=1e-11, atol=1e-13 * np.abs(want).max())
    assert np.count_nonzero(np.abs(out).sum(1)) == 8


def test_autotuner_weighs_whole_steps(gpu_ctx, monkeypatch):
    """The tuner picks the deposit pass by the deposit's time, but a pass
    other than the matrix-core and lane ones (which read a step's particles
    where the migration left them) wins only if whole steps are faster with
    it too; the chosen configuration then runs steps."""
    from amps_amd import autotune

    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [0.05, 0.05])
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.generate_particles(mesh.ncells * 8, [0.3, 0.3], seed=3)
    monkeypatch.setattr(autotune, "DEPOSIT_VARIANTS", ((1, 4), (2, 4)))
    opts = dict(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2)
    try:
        best, cached = autotune.tune_pic(gpu_ctx, st, opts, 0.05, "test", cache=None)
        print(f"tuned: {best}", flush=True)
        assert not cached and best["deposit_kernel"] in (1, 2)
        dep = best["deposit_ms"]
        if dep["k1c4"] < dep["k2c4"]:
            steps = best["step_ms"]
            assert set(steps) == {"k1c4", "k2c4"}
            assert best["deposit_kernel"] == (2 if steps["k2c4"] <= steps["k1c4"] else 1)
        else:
            assert best["deposit_kernel"] == 2 and not best["step_ms"]
        n0 = st.particle_count()
        st.run_steps(3, 0.05, **opts)
        assert 0 < st.particle_count() <= n0
        # the whole-step comparison itself
        var, steps = autotune.weigh_steps(gpu_ctx, st, opts, 0.05, ((1, 4), (2, 4)))
        print(f"steps: {steps}", flush=True)
        assert set(steps) == {(1, 4), (2, 4)} and all(t > 0 for t in steps.values())
        assert var == min(steps, key=steps.get)
    finally:  # the session context's defaults
        autotune.set_deposit_variant(gpu_ctx, 2, 12)
        gpu_ctx.set_param("tile", 8192)
        gpu_ctx.set_param("push_waves", 1)
        gpu_ctx.set_param("node_batch", 8)
        gpu_ctx.set_timing(False)
