"""The workloads at their stated sizes on the GPU, each checked against the
CPU oracle (or, at 2^27 particles, through size-independent identities):

* the plain FP64 Boris workload: 2^20 protons on 4x4x4 blocks, 100 steps;
* the ECSIM deposit at 16 + 16 particles per cell on the 2-level mesh;
* the full particle step on the 3-level mesh (levels 4-6, 7,456 blocks):
  graph-captured steps against the oracle at one particle per cell and
  species, and the full 2^27 population through its invariants;
* the batched cutoff map with the real 1e5-step cap;
* the learned surrogate: trained on the tracer's own output, the bracketed
  map against the full scan, and what it saves.
"""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import cutoff as CU
from amps_amd import mesh as M
from amps_amd import surrogate as SG
from amps_amd.pic import PicState, Species
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _by_id(p):
    return p.take(np.argsort(p.id))


def _same_state(g, o):
    g, o = _by_id(g), _by_id(o)
    assert np.array_equal(g.id, o.id)
    assert np.array_equal(g.x, o.x) and np.array_equal(g.v, o.v)
    assert np.array_equal(g.blk, o.blk) and np.array_equal(g.cell, o.cell)


def _rel_rows(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def test_plain_boris_2e20_protons_100_steps(gpu_ctx):
    """2^20 protons in [-1,1]^3 on 4x4x4 blocks of 8^3 cells, dipole +
    Bz 0.05 and E = (0.01, 0, 0) at the corners, v_th 0.01, dt from
    max omega_c dt = 0.1, 100 steps with reflection at the box and absorption
    at r < 0.2: positions, velocities, blocks and cells bit-exact by id."""
    mesh = M.uniform_mesh(2)
    E, B = M.corner_fields(mesh)
    parts = M.uniform_particles(mesh, 1 << 20, 0.01)
    dt = 0.1 / np.max(np.linalg.norm(B, axis=-1))
    sp = Species([1.0], [1.0], [dt])
    print(f"plain Boris 2^20 x 100 steps n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    om = O.Mesh(mesh)
    o = parts
    absorbed = 0
    for step in range(100):
        st.MoveParticles(mover=L.MOVER_LAPENTA, bc_mode=L.BC_REFLECT, r_sphere=0.2, compact=True)
        o, code = O.push(om, E, B, sp, o, mover=0, bc_mode=1, r_sphere=0.2)
        absorbed += int(np.sum(code == 5))
        o = o.take(np.flatnonzero((code == 3) | (code == -1)))
        if step % 25 == 24:
            g, _ = st.get_particles()
            assert g.n == o.n, step
            _same_state(g, o)
            print(f"  step {step + 1}: {g.n} particles, {absorbed} absorbed so far", flush=True)
    assert absorbed > 0


def test_ecsim_deposit_16_16_per_cell(gpu_ctx):
    """16 electrons + 16 ions per cell on the 2-level mesh (1.97 M
    particles): J and M within 1e-12 of the oracle per node (the default
    deposit), bit-exact in the exact mode; energy and CFL bit-exact."""
    mesh = M.central_octant_mesh()
    E, B = M.corner_fields(mesh)
    parts = M.per_cell_particles(mesh, [16, 16], [0.05, 0.005], drift=(0.01, 0.0, 0.0))
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    print(f"ECSIM deposit 16+16 per cell n={parts.n}", flush=True)
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.set_particles(parts)
    J, Mm = st.UpdateJMassMatrix(dt)
    g, _ = st.get_particles()
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    Jo, Mo = O.deposit(om, B, sp, dt, g, noc, len(keys))
    nj, nm = _rel_rows(J, Jo), _rel_rows(Mm, Mo)
    print(f"  max relative node error J {nj.max():.2e} M {nm.max():.2e}", flush=True)
    assert nj.max() <= 1e-12 and nm.max() <= 1e-12
    gpu_ctx.set_param("deposit_exact", 1)
    try:
        Je, Me = st.UpdateJMassMatrix(dt)
    finally:
        gpu_ctx.set_param("deposit_exact", 0)
    assert np.array_equal(Je, Jo) and np.array_equal(Me, Mo)
    e, cfl = st.ParticleEnergyCfl(dt)
    eo, cflo, _, _ = O.energy_cfl(om, sp, dt, g)
    assert e == eo and np.array_equal(cfl, cflo)


def _three_level_case():
    mesh = M.three_level_mesh()
    E, B = M.corner_fields(mesh)
    dt = 0.1 / (np.max(np.linalg.norm(B, axis=-1)) / 0.01)
    sp = Species([-1.0, 1.0], [0.01, 1.0], [dt, dt])
    return mesh, E, B, sp, dt


def test_three_level_graph_steps_against_oracle(gpu_ctx):
    """The headline step's mesh (16^3 root blocks refined twice: levels 4, 5
    and 6, corners where three levels meet) with one electron and one ion per
    cell (7.6 M particles): 4 graph-captured steps of Lapenta push +
    reflection + migration + deposit.  Ids, x, v, blocks and cells bit-exact
    by id against 4 oracle pushes; the node map equal to the oracle's; J and
    M within 1e-12 of the oracle's deposit over the device order."""
    mesh, E, B, sp, dt = _three_level_case()
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    st.generate_particles(mesh.ncells, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=12)
    p0, _ = st.get_particles()
    print(f"3-level mesh: {mesh.nblocks} blocks, {st.nnodes} nodes, {p0.n} particles", flush=True)
    om = O.Mesh(mesh)
    noc, keys = om.build_nodes()
    gnoc, gkeys = st.node_map()
    assert np.array_equal(noc, gnoc) and np.array_equal(keys, gkeys)
    st.run_steps(4, dt, bc_mode=L.BC_REFLECT, r_sphere=0.0, use_graph=True)
    o = p0
    for _ in range(4):
        o, code = O.push(om, E, B, sp, o, mover=0, bc_mode=1, r_sphere=0.0)
        assert np.all((code == 3) | (code == -1))
    g, _ = st.get_particles()
    key = g.blk.astype(np.int64) * 512 + g.cell
    assert np.all(np.diff(key) >= 0)
    assert g.n == o.n == p0.n
    _same_state(g, o)
    levels = mesh.level[g.blk]
    print(f"  particles per level {np.bincount(levels)[4:]}", flush=True)
    J, Mm = st.UpdateJMassMatrix(dt)
    Jo, Mo = O.deposit(om, B, sp, dt, g, noc, len(keys))
    worst_j = worst_m = 0.0
    for a in range(0, len(keys), 1 << 19):
        sl = slice(a, a + (1 << 19))
        worst_j = max(worst_j, _rel_rows(J[sl], Jo[sl]).max())
        worst_m = max(worst_m, _rel_rows(Mm[sl], Mo[sl]).max())
    print(f"  max relative node error J {worst_j:.2e} M {worst_m:.2e}", flush=True)
    assert worst_j <= 1e-12 and worst_m <= 1e-12


def test_full_size_step_invariants(gpu_ctx):
    """The benchmark population itself (2^27 particles, two species, the
    3-level mesh) after 2 graph-captured steps: nothing lost with reflecting
    walls, keys sorted, states finite, and the total current equal to the sum
    over particles of q w (alpha v) / V (the corner weights of each particle
    sum to one), alpha from B interpolated at each particle."""
    mesh, E, B, sp, dt = _three_level_case()
    st = PicState(mesh, E, B, sp, ctx=gpu_ctx)
    nps = 1 << 26
    st.generate_particles(nps, [0.05, 0.005], drift=(0.01, 0.0, 0.0), seed=4)
    print(f"full-size step: {2 * nps} particles", flush=True)
    st.run_steps(2, dt, bc_mode=L.BC_REFLECT, r_sphere=0.0, use_graph=True)
    J, _ = st.UpdateJMassMatrix(dt, copy_out="J")
    g, code = st.get_particles()
    assert g.n == 2 * nps
    key = g.blk.astype(np.int64) * 512 + g.cell
    assert np.all(np.diff(key) >= 0)
    assert np.all(np.isfinite(g.x)) and np.all(np.isfinite(g.v))
    want = np.zeros(3)
    om = O.Mesh(mesh)
    for a in range(0, g.n, 1 << 24):
        sl = slice(a, min(a + (1 << 24), g.n))
        x, v, b, s = g.x[:, sl], g.v[:, sl], g.blk[sl], g.spec[sl]
        lo, hi = mesh.blk_xmin[b], mesh.blk_xmax[b]
        dxs = (hi - lo) / 8
        loc = (x.T - lo) / dxs
        ix = np.clip(loc.astype(np.int64), 0, 7)
        f = loc - ix
        Bp = np.zeros((len(b), 3))
        for l in range(8):
            di, dj, dk = l >> 2, (l >> 1) & 1, l & 1
            w = (f[:, 0] if di else 1 - f[:, 0]) * (f[:, 1] if dj else 1 - f[:, 1]) * (f[:, 2] if dk else 1 - f[:, 2])
            c = (ix[:, 0] + di) + 9 * ((ix[:, 1] + dj) + 9 * (ix[:, 2] + dk))
            Bp += w[:, None] * B[b, c]
        q, m = sp.charge[s], sp.mass[s]
        h = q * dt / (2 * m)
        hB = h[:, None] * Bp
        vv = v.T
        # alpha v = (v + (hB.v) hB - hB x v) / (1 + |hB|^2)
        av = (vv + (hB * vv).sum(1)[:, None] * hB - np.cross(hB, vv)) / (1 + (hB * hB).sum(1))[:, None]
        V = np.prod(dxs, axis=1)
        want += ((q / V)[:, None] * av).sum(0)
    print(f"  total J {J.sum(0)} vs {want}", flush=True)
    assert np.allclose(J.sum(0), want, rtol=1e-10, atol=1e-10 * np.abs(want).max())


def _oparams(P):
    Q = O.TraceParams()
    for name, _t in L.TraceParams._fields_:
        v = getattr(P, name)
        if hasattr(v, "__len__"):
            for i in range(len(v)):
                getattr(Q, name)[i] = v[i]
        else:
            setattr(Q, name, v)
    return Q


def test_cutoff_map_real_step_cap(gpu_ctx):
    """The batched map's parameterisation with its real 1e5-step cap on 256
    (location, direction) pairs x 48 rigidities: FP64 states and cutoff bins
    bit-exact with the oracle, trajectories that hit the cap included; the
    FP32 map's agreement stated."""
    P = CU.config2_params()
    assert P.max_steps == 100000
    pts, lat, lon = CU.config2_points(nlon=16, nlat=16)
    zen, az = CU.config2_directions(8)
    sel = np.random.default_rng(2).integers(0, 8, len(pts))
    dirs = CU.arrival_dirs(pts, zen, az)[np.arange(len(pts)), sel]
    grid = CU.build_search_grid(0.1, 20.0, 48)
    print(f"map with the 1e5 cap: {len(pts)} pairs x 48", flush=True)
    sf, bf, steps = CU.cutoff_scan(gpu_ctx, P, pts, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    o = O.trace_batch(_oparams(P), np.repeat(pts, 48, axis=0), np.repeat(dirs, 48, axis=0), np.tile(grid, len(pts)))
    capped = int(np.sum(o["steps"] >= 100000))
    ost = (o["termination"] == 0).astype(np.int8).reshape(-1, 48)
    assert np.array_equal(sf, ost)
    assert np.array_equal(bf, np.array([O.upper_cutoff_bin(r) for r in ost]))
    assert steps == int(o["steps"].sum())
    s32, b32, _ = CU.cutoff_scan(gpu_ctx, P, pts, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP32)
    print(f"  {capped} trajectories at the cap; FP32 states equal {np.mean(s32 == sf):.4f}, "
          f"bins off by > 1: {np.mean(np.abs(b32.astype(int) - bf.astype(int)) > 1):.4f}", flush=True)
    assert capped > 0
    assert np.mean(s32 == sf) > 0.95


def test_surrogate_trained_on_tracer_brackets_the_map(gpu_ctx):
    """The learned part end to end: the network is trained on 2^16 cutoffs
    that the tracer computes (FP32 downward scans of random pairs), then the
    bracketed FP64 map of the 64x32x8 grid is compared with the plain full
    FP64 scan.  Where the bracket resolves the transition (no fallback) and
    nothing above the bracket is forbidden, the bins are the full scan's;
    fallbacks are rescanned in full; the bracketed map traces far fewer Boris
    steps than the full scan, and fewer than the same method with the
    untrained network, whose brackets mostly miss."""
    P = CU.config2_params()
    grid = CU.build_search_grid(0.1, 20.0, 48)
    X, T, _, tsteps = SG.training_set(gpu_ctx, P, 1 << 16, grid, seed=6)
    sur = SG.Surrogate(gpu_ctx, seed=0)
    sur.set_data(X, T)
    sur.train(1536, batch=4096, lr=1e-3)
    _, loss = sur.grad(4096)
    pts, lat, lon = CU.config2_points()
    zen, az = CU.config2_directions(8)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    feats = SG.features(np.repeat(lat, len(zen)), np.repeat(lon, len(zen)), np.tile(zen, len(pts)),
                        np.tile(az, len(pts)))
    sf, bf, st_full = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    bs, fb, pb, st_sur = SG.bracketed_map(gpu_ctx, sur, P, x0, dirs, feats, grid, half=3, precision=L.FP64)
    untrained = SG.Surrogate(gpu_ctx, seed=1)
    bu, fbu, _, st_un = SG.bracketed_map(gpu_ctx, untrained, P, x0, dirs, feats, grid, half=3, precision=L.FP64)
    lo, hi = SG.bracket_bins(pb, len(grid), 3)
    clean_above = np.array([np.all(sf[i, hi[i]:] == 1) for i in range(len(hi))])
    resolved = ~fb
    print(f"surrogate: loss {loss:.4f}, bin MAE {np.mean(np.abs(pb - np.where(bf < 0, 47, bf))):.2f}, "
          f"fallback {np.mean(fb):.4f} (untrained {np.mean(fbu):.4f}), steps {st_sur / st_full:.3f} of the full "
          f"scan (untrained {st_un / st_full:.3f}), bins equal {np.mean(bs == bf):.4f}", flush=True)
    assert np.array_equal(bs[fb], bf[fb])
    assert np.array_equal(bs[resolved & clean_above], bf[resolved & clean_above])
    assert np.mean(np.abs(bs.astype(int) - bf.astype(int)) > 1) <= 0.01
    assert st_sur < 0.35 * st_full
    assert st_sur < st_un and np.mean(fb) < np.mean(fbu)
