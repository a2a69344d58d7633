"""GPU parity of the persistent cutoff tracer against the CPU oracle.

FP64: termination, step count, time, path length, final x and p are
bit-exact with the oracle (legacy deck and the batched-map parameterisation,
plus the forced failure paths).  FP32: classification and cutoff-bin
agreement with the FP64 oracle, stated as rates."""
import numpy as np
import pytest

from amps_amd import _lib as L
from amps_amd import cutoff as CU
from oracle import oracle as O
from tests import tracer_setup as TS

pytestmark = pytest.mark.gpu


def _oparams(P):
    """The same parameter block as the oracle's struct."""
    Q = O.TraceParams()
    for name, _t in L.TraceParams._fields_:
        v = getattr(P, name)
        if hasattr(v, "__len__"):
            for i in range(len(v)):
                getattr(Q, name)[i] = v[i]
        else:
            setattr(Q, name, v)
    return Q


def _random_tasks(n, seed, rmin=0.1, rmax=20.0, rlo=1.05, rhi=1.3, re=CU.RE_DIPOLE):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 3))
    x *= (rng.uniform(rlo, rhi, n) * re / np.linalg.norm(x, axis=1))[:, None]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    # arrival directions mostly downward
    d -= 1.2 * x / np.linalg.norm(x, axis=1)[:, None]
    d /= np.linalg.norm(d, axis=1)[:, None]
    R = np.exp(rng.uniform(np.log(rmin), np.log(rmax), n))
    return x, d, R


def _same(g, o):
    for f in ["termination", "retry_count", "steps"]:
        assert np.array_equal(g[f], o[f]), f
    for f in ["time_s", "distance_m", "x", "p"]:
        assert np.array_equal(g[f], o[f]), f


@pytest.mark.parametrize("which", ["config2", "legacy"])
def test_trace_fp64_bitexact(gpu_ctx, which):
    if which == "config2":
        P = CU.config2_params()
        P.max_steps = 20000
    else:
        P = CU.legacy_params(dphi=0.15)
        P.max_time_s = 60.0
    x, d, R = _random_tasks(3000, 3)
    print(f"trace case {which} n={len(R)}", flush=True)
    g, steps = CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP64)
    o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    assert steps == int(o["steps"].sum())
    assert len(np.unique(o["termination"])) >= 2


def test_trace_forced_failures_match_oracle(gpu_ctx):
    """Host-forced failure paths: zero step (InvalidTimeStep -> one retry with
    dt = 1e-12 -> step limit), a tiny step cap, a start inside the inner sphere."""
    x, d, R = _random_tasks(64, 9)
    P = CU.legacy_params(adaptive=False, dt_trace=0.0, max_steps=50)
    g, _ = CU.trace_batch(gpu_ctx, P, x, d, R)
    o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    assert np.all(o["retry_count"] == 1) and np.all(o["termination"] == 4)
    P = CU.legacy_params(max_steps=3)
    g, _ = CU.trace_batch(gpu_ctx, P, x, d, R)
    o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    assert np.all(o["termination"] == 4)
    xi = x * 0.5
    P = CU.legacy_params()
    g, _ = CU.trace_batch(gpu_ctx, P, xi, d, R)
    o = O.trace_batch(_oparams(P), xi, d, R)
    _same(g, o)
    assert np.all(o["termination"] == 1) and np.all(o["steps"] == 0)


def test_cutoff_scan_full_vs_upper_and_oracle(gpu_ctx):
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=8, nlat=6)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    og = O.rigidity_grid(0.1, 20.0, 48)
    assert np.array_equal(grid, og)
    sf, bf, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL)
    su, bu, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_UPPER)
    assert np.array_equal(bf, bu)
    traced = su >= 0
    assert np.array_equal(su[traced], sf[traced])
    # oracle on every (location, bin)
    o = O.trace_batch(_oparams(P), np.repeat(x0, 48, axis=0), np.repeat(dirs, 48, axis=0),
                      np.tile(grid, len(x0)))
    ost = (o["termination"] == 0).astype(np.int8).reshape(-1, 48)
    assert np.array_equal(sf, ost)
    ob = np.array([O.upper_cutoff_bin(r) for r in ost])
    assert np.array_equal(bf, ob)


def test_bracketed_full_scan_matches_full_scan(gpu_ctx):
    """Bracket-restricted FULL scan (the surrogate's scan): the bracket's
    samples equal the full scan's on the same bins, and the bin is the full
    scan's bin wherever the bracket holds the transition and nothing above it
    is forbidden; otherwise the ABOVE/BELOW fallback codes appear."""
    from amps_amd import surrogate as SG
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=8, nlat=8)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    sf, bf, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL)
    rng = np.random.default_rng(9)
    guess = np.where(bf < 0, 47, bf) + rng.integers(-4, 5, len(bf))
    lo, hi = SG.bracket_bins(np.clip(guess, 0, 47), 48)
    sb, bb, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, lo=lo, hi=hi)
    inb = np.zeros_like(sf, dtype=bool)
    for i in range(len(lo)):
        inb[i, lo[i]:hi[i] + 1] = True
    assert np.array_equal(sb[inb], sf[inb])
    assert np.all(sb[~inb] == -1)
    for i in range(len(lo)):
        s = sf[i]
        clean_above = np.all(s[hi[i]:] == 1)
        if bb[i] >= 0:
            if clean_above:
                assert bb[i] == bf[i], i
        elif bb[i] == CU.BIN_ABOVE:
            assert s[hi[i]] != 1 and hi[i] < 47
        elif bb[i] == CU.BIN_BELOW:
            assert np.all(s[lo[i]:hi[i] + 1] == 1) and lo[i] > 0
        else:
            assert bb[i] == -1 and hi[i] == 47 and s[47] != 1
    print(f"bracket results: {np.mean(bb >= 0):.3f} resolved, {np.mean(bb == bf):.3f} equal to the full scan")


@pytest.mark.parametrize("precision", [L.FP64, L.FP32])
def test_one_launch_bracketed_scan_matches_two_launches(gpu_ctx, precision):
    """AMPS_SCAN_BRACKETED: the bracket's bins, and for every pair whose
    bracket misses the transition the whole grid, traced by one persistent
    launch.  Its bins equal the two-launch form (bracket scan, then a full
    rescan of the missed pairs), which are the full scan's bins where the
    bracket holds the transition; the missed pairs' samples equal the full
    scan's on every bin, the others are traced only inside their bracket.
    Brackets around perturbed guesses make about a third of the pairs miss."""
    from amps_amd import surrogate as SG
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=8, nlat=8)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    sf, bf, st_full = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=precision)
    rng = np.random.default_rng(13)
    guess = np.clip(np.where(bf < 0, 47, bf) + rng.integers(-6, 7, len(bf)), 0, 47)
    lo, hi = SG.bracket_bins(guess, 48)
    print(f"one-launch bracketed scan, precision {precision}, {len(lo)} pairs", flush=True)
    s1, b1, st1 = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_BRACKETED, lo=lo, hi=hi,
                                 precision=precision)
    # the two-launch form
    _, b2, st2 = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, lo=lo, hi=hi, precision=precision)
    miss = b2 < -1
    b2 = b2.copy()
    if miss.any():
        _, b2m, st2m = CU.cutoff_scan(gpu_ctx, P, x0[miss], dirs[miss], grid, mode=CU.SCAN_FULL,
                                      precision=precision)
        b2[miss] = b2m
        st2 += st2m
    assert 0.1 < miss.mean() < 0.9
    assert np.array_equal(b1, b2)
    assert np.array_equal(b1[miss], bf[miss])
    assert np.array_equal(s1[miss], sf[miss])
    inb = np.zeros_like(sf, dtype=bool)
    for i in range(len(lo)):
        inb[i, lo[i]:hi[i] + 1] = True
    ok = ~miss
    assert np.array_equal(s1[ok][inb[ok]], sf[ok][inb[ok]]) and np.all(s1[ok][~inb[ok]] == -1)
    assert st1 < st2 and st1 < st_full
    print(f"missed brackets {miss.mean():.3f}; steps one launch {st1}, two launches {st2}, full {st_full}")


@pytest.mark.parametrize("half", [0, 24])
def test_bracketed_scan_extreme_widths(gpu_ctx, half):
    """One-bin brackets (most directions rescan the whole grid) and brackets
    that are the whole grid (nothing to rescan): the one-launch bracketed
    scan still gives the full scan's bins and samples."""
    from amps_amd import surrogate as SG
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=6, nlat=4)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    sf, bf, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL)
    guess = np.clip(np.where(bf < 0, 47, bf) + np.random.default_rng(3).integers(-2, 3, len(bf)), 0, 47)
    lo, hi = SG.bracket_bins(guess, 48, half=half)
    if half == 24:
        lo, hi = np.zeros_like(lo), np.full_like(hi, 47)
    print(f"bracketed scan with brackets of {hi[0] - lo[0] + 1} bins", flush=True)
    sb, bb, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_BRACKETED, lo=lo, hi=hi)
    assert np.array_equal(bb, bf)
    traced = sb >= 0
    assert np.array_equal(sb[traced], sf[traced])
    if half == 24:
        assert traced.all()


def test_refine_cutoff_matches_reference_bisection(gpu_ctx):
    """Batched bisection (FP64) vs the reference's RefineCutoffTransition
    compiled from CutoffBandSearch.h, driven by the oracle tracer: the same
    cutoffs and iteration counts, bit for bit."""
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=6, nlat=4)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    sf, bf, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL)
    sel = np.flatnonzero((bf >= 1) & (sf[np.arange(len(bf)), np.maximum(bf - 1, 0)] == 0))[:20]
    assert len(sel) >= 5
    lo, hi = grid[bf[sel] - 1], grid[bf[sel]]
    cut, it, stt, _ = CU.refine_cutoff(gpu_ctx, P, x0[sel], dirs[sel], lo, hi)
    Q = _oparams(P)

    def classify(i):
        def f(R):
            t = int(O.trace_batch(Q, x0[i:i + 1], dirs[i:i + 1], np.array([R]))["termination"][0])
            return 1 if t == 0 else (0 if t <= 5 else 2)
        return f

    for k, i in enumerate(sel):
        c, n_it, un = O.ref_refine_cutoff(lo[k], hi[k], classify(i))
        assert not un and stt[k] == 0
        assert cut[k] == c and it[k] == n_it, (k, cut[k], c, it[k], n_it)
    print(f"refined {len(sel)} upper cutoffs, {int(it.sum())} midpoints")


def test_penumbra_scan_matches_reference_chain(gpu_ctx):
    """PENUMBRA_SCAN (FP64) for penumbral directions: the GPU chain (full
    scan, band topology, batched refinement of both transitions, batched
    effective-cutoff bisection) against the reference's AnalyzeCutoffBand-
    Samples, RefineCutoffTransition and IntegrateEffectiveCutoff compiled from
    CutoffBandSearch.h and driven by the oracle tracer: identical lower,
    upper and effective cutoffs."""
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=8, nlat=6)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    res = CU.penumbra_scan(gpu_ctx, P, x0, dirs, 0.1, 20.0, 24)
    grid, states = res["grid"], res["states"]
    pen = np.flatnonzero((res["nTransitions"] > 1) & (res["effective_GV"] >= 0))
    assert len(pen) >= 3, int(np.sum(res["nTransitions"] > 1))
    Q = _oparams(P)

    def classify(i):
        def f(R):
            t = int(O.trace_batch(Q, x0[i:i + 1], dirs[i:i + 1], np.array([R]))["termination"][0])
            return 1 if t == 0 else (0 if t <= 5 else 2)
        return f

    r = O.ref()
    for i in pen[:8]:
        st = states[i].astype(np.int32)
        topo = np.zeros(15, np.int32)
        r.ref_band_topology(len(st), O._p(st, O.C.c_int), O._p(topo, O.C.c_int))
        assert list(topo[:7]) == [int(res[f][i]) for f in CU.TOPOLOGY_FIELDS[:7]]
        # CutoffForDirectionPenumbraScan_GV (CutoffRigidityGridless.cpp:2733-2750)
        lo = up = -1.0
        un_lo = un_up = False
        if topo[9]:
            lo = grid[0]
        elif topo[0] >= 0 and topo[1] >= 0:
            lo, _, un_lo = O.ref_refine_cutoff(grid[topo[0]], grid[topo[1]], classify(i))
        if topo[7]:
            up = grid[0]
        elif topo[2] >= 0 and topo[3] >= 0:
            up, _, un_up = O.ref_refine_cutoff(grid[topo[2]], grid[topo[3]], classify(i))
        assert not un_lo and not un_up
        assert res["lower_GV"][i] == lo and res["upper_GV"][i] == up
        eff, width, n_ref, un = O.ref_effective_cutoff(grid, st, lo, up, classify(i))
        assert not un and res["effective_GV"][i] == eff, (i, res["effective_GV"][i], eff)
    print(f"penumbra: {len(pen)} of {len(x0)} directions with more than one transition")


def test_fp32_tracer_agreement(gpu_ctx):
    """FP32 arithmetic with FP64 position accumulation vs the FP64 oracle:
    per-trajectory classification and per-direction cutoff bin.  Transitions
    in the penumbra are chaotic: the FP64 map itself moves by more than a bin
    at about 5 % of directions when the start points move by 1e-12 (printed
    here as the floor), so the FP32 bound is 6 %, the full map's measured
    rate, not the FP64 map's 0 %."""
    P = CU.config2_params()
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=12, nlat=8)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    s64, b64, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    s32, b32, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP32)
    _, bp, _ = CU.cutoff_scan(gpu_ctx, P, x0 * (1.0 + 1e-12), dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    agree = np.mean(s64 == s32)
    off = np.abs(b64.astype(int) - b32.astype(int))
    frac_gt1 = np.mean(off > 1)
    floor = np.mean(np.abs(b64.astype(int) - bp.astype(int)) > 1)
    print(f"FP32 vs FP64: state agreement {agree:.5f}, bins off by >1: {frac_gt1:.5f} "
          f"(FP64 with start points moved by 1e-12: {floor:.5f})")
    assert agree >= 0.97
    assert frac_gt1 <= 0.06


@pytest.mark.parametrize("mover", ["HC4", "RK2", "RK4", "RK6"])
def test_trace_movers_fp64_bitexact(gpu_ctx, mover):
    """Every full-orbit mover of StepParticleChecked through the persistent
    kernel, bit-exact with the oracle: the legacy deck (inner-sphere hits
    through stage and macro-step chords, outer-box exits, time limits) with
    starts spread from 1.05 to 3 Re."""
    P = CU.legacy_params(dphi=0.15, mover=mover)
    P.max_time_s = 20.0
    x, d, R = _random_tasks(1500, 7 + CU.MOVERS[mover], rhi=3.0)
    print(f"trace mover {mover} n={len(R)}", flush=True)
    g, steps = CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP64)
    o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    assert steps == int(o["steps"].sum())
    terms = set(np.unique(o["termination"]).tolist())
    assert {0, 1} <= terms, terms


def test_trace_mover_reaches_scan(gpu_ctx):
    """The mover reaches the scan kernel: an RK4 FULL scan of vertical
    directions on the 9000 km shell equals the oracle's traces of the same
    samples, state for state."""
    P = CU.legacy_params(dphi=0.05, mover="RK4")
    P.max_time_s = 60.0
    pts = np.array([TS.shell_point(0.0, lat, 9000.0, TS.RE_AMPS) for lat in (0.0, 20.0, 45.0)])
    dirs = -pts / np.linalg.norm(pts, axis=1)[:, None]
    grid = CU.build_search_grid(0.5, 20.0, 24)
    states, cut, _ = CU.cutoff_scan(gpu_ctx, P, pts, dirs, grid, mode=CU.SCAN_FULL)
    for i in range(len(pts)):
        o = O.trace_batch(_oparams(P), np.tile(pts[i], (len(grid), 1)), np.tile(dirs[i], (len(grid), 1)), grid)
        st = np.where(o["termination"] == 0, 1, np.where(o["termination"] <= 5, 0, 2))
        assert np.array_equal(states[i], st), i
    assert 0 < int((states == 1).sum()) < states.size


def test_trace_fp32_hc4_agreement(gpu_ctx):
    """FP32 HC4 against the FP64 HC4 path: classification agreement over a
    full scan of the batched-map parameterisation."""
    P = CU.config2_params()
    P.mover = CU.MOVERS["HC4"]
    P.max_steps = 20000
    pts, lat, lon = CU.config2_points(nlon=8, nlat=6)
    zen, az = CU.config2_directions(4)
    dirs = CU.arrival_dirs(pts, zen, az).reshape(-1, 3)
    x0 = np.repeat(pts, len(zen), axis=0)
    grid = CU.build_search_grid(0.1, 20.0, 48)
    s64, b64, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP64)
    s32, b32, _ = CU.cutoff_scan(gpu_ctx, P, x0, dirs, grid, mode=CU.SCAN_FULL, precision=L.FP32)
    agree = np.mean(s64 == s32)
    print(f"FP32 vs FP64 HC4: state agreement {agree:.5f}")
    assert agree >= 0.97


def test_trace_rejects_unbuilt_movers(gpu_ctx):
    P = CU.legacy_params()
    x, d, R = _random_tasks(4, 1)
    P.mover = 5  # GC2
    with pytest.raises(RuntimeError, match="mover"):
        CU.trace_batch(gpu_ctx, P, x, d, R)
    P.mover = CU.MOVERS["RK4"]
    with pytest.raises(RuntimeError, match="FP32"):
        CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP32)


def _same_exit(g, o):
    assert np.array_equal(g["valid"], o["valid"])
    for f in ["x_exit_m", "v_exit_unit", "cos_alpha"]:
        assert np.array_equal(g[f], o[f]), f


@pytest.mark.parametrize("mover,retry", [("BORIS", False), ("RK4", True)])
def test_structured_trace_and_exit_states_bitexact(gpu_ctx, mover, retry):
    """The density / transmission policy (TraceTrajectorySharedEx): the first
    boundary event on every chord, the exit state of the allowed traces and,
    with DS_RETRY_UNRESOLVED, the retry of every unresolved trace, bit-exact
    with the oracle (terminations, steps, time, path, x, p, exit x, direction
    and pitch cosine)."""
    P = CU.density_params(retry_unresolved=retry, mover=mover, tilt_deg=10.0, bz_T=-5e-9)
    P.max_time_s = 30.0
    if retry:
        P.max_steps = 3000  # step limits that the retry (2x steps, 2x time) revisits
    x, d, R = _random_tasks(1500, 23, rhi=4.0)
    print(f"structured trace case mover={mover} retry={retry} n={len(R)}", flush=True)
    g, ge, steps = CU.trace_batch_ex(gpu_ctx, P, x, d, R)
    o, oe = O.trace_batch_ex(_oparams(P), x, d, R)
    _same(g, o)
    _same_exit(ge, oe)
    # the step count covers every step taken, the discarded first attempts too
    first = 0
    rt = np.flatnonzero(o["retry_count"] == 1)
    if len(rt):
        P0 = _oparams(P)
        P0.retry = 0
        first = int(O.trace_batch(P0, x[rt], d[rt], R[rt])["steps"].sum())
    assert steps == int(o["steps"].sum()) + first
    assert (o["termination"] == 0).sum() > 100 and (o["termination"] == 1).sum() > 100
    if retry:
        assert len(rt) > 10


def test_transmission_matches_oracle_counts(gpu_ctx):
    """Isotropic transmission per (point, rigidity): the library's counts are
    those of the oracle's traces of the same tasks, reduced over directions."""
    P = CU.density_params(tilt_deg=10.0)
    P.max_time_s = 30.0
    pts = np.array([TS.shell_point(lon, lat, 500.0, TS.RE_AMPS) for lon, lat in [(0, 0), (90, 40), (200, -65)]])
    R = np.array([0.3, 1.0, 3.0, 10.0])
    zen, az = CU.config2_directions(24)
    dirs = np.array([CU.arrival_dirs(p[None], zen, az).reshape(-1, 3) for p in pts])
    # one direction set for all points: the local frame of the first point
    dset = dirs[0]
    print(f"transmission case pts={len(pts)} nR={len(R)} ndir={len(dset)}", flush=True)
    t = CU.transmission(gpu_ctx, P, pts, R, dset)
    x0 = np.repeat(pts, len(R) * len(dset), axis=0)
    Rt = np.tile(np.repeat(R, len(dset)), len(pts))
    dt_ = np.tile(dset, (len(pts) * len(R), 1))
    o, _ = O.trace_batch_ex(_oparams(P), x0, dt_, Rt)
    term = o["termination"].reshape(len(pts), len(R), len(dset))
    assert np.array_equal(t["sampled"], np.full((len(pts), len(R)), len(dset)))
    assert np.array_equal(t["resolved"], (term <= 2).sum(-1))
    assert np.array_equal(t["weight_sum"], (term == 0).sum(-1).astype(float))
    for k in range(9):
        assert np.array_equal(t["term_counts"][..., k], (term == k).sum(-1))
    assert t["T"].min() >= 0.0 and t["T"].max() <= 1.0 and t["T"].max() > 0.5 and t["T"].min() < 0.5


@pytest.mark.parametrize("mover,policy", [("BORIS", 0), ("HC4", 1)])
def test_trap_detector_traces_bitexact(gpu_ctx, mover, policy):
    """MagneticallyTrappedForbidden: mirroring orbits started at 2-4 Re with
    the trap detector on (TrajectoryTrap::Detector after every accepted
    step), bit-exact with the oracle under the legacy and the structured
    policy, with trapped, forbidden, allowed and limit terminations present."""
    P = CU.make_params(trap=True, mover=mover, policy=policy, max_steps=60000)
    rng = np.random.default_rng(5 + policy)
    n = 600
    x = rng.normal(size=(n, 3))
    x[:, 2] *= 0.3
    x *= (rng.uniform(2.0, 4.0, n) * CU.RE_DIPOLE / np.linalg.norm(x, axis=1))[:, None]
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    R = np.exp(rng.uniform(np.log(0.01), np.log(1.0), n))
    print(f"trap trace case mover={mover} policy={policy} n={n}", flush=True)
    if policy == 1:
        g, ge, steps = CU.trace_batch_ex(gpu_ctx, P, x, d, R)
        o, oe = O.trace_batch_ex(_oparams(P), x, d, R)
        _same_exit(ge, oe)
    else:
        g, steps = CU.trace_batch(gpu_ctx, P, x, d, R, precision=L.FP64)
        o = O.trace_batch(_oparams(P), x, d, R)
    _same(g, o)
    t = o["termination"]
    assert (t == 2).sum() > 100 and (t == 1).sum() > 10 and (t == 0).sum() > 5
