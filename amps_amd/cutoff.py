"""Host mirror of the gridless cutoff-rigidity interface over the GPU tracer.

* ``trace_batch``   — batched TraceTrajectoryShared / TraceAllowedShared
  (CutoffRigidityGridless.h:245-315, .cpp:1246-1520).
* ``cutoff_scan``   — CutoffForDirectionUpperScan_GV on the
  BuildCutoffSearchGrid_GV grid (.cpp:2578-2834), full or downward scan,
  optionally restricted to a per-direction bracket of bins.
* ``build_search_grid``, ``uniform_sphere_dirs``, ``local_dirs`` — the grid
  and direction helpers (.cpp:2578-2617, 1652-1674, 1757-1800).
* ``config2_*``     — the batched-map workload: a 64x32 lat/lon grid at
  r = 1.05 Re, 8 arrival directions per point, 48 log-spaced rigidities
  0.1..20 GV, tilted dipole (10 deg, B0 = 3.1e-5 T) plus Bz = -5 nT, steps of
  0.01 gyro-period, r < 1 forbidden, r > 25 allowed, 1e5 steps forbidden.

All tracing runs in the HIP library.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as L

SPEED_OF_LIGHT = 299792458.0      # constants.h:40
ELECTRON_CHARGE = 1.602176565e-19  # constants.h:31
AMU = 1.66053906660e-27            # constants.AtomicData.h:11
RE_AMPS = 6371.0e3                 # _EARTH__RADIUS_, constants.PlanetaryData.h:54
RE_DIPOLE = 6371.2e3               # Re_km of DipoleInterface.h:63-64
B_EQ = 3.12e-5                     # DipoleInterface.h:71
PROTON_AMU = 1.007276466621        # DipoleInterface.h:67

TERM_NAMES = ["OUTER_BOUNDARY_ALLOWED", "INNER_BOUNDARY_FORBIDDEN",
              "MAGNETICALLY_TRAPPED_FORBIDDEN", "TIME_LIMIT", "STEP_LIMIT", "DISTANCE_LIMIT",
              "INVALID_TIME_STEP", "INVALID_FIELD", "NUMERICAL_FAILURE"]

# MoverType of the full-orbit movers (GridlessParticleMovers.h:300-310); the
# FP32 tracer runs BORIS and HC4.
MOVERS = {"BORIS": 0, "HC4": 1, "RK2": 2, "RK4": 3, "RK6": 4}

SCAN_FULL = 0
SCAN_UPPER = 1
SCAN_BRACKETED = 2  # the bracket's bins, the whole grid for pairs it misses, one launch
BIN_ABOVE = -2
BIN_BELOW = -3


def dipole_moment(scale=1.0):
    """M_E_Am2 of DipoleInterface.h:71-72 times the moment scale."""
    Re_m = 6371.2 * 1000.0
    return scale * (B_EQ * Re_m * Re_m * Re_m / 1e-7)


def make_params(*, tilt_deg=0.0, moment_scale=1.0, bz_T=0.0, mass_amu=PROTON_AMU, charge_e=1,
                reversed_charge=False, dt_trace=1.0, max_time=600.0, max_steps=300000,
                max_distance_Re=400.0, dphi=0.15, adaptive=True, dist_limiter=True, dt_floor=True,
                outer_sphere=False, box_Re=35.0, r_outer_Re=35.0, r_inner_Re=1.0,
                earth_radius_m=RE_AMPS, retry=True, mover="BORIS", field_model=0, dt_selector=0,
                policy=0, event_tol_m=1.0, trap=False, trap_min_mirror=8, trap_min_bounce=4,
                trap_outer_margin_Re=1.0, trap_radial_tol_Re=0.05, trap_energy_tol=1.0e-4, trap_deadband=1.0e-6):
    P = L.TraceParams()
    if mover not in MOVERS:
        raise ValueError("mover must be one of %s (guiding-centre and hybrid movers are not built)"
                         % ", ".join(MOVERS))
    P.mover = MOVERS[mover]
    P.field_model = int(field_model)
    P.dt_selector = int(dt_selector)
    P.policy = int(policy)
    P.event_tol_m = float(event_tol_m)
    # TrajectoryTrap::Config (util/TrajectoryTrapDetector.h:13-21), parser defaults
    P.trap = int(trap)
    P.trap_min_mirror, P.trap_min_bounce = int(trap_min_mirror), int(trap_min_bounce)
    P.trap_outer_margin_Re, P.trap_radial_tol_Re = float(trap_outer_margin_Re), float(trap_radial_tol_Re)
    P.trap_energy_tol, P.trap_deadband = float(trap_energy_tol), float(trap_deadband)
    q = charge_e * ELECTRON_CHARGE
    P.q_C = -q if reversed_charge else q
    P.qabs_C = abs(q)
    P.m0_kg = mass_amu * AMU
    P.dt_trace_s = dt_trace
    P.max_time_s = max_time
    P.max_steps = int(max_steps)
    P.max_distance_m = max_distance_Re * earth_radius_m if max_distance_Re > 0 else -1.0
    P.dphi_max = dphi
    P.adaptive = int(adaptive)
    P.dist_limiter = int(dist_limiter)
    P.dt_floor = int(dt_floor)
    P.outer_sphere = int(outer_sphere)
    P.retry = int(retry)  # True/1: numerical failures; 2: every unresolved trace
    for d in range(3):
        P.box_min_Re[d] = -box_Re
        P.box_max_Re[d] = box_Re
    P.r_outer_Re = r_outer_Re
    P.r_inner_Re = r_inner_Re
    P.earth_radius_m = earth_radius_m
    P.moment_Am2 = dipole_moment(moment_scale)
    th = tilt_deg * math.pi / 180.0   # SetTiltDeg, DipoleInterface.cpp:22-28
    P.m_hat[0] = math.sin(th)
    P.m_hat[1] = 0.0
    P.m_hat[2] = math.cos(th)
    P.B_ext_T[0] = 0.0
    P.B_ext_T[1] = 0.0
    P.B_ext_T[2] = bz_T
    return P


def legacy_params(**kw):
    """The srcEarth cutoff decks' defaults (C1/C12): DT_TRACE 1 s, 600 s,
    400 Re path cap, box +-35 Re, R_INNER 1, legacy 0.15 rad step control."""
    return make_params(**kw)


def config2_params():
    return make_params(tilt_deg=10.0, moment_scale=3.1 / 3.12, bz_T=-5e-9, dt_trace=1e30,
                       max_time=1e30, max_steps=100000, max_distance_Re=-1.0,
                       dphi=2.0 * math.pi * 0.01, dist_limiter=False, dt_floor=False,
                       outer_sphere=True, r_outer_Re=25.0, r_inner_Re=1.0,
                       earth_radius_m=RE_DIPOLE)


def density_params(retry_unresolved=False, **kw):
    """TraceTrajectoryShared(Ex) as the density / spectrum solver calls it
    (TraceDensityDirection, DensityGridless.cpp:849-879): the StructuredAccurate
    policy (gyro budget 0.15 rad, no boundary limiter or floor, exact chord
    events with boundaryEventTolerance 1 m), no retry unless
    DS_RETRY_UNRESOLVED, which retries every unresolved trace once."""
    kw.setdefault("dphi", 0.15)
    return make_params(policy=1, retry=2 if retry_unresolved else 0, **kw)


def trace_batch_ex(ctx, params, x0, dirs, R, precision=L.FP64):
    """trace_batch with the exit states of the allowed structured traces."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(R, np.float64)
    out = np.zeros(len(R), L.TRACE_RESULT_DTYPE)
    ex = np.zeros(len(R), L.TRACE_EXIT_DTYPE)
    steps = C.c_int64()
    L.check(ctx.lib.amps_gpu_trace_batch_ex(ctx.h, C.byref(params), precision, len(R),
                                            L._ptr(x0, C.c_double), L._ptr(dirs, C.c_double),
                                            L._ptr(R, C.c_double), out.ctypes.data_as(C.c_void_p),
                                            ex.ctypes.data_as(C.c_void_p), C.byref(steps)), "trace_batch_ex")
    return out, ex, steps.value


def transmission(ctx, params, x0, R_GV, dirs, precision=L.FP64):
    """Isotropic transmission T = weight_sum / resolved per (point, rigidity)
    over the arrival directions dirs (ComputeT_atEnergy, DensityGridless.cpp:
    915-925).  Returns a dict of [npts, nR] arrays and the step count."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(R_GV, np.float64).ravel()
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    g = (len(x0), len(R))
    ws = np.zeros(g)
    ns, nres, nret = (np.zeros(g, np.int32) for _ in range(3))
    tc = np.zeros(g + (9,), np.int32)
    steps = C.c_int64()
    L.check(ctx.lib.amps_gpu_transmission(ctx.h, C.byref(params), precision, len(x0), L._ptr(x0, C.c_double),
                                          len(R), L._ptr(R, C.c_double), len(dirs), L._ptr(dirs, C.c_double),
                                          L._ptr(ws, C.c_double), L._ptr(ns, C.c_int32), L._ptr(nres, C.c_int32),
                                          L._ptr(nret, C.c_int32), L._ptr(tc, C.c_int32), C.byref(steps)),
            "transmission")
    with np.errstate(invalid="ignore", divide="ignore"):
        T = np.where(nres > 0, ws / np.maximum(nres, 1), 0.0)
    return {"T": T, "weight_sum": ws, "sampled": ns, "resolved": nres, "retried": nret,
            "term_counts": tc, "steps": steps.value}


def mode3d_params(**kw):
    """TraceAllowed3D on the Mode3D mesh field (CutoffRigidityMode3D.cpp:
    1048-1300): field_model 1 and the Mode3D legacy step selector (0.15 rad
    gyro budget, box and inner distances in metres, limiter and floor on)."""
    kw.setdefault("dphi", 0.15)
    return make_params(field_model=1, dt_selector=1, **kw)


class Mode3DField:
    """The compact Mode3D field on a block mesh, resident in the context:
    B of every leaf's interior cells, [nb][512][3] with cell i + 8(j + 8k)
    (GlobalMagneticField.h:18-30, block id = Temp_ID)."""

    def __init__(self, ctx, mesh, B_cells):
        self.ctx, self.mesh = ctx, mesh
        d = L.MeshDesc()
        for i in range(3):
            d.xmin[i] = mesh.xmin[i]
            d.xmax[i] = mesh.xmax[i]
        d.lmax = mesh.lmax
        d.nblocks = mesh.nblocks
        lv, ijk = mesh.desc_arrays()
        L.check(ctx.lib.amps_gpu_set_mesh(ctx.h, C.byref(d), L._ptr(lv, C.c_int32), L._ptr(ijk, C.c_int64)),
                "set_mesh")
        self.set(B_cells)

    def set(self, B_cells):
        Bm = np.ascontiguousarray(B_cells, np.float64)
        if Bm.shape != (self.mesh.nblocks, 512, 3):
            raise ValueError("B_cells must be [nblocks][512][3]")
        L.check(self.ctx.lib.amps_gpu_set_mode3d_field(self.ctx.h, L._ptr(Bm, C.c_double)), "set_mode3d_field")

    def at(self, x):
        """B [n][3] at points x [n][3] (cMode3DMeshFieldEval::GetB_T)."""
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
        B = np.zeros_like(x)
        L.check(self.ctx.lib.amps_gpu_mode3d_field(self.ctx.h, len(x), L._ptr(x, C.c_double),
                                                   L._ptr(B, C.c_double)), "mode3d_field")
        return B


def build_search_grid(rmin, rmax, n):
    """BuildCutoffSearchGrid_GV, LOG spacing with exact endpoints."""
    lmin, lmax = math.log(rmin), math.log(rmax)
    g = np.array([math.exp((1.0 - i / (n - 1)) * lmin + (i / (n - 1)) * lmax) for i in range(n)])
    g[0], g[-1] = rmin, rmax
    return g


def uniform_sphere_dirs(n):
    """BuildUniformSphereDirs (CutoffRigidityGridless.cpp:1652-1674): n
    near-uniform unit vectors, [n, 3] (library host code)."""
    out = np.zeros((int(n), 3))
    L.check(L.load().amps_gpu_uniform_sphere_dirs(int(n), L._ptr(out, C.c_double)), "uniform_sphere_dirs")
    return out


def local_dirs(x0, lon_deg, lat_deg):
    """BuildLocalENU_GSM + LocalLonLatToDir_GSM (CutoffRigidityGridless.cpp:
    1757-1800) through the library: for every location x0[p] and sky
    direction (lon_deg[k] from local north toward east, lat_deg[k] toward up),
    the unit vector dirs[p, k]; frames[p] = (E, N, U) rows."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    lon = np.ascontiguousarray(np.atleast_1d(lon_deg), np.float64)
    lat = np.ascontiguousarray(np.atleast_1d(lat_deg), np.float64)
    if lon.shape != lat.shape:
        raise ValueError("lon_deg and lat_deg must have the same length")
    dirs = np.zeros((len(x0), len(lon), 3))
    frames = np.zeros((len(x0), 3, 3))
    L.check(L.load().amps_gpu_local_dirs(len(x0), L._ptr(x0, C.c_double), len(lon), L._ptr(lon, C.c_double),
                                         L._ptr(lat, C.c_double), L._ptr(dirs, C.c_double),
                                         L._ptr(frames, C.c_double)), "local_dirs")
    return dirs, frames


def local_dirs_paired(x0, lon_deg, lat_deg):
    """local_dirs for (location, sky direction) pairs: dirs[p] for x0[p],
    lon_deg[p], lat_deg[p]."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    lon = np.ascontiguousarray(lon_deg, np.float64).ravel()
    lat = np.ascontiguousarray(lat_deg, np.float64).ravel()
    if not (len(x0) == len(lon) == len(lat)):
        raise ValueError("x0, lon_deg and lat_deg must have the same length")
    dirs = np.zeros((len(x0), 3))
    L.check(L.load().amps_gpu_local_dirs_paired(len(x0), L._ptr(x0, C.c_double), L._ptr(lon, C.c_double),
                                                L._ptr(lat, C.c_double), L._ptr(dirs, C.c_double)),
            "local_dirs_paired")
    return dirs


def local_enu(x0):
    """BuildLocalENU_GSM (CutoffRigidityGridless.cpp:1757-1773): (E, N, U)."""
    _, f = local_dirs(x0, [0.0], [0.0])
    return f[0, 0], f[0, 1], f[0, 2]


def config2_points(nlon=64, nlat=32, r_Re=1.05, re_m=RE_DIPOLE):
    lon = (np.arange(nlon) + 0.5) * 360.0 / nlon
    lat = -90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat
    LAT, LON = np.meshgrid(lat, lon, indexing="ij")
    la, lo = np.radians(LAT.ravel()), np.radians(LON.ravel())
    r = r_Re * re_m
    x = np.stack([r * np.cos(la) * np.cos(lo), r * np.cos(la) * np.sin(lo), r * np.sin(la)], 1)
    return x, LAT.ravel(), LON.ravel()


def config2_directions(ndir=8, seed=5, max_zenith_deg=60.0):
    """Fixed (zenith, azimuth) set: the first is vertical, the rest uniform in
    cos(zenith) over [0, max_zenith] and in azimuth, from a seeded RNG."""
    rng = np.random.default_rng(seed)
    zen = np.zeros(ndir)
    az = np.zeros(ndir)
    cz = rng.uniform(math.cos(math.radians(max_zenith_deg)), 1.0, ndir - 1)
    zen[1:] = np.degrees(np.arccos(cz))
    az[1:] = rng.uniform(0.0, 360.0, ndir - 1)
    return zen, az


def arrival_dirs(x0, zen, az):
    """Arrival direction (unit, pointing down along the particle's motion) for
    every point and every (zenith, azimuth): the backtrace starts along
    LocalLonLatToDir(azimuth, 90 - zenith)."""
    d, _ = local_dirs(x0, np.asarray(az, np.float64), 90.0 - np.asarray(zen, np.float64))
    return -d


def trace_batch(ctx, params, x0, dirs, R, precision=L.FP64):
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    R = np.ascontiguousarray(R, np.float64)
    out = np.zeros(len(R), L.TRACE_RESULT_DTYPE)
    steps = C.c_int64()
    L.check(ctx.lib.amps_gpu_trace_batch(ctx.h, C.byref(params), precision, len(R),
                                         L._ptr(x0, C.c_double), L._ptr(dirs, C.c_double),
                                         L._ptr(R, C.c_double), out.ctypes.data_as(C.c_void_p),
                                         C.byref(steps)), "trace_batch")
    return out, steps.value


def cutoff_scan(ctx, params, x0, dirs, grid, mode=SCAN_UPPER, lo=None, hi=None, precision=L.FP64):
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    grid = np.ascontiguousarray(grid, np.float64)
    n, nR = len(x0), len(grid)
    states = np.zeros((n, nR), np.int8)
    bins = np.zeros(n, np.int32)
    steps = C.c_int64()
    lo_a = None if lo is None else np.ascontiguousarray(lo, np.int32)
    hi_a = None if hi is None else np.ascontiguousarray(hi, np.int32)
    L.check(ctx.lib.amps_gpu_cutoff_scan(ctx.h, C.byref(params), precision, mode, n,
                                         L._ptr(x0, C.c_double), L._ptr(dirs, C.c_double), nR,
                                         L._ptr(grid, C.c_double), L._ptr(lo_a, C.c_int32),
                                         L._ptr(hi_a, C.c_int32), L._ptr(states, C.c_int8),
                                         L._ptr(bins, C.c_int32), C.byref(steps)), "cutoff_scan")
    return states, bins, steps.value


def refine_cutoff(ctx, params, x0, dirs, r_forbid, r_allow, abs_tol=1.0e-3, rel_tol=1.0e-6,
                  trace_limit_forbidden=True, precision=L.FP64):
    """RefineCutoffTransition (CutoffBandSearch.h:212-242) for many brackets:
    bisection of [r_forbid, r_allow] per (location, direction), one batched
    trace launch per round.  Returns (cutoff_GV, iterations, status, steps);
    status 1 = a trace limit ended the bracket (with trace_limit_forbidden
    off).  A numerical failure after the retry raises, as the reference's
    classifier throws."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    lo = np.ascontiguousarray(r_forbid, np.float64)
    hi = np.ascontiguousarray(r_allow, np.float64)
    n = len(lo)
    cut = np.zeros(n)
    it = np.zeros(n, np.int32)
    stt = np.zeros(n, np.int32)
    steps = C.c_int64()
    L.check(ctx.lib.amps_gpu_refine_cutoff(ctx.h, C.byref(params), precision, n, L._ptr(x0, C.c_double),
                                           L._ptr(dirs, C.c_double), L._ptr(lo, C.c_double),
                                           L._ptr(hi, C.c_double), abs_tol, rel_tol,
                                           1 if trace_limit_forbidden else 0, L._ptr(cut, C.c_double),
                                           L._ptr(it, C.c_int32), L._ptr(stt, C.c_int32),
                                           C.byref(steps)), "refine_cutoff")
    if np.any(stt == 2):
        raise L.AmpsGpuError(f"refine_cutoff: numerical failure after retry in {int(np.sum(stt == 2))} brackets")
    return cut, it, stt, steps.value


def effective_cutoff(ctx, params, x0, dirs, grid, states, lower, upper, abs_tol=1.0e-3, rel_tol=1.0e-6,
                     trace_limit_forbidden=True, precision=L.FP64):
    """IntegrateEffectiveCutoff (CutoffBandSearch.h:284-366) for many
    (location, direction) pairs: the penumbra's allowed width between the
    refined lower and upper cutoffs, each straddling interval of the scan
    grid bisected (batched).  Returns (effective_GV, allowed_width_GV,
    refined, status, steps); effective is -1 where the integration is
    unresolved (status 1).  A numerical failure after the retry raises, as
    the reference's classifier throws."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    grid = np.ascontiguousarray(grid, np.float64)
    states = np.ascontiguousarray(states, np.int8).reshape(len(x0), len(grid))
    lo = np.ascontiguousarray(lower, np.float64)
    up = np.ascontiguousarray(upper, np.float64)
    n = len(lo)
    eff = np.zeros(n)
    width = np.zeros(n)
    ref = np.zeros(n, np.int32)
    stt = np.zeros(n, np.int32)
    steps = C.c_int64()
    L.check(ctx.lib.amps_gpu_effective_cutoff(ctx.h, C.byref(params), precision, n, L._ptr(x0, C.c_double),
                                              L._ptr(dirs, C.c_double), len(grid), L._ptr(grid, C.c_double),
                                              L._ptr(states, C.c_int8), L._ptr(lo, C.c_double),
                                              L._ptr(up, C.c_double), abs_tol, rel_tol,
                                              1 if trace_limit_forbidden else 0, L._ptr(eff, C.c_double),
                                              L._ptr(width, C.c_double), L._ptr(ref, C.c_int32),
                                              L._ptr(stt, C.c_int32), C.byref(steps)), "effective_cutoff")
    if np.any(stt == 2):
        raise L.AmpsGpuError(f"effective_cutoff: numerical failure after retry in {int(np.sum(stt == 2))} pairs")
    return eff, width, ref, stt, steps.value


TOPOLOGY_FIELDS = ("lowerForbiddenIndex", "lowerAllowedIndex", "upperForbiddenIndex", "upperAllowedIndex",
                   "nTransitions", "nAllowedIntervals", "nUnresolved", "allAllowed", "allForbidden",
                   "lowerBelowRange", "lowerAboveRange", "upperBelowRange", "upperAboveRange",
                   "lowerBracketUnresolved", "upperBracketUnresolved")


def band_topology(states):
    """AnalyzeCutoffBandSamples (CutoffBandSearch.h:96-195) for each row of
    states (0 physically forbidden, 1 allowed, 2 unresolved): [n, 15] int32
    in TOPOLOGY_FIELDS order."""
    st = np.ascontiguousarray(states, np.int8)
    st2 = st.reshape(-1, st.shape[-1])
    out = np.zeros((len(st2), 15), np.int32)
    L.check(L.load().amps_gpu_band_topology(len(st2), st2.shape[1], L._ptr(st2, C.c_int8),
                                            L._ptr(out, C.c_int32)), "band_topology")
    return out


def penumbra_scan(ctx, params, x0, dirs, rmin_gv, rmax_gv, nscan=48, abs_tol=1.0e-3, rel_tol=1.0e-6,
                  precision=L.FP64):
    """The gridless PENUMBRA_SCAN cutoff for many (location, direction)
    pairs (CutoffForDirectionPenumbraScan_GV, CutoffRigidityGridless.cpp:
    2704-2772): the full scan of the search grid, the band topology, the
    lower and upper transitions refined by bisection, and the effective
    cutoff integrated over the penumbra, each stage batched over all pairs.
    Trace limits count as forbidden (traceLimitPolicy FORBIDDEN).  Returns a
    dict of per-pair arrays: lower_GV, upper_GV, effective_GV (-1 where not
    determined) and the topology counts and flags."""
    x0 = np.ascontiguousarray(x0, np.float64).reshape(-1, 3)
    dirs = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    n = len(x0)
    out = {"lower_GV": np.full(n, -1.0), "upper_GV": np.full(n, -1.0), "effective_GV": np.full(n, -1.0)}
    if rmax_gv < rmin_gv:
        return out
    grid = build_search_grid(rmin_gv, rmax_gv, nscan)
    states, _, steps = cutoff_scan(ctx, params, x0, dirs, grid, mode=SCAN_FULL, precision=precision)
    topo = band_topology(states)
    for k, f in enumerate(TOPOLOGY_FIELDS):
        out[f] = topo[:, k].copy()
    lfi, lai, ufi, uai = topo[:, 0], topo[:, 1], topo[:, 2], topo[:, 3]
    lower_unres = topo[:, 13].astype(bool)
    upper_unres = topo[:, 14].astype(bool)
    out["lower_GV"][topo[:, 9] == 1] = grid[0]
    sel = np.flatnonzero((topo[:, 9] == 0) & (lfi >= 0) & (lai >= 0))
    if len(sel):
        c, _, stt, s2 = refine_cutoff(ctx, params, x0[sel], dirs[sel], grid[lfi[sel]], grid[lai[sel]], abs_tol,
                                      rel_tol, precision=precision)
        out["lower_GV"][sel] = c
        lower_unres[sel[stt != 0]] = True
        steps += s2
    out["upper_GV"][topo[:, 7] == 1] = grid[0]
    sel = np.flatnonzero((topo[:, 7] == 0) & (ufi >= 0) & (uai >= 0))
    if len(sel):
        c, _, stt, s2 = refine_cutoff(ctx, params, x0[sel], dirs[sel], grid[ufi[sel]], grid[uai[sel]], abs_tol,
                                      rel_tol, precision=precision)
        out["upper_GV"][sel] = c
        upper_unres[sel[stt != 0]] = True
        steps += s2
    out["lowerBracketUnresolved"] = lower_unres.astype(np.int32)
    out["upperBracketUnresolved"] = upper_unres.astype(np.int32)
    lo, up = out["lower_GV"], out["upper_GV"]
    sel = np.flatnonzero((lo >= 0.0) & (up >= lo) & (topo[:, 6] == 0) & ~lower_unres & ~upper_unres)
    if len(sel):
        eff, _, _, stt, s3 = effective_cutoff(ctx, params, x0[sel], dirs[sel], grid, states[sel], lo[sel], up[sel],
                                              abs_tol, rel_tol, precision=precision)
        ok = stt == 0
        out["effective_GV"][sel[ok]] = eff[ok]
        steps += s3
    out["steps"] = steps
    out["grid"] = grid
    out["states"] = states
    return out


def set_trace_param(ctx, name, value):
    L.check(ctx.lib.amps_gpu_trace_set_param(ctx.h, name.encode(), int(value)), f"trace param {name}")
