"""Shared tracer settings of the tests: constants, the C1-style legacy deck,
the config-2 parameterisation, shell points and an oracle cutoff search
(UPPER_SCAN + bisection, CutoffRigidityGridless.cpp:2578-2834)."""
import math

import numpy as np

from oracle import oracle as O

C = 299792458.0                  # SpeedOfLight, constants.h:40
QE = 1.602176565e-19             # ElectronCharge, constants.h:31
AMU = 1.66053906660e-27          # _AMU_, constants.AtomicData.h:11
RE_AMPS = 6371.0e3               # _EARTH__RADIUS_, constants.PlanetaryData.h:54
RE_DIPOLE = 6371.2e3             # Re_km of DipoleInterface.h:63-64
B_EQ = 3.12e-5                   # DipoleInterface.h:71


def dipole_moment(scale=1.0):
    Re_m = 6371.2 * 1000.0
    return scale * (B_EQ * Re_m * Re_m * Re_m / 1e-7)


def params(tilt_deg=0.0, moment_scale=1.0, bz_T=0.0, mass_amu=1.0073, charge_e=1):
    P = O.TraceParams()
    P.q_C = charge_e * QE
    P.qabs_C = abs(charge_e * QE)
    P.m0_kg = mass_amu * AMU
    P.dt_trace_s = 1.0
    P.max_time_s = 600.0
    P.max_steps = 300000
    P.max_distance_m = 400.0 * RE_AMPS
    P.dphi_max = 0.15
    P.adaptive = 1
    P.dist_limiter = 1
    P.dt_floor = 1
    P.outer_sphere = 0
    P.retry = 1
    for d in range(3):
        P.box_min_Re[d] = -35.0
        P.box_max_Re[d] = 35.0
    P.r_outer_Re = 35.0
    P.r_inner_Re = 1.0
    P.earth_radius_m = RE_AMPS
    P.moment_Am2 = dipole_moment(moment_scale)
    th = tilt_deg * math.pi / 180.0
    P.m_hat[0] = math.sin(th)
    P.m_hat[1] = 0.0
    P.m_hat[2] = math.cos(th)
    P.B_ext_T[0] = 0.0
    P.B_ext_T[1] = 0.0
    P.B_ext_T[2] = bz_T
    return P


def c1_params():
    """srcEarth/test/C1/AMPS_PARAM_C1_gridless.in: DT_TRACE 1, MAX_TRACE_TIME
    600, MAX_TRACE_DISTANCE 400 Re, box +-35 Re, R_INNER 1, MASS_AMU 1.0073."""
    return params()


def c1_rigidity_range(emin=1.0, emax=20000.0, mass_amu=1.0073):
    m0 = mass_amu * AMU
    out = []
    for E in (emin, emax):
        EJ = E * 1e6 * QE
        mc2 = m0 * C * C
        p = math.sqrt(EJ * EJ + 2 * EJ * mc2) / C
        out.append(p * C / QE / 1e9)
    return out[0], out[1]


def shell_point(lon_deg, lat_deg, alt_km, re_m):
    r = re_m + alt_km * 1000.0
    lon, lat = math.radians(lon_deg), math.radians(lat_deg)
    return np.array([r * math.cos(lat) * math.cos(lon), r * math.cos(lat) * math.sin(lon), r * math.sin(lat)])


def oracle_cutoff(P, x0, dirv, rmin, rmax, nscan, rel_tol=None):
    """UPPER_SCAN bin + bisection; the bracket closes at max(1e-3 GV, 1e-6 R)
    as the reference's default, or at rel_tol * R when given."""
    grid = O.rigidity_grid(rmin, rmax, nscan)
    res = O.trace_batch(P, np.tile(x0, (nscan, 1)), np.tile(dirv, (nscan, 1)), grid)
    allowed = (res["termination"] == 0).astype(np.int32)
    b = O.upper_cutoff_bin(allowed)
    if b < 0:
        return -1.0
    if b == 0:
        return rmin
    lo, hi = grid[b - 1], grid[b]
    tol = max(1e-3, 1e-6 * max(abs(lo), abs(hi))) if rel_tol is None else rel_tol * hi
    while hi - lo > tol:
        mid = 0.5 * (lo + hi)
        r = O.trace_batch(P, x0[None], dirv[None], np.array([mid]))
        if r["termination"][0] == 0:
            hi = mid
        else:
            lo = mid
    return hi


def geospace_mesh(half_Re=16.0, base=3, r_refine_Re=5.0, re_m=RE_AMPS):
    """Block mesh over [-half, half]^3 Re in metres: level-`base` blocks, those
    within r_refine of the origin refined once (a coarse/fine shell around the
    Earth for the Mode3D field tests)."""
    from amps_amd import mesh as M
    L = half_Re * re_m
    m = M.uniform_mesh(base, xmin=(-L, -L, -L), xmax=(L, L, L))
    h = 2.0 * L / (1 << base)

    def pred(lv, i, j, k):
        lo = -L + np.array([i, j, k]) * h
        c = np.clip(0.0, lo, lo + h)
        return float(np.sum(c * c)) < (r_refine_Re * re_m) ** 2

    lv, ijk = M.refine(m.level, m.ijk, pred)
    return M.BlockMesh(lv, ijk, xmin=(-L, -L, -L), xmax=(L, L, L), lmax=m.lmax)


def dipole_cells(mesh, P, soft_Re=0.5):
    """The analytic dipole + uniform field of P sampled at every interior cell
    centre, [nb][512][3] (r clamped to soft_Re inside the Earth)."""
    from amps_amd import mesh as M
    Re = P.earth_radius_m

    def f(X):
        r = np.linalg.norm(X, axis=1, keepdims=True)
        Xs = np.where(r < soft_Re * Re, X * (soft_Re * Re / np.maximum(r, 1e-30)), X)
        r = np.linalg.norm(Xs, axis=1, keepdims=True)
        m = P.moment_Am2 * np.array(P.m_hat[:])
        mdr = Xs @ m
        B = 1e-7 * (3.0 * Xs * mdr[:, None] / r ** 5 - m[None, :] / r ** 3)
        return B + np.array(P.B_ext_T[:])[None, :]

    return M.interior_center_fields(mesh, f)
