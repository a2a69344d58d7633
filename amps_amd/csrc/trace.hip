// trace.hip — batched backward cutoff tracing in one persistent kernel.
//
// Each lane owns one trajectory at a time.  When a lane's trajectory ends it
// writes its result and takes the next work item from a device queue: the
// wavefront ballots the idle lanes, one lane reserves that many items with a
// single atomicAdd, and mbcnt ranks hand them out.  Trajectories of very
// different length (tens to 1e5 steps) therefore never leave lanes idle
// waiting for the longest trajectory of a static batch.
//
// FP64 path: TraceTrajectoryImpl + TraceTrajectoryWithSingleRetry
// (CutoffRigidityGridless.cpp:1246-1496) with BorisStep/StepParticleChecked
// (GridlessParticleMovers.cpp:781-801, 1035-1046) in SI units, expression for
// expression, so it reproduces the CPU restatement bit for bit.
// FP32 path: the same loop in units of (Re, Re/c, m0 c) with FP32 field,
// momentum and rotation arithmetic; positions, time and path length are
// accumulated in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "amps_common.h"
#include "center_device.h"
#include "trace_common.h"

using namespace amps;

namespace {

constexpr double C_LIGHT = 299792458.0;  // SpeedOfLight, constants.h:40

__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dot3(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ bool fin3(const double* a) {
  return isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]);
}

// ---------------------------------------------------------------------
// FP64 stepper (SI units)
// ---------------------------------------------------------------------
struct Fp64Lane {
  double x[3], p[3], t, s, dtTrace, maxTime;
  int64_t n, maxSteps;
  int retry;
};
// the lane of the structured policy and the trap detector
struct Fp64LaneX : Fp64Lane {
  // exit state of an allowed structured trace (x_exit, v_exit, cos alpha)
  double ex[7];
  int ex_valid;
  // trap detector (TrajectoryTrap::Detector, util/TrajectoryTrapDetector.h:23-153)
  int tr_init, tr_trapped, tr_sign, tr_mirror, tr_bounce, tr_stable, tr_prev;
  double tr_pref, tr_pmin, tr_pmax, tr_rmin, tr_rmax, tr_prmin, tr_prmax, tr_margin;
};

// Dipole::GetB_Tesla (DipoleInterface.h:99-125) plus a uniform field
__device__ __forceinline__ void dipole64(const amps_trace_params& P, const double* xm, double* B) {
  const double mu0_over_4pi = 1.0e-7;
  const double x = xm[0], y = xm[1], z = xm[2];
  const double r2 = x * x + y * y + z * z;
  if (r2 <= 0.0) {
    B[0] = B[1] = B[2] = 0.0;
  } else {
    const double r = sqrt(r2), r3 = r2 * r, r5 = r3 * r2;
    const double M = P.moment_Am2;
    const double mx = M * P.m_hat[0], my = M * P.m_hat[1], mz = M * P.m_hat[2];
    const double mdotr = mx * x + my * y + mz * z;
    B[0] = mu0_over_4pi * (3.0 * x * mdotr / r5 - mx / r3);
    B[1] = mu0_over_4pi * (3.0 * y * mdotr / r5 - my / r3);
    B[2] = mu0_over_4pi * (3.0 * z * mdotr / r5 - mz / r3);
  }
  B[0] = B[0] + P.B_ext_T[0];
  B[1] = B[1] + P.B_ext_T[1];
  B[2] = B[2] + P.B_ext_T[2];
}

__device__ __forceinline__ double gamma64(const double* p, double m0) {
  const double p2 = dot3(p, p);
  const double mc = m0 * C_LIGHT;
  return sqrt(1.0 + p2 / (mc * mc));
}

// the tracer's field: the analytic dipole plus the uniform field (FM 0) or
// the Mode3D mesh field of the compact cell-centred array (FM 1)
template <int FM>
__device__ __forceinline__ void field64(const TraceView& a, const double* x, double* B) {
  if constexpr (FM == 1)
    mode3d_B(a.mesh, a.Bm, x, B);
  else
    dipole64(a.P, x, B);
}

// BorisStep, GridlessParticleMovers.cpp:781-801
// Bx: B at x when the step selector has already evaluated it (have_B)
template <int FM>
__device__ __forceinline__ void boris64(const TraceView& a, double* x, double* p, double dt, const double* Bx,
                                        bool have_B) {
  const amps_trace_params& P = a.P;
  double B[3];
  if (have_B) B[0] = Bx[0], B[1] = Bx[1], B[2] = Bx[2];
  else field64<FM>(a, x, B);
  const double m0 = P.m0_kg, q = P.q_C;
  const double gam = gamma64(p, m0);
  const double iv = 1.0 / (gam * m0);
  const double vv[3] = {iv * p[0], iv * p[1], iv * p[2]};
  const double h = 0.5 * dt;
#pragma unroll
  for (int d = 0; d < 3; d++) x[d] = x[d] + h * vv[d];
  const double tc = (q * dt) / (2.0 * gam * m0);
  const double t[3] = {tc * B[0], tc * B[1], tc * B[2]};
  const double t2 = dot3(t, t);
  const double sc = 2.0 / (1.0 + t2);
  const double s[3] = {sc * t[0], sc * t[1], sc * t[2]};
  const double pm[3] = {p[0], p[1], p[2]};
  const double pp[3] = {pm[0] + (pm[1] * t[2] - pm[2] * t[1]), pm[1] + (pm[2] * t[0] - pm[0] * t[2]),
                        pm[2] + (pm[0] * t[1] - pm[1] * t[0])};
  p[0] = pm[0] + (pp[1] * s[2] - pp[2] * s[1]);
  p[1] = pm[1] + (pp[2] * s[0] - pp[0] * s[2]);
  p[2] = pm[2] + (pp[0] * s[1] - pp[1] * s[0]);
  const double gam2 = gamma64(p, m0);
  const double iv2 = 1.0 / (gam2 * m0);
  const double v2[3] = {iv2 * p[0], iv2 * p[1], iv2 * p[2]};
#pragma unroll
  for (int d = 0; d < 3; d++) x[d] = x[d] + h * v2[d];
}

__device__ __forceinline__ bool inside_inner_m(const double* x, double r) {
  return (r > 0.0) && (dot3(x, x) <= r * r);
}
// SegmentHitsInnerSphere, GridlessParticleMovers.cpp:136-149
__device__ __forceinline__ bool seg_hits(const double* a, const double* b, double r) {
  if (r <= 0.0) return false;
  if (inside_inner_m(a, r) || inside_inner_m(b, r)) return true;
  const double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double dd = dot3(d, d);
  if (dd <= 0.0) return false;
  double t = -dot3(a, d) / dd;
  if (t < 0.0) t = 0.0;
  if (t > 1.0) t = 1.0;
  const double q[3] = {a[0] + t * d[0], a[1] + t * d[1], a[2] + t * d[2]};
  return dot3(q, q) <= r * r;
}

// SelectLegacyCutoffTraceDt3D, CutoffRigidityMode3D.cpp:881-936: the same
// rules with the box and inner-sphere distances in metres (no box limit
// outside the box; inner distance max(0, |x| - r_inner)); limiter and floor
// always on.
template <int FM>
__device__ __forceinline__ double select_dt3d(const TraceView& a, const double* x, const double* p, double dtTrace,
                                              double timeRemaining, double* B, bool& have_B) {
  const amps_trace_params& P = a.P;
  double dt = dtTrace;
  if (timeRemaining < dt) dt = timeRemaining;
  if (!P.adaptive) return dt;
  const double m0 = P.m0_kg;
  const double p2 = dot3(p, p);
  const double mc = m0 * C_LIGHT;
  const double gamma = sqrt(1.0 + p2 / (mc * mc));
  const double pMag = sqrt(dmax(0.0, p2));
  const double vMag = (gamma > 0.0 && m0 > 0.0) ? pMag / (gamma * m0) : 0.0;
  field64<FM>(a, x, B);
  have_B = true;
  const double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    const double omega = fabs(P.q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P.dphi_max / omega);
  }
  const double dxp = a.box_max_m[0] - x[0], dxm = x[0] - a.box_min_m[0];
  const double dyp = a.box_max_m[1] - x[1], dym = x[1] - a.box_min_m[1];
  const double dzp = a.box_max_m[2] - x[2], dzm = x[2] - a.box_min_m[2];
  const double dBox = dmin(dmin(dmin(dmin(dmin(dxp, dxm), dyp), dym), dzp), dzm);
  const double dBoxSafe = (dBox < 0.0) ? 1.0e300 : dBox;
  const double dInner = dmax(0.0, sqrt(dot3(x, x)) - a.r_inner_m);
  const double dNear = dmin(dBoxSafe, dInner);
  const double vFor = vMag;
  if (vFor > 0.0 && dNear < 1.0e299) dt = dmin(dt, 0.20 * dNear / vFor);
  const double safeSpeed = dmax(vFor, 1.0e-12);
  const double dtFloor = dmax(dmax(1.0e-12, 1.0e-9 * dmax(dtTrace, 1.0)), 100.0e3 / safeSpeed);
  dt = dmax(dtFloor, dt);
  if (timeRemaining > 0.0) dt = dmin(dt, timeRemaining);
  return dt;
}

// Full-orbit movers of StepParticleChecked (MoverType numbering).
constexpr int MV_BORIS = 0, MV_HC4 = 1, MV_RK2 = 2, MV_RK4 = 3, MV_RK6 = 4;

// HigueraCaryMagneticMidpointStep, GridlessParticleMovers.cpp:723-757: the
// Boris rotation with B sampled at the half-drift midpoint.
template <int FM>
__device__ __forceinline__ void hc_mid64(const TraceView& a, double* x, double* p, double dt) {
  const amps_trace_params& P = a.P;
  const double m0 = P.m0_kg, q = P.q_C;
  const double g0 = gamma64(p, m0);
  const double iv0 = 1.0 / (g0 * m0);
  const double h = 0.5 * dt;
  const double xm[3] = {x[0] + h * (iv0 * p[0]), x[1] + h * (iv0 * p[1]), x[2] + h * (iv0 * p[2])};
  double B[3];
  field64<FM>(a, xm, B);
  const double tc = (q * dt) / (2.0 * g0 * m0);
  const double t[3] = {tc * B[0], tc * B[1], tc * B[2]};
  const double t2 = dot3(t, t);
  const double sc = 2.0 / (1.0 + t2);
  const double s[3] = {sc * t[0], sc * t[1], sc * t[2]};
  const double pm[3] = {p[0], p[1], p[2]};
  const double pp[3] = {pm[0] + (pm[1] * t[2] - pm[2] * t[1]), pm[1] + (pm[2] * t[0] - pm[0] * t[2]),
                        pm[2] + (pm[0] * t[1] - pm[1] * t[0])};
  p[0] = pm[0] + (pp[1] * s[2] - pp[2] * s[1]);
  p[1] = pm[1] + (pp[2] * s[0] - pp[0] * s[2]);
  p[2] = pm[2] + (pp[0] * s[1] - pp[1] * s[0]);
  const double iv1 = 1.0 / (gamma64(p, m0) * m0);
#pragma unroll
  for (int d = 0; d < 3; d++) x[d] = xm[d] + h * (iv1 * p[d]);
}

// Yoshida triple-jump weights of HC4StepUnchecked (GridlessParticleMovers.cpp:
// 759-779): w1 = 1/(2 - 2^(1/3)), w0 = -2^(1/3)/(2 - 2^(1/3)).
constexpr double HC4_CBRT2 = 1.25992104989487316477;
constexpr double HC4_W1 = 1.0 / (2.0 - HC4_CBRT2);
constexpr double HC4_W0 = -HC4_CBRT2 / (2.0 - HC4_CBRT2);

template <int FM>
__device__ __forceinline__ void hc4_64(const TraceView& a, double* x, double* p, double dt) {
  hc_mid64<FM>(a, x, p, HC4_W1 * dt);
  hc_mid64<FM>(a, x, p, HC4_W0 * dt);
  hc_mid64<FM>(a, x, p, HC4_W1 * dt);
}

// Butcher tables of the RK2 / RK4 / RK6 cases, GridlessParticleMovers.cpp:
// 1061-1098 (midpoint, classical, Butcher's seven-stage sixth order).
template <int MV>
struct RKTable;
template <>
struct RKTable<MV_RK2> {
  static constexpr int NS = 2;
  static constexpr double a[2][2] = {{0.0, 0.0}, {0.5, 0.0}};
  static constexpr double b[2] = {0.0, 1.0};
};
template <>
struct RKTable<MV_RK4> {
  static constexpr int NS = 4;
  static constexpr double a[4][4] = {
      {0.0, 0.0, 0.0, 0.0}, {0.5, 0.0, 0.0, 0.0}, {0.0, 0.5, 0.0, 0.0}, {0.0, 0.0, 1.0, 0.0}};
  static constexpr double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
};
template <>
struct RKTable<MV_RK6> {
  static constexpr int NS = 7;
  static constexpr double a[7][7] = {{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
                                     {1.0 / 3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0},
                                     {0.0, 2.0 / 3.0, 0.0, 0.0, 0.0, 0.0, 0.0},
                                     {1.0 / 12.0, 1.0 / 3.0, -1.0 / 12.0, 0.0, 0.0, 0.0, 0.0},
                                     {-1.0 / 16.0, 9.0 / 8.0, -3.0 / 16.0, -3.0 / 8.0, 0.0, 0.0, 0.0},
                                     {0.0, 9.0 / 8.0, -3.0 / 8.0, -3.0 / 4.0, 0.5, 0.0, 0.0},
                                     {9.0 / 44.0, -9.0 / 11.0, 63.0 / 44.0, 18.0 / 11.0, 0.0, -16.0 / 11.0, 0.0}};
  static constexpr double b[7] = {11.0 / 120.0, 0.0, 27.0 / 40.0, 27.0 / 40.0, -4.0 / 15.0, -4.0 / 15.0, 11.0 / 120.0};
};

// RKStepGeneric, GridlessParticleMovers.cpp:655-692.  Every product of the
// table is formed, zeros included, so that the sums round as the reference's.
// A stage chord meeting the sphere leaves x at the stage point and p at its
// start value; the final chord leaves both at y1.
template <int MV, int FM>
__device__ __forceinline__ bool rk_checked64(const TraceView& a, double* x, double* p, double dt, double rIn,
                                             const double* Bx, bool have_B) {
  const amps_trace_params& P = a.P;
  using T = RKTable<MV>;
  double kx[T::NS][3], kp[T::NS][3];
  const double y0x[3] = {x[0], x[1], x[2]}, y0p[3] = {p[0], p[1], p[2]};
  double prev[3] = {x[0], x[1], x[2]};
#pragma unroll
  for (int i = 0; i < T::NS; i++) {
    double yx[3] = {y0x[0], y0x[1], y0x[2]}, yp[3] = {y0p[0], y0p[1], y0p[2]};
#pragma unroll
    for (int j = 0; j < i; j++) {
      const double c = dt * T::a[i][j];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        yx[d] = yx[d] + c * kx[j][d];
        yp[d] = yp[d] + c * kp[j][d];
      }
    }
    if (seg_hits(prev, yx, rIn)) {
#pragma unroll
      for (int d = 0; d < 3; d++) x[d] = yx[d];
      return false;
    }
    if (inside_inner_m(yx, rIn)) {
#pragma unroll
      for (int d = 0; d < 3; d++) x[d] = yx[d], p[d] = yp[d];
      return false;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) prev[d] = yx[d];
    // LorentzRhs, GridlessParticleMovers.cpp:151-157
    // the first stage sits at x itself (no table terms), where B may be known
    double B[3];
    if (i == 0 && have_B) B[0] = Bx[0], B[1] = Bx[1], B[2] = Bx[2];
    else field64<FM>(a, yx, B);
    const double iv = 1.0 / (gamma64(yp, P.m0_kg) * P.m0_kg);
    const double v[3] = {iv * yp[0], iv * yp[1], iv * yp[2]};
    kx[i][0] = v[0], kx[i][1] = v[1], kx[i][2] = v[2];
    kp[i][0] = P.q_C * (v[1] * B[2] - v[2] * B[1]);
    kp[i][1] = P.q_C * (v[2] * B[0] - v[0] * B[2]);
    kp[i][2] = P.q_C * (v[0] * B[1] - v[1] * B[0]);
  }
  double y1x[3] = {y0x[0], y0x[1], y0x[2]}, y1p[3] = {y0p[0], y0p[1], y0p[2]};
#pragma unroll
  for (int i = 0; i < T::NS; i++) {
    const double c = dt * T::b[i];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      y1x[d] = y1x[d] + c * kx[i][d];
      y1p[d] = y1p[d] + c * kp[i][d];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) x[d] = y1x[d], p[d] = y1p[d];
  return !seg_hits(prev, y1x, rIn) && !inside_inner_m(y1x, rIn);
}

// StepParticleChecked, GridlessParticleMovers.cpp:1035-1098 (full-orbit
// movers): false when the step met the inner sphere.
template <int MV, int FM>
__device__ __forceinline__ bool step_checked64(const TraceView& a, double* x, double* p, double dt, double rIn,
                                               const double* Bx, bool have_B) {
  if constexpr (MV == MV_BORIS) {
    const double x0[3] = {x[0], x[1], x[2]};
    boris64<FM>(a, x, p, dt, Bx, have_B);
    return !seg_hits(x0, x, rIn) && !inside_inner_m(x, rIn);
  } else if constexpr (MV == MV_HC4) {
    // four positive macro-steps, each chord checked; the composition's
    // negative inner stage is never classified
    const double dtSub = dt / 4.0;
    for (int i = 0; i < 4; i++) {
      const double x0[3] = {x[0], x[1], x[2]};
      hc4_64<FM>(a, x, p, dtSub);
      if (seg_hits(x0, x, rIn) || inside_inner_m(x, rIn)) return false;
    }
    return true;
  } else {
    return rk_checked64<MV, FM>(a, x, p, dt, rIn, Bx, have_B);
  }
}

// SelectLegacyCutoffTraceDt, CutoffRigidityGridless.cpp:1086-1152 (switchable parts)
template <int FM>
__device__ __forceinline__ double select_dt64(const TraceView& a, const double* x, const double* p, double dtTrace,
                                              double timeRemaining, double* B, bool& have_B) {
  const amps_trace_params& P = a.P;
  double dt = dtTrace;
  if (timeRemaining < dt) dt = timeRemaining;
  if (!P.adaptive) return dt;
  const double m0 = P.m0_kg;
  const double p2 = dot3(p, p);
  const double mc = m0 * C_LIGHT;
  const double gamma = sqrt(1.0 + p2 / (mc * mc));
  const double pMag = sqrt(dmax(0.0, p2));
  const double vMag = (gamma > 0.0 && m0 > 0.0) ? pMag / (gamma * m0) : 0.0;
  field64<FM>(a, x, B);
  have_B = true;
  const double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    const double omega = fabs(P.q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P.dphi_max / omega);
  }
  const double vFor = vMag;
  if (P.dist_limiter) {
    const double Re = P.earth_radius_m;
    const double xRe[3] = {x[0] / Re, x[1] / Re, x[2] / Re};
    const double rRe = sqrt(xRe[0] * xRe[0] + xRe[1] * xRe[1] + xRe[2] * xRe[2]);
    double dInner = 1.0e300;
    if (rRe > P.r_inner_Re) dInner = (rRe - P.r_inner_Re) * Re;
    double dBox = 1.0e300;
    if (P.outer_sphere) {
      if (rRe < P.r_outer_Re) dBox = (P.r_outer_Re - rRe) * Re;
    } else if (xRe[0] >= P.box_min_Re[0] && xRe[0] <= P.box_max_Re[0] && xRe[1] >= P.box_min_Re[1] &&
               xRe[1] <= P.box_max_Re[1] && xRe[2] >= P.box_min_Re[2] && xRe[2] <= P.box_max_Re[2]) {
      const double dxp = (P.box_max_Re[0] - xRe[0]) * Re, dxm = (xRe[0] - P.box_min_Re[0]) * Re;
      const double dyp = (P.box_max_Re[1] - xRe[1]) * Re, dym = (xRe[1] - P.box_min_Re[1]) * Re;
      const double dzp = (P.box_max_Re[2] - xRe[2]) * Re, dzm = (xRe[2] - P.box_min_Re[2]) * Re;
      dBox = dmin(dmin(dmin(dxp, dxm), dmin(dyp, dym)), dmin(dzp, dzm));
    }
    const double dNear = dmin(dInner, dBox);
    if (vFor > 0.0 && dNear < 1.0e299) dt = dmin(dt, 0.20 * dNear / vFor);
  }
  if (P.dt_floor) {
    const double safeSpeed = dmax(vFor, 1.0e-12);
    const double dtFloor = dmax(dmax(1.0e-12, 1.0e-9 * dmax(dtTrace, 1.0)), 100.0e3 / safeSpeed);
    dt = dmax(dtFloor, dt);
  }
  if (timeRemaining > 0.0) dt = dmin(dt, timeRemaining);
  return dt;
}

// SelectStructuredTraceDt (CutoffRigidityGridless.cpp:1154-1228) with
// FinalizeUpperBoundedStep (util/TrajectoryTimeStep.h:15-24): DT_TRACE capped
// by the gyro-angle budget; no boundary limiter, no floor.
__device__ __forceinline__ double finalize_upper_bounded(double dt, double trem) {
  if (!(dt > 0.0) || !isfinite(dt)) return __builtin_nan("");
  if (trem > 0.0) dt = dmin(dt, trem);
  if (!(dt > 0.0) || !isfinite(dt)) return __builtin_nan("");
  return dt;
}
template <int FM>
__device__ __forceinline__ double select_dt_structured(const TraceView& a, const double* x, const double* p,
                                                       double dtTrace, double trem, double* B, bool& have_B) {
  const amps_trace_params& P = a.P;
  double dt = dtTrace;
  if (!P.adaptive) return finalize_upper_bounded(dt, trem);
  const double m0 = P.m0_kg;
  const double p2 = dot3(p, p);
  const double mc = m0 * C_LIGHT;
  const double gamma = sqrt(1.0 + p2 / (mc * mc));
  field64<FM>(a, x, B);
  have_B = true;
  const double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    const double omega = fabs(P.q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P.dphi_max / omega);
  }
  return finalize_upper_bounded(dt, trem);
}

// TrajectoryBoundary::FindFirstEvent (util/TrajectoryBoundary.h:46-141): the
// first inner-sphere (type 1) or outer-box (type 2) crossing on the chord.
__device__ __forceinline__ void find_first_event(const double* a, const double* z, const double* bmin,
                                                 const double* bmax, double rInner, double tol, int& type,
                                                 double& frac, double* pos) {
  type = 0;
  frac = __builtin_inf();
  if (!fin3(a) || !fin3(z)) return;
  const double d[3] = {z[0] - a[0], z[1] - a[1], z[2] - a[2]};
  if (rInner > 0.0) {
    const double r0 = dmax(0.0, rInner + 0.0);
    if (a[0] * a[0] + a[1] * a[1] + a[2] * a[2] <= r0 * r0) {
      type = 1, frac = 0.0;
      pos[0] = a[0], pos[1] = a[1], pos[2] = a[2];
    } else {
      const double aa = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (aa > 0.0) {
        const double segmentLength = sqrt(aa);
        const double fractionTol = (segmentLength > 0.0) ? dmax(0.0, tol) / segmentLength : 0.0;
        const double bb = 2.0 * (a[0] * d[0] + a[1] * d[1] + a[2] * d[2]);
        const double cc = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] - rInner * rInner;
        double disc = bb * bb - 4.0 * aa * cc;
        if (!(disc < 0.0)) {
          disc = dmax(0.0, disc);
          const double root = sqrt(disc);
          const double cand[2] = {(-bb - root) / (2.0 * aa), (-bb + root) / (2.0 * aa)};
#pragma unroll
          for (int c = 0; c < 2; c++) {
            if (cand[c] < -fractionTol || cand[c] > 1.0 + fractionTol) continue;
            const double t = dmax(0.0, dmin(1.0, cand[c]));
            if (t >= frac) continue;
            type = 1, frac = t;
#pragma unroll
            for (int i = 0; i < 3; i++) pos[i] = a[i] + frac * d[i];
          }
        }
      }
    }
  }
  const bool ain = a[0] >= bmin[0] && a[0] <= bmax[0] && a[1] >= bmin[1] && a[1] <= bmax[1] && a[2] >= bmin[2] &&
                   a[2] <= bmax[2];
  if (!ain) {
    if (0.0 < frac) {
      type = 2, frac = 0.0;
      pos[0] = a[0], pos[1] = a[1], pos[2] = a[2];
    }
    return;
  }
  const bool zin = z[0] >= bmin[0] && z[0] <= bmax[0] && z[1] >= bmin[1] && z[1] <= bmax[1] && z[2] >= bmin[2] &&
                   z[2] <= bmax[2];
  if (zin) return;
  const double segmentLength = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  const double fractionTol = (segmentLength > 0.0) ? dmax(0.0, tol) / segmentLength : 0.0;
  for (int axis = 0; axis < 3; axis++) {
    if (d[axis] == 0.0) continue;
    const int side = (d[axis] > 0.0) ? 1 : 0;
    const double plane = side ? bmax[axis] : bmin[axis];
    const double cand = (plane - a[axis]) / d[axis];
    if (cand < -fractionTol || cand > 1.0 + fractionTol) continue;
    const double t = dmax(0.0, dmin(1.0, cand));
    if (t >= frac) continue;
    const double q[3] = {a[0] + t * d[0], a[1] + t * d[1], a[2] + t * d[2]};
    bool onFace = true;
    for (int j = 0; j < 3; j++) {
      if (j == axis) continue;
      if (q[j] < bmin[j] - tol || q[j] > bmax[j] + tol) {
        onFace = false;
        break;
      }
    }
    if (!onFace) continue;
    type = 2, frac = t;
    pos[0] = q[0], pos[1] = q[1], pos[2] = q[2];
  }
}

// PopulateOuterExit (CutoffRigidityGridless.cpp:1323-1358) into the lane:
// the event position, the direction of the momentum interpolated to the
// event, cos(pitch angle) against B a tolerance inside the box.  False for an
// invalid field or momentum (InvalidField).
template <int FM>
__device__ __forceinline__ bool exit_state(const TraceView& a, Fp64LaneX& L, const double* pos, double fraction,
                                           const double* xb, const double* xa, const double* pb, const double* pa) {
  const amps_trace_params& P = a.P;
#pragma unroll
  for (int d = 0; d < 3; d++) L.ex[d] = pos[d];
  const double f = dmax(0.0, dmin(1.0, fraction));
  const double pc[3] = {pb[0] + f * (pa[0] - pb[0]), pb[1] + f * (pa[1] - pb[1]), pb[2] + f * (pa[2] - pb[2])};
  const double p2n = dot3(pc, pc);
  const double mc = P.m0_kg * C_LIGHT;
  const double g = sqrt(1.0 + p2n / (mc * mc));
  if (!(g > 0.0) || !isfinite(g)) return false;
  const double iv = 1.0 / (g * P.m0_kg);
  double v[3] = {iv * pc[0], iv * pc[1], iv * pc[2]};
  const double nv = sqrt(dot3(v, v));
  if (nv > 0.0) {
    const double in = 1.0 / nv;
#pragma unroll
    for (int d = 0; d < 3; d++) v[d] = in * v[d];
  } else {
    v[0] = v[1] = v[2] = 0.0;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) L.ex[3 + d] = v[d];
  const double ch[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
  const double chl = sqrt(dot3(ch, ch));
  double inset = 0.0;
  if (chl > 0.0) inset = dmin(f, dmax(1.0, P.event_tol_m) / chl);
  const double fe = dmax(0.0, f - inset);
  const double xe[3] = {xb[0] + fe * ch[0], xb[1] + fe * ch[1], xb[2] + fe * ch[2]};
  double B[3];
  field64<FM>(a, xe, B);
  const double Bn = sqrt(dot3(B, B));
  if (!(Bn > 0.0) || !isfinite(Bn)) return false;
  const double ib = 1.0 / Bn;
  const double bh[3] = {ib * B[0], ib * B[1], ib * B[2]};
  L.ex[6] = dot3(v, bh);
  L.ex_valid = isfinite(L.ex[6]) ? 1 : 0;
  return L.ex_valid != 0;
}

// TrajectoryTrap::Detector::Update (util/TrajectoryTrapDetector.h:36-103):
// mirror points are sign changes of p.B beyond the deadband, two per bounce
// cycle; the orbit is trapped once enough cycles have a stable radial
// envelope, stay the margin inside the box and keep |p| to the tolerance.
__device__ __forceinline__ void trap_reset(Fp64LaneX& L) {
  L.tr_init = L.tr_trapped = L.tr_sign = L.tr_mirror = L.tr_bounce = L.tr_stable = L.tr_prev = 0;
  L.tr_pref = L.tr_pmax = L.tr_rmax = L.tr_prmin = L.tr_prmax = 0.0;
  L.tr_pmin = L.tr_rmin = L.tr_margin = __builtin_inf();
}
__device__ __forceinline__ double trap_margin(const TraceView& a, const double* x) {
  return dmin(dmin(dmin(x[0] - a.box_min_m[0], a.box_max_m[0] - x[0]), dmin(x[1] - a.box_min_m[1], a.box_max_m[1] - x[1])),
              dmin(x[2] - a.box_min_m[2], a.box_max_m[2] - x[2]));
}
__device__ __forceinline__ bool trap_update(const TraceView& a, Fp64LaneX& L, const double* x, const double* p,
                                            const double* B) {
  const amps_trace_params& P = a.P;
  if (L.tr_trapped) return true;
  if (!fin3(x) || !fin3(p) || !fin3(B)) return false;
  const double pMag = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double bMag = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (!(pMag > 0.0) || !(bMag > 0.0) || !isfinite(r)) return false;
  if (!L.tr_init) {
    L.tr_init = 1;
    L.tr_pref = pMag, L.tr_pmin = pMag, L.tr_pmax = pMag;
    L.tr_rmin = r, L.tr_rmax = r;
    L.tr_margin = dmin(L.tr_margin, trap_margin(a, x));
  } else {
    L.tr_pmin = dmin(L.tr_pmin, pMag);
    L.tr_pmax = dmax(L.tr_pmax, pMag);
    L.tr_rmin = dmin(L.tr_rmin, r);
    L.tr_rmax = dmax(L.tr_rmax, r);
    L.tr_margin = dmin(L.tr_margin, trap_margin(a, x));
  }
  const double mu = (p[0] * B[0] + p[1] * B[1] + p[2] * B[2]) / (pMag * bMag);
  const int sign = (mu > P.trap_deadband) ? 1 : ((mu < -P.trap_deadband) ? -1 : 0);
  if (sign == 0) return false;
  if (L.tr_sign == 0) {
    L.tr_sign = sign;
    return false;
  }
  if (sign == L.tr_sign) return false;
  L.tr_sign = sign;
  ++L.tr_mirror;
  if ((L.tr_mirror % 2) != 0) return false;
  ++L.tr_bounce;
  if (L.tr_prev) {
    const double ch = dmax(fabs(L.tr_rmin - L.tr_prmin), fabs(L.tr_rmax - L.tr_prmax));
    if (ch <= P.trap_radial_tol_Re * P.earth_radius_m) ++L.tr_stable;
    else L.tr_stable = 0;
  }
  L.tr_prmin = L.tr_rmin;
  L.tr_prmax = L.tr_rmax;
  L.tr_prev = 1;
  const double spread = (L.tr_pref > 0.0) ? (L.tr_pmax - L.tr_pmin) / L.tr_pref : __builtin_inf();
  L.tr_trapped = (L.tr_mirror >= P.trap_min_mirror && L.tr_bounce >= P.trap_min_bounce && L.tr_stable >= 2 &&
                  L.tr_margin >= P.trap_outer_margin_Re * P.earth_radius_m && spread <= P.trap_energy_tol)
                     ? 1
                     : 0;
  L.tr_rmin = r, L.tr_rmax = r;
  L.tr_margin = dmin(L.tr_margin, trap_margin(a, x));
  return L.tr_trapped != 0;
}

// X: the structured policy, exit states and the trap detector are compiled
// in (the legacy cutoff kernels stay lean: twice the wavefronts per SIMD)
template <int MV, int FM, bool X>
struct Fp64Stepper {
  using Lane = std::conditional_t<X, Fp64LaneX, Fp64Lane>;
  static constexpr bool HAS_EXIT = X;
  static __device__ __forceinline__ void init(const TraceView& a, Lane& L, const double* x0,
                                              const double* dir, double R, int attempt) {
    const amps_trace_params& P = a.P;
    const double pMag = (R * 1.0e9 * P.qabs_C) / C_LIGHT;
    double v0[3] = {-1.0 * dir[0], -1.0 * dir[1], -1.0 * dir[2]};
    if (X && P.policy == 1) {
      // TraceTrajectoryShared(Ex) (CutoffRigidityGridless.cpp:1563-1585)
      // normalise the start direction
      const double nv = sqrt(dot3(v0, v0));
      if (nv > 0.0) {
        const double in = 1.0 / nv;
#pragma unroll
        for (int d = 0; d < 3; d++) v0[d] = in * v0[d];
      } else {
        v0[0] = v0[1] = v0[2] = 0.0;
      }
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
      L.p[d] = pMag * v0[d];
      L.x[d] = x0[d];
    }
    L.t = 0.0;
    L.s = 0.0;
    L.n = 0;
    L.retry = attempt;
    if constexpr (X) {
      L.ex_valid = 0;
      trap_reset(L);
      if (P.trap) {
        double B0[3];
        field64<FM>(a, L.x, B0);
        trap_update(a, L, L.x, L.p, B0);
      }
    }
    if (attempt == 0) {
      L.dtTrace = P.dt_trace_s;
      L.maxSteps = P.max_steps;
      L.maxTime = P.max_time_s;
    } else {
      L.dtTrace = dmax(1.0e-12, 0.5 * P.dt_trace_s);
      L.maxSteps = (P.max_steps <= (int64_t)2147483647 / 2) ? 2 * P.max_steps : (int64_t)2147483647;
      L.maxTime = 2.0 * P.max_time_s;
    }
  }
  // one iteration of the trace loop; returns a termination code or -1
  static __device__ __forceinline__ int step(const TraceView& a, Lane& L) {
    const amps_trace_params& P = a.P;
    if (!(L.n < L.maxSteps && L.t < L.maxTime && (P.max_distance_m <= 0.0 || L.s < P.max_distance_m))) {
      if (L.n >= L.maxSteps) return 4;
      if (P.max_distance_m > 0.0 && L.s >= P.max_distance_m) return 5;
      if (L.t >= L.maxTime) return 3;
      return 8;
    }
    double* x = L.x;
    double* p = L.p;
    if (!fin3(x) || !fin3(p)) return 8;
    const double rIn = a.r_inner_m;
    if (rIn > 0.0 && x[0] * x[0] + x[1] * x[1] + x[2] * x[2] <= dmax(0.0, rIn + 0.0) * dmax(0.0, rIn + 0.0))
      return 1;
    bool inside;
    if (P.outer_sphere)
      inside = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] <= a.r_outer_m * a.r_outer_m);
    else
      inside = x[0] >= a.box_min_m[0] - 0.0 && x[0] <= a.box_max_m[0] + 0.0 && x[1] >= a.box_min_m[1] - 0.0 &&
               x[1] <= a.box_max_m[1] + 0.0 && x[2] >= a.box_min_m[2] - 0.0 && x[2] <= a.box_max_m[2] + 0.0;
    if (!inside) {
      // structured policy: an OuterBox event at the step start (fraction 0
      // on a zero chord)
      if constexpr (X)
        if (P.policy == 1 && a.exits) return exit_state<FM>(a, L, x, 0.0, x, x, p, p) ? 0 : 7;
      return 0;
    }
    // B at the start point: evaluated once by the adaptive step selector and
    // reused by the step (the same function of the same point)
    double Bx[3];
    bool have_B = false;
    double dt;
    if (X && P.policy == 1)
      dt = select_dt_structured<FM>(a, x, p, L.dtTrace, L.maxTime - L.t, Bx, have_B);
    else if (P.dt_selector == 1)
      dt = select_dt3d<FM>(a, x, p, L.dtTrace, L.maxTime - L.t, Bx, have_B);
    else
      dt = select_dt64<FM>(a, x, p, L.dtTrace, L.maxTime - L.t, Bx, have_B);
    if (!(dt > 0.0) || !isfinite(dt)) return 6;
    const double xb[3] = {x[0], x[1], x[2]};
    const double pb[3] = {p[0], p[1], p[2]};
    if (!step_checked64<MV, FM>(a, x, p, dt, rIn, Bx, have_B)) return 1;
    if (!fin3(x) || !fin3(p)) return 8;
    int etype = 0;
    double efrac = 0.0, epos[3] = {0.0, 0.0, 0.0};
    if constexpr (X)
      if (P.policy == 1) find_first_event(xb, x, a.box_min_m, a.box_max_m, rIn, P.event_tol_m, etype, efrac, epos);
    const double dd[3] = {x[0] - xb[0], x[1] - xb[1], x[2] - xb[2]};
    const double seg = sqrt(dot3(dd, dd));
    L.s += isfinite(seg) ? seg : 0.0;
    L.t += dt;
    ++L.n;
    if constexpr (X) {
      if (etype == 1) return 1;
      if (etype == 2) return (!a.exits || exit_state<FM>(a, L, epos, efrac, xb, x, pb, p)) ? 0 : 7;
      if (P.trap) {
        double Bt[3];
        field64<FM>(a, x, Bt);
        if (trap_update(a, L, x, p, Bt)) return 2;  // MagneticallyTrappedForbidden
      }
    }
    return -1;
  }
  static __device__ __forceinline__ void result(const TraceView& a, const Lane& L, int term,
                                                amps_trace_result& r) {
    r.termination = term;
    r.retry_count = L.retry;
    r.steps = L.n;
    r.time_s = L.t;
    r.distance_m = L.s;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      r.x[d] = L.x[d];
      r.p[d] = L.p[d];
    }
  }
  static __device__ __forceinline__ void exit_out(const Lane& L, int term, amps_trace_exit& e) {
    static_assert(X, "exit states come from the extended lane");
    const bool v = term == 0 && L.ex_valid;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      e.x_exit_m[d] = v ? L.ex[d] : 0.0;
      e.v_exit_unit[d] = v ? L.ex[3 + d] : 0.0;
    }
    e.cos_alpha = v ? L.ex[6] : 0.0;
    e.valid = v ? 1 : 0;
    e.pad_ = 0;
  }
};

// ---------------------------------------------------------------------
// FP32 stepper: units of length Re, time Re/c, momentum m0 c.  FP32 field,
// momentum and rotation; FP64 position/time/path accumulation.
// ---------------------------------------------------------------------
struct Fp32Lane {
  double x[3];
  float u[3];
  double t, s, dtTrace, maxTime;
  int64_t n, maxSteps;
  int retry;
};

__device__ __forceinline__ void bfield32(const TraceView& a, const double* xd, float* b) {
#pragma clang fp contract(fast)
  const float x = (float)xd[0], y = (float)xd[1], z = (float)xd[2];
  const float r2 = x * x + y * y + z * z;
  const float ir = rsqrtf(r2);
  const float ir2 = ir * ir;
  const float ir3 = ir2 * ir;
  const float mdr = a.f_mhat[0] * x + a.f_mhat[1] * y + a.f_mhat[2] * z;
  const float k5 = 3.0f * mdr * ir3 * ir2;
  b[0] = a.f_cb * (k5 * x - a.f_mhat[0] * ir3) + a.f_bext[0];
  b[1] = a.f_cb * (k5 * y - a.f_mhat[1] * ir3) + a.f_bext[1];
  b[2] = a.f_cb * (k5 * z - a.f_mhat[2] * ir3) + a.f_bext[2];
}

// Boris rotation of u by t = tc b, in place (FP32)
__device__ __forceinline__ void rotate32(float* u, const float* b, float tc) {
#pragma clang fp contract(fast)
  const float t[3] = {tc * b[0], tc * b[1], tc * b[2]};
  const float t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
  const float sc = 2.0f * __builtin_amdgcn_rcpf(1.0f + t2);
  const float s[3] = {sc * t[0], sc * t[1], sc * t[2]};
  const float pp[3] = {u[0] + (u[1] * t[2] - u[2] * t[1]), u[1] + (u[2] * t[0] - u[0] * t[2]),
                       u[2] + (u[0] * t[1] - u[1] * t[0])};
  const float un[3] = {u[0] + (pp[1] * s[2] - pp[2] * s[1]), u[1] + (pp[2] * s[0] - pp[0] * s[2]),
                       u[2] + (pp[0] * s[1] - pp[1] * s[0])};
  u[0] = un[0];
  u[1] = un[1];
  u[2] = un[2];
}

// MV: MV_BORIS or MV_HC4 (the FP32 path has no RK movers)
template <int MV>
struct Fp32Stepper {
  using Lane = Fp32Lane;
  static constexpr bool HAS_EXIT = false;
  static __device__ __forceinline__ void init(const TraceView& a, Lane& L, const double* x0,
                                              const double* dir, double R, int attempt) {
    const amps_trace_params& P = a.P;
    // |u| = p/(m0 c) with p = R*1e9*|q|/c
    const double umag = (R * 1.0e9 * P.qabs_C) / C_LIGHT / (P.m0_kg * C_LIGHT);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      L.u[d] = (float)(umag * (-1.0 * dir[d]));
      L.x[d] = x0[d] / a.len;
    }
    L.t = 0.0;
    L.s = 0.0;
    L.n = 0;
    L.retry = attempt;
    const double T = a.len / C_LIGHT;
    if (attempt == 0) {
      L.dtTrace = P.dt_trace_s / T;
      L.maxSteps = P.max_steps;
      L.maxTime = P.max_time_s / T;
    } else {
      L.dtTrace = dmax(1.0e-12, 0.5 * P.dt_trace_s) / T;
      L.maxSteps = (P.max_steps <= (int64_t)2147483647 / 2) ? 2 * P.max_steps : (int64_t)2147483647;
      L.maxTime = 2.0 * P.max_time_s / T;
    }
  }
  static __device__ __forceinline__ int step(const TraceView& a, Lane& L) {
#pragma clang fp contract(fast)
    const amps_trace_params& P = a.P;
    if (!(L.n < L.maxSteps && L.t < L.maxTime && (a.maxdist_n <= 0.0 || L.s < a.maxdist_n))) {
      if (L.n >= L.maxSteps) return 4;
      if (a.maxdist_n > 0.0 && L.s >= a.maxdist_n) return 5;
      if (L.t >= L.maxTime) return 3;
      return 8;
    }
    double* x = L.x;
    float* u = L.u;
    if (!fin3(x) || !isfinite(u[0]) || !isfinite(u[1]) || !isfinite(u[2])) return 8;
    const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    if (a.rin_n > 0.0 && r2 <= a.rin_n * a.rin_n) return 1;
    bool inside;
    if (P.outer_sphere)
      inside = r2 <= a.rout_n * a.rout_n;
    else
      inside = x[0] >= P.box_min_Re[0] && x[0] <= P.box_max_Re[0] && x[1] >= P.box_min_Re[1] &&
               x[1] <= P.box_max_Re[1] && x[2] >= P.box_min_Re[2] && x[2] <= P.box_max_Re[2];
    if (!inside) return 0;
    // time step (normalized)
    const float u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float igam = rsqrtf(1.0f + u2);
    float b[3];
    bfield32(a, x, b);
    double dt = L.dtTrace;
    const double trem = L.maxTime - L.t;
    if (trem < dt) dt = trem;
    if (P.adaptive) {
      const float bm2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
      if (bm2 > 0.0f) {
        // dphi / omega with omega = |B| / gamma, gamma = (1 + u^2) / sqrt(1 + u^2)
        dt = dmin(dt, (double)(a.f_dphi * ((1.0f + u2) * igam) * rsqrtf(bm2)));
      }
      if (P.dist_limiter || P.dt_floor) {
        const float vmag = sqrtf(u2) * igam;
        if (P.dist_limiter) {
          const double r = sqrt(r2);
          double dIn = 1.0e300, dOut = 1.0e300;
          if (r > a.rin_n) dIn = r - a.rin_n;
          if (P.outer_sphere) {
            if (r < a.rout_n) dOut = a.rout_n - r;
          } else if (inside) {
            dOut = dmin(dmin(dmin(P.box_max_Re[0] - x[0], x[0] - P.box_min_Re[0]),
                             dmin(P.box_max_Re[1] - x[1], x[1] - P.box_min_Re[1])),
                        dmin(P.box_max_Re[2] - x[2], x[2] - P.box_min_Re[2]));
          }
          const double dn = dmin(dIn, dOut);
          if (vmag > 0.0f && dn < 1.0e299) dt = dmin(dt, 0.20 * dn / (double)vmag);
        }
        if (P.dt_floor) {
          const double sv = dmax((double)vmag, 1.0e-12 / C_LIGHT);
          const double fl = dmax(dmax(a.floor_abs_n, a.floor_rel_n), a.floor_dist_n / sv);
          dt = dmax(fl, dt);
        }
      }
      if (trem > 0.0) dt = dmin(dt, trem);
    }
    if (!(dt > 0.0) || !isfinite(dt)) return 6;
    const float dtf = (float)dt;
    const double xb[3] = {x[0], x[1], x[2]};
    if constexpr (MV == MV_HC4) {
      // four positive macro-steps of the triple jump H2(w1) H2(w0) H2(w1),
      // H2 = half drift, B at the midpoint, rotation, half drift; each
      // macro-step chord is checked against the inner sphere
      const float dq = 0.25f * dtf;
      float ig = igam;
      for (int m = 0; m < 4; m++) {
        const double xm0[3] = {x[0], x[1], x[2]};
        const float rm = sqrtf((float)(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]));
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const float dk = (k == 1 ? (float)HC4_W0 : (float)HC4_W1) * dq;
          const float hk = 0.5f * dk;
#pragma unroll
          for (int d = 0; d < 3; d++) x[d] += (double)(hk * (u[d] * ig));
          float bm[3];
          bfield32(a, x, bm);
          rotate32(u, bm, dk * 0.5f * ig);
          ig = rsqrtf(1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
#pragma unroll
          for (int d = 0; d < 3; d++) x[d] += (double)(hk * (u[d] * ig));
        }
        const float dm[3] = {(float)(x[0] - xm0[0]), (float)(x[1] - xm0[1]), (float)(x[2] - xm0[2])};
        const float segm = sqrtf(dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2]);
        if (a.rin_n > 0.0 && rm - segm <= (float)a.rin_n * 1.0001f)
          if (seg_hits(xm0, x, a.rin_n) || inside_inner_m(x, a.rin_n)) return 1;
      }
    } else {
      // drift - kick - drift (BorisStep with B sampled at the start point)
      const float hf = 0.5f * dtf;
#pragma unroll
      for (int d = 0; d < 3; d++) x[d] += (double)(hf * (u[d] * igam));
      rotate32(u, b, dtf * 0.5f * igam);
      const float igam2 = rsqrtf(1.0f + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
#pragma unroll
      for (int d = 0; d < 3; d++) x[d] += (double)(hf * (u[d] * igam2));
    }
    const float dd[3] = {(float)(x[0] - xb[0]), (float)(x[1] - xb[1]), (float)(x[2] - xb[2])};
    const float seg = sqrtf(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2]);
    // the chord can reach the inner sphere only if it starts within its
    // length of the sphere; the exact FP64 test runs only then
    if constexpr (MV == MV_BORIS) {
      if (a.rin_n > 0.0 && sqrtf((float)r2) - seg <= (float)a.rin_n * 1.0001f)
        if (seg_hits(xb, x, a.rin_n) || inside_inner_m(x, a.rin_n)) return 1;
    }
    if (!fin3(x) || !isfinite(u[0]) || !isfinite(u[1]) || !isfinite(u[2])) return 8;
    L.s += isfinite(seg) ? (double)seg : 0.0;
    L.t += dt;
    ++L.n;
    return -1;
  }
  static __device__ __forceinline__ void result(const TraceView& a, const Lane& L, int term,
                                                amps_trace_result& r) {
    const double T = a.len / C_LIGHT;
    const double pu = a.P.m0_kg * C_LIGHT;
    r.termination = term;
    r.retry_count = L.retry;
    r.steps = L.n;
    r.time_s = L.t * T;
    r.distance_m = L.s * a.len;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      r.x[d] = L.x[d] * a.len;
      r.p[d] = (double)L.u[d] * pu;
    }
  }
};

// retry 1: numerical failures (TraceTrajectoryWithSingleRetry); retry 2:
// every unresolved termination (TraceDensityDirection, DensityGridless.cpp:849-879)
__device__ __forceinline__ bool retryable(int mode, int t) { return mode == 2 ? t > 2 : (t >= 6 && t <= 8); }
__device__ __forceinline__ int8_t state_of(int term) {
  // 1 allowed; 0 cutoff-forbidden (physical or trace limit); 2 numerical failure
  return term == 0 ? 1 : (term <= 5 ? 0 : 2);
}

// BRACKET mode phase of a location, read and published across workgroups.
// A lane that reads it only decides whether to trace read-only inputs, so the
// read needs no ordering (an acquire would invalidate the caches at every poll)
__device__ __forceinline__ int load_status(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// The cutoff bin of one location from its samples s[lo..hi]
// (CutoffForDirectionUpperScan_GV read off a sample vector): the allowed side
// of the highest forbidden -> allowed transition; -1 when the grid's top is
// forbidden.  Returns false when [lo, hi] does not hold the transition (top
// forbidden below the grid's top, or allowed down to a bottom above Rmin).
__device__ __forceinline__ bool bin_from_samples(const int8_t* s, int lo, int hi, int nR, int& b) {
  if (s[hi] != 1) {
    b = -1;
    return hi == nR - 1;
  }
  b = lo;
  for (int k = hi - 1; k >= lo; --k)
    if (s[k] != 1) {
      b = k + 1;
      break;
    }
  return !(b == lo && lo > 0);
}

// ---------------------------------------------------------------------
// Persistent kernel.  Work item semantics by mode:
//   TRACE: item i = trajectory i of an explicit (x0, dir, R) list
//   FULL : item i = (loc i / nbin, bin lo(loc) + i % nbin)   [states only]
//   UPPER: item i = loc i, bins walked downward from hi(loc) until the first
//          forbidden sample (CutoffForDirectionUpperScan_GV semantics)
//   BRACKET: items [0, nloc W) are FULL over each location's bracket of W
//          bins (the bins below its centre first); items [nloc W, nloc nR)
//          are the location's other bins, each traced only if the bracket
//          misses the transition.  The whole-grid rescan (phase 2) opens as
//          soon as the bracket's top bin comes out not allowed below the
//          grid's top, or, when the last bracket bin is done, if every bin
//          was allowed above Rmin; otherwise that last lane writes the
//          cutoff bin (phase 1: the other bins are skipped).  pend counts a
//          location's outstanding items (the rescan's are added before it
//          opens); the lane that retires the last one reads the bins off.  A
//          lane that draws an other-bin item of a location still in phase 0
//          holds it until the phase is known: every bracket item was drawn
//          before any other-bin item and is being traced, so every wait ends.
// ---------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(256) void k_trace_persistent(TraceView a) {
  typename S::Lane L;
  int64_t item = -1;
  int loc = 0, bin = 0, attempt = 0;
  bool has = false, done = false, wait = false;
  int64_t my_steps = 0;
  const int lane = __lane_id();
  const bool BRK = a.mode == TRACE_MODE_BRACKET;
  const int64_t nloc_full = (a.mode == TRACE_MODE_FULL) ? a.nitems / a.nbin_item : (BRK ? a.nloc : 1);
  for (;;) {
    if (BRK && wait) {
      const int st = load_status(a.status + loc);
      if (st == 2) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // after the rescan's count
        wait = false;
        S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
      } else if (st == 1) {
        wait = false;
        has = false;
      }
    }
    // draws; in BRACKET mode an item of a resolved location is skipped and
    // its lane draws again (a few rounds per pass)
    for (int round = 0; round < (BRK ? 4 : 1); round++) {
      const bool need = !has && !done;
      const uint64_t mask = __ballot(need);
      if (!mask) break;
      const int leader = __ffsll((long long)mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(a.queue, (unsigned long long)__popcll(mask));
      base = __shfl(base, leader);
      if (need) {
        const uint64_t below = (lane == 0) ? 0ull : (mask & ((1ull << lane) - 1ull));
        const unsigned long long my = base + (unsigned long long)__popcll(below);
        if ((int64_t)my < a.nitems) {
          item = (int64_t)my;
          attempt = 0;
          has = true;
          if (a.mode == TRACE_MODE_LIST) {
            loc = (int)item;
            S::init(a, L, a.x0 + 3 * item, a.dir + 3 * item, a.R[item], 0);
          } else if (a.mode == TRACE_MODE_FULL) {
            // bin-major: every location's lowest rigidities first, the
            // longest trajectories (quasi-trapped, up to the step cap)
            // therefore start first
            loc = (int)(item % nloc_full);
            bin = (a.lo ? a.lo[loc] : 0) + (int)(item / nloc_full);
            S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
          } else if (BRK) {
            loc = (int)(item % nloc_full);
            const int kk = (int)(item / nloc_full), lo = a.lo[loc];
            if (kk < a.nbin_item) {
              // the bins just below the bracket's centre first: the predicted
              // cutoff's forbidden side, where the quasi-trapped (longest)
              // trajectories are, then the centre and the allowed side
              const int half = a.nbin_item / 2;
              bin = lo + (kk < half ? half - 1 - kk : kk);
              S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
            } else {
              const int k2 = kk - a.nbin_item;
              bin = k2 < lo ? k2 : k2 + a.nbin_item;
              const int st = load_status(a.status + loc);
              if (st == 2) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
              } else if (st == 0)
                wait = true;
              else
                has = false;
            }
          } else {
            loc = (int)item;
            bin = a.hi ? a.hi[loc] : a.nR - 1;
            S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
          }
        } else {
          done = true;
        }
      }
    }
    if (__ballot(has) == 0ull) break;
    // a wavefront whose lanes all wait for bracket results yields its issue slots
    if (BRK && __ballot(has && !wait) == 0ull) __builtin_amdgcn_s_sleep(64);
    for (int k = 0; k < a.chunk && has && !wait; k++) {
      const int term = S::step(a, L);
      if (term < 0) continue;
      if (a.P.retry && attempt == 0 && retryable(a.P.retry, term)) {
        attempt = 1;
        my_steps += L.n;
        if (a.mode == TRACE_MODE_LIST)
          S::init(a, L, a.x0 + 3 * item, a.dir + 3 * item, a.R[item], 1);
        else
          S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 1);
        continue;
      }
      my_steps += L.n;
      if (a.mode == TRACE_MODE_LIST) {
        amps_trace_result r;
        S::result(a, L, term, r);
        a.out[item] = r;
        if constexpr (S::HAS_EXIT) {
          if (a.exits) {
            amps_trace_exit e;
            S::exit_out(L, term, e);
            a.exits[item] = e;
          }
        }
        has = false;
      } else if (a.mode == TRACE_MODE_FULL) {
        a.states[(int64_t)loc * a.nR + bin] = state_of(term);
        has = false;
      } else if (BRK) {
        int8_t* s = a.states + (int64_t)loc * a.nR;
        const int8_t sv = state_of(term);
        s[bin] = sv;
        has = false;
        const int top = a.lo[loc] + a.nbin_item - 1;
        if (bin == top && sv != 1 && top < a.nR - 1 && load_status(a.status + loc) == 0) {
          // the bracket's top is not allowed below the grid's top: the
          // transition lies above the bracket, whatever its other bins give,
          // so the rescan starts now (counted before this bin is retired)
          atomicAdd(a.pend + loc, a.nR - a.nbin_item);
          store_status(a.status + loc, 2);
        }
        // the sample (and any rescan count and phase) before this item retires
        __threadfence();
        if (atomicSub(a.pend + loc, 1) == 1) {
          // this location's phase is complete: every sample of it is visible
          __threadfence();
          const int phase = load_status(a.status + loc);
          const int lo = phase == 0 ? a.lo[loc] : 0;
          const int hi = phase == 0 ? lo + a.nbin_item - 1 : a.nR - 1;
          int b;
          if (bin_from_samples(s, lo, hi, a.nR, b)) {
            a.cutoff[loc] = b;
            if (phase == 0) store_status(a.status + loc, 1);
          } else {
            // allowed down to a bottom above Rmin: the other bins, then the whole grid
            __hip_atomic_store(a.pend + loc, a.nR - a.nbin_item, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            store_status(a.status + loc, 2);
          }
        }
      } else {
        const int8_t st = state_of(term);
        a.states[(int64_t)loc * a.nR + bin] = st;
        const int lo = a.lo ? a.lo[loc] : 0;
        if (st == 1 && bin > lo) {
          bin--;
          attempt = 0;
          S::init(a, L, a.x0 + 3 * (int64_t)loc, a.dir + 3 * (int64_t)loc, a.grid[bin], 0);
        } else {
          // first non-allowed sample from the top, or the bracket's bottom
          a.cutoff[loc] = (st == 1) ? bin : bin + 1;
          has = false;
        }
      }
    }
  }
  // total steps of this wavefront
  unsigned long long tot = (unsigned long long)my_steps;
  for (int off = 32; off > 0; off >>= 1) tot += __shfl_down(tot, off);
  if (lane == 0 && a.steps_total) atomicAdd(a.steps_total, tot);
}

}  // namespace

// =====================================================================
// host side
// =====================================================================
struct TraceState {
  unsigned long long* d_queue = nullptr;
  unsigned long long* d_steps = nullptr;
  double* d_x0 = nullptr;
  double* d_dir = nullptr;
  double* d_R = nullptr;
  double* d_grid = nullptr;
  int* d_lo = nullptr;
  int* d_hi = nullptr;
  int* d_cut = nullptr;
  int* d_pend = nullptr;
  int* d_status = nullptr;
  int8_t* d_states = nullptr;
  amps_trace_result* d_out = nullptr;
  amps_trace_exit* d_exits = nullptr;
  size_t b_x0 = 0, b_dir = 0, b_R = 0, b_lo = 0, b_hi = 0, b_cut = 0, b_states = 0, b_out = 0, b_exits = 0;
  size_t b_pend = 0, b_status = 0;
  int grid_waves = 0;  // persistent grid size (workgroups); 0 = 8 per CU
  int chunk = 32;      // steps between queue refills
};

namespace amps {
void trace_free(amps_ctx* c) {
  TraceState* t = (TraceState*)c->trace_state;
  if (!t) return;
  (void)hipFree(t->d_queue);
  (void)hipFree(t->d_steps);
  (void)hipFree(t->d_x0);
  (void)hipFree(t->d_dir);
  (void)hipFree(t->d_R);
  (void)hipFree(t->d_grid);
  (void)hipFree(t->d_lo);
  (void)hipFree(t->d_hi);
  (void)hipFree(t->d_cut);
  (void)hipFree(t->d_pend);
  (void)hipFree(t->d_status);
  (void)hipFree(t->d_states);
  (void)hipFree(t->d_out);
  (void)hipFree(t->d_exits);
  delete t;
  c->trace_state = nullptr;
}

static TraceState* tstate(amps_ctx* c) {
  if (!c->trace_state) {
    TraceState* t = new TraceState();
    if (hipMalloc(&t->d_queue, 64) != hipSuccess || hipMalloc(&t->d_steps, 64) != hipSuccess) {
      delete t;
      return nullptr;
    }
    c->trace_state = t;
  }
  return (TraceState*)c->trace_state;
}

static int ensure(void** p, size_t* have, size_t bytes) {
  if (bytes <= *have && *p) return 0;
  (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  HIP_TRY(hipMalloc(p, std::max<size_t>(bytes, 64)));
  *have = std::max<size_t>(bytes, 64);
  return 0;
}

static int check_params(const amps_trace_params* P) {
  if (!P) return fail(AMPS_E_ARG, "trace: null params");
  if (!(P->m0_kg > 0.0) || !(P->qabs_C > 0.0)) return fail(AMPS_E_ARG, "trace: mass and |charge| must be > 0");
  if (!(P->earth_radius_m > 0.0)) return fail(AMPS_E_ARG, "trace: earth_radius_m must be > 0");
  if (P->max_steps <= 0) return fail(AMPS_E_ARG, "trace: max_steps must be > 0");
  if (P->field_model < 0 || P->field_model > 1) return fail(AMPS_E_ARG, "trace: field_model must be 0 or 1");
  if (P->retry < 0 || P->retry > 2) return fail(AMPS_E_ARG, "trace: retry must be 0, 1 or 2");
  if (P->policy < 0 || P->policy > 1) return fail(AMPS_E_ARG, "trace: policy must be 0 (legacy) or 1 (structured)");
  if (P->trap && (P->trap_min_mirror < 0 || P->trap_min_bounce < 0 || P->outer_sphere))
    return fail(AMPS_E_ARG, "trace: trap detection needs non-negative minima and the box boundary");
  if (P->policy == 1 && (P->outer_sphere || P->dt_selector != 0))
    return fail(AMPS_E_ARG, "trace: the structured policy runs with the box boundary and its own step selector");
  if (P->dt_selector < 0 || P->dt_selector > 1) return fail(AMPS_E_ARG, "trace: dt_selector must be 0 or 1");
  if (P->mover < MV_BORIS || P->mover > MV_RK6)
    return fail(AMPS_E_ARG, "trace: mover must be 0 BORIS, 1 HC4, 2 RK2, 3 RK4 or 4 RK6 "
                            "(the guiding-centre and hybrid movers are not built)");
  return 0;
}

static void fill_view(amps_ctx* c, TraceState* t, const amps_trace_params* P, TraceView& a) {
  std::memset(&a, 0, sizeof(a));
  a.P = *P;
  const double Re = P->earth_radius_m;
  for (int d = 0; d < 3; d++) {
    a.box_min_m[d] = P->box_min_Re[d] * Re;
    a.box_max_m[d] = P->box_max_Re[d] * Re;
  }
  a.r_inner_m = P->r_inner_Re * Re;
  a.r_outer_m = P->r_outer_Re * Re;
  // normalized (FP32) constants: length unit Re, time unit Re/c
  a.len = Re;
  const double T = Re / C_LIGHT;
  const double kq = P->q_C * T / P->m0_kg;
  a.f_cb = (float)(kq * 1.0e-7 * P->moment_Am2 / (Re * Re * Re));
  for (int d = 0; d < 3; d++) {
    a.f_mhat[d] = (float)P->m_hat[d];
    a.f_bext[d] = (float)(kq * P->B_ext_T[d]);
  }
  a.f_dphi = (float)P->dphi_max;
  a.rin_n = P->r_inner_Re;
  a.rout_n = P->r_outer_Re;
  a.maxdist_n = (P->max_distance_m > 0.0) ? P->max_distance_m / Re : -1.0;
  a.floor_abs_n = 1.0e-12 / T;
  a.floor_rel_n = 1.0e-9 * std::max(P->dt_trace_s, 1.0) / T;
  a.floor_dist_n = 100.0e3 / Re;
  a.queue = t->d_queue;
  a.steps_total = t->d_steps;
  a.chunk = t->chunk;
}

template <int FM, bool X>
static void launch64(amps_ctx* c, const TraceView& a, dim3 g, dim3 blk) {
  switch (a.P.mover) {
    case MV_HC4: hipLaunchKernelGGL((k_trace_persistent<Fp64Stepper<MV_HC4, FM, X>>), g, blk, 0, c->stream, a); break;
    case MV_RK2: hipLaunchKernelGGL((k_trace_persistent<Fp64Stepper<MV_RK2, FM, X>>), g, blk, 0, c->stream, a); break;
    case MV_RK4: hipLaunchKernelGGL((k_trace_persistent<Fp64Stepper<MV_RK4, FM, X>>), g, blk, 0, c->stream, a); break;
    case MV_RK6: hipLaunchKernelGGL((k_trace_persistent<Fp64Stepper<MV_RK6, FM, X>>), g, blk, 0, c->stream, a); break;
    default: hipLaunchKernelGGL((k_trace_persistent<Fp64Stepper<MV_BORIS, FM, X>>), g, blk, 0, c->stream, a); break;
  }
}

static int launch(amps_ctx* c, TraceState* t, TraceView& a, int precision) {
  if (precision == AMPS_FP32 && a.P.mover != MV_BORIS && a.P.mover != MV_HC4)
    return fail(AMPS_E_ARG, "trace: the FP32 path runs the BORIS and HC4 movers only");
  if (precision == AMPS_FP32 && (a.P.field_model != 0 || a.P.dt_selector != 0 || a.P.policy != 0 || a.P.trap))
    return fail(AMPS_E_ARG, "trace: the FP32 path runs the analytic field and the legacy gridless policy only");
  if (a.P.field_model == 1) {
    if (!c->has_mode3d) return fail(AMPS_E_STATE, "trace: field_model 1 needs amps_gpu_set_mode3d_field");
    a.mesh = c->mv;
    a.Bm = c->d_Bm;
  }
  HIP_TRY(hipMemsetAsync(t->d_queue, 0, 8, c->stream));
  int grid = t->grid_waves;
  if (grid <= 0) {
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, c->device));
    grid = p.multiProcessorCount * 8;
  }
  tbegin(c);
  const dim3 g(grid), blk(256);
  if (precision == AMPS_FP32) {
    if (a.P.mover == MV_HC4)
      hipLaunchKernelGGL(k_trace_persistent<Fp32Stepper<MV_HC4>>, g, blk, 0, c->stream, a);
    else
      hipLaunchKernelGGL(k_trace_persistent<Fp32Stepper<MV_BORIS>>, g, blk, 0, c->stream, a);
  } else {
    const bool x = a.P.policy == 1 || a.P.trap;
    if (a.P.field_model == 1)
      x ? launch64<1, true>(c, a, g, blk) : launch64<1, false>(c, a, g, blk);
    else
      x ? launch64<0, true>(c, a, g, blk) : launch64<0, false>(c, a, g, blk);
  }
  HIP_TRY(hipGetLastError());
  tend(c, AMPS_K_TRACE);
  return 0;
}
}  // namespace amps

extern "C" int amps_gpu_trace_set_param(amps_ctx* c, const char* name, int64_t v) {
  if (!c || !name) return fail(AMPS_E_ARG, "trace_set_param: null argument");
  TraceState* t = tstate(c);
  if (!t) return fail(AMPS_E_HIP, "trace: allocation failed");
  std::string n(name);
  if (n == "workgroups") {
    if (v < 0 || v > (1 << 20)) return fail(AMPS_E_ARG, "workgroups out of range");
    t->grid_waves = (int)v;
  } else if (n == "chunk") {
    if (v < 1 || v > 100000) return fail(AMPS_E_ARG, "chunk out of range [1, 1e5]");
    t->chunk = (int)v;
  } else {
    return fail(AMPS_E_ARG, "unknown trace parameter " + n);
  }
  return AMPS_OK;
}

static int trace_list(amps_ctx* c, const amps_trace_params* P, int32_t precision, int64_t n, const double* x0,
                      const double* dir, const double* R_GV, amps_trace_result* out, amps_trace_exit* exits,
                      int64_t* steps_total) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (int r = check_params(P)) return r;
  if (precision != AMPS_FP64 && precision != AMPS_FP32) return fail(AMPS_E_ARG, "trace: bad precision");
  if (n < 0 || n > ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "trace: n out of range");
  if (n == 0) return AMPS_OK;
  if (!x0 || !dir || !R_GV || !out) return fail(AMPS_E_ARG, "trace: null array");
  (void)hipSetDevice(c->device);
  TraceState* t = tstate(c);
  if (!t) return fail(AMPS_E_HIP, "trace: allocation failed");
  if (int r = ensure((void**)&t->d_x0, &t->b_x0, sizeof(double) * 3 * n)) return r;
  if (int r = ensure((void**)&t->d_dir, &t->b_dir, sizeof(double) * 3 * n)) return r;
  if (int r = ensure((void**)&t->d_R, &t->b_R, sizeof(double) * n)) return r;
  if (int r = ensure((void**)&t->d_out, &t->b_out, sizeof(amps_trace_result) * n)) return r;
  if (exits)
    if (int r = ensure((void**)&t->d_exits, &t->b_exits, sizeof(amps_trace_exit) * n)) return r;
  HIP_TRY(hipMemcpyAsync(t->d_x0, x0, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(t->d_dir, dir, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(t->d_R, R_GV, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(t->d_steps, 0, 8, c->stream));
  TraceView a;
  fill_view(c, t, P, a);
  a.mode = TRACE_MODE_LIST;
  a.nitems = n;
  a.x0 = t->d_x0;
  a.dir = t->d_dir;
  a.R = t->d_R;
  a.out = t->d_out;
  a.exits = exits ? t->d_exits : nullptr;
  if (int r = launch(c, t, a, precision)) return r;
  HIP_TRY(hipMemcpyAsync(out, t->d_out, sizeof(amps_trace_result) * n, hipMemcpyDeviceToHost, c->stream));
  if (exits)
    HIP_TRY(hipMemcpyAsync(exits, t->d_exits, sizeof(amps_trace_exit) * n, hipMemcpyDeviceToHost, c->stream));
  unsigned long long st = 0;
  HIP_TRY(hipMemcpyAsync(&st, t->d_steps, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (steps_total) *steps_total = (int64_t)st;
  return AMPS_OK;
}

extern "C" int amps_gpu_trace_batch(amps_ctx* c, const amps_trace_params* P, int32_t precision,
                                    int64_t n, const double* x0, const double* dir,
                                    const double* R_GV, amps_trace_result* out,
                                    int64_t* steps_total) {
  return trace_list(c, P, precision, n, x0, dir, R_GV, out, nullptr, steps_total);
}

// TraceTrajectorySharedEx batched: the exit states of the allowed
// structured traces come back with the results.
extern "C" int amps_gpu_trace_batch_ex(amps_ctx* c, const amps_trace_params* P, int32_t precision,
                                       int64_t n, const double* x0, const double* dir,
                                       const double* R_GV, amps_trace_result* out, amps_trace_exit* exits,
                                       int64_t* steps_total) {
  if (!exits && n > 0) return fail(AMPS_E_ARG, "trace_batch_ex: null exits");
  return trace_list(c, P, precision, n, x0, dir, R_GV, out, exits, steps_total);
}

// ComputeWeightBlockAtEnergy's isotropic branch (DensityGridless.cpp:
// 880-912) for every (point, rigidity) at once: one launch traces all
// (point, rigidity, direction) tasks, the per-(point, rigidity) counts are
// then summed on the host in direction order.
extern "C" int amps_gpu_transmission(amps_ctx* c, const amps_trace_params* P, int32_t precision, int64_t npts,
                                     const double* x0, int32_t nR, const double* R_GV, int32_t ndir,
                                     const double* dirs, double* weight_sum, int32_t* sampled, int32_t* resolved,
                                     int32_t* retried, int32_t* term_counts, int64_t* steps_total) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (npts < 0 || nR < 0 || ndir < 0) return fail(AMPS_E_ARG, "transmission: negative sizes");
  const int64_t n = npts * (int64_t)nR * (int64_t)ndir;
  if (n > ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "transmission: more than 2^31 trajectories");
  if (n == 0) return AMPS_OK;
  if (!x0 || !R_GV || !dirs) return fail(AMPS_E_ARG, "transmission: null array");
  std::vector<double> tx((size_t)n * 3), td((size_t)n * 3), tr((size_t)n);
  for (int64_t p = 0; p < npts; p++)
    for (int32_t e = 0; e < nR; e++)
      for (int32_t k = 0; k < ndir; k++) {
        const size_t i = (size_t)((p * nR + e) * ndir + k);
        for (int d = 0; d < 3; d++) tx[3 * i + d] = x0[3 * p + d], td[3 * i + d] = dirs[3 * k + d];
        tr[i] = R_GV[e];
      }
  std::vector<amps_trace_result> out((size_t)n);
  if (int r = trace_list(c, P, precision, n, tx.data(), td.data(), tr.data(), out.data(), nullptr, steps_total))
    return r;
  for (int64_t g = 0; g < npts * nR; g++) {
    double ws = 0.0;
    int32_t ns = 0, nres = 0, nret = 0, tc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int32_t k = 0; k < ndir; k++) {
      const amps_trace_result& r = out[(size_t)(g * ndir + k)];
      ++ns;
      if (r.retry_count > 0) ++nret;
      if (r.termination >= 0 && r.termination < 9) ++tc[r.termination];
      if (r.termination > 2) continue;  // unresolved
      ++nres;
      if (r.termination == 0) ws += 1.0;
    }
    if (weight_sum) weight_sum[g] = ws;
    if (sampled) sampled[g] = ns;
    if (resolved) resolved[g] = nres;
    if (retried) retried[g] = nret;
    if (term_counts)
      for (int t2 = 0; t2 < 9; t2++) term_counts[9 * g + t2] = tc[t2];
  }
  return AMPS_OK;
}

// RefineCutoffTransition (CutoffBandSearch.h:212-242) for many brackets at
// once: every bisection round traces the midpoints of all still-open brackets
// in one persistent launch.  Midpoint, tolerance and update are the
// reference's: tol = max(abs_tol, rel_tol max(|lo|, |hi|)), mid = 0.5 (lo + hi),
// Allowed -> hi = mid, PhysicalForbidden -> lo = mid; a trace-limit
// termination is forbidden when trace_limit_forbidden is set (the
// "FORBIDDEN" traceLimitPolicy, CutoffRigidityGridless.cpp:2685-2694) and
// otherwise ends that bracket as unresolved; a numerical failure after the
// retry ends it with status 2 (the reference throws).
extern "C" int amps_gpu_refine_cutoff(amps_ctx* c, const amps_trace_params* P, int32_t precision,
                                      int64_t n, const double* x0, const double* dir,
                                      const double* R_forbid, const double* R_allow, double abs_tol,
                                      double rel_tol, int32_t trace_limit_forbidden, double* cutoff,
                                      int32_t* iterations, int32_t* status, int64_t* steps_total) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (n < 0 || n > ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "refine_cutoff: n out of range");
  if (n == 0) return AMPS_OK;
  if (!x0 || !dir || !R_forbid || !R_allow || !cutoff || !iterations || !status)
    return fail(AMPS_E_ARG, "refine_cutoff: null array");
  std::vector<double> lo(R_forbid, R_forbid + n), hi(R_allow, R_allow + n);
  std::vector<int64_t> act;
  for (int64_t i = 0; i < n; i++) {
    iterations[i] = 0;
    status[i] = 0;
    cutoff[i] = -1.0;
    if (!(hi[(size_t)i] > lo[(size_t)i])) continue;  // RefinedCutoffTransition{} (cutoff -1)
    act.push_back(i);
  }
  auto tol_of = [&](int64_t i) {
    return std::max(abs_tol, rel_tol * std::max(std::fabs(R_forbid[i]), std::fabs(R_allow[i])));
  };
  int64_t st_total = 0;
  std::vector<double> bx, bd, bR;
  std::vector<amps_trace_result> res;
  std::vector<int64_t> next;
  while (!act.empty()) {
    next.clear();
    bx.clear();
    bd.clear();
    bR.clear();
    for (int64_t i : act) {
      if (!((hi[(size_t)i] - lo[(size_t)i]) > tol_of(i))) {
        cutoff[i] = hi[(size_t)i];
        continue;
      }
      next.push_back(i);
      for (int d = 0; d < 3; d++) bx.push_back(x0[3 * i + d]), bd.push_back(dir[3 * i + d]);
      bR.push_back(0.5 * (lo[(size_t)i] + hi[(size_t)i]));
    }
    if (next.empty()) break;
    res.resize(next.size());
    int64_t st = 0;
    if (int r = amps_gpu_trace_batch(c, P, precision, (int64_t)next.size(), bx.data(), bd.data(), bR.data(),
                                     res.data(), &st))
      return r;
    st_total += st;
    act.clear();
    for (size_t k = 0; k < next.size(); k++) {
      const int64_t i = next[k];
      iterations[i]++;
      const int term = res[k].termination;
      const double mid = bR[k];
      if (term == 0) {
        hi[(size_t)i] = mid;
      } else if (term == 1 || term == 2 || (trace_limit_forbidden && term >= 3 && term <= 5)) {
        lo[(size_t)i] = mid;
      } else {
        status[i] = (term >= 3 && term <= 5) ? 1 : 2;
        cutoff[i] = hi[(size_t)i];
        continue;
      }
      act.push_back(i);
    }
  }
  if (steps_total) *steps_total = st_total;
  return AMPS_OK;
}

// AnalyzeCutoffBandSamples (CutoffBandSearch.h:96-195) for n state vectors:
// the lower and upper transition brackets, transition and allowed-interval
// counts and the range/unresolved flags.  An untraced sample (-1) counts as
// unresolved.
extern "C" int amps_gpu_band_topology(int64_t n, int32_t nR, const int8_t* states, int32_t* out) {
  if (n < 0 || nR < 0 || nR > 4096) return fail(AMPS_E_RANGE, "band_topology: n or nR out of range");
  if (n > 0 && (!states || !out)) return fail(AMPS_E_ARG, "band_topology: null array");
  constexpr int8_t FORB = 0, ALLOW = 1;
  for (int64_t i = 0; i < n; i++) {
    const int8_t* s = states + i * (int64_t)nR;
    auto st = [&](int k) { return (s[k] == FORB || s[k] == ALLOW) ? s[k] : (int8_t)2; };
    int lfi = -1, lai = -1, ufi = -1, uai = -1, first_allowed = -1, last_forbidden = -1;
    int ntr = 0, nint = 0, nunres = 0;
    bool all_allowed = false, all_forbidden = false, lower_below = false, lower_above = false;
    bool upper_below = false, upper_above = false, lower_unres = false, upper_unres = false;
    if (nR > 0) {
      bool in_allowed = false;
      for (int k = 0; k < nR; k++) {
        const int8_t v = st(k);
        if (v == ALLOW) {
          if (first_allowed < 0) first_allowed = k;
          if (!in_allowed) ++nint, in_allowed = true;
        } else {
          in_allowed = false;
          if (v == FORB) last_forbidden = k;
          else ++nunres;
        }
        if (k > 0) {
          const int8_t a = st(k - 1);
          if (a != 2 && v != 2 && a != v) ++ntr;
        }
      }
      all_allowed = first_allowed == 0 && last_forbidden < 0 && nunres == 0;
      all_forbidden = first_allowed < 0 && last_forbidden == nR - 1 && nunres == 0;
      if (first_allowed == 0) {
        lower_below = true;
      } else if (first_allowed > 0) {
        bool clean = true;
        for (int k = 0; k < first_allowed && clean; k++) clean = st(k) == FORB;
        if (clean) lfi = first_allowed - 1, lai = first_allowed;
        else lower_unres = true;
      } else if (nunres > 0) {
        lower_unres = true;
      } else {
        lower_above = true;
      }
      if (last_forbidden < 0) {
        if (nunres > 0) upper_unres = true;
        else upper_below = true;
      } else if (last_forbidden == nR - 1) {
        upper_above = true;
      } else {
        bool clean = true;
        for (int k = last_forbidden + 1; k < nR && clean; k++) clean = st(k) == ALLOW;
        if (clean) ufi = last_forbidden, uai = last_forbidden + 1;
        else upper_unres = true;
      }
    }
    int32_t* o = out + 15 * i;
    o[0] = lfi, o[1] = lai, o[2] = ufi, o[3] = uai, o[4] = ntr, o[5] = nint, o[6] = nunres;
    o[7] = all_allowed, o[8] = all_forbidden, o[9] = lower_below, o[10] = lower_above;
    o[11] = upper_below, o[12] = upper_above, o[13] = lower_unres, o[14] = upper_unres;
  }
  return AMPS_OK;
}

// IntegrateEffectiveCutoff (CutoffBandSearch.h:284-366) for n (location,
// direction) pairs.  Every coarse interval of [lower, upper] whose end states
// differ is bisected; the bisections of all pairs advance together, one
// batched trace of their midpoints per round.  The widths are then summed per
// pair in interval order, stopping where the reference returns early (an
// unresolved end state or midpoint), so counts and widths match the serial
// routine.
extern "C" int amps_gpu_effective_cutoff(amps_ctx* c, const amps_trace_params* P, int32_t precision,
                                         int64_t n, const double* x0, const double* dir, int32_t nR,
                                         const double* grid, const int8_t* states, const double* lower,
                                         const double* upper, double abs_tol, double rel_tol,
                                         int32_t trace_limit_forbidden, double* effective,
                                         double* allowed_width, int32_t* refined, int32_t* status,
                                         int64_t* steps_total) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (n < 0 || n > ((int64_t)1 << 28) || nR < 0 || nR > 4096)
    return fail(AMPS_E_RANGE, "effective_cutoff: n or nR out of range");
  if (steps_total) *steps_total = 0;
  if (n == 0) return AMPS_OK;
  if (!x0 || !dir || !grid || !states || !lower || !upper || !effective || !allowed_width || !refined || !status)
    return fail(AMPS_E_ARG, "effective_cutoff: null array");
  constexpr int8_t FORB = 0, ALLOW = 1, UNRES = 2;  // CutoffSampleState
  bool grid_ok = nR >= 2;
  for (int k = 0; k + 1 < nR; k++) grid_ok = grid_ok && grid[k + 1] > grid[k];
  // one bisection per straddling interval that the serial routine reaches
  struct Bis {
    int64_t pair;
    int k;
    double lo, hi, tol;
    int8_t low_state;
    int count;
    bool unresolved, done;
  };
  std::vector<Bis> bis;
  std::vector<int> stop_at((size_t)n, -1);  // first interval with an unresolved end state
  for (int64_t i = 0; i < n; i++) {
    effective[i] = -1.0;
    allowed_width[i] = 0.0;
    refined[i] = 0;
    status[i] = 0;
    const double lo_i = lower[i], up_i = upper[i];
    if (!grid_ok || !(lo_i >= 0.0) || !(up_i >= lo_i) || up_i == lo_i) continue;
    const int8_t* st = states + i * (int64_t)nR;
    for (int k = 0; k + 1 < nR; k++) {
      const double a = std::max(grid[k], lo_i), b = std::min(grid[k + 1], up_i);
      if (!(b > a)) continue;
      const int8_t l = st[k], r = st[k + 1];
      if (l == UNRES || r == UNRES || l < 0 || r < 0) {
        stop_at[(size_t)i] = k;
        break;
      }
      if (l == r) continue;
      const double tol = std::max(abs_tol, rel_tol * std::max(std::fabs(grid[k]), std::fabs(grid[k + 1])));
      bis.push_back({i, k, grid[k], grid[k + 1], tol, l, 0, false, false});
    }
  }
  int64_t st_total = 0;
  std::vector<double> bx, bd, bR;
  std::vector<amps_trace_result> res;
  std::vector<size_t> act;
  for (size_t q = 0; q < bis.size(); q++) act.push_back(q);
  while (!act.empty()) {
    std::vector<size_t> run;
    bx.clear(), bd.clear(), bR.clear();
    for (size_t q : act) {
      Bis& B = bis[q];
      if (!((B.hi - B.lo) > B.tol)) {
        B.done = true;
        continue;
      }
      run.push_back(q);
      for (int d = 0; d < 3; d++) bx.push_back(x0[3 * B.pair + d]), bd.push_back(dir[3 * B.pair + d]);
      bR.push_back(0.5 * (B.lo + B.hi));
    }
    if (run.empty()) break;
    res.resize(run.size());
    int64_t st = 0;
    if (int r = amps_gpu_trace_batch(c, P, precision, (int64_t)run.size(), bx.data(), bd.data(), bR.data(),
                                     res.data(), &st))
      return r;
    st_total += st;
    act.clear();
    for (size_t k = 0; k < run.size(); k++) {
      Bis& B = bis[run[k]];
      B.count++;
      const int term = res[k].termination;
      int8_t ms;
      if (term == 0) ms = ALLOW;
      else if (term == 1 || term == 2 || (trace_limit_forbidden && term >= 3 && term <= 5)) ms = FORB;
      else ms = UNRES;
      if (ms == UNRES) {
        B.unresolved = true;
        B.done = true;
        if (term >= 6) status[B.pair] = 2;  // numerical failure after retry (the reference throws)
        continue;
      }
      if (ms == B.low_state) B.lo = bR[k];
      else B.hi = bR[k];
      act.push_back(run[k]);
    }
  }
  // serial accumulation in interval order
  size_t q = 0;
  for (int64_t i = 0; i < n; i++) {
    const double lo_i = lower[i], up_i = upper[i];
    if (!grid_ok || !(lo_i >= 0.0) || !(up_i >= lo_i)) continue;
    if (up_i == lo_i) {
      effective[i] = lo_i;
      continue;
    }
    const int8_t* st = states + i * (int64_t)nR;
    bool unres = false;
    double width = 0.0;
    int count = 0;
    for (int k = 0; k + 1 < nR && !unres; k++) {
      const double a = std::max(grid[k], lo_i), b = std::min(grid[k + 1], up_i);
      if (!(b > a)) continue;
      if (k == stop_at[(size_t)i]) {
        unres = true;
        break;
      }
      const int8_t l = st[k], r = st[k + 1];
      if (l == r) {
        if (l == ALLOW) width += b - a;
        continue;
      }
      const Bis& B = bis[q++];
      count += B.count;
      if (B.unresolved) {
        unres = true;
        break;
      }
      const double tr = 0.5 * (B.lo + B.hi);
      if (l == ALLOW)
        width += std::max(0.0, std::min(tr, b) - a);
      else
        width += std::max(0.0, b - std::max(tr, a));
    }
    // the bisections queued beyond the stop (none: the queue stops at it) are skipped
    while (q < bis.size() && bis[q].pair == i) q++;
    allowed_width[i] = width;
    refined[i] = count;
    if (unres) {
      if (status[i] == 0) status[i] = 1;
      continue;
    }
    effective[i] = std::max(lo_i, std::min(up_i, up_i - width));
  }
  if (steps_total) *steps_total = st_total;
  return AMPS_OK;
}

extern "C" int amps_gpu_cutoff_scan(amps_ctx* c, const amps_trace_params* P, int32_t precision,
                                    int32_t mode, int64_t nloc, const double* x0,
                                    const double* dir, int32_t nR, const double* grid,
                                    const int32_t* bin_lo, const int32_t* bin_hi,
                                    int8_t* states, int32_t* cutoff_bin, int64_t* steps_total) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (int r = check_params(P)) return r;
  if (precision != AMPS_FP64 && precision != AMPS_FP32) return fail(AMPS_E_ARG, "cutoff_scan: bad precision");
  if (mode != AMPS_SCAN_FULL && mode != AMPS_SCAN_UPPER && mode != AMPS_SCAN_BRACKETED)
    return fail(AMPS_E_ARG, "cutoff_scan: bad mode");
  if (mode == AMPS_SCAN_BRACKETED && !bin_lo) return fail(AMPS_E_ARG, "cutoff_scan: BRACKETED mode needs the brackets");
  if (nloc < 0 || nloc > ((int64_t)1 << 28) || nR < 2 || nR > 4096)
    return fail(AMPS_E_RANGE, "cutoff_scan: nloc or nR out of range");
  if (nloc == 0) return AMPS_OK;
  if (!x0 || !dir || !grid || !states || !cutoff_bin) return fail(AMPS_E_ARG, "cutoff_scan: null array");
  if ((bin_lo == nullptr) != (bin_hi == nullptr)) return fail(AMPS_E_ARG, "cutoff_scan: give both bracket arrays or neither");
  int nbin_item = nR;
  if (bin_lo) {
    for (int64_t i = 0; i < nloc; i++)
      if (bin_lo[i] < 0 || bin_hi[i] >= nR || bin_lo[i] > bin_hi[i])
        return fail(AMPS_E_ARG, "cutoff_scan: bracket outside [0, nR)");
    if (mode != AMPS_SCAN_UPPER) {
      const int w = bin_hi[0] - bin_lo[0] + 1;
      for (int64_t i = 0; i < nloc; i++)
        if (bin_hi[i] - bin_lo[i] + 1 != w)
          return fail(AMPS_E_ARG, "cutoff_scan: FULL and BRACKETED modes need equal bracket widths");
      nbin_item = w;
    }
  }
  (void)hipSetDevice(c->device);
  TraceState* t = tstate(c);
  if (!t) return fail(AMPS_E_HIP, "cutoff_scan: allocation failed");
  if (int r = ensure((void**)&t->d_x0, &t->b_x0, sizeof(double) * 3 * nloc)) return r;
  if (int r = ensure((void**)&t->d_dir, &t->b_dir, sizeof(double) * 3 * nloc)) return r;
  if (int r = ensure((void**)&t->d_lo, &t->b_lo, sizeof(int) * nloc)) return r;
  if (int r = ensure((void**)&t->d_hi, &t->b_hi, sizeof(int) * nloc)) return r;
  if (int r = ensure((void**)&t->d_cut, &t->b_cut, sizeof(int) * nloc)) return r;
  if (int r = ensure((void**)&t->d_states, &t->b_states, (size_t)nloc * nR)) return r;
  if (!t->d_grid) HIP_TRY(hipMalloc(&t->d_grid, sizeof(double) * 4096));
  HIP_TRY(hipMemcpyAsync(t->d_x0, x0, sizeof(double) * 3 * nloc, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(t->d_dir, dir, sizeof(double) * 3 * nloc, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(t->d_grid, grid, sizeof(double) * nR, hipMemcpyHostToDevice, c->stream));
  if (bin_lo) {
    HIP_TRY(hipMemcpyAsync(t->d_lo, bin_lo, sizeof(int) * nloc, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(t->d_hi, bin_hi, sizeof(int) * nloc, hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemsetAsync(t->d_states, -1, nloc * nR, c->stream));
  HIP_TRY(hipMemsetAsync(t->d_cut, 0, sizeof(int) * nloc, c->stream));
  HIP_TRY(hipMemsetAsync(t->d_steps, 0, 8, c->stream));
  std::vector<int> pend;
  if (mode == AMPS_SCAN_BRACKETED) {
    // every location starts in its bracket phase with W outstanding bins
    if (int r = ensure((void**)&t->d_pend, &t->b_pend, sizeof(int) * nloc)) return r;
    if (int r = ensure((void**)&t->d_status, &t->b_status, sizeof(int) * nloc)) return r;
    pend.assign((size_t)nloc, nbin_item);
    HIP_TRY(hipMemcpyAsync(t->d_pend, pend.data(), sizeof(int) * nloc, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemsetAsync(t->d_status, 0, sizeof(int) * nloc, c->stream));
  }
  TraceView a;
  fill_view(c, t, P, a);
  a.mode = (mode == AMPS_SCAN_FULL) ? TRACE_MODE_FULL : (mode == AMPS_SCAN_UPPER ? TRACE_MODE_UPPER : TRACE_MODE_BRACKET);
  a.nitems = (mode == AMPS_SCAN_FULL) ? nloc * nbin_item : (mode == AMPS_SCAN_UPPER ? nloc : nloc * nR);
  a.nloc = nloc;
  a.pend = t->d_pend;
  a.status = t->d_status;
  a.nbin_item = nbin_item;
  a.nR = nR;
  a.x0 = t->d_x0;
  a.dir = t->d_dir;
  a.grid = t->d_grid;
  a.lo = bin_lo ? t->d_lo : nullptr;
  a.hi = bin_lo ? t->d_hi : nullptr;
  a.states = t->d_states;
  a.cutoff = t->d_cut;
  if (int r = launch(c, t, a, precision)) return r;
  HIP_TRY(hipMemcpyAsync(states, t->d_states, nloc * nR, hipMemcpyDeviceToHost, c->stream));
  std::vector<int> cut((size_t)nloc);
  HIP_TRY(hipMemcpyAsync(cut.data(), t->d_cut, sizeof(int) * nloc, hipMemcpyDeviceToHost, c->stream));
  unsigned long long st = 0;
  HIP_TRY(hipMemcpyAsync(&st, t->d_steps, 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (steps_total) *steps_total = (int64_t)st;
  // cutoff bins
  for (int64_t i = 0; i < nloc; i++) {
    const int8_t* s = states + i * nR;
    const int lo = bin_lo ? bin_lo[i] : 0, hi = bin_hi ? bin_hi[i] : nR - 1;
    if (mode == AMPS_SCAN_BRACKETED) {
      cutoff_bin[i] = cut[(size_t)i];  // read off on the device, rescan included
    } else if (mode == AMPS_SCAN_UPPER) {
      int b = cut[(size_t)i];
      // bracket diagnostics: top sample forbidden below Rmax, or allowed down to a bottom above Rmin
      if (s[hi] != 1) b = (hi == nR - 1) ? -1 : AMPS_BIN_ABOVE;
      else if (b == lo && lo > 0 && s[lo] == 1) b = AMPS_BIN_BELOW;
      cutoff_bin[i] = b;
    } else {
      // UPPER_SCAN result from a full sample vector (CutoffForDirectionUpperScan_GV)
      int b;
      if (s[hi] != 1) b = (hi == nR - 1) ? -1 : AMPS_BIN_ABOVE;
      else {
        b = lo;
        for (int k = hi - 1; k >= lo; --k)
          if (s[k] != 1) {
            b = k + 1;
            break;
          }
        if (b == lo && lo > 0) b = AMPS_BIN_BELOW;
      }
      cutoff_bin[i] = b;
    }
  }
  return AMPS_OK;
}
