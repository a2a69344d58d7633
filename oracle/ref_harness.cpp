// ref_harness.cpp — C entry points over the reference's own gridless mover,
// dipole and cutoff-band sources, compiled from /root/reference by
// oracle/Makefile into oracle/_ref/.  Test infrastructure only: used to pin
// the oracle's restatement, never by the product path.
#include <vector>

#include "CutoffBandSearch.h"
#include "DipoleInterface.h"
#include "GridlessParticleMovers.h"
#include "TrajectoryBoundary.h"
#include "TrajectoryTimeStep.h"
#include "TrajectoryTrapDetector.h"

namespace {
struct DipolePlusUniform : public IGridlessFieldEvaluator {
  double bext[3] = {0.0, 0.0, 0.0};
  void GetB_T(const V3& x, V3& B) const override {
    const double xm[3] = {x.x, x.y, x.z};
    double b[3];
    Earth::GridlessMode::Dipole::GetB_Tesla(xm, b);
    B = V3{b[0] + bext[0], b[1] + bext[1], b[2] + bext[2]};
  }
};
}  // namespace

extern "C" {

void ref_set_dipole(double moment_scale, double tilt_deg) {
  Earth::GridlessMode::Dipole::SetMomentScale(moment_scale);
  Earth::GridlessMode::Dipole::SetTiltDeg(tilt_deg);
}

void ref_get_dipole(double* moment_Am2, double m_hat[3]) {
  using namespace Earth::GridlessMode::Dipole;
  *moment_Am2 = gParams.momentScale_Me * M_E_Am2;
  for (int d = 0; d < 3; d++) m_hat[d] = gParams.m_hat[d];
}

void ref_field(const double x[3], const double bext[3], double B[3]) {
  DipolePlusUniform f;
  for (int d = 0; d < 3; d++) f.bext[d] = bext[d];
  V3 b;
  f.GetB_T(V3{x[0], x[1], x[2]}, b);
  B[0] = b.x, B[1] = b.y, B[2] = b.z;
}

// n consecutive BorisStep calls; x,p updated in place (3*n outputs per call
// are written to traj when traj != nullptr).
void ref_boris_steps(int n, double x[3], double p[3], double q, double m0, double dt,
                     const double bext[3], double* traj) {
  DipolePlusUniform f;
  for (int d = 0; d < 3; d++) f.bext[d] = bext[d];
  V3 xv{x[0], x[1], x[2]}, pv{p[0], p[1], p[2]};
  for (int i = 0; i < n; i++) {
    BorisStep(xv, pv, q, m0, dt, f);
    if (traj) {
      traj[6 * i + 0] = xv.x, traj[6 * i + 1] = xv.y, traj[6 * i + 2] = xv.z;
      traj[6 * i + 3] = pv.x, traj[6 * i + 4] = pv.y, traj[6 * i + 5] = pv.z;
    }
  }
  x[0] = xv.x, x[1] = xv.y, x[2] = xv.z;
  p[0] = pv.x, p[1] = pv.y, p[2] = pv.z;
}

int ref_step_checked(double x[3], double p[3], double q, double m0, double dt,
                     const double bext[3], double r_inner) {
  DipolePlusUniform f;
  for (int d = 0; d < 3; d++) f.bext[d] = bext[d];
  V3 xv{x[0], x[1], x[2]}, pv{p[0], p[1], p[2]};
  bool ok = StepParticleChecked(MoverType::BORIS, xv, pv, q, m0, dt, f, r_inner);
  x[0] = xv.x, x[1] = xv.y, x[2] = xv.z;
  p[0] = pv.x, p[1] = pv.y, p[2] = pv.z;
  return ok ? 1 : 0;
}

// n consecutive StepParticle(mover) calls (MoverType numbering: 0 BORIS,
// 1 HC4, 2 RK2, 3 RK4, 4 RK6), trajectory as in ref_boris_steps.
void ref_mover_steps(int mover, int n, double x[3], double p[3], double q, double m0, double dt,
                     const double bext[3], double* traj) {
  DipolePlusUniform f;
  for (int d = 0; d < 3; d++) f.bext[d] = bext[d];
  V3 xv{x[0], x[1], x[2]}, pv{p[0], p[1], p[2]};
  for (int i = 0; i < n; i++) {
    StepParticle(static_cast<MoverType>(mover), xv, pv, q, m0, dt, f);
    if (traj) {
      traj[6 * i + 0] = xv.x, traj[6 * i + 1] = xv.y, traj[6 * i + 2] = xv.z;
      traj[6 * i + 3] = pv.x, traj[6 * i + 4] = pv.y, traj[6 * i + 5] = pv.z;
    }
  }
  x[0] = xv.x, x[1] = xv.y, x[2] = xv.z;
  p[0] = pv.x, p[1] = pv.y, p[2] = pv.z;
}

// StepParticleChecked(mover) once; x,p updated in place.
int ref_step_checked_mover(int mover, double x[3], double p[3], double q, double m0, double dt,
                           const double bext[3], double r_inner) {
  DipolePlusUniform f;
  for (int d = 0; d < 3; d++) f.bext[d] = bext[d];
  V3 xv{x[0], x[1], x[2]}, pv{p[0], p[1], p[2]};
  bool ok = StepParticleChecked(static_cast<MoverType>(mover), xv, pv, q, m0, dt, f, r_inner);
  x[0] = xv.x, x[1] = xv.y, x[2] = xv.z;
  p[0] = pv.x, p[1] = pv.y, p[2] = pv.z;
  return ok ? 1 : 0;
}

// TrajectoryBoundary::FindFirstEvent on the chord a -> z.
void ref_find_first_event(const double a[3], const double z[3], const double bmin[3], const double bmax[3],
                          double r_inner, double tol, int* type, double* fraction, double pos[3], int* face) {
  Earth::TrajectoryBoundary::Box b;
  for (int d = 0; d < 3; d++) b.min[d] = bmin[d], b.max[d] = bmax[d];
  b.innerRadius = r_inner;
  const Earth::TrajectoryBoundary::Event e = Earth::TrajectoryBoundary::FindFirstEvent(a, z, b, tol);
  *type = static_cast<int>(e.type);
  *fraction = e.fraction;
  for (int d = 0; d < 3; d++) pos[d] = e.position[d];
  *face = e.outerFace;
}

// TrajectoryTrap::Detector::Update over n consecutive (x, p, B) samples;
// trapped[i], mirror[i], bounce[i] after sample i.
void ref_trap_sequence(int n, const double* x, const double* p, const double* B, int min_mirror, int min_bounce,
                       double outer_margin_m, double radial_tol_m, double energy_tol, double deadband,
                       const double bmin[3], const double bmax[3], int* trapped, int* mirror, int* bounce) {
  Earth::TrajectoryTrap::Config c;
  c.enabled = true;
  c.minMirrorPoints = min_mirror;
  c.minBounceCycles = min_bounce;
  c.outerMargin_m = outer_margin_m;
  c.radialEnvelopeTolerance_m = radial_tol_m;
  c.energyRelativeTolerance = energy_tol;
  c.parallelDeadband = deadband;
  Earth::TrajectoryBoundary::Box b;
  for (int d = 0; d < 3; d++) b.min[d] = bmin[d], b.max[d] = bmax[d];
  b.innerRadius = 0.0;
  Earth::TrajectoryTrap::Detector det(c, b);
  for (int i = 0; i < n; i++) {
    trapped[i] = det.Update(x + 3 * i, p + 3 * i, B + 3 * i) ? 1 : 0;
    mirror[i] = det.mirrorPoints();
    bounce[i] = det.bounceCycles();
  }
}

double ref_finalize_step(double dt, double time_remaining) {
  return Earth::TrajectoryTimeStep::FinalizeUpperBoundedStep(dt, time_remaining);
}

// AnalyzeCutoffBandSamples over states numbered as CutoffSampleState
// (0 physical forbidden, 1 allowed, 2 unresolved).
void ref_band_topology(int n, const int* states, int* out) {
  std::vector<EarthUtil::CutoffSampleState> s((size_t)n);
  for (int i = 0; i < n; i++)
    s[(size_t)i] = static_cast<EarthUtil::CutoffSampleState>(states[i]);
  EarthUtil::CutoffBandTopology t = EarthUtil::AnalyzeCutoffBandSamples(s);
  out[0] = t.lowerForbiddenIndex;
  out[1] = t.lowerAllowedIndex;
  out[2] = t.upperForbiddenIndex;
  out[3] = t.upperAllowedIndex;
  out[4] = t.nTransitions;
  out[5] = t.nAllowedIntervals;
  out[6] = t.nUnresolved;
  out[7] = t.allAllowed ? 1 : 0;
  out[8] = t.allForbidden ? 1 : 0;
  out[9] = t.lowerBelowRange ? 1 : 0;
  out[10] = t.lowerAboveRange ? 1 : 0;
  out[11] = t.upperBelowRange ? 1 : 0;
  out[12] = t.upperAboveRange ? 1 : 0;
  out[13] = t.lowerBracketUnresolved ? 1 : 0;
  out[14] = t.upperBracketUnresolved ? 1 : 0;
}

// RefineCutoffTransition driven by a caller-supplied classifier (0 physical
// forbidden, 1 allowed, 2 unresolved).
typedef int (*ref_classify_fn)(double R_GV, void* user);
void ref_refine_cutoff(double r_forbid, double r_allow, double abs_tol, double rel_tol,
                       ref_classify_fn classify, void* user, double* cutoff, int* iterations,
                       int* unresolved) {
  auto cls = [&](double R) { return static_cast<EarthUtil::CutoffSampleState>(classify(R, user)); };
  const EarthUtil::RefinedCutoffTransition r =
      EarthUtil::RefineCutoffTransition(r_forbid, r_allow, cls, abs_tol, rel_tol);
  *cutoff = r.cutoff_GV;
  *iterations = r.iterations;
  *unresolved = r.unresolved ? 1 : 0;
}

// IntegrateEffectiveCutoff over the grid and states given, with the same
// classifier convention.
void ref_effective_cutoff(int n, const double* grid, const int* states, double lower, double upper,
                          double abs_tol, double rel_tol, ref_classify_fn classify, void* user,
                          double* cutoff, double* width, int* refined, int* unresolved) {
  auto cls = [&](double R) { return static_cast<EarthUtil::CutoffSampleState>(classify(R, user)); };
  std::vector<double> g(grid, grid + n);
  std::vector<EarthUtil::CutoffSampleState> st(n);
  for (int i = 0; i < n; i++) st[i] = static_cast<EarthUtil::CutoffSampleState>(states[i]);
  const EarthUtil::EffectiveCutoffIntegration r =
      EarthUtil::IntegrateEffectiveCutoff(g, st, lower, upper, cls, abs_tol, rel_tol);
  *cutoff = r.cutoff_GV;
  *width = r.allowedWidth_GV;
  *refined = r.refinedTransitions;
  *unresolved = r.unresolved ? 1 : 0;
}
}
