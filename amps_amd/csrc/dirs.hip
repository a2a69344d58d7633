// dirs.hip — host-side direction and location helpers of the gridless cutoff
// solver: the Fibonacci direction set and the local east-north-up frame that
// turn a (location, sky longitude/latitude) pair into the tracer's arrival
// direction.  Plain host code: the caller builds its task arrays with them.
#include <cmath>
#include <cstdint>

#include "amps_common.h"

using namespace amps;

namespace {

struct V3d {
  double x, y, z;
};
// the vector helpers of GridlessParticleMovers.h:289-295
inline V3d vadd(const V3d& a, const V3d& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3d vmul(double s, const V3d& a) { return {s * a.x, s * a.y, s * a.z}; }
inline double vdot(const V3d& a, const V3d& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3d vcross(const V3d& a, const V3d& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double vnorm(const V3d& a) { return std::sqrt(vdot(a, a)); }
inline V3d vunit(const V3d& a) {
  const double n = vnorm(a);
  return (n > 0) ? vmul(1.0 / n, a) : V3d{0, 0, 0};
}

}  // namespace

// BuildUniformSphereDirs (srcEarth/gridless/CutoffRigidityGridless.cpp:
// 1652-1674): z = 1 - 2 (k + 1/2) / n uniform in [-1, 1], the azimuth advanced
// by the golden angle 2 pi (1 - 1/phi).
extern "C" int amps_gpu_uniform_sphere_dirs(int32_t n, double* dirs) {
  if (n <= 0) return fail(AMPS_E_ARG, "amps_gpu_uniform_sphere_dirs requires n > 0");
  if (!dirs) return fail(AMPS_E_ARG, "amps_gpu_uniform_sphere_dirs: null array");
  const double goldenRatio = 0.5 * (1.0 + std::sqrt(5.0));
  const double goldenAngle = 2.0 * M_PI * (1.0 - 1.0 / goldenRatio);
  for (int k = 0; k < n; ++k) {
    const double z = 1.0 - 2.0 * (static_cast<double>(k) + 0.5) / static_cast<double>(n);
    const double r = std::sqrt(std::fmax(0.0, 1.0 - z * z));
    const double phi = goldenAngle * static_cast<double>(k);
    const V3d u = vunit(V3d{r * std::cos(phi), r * std::sin(phi), z});
    dirs[3 * k] = u.x, dirs[3 * k + 1] = u.y, dirs[3 * k + 2] = u.z;
  }
  return AMPS_OK;
}

// BuildLocalENU_GSM and LocalLonLatToDir_GSM (:1757-1800): U = x/|x|,
// E = unit(Z x U) (X x U within 1e-12 of the pole), N = unit(U x E); the
// direction of sky longitude lon (from +N toward +E) and latitude lat (toward
// +U) is unit(cos(lat) sin(lon) E + cos(lat) cos(lon) N + sin(lat) U).
namespace {

struct Enu {
  V3d E, N, U;
};

Enu local_enu(const double* x) {
  const V3d U = vunit(V3d{x[0], x[1], x[2]});
  V3d E = vcross(V3d{0.0, 0.0, 1.0}, U);
  if (vnorm(E) < 1.0e-12) E = vcross(V3d{1.0, 0.0, 0.0}, U);
  E = vunit(E);
  const V3d N = vunit(vcross(U, E));
  return {E, N, U};
}

void sky_dir(const Enu& f, double lon_deg, double lat_deg, double* o) {
  const double lon = lon_deg * M_PI / 180.0;
  const double lat = lat_deg * M_PI / 180.0;
  const double cl = std::cos(lat);
  const double sl = std::sin(lat);
  const double dE = cl * std::sin(lon);
  const double dN = cl * std::cos(lon);
  const double dU = sl;
  const V3d d = vunit(vadd(vadd(vmul(dE, f.E), vmul(dN, f.N)), vmul(dU, f.U)));
  o[0] = d.x, o[1] = d.y, o[2] = d.z;
}

}  // namespace

extern "C" int amps_gpu_local_dirs(int64_t npts, const double* x0, int32_t ndir, const double* lon_deg,
                                   const double* lat_deg, double* dirs, double* frames) {
  if (npts < 0 || ndir < 0) return fail(AMPS_E_ARG, "amps_gpu_local_dirs: negative count");
  if ((npts > 0 && !x0) || (ndir > 0 && (!lon_deg || !lat_deg || !dirs)))
    return fail(AMPS_E_ARG, "amps_gpu_local_dirs: null array");
  for (int64_t p = 0; p < npts; p++) {
    const Enu f = local_enu(x0 + 3 * p);
    if (frames) {
      double* g = frames + 9 * p;
      g[0] = f.E.x, g[1] = f.E.y, g[2] = f.E.z, g[3] = f.N.x, g[4] = f.N.y, g[5] = f.N.z;
      g[6] = f.U.x, g[7] = f.U.y, g[8] = f.U.z;
    }
    for (int k = 0; k < ndir; k++) sky_dir(f, lon_deg[k], lat_deg[k], dirs + 3 * (p * ndir + k));
  }
  return AMPS_OK;
}

extern "C" int amps_gpu_local_dirs_paired(int64_t n, const double* x0, const double* lon_deg, const double* lat_deg,
                                          double* dirs) {
  if (n < 0) return fail(AMPS_E_ARG, "amps_gpu_local_dirs_paired: negative count");
  if (n > 0 && (!x0 || !lon_deg || !lat_deg || !dirs))
    return fail(AMPS_E_ARG, "amps_gpu_local_dirs_paired: null array");
  for (int64_t p = 0; p < n; p++) sky_dir(local_enu(x0 + 3 * p), lon_deg[p], lat_deg[p], dirs + 3 * p);
  return AMPS_OK;
}
