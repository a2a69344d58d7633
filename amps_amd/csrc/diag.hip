// diag.hip — the particle energy and CFL reductions of ECSIM::UpdateJMassMatrix.
//
// ProcessCell accumulates, per particle of a cell (pic_field_solver_ecsim.cpp:
// 2229-2238), vmean[spec] += sqrt(vsqr) dt and ParticleEnergyCell +=
// 0.5 mass vsqr with the weighted mass; at the end of the cell
// (:2355-2358) cflCell[spec] = vmean[spec] / (count[spec] |dx|).
// UpdateJMassMatrix then sums the energies and takes each species' maximum
// cflCell over all cells and threads (:3925-3929), and over the MPI ranks
// (:3972-3987).  Here one wavefront per block evaluates its 512 cells (lane l
// takes cells l, l+64, ..., each cell's particles in device order), the
// block's cell energies are summed in cell order, and a final single-lane
// pass sums the blocks in block order: a fixed order, so the result does not
// depend on scheduling.  A cell without particles of a species gives 0/0,
// which never wins the maximum (the reference's comparison also fails on it).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "amps_common.h"

using namespace amps;

// Block partials of the allocation-free variant: [nb][1 + nspec] doubles for
// at most NOALLOC_MAX_BLOCKS blocks of MAX_SPEC species (2.25 MiB).  Every
// GPU the library runs on holds its own copy until the process ends, and
// every call on that GPU shares it: calls that overlap on one GPU race.  Each
// call writes every entry it reads, so nothing has to start at zero.
constexpr int NOALLOC_MAX_BLOCKS = 32768;
static __device__ __align__(256) double scratch[(size_t)NOALLOC_MAX_BLOCKS * (1 + MAX_SPEC)];

namespace {

// device address of this GPU's scratch array (host code must use it, never
// the array's name), looked up once per device at context creation
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

__global__ __launch_bounds__(64) void k_energy_cfl_block(MeshView m, Species sp, double dt,
                                                          const int* __restrict__ offset, int64_t cap,
                                                          const double* __restrict__ v, int vf32,
                                                          const uint8_t* __restrict__ spec,
                                                          double* __restrict__ blk_out /* [nb][1 + nspec] */) {
  __shared__ double ecell[NCELL];
  __shared__ double cmax[64][MAX_SPEC];
  const int b = blockIdx.x, lane = threadIdx.x;
  const int ns = sp.n;
  double dx[3];
#pragma unroll
  for (int d = 0; d < 3; d++) dx[d] = (m.blk_xmax[3 * b + d] - m.blk_xmin[3 * b + d]) / NC;
  const double ldx = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
  double mx[MAX_SPEC];
  for (int s = 0; s < MAX_SPEC; s++) mx[s] = 0.0;
  for (int lc = lane; lc < NCELL; lc += 64) {
    const int key = b * NCELL + lc;
    const int p0 = offset[key], p1 = offset[key + 1];
    double e = 0.0, vm[MAX_SPEC];
    int cnt[MAX_SPEC];
    for (int s = 0; s < MAX_SPEC; s++) vm[s] = 0.0, cnt[s] = 0;
    for (int p = p0; p < p1; p++) {
      double vx, vy, vz;
      if (vf32) {
        const float* vf = (const float*)v;
        vx = vf[p], vy = vf[cap + p], vz = vf[2 * cap + p];
      } else {
        vx = v[p], vy = v[cap + p], vz = v[2 * cap + p];
      }
      const int s = spec[p];
      const double vsqr = vx * vx + vy * vy + vz * vz;
      const double ms = sp.mass[s] * sp.weight[s];
      const double sv = sqrt(vsqr) * dt;
      // dynamic species index into registers: unrolled selects
#pragma unroll
      for (int q = 0; q < MAX_SPEC; q++)
        if (q == s) vm[q] += sv, cnt[q]++;
      e += 0.5 * ms * vsqr;
    }
    ecell[lc] = e;
#pragma unroll
    for (int q = 0; q < MAX_SPEC; q++) {
      const double cc = vm[q] / (cnt[q] * ldx);
      if (mx[q] < cc) mx[q] = cc;
    }
  }
#pragma unroll
  for (int q = 0; q < MAX_SPEC; q++) cmax[lane][q] = mx[q];
  __syncthreads();
  if (lane == 0) {
    double eb = 0.0;
    for (int lc = 0; lc < NCELL; lc++) eb += ecell[lc];
    blk_out[(size_t)b * (1 + ns)] = eb;
  } else if (lane <= ns) {
    const int q = lane - 1;
    double mq = 0.0;
    for (int l = 0; l < 64; l++)
      if (mq < cmax[l][q]) mq = cmax[l][q];
    blk_out[(size_t)b * (1 + ns) + lane] = mq;
  }
}

__global__ void k_energy_cfl_total(int nb, int ns, const double* __restrict__ blk_out, double* out) {
  const int t = threadIdx.x;
  if (t == 0) {
    double e = 0.0;
    for (int b = 0; b < nb; b++) e += blk_out[(size_t)b * (1 + ns)];
    out[0] = e;
  } else if (t <= ns) {
    double mq = 0.0;
    for (int b = 0; b < nb; b++) {
      const double c = blk_out[(size_t)b * (1 + ns) + t];
      if (mq < c) mq = c;
    }
    out[t] = mq;
  }
}

}  // namespace

namespace amps {
// called by amps_gpu_create with the context's device current
int energy_scratch_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) return fail(AMPS_E_ARG, "energy_scratch_init: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return AMPS_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return AMPS_OK;
}
}  // namespace amps

// The reductions of amps_gpu_particle_energy_cfl with no allocation: the block
// partials go to the library's static scratch array and the result to the
// caller's device array out[1 + nspec] (out[0] energy, out[1 + s] the CFL of
// species s), on the context's stream, without synchronising.
extern "C" int amps_gpu_particle_energy_cfl_noalloc(amps_ctx* c, double dt, double* out) {
  if (!c || !out) return fail(AMPS_E_ARG, "amps_gpu_particle_energy_cfl_noalloc: null argument");
  if (!c->has_mesh || !c->has_species)
    return fail(AMPS_E_STATE, "amps_gpu_particle_energy_cfl_noalloc: mesh and species are required");
  const int nb = c->mv.nb, ns = c->sp.n;
  if (nb > NOALLOC_MAX_BLOCKS)
    return fail(AMPS_E_RANGE, "amps_gpu_particle_energy_cfl_noalloc: " + std::to_string(nb) +
                                  " blocks, more than the scratch array's " + std::to_string(NOALLOC_MAX_BLOCKS));
  if (c->cap > 0 && !c->offsets_valid)
    return fail(AMPS_E_STATE, "amps_gpu_particle_energy_cfl_noalloc: particles not sorted (call amps_gpu_compact)");
  double* sc = (c->device >= 0 && c->device < MAX_DEVICES) ? g_scratch_ptr[c->device] : nullptr;
  if (!sc) return fail(AMPS_E_STATE, "amps_gpu_particle_energy_cfl_noalloc: scratch not set up on this device");
  (void)hipSetDevice(c->device);
  if (c->cap == 0) {
    // no particle arrays: energy and every CFL are 0
    hipLaunchKernelGGL(k_energy_cfl_total, dim3(1), dim3(64), 0, c->stream, 0, ns, sc, out);
    HIP_TRY(hipGetLastError());
    return AMPS_OK;
  }
  const int k = c->cur;
  hipLaunchKernelGGL(k_energy_cfl_block, dim3((unsigned)nb), dim3(64), 0, c->stream, c->mv, c->sp, dt, c->d_offset,
                     c->cap, c->d_v[k], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], sc);
  hipLaunchKernelGGL(k_energy_cfl_total, dim3(1), dim3(64), 0, c->stream, nb, ns, sc, out);
  HIP_TRY(hipGetLastError());
  return AMPS_OK;
}

extern "C" int amps_gpu_particle_energy_cfl(amps_ctx* c, double dt, double* energy, double* cfl) {
  if (!c || !energy) return fail(AMPS_E_ARG, "amps_gpu_particle_energy_cfl: null argument");
  if (!c->has_mesh || !c->has_species)
    return fail(AMPS_E_STATE, "amps_gpu_particle_energy_cfl: mesh and species are required");
  (void)hipSetDevice(c->device);
  const int nb = c->mv.nb, ns = c->sp.n;
  *energy = 0.0;
  for (int s = 0; s < ns && cfl; s++) cfl[s] = 0.0;
  if (c->cap == 0) return AMPS_OK;
  if (!c->offsets_valid)
    if (int r = amps_gpu_compact(c)) return r;
  double* d_blk = nullptr;
  double* d_out = nullptr;
  HIP_TRY(hipMalloc(&d_blk, sizeof(double) * (size_t)nb * (1 + ns)));
  if (hipMalloc(&d_out, sizeof(double) * (1 + MAX_SPEC)) != hipSuccess) {
    (void)hipFree(d_blk);
    return fail(AMPS_E_HIP, "amps_gpu_particle_energy_cfl: out of device memory");
  }
  const int k = c->cur;
  hipLaunchKernelGGL(k_energy_cfl_block, dim3((unsigned)nb), dim3(64), 0, c->stream, c->mv, c->sp, dt, c->d_offset,
                     c->cap, c->d_v[k], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], d_blk);
  hipLaunchKernelGGL(k_energy_cfl_total, dim3(1), dim3(64), 0, c->stream, nb, ns, d_blk, d_out);
  double h[1 + MAX_SPEC];
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h, d_out, sizeof(double) * (1 + ns), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_blk);
  (void)hipFree(d_out);
  if (e != hipSuccess) return hip_fail(e, "amps_gpu_particle_energy_cfl");
  *energy = h[0];
  for (int s = 0; s < ns && cfl; s++) cfl[s] = h[1 + s];
  return AMPS_OK;
}
