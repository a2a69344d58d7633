// ecsim.hip — the mass-matrix term of the ECSIM field equation applied on
// the device to a corner-node E, from the J/M that the deposit left
// resident (one NODE_STRIDE row per unique node: J[3], M[243]).
//
// For node n and component c the reference accumulates, over its 27
// neighbours ii = a(di) + 3a(dj) + 9a(dk), a(0,-1,+1) = (0,1,2)
// (GetStencil, pic_field_solver_ecsim.cpp:1061-1074):
//   res += (Ex[ii] M[9ii+3c] + Ey[ii] M[9ii+3c+1] + Ez[ii] M[9ii+3c+2]) * c1 * c2
// with M indexed through MassMatrixOffsetTable (:621-639) and
// c1 = -4 pi dt theta, c2 = E_conv in UpdateRhs_Legacy (:6116-6124).  The
// same products are the mass part of every matrix element that
// UpdateMatrixElement forms (:6004-6016), so this is also the mass part of
// the operator the GMRES iterations apply.
//
// Neighbour ii of a node is the node one cell of the node's finest incident
// cell size away on the level-lmax lattice (inside one block the block's own
// cell size, as GetCornerNode(i+-1, j+-1, k+-1) of the owning block); a
// neighbour outside the domain contributes nothing (a boundary row in the
// reference's system).  The terms are summed in the reference's order onto
// the caller's partial row when one is given (accumulate), so the result is
// bit-exact with the reference's res after its mass-matrix loop.  The M rows are read once by consecutive
// lanes, the E of the neighbours through the caches: the kernel is HBM-bound
// at 1944 + 108 + 24 + 24 bytes per node.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "amps_common.h"

using namespace amps;

namespace {

// Eight nodes per workgroup, two per wavefront: lane s (0..26) of each half
// reads slot s of its node's M row (consecutive lanes, 72 bytes each) and
// the E of that slot's neighbour, and writes the three slot terms to LDS;
// then 24 lanes, one per (node, component), add the 27 terms in slot order,
// as the reference sums them.
__global__ __launch_bounds__(256) void k_ecsim_mass(const double* __restrict__ rows, const int* __restrict__ nbr,
                                                    const double* __restrict__ E, double c1, double c2,
                                                    int64_t nnodes, const double* __restrict__ seed,
                                                    double* __restrict__ out) {
  __shared__ double term[8][27][3];
  __shared__ uint64_t present[4];
  // workgroups go round-robin over the 8 XCDs: give each XCD one contiguous
  // eighth of the nodes, so the neighbours' E stay in that XCD's L2 (the grid
  // is a multiple of 8; a whole workgroup past the end leaves at once)
  const int64_t per = gridDim.x >> 3;
  const int64_t chunk = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (chunk * 8 >= nnodes) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, s = lane & 31;
  const int64_t n = chunk * 8 + wave * 2 + half;
  const bool act = n < nnodes && s < 27;
  int j = -1;
  if (act) {
    j = nbr[n * 27 + s];
    double t[3] = {0.0, 0.0, 0.0};
    if (j >= 0) {
      const double* m = rows + n * NODE_STRIDE + 3 + 9 * s;
      const double* e = E + 3 * (int64_t)j;
      const double ex = e[0], ey = e[1], ez = e[2];
#pragma unroll
      for (int c = 0; c < 3; c++) t[c] = ((ex * m[3 * c] + ey * m[3 * c + 1]) + ez * m[3 * c + 2]) * c1 * c2;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) term[wave * 2 + half][s][c] = t[c];
  }
  const uint64_t pm = __ballot(act && j >= 0);
  if (lane == 0) present[wave] = pm;
  __syncthreads();
  if (threadIdx.x < 24) {
    const int ln = threadIdx.x / 3, c = threadIdx.x - 3 * ln;
    const int64_t nn = chunk * 8 + ln;
    if (nn < nnodes) {
      const uint64_t p = present[ln >> 1] >> (32 * (ln & 1));
      // res starts from the part of the row the caller has already summed
      // (UpdateRhs_Legacy adds the 27 mass terms onto it one at a time)
      double res = seed ? seed[3 * nn + c] : 0.0;
      for (int q = 0; q < 27; q++)
        if ((p >> q) & 1) res += term[ln][q][c];
      out[3 * nn + c] = res;
    }
  }
}

}  // namespace

namespace amps {
// neighbour table [nnodes][27] on the level-lmax lattice (-1: outside)
int build_nbr(amps_ctx* c) {
  if (c->d_ecsim_nbr) return 0;
  const int nb = c->mv.nb, lmax = c->mv.lmax;
  const int64_t S = ((int64_t)1 << lmax) + 1;
  const int64_t nn = c->nnodes;
  try {
  std::vector<int64_t> h((size_t)nn, INT64_MAX);
  for (int b = 0; b < nb; b++) {
    const int64_t cs = ((int64_t)1 << (lmax - c->h_level[(size_t)b])) / NC;
    for (int q = 0; q < NCORNER; q++) {
      int64_t& hn = h[(size_t)c->h_node_of_corner[(size_t)b * NCORNER + q]];
      hn = std::min(hn, cs);
    }
  }
  const std::vector<int64_t>& keys = c->h_node_keys;
  std::vector<int> nbr((size_t)nn * 27, -1);
  static const int A[3] = {0, -1, 1};
  for (int64_t n = 0; n < nn; n++) {
    const int64_t k = keys[(size_t)n];
    const int64_t g[3] = {k % S, (k / S) % S, k / (S * S)};
    for (int s = 0; s < 27; s++) {
      const int64_t o[3] = {g[0] + A[s % 3] * h[(size_t)n], g[1] + A[(s / 3) % 3] * h[(size_t)n],
                            g[2] + A[s / 9] * h[(size_t)n]};
      if (o[0] < 0 || o[1] < 0 || o[2] < 0 || o[0] >= S || o[1] >= S || o[2] >= S) continue;
      const int64_t kk = (o[2] * S + o[1]) * S + o[0];
      const auto it = std::lower_bound(keys.begin(), keys.end(), kk);
      if (it != keys.end() && *it == kk) nbr[(size_t)n * 27 + s] = (int)(it - keys.begin());
    }
  }
  HIP_TRY(hipMalloc(&c->d_ecsim_nbr, sizeof(int) * std::max<size_t>(27, nbr.size())));
  HIP_TRY(hipMemcpy(c->d_ecsim_nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  } catch (const std::bad_alloc&) {
    return fail(AMPS_E_RANGE, "amps_gpu_ecsim_mass_apply: out of host memory for the neighbour table");
  }
  return 0;
}

// one refinement level: the reference's ECSIM solver works on a uniform
// block mesh (UpdateJMassMatrix takes the resolution of BlockTable[0],
// pic_field_solver_ecsim.cpp:3566); at a coarse/fine node the M slots of
// cells of two sizes would refer to different neighbours
bool single_level(const amps_ctx* c) {
  for (size_t b = 1; b < c->h_level.size(); b++)
    if (c->h_level[b] != c->h_level[0]) return false;
  return true;
}

}  // namespace amps

extern "C" int amps_gpu_ecsim_mass_apply(amps_ctx* c, const double* E_nodes, double c1, double c2, double* out,
                                         int32_t accumulate, int32_t reps) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (!c->has_mesh || !c->d_nodeout)
    return fail(AMPS_E_STATE, "amps_gpu_ecsim_mass_apply: run amps_gpu_update_j_mass_matrix first");
  if (!E_nodes || !out) return fail(AMPS_E_ARG, "amps_gpu_ecsim_mass_apply: null array");
  if (reps < 1 || reps > 1000) return fail(AMPS_E_RANGE, "amps_gpu_ecsim_mass_apply: reps must be in 1..1000");
  if (!single_level(c))
    return fail(AMPS_E_STATE, "amps_gpu_ecsim_mass_apply: the ECSIM field equation needs a single-level mesh "
                              "(the reference solver runs on one resolution level)");
  const int64_t nn = c->nnodes;
  if (nn == 0) return AMPS_OK;
  if (nn > ((int64_t)1 << 27)) return fail(AMPS_E_RANGE, "amps_gpu_ecsim_mass_apply: more than 2^27 nodes");
  (void)hipSetDevice(c->device);
  if (int r = build_nbr(c)) return r;
  double *dE = nullptr, *dO = nullptr, *dS = nullptr;
  int rc = AMPS_OK;
  hipError_t e = hipMalloc(&dE, sizeof(double) * 3 * nn);
  if (e == hipSuccess) e = hipMalloc(&dO, sizeof(double) * 3 * nn);
  if (e == hipSuccess && accumulate) e = hipMalloc(&dS, sizeof(double) * 3 * nn);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dE, E_nodes, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess && dS) e = hipMemcpyAsync(dS, out, sizeof(double) * 3 * nn, hipMemcpyHostToDevice, c->stream);
  // 4 wavefronts of 2 nodes per workgroup, the workgroup count rounded up to 8
  const dim3 grid((unsigned)((((nn + 7) / 8) + 7) / 8 * 8));
  for (int r = 0; r < reps && e == hipSuccess; r++) {
    tbegin(c);
    hipLaunchKernelGGL(k_ecsim_mass, grid, dim3(256), 0, c->stream, c->d_nodeout, c->d_ecsim_nbr, dE, c1, c2, nn,
                       dS, dO);
    e = hipGetLastError();
    tend(c, AMPS_K_ECSIM);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dO, sizeof(double) * 3 * nn, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) rc = hip_fail(e, "amps_gpu_ecsim_mass_apply");
  (void)hipFree(dE);
  (void)hipFree(dO);
  (void)hipFree(dS);
  return rc;
}
