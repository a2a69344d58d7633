// ecsim_field.hip — the implicit ECSIM field equation on the device, over the
// J and M that the deposit leaves resident (one NODE_STRIDE row per node).
//
// The unknown is dE = (E^{n+theta} - E^n) E_conv at the corner nodes of a
// single-level mesh.  Per node n and component iVar (GetStencil_Legacy,
// src/pic/pic_field_solver_ecsim.cpp:740-1228; UpdateMatrixElement
// :6004-6016):
//   (A x)[n][iVar] = sum over iVarIndex, then over the 27 neighbour slots s of
//                    (K[iVar][iVarIndex][s] + M_n[9s + 3iVar + iVarIndex] 4 pi dt theta)
//                    x[nbr(n, s)][iVarIndex]
// with K = identity - coeff^2 Laplacian + coeff coeff grad-div (the
// 27-point stencils the reference composes from edge averages, :7185-7300,
// coeff = c dt theta / dx); neighbours outside the domain or on its boundary
// are dropped (:1032), and a corner on the domain boundary is a row of the
// identity with a zero right-hand side (:782-799).  The right-hand side is
// UpdateRhs_Legacy (:6040-6137): the neighbours' E^n times the opposite
// Laplacian / grad-div coefficients, the mass term times -4 pi dt theta, the
// node's J times -4 pi dt theta and the 16 cell-centred B terms of curl B
// (:1234-1375).  The coefficient tables are built on the host in the
// reference's loop order and every sum runs in the reference's order, so the
// operator, the right-hand side and the B update are bit-exact with the
// oracle's restatement.
//
// The system is solved by restarted GMRES (modified Gram-Schmidt, zero
// initial guess as SetInitialGuess :6296-6301, stop at ||r|| <= tol ||b||):
// vectors and the Krylov basis stay on the device, inner products are
// fixed-order two-stage reductions (the same bits on every run), and only the
// Hessenberg column of each step comes to the host for the Givens update.
// ProcessFinalSolution (:6311-6318), UpdateE (:5909-5960) and UpdateB
// (:5160-5270) complete a time step.  The reference's solver is SWMF's
// linear_solver_wrapper("GMRES") (srcInterface/LinearSystemCornerNode.h:3282),
// not part of the tree: its iteration count is not reproduced, its converged
// solution is (to the tolerance).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <array>
#include <vector>

#include "amps_common.h"

using namespace amps;

namespace {

constexpr double PI_REF = 3.14159265358979323846;  // Pi of src/general/constants.h
constexpr int RED_BLOCKS = 1024;
constexpr int RED_THREADS = 256;

// curl-B entries of UpdateRhs_Legacy per row component iVar: (coefficient
// axis, sign, ci, cj, ck) of the cell node + (ci, cj, ck); the first 8 take
// B component (iVar + 2) % 3, the next 8 (iVar + 4) % 3 (:1234-1375, groups
// of 4 in the reference's ii-outer / jj-inner order)
__constant__ int c_curlb[3][16][5] = {
    {{1, 1, -1, 0, -1}, {1, 1, -1, 0, 0}, {1, 1, 0, 0, -1}, {1, 1, 0, 0, 0}, {1, -1, -1, -1, -1},
     {1, -1, -1, -1, 0}, {1, -1, 0, -1, -1}, {1, -1, 0, -1, 0}, {2, -1, -1, -1, 0}, {2, -1, -1, 0, 0},
     {2, -1, 0, -1, 0}, {2, -1, 0, 0, 0}, {2, 1, -1, -1, -1}, {2, 1, -1, 0, -1}, {2, 1, 0, -1, -1},
     {2, 1, 0, 0, -1}},
    {{2, 1, -1, -1, 0}, {2, 1, -1, 0, 0}, {2, 1, 0, -1, 0}, {2, 1, 0, 0, 0}, {2, -1, -1, -1, -1},
     {2, -1, -1, 0, -1}, {2, -1, 0, -1, -1}, {2, -1, 0, 0, -1}, {0, -1, 0, -1, -1}, {0, -1, 0, 0, -1},
     {0, -1, 0, -1, 0}, {0, -1, 0, 0, 0}, {0, 1, -1, -1, -1}, {0, 1, -1, 0, -1}, {0, 1, -1, -1, 0},
     {0, 1, -1, 0, 0}},
    {{0, 1, 0, -1, -1}, {0, 1, 0, 0, -1}, {0, 1, 0, -1, 0}, {0, 1, 0, 0, 0}, {0, -1, -1, -1, -1},
     {0, -1, -1, 0, -1}, {0, -1, -1, -1, 0}, {0, -1, -1, 0, 0}, {1, -1, -1, 0, -1}, {1, -1, 0, 0, -1},
     {1, -1, -1, 0, 0}, {1, -1, 0, 0, 0}, {1, 1, -1, -1, -1}, {1, 1, 0, -1, -1}, {1, 1, -1, -1, 0},
     {1, 1, 0, -1, 0}}};

struct Coefs {
  double K[3][3][27];  // operator constants [iVar][iVarIndex][slot]
  double R[3][3][27];  // E^n coefficients of the right-hand side
  double coeff[3];     // c dt theta / dx
};

int slot_of(int di, int dj, int dk) {
  auto a = [](int d) { return d == 0 ? 0 : (d == -1 ? 1 : 2); };
  return a(di) + 3 * a(dj) + 9 * a(dk);
}

// The Laplacian and grad-div stencils as the reference composes them; every
// entry is a dyadic fraction, so the merge order does not change a bit.
using St = std::map<std::array<int, 3>, double>;
void st_add(St& st, const St& o, int di, int dj, int dk, double f) {
  for (const auto& e : o) st[{e.first[0] + di, e.first[1] + dj, e.first[2] + dk}] += f * e.second;
}

void stencils(St lap[3], St gd[3][3]) {
  auto edge = [](std::vector<int> ir, std::vector<int> jr, std::vector<int> kr) {
    St s;
    for (int i : ir)
      for (int j : jr)
        for (int k : kr) s[{i, j, k}] = 0.25;
    return s;
  };
  const St xe0 = edge({0}, {-1, 0}, {-1, 0}), xe4 = edge({-1, 0}, {0}, {-1, 0}), xe8 = edge({-1, 0}, {-1, 0}, {0});
  St f0;
  st_add(f0, xe0, 1, 0, 0, 1.0);
  st_add(f0, xe0, 0, 0, 0, -2.0);
  st_add(f0, xe0, -1, 0, 0, 1.0);
  St dxx;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) st_add(dxx, f0, 0, i, j, 0.25);
  St dyy, dzz;
  for (const auto& e : dxx) {
    dyy[{e.first[1], e.first[0], e.first[2]}] = e.second;
    dzz[{e.first[2], e.first[1], e.first[0]}] = e.second;
  }
  St dy0 = xe4, dz0 = xe8;
  st_add(dy0, xe4, 0, -1, 0, -1.0);
  st_add(dz0, xe8, 0, 0, -1, -1.0);
  St dxy, dxz, dyz;
  for (int l = 0; l < 2; l++)
    for (int m = 0; m < 2; m++) {
      st_add(dxy, dy0, 1, l, m, 0.25);
      st_add(dxy, dy0, 0, l, m, -0.25);
      st_add(dxz, dz0, 1, l, m, 0.25);
      st_add(dxz, dz0, 0, l, m, -0.25);
      st_add(dyz, dz0, l, 1, m, 0.25);
      st_add(dyz, dz0, l, 0, m, -0.25);
    }
  lap[0] = dxx, lap[1] = dyy, lap[2] = dzz;
  gd[0][0] = dxx, gd[0][1] = dxy, gd[0][2] = dxz;
  gd[1][0] = dxy, gd[1][1] = dyy, gd[1][2] = dyz;
  gd[2][0] = dxz, gd[2][1] = dyz, gd[2][2] = dzz;
}

// the row tables in the reference's accumulation order (GetStencil_Legacy
// :873-910 for the matrix, :1152-1185 for the right-hand side)
Coefs coefficients(const double dx[3], const amps_ecsim_params& p) {
  Coefs C{};
  double coeffSqr[3];
  for (int d = 0; d < 3; d++) {
    const double dxd = dx[d] * p.length_conv;
    C.coeff[d] = p.cDt / dxd * p.theta;
    coeffSqr[d] = C.coeff[d] * C.coeff[d];
  }
  St lap[3], gd[3][3];
  stencils(lap, gd);
  for (int iVar = 0; iVar < 3; iVar++) {
    for (int idim = 0; idim < 3; idim++)
      for (const auto& e : lap[idim]) {
        const int s = slot_of(e.first[0], e.first[1], e.first[2]);
        C.K[iVar][iVar][s] -= e.second * coeffSqr[idim];
        C.R[iVar][iVar][s] += e.second * coeffSqr[idim];
      }
    C.K[iVar][iVar][0] += 1;
    for (int iv = 0; iv < 3; iv++)
      for (const auto& e : gd[iVar][iv]) {
        const int s = slot_of(e.first[0], e.first[1], e.first[2]);
        C.K[iVar][iv][s] += (1 - 0.0) * e.second * C.coeff[iVar] * C.coeff[iv];
        C.R[iVar][iv][s] -= (1 - 0.0) * e.second * C.coeff[iVar] * C.coeff[iv];
      }
  }
  return C;
}

// ---------------------------------------------------------------------------
// kernels: one thread per (node, component) row
// ---------------------------------------------------------------------------
// Eight nodes per workgroup, two per wavefront (as the mass term of
// ecsim.hip): lane s (0..26) of each half reads slot s of its node's M row (72
// contiguous bytes, consecutive lanes) and its neighbour's x, and writes the
// 9 products (K + M 4 pi dt theta) x of the slot to LDS; then 24 lanes, one
// per (node, component) row, add the row's 81 products in the row's element
// order (iVarIndex outer, slot inner), skipping the dropped neighbours.
__global__ __launch_bounds__(256) void k_ecsim_op(int64_t nn, const double* __restrict__ rows,
                                                  const int* __restrict__ nbr, const uint8_t* __restrict__ bnd,
                                                  const Coefs* __restrict__ C, double fpdt,
                                                  const double* __restrict__ x, double* __restrict__ y) {
  __shared__ double term[8][3][3][27];
  __shared__ uint64_t present[4];
  const int64_t per = gridDim.x >> 3;
  const int64_t chunk = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (chunk * 8 >= nn) return;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int half = lane >> 5, s = lane & 31;
  const int ln = wave * 2 + half;
  const int64_t n = chunk * 8 + ln;
  const bool act = n < nn && s < 27;
  bool keep = false;
  if (act) {
    const int j = nbr[n * 27 + s];
    keep = j >= 0 && !bnd[j];
    if (keep) {
      const double* m = rows + n * NODE_STRIDE + 3 + 9 * s;
      const double* xj = x + 3 * (int64_t)j;
      const double xv[3] = {xj[0], xj[1], xj[2]};
#pragma unroll
      for (int iVar = 0; iVar < 3; iVar++)
#pragma unroll
        for (int iv = 0; iv < 3; iv++) {
          const double val = C->K[iVar][iv][s] + m[3 * iVar + iv] * fpdt;
          term[ln][iVar][iv][s] = val * xv[iv];
        }
    }
  }
  const uint64_t pm = __ballot(keep);
  if (lane == 0) present[wave] = pm;
  __syncthreads();
  if (threadIdx.x < 24) {
    const int r = threadIdx.x / 3, iVar = threadIdx.x - 3 * r;
    const int64_t nr = chunk * 8 + r;
    if (nr < nn) {
      double out;
      if (bnd[nr]) {
        out = x[3 * nr + iVar];
      } else {
        const uint64_t p = present[r >> 1] >> (32 * (r & 1));
        double acc = 0.0;
        for (int iv = 0; iv < 3; iv++)
          for (int q = 0; q < 27; q++)
            if ((p >> q) & 1) acc = acc + term[r][iVar][iv][q];
        out = acc;
      }
      y[3 * nr + iVar] = out;
    }
  }
}

__global__ __launch_bounds__(256) void k_ecsim_rhs(int64_t nn, const double* __restrict__ rows,
                                                   const int* __restrict__ nbr, const uint8_t* __restrict__ bnd,
                                                   const int* __restrict__ cells8, const Coefs* __restrict__ C,
                                                   double dtTotal, double theta, double E_conv, double B_conv,
                                                   const double* __restrict__ En, const double* __restrict__ Bc,
                                                   double* __restrict__ out) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= 3 * nn) return;
  const int64_t n = r / 3;
  const int iVar = (int)(r - 3 * n);
  if (bnd[n]) {
    out[r] = 0.0;
    return;
  }
  const double fpdt = 4 * PI_REF * dtTotal * theta;
  const int* nb = nbr + n * 27;
  double res = 0.0;
  for (int s = 0; s < 27; s++) {
    const double* e = En + 3 * (int64_t)nb[s];
    res += (e[0] * C->R[iVar][0][s] + e[1] * C->R[iVar][1][s] + e[2] * C->R[iVar][2][s]) * E_conv;
  }
  const double* row = rows + n * NODE_STRIDE;
  const double* M = row + 3;
  for (int s = 0; s < 27; s++) {
    const double* e = En + 3 * (int64_t)nb[s];
    const double* m = M + 9 * s + 3 * iVar;
    res += (e[0] * m[0] + e[1] * m[1] + e[2] * m[2]) * (-fpdt) * E_conv;
  }
  res += row[iVar] * (-4 * PI_REF * dtTotal * theta);
  const int* cl = cells8 + n * 8;
  for (int t = 0; t < 16; t++) {
    const int* e = c_curlb[iVar][t];
    const int comp = t < 8 ? (iVar + 2) % 3 : (iVar + 4) % 3;
    const int cell = cl[(e[2] + 1) + 2 * (e[3] + 1) + 4 * (e[4] + 1)];
    const double c4 = 0.25 * C->coeff[e[0]];
    res += Bc[3 * (int64_t)cell + comp] * (e[1] > 0 ? c4 : -c4) * B_conv;
  }
  out[r] = res;
}

// E^{n+theta} = dE / E_conv + E^n; E^{n+1} = (E^{n+theta} - (1 - theta) E^n) / theta
__global__ __launch_bounds__(256) void k_ecsim_update_e(int64_t n3, const double* __restrict__ En,
                                                        const double* __restrict__ dE, double theta, double E_conv,
                                                        double* __restrict__ Eth, double* __restrict__ E1) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n3) return;
  const double th = dE[i] / E_conv + En[i];
  Eth[i] = th;
  E1[i] = (th - (1.0 - theta) * En[i]) / theta;
}

// UpdateB for every cell: one thread per cell, the 2x2 face averages of the
// corner E^{n+theta} (coefficients coeff4 = c dt / dx / 4, no theta)
__global__ __launch_bounds__(256) void k_ecsim_update_b(int64_t ncell, const int* __restrict__ noc,
                                                        const double* __restrict__ Eth,
                                                        const double* __restrict__ Bc, double c40, double c41,
                                                        double c42, double E_conv, double B_conv,
                                                        double* __restrict__ Bout) {
  const int64_t cid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (cid >= ncell) return;
  const int64_t b = cid / NCELL;
  const int lc = (int)(cid - b * NCELL);
  const int i = lc % NC, j = (lc / NC) % NC, k = lc / (NC * NC);
  double Ex[2][2][2], Ey[2][2][2], Ez[2][2][2];
  for (int kk = 0; kk < 2; kk++)
    for (int jj = 0; jj < 2; jj++)
      for (int ii = 0; ii < 2; ii++) {
        const int64_t nd = noc[b * NCORNER + (i + ii) + NN * ((j + jj) + NN * (k + kk))];
        Ex[ii][jj][kk] = Eth[3 * nd] * E_conv;
        Ey[ii][jj][kk] = Eth[3 * nd + 1] * E_conv;
        Ez[ii][jj][kk] = Eth[3 * nd + 2] * E_conv;
      }
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
  for (int p = 0; p < 2; p++)
    for (int q = 0; q < 2; q++) t0 += -c41 * (Ez[p][1][q] - Ez[p][0][q]) + c42 * (Ey[p][q][1] - Ey[p][q][0]);
  for (int p = 0; p < 2; p++)
    for (int q = 0; q < 2; q++) t1 += -c42 * (Ex[p][q][1] - Ex[p][q][0]) + c40 * (Ez[1][p][q] - Ez[0][p][q]);
  for (int p = 0; p < 2; p++)
    for (int q = 0; q < 2; q++) t2 += -c40 * (Ey[1][p][q] - Ey[0][p][q]) + c41 * (Ex[p][1][q] - Ex[p][0][q]);
  Bout[3 * cid] = Bc[3 * cid] + t0 / B_conv;
  Bout[3 * cid + 1] = Bc[3 * cid + 1] + t1 / B_conv;
  Bout[3 * cid + 2] = Bc[3 * cid + 2] + t2 / B_conv;
}

// ---------------------------------------------------------------------------
// Krylov helpers: fixed-order reductions and vector updates with device
// scalars (no host round trip inside the Gram-Schmidt loop)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double s[RED_THREADS];
  s[threadIdx.x] = v;
  __syncthreads();
  for (int o = RED_THREADS / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  return s[0];
}

__global__ __launch_bounds__(RED_THREADS) void k_dot_partial(int64_t n, const double* __restrict__ a,
                                                             const double* __restrict__ b, double* __restrict__ part) {
  double v = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)RED_THREADS + threadIdx.x; i < n; i += (int64_t)RED_BLOCKS * RED_THREADS)
    v += a[i] * b[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// out = sum of the partials; sqrt_out: out = sqrt of it
__global__ __launch_bounds__(RED_THREADS) void k_dot_final(const double* __restrict__ part, double* __restrict__ out,
                                                           int sqrt_out) {
  double v = 0.0;
  for (int i = threadIdx.x; i < RED_BLOCKS; i += RED_THREADS) v += part[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) *out = sqrt_out ? sqrt(t) : t;
}

// y -= (*h) x
__global__ __launch_bounds__(256) void k_axpy_neg(int64_t n, const double* __restrict__ h,
                                                  const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] -= (*h) * x[i];
}

// y = x / (*h)
__global__ __launch_bounds__(256) void k_scale_inv(int64_t n, const double* __restrict__ h,
                                                   const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i] / (*h);
}

// y = a - b
__global__ __launch_bounds__(256) void k_sub(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                             double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = a[i] - b[i];
}

// x += sum_j c[j] V[j]
__global__ __launch_bounds__(256) void k_update_x(int64_t n, int k, const double* __restrict__ c,
                                                  const double* __restrict__ V, double* __restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = x[i];
  for (int j = 0; j < k; j++) v += c[j] * V[(int64_t)j * n + i];
  x[i] = v;
}

__global__ __launch_bounds__(256) void k_fill(int64_t n, double* __restrict__ y, double v) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

dim3 grid_of(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// ---------------------------------------------------------------------------
// host: tables, checks, work space
// ---------------------------------------------------------------------------
int check_ready(amps_ctx* c, const amps_ecsim_params* p, const char* who) {
  if (!c || !p) return fail(AMPS_E_ARG, std::string(who) + ": null argument");
  if (!c->has_mesh || !c->d_nodeout)
    return fail(AMPS_E_STATE, std::string(who) + ": run amps_gpu_update_j_mass_matrix first (J and M)");
  if (!single_level(c))
    return fail(AMPS_E_STATE, std::string(who) + ": the ECSIM field equation needs a single-level mesh "
                                                 "(the reference solver runs on one resolution level)");
  if (!(p->theta > 0.0) || !(p->E_conv != 0.0) || !(p->B_conv != 0.0) || !(p->length_conv > 0.0))
    return fail(AMPS_E_ARG, std::string(who) + ": theta, E_conv, B_conv and length_conv must be nonzero");
  if (c->nnodes > ((int64_t)1 << 27)) return fail(AMPS_E_RANGE, std::string(who) + ": more than 2^27 nodes");
  return 0;
}

int build_tables(amps_ctx* c) {
  if (int r = build_nbr(c)) return r;
  if (c->d_ecsim_bnd) return 0;
  const int lmax = c->mv.lmax;
  const int64_t S = ((int64_t)1 << lmax) + 1;
  const int64_t nn = c->nnodes;
  try {
    std::vector<uint8_t> bnd((size_t)nn, 0);
    for (int64_t n = 0; n < nn; n++) {
      const int64_t k = c->h_node_keys[(size_t)n];
      const int64_t g[3] = {k % S, (k / S) % S, k / (S * S)};
      for (int d = 0; d < 3; d++)
        if (g[d] == 0 || g[d] == S - 1) bnd[(size_t)n] = 1;
    }
    // the cell whose lower corner is each node, then the 8 cells around
    std::vector<int> lower((size_t)nn, -1);
    const int nb = c->mv.nb;
    for (int b = 0; b < nb; b++)
      for (int kk = 0; kk < NC; kk++)
        for (int jj = 0; jj < NC; jj++)
          for (int ii = 0; ii < NC; ii++) {
            const int nd = c->h_node_of_corner[(size_t)b * NCORNER + ii + NN * (jj + NN * kk)];
            lower[(size_t)nd] = b * NCELL + ii + NC * (jj + NC * kk);
          }
    std::vector<int> nbr((size_t)nn * 27);
    HIP_TRY(hipMemcpy(nbr.data(), c->d_ecsim_nbr, sizeof(int) * nbr.size(), hipMemcpyDeviceToHost));
    std::vector<int> cells8((size_t)nn * 8, -1);
    for (int64_t n = 0; n < nn; n++)
      for (int ck = -1; ck <= 0; ck++)
        for (int cj = -1; cj <= 0; cj++)
          for (int ci = -1; ci <= 0; ci++) {
            const int m = nbr[(size_t)n * 27 + slot_of(ci, cj, ck)];
            // a boundary row reads no cell; a missing one is read as cell 0 and never used
            cells8[(size_t)n * 8 + (ci + 1) + 2 * (cj + 1) + 4 * (ck + 1)] =
                m >= 0 && lower[(size_t)m] >= 0 ? lower[(size_t)m] : (bnd[(size_t)n] ? 0 : -1);
          }
    for (int64_t n = 0; n < nn; n++)
      if (!bnd[(size_t)n])
        for (int q = 0; q < 8; q++)
          if (cells8[(size_t)n * 8 + q] < 0)
            return fail(AMPS_E_STATE, "ecsim: an interior corner without its 8 cells (the mesh does not cover the box)");
    HIP_TRY(hipMalloc(&c->d_ecsim_bnd, std::max<size_t>(1, bnd.size())));
    HIP_TRY(hipMemcpy(c->d_ecsim_bnd, bnd.data(), bnd.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_ecsim_cells8, sizeof(int) * std::max<size_t>(8, cells8.size())));
    HIP_TRY(hipMemcpy(c->d_ecsim_cells8, cells8.data(), sizeof(int) * cells8.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_ecsim_noc, sizeof(int) * std::max<size_t>(1, c->h_node_of_corner.size())));
    HIP_TRY(hipMemcpy(c->d_ecsim_noc, c->h_node_of_corner.data(), sizeof(int) * c->h_node_of_corner.size(),
                      hipMemcpyHostToDevice));
  } catch (const std::bad_alloc&) {
    return fail(AMPS_E_RANGE, "ecsim: out of host memory for the field-solver tables");
  }
  return 0;
}

int ensure_work(amps_ctx* c, size_t len) {
  if (c->ecsim_work_len >= len) return 0;
  (void)hipFree(c->d_ecsim_work);
  c->d_ecsim_work = nullptr;
  c->ecsim_work_len = 0;
  HIP_TRY(hipMalloc(&c->d_ecsim_work, sizeof(double) * len));
  c->ecsim_work_len = len;
  return 0;
}

void cell_size(const amps_ctx* c, double dx[3]) {
  for (int d = 0; d < 3; d++) dx[d] = (c->h_xmax[(size_t)d] - c->h_xmin[(size_t)d]) / NC;
}

// a device copy of the row tables (the tail of the work space)
int upload_coefs(amps_ctx* c, const amps_ecsim_params* p, Coefs* dC) {
  double dx[3];
  cell_size(c, dx);
  const Coefs C = coefficients(dx, *p);
  HIP_TRY(hipMemcpyAsync(dC, &C, sizeof(Coefs), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// layout of the work space (doubles): En, Bc or rhs, x, y/w, dE, Eth, E1, Bout,
// reductions, scalars, coefficient tables, then the Krylov basis
struct Work {
  double *En, *Bc, *rhs, *x, *w, *Eth, *E1, *Bout, *part, *scal, *hcol, *yv;
  Coefs* C;
  double* V;
};
size_t work_len(int64_t nn, int64_t ncell, int restart) {
  const size_t n3 = 3 * (size_t)nn, c3 = 3 * (size_t)ncell;
  const size_t coef = (sizeof(Coefs) + 7) / 8;
  return n3 * 6 + c3 * 2 + RED_BLOCKS + 8 + 2 * ((size_t)restart + 2) + coef + n3 * ((size_t)restart + 1);
}
Work carve(amps_ctx* c, int restart) {
  const size_t n3 = 3 * (size_t)c->nnodes, c3 = 3 * (size_t)c->mv.nb * NCELL;
  double* p = c->d_ecsim_work;
  Work w;
  w.En = p, p += n3;
  w.rhs = p, p += n3;
  w.x = p, p += n3;
  w.w = p, p += n3;
  w.Eth = p, p += n3;
  w.E1 = p, p += n3;
  w.Bc = p, p += c3;
  w.Bout = p, p += c3;
  w.part = p, p += RED_BLOCKS;
  w.scal = p, p += 8;
  w.hcol = p, p += restart + 2;
  w.yv = p, p += restart + 2;
  w.C = reinterpret_cast<Coefs*>(p), p += (sizeof(Coefs) + 7) / 8;
  w.V = p;
  return w;
}

void launch_op(amps_ctx* c, const Work& W, double fpdt, const double* x, double* y) {
  const int64_t nn = c->nnodes;
  // 8 nodes per workgroup, the workgroup count rounded up to 8 (XCD chunks)
  const dim3 grid((unsigned)((((nn + 7) / 8) + 7) / 8 * 8));
  hipLaunchKernelGGL(k_ecsim_op, grid, dim3(256), 0, c->stream, nn, c->d_nodeout, c->d_ecsim_nbr, c->d_ecsim_bnd,
                     W.C, fpdt, x, y);
}

void launch_dot(amps_ctx* c, const Work& W, const double* a, const double* b, double* out, bool sqrt_out) {
  const int64_t n = 3 * c->nnodes;
  hipLaunchKernelGGL(k_dot_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, n, a, b, W.part);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(RED_THREADS), 0, c->stream, W.part, out, sqrt_out ? 1 : 0);
}

// restarted GMRES on the device: W.rhs -> W.x
int gmres(amps_ctx* c, const Work& W, double fpdt, double tol, int max_iter, int restart, int* iters,
          double* relres) {
  const int64_t n = 3 * c->nnodes;
  const dim3 g = grid_of(n);
  hipLaunchKernelGGL(k_fill, g, dim3(256), 0, c->stream, n, W.x, 0.0);
  double bnorm = 0.0;
  launch_dot(c, W, W.rhs, W.rhs, W.scal, true);
  HIP_TRY(hipMemcpyAsync(&bnorm, W.scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  int it = 0;
  *iters = 0;
  *relres = 0.0;
  if (bnorm == 0.0) return 0;
  std::vector<double> H((size_t)(restart + 1) * restart, 0.0), cs(restart), sn(restart), gv(restart + 1);
  std::vector<double> col(restart + 2);
  double res = bnorm;
  while (it < max_iter) {
    // r = b - A x into V[0], beta = ||r||
    double* V0 = W.V;
    launch_op(c, W, fpdt, W.x, W.w);
    hipLaunchKernelGGL(k_sub, g, dim3(256), 0, c->stream, n, W.rhs, W.w, V0);
    launch_dot(c, W, V0, V0, W.scal, true);
    double beta = 0.0;
    HIP_TRY(hipMemcpyAsync(&beta, W.scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    res = beta;
    if (beta <= tol * bnorm) break;
    hipLaunchKernelGGL(k_scale_inv, g, dim3(256), 0, c->stream, n, W.scal, V0, V0);
    std::fill(gv.begin(), gv.end(), 0.0);
    gv[0] = beta;
    int k = 0;
    for (; k < restart && it < max_iter; k++) {
      double* Vk = W.V + (int64_t)k * n;
      double* Vk1 = W.V + (int64_t)(k + 1) * n;
      launch_op(c, W, fpdt, Vk, W.w);
      for (int j = 0; j <= k; j++) {
        launch_dot(c, W, W.w, W.V + (int64_t)j * n, W.hcol + j, false);
        hipLaunchKernelGGL(k_axpy_neg, g, dim3(256), 0, c->stream, n, W.hcol + j, W.V + (int64_t)j * n, W.w);
      }
      launch_dot(c, W, W.w, W.w, W.hcol + k + 1, true);
      hipLaunchKernelGGL(k_scale_inv, g, dim3(256), 0, c->stream, n, W.hcol + k + 1, W.w, Vk1);
      HIP_TRY(hipMemcpyAsync(col.data(), W.hcol, sizeof(double) * (k + 2), hipMemcpyDeviceToHost, c->stream));
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(c->stream));
      it++;
      // Givens rotations of the new column
      for (int j = 0; j < k; j++) {
        const double t = cs[j] * col[j] + sn[j] * col[j + 1];
        col[j + 1] = -sn[j] * col[j] + cs[j] * col[j + 1];
        col[j] = t;
      }
      const double hk = col[k], hk1 = col[k + 1];
      const double rr = std::hypot(hk, hk1);
      cs[k] = rr > 0 ? hk / rr : 1.0;
      sn[k] = rr > 0 ? hk1 / rr : 0.0;
      col[k] = rr;
      col[k + 1] = 0.0;
      gv[k + 1] = -sn[k] * gv[k];
      gv[k] = cs[k] * gv[k];
      for (int j = 0; j <= k; j++) H[(size_t)j * restart + k] = col[j];
      res = std::fabs(gv[k + 1]);
      if (res <= tol * bnorm || hk1 == 0.0) {
        k++;
        break;
      }
    }
    // y = H^{-1} g over the k columns, then x += V y
    std::vector<double> y(k);
    for (int j = k - 1; j >= 0; j--) {
      double s = gv[j];
      for (int l = j + 1; l < k; l++) s -= H[(size_t)j * restart + l] * y[l];
      y[j] = s / H[(size_t)j * restart + j];
    }
    if (k > 0) {
      HIP_TRY(hipMemcpyAsync(W.yv, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->stream));
      hipLaunchKernelGGL(k_update_x, g, dim3(256), 0, c->stream, n, k, W.yv, W.V, W.x);
    }
    if (res <= tol * bnorm) break;
  }
  // the true residual of the returned solution
  launch_op(c, W, fpdt, W.x, W.w);
  hipLaunchKernelGGL(k_sub, g, dim3(256), 0, c->stream, n, W.rhs, W.w, W.w);
  launch_dot(c, W, W.w, W.w, W.scal, true);
  double rn = 0.0;
  HIP_TRY(hipMemcpyAsync(&rn, W.scal, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *iters = it;
  *relres = rn / bnorm;
  return 0;
}

int prepare(amps_ctx* c, const amps_ecsim_params* p, int restart, Work* W, const char* who) {
  if (int r = check_ready(c, p, who)) return r;
  (void)hipSetDevice(c->device);
  if (int r = build_tables(c)) return r;
  if (int r = ensure_work(c, work_len(c->nnodes, (int64_t)c->mv.nb * NCELL, restart))) return r;
  *W = carve(c, restart);
  return upload_coefs(c, p, W->C);
}

double fpdt_of(const amps_ecsim_params* p) { return 4 * PI_REF * p->dt * p->theta; }

}  // namespace

namespace amps {
void ecsim_field_free(amps_ctx* c) {
  (void)hipFree(c->d_ecsim_bnd);
  (void)hipFree(c->d_ecsim_cells8);
  (void)hipFree(c->d_ecsim_noc);
  (void)hipFree(c->d_ecsim_work);
  c->d_ecsim_bnd = nullptr;
  c->d_ecsim_cells8 = nullptr;
  c->d_ecsim_noc = nullptr;
  c->d_ecsim_work = nullptr;
  c->ecsim_work_len = 0;
}
}  // namespace amps

extern "C" int amps_gpu_ecsim_operator(amps_ctx* c, const amps_ecsim_params* p, const double* x, double* y,
                                       int32_t reps) {
  if (!x || !y) return fail(AMPS_E_ARG, "amps_gpu_ecsim_operator: null array");
  if (reps < 1 || reps > 1000) return fail(AMPS_E_RANGE, "amps_gpu_ecsim_operator: reps must be in 1..1000");
  Work W;
  if (int r = prepare(c, p, 0, &W, "amps_gpu_ecsim_operator")) return r;
  const size_t n3 = 3 * (size_t)c->nnodes;
  HIP_TRY(hipMemcpyAsync(W.x, x, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  for (int i = 0; i < reps; i++) {
    tbegin(c);
    launch_op(c, W, fpdt_of(p), W.x, W.w);
    tend(c, AMPS_K_ECSIM);
  }
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(y, W.w, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int amps_gpu_ecsim_rhs(amps_ctx* c, const amps_ecsim_params* p, const double* E_n, const double* B_c,
                                  double* rhs) {
  if (!E_n || !B_c || !rhs) return fail(AMPS_E_ARG, "amps_gpu_ecsim_rhs: null array");
  Work W;
  if (int r = prepare(c, p, 0, &W, "amps_gpu_ecsim_rhs")) return r;
  const int64_t nn = c->nnodes;
  const size_t n3 = 3 * (size_t)nn, c3 = 3 * (size_t)c->mv.nb * NCELL;
  HIP_TRY(hipMemcpyAsync(W.En, E_n, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(W.Bc, B_c, sizeof(double) * c3, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_ecsim_rhs, grid_of(3 * nn), dim3(256), 0, c->stream, nn, c->d_nodeout, c->d_ecsim_nbr,
                     c->d_ecsim_bnd, c->d_ecsim_cells8, W.C, p->dt, p->theta, p->E_conv, p->B_conv, W.En, W.Bc,
                     W.rhs);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(rhs, W.rhs, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int amps_gpu_ecsim_solve(amps_ctx* c, const amps_ecsim_params* p, const double* rhs, double tol,
                                    int32_t max_iter, int32_t restart, double* dE, int32_t* iters,
                                    double* rel_residual) {
  if (!rhs || !dE || !iters || !rel_residual) return fail(AMPS_E_ARG, "amps_gpu_ecsim_solve: null argument");
  if (!(tol > 0.0) || max_iter < 1 || restart < 1 || restart > 400)
    return fail(AMPS_E_ARG, "amps_gpu_ecsim_solve: tol > 0, max_iter >= 1 and restart in 1..400 are required");
  Work W;
  if (int r = prepare(c, p, restart, &W, "amps_gpu_ecsim_solve")) return r;
  const size_t n3 = 3 * (size_t)c->nnodes;
  HIP_TRY(hipMemcpyAsync(W.rhs, rhs, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  int it = 0;
  double rr = 0.0;
  if (int r = gmres(c, W, fpdt_of(p), tol, max_iter, restart, &it, &rr)) return r;
  HIP_TRY(hipMemcpyAsync(dE, W.x, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *iters = it;
  *rel_residual = rr;
  return AMPS_OK;
}

extern "C" int amps_gpu_ecsim_update_b(amps_ctx* c, const amps_ecsim_params* p, const double* E_theta,
                                       const double* B_c, double* B_out) {
  if (!E_theta || !B_c || !B_out) return fail(AMPS_E_ARG, "amps_gpu_ecsim_update_b: null array");
  Work W;
  if (int r = prepare(c, p, 0, &W, "amps_gpu_ecsim_update_b")) return r;
  const size_t n3 = 3 * (size_t)c->nnodes;
  const int64_t ncell = (int64_t)c->mv.nb * NCELL;
  HIP_TRY(hipMemcpyAsync(W.Eth, E_theta, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(W.Bc, B_c, sizeof(double) * 3 * ncell, hipMemcpyHostToDevice, c->stream));
  double dx[3];
  cell_size(c, dx);
  double c4[3];
  for (int d = 0; d < 3; d++) c4[d] = 0.25 * (p->cDt / (dx[d] * p->length_conv));
  hipLaunchKernelGGL(k_ecsim_update_b, grid_of(ncell), dim3(256), 0, c->stream, ncell, c->d_ecsim_noc, W.Eth,
                     W.Bc, c4[0], c4[1], c4[2], p->E_conv, p->B_conv, W.Bout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(B_out, W.Bout, sizeof(double) * 3 * ncell, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int amps_gpu_ecsim_time_step(amps_ctx* c, const amps_ecsim_params* p, double* E, double* E_theta,
                                        double* B_c, double tol, int32_t max_iter, int32_t restart, int32_t* iters,
                                        double* rel_residual) {
  if (!E || !B_c || !iters || !rel_residual) return fail(AMPS_E_ARG, "amps_gpu_ecsim_time_step: null argument");
  if (!(tol > 0.0) || max_iter < 1 || restart < 1 || restart > 400)
    return fail(AMPS_E_ARG, "amps_gpu_ecsim_time_step: tol > 0, max_iter >= 1 and restart in 1..400 are required");
  Work W;
  if (int r = prepare(c, p, restart, &W, "amps_gpu_ecsim_time_step")) return r;
  const int64_t nn = c->nnodes;
  const size_t n3 = 3 * (size_t)nn;
  const int64_t ncell = (int64_t)c->mv.nb * NCELL;
  HIP_TRY(hipMemcpyAsync(W.En, E, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(W.Bc, B_c, sizeof(double) * 3 * ncell, hipMemcpyHostToDevice, c->stream));
  // right-hand side, GMRES, E^{n+theta} and E^{n+1}, B^{n+1}
  hipLaunchKernelGGL(k_ecsim_rhs, grid_of(3 * nn), dim3(256), 0, c->stream, nn, c->d_nodeout, c->d_ecsim_nbr,
                     c->d_ecsim_bnd, c->d_ecsim_cells8, W.C, p->dt, p->theta, p->E_conv, p->B_conv, W.En, W.Bc,
                     W.rhs);
  int it = 0;
  double rr = 0.0;
  if (int r = gmres(c, W, fpdt_of(p), tol, max_iter, restart, &it, &rr)) return r;
  hipLaunchKernelGGL(k_ecsim_update_e, grid_of((int64_t)n3), dim3(256), 0, c->stream, (int64_t)n3, W.En, W.x,
                     p->theta, p->E_conv, W.Eth, W.E1);
  double dx[3];
  cell_size(c, dx);
  double c4[3];
  for (int d = 0; d < 3; d++) c4[d] = 0.25 * (p->cDt / (dx[d] * p->length_conv));
  hipLaunchKernelGGL(k_ecsim_update_b, grid_of(ncell), dim3(256), 0, c->stream, ncell, c->d_ecsim_noc, W.Eth,
                     W.Bc, c4[0], c4[1], c4[2], p->E_conv, p->B_conv, W.Bout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(E, W.E1, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  if (E_theta) HIP_TRY(hipMemcpyAsync(E_theta, W.Eth, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(B_c, W.Bout, sizeof(double) * 3 * ncell, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *iters = it;
  *rel_residual = rr;
  return AMPS_OK;
}
