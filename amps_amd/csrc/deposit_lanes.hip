// deposit_lanes.hip — ECSIM deposit cell pass on the vector ALUs, eight
// lanes per cell.
//
// ProcessCell (pic_field_solver_ecsim.cpp:1882-2438) in the moment form: per
// cell, J[c][d] = sum_p W_c (q alpha v)[d] and the 27 mass-matrix moments
// M[t][k] = sum_p mx[tx] my[ty] mz[tz] (MC alpha[k]), with
// mx = ((1-fx)^2, fx(1-fx), fx^2) (my, mz alike), t = tx + 3 ty + 9 tz and
// MC = q (q dt / 2m) / V; the node pass turns the moments into the IndexMatrix
// slots of each corner.
//
// One wavefront per cube of 2x2x2 cells of a block, no workgroup barriers:
//  * build: lane (cell c, slot j) turns the cell's particle 8 b + j of batch
//    b into a record in the wavefront's LDS — the products mx[tx] my[ty],
//    mz[tz], MC alpha[k], the linear factors (1 - f, f) and q alpha v — with
//    the next batch's particle already in flight;
//  * accumulate: lane (c, l) owns component k = l of all 27 moments of cell
//    c, the moments (tx ty = l, tz) of component 8 (and (tx ty = 8, tz = l)
//    for l < 3), and J of corner l; it walks the cell's records of the batch,
//    all eight lanes of a cell reading the same record (broadcast).
// Per particle this is 27 + 4 + 3 fused multiply-adds per lane and the
// prologue once per particle, instead of 512 FMA of a 16x32 matrix tile.  A
// slot without a particle holds a zero record, so every lane walks all PB
// slots of a batch unbranched, two records per step.  The register budget is
// a launch parameter (2 or 3 wavefronts per SIMD; at 3 the record loop
// spills).  The finished cells' records (J / V, then M[27][9]) leave through
// LDS as whole 128-byte lines.
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

namespace {

constexpr int NFL = 30;                  // record fields
constexpr int F_MXY = 0;                 // mx[tx] my[ty] at tx + 3 ty
constexpr int F_MZ = 9;                  // mz[tz]
constexpr int F_MA = 12;                 // MC alpha[k]
constexpr int F_LF = 21;                 // (1 - f, f) of x, y, z
constexpr int F_QV = 27;                 // q alpha v
constexpr int CSR = CELL_STRIDE_MOMENT;  // cell record stride (267 values + padding)

// InterpolationCoefficientTable corner l -> offset bit of axis d
__device__ __forceinline__ int cbit(int l, int d) {
  return d == 0 ? ((l == 1) | (l == 2) | (l == 5) | (l == 6)) : d == 1 ? ((l == 2) | (l == 3) | (l == 6) | (l == 7))
                                                                       : (l >= 4);
}

__device__ __forceinline__ void wave_lds_sync() { asm volatile("" ::: "memory"); }

template <int PB, bool BC, int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE))) void k_deposit_cell_lanes(MeshView m, Species sp, double dtTotal,
                                                           const double* __restrict__ B,
                                                           const int* __restrict__ offset, int64_t cap,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ v, int vf32,
                                                           const uint8_t* __restrict__ spec,
                                                           double* __restrict__ cellpart,
                                                           const double* __restrict__ Bc,
                                                           const int* __restrict__ src) {
  constexpr int RS = 8 * (PB + 1);  // slots per field: cell c, particle j at c (PB + 1) + j
  static_assert(NFL * RS >= 4 * CSR, "the record buffer stages four cell records");
  __shared__ double rec[NFL * RS];
  __shared__ double sB[27 * 3];
  __shared__ double sconst[MAX_SPEC][3];  // chargeQ, QdT_over_2m, matrixConst
  const int tile = xcd_tile();            // consecutive cubes of a block on one XCD
  const int b = tile >> 6, q = tile & 63;
  const int qx = 2 * (q & 3), qy = 2 * ((q >> 2) & 3), qz = 2 * (q >> 4);
  const int lane = threadIdx.x;
  const int c = lane >> 3, l = lane & 7;
  const int keybase = b * NCELL;
  auto cellkey = [&](int cc) {
    return keybase + (qx + (cc & 1)) + NC * ((qy + ((cc >> 1) & 1)) + NC * (qz + (cc >> 2)));
  };
  if constexpr (!BC) {
    const double* Bb = B + (size_t)b * NCORNER * 3;
    for (int i = lane; i < 81; i += 64) {
      const int cidx = i / 3, d = i % 3;
      const int a0 = cidx % 3, a1 = (cidx / 3) % 3, a2 = cidx / 9;
      sB[i] = Bb[3 * ((qx + a0) + NN * ((qy + a1) + NN * (qz + a2))) + d];
    }
  }
  BlockGeom g;
  load_block_geom(m, b, g);
  double CellVolume = 1;
  for (int d = 0; d < 3; d++) CellVolume *= (g.xmax[d] - g.xmin[d]) / NC * 1.0;
  double inv_dxs[3];
#pragma unroll
  for (int d = 0; d < 3; d++) inv_dxs[d] = 1.0 / g.dxs[d];
  if (lane < sp.n) {
    // per-species factors of ProcessCell, same operations in the same order
    const int s = lane;
    double chargeQ = sp.charge[s], mass = sp.mass[s];
    chargeQ *= sp.weight[s];
    mass *= sp.weight[s];
    const double QdT_over_m = chargeQ * dtTotal / mass;
    const double QdT_over_2m = 0.5 * QdT_over_m;
    sconst[s][0] = chargeQ;
    sconst[s][1] = QdT_over_2m;
    sconst[s][2] = chargeQ * QdT_over_2m / CellVolume;
  }
  __syncthreads();  // one wavefront: orders the staging above before its use

  const int key = cellkey(c);
  const int p0 = offset[key];
  const int n = offset[key + 1] - p0;
  int nmax = n;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) nmax = max(nmax, __shfl_xor(nmax, o));
  const int nbatch = (nmax + PB - 1) / PB;
  const int iX[3] = {qx + (c & 1), qy + ((c >> 1) & 1), qz + (c >> 2)};

  // the particle this lane builds next (builders: l < PB)
  double px[3] = {0.0, 0.0, 0.0}, pv[3] = {0.0, 0.0, 0.0};
  int ps = 0;
  auto fetch = [&](int rank) {
    if (l < PB && rank < n) {
      // src: the particles are still where the last compaction found them,
      // particle p of the new order at src[p]
      const int p = src ? src[p0 + rank] : p0 + rank;
      px[0] = x[p], px[1] = x[cap + p], px[2] = x[2 * cap + p];
      if (vf32) {
        const float* vf = (const float*)v;
        pv[0] = vf[p], pv[1] = vf[cap + p], pv[2] = vf[2 * cap + p];
      } else {
        pv[0] = v[p], pv[1] = v[cap + p], pv[2] = v[2 * cap + p];
      }
      ps = spec[p];
    }
  };

  // this lane's accumulators
  double acc[27];
#pragma unroll
  for (int t = 0; t < 27; t++) acc[t] = 0.0;
  double ax[4] = {0.0, 0.0, 0.0, 0.0};  // component 8: (tx ty = l, tz = 0..2), then (tx ty = 8, tz = l)
  double aj[3] = {0.0, 0.0, 0.0};       // J of corner l
  const int bx = cbit(l, 0), by = cbit(l, 1), bz = cbit(l, 2);
  const int lz = l < 3 ? l : 0;
  const double has8 = l < 3 ? 1.0 : 0.0;

  fetch(l);
  for (int bt = 0; bt < nbatch; bt++) {
    if (l < PB && PB * bt + l >= n) {
      // no particle in this slot: a zero record adds nothing, so every lane
      // walks all PB slots of the batch without a branch
      double* R = rec + c * (PB + 1) + l;
#pragma unroll
      for (int f = 0; f < NFL; f++) R[f * RS] = 0.0;
    } else if (l < PB) {
      // the record: the reference's expressions, contracted to FMA where
      // the compiler finds them (this pass is the tolerance-bound one)
#pragma clang fp contract(fast)
      const int s = ps;
      double Lf[6], xn[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double xd = px[d];
        if (fabs(xd - g.xmax[d]) < 1e-10 * g.dxs[d]) xd = g.xmax[d] - 1e-10 * g.dxs[d];
        xn[d] = xd;
        const double t = (xd - g.xmin[d]) * inv_dxs[d] - iX[d];
        const double f = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        Lf[2 * d] = 1.0 - f;
        Lf[2 * d + 1] = f;
      }
      double Bv[3] = {0.0, 0.0, 0.0};
      if constexpr (BC) {
        center_B(m, Bc, xn, b, Bv);  // centre-based B, pic_field_solver_ecsim.cpp:2097-2129
      } else {
        double wn[8], norm = 0.0;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          wn[k] = Lf[k >> 2] * Lf[2 + ((k >> 1) & 1)] * Lf[4 + (k & 1)];
          norm += wn[k];
        }
        const double rn = 1.0 / norm;
#pragma unroll
        for (int k = 0; k < 8; k++) {
          const double wl = wn[k] * rn;
          const int cc = ((c & 1) + (k >> 2)) + 3 * ((((c >> 1) & 1) + ((k >> 1) & 1)) + 3 * ((c >> 2) + (k & 1)));
#pragma unroll
          for (int d = 0; d < 3; d++) Bv[d] += wl * sB[3 * cc + d];
        }
      }
      const double QdT_over_2m = sconst[s][1];
      const double QdT_over_2m_squared = QdT_over_2m * QdT_over_2m;
      double BB[3][3], P[3], alpha[9];
#pragma unroll
      for (int ii = 0; ii < 3; ii++) {
        P[ii] = -QdT_over_2m * Bv[ii];
#pragma unroll
        for (int jj = 0; jj <= ii; jj++) {
          BB[ii][jj] = QdT_over_2m_squared * Bv[ii] * Bv[jj];
          BB[jj][ii] = BB[ii][jj];
        }
      }
      const double c0 = 1.0 / (1.0 + QdT_over_2m_squared * (Bv[0] * Bv[0] + Bv[1] * Bv[1] + Bv[2] * Bv[2]));
      alpha[0] = c0 * (1.0 + BB[0][0]);
      alpha[1] = c0 * (-P[2] + BB[0][1]);
      alpha[2] = c0 * (P[1] + BB[0][2]);
      alpha[3] = c0 * (P[2] + BB[1][0]);
      alpha[4] = c0 * (1.0 + BB[1][1]);
      alpha[5] = c0 * (-P[0] + BB[1][2]);
      alpha[6] = c0 * (-P[1] + BB[2][0]);
      alpha[7] = c0 * (P[0] + BB[2][1]);
      alpha[8] = c0 * (1.0 + BB[2][2]);
      const double chargeQ = sconst[s][0], MC = sconst[s][2];
      double* R = rec + c * (PB + 1) + l;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double vr = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; jj++) vr += alpha[3 * i + jj] * pv[jj];
        R[(F_QV + i) * RS] = chargeQ * vr;
      }
      double mo[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        mo[d][0] = Lf[2 * d] * Lf[2 * d];
        mo[d][1] = Lf[2 * d] * Lf[2 * d + 1];
        mo[d][2] = Lf[2 * d + 1] * Lf[2 * d + 1];
      }
#pragma unroll
      for (int ty = 0; ty < 3; ty++)
#pragma unroll
        for (int tx = 0; tx < 3; tx++) R[(F_MXY + tx + 3 * ty) * RS] = mo[0][tx] * mo[1][ty];
#pragma unroll
      for (int tz = 0; tz < 3; tz++) R[(F_MZ + tz) * RS] = mo[2][tz];
#pragma unroll
      for (int k = 0; k < 9; k++) R[(F_MA + k) * RS] = MC * alpha[k];
#pragma unroll
      for (int i = 0; i < 6; i++) R[(F_LF + i) * RS] = Lf[i];
    }
    fetch(PB * (bt + 1) + l);  // in flight while the batch is accumulated
    wave_lds_sync();           // the DS ops of a wavefront execute in issue order
#pragma unroll 2
    for (int r = 0; r < PB; r++) {
      {
        const double* R = rec + c * (PB + 1) + r;
        const double mak = R[(F_MA + l) * RS], ma8 = R[(F_MA + 8) * RS];
        double a[3], a8[3];
#pragma unroll
        for (int tz = 0; tz < 3; tz++) {
          const double mz = R[(F_MZ + tz) * RS];
          a[tz] = mak * mz;
          a8[tz] = ma8 * mz;
        }
#pragma unroll
        for (int i = 0; i < 9; i++) {
          const double mxy = R[(F_MXY + i) * RS];
#pragma unroll
          for (int tz = 0; tz < 3; tz++) acc[i + 9 * tz] = fma(mxy, a[tz], acc[i + 9 * tz]);
        }
        // component 8: tx ty = l for every tz; tx ty = 8 at tz = l (l < 3)
        const double mxyl = R[(F_MXY + l) * RS];
#pragma unroll
        for (int tz = 0; tz < 3; tz++) ax[tz] = fma(mxyl, a8[tz], ax[tz]);
        const double w8 = R[(F_MXY + 8) * RS] * has8;
        ax[3] = fma(w8, ma8 * R[(F_MZ + lz) * RS], ax[3]);
        // J of corner l
        const double wc = R[(F_LF + bx) * RS] * (R[(F_LF + 2 + by) * RS] * R[(F_LF + 4 + bz) * RS]);
#pragma unroll
        for (int d = 0; d < 3; d++) aj[d] = fma(wc, R[(F_QV + d) * RS], aj[d]);
      }
    }
    wave_lds_sync();  // the records are read before the next batch overwrites them
  }

  // the cells' records through LDS, four cells at a time, as whole lines
  const double invV = 1.0 / CellVolume;
#pragma unroll
  for (int half = 0; half < 2; half++) {
    if ((c >> 2) == half) {
      double* S = rec + (c & 3) * CSR;
#pragma unroll
      for (int d = 0; d < 3; d++) S[3 * l + d] = aj[d] * invV;
#pragma unroll
      for (int t = 0; t < 27; t++) S[24 + 9 * t + l] = acc[t];
#pragma unroll
      for (int tz = 0; tz < 3; tz++) S[24 + 9 * (l + 9 * tz) + 8] = ax[tz];
      if (l < 3) S[24 + 9 * (8 + 9 * l) + 8] = ax[3];
      if (l < CSR - CELL_VALS_MOMENT) S[CELL_VALS_MOMENT + l] = 0.0;
    }
    wave_lds_sync();
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
      double* dst = cellpart + (size_t)cellkey(4 * half + cc) * CSR;
      const double* S = rec + cc * CSR;
      for (int i = lane; i < CSR; i += 64) dst[i] = S[i];
    }
    wave_lds_sync();
  }
}

template <int PB, bool BC, int WPE>
void launch(amps_ctx* c, double dt) {
  // records a step's migration left: read from the previous buffer through
  // its source indices
  const int k = c->cur, r = c->pending_gather ? 1 - k : k;
  hipLaunchKernelGGL((k_deposit_cell_lanes<PB, BC, WPE>), dim3((unsigned)c->mv.nb * 64), dim3(64), 0, c->stream,
                     c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[r], c->d_v[r],
                     c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[r], c->d_cellpart, c->d_Bc,
                     c->pending_gather ? c->d_tmpsrc : nullptr);
}

template <int PB>
void launch_pb(amps_ctx* c, double dt) {
  // the register budget: 2 or 3 wavefronts per SIMD (the centre-B variant
  // calls the non-inlined stencil and takes what it needs)
  if (c->b_center) launch<PB, true, 1>(c, dt);
  else if (c->dep_waves == 2) launch<PB, false, 2>(c, dt);
  else launch<PB, false, 3>(c, dt);
}

}  // namespace

namespace amps {
int launch_deposit_cell_lanes(amps_ctx* c, double dt) {
  switch (c->dep_nch) {
    case 4: launch_pb<4>(c, dt); break;
    case 8: launch_pb<8>(c, dt); break;
    default: return fail(AMPS_E_ARG, "deposit_kernel 3 runs with deposit_chunks 4 or 8 (particles per cell per batch)");
  }
  return 0;
}
}  // namespace amps
