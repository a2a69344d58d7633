// deposit.hip — ECSIM deposit cell pass, record-batched.
//
// One workgroup per 4x4x4 cell sub-tile of a block (64 cells) with NCH
// wavefronts.  The sub-tile's particles are taken in batches of NCH particles
// per cell.  A batch first turns each particle into a record — its corner
// weights (InitStencil), alpha matrix (Lapenta Eq. D.3), alpha v, charge and
// mass-matrix factor — built once per particle by one thread (whose next
// particle is loaded while the batch is accumulated) and kept in LDS as
// structure-of-arrays [field][rank][cell], so the reads below are
// bank-conflict free.  Lane l of wavefront w then accumulates output chunk
// w of cell l (J and the 36 lower-triangular corner-pair 3x3 blocks of
// ProcessCell, pic_field_solver_ecsim.cpp:2241-2342) from its cell's records.
//
// With EXACT every record value is computed by the same expressions, in the
// same order, as the reference's ProcessCell (separate multiply and add), and
// each cell's particles are summed in list order, so the cell partials are
// bit-identical to a serial evaluation.  Without EXACT the stencil uses the
// block's reciprocal cell size and one reciprocal of the weight sum, and the
// sums use FMA (results within a few ulp per term).
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

namespace {

// Record fields.  EXACT: table[8] alpha[9] chargeQ matrixConst vRot[3].
// Moment form: the linear factors L[2d + bit] (1 - f_d, f_d) of the three axes
// instead of the 8 corner weights.
template <bool EXACT>
struct Rec {
  static constexpr int NF = EXACT ? 22 : 20;
  static constexpr int TAB = 0, LIN = 0;
  static constexpr int ALPHA = EXACT ? 8 : 6;
  static constexpr int CQ = ALPHA + 9, MC = ALPHA + 10, VROT = ALPHA + 11;
  static constexpr int VALS = EXACT ? CELL_VALS : CELL_VALS_MOMENT;
  static constexpr int STRIDE = EXACT ? CELL_VALS : CELL_STRIDE_MOMENT;
};

// corner c (InterpolationCoefficientTable order) -> its x, y, z offset bits
constexpr int cbit(int c, int d) {
  return d == 0 ? (c == 1 || c == 2 || c == 5 || c == 6) : d == 1 ? (c == 2 || c == 3 || c == 6 || c == 7) : (c >= 4);
}

template <int O0, int O1, int RB, bool EXACT>
__device__ __forceinline__ void accum_records(double* acc, const double* rec, int lane, int nr) {
  using R_ = Rec<EXACT>;
  for (int r = 0; r < nr; r++) {
    const double* R = rec + r * 64 + lane;
#define FLD(f) R[(f) * RB * 64]
    const double chargeQ = FLD(R_::CQ), matrixConst = FLD(R_::MC);
    if constexpr (EXACT) {
      double tab[8];
#pragma unroll
      for (int i = 0; i < 8; i++) tab[i] = FLD(R_::TAB + i);
      double twc[8];
#pragma unroll
      for (int i = 0; i < 8; i++) twc[i] = matrixConst * tab[i];
#pragma unroll
      for (int o = O0; o < O1; o++) {
        if (o < 24) {
          const double t = chargeQ * tab[o / 3];
          acc[o - O0] += t * FLD(R_::VROT + o % 3);
        } else {
          const int pr = (o - 24) / 9, k = (o - 24) % 9;
          const double twp = tab[pair_j(pr)] * twc[pair_i(pr)];
          acc[o - O0] += FLD(R_::ALPHA + k) * twp;
        }
      }
    } else {
      double Lf[6];
#pragma unroll
      for (int i = 0; i < 6; i++) Lf[i] = FLD(R_::LIN + i);
#pragma unroll
      for (int o = O0; o < O1; o++) {
        if (o < 24) {
          const int c = o / 3;
          const double w = Lf[cbit(c, 0)] * Lf[2 + cbit(c, 1)] * Lf[4 + cbit(c, 2)];
          acc[o - O0] = fma(chargeQ * w, FLD(R_::VROT + o % 3), acc[o - O0]);
        } else {
          const int t = (o - 24) / 9, k = (o - 24) % 9;
          const int s0 = t % 3, s1 = (t / 3) % 3, s2 = t / 9;
          const double g0 = Lf[s0 == 2 ? 1 : 0] * Lf[s0 == 0 ? 0 : 1];
          const double g1 = Lf[2 + (s1 == 2 ? 1 : 0)] * Lf[2 + (s1 == 0 ? 0 : 1)];
          const double g2 = Lf[4 + (s2 == 2 ? 1 : 0)] * Lf[4 + (s2 == 0 ? 0 : 1)];
          acc[o - O0] = fma(FLD(R_::ALPHA + k), ((matrixConst * g0) * g1) * g2, acc[o - O0]);
        }
      }
    }
#undef FLD
  }
}

template <int NCH, int RB, bool EXACT, int W = 0>
__device__ __forceinline__ void dispatch_accum(int w, double* acc, const double* rec, int lane, int nr) {
  if constexpr (W < NCH) {
    constexpr int V = Rec<EXACT>::VALS;
    constexpr int CH = (V + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < V) ? O0 + CH : V;
    if (w == W)
      accum_records<O0, O1, RB, EXACT>(acc, rec, lane, nr);
    else
      dispatch_accum<NCH, RB, EXACT, W + 1>(w, acc, rec, lane, nr);
  }
}

template <int NCH, bool EXACT, int W = 0>
__device__ __forceinline__ void dispatch_stage(int w, int w0, const double* acc, double* stage, int SW,
                                               int lane, double CellVolume) {
  if constexpr (W < NCH) {
    constexpr int V = Rec<EXACT>::VALS;
    constexpr int CH = (V + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < V) ? O0 + CH : V;
    if (w == W) {
      double* dst = stage + lane * SW + (W - w0) * CH;
#pragma unroll
      for (int o = O0; o < O1; o++) dst[o - O0] = (o < 24) ? acc[o - O0] / CellVolume : acc[o - O0];
    } else {
      dispatch_stage<NCH, EXACT, W + 1>(w, w0, acc, stage, SW, lane, CellVolume);
    }
  }
}

// RB = NCH: every thread builds one record per batch, so a batch holds NCH
// particles of each of the 64 cells.
template <int NCH, bool EXACT, bool BC>
__global__ __launch_bounds__(64 * NCH) void k_deposit_cell_rec(MeshView m, Species sp, double dtTotal,
                                                                const double* __restrict__ B,
                                                                const int* __restrict__ offset,
                                                                int64_t cap,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ v,
                                                                int vf32,
                                                                const uint8_t* __restrict__ spec,
                                                                double* __restrict__ cellpart,
                                                                const double* __restrict__ Bc) {
  using R_ = Rec<EXACT>;
  constexpr int RB = NCH;
  constexpr int V = R_::VALS;  // values per cell record
  constexpr int CH = (V + NCH - 1) / NCH;
  constexpr int FS = RB * 64;  // field stride of a record batch
  constexpr int REC = R_::NF * FS;
  // output staging reuses the record buffer: WPR waves' chunks of all 64 cells per round
  constexpr int WPR = (REC / (64 * (CH + 1)) < NCH) ? REC / (64 * (CH + 1)) : NCH;
  static_assert(WPR >= 1, "record buffer too small to stage one wave's output");
  constexpr int SW = (WPR * CH) | 1;  // odd stride (in doubles) between the staged cells
  __shared__ double sB[125 * 3];
  __shared__ double rec[REC];
  __shared__ double sconst[MAX_SPEC][3];  // chargeQ, QdT_over_2m, matrixConst
  __shared__ int maxcnt;
  const int tile = xcd_tile();  // the 8 sub-tiles of a block on one XCD
  const int b = tile >> 3;
  const int st = tile & 7;
  const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4, sz = (st >> 2) * 4;
  const double* Bb = B + (size_t)b * NCORNER * 3;
  if constexpr (!BC) {
    for (int i = threadIdx.x; i < 125 * 3; i += blockDim.x) {
      const int cidx = i / 3, d = i % 3;
      const int a0 = cidx % 5, a1 = (cidx / 5) % 5, a2 = cidx / 25;
      sB[i] = Bb[3 * ((sx + a0) + NN * ((sy + a1) + NN * (sz + a2))) + d];
    }
  }
  BlockGeom g;
  load_block_geom(m, b, g);
  double CellVolume = 1;
  for (int d = 0; d < 3; d++) CellVolume *= (g.xmax[d] - g.xmin[d]) / NC * 1.0;
  double inv_dxs[3];
#pragma unroll
  for (int d = 0; d < 3; d++) inv_dxs[d] = EXACT ? 0.0 : 1.0 / g.dxs[d];
  if (threadIdx.x < sp.n) {
    // per-species factors of ProcessCell, same operations in the same order
    const int s = threadIdx.x;
    double chargeQ = sp.charge[s], mass = sp.mass[s];
    chargeQ *= sp.weight[s];
    mass *= sp.weight[s];
    const double QdT_over_m = chargeQ * dtTotal / mass;
    const double QdT_over_2m = 0.5 * QdT_over_m;
    sconst[s][0] = chargeQ;
    sconst[s][1] = QdT_over_2m;
    sconst[s][2] = chargeQ * QdT_over_2m / CellVolume;
  }
  if (threadIdx.x == 0) maxcnt = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int keybase = b * NCELL;
  // accumulating role: cell = lane
  const int akey = keybase + (sx + (lane & 3)) + NC * ((sy + ((lane >> 2) & 3)) + NC * (sz + (lane >> 4)));
  const int acnt = offset[akey + 1] - offset[akey];
  if (w == 0) atomicMax(&maxcnt, acnt);
  // record-building role: thread -> (cell, rank within the batch)
  const int bcell = threadIdx.x / RB, brank = threadIdx.x % RB;
  const int bkey = keybase + (sx + (bcell & 3)) + NC * ((sy + ((bcell >> 2) & 3)) + NC * (sz + (bcell >> 4)));
  const int bp0 = offset[bkey], bcnt = offset[bkey + 1] - bp0;
  // the particle of the next batch, loaded one batch ahead
  double nx[3] = {0.0, 0.0, 0.0}, nv[3] = {0.0, 0.0, 0.0};
  int ns = 0;
  auto fetch = [&](int rank) {
    if (rank < bcnt) {
      const int p = bp0 + rank;
      nx[0] = x[p], nx[1] = x[cap + p], nx[2] = x[2 * cap + p];
      if (vf32) {
        const float* vf = (const float*)v;
        nv[0] = vf[p], nv[1] = vf[cap + p], nv[2] = vf[2 * cap + p];
      } else {
        nv[0] = v[p], nv[1] = v[cap + p], nv[2] = v[2 * cap + p];
      }
      ns = spec[p];
    }
  };
  fetch(brank);
  __syncthreads();
  const int nbatch = (maxcnt + RB - 1) / RB;
  double acc[CH];
#pragma unroll
  for (int o = 0; o < CH; o++) acc[o] = 0.0;
  for (int bt = 0; bt < nbatch; bt++) {
    const int rank = bt * RB + brank;
    if (rank < bcnt) {
      double xI[3] = {nx[0], nx[1], nx[2]};
      const double vI[3] = {nv[0], nv[1], nv[2]};
      const int s = ns;
      double wn[8], table[8], Lf[6];
      double xn[3] = {xI[0], xI[1], xI[2]};  // the position the stencils see (nudged)
      int iX[3];
      if constexpr (EXACT) {
        corner_stencil(g, xI, wn, table, iX);
#pragma unroll
        for (int d = 0; d < 3; d++) xn[d] = xI[d];
        iX[0] = clampi(iX[0], sx, sx + 3);
        iX[1] = clampi(iX[1], sy, sy + 3);
        iX[2] = clampi(iX[2], sz, sz + 3);
      } else {
        // the same weights with the block's reciprocal cell size, relative
        // to the particle's own cell (a few ulp from the divided form)
        iX[0] = sx + (bcell & 3), iX[1] = sy + ((bcell >> 2) & 3), iX[2] = sz + (bcell >> 4);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          double xd = xI[d];
          if (fabs(xd - g.xmax[d]) < 1e-10 * g.dxs[d]) xd = g.xmax[d] - 1e-10 * g.dxs[d];
          xn[d] = xd;
          const double t = (xd - g.xmin[d]) * inv_dxs[d] - iX[d];
          const double f = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
          Lf[2 * d] = 1.0 - f;
          Lf[2 * d + 1] = f;
        }
        double norm = 0.0;
#pragma unroll
        for (int l = 0; l < 8; l++) {
          wn[l] = Lf[l >> 2] * Lf[2 + ((l >> 1) & 1)] * Lf[4 + (l & 1)];
          norm += wn[l];
        }
        const double rn = 1.0 / norm;
#pragma unroll
        for (int l = 0; l < 8; l++) wn[l] *= rn;
      }
      double Bv[3] = {0.0, 0.0, 0.0};
      if constexpr (BC) {
        center_B(m, Bc, xn, b, Bv);  // centre-based B, pic_field_solver_ecsim.cpp:2097-2129
      } else {
#pragma unroll
        for (int l = 0; l < 8; l++) {
          const int c = (iX[0] - sx + (l >> 2)) + 5 * ((iX[1] - sy + ((l >> 1) & 1)) + 5 * (iX[2] - sz + (l & 1)));
#pragma unroll
          for (int d = 0; d < 3; d++) Bv[d] += wn[l] * sB[3 * c + d];
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
      double* R = rec + brank * 64 + bcell;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double vr = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; jj++) vr += alpha[3 * i + jj] * vI[jj];
        R[(R_::VROT + i) * FS] = vr;
      }
      if constexpr (EXACT) {
#pragma unroll
        for (int i = 0; i < 8; i++) R[(R_::TAB + i) * FS] = table[i];
      } else {
#pragma unroll
        for (int i = 0; i < 6; i++) R[(R_::LIN + i) * FS] = Lf[i];
      }
#pragma unroll
      for (int k = 0; k < 9; k++) R[(R_::ALPHA + k) * FS] = alpha[k];
      R[R_::CQ * FS] = sconst[s][0];
      R[R_::MC * FS] = sconst[s][2];
    }
    fetch(rank + RB);  // in flight while the batch is accumulated
    __syncthreads();
    int nr = acnt - bt * RB;
    nr = nr < 0 ? 0 : (nr > RB ? RB : nr);
    dispatch_accum<NCH, RB, EXACT>(w, acc, rec, lane, nr);
    __syncthreads();
  }
  // write the 64 cell records through LDS, WPR waves' chunks per round, so
  // that each cell's values leave as contiguous runs
  for (int w0 = 0; w0 < NCH; w0 += WPR) {
    const int w1 = (w0 + WPR < NCH) ? w0 + WPR : NCH;
    const int v0 = w0 * CH, v1 = (w1 * CH < V) ? w1 * CH : V;
    const int nv1 = v1 - v0;
    if (w >= w0 && w < w1) dispatch_stage<NCH, EXACT>(w, w0, acc, rec, SW, lane, CellVolume);
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * nv1; i += blockDim.x) {
      const int c = i / nv1, j = i - c * nv1;
      const int key = keybase + (sx + (c & 3)) + NC * ((sy + ((c >> 2) & 3)) + NC * (sz + (c >> 4)));
      cellpart[(size_t)key * R_::STRIDE + v0 + j] = rec[c * SW + j];
    }
    __syncthreads();
  }
}

template <int NCH, bool EX, bool BC>
void launch_rec(amps_ctx* c, double dt) {
  const int k = c->cur;
  hipLaunchKernelGGL((k_deposit_cell_rec<NCH, EX, BC>), dim3((unsigned)c->mv.nb * 8), dim3(64 * NCH), 0,
                     c->stream, c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[k], c->d_v[k],
                     c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], c->d_cellpart, c->d_Bc);
}

template <int NCH>
void launch_rec2(amps_ctx* c, double dt) {
  const bool ex = c->dep_exact != 0, bc = c->b_center != 0;
  if (ex) bc ? launch_rec<NCH, true, true>(c, dt) : launch_rec<NCH, true, false>(c, dt);
  else bc ? launch_rec<NCH, false, true>(c, dt) : launch_rec<NCH, false, false>(c, dt);
}

}  // namespace

namespace amps {
int launch_deposit_cell_rec(amps_ctx* c, double dt, int nch) {
  switch (nch) {
    case 4: launch_rec2<4>(c, dt); break;
    case 6: launch_rec2<6>(c, dt); break;
    case 12: launch_rec2<12>(c, dt); break;
    default:
      return fail(AMPS_E_ARG, "deposit_kernel 1 runs with deposit_chunks 4, 6 or 12");
  }
  return 0;
}
}  // namespace amps
