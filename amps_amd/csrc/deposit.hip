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
// Every record value is computed by the same expressions, in the same order,
// as the reference's ProcessCell, and each cell's particles are summed in
// list order, so with EXACT (separate multiply and add) the cell partials are
// bit-identical to a serial evaluation; without EXACT the sums use FMA.
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "pic_device.h"

using namespace amps;

namespace {

constexpr int NF = 22;  // record fields: table[8] alpha[9] chargeQ matrixConst vRot[3]
constexpr int F_TAB = 0, F_ALPHA = 8, F_CQ = 17, F_MC = 18, F_VROT = 19;

template <int O0, int O1, int RB, bool EXACT>
__device__ __forceinline__ void accum_records(double* acc, const double* rec, int lane, int nr) {
  for (int r = 0; r < nr; r++) {
    const double* R = rec + r * 64 + lane;
#define FLD(f) R[(f) * RB * 64]
    double tab[8];
#pragma unroll
    for (int i = 0; i < 8; i++) tab[i] = FLD(F_TAB + i);
    const double chargeQ = FLD(F_CQ), matrixConst = FLD(F_MC);
    double twc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) twc[i] = matrixConst * tab[i];
#pragma unroll
    for (int o = O0; o < O1; o++) {
      if (o < 24) {
        const double t = chargeQ * tab[o / 3];
        if (EXACT) acc[o - O0] += t * FLD(F_VROT + o % 3);
        else acc[o - O0] = fma(t, FLD(F_VROT + o % 3), acc[o - O0]);
      } else {
        const int pr = (o - 24) / 9, k = (o - 24) % 9;
        const double twp = tab[pair_j(pr)] * twc[pair_i(pr)];
        if (EXACT) acc[o - O0] += FLD(F_ALPHA + k) * twp;
        else acc[o - O0] = fma(FLD(F_ALPHA + k), twp, acc[o - O0]);
      }
    }
#undef FLD
  }
}

template <int NCH, int RB, bool EXACT, int W = 0>
__device__ __forceinline__ void dispatch_accum(int w, double* acc, const double* rec, int lane, int nr) {
  if constexpr (W < NCH) {
    constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < CELL_VALS) ? O0 + CH : CELL_VALS;
    if (w == W)
      accum_records<O0, O1, RB, EXACT>(acc, rec, lane, nr);
    else
      dispatch_accum<NCH, RB, EXACT, W + 1>(w, acc, rec, lane, nr);
  }
}

template <int NCH, int W = 0>
__device__ __forceinline__ void dispatch_stage(int w, int w0, const double* acc, double* stage, int SW,
                                               int lane, double CellVolume) {
  if constexpr (W < NCH) {
    constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < CELL_VALS) ? O0 + CH : CELL_VALS;
    if (w == W) {
      double* dst = stage + lane * SW + (W - w0) * CH;
#pragma unroll
      for (int o = O0; o < O1; o++) dst[o - O0] = (o < 24) ? acc[o - O0] / CellVolume : acc[o - O0];
    } else {
      dispatch_stage<NCH, W + 1>(w, w0, acc, stage, SW, lane, CellVolume);
    }
  }
}

// RB = NCH: every thread builds one record per batch, so a batch holds NCH
// particles of each of the 64 cells.
template <int NCH, bool EXACT>
__global__ __launch_bounds__(64 * NCH) void k_deposit_cell_rec(MeshView m, Species sp, double dtTotal,
                                                                const double* __restrict__ B,
                                                                const int* __restrict__ offset,
                                                                int64_t cap,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ v,
                                                                int vf32,
                                                                const uint8_t* __restrict__ spec,
                                                                double* __restrict__ cellpart) {
  constexpr int RB = NCH;
  constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
  constexpr int FS = RB * 64;  // field stride of a record batch
  constexpr int REC = NF * FS;
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
  for (int i = threadIdx.x; i < 125 * 3; i += blockDim.x) {
    const int cidx = i / 3, d = i % 3;
    const int a0 = cidx % 5, a1 = (cidx / 5) % 5, a2 = cidx / 25;
    sB[i] = Bb[3 * ((sx + a0) + NN * ((sy + a1) + NN * (sz + a2))) + d];
  }
  BlockGeom g;
  load_block_geom(m, b, g);
  double CellVolume = 1;
  for (int d = 0; d < 3; d++) CellVolume *= (g.xmax[d] - g.xmin[d]) / NC * 1.0;
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
      double wn[8], table[8];
      int iX[3];
      corner_stencil(g, xI, wn, table, iX);
      iX[0] = clampi(iX[0], sx, sx + 3);
      iX[1] = clampi(iX[1], sy, sy + 3);
      iX[2] = clampi(iX[2], sz, sz + 3);
      double Bv[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int l = 0; l < 8; l++) {
        const int c = (iX[0] - sx + (l >> 2)) + 5 * ((iX[1] - sy + ((l >> 1) & 1)) + 5 * (iX[2] - sz + (l & 1)));
#pragma unroll
        for (int d = 0; d < 3; d++) Bv[d] += wn[l] * sB[3 * c + d];
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
        R[(F_VROT + i) * FS] = vr;
      }
#pragma unroll
      for (int i = 0; i < 8; i++) R[(F_TAB + i) * FS] = table[i];
#pragma unroll
      for (int k = 0; k < 9; k++) R[(F_ALPHA + k) * FS] = alpha[k];
      R[F_CQ * FS] = sconst[s][0];
      R[F_MC * FS] = sconst[s][2];
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
    const int v0 = w0 * CH, v1 = (w1 * CH < CELL_VALS) ? w1 * CH : CELL_VALS;
    const int nv1 = v1 - v0;
    if (w >= w0 && w < w1) dispatch_stage<NCH>(w, w0, acc, rec, SW, lane, CellVolume);
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * nv1; i += blockDim.x) {
      const int c = i / nv1, j = i - c * nv1;
      const int key = keybase + (sx + (c & 3)) + NC * ((sy + ((c >> 2) & 3)) + NC * (sz + (c >> 4)));
      cellpart[(size_t)key * CELL_VALS + v0 + j] = rec[c * SW + j];
    }
    __syncthreads();
  }
}

template <int NCH, bool EX>
void launch_rec(amps_ctx* c, double dt) {
  const int k = c->cur;
  hipLaunchKernelGGL((k_deposit_cell_rec<NCH, EX>), dim3((unsigned)c->mv.nb * 8), dim3(64 * NCH), 0,
                     c->stream, c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[k], c->d_v[k],
                     c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], c->d_cellpart);
}

}  // namespace

namespace amps {
int launch_deposit_cell_rec(amps_ctx* c, double dt) {
  const bool ex = c->dep_exact != 0;
  switch (c->dep_nch) {
    case 4: ex ? launch_rec<4, true>(c, dt) : launch_rec<4, false>(c, dt); break;
    case 6: ex ? launch_rec<6, true>(c, dt) : launch_rec<6, false>(c, dt); break;
    case 12: ex ? launch_rec<12, true>(c, dt) : launch_rec<12, false>(c, dt); break;
    default:
      return fail(AMPS_E_ARG, "deposit_kernel 1 runs with deposit_chunks 4, 6 or 12");
  }
  return 0;
}
}  // namespace amps
