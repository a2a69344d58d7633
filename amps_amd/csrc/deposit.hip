// deposit.hip — ECSIM deposit cell pass, record-batched.
//
// One workgroup per 4x4x4 cell sub-tile of a block (64 cells) with NCH
// wavefronts.  The sub-tile's particles are taken in batches of RB particles
// per cell.  A batch first turns each particle into a record — its corner
// weights (InitStencil), alpha matrix (Lapenta Eq. D.3), alpha v, charge and
// mass-matrix factor — built once per particle by one of the first 64 RB
// threads and kept in LDS as structure-of-arrays [field][rank][cell], so the reads below
// are bank-conflict free.  Lane l of wavefront w then accumulates output chunk
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
__device__ __forceinline__ void dispatch_store(int w, const double* acc, double* out, double CellVolume) {
  if constexpr (W < NCH) {
    constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < CELL_VALS) ? O0 + CH : CELL_VALS;
    if (w == W) {
#pragma unroll
      for (int o = O0; o < O1; o++) out[o] = (o < 24) ? acc[o - O0] / CellVolume : acc[o - O0];
    } else {
      dispatch_store<NCH, W + 1>(w, acc, out, CellVolume);
    }
  }
}

template <int NCH, int RB, bool EXACT>
__global__ __launch_bounds__(64 * NCH) void k_deposit_cell_rec(MeshView m, Species sp, double dtTotal,
                                                                const double* __restrict__ B,
                                                                const int* __restrict__ offset,
                                                                int64_t cap,
                                                                const double* __restrict__ x,
                                                                const double* __restrict__ v,
                                                                int vf32,
                                                                const uint8_t* __restrict__ spec,
                                                                double* __restrict__ cellpart) {
  constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
  constexpr int FS = RB * 64;  // field stride of a record batch
  __shared__ double sB[125 * 3];
  __shared__ double rec[NF * FS];
  __shared__ double sconst[MAX_SPEC][3];  // chargeQ, QdT_over_2m, matrixConst
  __shared__ int maxcnt;
  const int b = blockIdx.x >> 3;
  const int st = blockIdx.x & 7;
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
  // record-building role: thread -> (cell, rank within the batch); threads
  // beyond 64 RB only accumulate
  const bool builder = threadIdx.x < 64 * RB;
  const int bcell = threadIdx.x / RB, brank = threadIdx.x % RB;
  const int bkey = keybase + (sx + (bcell & 3)) + NC * ((sy + ((bcell >> 2) & 3)) + NC * (sz + (bcell >> 4)));
  const int bp0 = builder ? offset[bkey] : 0, bcnt = builder ? offset[bkey + 1] - bp0 : 0;
  __syncthreads();
  const int nbatch = (maxcnt + RB - 1) / RB;
  double acc[CH];
#pragma unroll
  for (int o = 0; o < CH; o++) acc[o] = 0.0;
  for (int bt = 0; bt < nbatch; bt++) {
    const int rank = bt * RB + brank;
    if (rank < bcnt) {
      const int p = bp0 + rank;
      double xI[3] = {x[p], x[cap + p], x[2 * cap + p]};
      double vI[3];
      if (vf32) {
        const float* vf = (const float*)v;
        vI[0] = vf[p], vI[1] = vf[cap + p], vI[2] = vf[2 * cap + p];
      } else {
        vI[0] = v[p], vI[1] = v[cap + p], vI[2] = v[2 * cap + p];
      }
      const int s = spec[p];
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
    __syncthreads();
    int nr = acnt - bt * RB;
    nr = nr < 0 ? 0 : (nr > RB ? RB : nr);
    dispatch_accum<NCH, RB, EXACT>(w, acc, rec, lane, nr);
    __syncthreads();
  }
  dispatch_store<NCH>(w, acc, cellpart + (size_t)akey * CELL_VALS, CellVolume);
}

template <int NCH, int RB, bool EX>
void launch_rec(amps_ctx* c, double dt) {
  const int k = c->cur;
  hipLaunchKernelGGL((k_deposit_cell_rec<NCH, RB, EX>), dim3((unsigned)c->mv.nb * 8), dim3(64 * NCH), 0,
                     c->stream, c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[k], c->d_v[k],
                     c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], c->d_cellpart);
}

}  // namespace

namespace amps {
int launch_deposit_cell_rec(amps_ctx* c, double dt) {
  const bool ex = c->dep_exact != 0;
  switch (c->dep_nch) {
    case 4: ex ? launch_rec<4, 4, true>(c, dt) : launch_rec<4, 4, false>(c, dt); break;
    case 6: ex ? launch_rec<6, 4, true>(c, dt) : launch_rec<6, 4, false>(c, dt); break;
    case 12: ex ? launch_rec<12, 4, true>(c, dt) : launch_rec<12, 4, false>(c, dt); break;
    default:
      return fail(AMPS_E_ARG, "deposit_kernel 1 runs with deposit_chunks 4, 6 or 12");
  }
  return 0;
}
}  // namespace amps
