// deposit_mfma.hip — ECSIM deposit cell pass on the FP64 matrix cores.
//
// The mass-matrix part of ProcessCell (pic_field_solver_ecsim.cpp:2241-2342)
// is, per cell, a sum over the cell's particles of rank-one updates.  In the
// moment form each particle contributes
//     M[t][k] += (mx[tx] my[ty]) * (mz[tz] MC alpha[k])
// with mx = ((1-fx)^2, fx(1-fx), fx^2) (and my, mz alike), t = tx + 3ty + 9tz
// and MC = q (q dt / 2m) / V: a 9 x 27 product over the particles.  The
// current is
//     J[c][d] += (wx[bx] q alphaV[d]) * (wy[by] wz[bz])
// a 6 x 4 product.  Both fit one 16 x 32 tile of A (16 x np) times B (np x 32):
//     A rows 0-8:  mx[tx] my[ty]               B cols 0-26:  mz[tz] MC alpha[k]
//     A rows 9-14: wx[bx] q alphaV[d]          B cols 27-30: wy[by] wz[bz]
// so a cell is two v_mfma_f64_16x16x4f64 per 4 particles, accumulated in
// registers, with the parts of the 16 x 32 result outside these blocks unused.
//
// One workgroup per 4x4x4 sub-tile of a block; wavefront w takes the 16
// cells of z-layer w.  Its cells' particles are consecutive runs of the
// key-sorted arrays; the wavefront walks their concatenation 64 particles at a
// time: each lane turns one particle into its operand record (corner stencil,
// B from the sub-tile's corners staged in LDS, alpha, alpha v) in LDS, then the
// wavefront issues the MFMA steps cell by cell and writes a finished cell's
// 267 values (J[8][3] / V, then M[27][9]) to its partial-sum record.  The next
// chunk's particles are loaded while the current one is accumulated.
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int MW = 4;       // wavefronts per workgroup: one per z-layer of the sub-tile
constexpr int RS = 33;      // record stride in doubles (odd: lane-per-record stores hit distinct banks)
constexpr int F_A = 0;      // A operand rows 0-15 (row 15 stays zero)
constexpr int F_ZERO = 15;
constexpr int F_MA = 16;    // MC alpha[k], k = 0..8
constexpr int F_MZ = 25;    // mz[tz], tz = 0..2
constexpr int F_WYZ = 28;   // wy[by] wz[bz] at by + 2bz
constexpr int F_ONE = 32;   // 1.0, written once
constexpr int CS = CELL_STRIDE_MOMENT;  // cell record stride: 267 values + padding

// LDS ordering between the lanes of one wavefront: DS instructions of a
// wavefront complete in order; the wait and the clobber keep the compiler
// from moving LDS accesses across.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// a value every lane of the wavefront holds, kept in a scalar register
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// InterpolationCoefficientTable corner index of offset bits (bx, by, bz)
__device__ __forceinline__ int corner_of(int bx, int by, int bz) { return 4 * bz + (by ? (bx ? 2 : 3) : bx); }

template <bool BC>
__global__ __launch_bounds__(64 * MW) void k_deposit_cell_mfma(MeshView m, Species sp, double dtTotal,
                                                                 const double* __restrict__ B,
                                                                 const int* __restrict__ offset,
                                                                 int64_t cap,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ v,
                                                                 int vf32,
                                                                 const uint8_t* __restrict__ spec,
                                                                 double* __restrict__ cellpart,
                                                                 const double* __restrict__ Bc) {
  __shared__ double sB[125 * 3];
  __shared__ double rec[MW][64 * RS];
  __shared__ double sconst[MAX_SPEC][3];  // chargeQ, QdT_over_2m, matrixConst
  __shared__ int sfirst[MW][16], scum[MW][17];
  __shared__ double sstage[MW][CS];  // a finished cell's record, written out in full lines
  const int tile = xcd_tile();  // the 8 sub-tiles of a block on one XCD
  const int b = tile >> 3;
  const int st = tile & 7;
  const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4, sz = (st >> 2) * 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int zc = sz + w;
  const int keybase = b * NCELL;
  if constexpr (!BC) {
    const double* Bb = B + (size_t)b * NCORNER * 3;
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
  for (int d = 0; d < 3; d++) inv_dxs[d] = 1.0 / g.dxs[d];
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
  // this wavefront's 16 cells (cell c = cx + 4cy of layer zc): first particle, running counts
  if (lane < 16) {
    const int key = keybase + (sx + (lane & 3)) + NC * ((sy + (lane >> 2)) + NC * zc);
    sfirst[w][lane] = offset[key];
  }
  if (lane < 17) {
    int cum = 0;
    for (int c = 0; c < lane; c++) {
      const int key = keybase + (sx + (c & 3)) + NC * ((sy + (c >> 2)) + NC * zc);
      cum += offset[key + 1] - offset[key];
    }
    scum[w][lane] = cum;
  }
  __syncthreads();
  const int* cum = scum[w];
  const int total = uniform(cum[16]);
  double* R0 = rec[w];

  // operand selectors of this lane: A row ia = j and B columns j, 16 + j of particle slot lane / 16
  const int j = lane & 15, slot = lane >> 4;
  const int f0a = F_MZ + j / 9, f0b = F_MA + j % 9;
  int f1a, f1b;
  {
    const int col = 16 + j;
    if (col < 27) {
      f1a = F_MZ + col / 9, f1b = F_MA + col % 9;
    } else if (col < 31) {
      f1a = F_WYZ + (col - 27), f1b = F_ONE;
    } else {
      f1a = F_ZERO, f1b = F_ZERO;
    }
  }
  // destinations of the 8 accumulator values; -1: unused.  v_mfma_f64_16x16x4f64
  // holds A[i][k] in lane i + 16k, B[k][j] in lane j + 16k and D[i][j] in
  // lane 16 (i % 4) + j, accumulator register i / 4.
  int oidx[8];
  bool ojc[8];
#pragma unroll
  for (int T = 0; T < 2; T++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = slot + 4 * r, col = 16 * T + j;
      int o = -1;
      bool jc = false;
      if (i < 9 && col < 27) {
        o = 24 + 9 * (i + 9 * (col / 9)) + col % 9;
      } else if (i >= 9 && i < 15 && col >= 27 && col < 31) {
        const int bx = (i - 9) / 3, d = (i - 9) % 3, by = (col - 27) & 1, bz = (col - 27) >> 1;
        o = 3 * corner_of(bx, by, bz) + d;
        jc = true;
      }
      oidx[4 * T + r] = o;
      ojc[4 * T + r] = jc;
    }
  }
  const double invV = 1.0 / CellVolume;
  R0[lane * RS + F_ONE] = 1.0;
  double* S = sstage[w];
  if (lane < CS - CELL_VALS_MOMENT) S[CELL_VALS_MOMENT + lane] = 0.0;

  // the particle of the next chunk, loaded one chunk ahead
  double nx[3] = {0.0, 0.0, 0.0}, nv[3] = {0.0, 0.0, 0.0};
  int ns = 0, nc = 0;
  auto fetch = [&](int pos) {
    if (pos < total) {
      int c = 0;
#pragma unroll
      for (int k = 1; k < 16; k++) c += (pos >= cum[k]) ? 1 : 0;
      const int p = sfirst[w][c] + (pos - cum[c]);
      nc = c;
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

  dbl4 C0 = {0.0, 0.0, 0.0, 0.0}, C1 = C0;
  int cur = 0;  // the cell being accumulated (uniform)
  auto flush = [&](int c) {
    const int key = keybase + (sx + (c & 3)) + NC * ((sy + (c >> 2)) + NC * zc);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (oidx[r] >= 0) S[oidx[r]] = ojc[r] ? C0[r] * invV : C0[r];
      if (oidx[4 + r] >= 0) S[oidx[4 + r]] = ojc[4 + r] ? C1[r] * invV : C1[r];
    }
    C0 = dbl4{0.0, 0.0, 0.0, 0.0};
    C1 = C0;
    wave_lds_sync();
    // 2176 bytes from a 128-byte boundary: every store covers whole lines
    double* dst = cellpart + (size_t)key * CS;
#pragma unroll
    for (int i = lane; i < CS; i += 64) dst[i] = S[i];
    wave_lds_sync();
  };
  // cells with no particles before the first particle
  while (cur < 16 && uniform(cum[cur + 1]) == 0) flush(cur++);

  fetch(lane);
  for (int base = 0; base < total; base += 64) {
    const int pos = base + lane;
    if (pos < total) {
      const double xI[3] = {nx[0], nx[1], nx[2]};
      const double vI[3] = {nv[0], nv[1], nv[2]};
      const int s = ns;
      const int iX[3] = {sx + (nc & 3), sy + (nc >> 2), zc};
      double Lf[6], xn[3];
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
      double Bv[3] = {0.0, 0.0, 0.0};
      if constexpr (BC) {
        center_B(m, Bc, xn, b, Bv);  // centre-based B, pic_field_solver_ecsim.cpp:2097-2129
      } else {
        double wn[8], norm = 0.0;
#pragma unroll
        for (int l = 0; l < 8; l++) {
          wn[l] = Lf[l >> 2] * Lf[2 + ((l >> 1) & 1)] * Lf[4 + (l & 1)];
          norm += wn[l];
        }
        const double rn = 1.0 / norm;
#pragma unroll
        for (int l = 0; l < 8; l++) {
          const double wl = wn[l] * rn;
          const int c = (iX[0] - sx + (l >> 2)) + 5 * ((iX[1] - sy + ((l >> 1) & 1)) + 5 * (iX[2] - sz + (l & 1)));
#pragma unroll
          for (int d = 0; d < 3; d++) Bv[d] += wl * sB[3 * c + d];
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
      double qv[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double vr = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; jj++) vr += alpha[3 * i + jj] * vI[jj];
        qv[i] = chargeQ * vr;
      }
      double mo[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        mo[d][0] = Lf[2 * d] * Lf[2 * d];
        mo[d][1] = Lf[2 * d] * Lf[2 * d + 1];
        mo[d][2] = Lf[2 * d + 1] * Lf[2 * d + 1];
      }
      double* R = R0 + lane * RS;
#pragma unroll
      for (int ty = 0; ty < 3; ty++)
#pragma unroll
        for (int tx = 0; tx < 3; tx++) R[F_A + tx + 3 * ty] = mo[0][tx] * mo[1][ty];
#pragma unroll
      for (int bx = 0; bx < 2; bx++)
#pragma unroll
        for (int d = 0; d < 3; d++) R[F_A + 9 + 3 * bx + d] = Lf[bx] * qv[d];
      R[F_ZERO] = 0.0;
#pragma unroll
      for (int k = 0; k < 9; k++) R[F_MA + k] = MC * alpha[k];
#pragma unroll
      for (int tz = 0; tz < 3; tz++) R[F_MZ + tz] = mo[2][tz];
#pragma unroll
      for (int bz = 0; bz < 2; bz++)
#pragma unroll
        for (int by = 0; by < 2; by++) R[F_WYZ + by + 2 * bz] = Lf[2 + by] * Lf[4 + bz];
    }
    fetch(pos + 64);  // in flight while this chunk is accumulated
    wave_lds_sync();
    const int cend = (base + 64 < total) ? base + 64 : total;
    int q = base;
    while (q < cend) {
      const int ce = uniform(cum[cur + 1]);
      const int se = ce < cend ? ce : cend;
      // full steps of 4 particles, the operands of step t + 1 loaded while step t runs
      const int nfull = (se - q) >> 2;
      const double* R = R0 + (q - base + slot) * RS;
      if (nfull > 0) {
        double a = R[F_A + j], b0 = R[f0a] * R[f0b], b1 = R[f1a] * R[f1b];
        for (int t = 1; t < nfull; t++) {
          R += 4 * RS;
          const double a2 = R[F_A + j], b02 = R[f0a] * R[f0b], b12 = R[f1a] * R[f1b];
          C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, C0, 0, 0, 0);
          C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, C1, 0, 0, 0);
          a = a2, b0 = b02, b1 = b12;
        }
        C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, C0, 0, 0, 0);
        C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, C1, 0, 0, 0);
        R += 4 * RS;
      }
      if (q + 4 * nfull < se) {
        // the cell's last 1-3 particles in this chunk: the other slots add zero
        double a = 0.0, b0 = 0.0, b1 = 0.0;
        if (q + 4 * nfull + slot < se) {
          a = R[F_A + j];
          b0 = R[f0a] * R[f0b];
          b1 = R[f1a] * R[f1b];
        }
        C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, C0, 0, 0, 0);
        C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, C1, 0, 0, 0);
      }
      q = se;
      if (q == ce) {
        flush(cur++);
        while (cur < 16 && uniform(cum[cur + 1]) == q) flush(cur++);  // empty cells
      }
    }
    wave_lds_sync();  // the records are read before the next chunk overwrites them
  }
}

template <bool BC>
void launch(amps_ctx* c, double dt) {
  const int k = c->cur;
  hipLaunchKernelGGL((k_deposit_cell_mfma<BC>), dim3((unsigned)c->mv.nb * 8), dim3(64 * MW), 0, c->stream, c->mv,
                     c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[k], c->d_v[k],
                     c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], c->d_cellpart, c->d_Bc);
}

}  // namespace

namespace amps {
int launch_deposit_cell_mfma(amps_ctx* c, double dt) {
  if (c->b_center)
    launch<true>(c, dt);
  else
    launch<false>(c, dt);
  return 0;
}
}  // namespace amps
