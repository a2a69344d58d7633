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
//
// PAIRED: eight wavefronts, two per SIMD (a z-layer's two row pairs), run
// in opposite phases locked by workgroup barriers: while one builds the
// records of its next chunk (VALU), the other issues the matrix products of
// its current one (MFMA), so the two units overlap instead of taking turns.
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

namespace {

typedef double dbl4 __attribute__((ext_vector_type(4)));

// Records are kept field-major, rec[f * FS + particle]: the record stores
// (lane = particle) are contiguous, and with FS = 66 the MFMA operand reads
// (16 fields x 2 particles per 32-lane pass) fall on distinct banks.
constexpr int NFLD = 32;
constexpr int FS = 66;
constexpr int F_A = 0;      // A operand rows 0-14
constexpr int F_ONE = 15;   // 1: A row 15 and the factor of the wy wz columns
constexpr int F_MA = 16;    // MC alpha[k], k = 0..8
constexpr int F_MZ = 25;    // mz[tz], tz = 0..2
constexpr int F_WYZ = 28;   // wy[by] wz[bz] at by + 2bz
constexpr int CS = CELL_STRIDE_MOMENT;  // cell record stride: 267 values + padding

// LDS ordering between the lanes of one wavefront: the DS instructions of a
// wavefront execute in issue order, so a compiler barrier that keeps LDS
// accesses on their side is all that is needed.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("" ::: "memory"); }

// a value every lane of the wavefront holds, kept in a scalar register
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// InterpolationCoefficientTable corner index of offset bits (bx, by, bz)
__device__ __forceinline__ int corner_of(int bx, int by, int bz) { return 4 * bz + (by ? (bx ? 2 : 3) : bx); }

template <bool PAIRED>
constexpr int waves_of() { return PAIRED ? 8 : 4; }

// VF32: FP32 velocities (widened when the record is built, not at the load:
// a conversion there would wait for the prefetched data)
template <bool BC, bool PAIRED, bool FUSED, bool VF32>
__global__ __launch_bounds__(64 * waves_of<PAIRED>()) void k_deposit_cell_mfma(MeshView m, Species sp, double dtTotal,
                                                                 const double* __restrict__ B,
                                                                 const int* __restrict__ offset,
                                                                 int64_t cap,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ v,
                                                                 int /*vf32*/,
                                                                 const uint8_t* __restrict__ spec,
                                                                 double* __restrict__ cellpart,
                                                                 const double* __restrict__ Bc,
                                                                 const int* __restrict__ src) {
  constexpr int MW = waves_of<PAIRED>();
  constexpr int NCW = PAIRED ? 8 : 16;  // cells per wavefront
  __shared__ double sB[125 * 3];
  // 8 doubles of padding: the operand loads run up to two steps (8
  // particles) past a segment's last particle, and are not used there
  __shared__ double rec[MW][NFLD * FS + 8];
  __shared__ double sconst[MAX_SPEC][3];  // chargeQ, QdT_over_2m, matrixConst
  __shared__ int sfirst[MW][NCW], scum[MW][NCW + 1];
  __shared__ double sstage[MW][CS];  // a finished cell's record, written out in full lines
  __shared__ int smaxchunks;
  const int tile = xcd_tile();  // the 8 sub-tiles of a block on one XCD
  const int b = tile >> 3;
  const int st = tile & 7;
  const int sx = (st & 1) * 4, sy = ((st >> 1) & 1) * 4, sz = (st >> 2) * 4;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // wavefront w: z-layer w % 4, rows row0.. of it (cell c: x = c % 4, y = row0 + c / 4)
  const int zc = sz + (w & 3);
  const int row0 = PAIRED ? 2 * (w >> 2) : 0;
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
  auto cellkey = [&](int c) { return keybase + (sx + (c & 3)) + NC * ((sy + row0 + (c >> 2)) + NC * zc); };
  // this wavefront's cells: first particle, running counts
  {
    // counts of lanes' cells, then an inclusive scan across the lanes
    const int ck = cellkey(lane < NCW ? lane : 0);
    const int o0 = offset[ck], cnt = lane < NCW ? offset[ck + 1] - o0 : 0;
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < NCW; d <<= 1) {
      const int t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane < NCW) {
      sfirst[w][lane] = o0;
      scum[w][lane] = incl - cnt;
    }
    if (lane == NCW - 1) scum[w][NCW] = incl;
  }
  if (threadIdx.x == 0) smaxchunks = 0;
  __syncthreads();
  const int* cum = scum[w];
  const int total = uniform(cum[NCW]);
  double* R0 = rec[w];

  // operand selectors of this lane: A row ia = j and B columns j, 16 + j of particle slot lane / 16
  const int j = lane & 15, slot = lane >> 4;
  const int f0a = F_MZ + j / 9, f0b = F_MA + j % 9;
  int f1a, f1b;
  {
    // A row 15 and B column 31 hold ones; the results they touch are not kept
    const int col = 16 + j;
    if (col < 27) {
      f1a = F_MZ + col / 9, f1b = F_MA + col % 9;
    } else if (col < 31) {
      f1a = F_WYZ + (col - 27), f1b = F_ONE;
    } else {
      f1a = F_ONE, f1b = F_ONE;
    }
  }
  const int oa = (F_A + j) * FS, o0a = f0a * FS, o0b = f0b * FS, o1a = f1a * FS, o1b = f1b * FS;
  // destinations of the 8 accumulator values; -1: unused.  v_mfma_f64_16x16x4f64
  // holds A[i][k] in lane i + 16k, B[k][j] in lane j + 16k and D[i][j] in
  // lane 16 (i % 4) + j, accumulator register i / 4.
  double* S = sstage[w];
  int oidx[8];
#pragma unroll
  for (int T = 0; T < 2; T++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = slot + 4 * r, col = 16 * T + j;
      int o = -1;
      if (i < 9 && col < 27) {
        o = 24 + 9 * (i + 9 * (col / 9)) + col % 9;
      } else if (i >= 9 && i < 15 && col >= 27 && col < 31) {
        const int bx = (i - 9) / 3, d = (i - 9) % 3, by = (col - 27) & 1, bz = (col - 27) >> 1;
        o = 3 * corner_of(bx, by, bz) + d;
      }
      oidx[4 * T + r] = o;
    }
  }
  const double invV = 1.0 / CellVolume;
  if (lane < CS - CELL_VALS_MOMENT) S[CELL_VALS_MOMENT + lane] = 0.0;

  // each lane's particles of the next two chunks, loaded ahead
  struct Pre {
    double x[3], v[3];
    float vf[3];
    int s, c;
  };
  // lane k < NCW: where cell k starts in the concatenation, and the offset
  // from a position in the concatenation to the particle's index
  const int cumk = lane < NCW ? cum[lane] : 0;
  const int dk = lane < NCW ? sfirst[w][lane] - cum[lane] : 0;
  // the particles at positions pbase + lane (called by the whole wavefront):
  // a position's cell is the last one that starts at or before it
  auto fetch = [&](int pbase, Pre& n) {
    const int pos = pbase + lane;
    int c = __popcll(__ballot(lane >= 1 && lane < NCW && cumk <= pbase));
    uint64_t starts = __ballot(lane >= 1 && lane < NCW && cumk > pbase && cumk <= pbase + 63);
    while (starts) {
      const int k = __builtin_ctzll(starts);
      starts &= starts - 1;
      c += (pos >= __builtin_amdgcn_readlane(cumk, k)) ? 1 : 0;
    }
    const int p = pos + __shfl(dk, c, 64);
    if (pos < total) {
      n.c = c;
      // FUSED: the particles are still where the last compaction found
      // them, particle p of the new order at src[p]
      const int q = FUSED ? src[p] : p;
      n.x[0] = x[q], n.x[1] = x[cap + q], n.x[2] = x[2 * cap + q];
      if constexpr (VF32) {
        const float* vf = (const float*)v;
        n.vf[0] = vf[q], n.vf[1] = vf[cap + q], n.vf[2] = vf[2 * cap + q];
      } else {
        n.v[0] = v[q], n.v[1] = v[cap + q], n.v[2] = v[2 * cap + q];
      }
      n.s = spec[q];
    }
  };
  Pre n1 = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.f, 0.f, 0.f}, 0, 0}, n2 = n1;

  // the two 16 x 16 halves of the cell's 16 x 32 tile
  dbl4 C0 = {0.0, 0.0, 0.0, 0.0}, C1 = C0;
  int cur = 0;  // the cell being accumulated (uniform)
  auto flush = [&](int c) {
    const int key = cellkey(c);
#pragma unroll
    for (int r = 0; r < 4; r++) {
      if (oidx[r] >= 0) S[oidx[r]] = C0[r];
      if (oidx[4 + r] >= 0) S[oidx[4 + r]] = C1[r];
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
  while (cur < NCW && uniform(cum[cur + 1]) == 0) flush(cur++);

  fetch(0, n1);
  fetch(64, n2);
  // the records of the chunk starting at particle base (VALU) from its
  // particles in nb; the particles two chunks ahead are then loaded into nb
  // (chunks alternate between n1 and n2)
  auto build = [&](int base, Pre& nb) {
    const int pos = base + lane;
    if (pos < total) {
      // the record: the reference's expressions, contracted to FMA where
      // the compiler finds them (this pass is the tolerance-bound one)
#pragma clang fp contract(fast)
      const double xI[3] = {nb.x[0], nb.x[1], nb.x[2]};
      const double vI[3] = {VF32 ? (double)nb.vf[0] : nb.v[0], VF32 ? (double)nb.vf[1] : nb.v[1],
                            VF32 ? (double)nb.vf[2] : nb.v[2]};
      const int s = nb.s;
      const int nc = nb.c;
      const int iX[3] = {sx + (nc & 3), sy + row0 + (nc >> 2), zc};
      double Lf[6], xn[3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        double xd = xI[d];
        if (fabs(xd - g.xmax[d]) < 1e-10 * g.dxs[d]) xd = g.xmax[d] - 1e-10 * g.dxs[d];
        xn[d] = xd;
        const double t = (xd - g.xmin[d]) * inv_dxs[d] - iX[d];
        const double f = fmin(fmax(t, 0.0), 1.0);
        Lf[2 * d] = 1.0 - f;
        Lf[2 * d + 1] = f;
      }
      double Bv[3] = {0.0, 0.0, 0.0};
      if constexpr (BC) {
        center_B(m, Bc, xn, b, Bv);  // centre-based B, pic_field_solver_ecsim.cpp:2097-2129
      } else {
        // the corner weights sum to 1 (InitStencil's normalization changes
        // them by rounding only)
        double lxy[4];
#pragma unroll
        for (int l = 0; l < 4; l++) lxy[l] = Lf[l >> 1] * Lf[2 + (l & 1)];
        // the cell's lower corner in the staged 5 x 5 x 5 corners; corner l
        // at a fixed offset from it
        const double* sBc = sB + 3 * ((nc & 3) + 5 * (row0 + (nc >> 2)) + 25 * (w & 3));
#pragma unroll
        for (int l = 0; l < 8; l++) {
          const double wl = lxy[l >> 1] * Lf[4 + (l & 1)];
          const int c = (l >> 2) + 5 * ((l >> 1) & 1) + 25 * (l & 1);
#pragma unroll
          for (int d = 0; d < 3; d++) Bv[d] += wl * sBc[3 * c + d];
        }
      }
      // alpha = c0 (I + P x + BB) with P = -QdT_over_2m B and BB = P P^T,
      // kept as its numerators a[]; c0 goes into the two factors that use it
      // (MC alpha, and q alpha v, which also takes the 1 / V of J)
      double qb[3];
#pragma unroll
      for (int ii = 0; ii < 3; ii++) qb[ii] = sconst[s][1] * Bv[ii];
      const double c0 = 1.0 / (1.0 + (qb[0] * qb[0] + qb[1] * qb[1] + qb[2] * qb[2]));
      double alpha[9];
      alpha[0] = 1.0 + qb[0] * qb[0];
      alpha[1] = qb[2] + qb[0] * qb[1];
      alpha[2] = -qb[1] + qb[0] * qb[2];
      alpha[3] = -qb[2] + qb[1] * qb[0];
      alpha[4] = 1.0 + qb[1] * qb[1];
      alpha[5] = qb[0] + qb[1] * qb[2];
      alpha[6] = qb[1] + qb[2] * qb[0];
      alpha[7] = -qb[0] + qb[2] * qb[1];
      alpha[8] = 1.0 + qb[2] * qb[2];
      const double MC = sconst[s][2] * c0, qc = sconst[s][0] * c0 * invV;
      double qv[3];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        double vr = 0.0;
#pragma unroll
        for (int jj = 0; jj < 3; jj++) vr += alpha[3 * i + jj] * vI[jj];
        qv[i] = qc * vr;
      }
      double mo[3][3];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        mo[d][0] = Lf[2 * d] * Lf[2 * d];
        mo[d][1] = Lf[2 * d] * Lf[2 * d + 1];
        mo[d][2] = Lf[2 * d + 1] * Lf[2 * d + 1];
      }
      double* R = R0 + lane;
#define FLD(f) R[(f) * FS]
#pragma unroll
      for (int ty = 0; ty < 3; ty++)
#pragma unroll
        for (int tx = 0; tx < 3; tx++) FLD(F_A + tx + 3 * ty) = mo[0][tx] * mo[1][ty];
#pragma unroll
      for (int bx = 0; bx < 2; bx++)
#pragma unroll
        for (int d = 0; d < 3; d++) FLD(F_A + 9 + 3 * bx + d) = Lf[bx] * qv[d];
      FLD(F_ONE) = 1.0;
#pragma unroll
      for (int k = 0; k < 9; k++) FLD(F_MA + k) = MC * alpha[k];
#pragma unroll
      for (int tz = 0; tz < 3; tz++) FLD(F_MZ + tz) = mo[2][tz];
#pragma unroll
      for (int bz = 0; bz < 2; bz++)
#pragma unroll
        for (int by = 0; by < 2; by++) FLD(F_WYZ + by + 2 * bz) = Lf[2 + by] * Lf[4 + bz];
#undef FLD
    }
    fetch(base + 128, nb);  // in flight while two chunks are accumulated
  };
  // the matrix products of that chunk, cell by cell, and the finished cells' records
  auto accumulate = [&](int base) {
    const int cend = (base + 64 < total) ? base + 64 : total;
    int q = base;
    while (q < cend) {
      const int ce = uniform(cum[cur + 1]);
      const int se = ce < cend ? ce : cend;
      // full steps of 4 particles, the operands of step t + 1 loaded while step t runs
      const int nfull = (se - q) >> 2;
      const double* R = R0 + (q - base + slot);
      if (nfull > 0) {
        // the raw operands of two steps in flight ahead of the products (the
        // LDS latency under load exceeds one step's two MFMAs); the factors
        // are multiplied just before use
        struct Ops {
          double a, x0, y0, x1, y1;
        };
        // one pointer per operand field, advanced by two steps per iteration
        const double *Ra = R + oa, *Rx0 = R + o0a, *Ry0 = R + o0b, *Rx1 = R + o1a, *Ry1 = R + o1b;
        auto ld = [&](int k, Ops& o) {
          o.a = Ra[k];
          o.x0 = Rx0[k];
          o.y0 = Ry0[k];
          o.x1 = Rx1[k];
          o.y1 = Ry1[k];
        };
        auto step = [&](const Ops& o) {
          const double b0 = o.x0 * o.y0, b1 = o.x1 * o.y1;
          C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a, b0, C0, 0, 0, 0);
          C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a, b1, C1, 0, 0, 0);
        };
        // two operand sets, unrolled by two so that neither is copied while
        // its loads are in flight; each is reloaded right after its products
        // (the loads of the two steps past the last one read records past
        // the segment, inside rec[], and are not used)
        Ops pa, pb;
        ld(0, pa);
        ld(4, pb);
        __builtin_amdgcn_sched_barrier(0);
        int t = 0;
        for (; t + 1 < nfull; t += 2) {
          // the scheduling barriers keep each reload where it is issued
          step(pa);
          Ra += 8, Rx0 += 8, Ry0 += 8, Rx1 += 8, Ry1 += 8;
          ld(0, pa);
          __builtin_amdgcn_sched_barrier(0);
          step(pb);
          ld(4, pb);
          __builtin_amdgcn_sched_barrier(0);
        }
        if (t < nfull) step(pa);
        R += 4 * nfull;
      }
      if (q + 4 * nfull < se) {
        // the cell's last 1-3 particles in this chunk: the other slots add zero
        double a = 0.0, b0 = 0.0, b1 = 0.0;
        if (q + 4 * nfull + slot < se) {
          a = R[oa];
          b0 = R[o0a] * R[o0b];
          b1 = R[o1a] * R[o1b];
        }
        C0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b0, C0, 0, 0, 0);
        C1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b1, C1, 0, 0, 0);
      }
      q = se;
      if (q == ce) {
        flush(cur++);
        while (cur < NCW && uniform(cum[cur + 1]) == q) flush(cur++);  // empty cells
      }
    }
  };
  if constexpr (!PAIRED) {
    for (int base = 0; base < total; base += 128) {
      build(base, n1);
      wave_lds_sync();
      accumulate(base);
      wave_lds_sync();  // the records are read before the next chunk overwrites them
      if (base + 64 < total) {
        build(base + 64, n2);
        wave_lds_sync();
        accumulate(base + 64);
        wave_lds_sync();
      }
    }
  } else {
    // phase p: waves 0-3 build chunk p/2 at even p and accumulate it at odd
    // p; waves 4-7 (the other wavefront of each SIMD) one phase later
    const int nch = (total + 63) / 64;
    if (lane == 0) atomicMax(&smaxchunks, nch);
    __syncthreads();
    const int grp = w >> 2;
    const int nphase = 2 * smaxchunks + 1;
    for (int ph = 0; ph < nphase; ph++) {
      const int k = ph - grp;
      if (k >= 0 && (k >> 1) < nch) {
        if ((k & 1) == 0) {
          if ((k >> 1) & 1)
            build(64 * (k >> 1), n2);
          else
            build(64 * (k >> 1), n1);
        } else
          accumulate(64 * (k >> 1));
      }
      __syncthreads();
    }
  }
}

template <bool BC, bool PAIRED, bool FUSED, bool VF32>
void launch(amps_ctx* c, double dt) {
  const int k = c->cur;
  // FUSED: the particles are read from the previous buffer through the
  // migration's source indices
  const int r = FUSED ? 1 - k : k;
  hipLaunchKernelGGL((k_deposit_cell_mfma<BC, PAIRED, FUSED, VF32>), dim3((unsigned)c->mv.nb * 8),
                     dim3(64 * waves_of<PAIRED>()), 0, c->stream, c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap,
                     c->d_x[r], c->d_v[r], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[r], c->d_cellpart, c->d_Bc,
                     c->d_tmpsrc);
}

template <bool BC, bool PAIRED, bool FUSED>
void launch_v(amps_ctx* c, double dt, bool vf32) {
  vf32 ? launch<BC, PAIRED, FUSED, true>(c, dt) : launch<BC, PAIRED, FUSED, false>(c, dt);
}

}  // namespace

namespace amps {
int launch_deposit_cell_mfma(amps_ctx* c, double dt) {
  const bool paired = c->dep_paired != 0;
  const bool vf32 = c->precision == AMPS_FP32;
  if (c->pending_gather) {
    // the records of the last compaction are still where it found them: the
    // pass reads them there (set only with the single-phase pass)
    if (paired) return fail(AMPS_E_STATE, "deposit: records left by a step's migration need the single-phase cell pass");
    c->b_center ? launch_v<true, false, true>(c, dt, vf32) : launch_v<false, false, true>(c, dt, vf32);
    return 0;
  }
  if (c->b_center)
    paired ? launch_v<true, true, false>(c, dt, vf32) : launch_v<true, false, false>(c, dt, vf32);
  else
    paired ? launch_v<false, true, false>(c, dt, vf32) : launch_v<false, false, false>(c, dt, vf32);
  return 0;
}
}  // namespace amps
