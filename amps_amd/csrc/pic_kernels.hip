// pic_kernels.hip — push, migration and ECSIM deposit kernels for gfx950.
//
// Built with -ffp-contract=off: every FP64 expression of the FP64 path rounds
// as written, in the reference's order, so positions, velocities, block and
// cell ids reproduce the CPU restatement bit for bit.  Deliberate fused
// multiply-adds are written as fma().
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "center_device.h"
#include "pic_device.h"

using namespace amps;

// ======================================================================
// Push: one workgroup per tile of particles of a single block; the block's
// corner E and B (729 x 3 doubles each) are staged in LDS (SetBlock_E/B,
// pic_mover.cpp:84-186), particles are read and written coalesced (SoA).
// ======================================================================
__device__ __forceinline__ bool alive_code(int c) {
  return c == AMPS_CODE_FINISHED || c == AMPS_CODE_REFLECTED;
}

struct RunInfo {
  bool head;
  int head_lane;
  int len;  // valid on head lanes
};

// Runs of equal valid keys inside the wavefront.
__device__ __forceinline__ RunInfo wave_runs(bool valid, int k) {
  const int lane = __lane_id();
  int prev = __shfl_up(k, 1);
  bool pvalid = __shfl_up((int)valid, 1) != 0;
  if (lane == 0) pvalid = false;
  const bool head = valid && (!pvalid || prev != k);
  const uint64_t heads = __ballot(head);
  const uint64_t breaks = __ballot(head || !valid);
  RunInfo r;
  r.head = head;
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  const uint64_t hm = heads & upto;
  r.head_lane = hm ? (63 - __clzll((long long)hm)) : lane;
  const uint64_t after = (lane == 63) ? 0ull : (breaks & ~upto);
  const int nxt = after ? (__ffsll((long long)after) - 1) : 64;
  r.len = nxt - lane;
  return r;
}

// A pass's movers of the push's classification, listed a pass later, when
// their rank (the cursor atomic's return, on the run's head lane) has had the
// next pass's loads to arrive.
struct MigPend {
  int raw;        // the atomic's return (head lanes)
  int head_lane;  // the lane holding the run's return
  int off;        // lane - head_lane: the rank within the run
  int knew;
  int p;
  bool mover;
};

// the movers of a pass: each at its rank in its key's slab, or past it in
// the overflow list (made by every lane of the wavefront)
__device__ __forceinline__ void mig_list(const PushMig& pm, const MigPend& pd) {
  const int lane = __lane_id();
  const int rank = __shfl(pd.raw, pd.head_lane) + pd.off;
  const bool ov = pd.mover && rank >= MIG_SLAB;
  if (pd.mover && !ov) pm.slab[rank * pm.nkeys + pd.knew] = pd.p;
  const uint64_t om = __ballot(ov);
  if (om) {
    const int lead = __ffsll((long long)om) - 1;
    int ob = 0;
    if (lane == lead) ob = atomicAdd(pm.novf, __popcll(om));
    ob = __shfl(ob, lead);
    if (ov) pm.ovf[ob + __popcll(om & ((1ull << lane) - 1ull))] = make_int2(pd.p, rank);
  }
}

// The push's count for the migration that follows, made by every lane of
// the wavefront (p: the lane's position, start..p1 the tile's).  Without
// pm.lay every survivor counts in count[knew], one atomic per run of equal
// keys.  With it (the passes 64-aligned, so that a wavefront's pass is one
// stayer mask word) the push classifies: a stayer (its new key's old run,
// sLay[] of the tile's block, holds it) counts in count[] and sets its mask
// bit, stored plainly where the word lies inside the tile and by atomics at
// the tile's ends, which neighbouring tiles share; a mover takes its rank
// among its new key's movers from the cursor[] atomic, and the previous
// pass's movers are listed (mig_list; the last pass's after the loop).
__device__ __forceinline__ void push_count(int* __restrict__ count, const PushMig& pm, const int* sLay, int b,
                                           int p, int start, int p1, bool alive, int knew, MigPend& pd) {
  if (!pm.lay) {
    const RunInfo r = wave_runs(alive, knew);
    if (r.head) atomicAdd(&count[knew], r.len);
    return;
  }
  const int lane = __lane_id();
  const int cb = knew - b * NCELL;
  const bool stay = alive && cb >= 0 && cb < NCELL && p >= sLay[cb] && p < sLay[cb + 1];
  const bool mover = alive && !stay;
  const RunInfo rs = wave_runs(stay, knew);
  if (rs.head) atomicAdd(&count[knew], rs.len);
  const uint64_t sm = __ballot(stay);
  if (lane == 0 && sm) {
    const int64_t w = (int64_t)p >> 6;  // lane 0 holds the word's first position
    if ((w << 6) >= start && (w << 6) + 64 <= p1) {
      pm.wmask[w] = sm;
      pm.wcnt[w] = __popcll(sm);
    } else {
      atomicOr((unsigned long long*)(pm.wmask + w), (unsigned long long)sm);
      atomicAdd(pm.wcnt + w, __popcll(sm));
    }
  }
  mig_list(pm, pd);
  const RunInfo rm = wave_runs(mover, knew);
  pd.raw = 0;
  if (rm.head) pd.raw = atomicAdd(&pm.cursor[knew], rm.len);
  pd.head_lane = rm.head_lane;
  pd.off = lane - rm.head_lane;
  pd.knew = knew;
  pd.p = p;
  pd.mover = mover;
}

// W: the occupancy the register allocation aims at (waves per SIMD).
template <int MOVER, bool BC, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_push_f64(MeshView m, Faces faces, Species sp,
                                                   const double* __restrict__ E,
                                                   const double* __restrict__ B,
                                                   const Tile* __restrict__ tiles,
                                                   const int* __restrict__ ntiles, int bc_mode,
                                                   double r_sphere, int64_t cap, double* x,
                                                   double* v, const uint8_t* __restrict__ spec,
                                                   int* key, int8_t* code, const double* __restrict__ Bc,
                                                   int* __restrict__ count, PushSrc ps, PushMig pm) {
  __shared__ double sE[NCORNER * 3];
  __shared__ double sB[NCORNER * 3];
  __shared__ int sLay[NCELL + 1];
  const int t = blockIdx.x;
  if (t >= *ntiles) return;
  const Tile tl = tiles[t];
  const int b = tl.block;
  const double* Eb = E + (size_t)b * NCORNER * 3;
  const double* Bb = B + (size_t)b * NCORNER * 3;
  for (int i = threadIdx.x; i < NCORNER * 3; i += blockDim.x) {
    sE[i] = Eb[i];
    sB[i] = Bb[i];
  }
  if (pm.lay)
    for (int i = threadIdx.x; i <= NCELL; i += blockDim.x) sLay[i] = pm.lay[(int64_t)b * NCELL + i];
  __syncthreads();
  BlockGeom g;
  load_block_geom(m, b, g);
  BlockLat lat;
  load_block_lat(m, b, lat);
  const int p1 = tl.start + tl.count;
  // every lane runs the same number of passes, so that the wavefront's run
  // aggregation of the migration count sees all 64 lanes; the passes start
  // 64-aligned when the push classifies.  Where the next pass reads (the last
  // compaction's source index) is loaded a pass ahead.
  const int pa = pm.lay ? (tl.start & ~63) : tl.start;
  MigPend pd = {0, 0, 0, 0, 0, false};
  const int pf = pa + (int)threadIdx.x;
  int qn = (ps.src && pf >= tl.start && pf < p1) ? ps.src[pf] : 0;
  for (int p0 = pa; p0 < p1; p0 += blockDim.x) {
    const int p = p0 + threadIdx.x;
    bool alive = false;
    int knew = -1;
    const int qs = qn;
    if (ps.src && p + (int)blockDim.x < p1) qn = ps.src[p + blockDim.x];
    if (p >= tl.start && p < p1) {
    // read in place, or where the last compaction left the particle
    const int q = ps.src ? qs : p;
    const double* xr = ps.src ? ps.x : x;
    const double* vr = ps.src ? (const double*)ps.v : v;
    double xp[3] = {xr[q], xr[cap + q], xr[2 * cap + q]};
    double vp[3] = {vr[q], vr[cap + q], vr[2 * cap + q]};
    const int s = ps.src ? ps.spec[q] : spec[p];
    if (ps.src) {
      ps.spec_out[p] = (uint8_t)s;
      ps.id_out[p] = ps.id[q];
    }
    int nb_out = b, cell_out = 0;
    int c;
    if constexpr (MOVER == AMPS_MOVER_RELATIVISTIC)
      c = push_particle_rel(m, faces, b, g, sE, sB, E, B, sp.charge[s], sp.mass[s], sp.dt[s], bc_mode,
                            r_sphere, xp, vp, &nb_out, &cell_out);
    else
      c = push_particle<MOVER, BC>(m, faces, g, lat, sE, sB, sp.charge[s], sp.mass[s], sp.dt[s], bc_mode,
                                   r_sphere, xp, vp, &nb_out, &cell_out, Bc);
    x[p] = xp[0];
    x[cap + p] = xp[1];
    x[2 * cap + p] = xp[2];
    v[p] = vp[0];
    v[cap + p] = vp[1];
    v[2 * cap + p] = vp[2];
    alive = alive_code(c);
    knew = nb_out * NCELL + cell_out;
    // an exit of the USER_FUNCTION mode keeps the block holding its face state
    if (alive || c >= AMPS_CODE_EXIT_FACE0) key[p] = knew;
    code[p] = (int8_t)c;
    }
    // the migration's count (null when the caller counts separately)
    if (count) push_count(count, pm, sLay, b, p, tl.start, p1, alive, knew, pd);
  }
  if (count && pm.lay) mig_list(pm, pd);
}

#define INST_PUSH64(MV, BCV, W)                                                                  \
  template __global__ void k_push_f64<MV, BCV, W>(MeshView, Faces, Species, const double*, const double*, \
                                                  const Tile*, const int*, int, double, int64_t, double*, \
                                                  double*, const uint8_t*, int*, int8_t*, const double*, int*, \
                                                  PushSrc, PushMig);
INST_PUSH64(AMPS_MOVER_LAPENTA, false, 1)
INST_PUSH64(AMPS_MOVER_BORIS, false, 1)
INST_PUSH64(AMPS_MOVER_LAPENTA, false, 4)
INST_PUSH64(AMPS_MOVER_BORIS, false, 4)
INST_PUSH64(AMPS_MOVER_RELATIVISTIC, false, 1)
INST_PUSH64(AMPS_MOVER_LAPENTA, true, 1)
INST_PUSH64(AMPS_MOVER_BORIS, true, 1)

// FP32 push: float E/B stencil in LDS (17.5 KB), float velocities, FP64
// positions.
template <int MOVER, int W>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) void k_push_f32(MeshView m, Faces faces, Species sp,
                                                   const double* __restrict__ E,
                                                   const double* __restrict__ B,
                                                   const Tile* __restrict__ tiles,
                                                   const int* __restrict__ ntiles, int bc_mode,
                                                   double r_sphere, int64_t cap, double* x,
                                                   float* v, const uint8_t* __restrict__ spec,
                                                   int* key, int8_t* code, int* __restrict__ count, PushSrc ps,
                                                   PushMig pm) {
  __shared__ float sE[NCORNER * 3];
  __shared__ float sB[NCORNER * 3];
  __shared__ int sLay[NCELL + 1];
  const int t = blockIdx.x;
  if (t >= *ntiles) return;
  const Tile tl = tiles[t];
  const int b = tl.block;
  const double* Eb = E + (size_t)b * NCORNER * 3;
  const double* Bb = B + (size_t)b * NCORNER * 3;
  for (int i = threadIdx.x; i < NCORNER * 3; i += blockDim.x) {
    sE[i] = (float)Eb[i];
    sB[i] = (float)Bb[i];
  }
  if (pm.lay)
    for (int i = threadIdx.x; i <= NCELL; i += blockDim.x) sLay[i] = pm.lay[(int64_t)b * NCELL + i];
  __syncthreads();
  BlockGeom g;
  load_block_geom(m, b, g);
  BlockLat lat;
  load_block_lat(m, b, lat);
  const int p1 = tl.start + tl.count;
  const int pa = pm.lay ? (tl.start & ~63) : tl.start;
  MigPend pd = {0, 0, 0, 0, 0, false};
  const int pf = pa + (int)threadIdx.x;
  int qn = (ps.src && pf >= tl.start && pf < p1) ? ps.src[pf] : 0;
  for (int p0 = pa; p0 < p1; p0 += blockDim.x) {
    const int p = p0 + threadIdx.x;
    bool alive = false;
    int knew = -1;
    const int qs = qn;
    if (ps.src && p + (int)blockDim.x < p1) qn = ps.src[p + blockDim.x];
    if (p >= tl.start && p < p1) {
    const int q = ps.src ? qs : p;
    const double* xr = ps.src ? ps.x : x;
    const float* vr = ps.src ? (const float*)ps.v : v;
    double xp[3] = {xr[q], xr[cap + q], xr[2 * cap + q]};
    float vp[3] = {vr[q], vr[cap + q], vr[2 * cap + q]};
    const int s = ps.src ? ps.spec[q] : spec[p];
    if (ps.src) {
      ps.spec_out[p] = (uint8_t)s;
      ps.id_out[p] = ps.id[q];
    }
    int nb_out = b, cell_out = 0;
    const int c = push_particle_f32<MOVER>(m, faces, g, lat, sE, sB, (float)sp.charge[s], (float)sp.mass[s],
                                           sp.dt[s], bc_mode, r_sphere, xp, vp, &nb_out, &cell_out);
    x[p] = xp[0];
    x[cap + p] = xp[1];
    x[2 * cap + p] = xp[2];
    v[p] = vp[0];
    v[cap + p] = vp[1];
    v[2 * cap + p] = vp[2];
    alive = alive_code(c);
    knew = nb_out * NCELL + cell_out;
    // an exit of the USER_FUNCTION mode keeps the block holding its face state
    if (alive || c >= AMPS_CODE_EXIT_FACE0) key[p] = knew;
    code[p] = (int8_t)c;
    }
    if (count) push_count(count, pm, sLay, b, p, tl.start, p1, alive, knew, pd);
  }
  if (count && pm.lay) mig_list(pm, pd);
}

#define INST_PUSH32(MV, W)                                                                        \
  template __global__ void k_push_f32<MV, W>(MeshView, Faces, Species, const double*, const double*, \
                                             const Tile*, const int*, int, double, int64_t, double*,  \
                                             float*, const uint8_t*, int*, int8_t*, int*, PushSrc,   \
                                             PushMig);
INST_PUSH32(AMPS_MOVER_LAPENTA, 1)
INST_PUSH32(AMPS_MOVER_BORIS, 1)
INST_PUSH32(AMPS_MOVER_LAPENTA, 4)
INST_PUSH32(AMPS_MOVER_BORIS, 4)

// ======================================================================
// Migration: counting sort by key = block*512 + cell with wavefront
// run-length aggregation (ballot + lane masks) so that one atomic serves a
// whole run of equal keys; deleted particles are dropped.
// ======================================================================
__global__ __launch_bounds__(256) void k_fill_int(int* a, int64_t n, int val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    a[i] = val;
}

__global__ __launch_bounds__(256) void k_count_keys(const int* __restrict__ n_dev,
                                                     const int* __restrict__ key,
                                                     const int8_t* __restrict__ code,
                                                     int* count) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int n = *n_dev;
  const bool valid = (p < n) && alive_code(code[p]);
  const int k = valid ? key[p] : -1;
  const RunInfo r = wave_runs(valid, k);
  if (r.head) atomicAdd(&count[k], r.len);
}

__global__ __launch_bounds__(256) void k_scatter(const int* __restrict__ n_dev,
                                                  const int* __restrict__ key,
                                                  const int8_t* __restrict__ code,
                                                  const int* __restrict__ offset, int* cursor,
                                                  int64_t cap, const double* __restrict__ xa,
                                                  const double* __restrict__ va, int vf32,
                                                  const uint8_t* __restrict__ sa,
                                                  const int64_t* __restrict__ ia, double* xb,
                                                  double* vb, uint8_t* sb, int* kb,
                                                  int64_t* ib) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int n = *n_dev;
  const bool valid = (p < n) && alive_code(code[p]);
  const int k = valid ? key[p] : -1;
  const RunInfo r = wave_runs(valid, k);
  int base = 0;
  if (r.head) base = atomicAdd(&cursor[k], r.len);
  base = __shfl(base, r.head_lane);
  if (!valid) return;
  const int64_t d = (int64_t)offset[k] + base + (__lane_id() - r.head_lane);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    xb[c * cap + d] = xa[c * cap + p];
    if (vf32)
      ((float*)vb)[c * cap + d] = ((const float*)va)[c * cap + p];
    else
      vb[c * cap + d] = va[c * cap + p];
  }
  sb[d] = sa[p];
  kb[d] = k;
  ib[d] = ia[p];
}

// ----------------------------------------------------------------------
// Deterministic migration.  The particles enter a compaction in the layout of
// the previous one (sorted by key; lay[] its offsets), so the particles that
// keep their key (stayers) are exactly those that still sit inside their
// key's old run.  Each key's new run takes its stayers first, in array order,
// then the particles that moved in (movers), in array order, whatever order
// the atomics land in:
//   classify: one stayer mask and count per wavefront; each mover lists its
//             position at the end of its new key's region of tmpsrc, where
//             the key's movers go (backwards from the end, in atomic order);
//   order:    per key, the movers' positions sorted in place (one lane per
//             key, a workgroup merge sort for keys with many movers), which
//             makes tmpsrc over the movers' destinations the source index of
//             each; when the records are moved by scatter, each mover's
//             destination is also written to dest[];
//   scatter:  a stayer's destination is its key's offset plus the stayers
//             between its key's old run start and itself, from the scanned
//             wavefront counts and masks; a mover's is dest[p].
// ----------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mig_classify(const int* __restrict__ n_dev, const int* __restrict__ key,
                                                       const int8_t* __restrict__ code, const int* __restrict__ lay,
                                                       int has_layout, const int* __restrict__ offset, int* cursor,
                                                       int* __restrict__ tmpsrc, int* __restrict__ wcnt,
                                                       uint64_t* __restrict__ wmask) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int n = *n_dev;
  const bool valid = (p < n) && alive_code(code[p]);
  const int k = valid ? key[p] : -1;
  const bool st = valid && has_layout && p >= lay[k] && p < lay[k + 1];
  const uint64_t m = __ballot(st);
  const bool mover = valid && !st;
  const RunInfo ru = wave_runs(mover, mover ? k : -1);
  int base = 0;
  if (ru.head) base = atomicAdd(&cursor[k], ru.len);
  base = __shfl(base, ru.head_lane);
  if (__lane_id() == 0) {
    wcnt[p >> 6] = __popcll(m);
    wmask[p >> 6] = m;
  }
  if (mover) tmpsrc[(int64_t)offset[k + 1] - 1 - (base + (__lane_id() - ru.head_lane))] = (int)p;
}

// the movers the push listed past its slab: rank r of key k's m movers at
// offset[k + 1] - m + r of tmpsrc (the key's movers in rank order)
__global__ __launch_bounds__(256) void k_mig_place(const int* __restrict__ novf, const int2* __restrict__ ovf,
                                                    const int* __restrict__ key, const int* __restrict__ offset,
                                                    const int* __restrict__ cursor, int* __restrict__ tmpsrc) {
  const int n = *novf;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int2 e = ovf[i];
    const int k = key[e.x];
    tmpsrc[(int64_t)offset[k + 1] - cursor[k] + e.y] = e.x;
  }
}

// a[i] += b[i]
__global__ __launch_bounds__(256) void k_add_ints(int64_t n, int* __restrict__ a, const int* __restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] += b[i];
}

// the destination base of key k's stayers: its new offset less the stayers
// (of any key) before its old run (a key with an empty old run has none)
__device__ __forceinline__ void key_base(int64_t k, const int* __restrict__ offset, const int* __restrict__ lay,
                                         const int* __restrict__ wscan, const uint64_t* __restrict__ wmask,
                                         int* __restrict__ kbase) {
  const int q = lay[k];
  if (lay[k + 1] > q)
    kbase[k] = offset[k] - (wscan[q >> 6] + __popcll(wmask[q >> 6] & ((1ull << (q & 63)) - 1ull)));
}

// one lane's insertion sort of a key's m movers at q = tmpsrc + d0 (nearly
// ascending, as the atomics ran, when it starts: linear time), and with
// write_dest each mover's destination in dest[]
__device__ __forceinline__ void sort_movers(int* __restrict__ q, int m, int d0, int write_dest,
                                            int* __restrict__ dest) {
  for (int i = 1; i < m; i++) {
    const int t = q[i];
    int j = i - 1;
    while (j >= 0 && q[j] > t) {
      q[j + 1] = q[j];
      j--;
    }
    q[j + 1] = t;
  }
  if (write_dest)
    for (int i = 0; i < m; i++) dest[q[i]] = d0 + i;
}

// per key: the movers' source positions (the last cursor[k] entries of the
// key's region of tmpsrc, its movers' destinations) in ascending order, and
// with write_dest each mover's destination in dest[].  Up to `small` (at most
// 64) are sorted by the key's own lane; keys with more are listed in big[]
// (count in big[nkeys]) for k_mig_order_big.
__global__ __launch_bounds__(256) void k_mig_order(int64_t nkeys, int small, int write_dest,
                                                    const int* __restrict__ offset,
                                                    const int* __restrict__ cursor, int* __restrict__ tmpsrc,
                                                    int* __restrict__ dest, int* __restrict__ big, int has_layout,
                                                    const int* __restrict__ lay, const int* __restrict__ wscan,
                                                    const uint64_t* __restrict__ wmask, int* __restrict__ kbase) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  if (has_layout) key_base(k, offset, lay, wscan, wmask, kbase);
  const int m = cursor[k];
  if (m == 0) return;
  if (m > small) {
    big[atomicAdd(&big[nkeys], 1)] = (int)k;
    return;
  }
  const int d0 = offset[k + 1] - m;
  // the classification listed them backwards: reversed, they are nearly
  // ascending
  int* q = tmpsrc + d0;
  for (int i = 0, j = m - 1; i < j; i++, j--) {
    const int t = q[i];
    q[i] = q[j];
    q[j] = t;
  }
  sort_movers(q, m, d0, write_dest, dest);
}

// The same for movers the push listed (PushMig): the key's lane places its
// slab entries beside the entries k_mig_place put past them (rank r at
// d0 + r) and sorts them, or past `small` lists the key for k_mig_order_big.
__global__ __launch_bounds__(256) void k_mig_order_slab(int64_t nkeys, int small, int write_dest,
                                                         const int* __restrict__ slab,
                                                         const int* __restrict__ offset,
                                                         const int* __restrict__ cursor, int* __restrict__ tmpsrc,
                                                         int* __restrict__ dest, int* __restrict__ big,
                                                         const int* __restrict__ lay, const int* __restrict__ wscan,
                                                         const uint64_t* __restrict__ wmask,
                                                         int* __restrict__ kbase) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= nkeys) return;
  key_base(k, offset, lay, wscan, wmask, kbase);
  const int m = cursor[k];
  if (m == 0) return;
  const int d0 = offset[k + 1] - m;
  const int ms = min(m, MIG_SLAB);
  for (int j = 0; j < ms; j++) tmpsrc[d0 + j] = slab[j * nkeys + k];
  if (m > small) {
    big[atomicAdd(&big[nkeys], 1)] = (int)k;
    return;
  }
  sort_movers(tmpsrc + d0, m, d0, write_dest, dest);
}

// number of entries of the ascending run r[0, n) below t
__device__ __forceinline__ int count_below(const int* r, int n, int t) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (r[mid] < t)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// the keys with many movers, one workgroup each (a grid-stride loop over the
// list): a bottom-up merge sort of the distinct source positions, every
// element placed by a binary search in the partner run, ping-ponging between
// tmpsrc and a spare array of the same size; the sorted run ends in tmpsrc
__global__ __launch_bounds__(256) void k_mig_order_big(int64_t nkeys, int write_dest,
                                                        const int* __restrict__ offset,
                                                        const int* __restrict__ cursor, int* tmpsrc, int* spare,
                                                        int* __restrict__ dest, const int* __restrict__ big) {
  const int nbig = big[nkeys];
  for (int e = blockIdx.x; e < nbig; e += gridDim.x) {
    const int k = big[e];
    const int m = cursor[k];
    const int d0 = offset[k + 1] - m;
    int* src = tmpsrc + d0;
    int* dst = spare + d0;
    for (int w = 1; w < m; w *= 2) {
      for (int i = threadIdx.x; i < m; i += blockDim.x) {
        const int r = i / w, s = (r >> 1) * 2 * w;
        const int t = src[i];
        int pos;
        if ((r & 1) == 0) {
          const int b0 = min(s + w, m), bn = min(s + 2 * w, m) - b0;
          pos = (i - s) + count_below(src + b0, bn, t);
        } else {
          pos = (i - s - w) + count_below(src + s, w, t);
        }
        dst[s + pos] = t;
      }
      __syncthreads();
      int* tmp = src;
      src = dst;
      dst = tmp;
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const int t = src[i];
      if (src != tmpsrc + d0) tmpsrc[d0 + i] = t;
      if (write_dest) dest[t] = d0 + i;
    }
    __syncthreads();
  }
}

// Where particle p (key k) goes: a stayer after the stayers of its key that
// precede it (the stayers before it, from its wave's stayer mask and the
// scanned per-wave counts, past its key's base from k_mig_order), a mover
// where k_mig_order put it.
__device__ __forceinline__ int64_t mig_dest(int64_t p, int k, const int* __restrict__ kbase,
                                            const int* __restrict__ wscan, const uint64_t* __restrict__ wmask,
                                            const int* __restrict__ dest) {
  const int lane = (int)(p & 63);
  const uint64_t mw = wmask[p >> 6];
  if ((mw >> lane) & 1ull) return (int64_t)kbase[k] + wscan[p >> 6] + __popcll(mw & ((1ull << lane) - 1ull));
  return dest[p];
}

// The same destinations without moving the records: src[d] = p, for a
// deposit cell pass that reads the particles where they are and a push that
// moves them (the push writes every surviving particle's key; the keys stay
// in the previous arrays until then, where k_mig_move finds them).  The
// movers' entries of src are the sorted lists k_mig_order left in place, so
// only the stayers are written here.
__global__ __launch_bounds__(256) void k_mig_dest(const int* __restrict__ n_dev, const int* __restrict__ key,
                                                  const int* __restrict__ kbase, const int* __restrict__ wscan,
                                                  const uint64_t* __restrict__ wmask, int64_t cap,
                                                  int* __restrict__ src) {
  // four consecutive positions per thread (one mask word per 16 threads), so
  // that each wavefront has four times the loads in flight
  const int64_t p0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4;
  const int n = *n_dev;
  if (p0 >= n) return;
  const uint64_t mw = wmask[p0 >> 6];
  const int sh = (int)(p0 & 63);
  const unsigned bits = (unsigned)(mw >> sh) & 15u;
  if (!bits) return;
  int kk[4];
  if (p0 + 4 <= cap) {
    const int4 k4 = *(const int4*)(key + p0);
    kk[0] = k4.x;
    kk[1] = k4.y;
    kk[2] = k4.z;
    kk[3] = k4.w;
  } else {
#pragma unroll
    for (int j = 0; j < 4; j++) kk[j] = (p0 + j < cap) ? key[p0 + j] : 0;
  }
  // a stayer is alive and kept its key (its bit: no position at or past n has one)
  const int before = wscan[p0 >> 6] + __popcll(mw & ((1ull << sh) - 1ull));
  int r = 0;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    if ((bits >> j) & 1u) {
      src[(int64_t)kbase[kk[j]] + before + r] = (int)(p0 + j);
      r++;
    }
  }
}

// The record move that k_mig_dest left: particle d of the new order from
// src[d] of the previous arrays.
__global__ __launch_bounds__(256) void k_mig_move(const int* __restrict__ n_dev, const int* __restrict__ src,
                                                  int64_t cap, const double* __restrict__ xa,
                                                  const double* __restrict__ va, int vf32,
                                                  const uint8_t* __restrict__ sa, const int64_t* __restrict__ ia,
                                                  const int* __restrict__ ka, double* __restrict__ xb,
                                                  double* __restrict__ vb, uint8_t* __restrict__ sb,
                                                  int64_t* __restrict__ ib, int* __restrict__ kb) {
  const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (d >= *n_dev) return;
  const int64_t p = src[d];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    xb[c * cap + d] = xa[c * cap + p];
    if (vf32)
      ((float*)vb)[c * cap + d] = ((const float*)va)[c * cap + p];
    else
      vb[c * cap + d] = va[c * cap + p];
  }
  sb[d] = sa[p];
  ib[d] = ia[p];
  kb[d] = ka[p];
}

__global__ __launch_bounds__(256) void k_mig_scatter(const int* __restrict__ n_dev, const int* __restrict__ key,
                                                      const int8_t* __restrict__ code, const int* __restrict__ kbase,
                                                      const int* __restrict__ wscan,
                                                      const uint64_t* __restrict__ wmask,
                                                      const int* __restrict__ dest, int64_t cap,
                                                      const double* __restrict__ xa, const double* __restrict__ va,
                                                      int vf32, const uint8_t* __restrict__ sa,
                                                      const int64_t* __restrict__ ia, double* __restrict__ xb,
                                                      double* __restrict__ vb, uint8_t* __restrict__ sb,
                                                      int* __restrict__ kb, int64_t* __restrict__ ib) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int n = *n_dev;
  const bool valid = (p < n) && alive_code(code[p]);
  if (!valid) return;
  const int k = key[p];
  const int64_t d = mig_dest(p, k, kbase, wscan, wmask, dest);
#pragma unroll
  for (int c = 0; c < 3; c++) {
    xb[c * cap + d] = xa[c * cap + p];
    if (vf32)
      ((float*)vb)[c * cap + d] = ((const float*)va)[c * cap + p];
    else
      vb[c * cap + d] = va[c * cap + p];
  }
  sb[d] = sa[p];
  kb[d] = k;
  ib[d] = ia[p];
}

__global__ __launch_bounds__(256) void k_copy_ints(int64_t n, const int* __restrict__ src, int* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_copy_total(const int* __restrict__ src, int* dst) { *dst = *src; }

// tiles of at most T particles, never straddling a block
__global__ __launch_bounds__(256) void k_tile_count(int nb, int T, const int* __restrict__ offset,
                                                     int* tcount) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) {
    const int cnt = offset[(b + 1) * NCELL] - offset[b * NCELL];
    tcount[b] = (cnt + T - 1) / T;
  } else if (b == nb) {
    tcount[nb] = 0;
  }
}

__global__ __launch_bounds__(256) void k_tile_emit(int nb, int T, const int* __restrict__ offset,
                                                    const int* __restrict__ toff, Tile* tiles,
                                                    int* ntiles) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b == 0) *ntiles = toff[nb];
  if (b >= nb) return;
  const int start = offset[b * NCELL];
  const int cnt = offset[(b + 1) * NCELL] - start;
  const int t0 = toff[b];
  for (int t = 0; t * T < cnt; t++) {
    Tile tl;
    tl.block = b;
    tl.start = start + t * T;
    tl.count = min(T, cnt - t * T);
    tl.pad = 0;
    tiles[t0 + t] = tl;
  }
}

// ======================================================================
// Device-wide exclusive scan (in place), three phases.
// ======================================================================
constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tile(int* data, int64_t n, int* sums) {
  __shared__ int s[SCAN_BLOCK];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
  int vals[SCAN_ITEMS];
  int tot = 0;
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    const int64_t j = base + i;
    vals[i] = (j < n) ? data[j] : 0;
    tot += vals[i];
  }
  s[threadIdx.x] = tot;
  __syncthreads();
  for (int off = 1; off < SCAN_BLOCK; off <<= 1) {
    int t = (threadIdx.x >= off) ? s[threadIdx.x - off] : 0;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  int run = s[threadIdx.x] - tot;  // exclusive prefix of this thread
  if (threadIdx.x == SCAN_BLOCK - 1) sums[blockIdx.x] = s[threadIdx.x];
#pragma unroll
  for (int i = 0; i < SCAN_ITEMS; i++) {
    const int64_t j = base + i;
    if (j < n) data[j] = run;
    run += vals[i];
  }
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_add(int* data, int64_t n,
                                                          const int* __restrict__ sums) {
  const int add = sums[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  for (int i = threadIdx.x; i < SCAN_TILE; i += SCAN_BLOCK) {
    const int64_t j = base + i;
    if (j < n) data[j] += add;
  }
}

namespace amps {
static int scan_rec(amps_ctx* c, int* data, int64_t n, int* tmp) {
  const int64_t nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, c->stream, data, n, tmp);
  if (nt > 1) {
    int r = scan_rec(c, tmp, nt, tmp + nt);
    if (r) return r;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nt), dim3(SCAN_BLOCK), 0, c->stream, data, n,
                       tmp);
  }
  return 0;
}

int scan_exclusive(amps_ctx* c, int* data, int64_t n, int* /*total_out*/) {
  // temporary space: sum over levels of ceil(n / TILE^k) < n/(TILE-1) + levels
  int64_t need = 0, m = n;
  while (m > 1) {
    m = (m + SCAN_TILE - 1) / SCAN_TILE;
    need += m;
  }
  need += 8;
  if (need > c->scan_tmp_len) {
    if (c->d_scan_tmp) (void)hipFree(c->d_scan_tmp);
    c->d_scan_tmp = nullptr;
    HIP_TRY(hipMalloc(&c->d_scan_tmp, sizeof(int) * need));
    c->scan_tmp_len = need;
  }
  return scan_rec(c, data, n, c->d_scan_tmp);
}
}  // namespace amps

// ======================================================================
// ECSIM deposit, phase 1: per-cell accumulation (ProcessCell,
// pic_field_solver_ecsim.cpp:1882-2438).  One workgroup per 4x4x4 cell
// sub-tile of a block; wave w owns output chunk w of the 348 per-cell
// values and lane l owns cell l of the sub-tile, walking that cell's
// particles in order with the chunk's accumulators held in registers.
// ======================================================================
template <int NCH, bool EXACT>
__global__ __launch_bounds__(64 * NCH) void k_deposit_cell(MeshView m, Species sp, double dtTotal,
                                                            const double* __restrict__ B,
                                                            const int* __restrict__ offset,
                                                            int64_t cap,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ v, int vf32,
                                                            const uint8_t* __restrict__ spec,
                                                            double* __restrict__ cellpart) {
  __shared__ double sB[125 * 3];
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
  __syncthreads();
  BlockGeom g;
  load_block_geom(m, b, g);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ci = sx + (lane & 3), cj = sy + ((lane >> 2) & 3), ck = sz + (lane >> 4);
  const int key = b * NCELL + ci + NC * (cj + NC * ck);
  const int p0 = offset[key], p1 = offset[key + 1];
  double dxc[3], CellVolume = 1;
  for (int d = 0; d < 3; d++) dxc[d] = (g.xmax[d] - g.xmin[d]) / NC * 1.0;
  for (int d = 0; d < 3; d++) CellVolume *= dxc[d];
  double* out = cellpart + (size_t)key * CELL_VALS;
  dispatch_chunk<NCH, EXACT>(w, g, sp, dtTotal, sB, sx, sy, sz, p0, p1, cap, x, v, vf32, spec,
                             CellVolume, out);
}

#define INST_DEP(NCH, EX)                                                                  \
  template __global__ void k_deposit_cell<NCH, EX>(MeshView, Species, double, const double*, \
                                                    const int*, int64_t, const double*,     \
                                                    const double*, int, const uint8_t*, double*);
INST_DEP(2, true)
INST_DEP(2, false)
INST_DEP(3, true)
INST_DEP(3, false)
INST_DEP(4, true)
INST_DEP(4, false)
INST_DEP(6, true)
INST_DEP(6, false)

// ======================================================================
// ECSIM deposit, phase 2: corner flush.  One wavefront per unique node sums
// the contributions of its incident cells in the order of its incidence row
// into J (3) and the 27x9 mass matrix row block (IndexMatrix,
// pic_field_solver_ecsim.cpp:1377-1379).  The exact records (MOMENT false)
// come with rows in ascending cell order, the order of the reference's
// serial cell loop, so the sums reproduce it bit for bit; the moment
// records with rows in octant order (k_deposit_node_staged's).
// ======================================================================
// c_soff[q][val]: where output value val (J[3], then 27 slots x 9) of a node
// that is local corner q of an incident cell sits in that cell's record
// (J of corner q at 3q; the pair block (max(q,j), min(q,j)) at
// 24 + 9 (hi (hi+1)/2 + lo)), -1 where the cell does not touch that
// neighbour slot.  Built on the host from IndexMatrix.
// c_soff_m: the same for the moment record (pair (q, j) -> moment of the
// per-axis bit sums of q and j, at 24 + 9 t).
__constant__ __align__(16) short c_soff[8][256];
__constant__ __align__(16) short c_soff_m[8][256];

template <bool MOMENT>
__global__ __launch_bounds__(256) void k_deposit_node(int64_t nnodes, int npw,
                                                       const int* __restrict__ node_order,
                                                       const int* __restrict__ inc8,
                                                       const double* __restrict__ cellpart,
                                                       double* __restrict__ nodeout) {
  constexpr int STRIDE = MOMENT ? CELL_STRIDE_MOMENT : CELL_VALS;
  __shared__ __align__(16) short soff[8][256];
  static_assert(sizeof(soff) == 256 * 16, "one 16-byte row per thread");
  reinterpret_cast<int4*>(&soff[0][0])[threadIdx.x] =
      reinterpret_cast<const int4*>(MOMENT ? &c_soff_m[0][0] : &c_soff[0][0])[threadIdx.x];
  __syncthreads();
  // consecutive nodes share incident cells: keep them on one XCD.  The node,
  // its incident cells and their records are the same for the whole
  // wavefront, so their addresses live in scalar registers and each lane adds
  // only its 32-bit offset.
  // Each wavefront takes npw consecutive nodes of the processing order.
  const int64_t first = ((int64_t)xcd_tile() * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6)) * npw;
  const int64_t last = (first + npw < nnodes) ? first + npw : nnodes;
  const int lane = threadIdx.x & 63;
  for (int64_t slot = first; slot < last; slot++) {
  int ics[8];
#pragma unroll
  for (int t = 0; t < 8; t++) ics[t] = inc8[slot * 8 + t];
  const int64_t node = node_order[slot];
  // all loads first, then the sums in row order (x + -0.0 == x, so absent
  // slots change nothing)
  double vv[8][4];
#pragma unroll
  for (int t = 0; t < 8; t++) {
    const int ic = ics[t];
    if (ic >= 0) {
      const short* so = soff[ic & 7];
      const double* P = cellpart + (size_t)(ic >> 3) * STRIDE;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int off = so[lane + 64 * r];
        const double val = P[off < 0 ? 0 : off];
        vv[t][r] = (off >= 0) ? val : -0.0;
      }
    }
  }
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 8; t++)
    if (ics[t] >= 0) {
#pragma unroll
      for (int r = 0; r < 4; r++) acc[r] += vv[t][r];
    }
  // the whole 2 KB record, padding included: four full-line stores
  double* o = nodeout + (size_t)node * NODE_STRIDE;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int val = lane + 64 * r;
    o[val] = (val < NODE_VALS) ? acc[r] : 0.0;
  }
  }
}

template __global__ void k_deposit_node<false>(int64_t, int, const int*, const int*, const double*, double*);
template __global__ void k_deposit_node<true>(int64_t, int, const int*, const int*, const double*, double*);

// Node gather for the moment record, staged.  The incidence rows come in
// octant order (slot t: the cell on the node's side (t & 1, t >> 1 & 1,
// t >> 2), of which the node is a fixed corner; an empty octant points at a
// record of zeros).  A node that is corner q of a cell needs 75 of the cell's
// 267 values: J of corner q and the moments t = b_q + b_j of its 8 neighbours
// j, which lie in four runs of 18 consecutive values.  The wavefront first
// copies the 8 x 75 values its node needs into LDS with loads whose lanes
// walk those runs (a handful of cache lines per load instead of one per
// lane), then each lane sums its outputs from LDS in octant order.  Which
// value a lane loads and which LDS slot each of its outputs reads depend on
// the lane alone, so they are worked out once per wavefront.  Same operands
// in the same order as k_deposit_node<true> over the same rows: bit-identical.
// c_gtab[t][w]: offset in the record of octant t's cell of gathered value w
// (w < 72: run w / 18 of the 2x2x2 moment block above the node's corner,
// w >= 72: J of that corner).
// c_wtab[t][val]: the gathered value that output val takes from octant t's
// cell, -1 where that cell does not touch the slot (val >= 246: padding).
__constant__ __align__(16) short c_gtab[8][80];
__constant__ __align__(16) short c_wtab[8][256];

constexpr int NODE_G = 75;  // values one incident cell gives a node

__global__ __launch_bounds__(256) void k_deposit_node_staged(int64_t nnodes, int npw,
                                                             const int* __restrict__ node_order,
                                                             const int* __restrict__ inc8,
                                                             const double* __restrict__ cellpart,
                                                             double* __restrict__ nodeout) {
  __shared__ double G[4][8 * NODE_G + 1];  // the gathered values, then a zero
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  double* Gw = G[wv];
  // npw > 0: workgroup g takes nodes [4 g npw, 4 (g+1) npw), its wavefronts
  // interleaved (wavefront w: the w-th, (w+4)-th, ... of them), so that the
  // four nodes a workgroup has in flight are consecutive and read the cells
  // they share (a node and the next one share four) while those are cached.
  // npw == 0 (a resident grid): XCD x takes the x-th eighth of the nodes and
  // its G wavefronts stride over it, wavefront i taking i, i + G, i + 2G, ...
  int64_t first, last, stride;
  if (npw > 0) {
    const int64_t g0 = (int64_t)xcd_tile() * 4 * npw;
    first = g0 + wv;
    last = (g0 + 4 * npw < nnodes) ? g0 + 4 * npw : nnodes;
    stride = 4;
  } else {
    const int64_t part = (nnodes + N_XCD - 1) / N_XCD;
    const int64_t x0 = (int64_t)(blockIdx.x % N_XCD) * part;
    first = x0 + (int64_t)(blockIdx.x / N_XCD) * 4 + wv;
    last = (x0 + part < nnodes) ? x0 + part : nnodes;
    stride = (int64_t)(gridDim.x / N_XCD) * 4;
  }
  if (first >= last) return;
  // this lane's gathered values e = lane + 64 i: value e % 75 of octant e / 75;
  // step i covers octants t0(i) and t0(i) + 1, hib bit i: the second one
  constexpr int NI = (8 * NODE_G + 63) / 64;
  int goff[NI];
  uint32_t hib = 0;
#pragma unroll
  for (int i = 0; i < NI; i++) {
    const int e = lane + 64 * i;
    const int t0 = (64 * i) / NODE_G;
    const int t = e < 8 * NODE_G ? e / NODE_G : t0;
    goff[i] = e < 8 * NODE_G ? c_gtab[t][e - NODE_G * t] : 0;
    hib |= (t != t0 ? 1u : 0u) << i;
  }
  // the LDS slot each output val = lane + 64 r takes from octant t (the zero
  // where the octant's cell does not touch it)
  const double* src[8][4];
#pragma unroll
  for (int t = 0; t < 8; t++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int w = c_wtab[t][lane + 64 * r];
      src[t][r] = Gw + (w < 0 ? 8 * NODE_G : NODE_G * t + w);
    }
  if (lane == 0) Gw[8 * NODE_G] = 0.0;
  // The next node's gathered values are loaded into registers while the
  // current node is summed (two buffers, the loop unrolled by two so that
  // neither is copied while its loads are in flight; past the last node the
  // last one is loaded again).
  struct Buf {
    double v[NI];
    int64_t node;
  };
  auto gather = [&](int64_t slot, Buf& B) {
    int ics[8];
#pragma unroll
    for (int t = 0; t < 8; t++) ics[t] = inc8[slot * 8 + t];
    B.node = node_order[slot];
#pragma unroll
    for (int i = 0; i < NI; i++) {
      // the cell of octant t0 or t0 + 1, chosen by masking (not by indexing
      // ics[] with a lane-dependent subscript)
      const int t0 = (64 * i) / NODE_G;
      const int clo = ics[t0] >> 3, chi = ics[t0 + 1 < 8 ? t0 + 1 : 7] >> 3;
      const int cell = clo + ((chi - clo) & -(int)((hib >> i) & 1));
      B.v[i] = cellpart[(size_t)cell * CELL_STRIDE_MOMENT + goff[i]];
    }
  };
  auto finish = [&](const Buf& B) {
#pragma unroll
    for (int i = 0; i < NI; i++) {
      const int e = lane + 64 * i;
      if (e < 8 * NODE_G) Gw[e] = B.v[i];
    }
    asm volatile("" ::: "memory");  // the DS ops of a wavefront execute in issue order
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
      for (int r = 0; r < 4; r++) acc[r] += *src[t][r];
    asm volatile("" ::: "memory");
    // node records are not read again by this step: non-temporal stores keep
    // them from displacing the cell records still being gathered
    double* o = nodeout + (size_t)B.node * NODE_STRIDE;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int val = lane + 64 * r;
      __builtin_nontemporal_store((val < NODE_VALS) ? acc[r] : 0.0, &o[val]);
    }
  };
  const int64_t final_slot = first + (last - 1 - first) / stride * stride;
  auto clampslot = [&](int64_t s2) { return s2 < final_slot ? s2 : final_slot; };
  Buf A, Bb;
  gather(first, A);
  for (int64_t slot = first; slot < last; slot += 2 * stride) {
    gather(clampslot(slot + stride), Bb);
    finish(A);
    if (slot + stride >= last) break;
    gather(clampslot(slot + 2 * stride), A);
    finish(Bb);
  }
}

namespace amps {
int upload_inv_index_matrix(hipStream_t s) {
  // IndexMatrix, pic_field_solver_ecsim.cpp:1377-1379
  static const int IM[8][8] = {{0, 2, 8, 6, 18, 20, 26, 24}, {1, 0, 6, 7, 19, 18, 24, 25},
                               {4, 3, 0, 1, 22, 21, 18, 19}, {3, 5, 2, 0, 21, 23, 20, 18},
                               {9, 11, 17, 15, 0, 2, 8, 6},  {10, 9, 15, 16, 1, 0, 6, 7},
                               {13, 12, 9, 10, 4, 3, 0, 1},  {12, 14, 11, 9, 3, 5, 2, 0}};
  // corner offset bits (x, y, z) in InterpolationCoefficientTable order
  static const int CB[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                               {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  short soff[8][256], soffm[8][256];
  for (int q = 0; q < 8; q++) {
    int inv[27];
    for (int s2 = 0; s2 < 27; s2++) inv[s2] = -1;
    for (int j = 0; j < 8; j++) inv[IM[q][j]] = j;
    for (int val = 0; val < 256; val++) {
      int off = -1, offm = -1;
      if (val < 3) {
        off = offm = 3 * q + val;
      } else if (val < NODE_VALS) {
        const int s2 = (val - 3) / 9, k = (val - 3) % 9, j = inv[s2];
        if (j >= 0) {
          const int hi = q > j ? q : j, lo = q > j ? j : q;
          off = 24 + 9 * (hi * (hi + 1) / 2 + lo) + k;
          const int t = (CB[q][0] + CB[j][0]) + 3 * (CB[q][1] + CB[j][1]) + 9 * (CB[q][2] + CB[j][2]);
          offm = 24 + 9 * t + k;
        }
      }
      soff[q][val] = (short)off;
      soffm[q][val] = (short)offm;
    }
  }
  // staged gather tables for the moment record, by octant t of the cell
  // around the node (the node is corner q of it, the corner whose offset
  // bits are 1 - the octant's bits)
  short gtab[8][80], wtab[8][256];
  for (int t = 0; t < 8; t++) {
    int q = 0;
    while (CB[q][0] != 1 - (t & 1) || CB[q][1] != 1 - ((t >> 1) & 1) || CB[q][2] != 1 - (t >> 2)) q++;
    for (int w = 0; w < 80; w++) {
      int off = 0;
      if (w < 72) {
        const int run = w / 18, o = w % 18;
        off = 24 + 9 * (CB[q][0] + 3 * (CB[q][1] + (run & 1)) + 9 * (CB[q][2] + (run >> 1))) + o;
      } else if (w < 75) {
        off = 3 * q + (w - 72);
      }
      gtab[t][w] = (short)off;
    }
    int inv[27];
    for (int s2 = 0; s2 < 27; s2++) inv[s2] = -1;
    for (int j = 0; j < 8; j++) inv[IM[q][j]] = j;
    for (int val = 0; val < 256; val++) {
      int w = -1;
      if (val < 3) {
        w = 72 + val;
      } else if (val < NODE_VALS) {
        const int s2 = (val - 3) / 9, k = (val - 3) % 9, j = inv[s2];
        if (j >= 0) w = 18 * (CB[j][1] + 2 * CB[j][2]) + 9 * CB[j][0] + k;
      }
      wtab[t][val] = (short)w;
    }
  }
  HIP_TRY(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_gtab), gtab, sizeof(gtab), 0, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_wtab), wtab, sizeof(wtab), 0, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_soff), soff, sizeof(soff), 0, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_soff_m), soffm, sizeof(soffm), 0, hipMemcpyHostToDevice, s));
  return 0;
}
}  // namespace amps

// ======================================================================
// Synthetic population on the device: counter-based RNG (SplitMix64 of
// seed and particle index), equal counts per cell with the first cells
// taking the remainder, uniform positions, Maxwellian velocities.
// ======================================================================
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }


__global__ __launch_bounds__(256) void k_generate(MeshView m, GenArgs ga, int64_t cap, double* x,
                                                   double* v, uint8_t* spec, int* key, int64_t* id,
                                                   int8_t* code) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ga.ntotal) return;
  const int64_t ns = ga.nspec;
  const int64_t big = ns * ga.rem * (ga.base + 1);
  int64_t c, j;
  if (i < big) {
    c = i / (ns * (ga.base + 1));
    j = i - c * ns * (ga.base + 1);
  } else {
    c = ga.rem + (i - big) / (ns * ga.base);
    j = i - big - (c - ga.rem) * ns * ga.base;
  }
  const int64_t nc = ga.base + (c < ga.rem ? 1 : 0);
  const int s = (int)(j / nc);
  const int b = (int)(c / NCELL), lc = (int)(c % NCELL);
  const int ci[3] = {lc % NC, (lc / NC) % NC, lc / (NC * NC)};
  double xp[3], vp[3];
  const uint64_t key0 = splitmix64(ga.seed ^ ((uint64_t)i * 0x632BE59BD9B4E019ull));
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double dx = m.blk_dxs[3 * b + d];
    const double lo = m.blk_xmin[3 * b + d] + ci[d] * dx;
    double xx = lo + u01(splitmix64(key0 + d)) * dx;
    if (!(xx < lo + dx)) xx = lo;
    xp[d] = xx;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double u1 = 1.0 - u01(splitmix64(key0 + 3 + 2 * d));
    const double u2 = u01(splitmix64(key0 + 4 + 2 * d));
    vp[d] = ga.drift[d] + ga.vth[s] * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  int nb = find_block(m, xp);
  int cell = (nb >= 0) ? find_cell(m, xp, nb) : -1;
  if (nb < 0 || cell < 0) {
    nb = b;
    cell = lc;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    x[d * cap + i] = xp[d];
    if (ga.vf32)
      ((float*)v)[d * cap + i] = (float)vp[d];
    else
      v[d * cap + i] = vp[d];
  }
  spec[i] = (uint8_t)s;
  key[i] = nb * NCELL + cell;
  id[i] = i;
  code[i] = AMPS_CODE_FINISHED;
}

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ a, float4* __restrict__ b,
                                                     int64_t n4) {
  // each workgroup copies consecutive 16 KB blocks: four 16-byte loads in
  // flight per lane before the stores, non-temporal both ways (the fastest
  // of the strided/blocked, 4/8-deep, cached/non-temporal forms measured)
  typedef float v4f __attribute__((ext_vector_type(4)));
  const v4f* av = reinterpret_cast<const v4f*>(a);
  v4f* bv = reinterpret_cast<v4f*>(b);
  constexpr int U = 4;
  const int64_t per = 256 * U;
  for (int64_t base = (int64_t)blockIdx.x * per; base < n4; base += (int64_t)gridDim.x * per) {
    v4f t[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + u * 256 + threadIdx.x;
      if (i < n4) t[u] = __builtin_nontemporal_load(av + i);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int64_t i = base + u * 256 + threadIdx.x;
      if (i < n4) __builtin_nontemporal_store(t[u], bv + i);
    }
  }
}
