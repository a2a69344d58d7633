// exits.hip — the exit list of the USER_FUNCTION boundary mode.
//
// In _PIC_PARTICLE_DOMAIN_BOUNDARY_INTERSECTION_PROCESSING_MODE__USER_FUNCTION_
// the reference mover finds the external face a particle left through, moves
// xInit/vInit onto it and calls ProcessOutsideDomainParticles(ptr, xInit,
// vInit, nIntersectionFace, newNode) on the spot (pic_mover_boris.cpp:
// 1156-1231, pic.h:6043-6044).  The push kernels here leave such a particle
// with code AMPS_CODE_EXIT_FACE0 + face, the relocated state in its x/v slots
// and the block holding that state in its key; after the push, three small
// kernels append those particles to a device list in device order, so that
// the host callback needs only the exits, not the whole population:
//   k_exit_count  — EX_WG workgroups, each counts the exits of one contiguous
//                   chunk of the particle array;
//   k_exit_offsets — one workgroup turns the counts into list offsets after
//                   the records already in the list (deterministic order);
//   k_exit_write  — each workgroup writes its exits in array order (wave
//                   ballots and a workgroup prefix over the waves).
// No host synchronisation: the step loop with this boundary mode can still be
// captured in a hipGraph, and the list then accumulates over the steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "amps_common.h"

using namespace amps;

namespace {

constexpr int EX_WG = 1024;   // workgroups of the count and write kernels
constexpr int EX_THREADS = 256;

__device__ __forceinline__ bool is_exit(int8_t code) {
  return code >= AMPS_CODE_EXIT_FACE0 && code < AMPS_CODE_EXIT_FACE0 + 6;
}

// chunk of the particle array that workgroup w covers: whole 256-particle
// steps, so every thread of a workgroup runs the same number of steps
__device__ __forceinline__ void ex_chunk(int n, int w, int& lo, int& hi) {
  const int steps = (n + EX_THREADS - 1) / EX_THREADS;
  const int per = (steps + EX_WG - 1) / EX_WG;
  lo = min(n, w * per * EX_THREADS);
  hi = min(n, (w + 1) * per * EX_THREADS);
}

__global__ __launch_bounds__(EX_THREADS) void k_exit_count(const int* __restrict__ n_dev,
                                                           const int8_t* __restrict__ code, int* wg) {
  __shared__ int part[EX_THREADS / 64];
  int lo, hi;
  ex_chunk(*n_dev, blockIdx.x, lo, hi);
  int cnt = 0;
  for (int p = lo + (int)threadIdx.x; p < hi; p += EX_THREADS) cnt += is_exit(code[p]) ? 1 : 0;
  // wave sums, then the workgroup's
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_down(cnt, o);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < EX_THREADS / 64; i++) t += part[i];
    wg[blockIdx.x] = t;
  }
}

// exclusive scan of the EX_WG counts (one workgroup of EX_WG threads), offset
// by the records already in the list; the list count advances by the total
// (records past the capacity are counted as dropped)
__global__ __launch_bounds__(EX_WG) void k_exit_offsets(int* wg, int64_t* state, int64_t capacity) {
  __shared__ int s[EX_WG];
  const int t = threadIdx.x;
  const int v = wg[t];
  s[t] = v;
  __syncthreads();
  for (int off = 1; off < EX_WG; off <<= 1) {
    const int a = (t >= off) ? s[t - off] : 0;
    __syncthreads();
    s[t] += a;
    __syncthreads();
  }
  const int64_t base = state[0];
  wg[t] = (int)(base + s[t] - v);
  if (t == EX_WG - 1) {
    const int64_t end = base + s[t];
    state[0] = min(end, capacity);
    if (end > capacity) state[1] += end - capacity;
  }
}

__global__ __launch_bounds__(EX_THREADS) void k_exit_write(const int* __restrict__ n_dev,
                                                           const int8_t* __restrict__ code,
                                                           const int* __restrict__ key,
                                                           const int* __restrict__ wg, int64_t cap,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ v, int vf32,
                                                           const uint8_t* __restrict__ spec,
                                                           const int64_t* __restrict__ id, int64_t capacity,
                                                           int64_t* ex_id, double* ex_x, double* ex_v,
                                                           int* ex_meta) {
  __shared__ int wsum[EX_THREADS / 64];
  int lo, hi;
  ex_chunk(*n_dev, blockIdx.x, lo, hi);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int64_t run = wg[blockIdx.x];
  for (int p0 = lo; p0 < hi; p0 += EX_THREADS) {
    const int p = p0 + (int)threadIdx.x;
    const int8_t cd = p < hi ? code[p] : (int8_t)0;
    const bool ex = p < hi && is_exit(cd);
    const uint64_t m = __ballot(ex);
    if (lane == 0) wsum[wave] = __popcll(m);
    __syncthreads();
    int before = 0, total = 0;
    for (int i = 0; i < EX_THREADS / 64; i++) {
      before += (i < wave) ? wsum[i] : 0;
      total += wsum[i];
    }
    __syncthreads();
    if (ex) {
      const uint64_t below = (lane == 0) ? 0ull : (m & ((1ull << lane) - 1ull));
      const int64_t d = run + before + __popcll(below);
      if (d < capacity) {
        ex_id[d] = id[p];
#pragma unroll
        for (int c = 0; c < 3; c++) {
          ex_x[3 * d + c] = x[c * cap + p];
          ex_v[3 * d + c] = vf32 ? (double)((const float*)v)[c * cap + p] : v[c * cap + p];
        }
        const int k = key[p];
        ex_meta[3 * d] = spec[p];
        ex_meta[3 * d + 1] = cd - AMPS_CODE_EXIT_FACE0;
        ex_meta[3 * d + 2] = k < 0 ? -1 : k / NCELL;
      }
    }
    run += total;
  }
}

int ensure_exit_list(amps_ctx* c) {
  if (c->exit_alloc == c->exit_cap && c->d_ex_state) return 0;
  exits_free(c);
  const int64_t n = c->exit_cap;
  HIP_TRY(hipMalloc(&c->d_ex_id, sizeof(int64_t) * n));
  HIP_TRY(hipMalloc(&c->d_ex_x, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&c->d_ex_v, sizeof(double) * 3 * n));
  HIP_TRY(hipMalloc(&c->d_ex_meta, sizeof(int) * 3 * n));
  HIP_TRY(hipMalloc(&c->d_ex_wg, sizeof(int) * EX_WG));
  HIP_TRY(hipMalloc(&c->d_ex_state, sizeof(int64_t) * 2));
  HIP_TRY(hipMemsetAsync(c->d_ex_state, 0, sizeof(int64_t) * 2, c->stream));
  c->exit_alloc = n;
  return 0;
}

}  // namespace

namespace amps {

void exits_free(amps_ctx* c) {
  (void)hipFree(c->d_ex_id);
  (void)hipFree(c->d_ex_x);
  (void)hipFree(c->d_ex_v);
  (void)hipFree(c->d_ex_meta);
  (void)hipFree(c->d_ex_wg);
  (void)hipFree(c->d_ex_state);
  c->d_ex_id = nullptr;
  c->d_ex_x = c->d_ex_v = nullptr;
  c->d_ex_meta = c->d_ex_wg = nullptr;
  c->d_ex_state = nullptr;
  c->exit_alloc = 0;
}

int collect_exits(amps_ctx* c) {
  if (!c->d_ex_state) return fail(AMPS_E_STATE, "exit list not allocated");
  const int k = c->cur;
  hipLaunchKernelGGL(k_exit_count, dim3(EX_WG), dim3(EX_THREADS), 0, c->stream, c->d_n, c->d_code, c->d_ex_wg);
  hipLaunchKernelGGL(k_exit_offsets, dim3(1), dim3(EX_WG), 0, c->stream, c->d_ex_wg, c->d_ex_state,
                     c->exit_alloc);
  hipLaunchKernelGGL(k_exit_write, dim3(EX_WG), dim3(EX_THREADS), 0, c->stream, c->d_n, c->d_code, c->d_key[k],
                     c->d_ex_wg, c->cap, c->d_x[k], c->d_v[k], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k],
                     c->d_id[k], c->exit_alloc, c->d_ex_id, c->d_ex_x, c->d_ex_v, c->d_ex_meta);
  HIP_TRY(hipGetLastError());
  return 0;
}

// called before a USER-mode push: the list exists (allocated outside any capture)
int prepare_exits(amps_ctx* c) { return ensure_exit_list(c); }

}  // namespace amps

extern "C" int amps_gpu_get_exits(amps_ctx* c, int64_t cap, int64_t* n_exits, int64_t* id, double* x, double* v,
                                  int32_t* spec, int32_t* face, int32_t* blk, int32_t clear) {
  if (!c || !n_exits) return fail(AMPS_E_ARG, "amps_gpu_get_exits: null argument");
  if (cap < 0) return fail(AMPS_E_ARG, "amps_gpu_get_exits: negative capacity");
  (void)hipSetDevice(c->device);
  *n_exits = 0;
  if (!c->d_ex_state) return AMPS_OK;  // no USER-mode push yet
  int64_t st[2] = {0, 0};
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemcpy(st, c->d_ex_state, sizeof(st), hipMemcpyDeviceToHost));
  *n_exits = st[0];
  const int64_t m = std::min(st[0], cap);
  if (m > 0) {
    if (id) HIP_TRY(hipMemcpy(id, c->d_ex_id, sizeof(int64_t) * m, hipMemcpyDeviceToHost));
    if (x) HIP_TRY(hipMemcpy(x, c->d_ex_x, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
    if (v) HIP_TRY(hipMemcpy(v, c->d_ex_v, sizeof(double) * 3 * m, hipMemcpyDeviceToHost));
    if (spec || face || blk) {
      std::vector<int> meta((size_t)(3 * m));
      HIP_TRY(hipMemcpy(meta.data(), c->d_ex_meta, sizeof(int) * 3 * m, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < m; i++) {
        if (spec) spec[i] = meta[(size_t)(3 * i)];
        if (face) face[i] = meta[(size_t)(3 * i + 1)];
        if (blk) blk[i] = meta[(size_t)(3 * i + 2)];
      }
    }
  }
  if (clear) {
    const int64_t zero[2] = {0, 0};
    HIP_TRY(hipMemcpy(c->d_ex_state, zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  if (st[1] > 0)
    return fail(AMPS_E_RANGE, "amps_gpu_get_exits: " + std::to_string(st[1]) +
                                  " exits dropped: the list was full (raise \"exit_capacity\")");
  return AMPS_OK;
}
