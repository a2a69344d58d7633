// restart.hip — the particle part of AMPS restart files and ParticleBuffer
// records <-> the device's structure of arrays.
//
// A reference particle is one ParticleDataLength-byte record whose mandatory
// fields sit at the offsets of picParticleDataMacro.h (next and prev list
// links, species id, velocity, position, optionally the weight correction).
// SaveParticleDataBlock (pic_restart.cpp:312-370) writes, per block with
// particles, the per-cell counts as int ParticleNumberTable[i][j][k] (cell
// i + 8(j + 8k) at position 64i + 8j + k) and then the records cell by cell in
// that same i, j, k order, each cell's list from its last particle to its
// first; ReadParticleDataBlock (:430-512) pushes every record read onto the head of its
// cell's list, which restores the list order.  Here the device order within a
// cell is the list order from the head, so export writes it reversed and
// import reverses it back.  Both directions are one permutation kernel over
// the particles, the positions coming from two prefix sums (device key
// order and restart order) built on the host from the per-cell counts.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "amps_common.h"

using namespace amps;

namespace {

// restart position of cell c within its block: 64 i + 8 j + k
__host__ __device__ inline int restart_rank(int c) { return 64 * (c % NC) + 8 * ((c / NC) % NC) + c / (NC * NC); }
__host__ __device__ inline int cell_of_rank(int r) { return (r / 64) + NC * ((r / 8) % 8 + NC * (r % 8)); }

template <bool ALIGNED>
__device__ inline void put_f64(unsigned char* rec, int off, double v) {
  if constexpr (ALIGNED) {
    *reinterpret_cast<double*>(rec + off) = v;
  } else {
    unsigned char b[8];
    __builtin_memcpy(b, &v, 8);
#pragma unroll
    for (int i = 0; i < 8; i++) rec[off + i] = b[i];
  }
}

template <bool ALIGNED>
__device__ inline double get_f64(const unsigned char* rec, int off) {
  if constexpr (ALIGNED) {
    return *reinterpret_cast<const double*>(rec + off);
  } else {
    unsigned char b[8];
#pragma unroll
    for (int i = 0; i < 8; i++) b[i] = rec[off + i];
    double v;
    __builtin_memcpy(&v, b, 8);
    return v;
  }
}

struct Layout {
  int bytes, spec_off, spec_bytes, v_off, x_off, w_off;
};

// particle p (device order) -> its record in restart order
template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_export(int64_t n, Layout L, const int* __restrict__ key,
                                                const int* __restrict__ offset, const int64_t* __restrict__ rbase,
                                                const double* __restrict__ x, const double* __restrict__ v,
                                                int vf32, const uint8_t* __restrict__ spec, int64_t cap,
                                                unsigned char* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int k = key[p];
  const int b = k / NCELL, c = k % NCELL;
  const int64_t first = offset[k], cnt = offset[k + 1] - first;
  const int64_t dst = rbase[(int64_t)b * NCELL + restart_rank(c)] + (cnt - 1 - (p - first));
  unsigned char* rec = out + dst * L.bytes;
  // the list links are meaningless outside the buffer: -1
  const int64_t none = -1;
  if constexpr (ALIGNED) {
    reinterpret_cast<int64_t*>(rec)[0] = none;
    reinterpret_cast<int64_t*>(rec)[1] = none;
  } else {
    put_f64<false>(rec, 0, __longlong_as_double(none));
    put_f64<false>(rec, 8, __longlong_as_double(none));
  }
  const int s = spec[p];
  if (L.spec_bytes == 1) {
    rec[L.spec_off] = (unsigned char)s;
  } else {
    for (int i = 0; i < L.spec_bytes; i++) rec[L.spec_off + i] = (unsigned char)(i == 0 ? s : 0);
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double vd = vf32 ? (double)((const float*)v)[d * cap + p] : v[d * cap + p];
    put_f64<ALIGNED>(rec, L.v_off + 8 * d, vd);
    put_f64<ALIGNED>(rec, L.x_off + 8 * d, x[d * cap + p]);
  }
  if (L.w_off >= 0) put_f64<ALIGNED>(rec, L.w_off, 1.0);
}

// record e (restart order) -> its device position
template <bool ALIGNED>
__global__ __launch_bounds__(256) void k_import(int64_t n, Layout L, int64_t nseg, const int64_t* __restrict__ rbase,
                                                const int* __restrict__ offset, const unsigned char* __restrict__ in,
                                                double* __restrict__ x, double* __restrict__ v, int vf32,
                                                uint8_t* __restrict__ spec, int* __restrict__ key,
                                                int64_t* __restrict__ id, int64_t cap) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  // the (block, restart rank) segment holding record e: last rbase[s] <= e
  int64_t lo = 0, hi = nseg;  // rbase[0] = 0 <= e < rbase[nseg] = n
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rbase[mid] <= e) lo = mid;
    else hi = mid;
  }
  const int b = (int)(lo / NCELL), r = (int)(lo % NCELL);
  const int k = b * NCELL + cell_of_rank(r);
  const int64_t first = offset[k], cnt = offset[k + 1] - first;
  const int64_t p = first + (cnt - 1 - (e - rbase[lo]));
  const unsigned char* rec = in + e * L.bytes;
  spec[p] = rec[L.spec_off];  // little endian: the low byte carries the id (< 256, checked)
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double vd = get_f64<ALIGNED>(rec, L.v_off + 8 * d);
    if (vf32)
      ((float*)v)[d * cap + p] = (float)vd;
    else
      v[d * cap + p] = vd;
    x[d * cap + p] = get_f64<ALIGNED>(rec, L.x_off + 8 * d);
  }
  key[p] = k;
  id[p] = e;
}

int check_layout(const amps_aos_layout* l, Layout& L, bool& aligned) {
  if (!l) return fail(AMPS_E_ARG, "particle record layout: null");
  L = {l->record_bytes, l->species_offset, l->species_bytes, l->velocity_offset, l->position_offset,
       l->weight_offset};
  auto inside = [&](int off, int len) { return off >= 16 && off + len <= L.bytes; };
  if (L.bytes < 16 + 1 + 48 || L.bytes > 4096) return fail(AMPS_E_ARG, "particle record layout: record_bytes out of range");
  if (L.spec_bytes != 1 && L.spec_bytes != 4 && L.spec_bytes != 8)
    return fail(AMPS_E_ARG, "particle record layout: species_bytes must be 1, 4 or 8");
  if (!inside(L.spec_off, L.spec_bytes) || !inside(L.v_off, 24) || !inside(L.x_off, 24) ||
      (L.w_off >= 0 && !inside(L.w_off, 8)))
    return fail(AMPS_E_ARG, "particle record layout: a field lies outside the record or on the list links");
  aligned = L.bytes % 8 == 0 && L.v_off % 8 == 0 && L.x_off % 8 == 0 && (L.w_off < 0 || L.w_off % 8 == 0);
  return 0;
}

}  // namespace

extern "C" int amps_gpu_get_particles_restart(amps_ctx* c, const amps_aos_layout* layout, int32_t* counts,
                                              void* records, int64_t cap_records, int64_t* n_records) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  Layout L;
  bool aligned = false;
  if (int r = check_layout(layout, L, aligned)) return r;
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_get_particles_restart: no mesh");
  (void)hipSetDevice(c->device);
  if (c->cap > 0 && !c->offsets_valid)
    if (int r = amps_gpu_compact(c)) return r;
  const int64_t n = c->cap > 0 ? amps_gpu_particle_count(c) : 0;
  if (n < 0) return fail(AMPS_E_HIP, "amps_gpu_get_particles_restart: cannot read the particle count");
  if (n_records) *n_records = n;
  const int nb = c->mv.nb;
  const int64_t nkeys = (int64_t)nb * NCELL;
  std::vector<int> off((size_t)nkeys + 1, 0);
  if (n > 0) HIP_TRY(hipMemcpy(off.data(), c->d_offset, sizeof(int) * (nkeys + 1), hipMemcpyDeviceToHost));
  // restart-order counts and prefix (block-major, then 64 i + 8 j + k)
  std::vector<int64_t> rb((size_t)nkeys + 1, 0), rbase_key((size_t)nkeys, 0);
  for (int b = 0; b < nb; b++)
    for (int r = 0; r < NCELL; r++) {
      const int64_t k = (int64_t)b * NCELL + cell_of_rank(r);
      const int cnt = off[(size_t)k + 1] - off[(size_t)k];
      if (counts) counts[(size_t)b * NCELL + r] = cnt;
      rb[(size_t)b * NCELL + r + 1] = rb[(size_t)b * NCELL + r] + cnt;
    }
  if (!records || n == 0) return AMPS_OK;
  if (cap_records < n) return fail(AMPS_E_RANGE, "amps_gpu_get_particles_restart: record buffer too small");
  for (int b = 0; b < nb; b++)
    for (int cc = 0; cc < NCELL; cc++) rbase_key[(size_t)b * NCELL + restart_rank(cc)] = rb[(size_t)b * NCELL + restart_rank(cc)];
  int64_t* d_rb = nullptr;
  unsigned char* d_out = nullptr;
  HIP_TRY(hipMalloc(&d_rb, sizeof(int64_t) * nkeys));
  HIP_TRY(hipMalloc(&d_out, (size_t)n * L.bytes));
  HIP_TRY(hipMemcpyAsync(d_rb, rbase_key.data(), sizeof(int64_t) * nkeys, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemsetAsync(d_out, 0, (size_t)n * L.bytes, c->stream));
  const int k = c->cur;
  const dim3 grid((unsigned)((n + 255) / 256));
  const int vf32 = c->precision == AMPS_FP32 ? 1 : 0;
  if (aligned)
    hipLaunchKernelGGL(k_export<true>, grid, dim3(256), 0, c->stream, n, L, c->d_key[k], c->d_offset, d_rb,
                       c->d_x[k], c->d_v[k], vf32, c->d_spec[k], c->cap, d_out);
  else
    hipLaunchKernelGGL(k_export<false>, grid, dim3(256), 0, c->stream, n, L, c->d_key[k], c->d_offset, d_rb,
                       c->d_x[k], c->d_v[k], vf32, c->d_spec[k], c->cap, d_out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(records, d_out, (size_t)n * L.bytes, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  (void)hipFree(d_rb);
  (void)hipFree(d_out);
  return AMPS_OK;
}

extern "C" int amps_gpu_set_particles_restart(amps_ctx* c, const amps_aos_layout* layout, const int32_t* counts,
                                              int64_t n_records, const void* records, int32_t precision) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  Layout L;
  bool aligned = false;
  if (int r = check_layout(layout, L, aligned)) return r;
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_set_particles_restart: set the mesh first");
  if (precision != AMPS_FP64 && precision != AMPS_FP32)
    return fail(AMPS_E_ARG, "amps_gpu_set_particles_restart: precision must be AMPS_FP64 or AMPS_FP32");
  if (n_records < 0 || n_records >= ((int64_t)1 << 31) - 1)
    return fail(AMPS_E_RANGE, "amps_gpu_set_particles_restart: record count out of [0, 2^31-1)");
  if (!counts || (n_records > 0 && !records)) return fail(AMPS_E_ARG, "amps_gpu_set_particles_restart: null array");
  const int nb = c->mv.nb;
  const int64_t nkeys = (int64_t)nb * NCELL;
  std::vector<int64_t> rb((size_t)nkeys + 1, 0);
  std::vector<int> off((size_t)nkeys + 1, 0), cntk((size_t)nkeys, 0);
  for (int b = 0; b < nb; b++)
    for (int r = 0; r < NCELL; r++) {
      const int cnt = counts[(size_t)b * NCELL + r];
      if (cnt < 0) return fail(AMPS_E_ARG, "amps_gpu_set_particles_restart: negative cell count");
      rb[(size_t)b * NCELL + r + 1] = rb[(size_t)b * NCELL + r] + cnt;
      cntk[(size_t)b * NCELL + cell_of_rank(r)] = cnt;
    }
  if (rb[(size_t)nkeys] != n_records)
    return fail(AMPS_E_ARG, "amps_gpu_set_particles_restart: the cell counts do not add up to the record count");
  for (int64_t k = 0; k < nkeys; k++) off[(size_t)k + 1] = off[(size_t)k] + cntk[(size_t)k];
  const unsigned char* rec = static_cast<const unsigned char*>(records);
  const int nsp = std::max(1, c->sp.n);
  for (int64_t e = 0; e < n_records; e++) {
    int64_t s = 0;
    std::memcpy(&s, rec + e * L.bytes + L.spec_off, (size_t)L.spec_bytes);  // little endian
    if (s < 0 || s >= nsp || s > 255)
      return fail(AMPS_E_ARG, "amps_gpu_set_particles_restart: species id out of range");
  }
  (void)hipSetDevice(c->device);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (int r = alloc_particles(c, n_records)) return r;
  const int64_t cap = c->cap;
  c->cur = 0;
  HIP_TRY(hipMemset(c->d_code, AMPS_CODE_FINISHED, cap));
  HIP_TRY(hipMemcpy(c->d_offset, off.data(), sizeof(int) * (nkeys + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_off_layout, off.data(), sizeof(int) * (nkeys + 1), hipMemcpyHostToDevice));
  c->layout_valid = true;
  if (n_records > 0) {
    int64_t* d_rb = nullptr;
    unsigned char* d_in = nullptr;
    HIP_TRY(hipMalloc(&d_rb, sizeof(int64_t) * (nkeys + 1)));
    HIP_TRY(hipMalloc(&d_in, (size_t)n_records * L.bytes));
    HIP_TRY(hipMemcpy(d_rb, rb.data(), sizeof(int64_t) * (nkeys + 1), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_in, rec, (size_t)n_records * L.bytes, hipMemcpyHostToDevice));
    const dim3 grid((unsigned)((n_records + 255) / 256));
    const int vf32 = precision == AMPS_FP32 ? 1 : 0;
    if (aligned)
      hipLaunchKernelGGL(k_import<true>, grid, dim3(256), 0, c->stream, n_records, L, nkeys, d_rb, c->d_offset,
                         d_in, c->d_x[0], c->d_v[0], vf32, c->d_spec[0], c->d_key[0], c->d_id[0], cap);
    else
      hipLaunchKernelGGL(k_import<false>, grid, dim3(256), 0, c->stream, n_records, L, nkeys, d_rb, c->d_offset,
                         d_in, c->d_x[0], c->d_v[0], vf32, c->d_spec[0], c->d_key[0], c->d_id[0], cap);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(d_rb);
    (void)hipFree(d_in);
  }
  const int nn = (int)n_records;
  HIP_TRY(hipMemcpy(c->d_n, &nn, sizeof(int), hipMemcpyHostToDevice));
  c->n = n_records;
  c->precision = precision;
  c->offsets_valid = true;
  if (int r = build_tiles(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}
