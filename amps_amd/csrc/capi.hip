// capi.hip — C ABI of the particle hot path: context, mesh, fields,
// particles, mover, migration and ECSIM deposit orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "amps_common.h"

using namespace amps;

// kernels (pic_kernels.hip)
template <int MOVER, bool BC, int W>
__global__ void k_push_f64(MeshView, Faces, Species, const double*, const double*, const Tile*,
                           const int*, int, double, int64_t, double*, double*, const uint8_t*,
                           int*, int8_t*, const double*, int*, PushSrc, PushMig);
template <int MOVER, int W>
__global__ void k_push_f32(MeshView, Faces, Species, const double*, const double*, const Tile*,
                           const int*, int, double, int64_t, double*, float*, const uint8_t*, int*,
                           int8_t*, int*, PushSrc, PushMig);
__global__ void k_fill_int(int*, int64_t, int);
__global__ void k_count_keys(const int*, const int*, const int8_t*, int*);
__global__ void k_scatter(const int*, const int*, const int8_t*, const int*, int*, int64_t,
                          const double*, const double*, int, const uint8_t*, const int64_t*,
                          double*, double*, uint8_t*, int*, int64_t*);
__global__ void k_copy_total(const int*, int*);
__global__ void k_mig_classify(const int*, const int*, const int8_t*, const int*, int, const int*, int*, int*, int*,
                               uint64_t*);
__global__ void k_mig_order(int64_t, int, int, const int*, const int*, int*, int*, int*, int, const int*,
                            const int*, const uint64_t*, int*);
__global__ void k_mig_order_slab(int64_t, int, int, const int*, const int*, const int*, int*, int*, int*,
                                 const int*, const int*, const uint64_t*, int*);
__global__ void k_mig_place(const int*, const int2*, const int*, const int*, const int*, int*);
__global__ void k_add_ints(int64_t, int*, const int*);
__global__ void k_mig_order_big(int64_t, int, const int*, const int*, int*, int*, int*, const int*);
__global__ void k_mig_scatter(const int*, const int*, const int8_t*, const int*, const int*, const uint64_t*,
                              const int*, int64_t, const double*, const double*, int, const uint8_t*,
                              const int64_t*, double*, double*, uint8_t*, int*, int64_t*);
__global__ void k_mig_dest(const int*, const int*, const int*, const int*, const uint64_t*, int64_t, int*);
__global__ void k_mig_move(const int*, const int*, int64_t, const double*, const double*, int, const uint8_t*,
                           const int64_t*, const int*, double*, double*, uint8_t*, int64_t*, int*);
__global__ void k_copy_ints(int64_t, const int*, int*);
__global__ void k_tile_count(int, int, const int*, int*);
__global__ void k_tile_emit(int, int, const int*, const int*, Tile*, int*);
template <int NCH, bool EXACT>
__global__ void k_deposit_cell(MeshView, Species, double, const double*, const int*, int64_t,
                               const double*, const double*, int, const uint8_t*, double*);
template <bool MOMENT>
__global__ void k_deposit_node(int64_t, int, const int*, const int*, const double*, double*);
__global__ void k_deposit_node_staged(int64_t, int, const int*, const int*, const double*, double*);
__global__ void k_generate(MeshView, GenArgs, int64_t, double*, double*, uint8_t*, int*, int64_t*, int8_t*);
__global__ void k_copy_probe(const float4*, float4*, int64_t);
namespace amps {
int upload_inv_index_matrix(hipStream_t s);
}

static thread_local std::string g_err;

namespace amps {
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
  return AMPS_E_HIP;
}
void tbegin(amps_ctx* c) {
  if (c->timing) (void)hipEventRecord(c->ev0, c->stream);
}
void tend(amps_ctx* c, int k) {
  if (!c->timing) return;
  (void)hipEventRecord(c->ev1, c->stream);
  (void)hipEventSynchronize(c->ev1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
  c->ms[k] += ms;
  c->launches[k] += 1;
}
}  // namespace amps

extern "C" const char* amps_gpu_last_error(void) { return g_err.c_str(); }

static void free_particles(amps_ctx* c) {
  for (int i = 0; i < 2; i++) {
    (void)hipFree(c->d_x[i]);
    (void)hipFree(c->d_v[i]);
    (void)hipFree(c->d_spec[i]);
    (void)hipFree(c->d_key[i]);
    (void)hipFree(c->d_id[i]);
    c->d_x[i] = c->d_v[i] = nullptr;
    c->d_spec[i] = nullptr;
    c->d_key[i] = nullptr;
    c->d_id[i] = nullptr;
  }
  (void)hipFree(c->d_code);
  c->d_code = nullptr;
  (void)hipFree(c->d_code2);
  c->d_code2 = nullptr;
  (void)hipFree(c->d_dest);
  (void)hipFree(c->d_tmpsrc);
  (void)hipFree(c->d_wcnt);
  (void)hipFree(c->d_wmask);
  (void)hipFree(c->d_mig_ovf);
  c->d_dest = c->d_tmpsrc = c->d_wcnt = nullptr;
  c->d_wmask = nullptr;
  c->d_mig_ovf = nullptr;
  c->layout_valid = false;
  c->cap = 0;
  c->n = 0;
}

static void free_mesh(amps_ctx* c) {
  (void)hipFree(c->d_xmin);
  (void)hipFree(c->d_xmax);
  (void)hipFree(c->d_dxc);
  (void)hipFree(c->d_dxs);
  (void)hipFree(c->d_leaf);
  (void)hipFree(c->d_blk_lat);
  c->d_blk_lat = nullptr;
  (void)hipFree(c->d_level);
  (void)hipFree(c->d_nbinfo);
  (void)hipFree(c->d_Bc);
  c->d_level = c->d_nbinfo = nullptr;
  c->d_Bc = nullptr;
  c->has_center = false;
  (void)hipFree(c->d_Bm);
  c->d_Bm = nullptr;
  c->has_mode3d = false;
  (void)hipFree(c->d_inc8);
  (void)hipFree(c->d_inc8_oct);
  (void)hipFree(c->d_node_order);
  c->d_node_order = nullptr;
  (void)hipFree(c->d_offset);
  (void)hipFree(c->d_off_layout);
  (void)hipFree(c->d_big);
  (void)hipFree(c->d_kbase);
  (void)hipFree(c->d_mig_slab);
  c->d_off_layout = c->d_big = c->d_kbase = c->d_mig_slab = nullptr;
  c->layout_valid = false;
  (void)hipFree(c->d_cursor);
  (void)hipFree(c->d_tcount);
  (void)hipFree(c->d_cellpart);
  (void)hipFree(c->d_nodeout);
  (void)hipFree(c->d_ecsim_nbr);
  ecsim_field_free(c);
  c->d_ecsim_nbr = nullptr;
  (void)hipFree(c->d_E);
  (void)hipFree(c->d_B);
  c->d_xmin = c->d_xmax = c->d_dxc = c->d_dxs = nullptr;
  c->d_leaf = c->d_inc8 = c->d_inc8_oct = nullptr;
  c->d_offset = c->d_cursor = c->d_tcount = nullptr;
  c->d_cellpart = c->d_nodeout = nullptr;
  c->d_E = c->d_B = nullptr;
  c->has_mesh = c->has_fields = false;
}

extern "C" int amps_gpu_create(int device, amps_ctx** out) {
  if (!out) return fail(AMPS_E_ARG, "amps_gpu_create: out is NULL");
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(AMPS_E_ARG, "amps_gpu_create: bad device index");
  HIP_TRY(hipSetDevice(device));
  amps_ctx* c = new amps_ctx();
  c->device = device;
  if (int r = energy_scratch_init(device)) {
    delete c;
    return r;
  }
  HIP_TRY(hipDeviceGetAttribute(&c->n_cu, hipDeviceAttributeMultiprocessorCount, device));
  HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&c->ev0));
  HIP_TRY(hipEventCreate(&c->ev1));
  HIP_TRY(hipMalloc(&c->d_n, sizeof(int) * 4));
  HIP_TRY(hipMalloc(&c->d_ntiles, sizeof(int) * 4));
  if (int r = upload_inv_index_matrix(c->stream)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  *out = c;
  return AMPS_OK;
}

extern "C" int amps_gpu_destroy(amps_ctx* c) {
  if (!c) return AMPS_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_particles(c);
  free_mesh(c);
  trace_free(c);
  mlp_free(c);
  exits_free(c);
  (void)hipFree(c->d_n);
  (void)hipFree(c->d_ntiles);
  (void)hipFree(c->d_tiles);
  (void)hipFree(c->d_scan_tmp);
  if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
  if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return AMPS_OK;
}

extern "C" int amps_gpu_synchronize(amps_ctx* c) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int amps_gpu_device_info(amps_ctx* c, char* name, int* n_cu, int64_t* hbm) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, c->device));
  if (name) {
    std::snprintf(name, 64, "%s", p.gcnArchName);
  }
  if (n_cu) *n_cu = p.multiProcessorCount;
  if (hbm) *hbm = (int64_t)p.totalGlobalMem;
  return AMPS_OK;
}

// chunk counts each deposit cell-pass kernel is instantiated for
static bool deposit_variant_ok(int kernel, int64_t nch) {
  // kernel 2 keeps the chunk count for its exact mode, which runs kernel 1
  if (kernel == 1 || kernel == 2) return nch == 4 || nch == 6 || nch == 12;
  // kernel 3: particles per cell per record batch (its exact mode runs kernel 1 with 4 chunks)
  if (kernel == 3) return nch == 4 || nch == 8;
  return nch == 2 || nch == 3 || nch == 4 || nch == 6;
}

extern "C" int amps_gpu_set_param(amps_ctx* c, const char* name, int64_t value) {
  if (!c || !name) return fail(AMPS_E_ARG, "amps_gpu_set_param: null argument");
  std::string n(name);
  if (n == "tile") {
    if (value < 64 || value > (1 << 20)) return fail(AMPS_E_ARG, "tile out of range [64, 2^20]");
    c->tile = (int)value;
    c->tiles_valid = false;
  } else if (n == "deposit_chunks") {
    if (!deposit_variant_ok(c->dep_kernel, value))
      return fail(AMPS_E_ARG, "deposit_chunks must be 2, 3, 4 or 6 (deposit_kernel 0), 4, 6 or 12 "
                              "(deposit_kernel 1 and 2) or 4 or 8 (deposit_kernel 3)");
    c->dep_nch = (int)value;
  } else if (n == "push_waves") {
    if (value != 1 && value != 4) return fail(AMPS_E_ARG, "push_waves must be 1 or 4");
    c->push_waves = (int)value;
  } else if (n == "deterministic") {
    if (value != 0 && value != 1) return fail(AMPS_E_ARG, "deterministic must be 0 or 1");
    c->deterministic = (int)value;
  } else if (n == "migration_sort_small") {
    if (value < 1 || value > 64) return fail(AMPS_E_ARG, "migration_sort_small must be in [1, 64]");
    c->mig_small = (int)value;
  } else if (n == "deposit_waves") {
    if (value != 2 && value != 3) return fail(AMPS_E_ARG, "deposit_waves must be 2 or 3");
    c->dep_waves = (int)value;
  } else if (n == "deposit_paired") {
    if (value != 0 && value != 1) return fail(AMPS_E_ARG, "deposit_paired must be 0 or 1");
    c->dep_paired = (int)value;
  } else if (n == "node_staged") {
    if (value != 0 && value != 1) return fail(AMPS_E_ARG, "node_staged must be 0 or 1");
    c->node_staged = (int)value;
  } else if (n == "node_batch") {
    if (value < 0 || value > 64) return fail(AMPS_E_ARG, "node_batch must be in [0, 64]");
    c->node_batch = (int)value;
  } else if (n == "exit_capacity") {
    if (value < 1 || value > (1 << 30)) return fail(AMPS_E_ARG, "exit_capacity must be in [1, 2^30]");
    // the list is reallocated (emptied) before the next USER_FUNCTION push
    c->exit_cap = value;
  } else if (n == "deposit_exact") {
    c->dep_exact = value ? 1 : 0;
  } else if (n == "b_center") {
    if (value != 0 && value != 1) return fail(AMPS_E_ARG, "b_center must be 0 or 1");
    c->b_center = (int)value;
  } else if (n == "deposit_kernel") {
    if (value < 0 || value > 3) return fail(AMPS_E_ARG, "deposit_kernel must be 0, 1, 2 or 3");
    if (!deposit_variant_ok((int)value, c->dep_nch))
      return fail(AMPS_E_ARG, "deposit_kernel: set deposit_chunks to a count this kernel runs with first");
    c->dep_kernel = (int)value;
  } else {
    return fail(AMPS_E_ARG, "unknown parameter " + n);
  }
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// mesh
// ---------------------------------------------------------------------
static void build_faces(const double gxmin[3], const double gxmax[3], Faces& F) {
  // ExternalBoundaryFaceTable and its Init, pic_mover.cpp:24-28, 60-74
  static const double nrm[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  static const int nX0[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}};
  static const double e0[6][3] = {{0, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}};
  static const double e1[6][3] = {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 0}, {0, 1, 0}};
  for (int nf = 0; nf < 6; nf++) {
    double cE0 = 0.0, cE1 = 0.0;
    for (int d = 0; d < 3; d++) {
      F.f[nf].norm[d] = nrm[nf][d];
      F.f[nf].e0[d] = e0[nf][d];
      F.f[nf].e1[d] = e1[nf][d];
      F.f[nf].x0[d] = (nX0[nf][d] == 0) ? gxmin[d] : gxmax[d];
      double a = ((e0[nf][d] + nX0[nf][d] < 0.5) ? gxmin[d] : gxmax[d]) - F.f[nf].x0[d];
      double b = ((e1[nf][d] + nX0[nf][d] < 0.5) ? gxmin[d] : gxmax[d]) - F.f[nf].x0[d];
      cE0 += std::pow(a, 2);
      cE1 += std::pow(b, 2);
    }
    F.f[nf].lE0 = std::sqrt(cE0);
    F.f[nf].lE1 = std::sqrt(cE1);
  }
}

extern "C" int amps_gpu_set_mesh(amps_ctx* c, const amps_mesh_desc* desc, const int32_t* level,
                                 const int64_t* ijk) {
  if (!c || !desc || !level || !ijk) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: null argument");
  const int nb = desc->nblocks, lmax = desc->lmax;
  if (nb <= 0) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: nblocks must be > 0");
  if (lmax < 3 || lmax > 20) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: lmax out of [3,20]");
  if ((int64_t)nb * NCELL >= ((int64_t)1 << 30))
    return fail(AMPS_E_RANGE, "amps_gpu_set_mesh: too many blocks for 32-bit cell keys");
  for (int d = 0; d < 3; d++)
    if (!(desc->xmax[d] > desc->xmin[d])) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: empty domain");
  (void)hipSetDevice(c->device);
  HIP_TRY(hipStreamSynchronize(c->stream));
  free_mesh(c);
  free_particles(c);
  int lf = 0;
  for (int b = 0; b < nb; b++) {
    if (level[b] < 0 || level[b] > lmax - 3)
      return fail(AMPS_E_ARG, "amps_gpu_set_mesh: block level out of [0, lmax-3]");
    lf = std::max(lf, (int)level[b]);
  }
  if (lf > 9) return fail(AMPS_E_RANGE, "amps_gpu_set_mesh: finest block level above 9");
  const int nf = 1 << lf;
  MeshView& m = c->mv;
  m.lmax = lmax;
  m.nb = nb;
  m.lf = lf;
  m.nf = nf;
  for (int d = 0; d < 3; d++) {
    m.gxmin[d] = desc->xmin[d];
    m.gxmax[d] = desc->xmax[d];
    m.dx_root[d] = desc->xmax[d] - desc->xmin[d];
    m.dx_maxref[d] = (desc->xmax[d] - desc->xmin[d]) / (1 << lmax);
  }
  // EPS, meshAMRgeneric.h:2340-2352
  m.eps = 0.0001 * m.dx_root[0] / (double)NC / (1 << lmax);
  for (int d = 1; d < 3; d++) {
    double e = 0.0001 * m.dx_root[d] / (double)NC / (1 << lmax);
    if (m.eps > e) m.eps = e;
  }
  build_faces(m.gxmin, m.gxmax, c->faces);
  // leaves must tile the root exactly
  std::vector<int> leaf((size_t)nf * nf * nf, -1);
  std::vector<double> xmin(3 * (size_t)nb), xmax(3 * (size_t)nb), dxc(3 * (size_t)nb),
      dxs(3 * (size_t)nb);
  for (int b = 0; b < nb; b++) {
    const int L = level[b];
    const int64_t span = (int64_t)1 << (lf - L);
    int64_t i0[3];
    for (int d = 0; d < 3; d++) {
      const int64_t q = ijk[3 * (size_t)b + d];
      if (q < 0 || q >= ((int64_t)1 << L)) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: block index out of range");
      i0[d] = q * span;
      const int64_t xg = q << (lmax - L), size = (int64_t)1 << (lmax - L);
      xmin[3 * b + d] = m.gxmin[d] + (double)xg * m.dx_maxref[d];
      xmax[3 * b + d] = m.gxmin[d] + (double)(xg + size) * m.dx_maxref[d];
      dxs[3 * b + d] = (xmax[3 * b + d] - xmin[3 * b + d]) / NC;
      dxc[3 * b + d] = m.dx_root[d] / (1 << L) / (double)NC;
    }
    for (int64_t k = 0; k < span; k++)
      for (int64_t j = 0; j < span; j++)
        for (int64_t i = 0; i < span; i++) {
          int& slot = leaf[((size_t)(i0[2] + k) * nf + (i0[1] + j)) * nf + (i0[0] + i)];
          if (slot >= 0) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: overlapping blocks");
          slot = b;
        }
  }
  for (size_t i = 0; i < leaf.size(); i++)
    if (leaf[i] < 0) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: blocks do not cover the domain");
  // unique corner nodes by lattice key
  const int64_t S = ((int64_t)1 << lmax) + 1;
  const int64_t total = (int64_t)nb * NCORNER;
  std::vector<int64_t> keys((size_t)total);
  for (int b = 0; b < nb; b++) {
    const int L = level[b];
    const int64_t cs = ((int64_t)1 << (lmax - L)) / NC;
    const int64_t g0 = ijk[3 * (size_t)b] << (lmax - L), g1 = ijk[3 * (size_t)b + 1] << (lmax - L),
                  g2 = ijk[3 * (size_t)b + 2] << (lmax - L);
    for (int k = 0; k < NN; k++)
      for (int j = 0; j < NN; j++)
        for (int i = 0; i < NN; i++)
          keys[(size_t)b * NCORNER + i + NN * (j + NN * k)] =
              ((g2 + k * cs) * S + (g1 + j * cs)) * S + (g0 + i * cs);
  }
  std::vector<int64_t> uk(keys);
  std::sort(uk.begin(), uk.end());
  uk.erase(std::unique(uk.begin(), uk.end()), uk.end());
  c->nnodes = (int64_t)uk.size();
  if (c->nnodes >= ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "too many nodes");
  c->h_node_of_corner.resize((size_t)total);
  for (int64_t i = 0; i < total; i++)
    c->h_node_of_corner[(size_t)i] =
        (int32_t)(std::lower_bound(uk.begin(), uk.end(), keys[(size_t)i]) - uk.begin());
  c->h_node_keys = uk;
  // node -> (cell, corner) incidences in ascending global cell order
  static const int CornerOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                      {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<int> inc_start((size_t)c->nnodes + 1, 0);
  for (int b = 0; b < nb; b++)
    for (int lc = 0; lc < NCELL; lc++) {
      const int ci = lc % NC, cj = (lc / NC) % NC, ck = lc / (NC * NC);
      for (int q = 0; q < 8; q++) {
        const int corner = (ci + CornerOff[q][0]) + NN * ((cj + CornerOff[q][1]) + NN * (ck + CornerOff[q][2]));
        inc_start[(size_t)c->h_node_of_corner[(size_t)b * NCORNER + corner] + 1]++;
      }
    }
  for (int64_t i = 0; i < c->nnodes; i++) inc_start[(size_t)i + 1] += inc_start[(size_t)i];
  std::vector<int> inc((size_t)inc_start.back());
  std::vector<int> fillp(inc_start.begin(), inc_start.end() - 1);
  for (int b = 0; b < nb; b++)
    for (int lc = 0; lc < NCELL; lc++) {
      const int ci = lc % NC, cj = (lc / NC) % NC, ck = lc / (NC * NC);
      const int gid = b * NCELL + lc;
      for (int q = 0; q < 8; q++) {
        const int corner = (ci + CornerOff[q][0]) + NN * ((cj + CornerOff[q][1]) + NN * (ck + CornerOff[q][2]));
        const int node = c->h_node_of_corner[(size_t)b * NCORNER + corner];
        inc[(size_t)fillp[(size_t)node]++] = gid * 8 + q;
      }
    }
  // node-gather order: by first incident cell, so consecutive workgroups read
  // the partials of the same and neighbouring blocks while they are cached
  std::vector<int> node_order((size_t)c->nnodes);
  for (int64_t i = 0; i < c->nnodes; i++) node_order[(size_t)i] = (int)i;
  std::stable_sort(node_order.begin(), node_order.end(), [&](int a, int b2) {
    return inc[(size_t)inc_start[(size_t)a]] < inc[(size_t)inc_start[(size_t)b2]];
  });
  HIP_TRY(hipMalloc(&c->d_node_order, sizeof(int) * std::max<size_t>(1, node_order.size())));
  HIP_TRY(hipMemcpy(c->d_node_order, node_order.data(), sizeof(int) * node_order.size(), hipMemcpyHostToDevice));
  // fixed-width incidence rows in processing order: one 32-byte row per node
  // (a node is a corner of at most one leaf cell per octant around it)
  std::vector<int> inc8((size_t)c->nnodes * 8, -1);
  for (int64_t sl = 0; sl < c->nnodes; sl++) {
    const int node = node_order[(size_t)sl];
    const int e0 = inc_start[(size_t)node], e1 = inc_start[(size_t)node + 1];
    if (e1 - e0 > 8) return fail(AMPS_E_ARG, "amps_gpu_set_mesh: a corner node with more than 8 incident cells");
    for (int e = e0; e < e1; e++) inc8[(size_t)sl * 8 + (e - e0)] = inc[(size_t)e];
  }
  HIP_TRY(hipMalloc(&c->d_inc8, sizeof(int) * std::max<size_t>(8, inc8.size())));
  if (!inc8.empty()) HIP_TRY(hipMemcpy(c->d_inc8, inc8.data(), sizeof(int) * inc8.size(), hipMemcpyHostToDevice));
  // the same rows by octant: a cell of which the node is corner q lies on the
  // node's side 1 - (corner offset bits of q); one cell per octant at most.
  // The zero record lies past both record layouts of the cell buffer.
  c->zero_rec = ((int64_t)nb * NCELL * CELL_VALS + CELL_STRIDE_MOMENT - 1) / CELL_STRIDE_MOMENT;
  if (c->zero_rec * 8 + 8 >= ((int64_t)1 << 31)) return fail(AMPS_E_RANGE, "too many cells for the node gather");
  int q_of_oct[8];
  for (int q = 0; q < 8; q++)
    q_of_oct[(1 - CornerOff[q][0]) + 2 * (1 - CornerOff[q][1]) + 4 * (1 - CornerOff[q][2])] = q;
  std::vector<int> inc8o((size_t)c->nnodes * 8);
  for (int64_t sl = 0; sl < c->nnodes; sl++) {
    int* row = &inc8o[(size_t)sl * 8];
    for (int t = 0; t < 8; t++) row[t] = (int)(c->zero_rec * 8) + q_of_oct[t];
    for (int e = 0; e < 8; e++) {
      const int ic = inc8[(size_t)sl * 8 + e];
      if (ic < 0) continue;
      const int q = ic & 7;
      const int t = (1 - CornerOff[q][0]) + 2 * (1 - CornerOff[q][1]) + 4 * (1 - CornerOff[q][2]);
      if ((row[t] >> 3) != c->zero_rec)
        return fail(AMPS_E_ARG, "amps_gpu_set_mesh: two cells in one octant of a corner node");
      row[t] = ic;
    }
  }
  HIP_TRY(hipMalloc(&c->d_inc8_oct, sizeof(int) * std::max<size_t>(8, inc8o.size())));
  if (!inc8o.empty())
    HIP_TRY(hipMemcpy(c->d_inc8_oct, inc8o.data(), sizeof(int) * inc8o.size(), hipMemcpyHostToDevice));
  // device upload
  HIP_TRY(hipMalloc(&c->d_xmin, sizeof(double) * 3 * nb));
  HIP_TRY(hipMalloc(&c->d_xmax, sizeof(double) * 3 * nb));
  HIP_TRY(hipMalloc(&c->d_dxc, sizeof(double) * 3 * nb));
  HIP_TRY(hipMalloc(&c->d_dxs, sizeof(double) * 3 * nb));
  HIP_TRY(hipMalloc(&c->d_leaf, sizeof(int) * leaf.size()));
  HIP_TRY(hipMemcpy(c->d_xmin, xmin.data(), sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_xmax, xmax.data(), sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_dxc, dxc.data(), sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_dxs, dxs.data(), sizeof(double) * 3 * nb, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_leaf, leaf.data(), sizeof(int) * leaf.size(), hipMemcpyHostToDevice));
  {
    // block origin and extent on the level-lmax lattice (find_block_in)
    std::vector<int> lat((size_t)nb * 4);
    for (int b = 0; b < nb; b++) {
      const int sh = m.lmax - level[b];
      for (int d = 0; d < 3; d++) lat[4 * (size_t)b + d] = (int)(ijk[3 * (size_t)b + d] << sh);
      lat[4 * (size_t)b + 3] = 1 << sh;
    }
    HIP_TRY(hipMalloc(&c->d_blk_lat, sizeof(int) * lat.size()));
    HIP_TRY(hipMemcpy(c->d_blk_lat, lat.data(), sizeof(int) * lat.size(), hipMemcpyHostToDevice));
  }
  const int64_t nkeys = (int64_t)nb * NCELL;
  HIP_TRY(hipMalloc(&c->d_offset, sizeof(int) * (nkeys + 1)));
  HIP_TRY(hipMalloc(&c->d_off_layout, sizeof(int) * (nkeys + 1)));
  HIP_TRY(hipMalloc(&c->d_big, sizeof(int) * (nkeys + 1)));
  HIP_TRY(hipMalloc(&c->d_kbase, sizeof(int) * std::max<int64_t>(nkeys, 1)));
  HIP_TRY(hipMalloc(&c->d_cursor, sizeof(int) * nkeys));
  HIP_TRY(hipMalloc(&c->d_mig_slab, sizeof(int) * (nkeys * MIG_SLAB + 1)));
  c->layout_valid = false;
  HIP_TRY(hipMalloc(&c->d_tcount, sizeof(int) * (nb + 1)));
  m.blk_xmin = c->d_xmin;
  m.blk_xmax = c->d_xmax;
  m.blk_dxc = c->d_dxc;
  m.blk_dxs = c->d_dxs;
  m.leaf_of = c->d_leaf;
  m.blk_lat = c->d_blk_lat;
  // neighbours of every block for the cell-centred stencil: the block just
  // across each face, edge and corner (points 1/64 of the block outside it),
  // and the min/max level over sub-face, sub-edge and corner samples
  // (SetNeibRefinmentLevelLimits, meshAMRgeneric.h:1018-1046)
  {
    auto host_find = [&](const double p[3]) -> int {
      int ix[3];
      for (int d = 0; d < 3; d++) {
        const double t = std::floor((p[d] - m.gxmin[d]) / m.dx_maxref[d]);
        if (!(t >= 0.0 && t < (double)((int64_t)1 << lmax))) return -1;
        ix[d] = (int)t;
      }
      const int sh = lmax - lf;
      return leaf[((size_t)(ix[2] >> sh) * nf + (ix[1] >> sh)) * nf + (ix[0] >> sh)];
    };
    static const int EdgeAxis[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
    static const int EdgeSide[12][3] = {{-1, 0, 0}, {-1, 0, 1}, {-1, 1, 1}, {-1, 1, 0},
                                        {0, -1, 0}, {1, -1, 0}, {1, -1, 1}, {0, -1, 1},
                                        {0, 0, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}};
    std::vector<int> nbinfo((size_t)nb * 28);
    for (int b = 0; b < nb; b++) {
      int* info = &nbinfo[(size_t)b * 28];
      double lo[3], hi[3], dl[3];
      for (int d = 0; d < 3; d++) {
        lo[d] = xmin[3 * (size_t)b + d], hi[d] = xmax[3 * (size_t)b + d];
        dl[d] = (hi[d] - lo[d]) / 64.0;
      }
      int mn = -1, mx = -1;
      auto seen = [&](int bb) {
        if (bb < 0) return;
        const int L = level[bb];
        if (mn == -1 || mn > L) mn = L;
        if (mx < L) mx = L;
      };
      for (int f = 0; f < 6; f++) {
        const int a = f / 2, side = f % 2;
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = 0.5 * (lo[d] + hi[d]);
        p[a] = side ? hi[a] + dl[a] : lo[a] - dl[a];
        info[2 + f] = host_find(p);
        const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
        for (int q = 0; q < 4; q++) {
          p[a1] = lo[a1] + (hi[a1] - lo[a1]) * ((q & 1) ? 0.75 : 0.25);
          p[a2] = lo[a2] + (hi[a2] - lo[a2]) * ((q & 2) ? 0.75 : 0.25);
          seen(host_find(p));
        }
      }
      for (int e = 0; e < 12; e++) {
        double p[3];
        const int ax = EdgeAxis[e];
        for (int d = 0; d < 3; d++) {
          if (d == ax) p[d] = 0.5 * (lo[d] + hi[d]);
          else p[d] = EdgeSide[e][d] ? hi[d] + dl[d] : lo[d] - dl[d];
        }
        info[8 + e] = host_find(p);
        for (int q = 0; q < 2; q++) {
          p[ax] = lo[ax] + (hi[ax] - lo[ax]) * (q ? 0.75 : 0.25);
          seen(host_find(p));
        }
      }
      for (int cc = 0; cc < 8; cc++) {
        double p[3];
        for (int d = 0; d < 3; d++) p[d] = ((cc >> d) & 1) ? hi[d] + dl[d] : lo[d] - dl[d];
        info[20 + cc] = host_find(p);
        seen(info[20 + cc]);
      }
      info[0] = mn;
      info[1] = mx;
    }
    HIP_TRY(hipMalloc(&c->d_level, sizeof(int) * nb));
    HIP_TRY(hipMalloc(&c->d_nbinfo, sizeof(int) * nbinfo.size()));
    HIP_TRY(hipMemcpy(c->d_level, level, sizeof(int) * nb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_nbinfo, nbinfo.data(), sizeof(int) * nbinfo.size(), hipMemcpyHostToDevice));
    m.blk_level = c->d_level;
    m.nbinfo = c->d_nbinfo;
  }
  c->h_level.assign(level, level + nb);
  c->h_ijk.assign(ijk, ijk + 3 * (size_t)nb);
  c->h_xmin = xmin;
  c->h_xmax = xmax;
  c->has_mesh = true;
  c->offsets_valid = false;
  c->tiles_valid = false;
  return AMPS_OK;
}

extern "C" int64_t amps_gpu_node_count(amps_ctx* c) { return (c && c->has_mesh) ? c->nnodes : -1; }

extern "C" int amps_gpu_get_node_map(amps_ctx* c, int32_t* node_of_corner, int64_t* keys) {
  if (!c || !c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_get_node_map: no mesh");
  if (node_of_corner)
    std::memcpy(node_of_corner, c->h_node_of_corner.data(), sizeof(int32_t) * c->h_node_of_corner.size());
  if (keys) std::memcpy(keys, c->h_node_keys.data(), sizeof(int64_t) * c->h_node_keys.size());
  return AMPS_OK;
}

extern "C" int amps_gpu_set_fields(amps_ctx* c, const double* E, const double* B) {
  if (!c || !E || !B) return fail(AMPS_E_ARG, "amps_gpu_set_fields: null argument");
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_set_fields: set the mesh first");
  const size_t nbytes = sizeof(double) * (size_t)c->mv.nb * NCORNER * 3;
  if (!c->d_E) HIP_TRY(hipMalloc(&c->d_E, nbytes));
  if (!c->d_B) HIP_TRY(hipMalloc(&c->d_B, nbytes));
  HIP_TRY(hipMemcpyAsync(c->d_E, E, nbytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->d_B, B, nbytes, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->has_fields = true;
  return AMPS_OK;
}

extern "C" int amps_gpu_set_species(amps_ctx* c, int32_t nspec, const double* charge,
                                    const double* mass, const double* dt, const double* weight) {
  if (!c || !charge || !mass || !dt || !weight) return fail(AMPS_E_ARG, "amps_gpu_set_species: null argument");
  if (nspec <= 0 || nspec > MAX_SPEC) return fail(AMPS_E_ARG, "amps_gpu_set_species: nspec out of [1,8]");
  c->sp.n = nspec;
  for (int s = 0; s < nspec; s++) {
    if (!(mass[s] > 0.0)) return fail(AMPS_E_ARG, "amps_gpu_set_species: mass must be > 0");
    c->sp.charge[s] = charge[s];
    c->sp.mass[s] = mass[s];
    c->sp.dt[s] = dt[s];
    c->sp.weight[s] = weight[s];
  }
  c->has_species = true;
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// particles
// ---------------------------------------------------------------------
int amps::alloc_particles(amps_ctx* c, int64_t n) {
  if (n <= c->cap) return 0;
  free_particles(c);
  const int64_t cap = std::max<int64_t>(n, 64);
  for (int i = 0; i < 2; i++) {
    HIP_TRY(hipMalloc(&c->d_x[i], sizeof(double) * 3 * cap));
    HIP_TRY(hipMalloc(&c->d_v[i], sizeof(double) * 3 * cap));
    HIP_TRY(hipMalloc(&c->d_spec[i], cap));
    HIP_TRY(hipMalloc(&c->d_key[i], sizeof(int) * cap));
    HIP_TRY(hipMalloc(&c->d_id[i], sizeof(int64_t) * cap));
  }
  HIP_TRY(hipMalloc(&c->d_code, (cap + 4) & ~(int64_t)3));
  HIP_TRY(hipMalloc(&c->d_code2, (cap + 4) & ~(int64_t)3));
  HIP_TRY(hipMalloc(&c->d_dest, sizeof(int) * cap));
  HIP_TRY(hipMalloc(&c->d_tmpsrc, sizeof(int) * cap));
  {
    const int64_t nw = (cap + 255) / 256 * 4;
    HIP_TRY(hipMalloc(&c->d_wcnt, sizeof(int) * (nw + 1)));
    HIP_TRY(hipMalloc(&c->d_wmask, sizeof(uint64_t) * nw));
  }
  HIP_TRY(hipMalloc(&c->d_mig_ovf, sizeof(int2) * cap));
  c->cap = cap;
  c->layout_valid = false;
  return 0;
}

int amps::build_tiles(amps_ctx* c) {
  const int nb = c->mv.nb;
  const int64_t need = nb + c->cap / c->tile + 2;
  if (need > c->max_tiles) {
    (void)hipFree(c->d_tiles);
    c->d_tiles = nullptr;
    HIP_TRY(hipMalloc(&c->d_tiles, sizeof(Tile) * need));
    c->max_tiles = need;
  }
  hipLaunchKernelGGL(k_tile_count, dim3((nb + 256) / 256), dim3(256), 0, c->stream, nb, c->tile,
                     c->d_offset, c->d_tcount);
  if (int r = scan_exclusive(c, c->d_tcount, nb + 1, nullptr)) return r;
  hipLaunchKernelGGL(k_tile_emit, dim3((nb + 256) / 256), dim3(256), 0, c->stream, nb, c->tile,
                     c->d_offset, c->d_tcount, c->d_tiles, c->d_ntiles);
  c->tiles_valid = true;
  return AMPS_OK;
}

extern "C" int amps_gpu_set_particles(amps_ctx* c, int64_t n, const double* x, const double* v,
                                      const int32_t* spec, const int32_t* blk,
                                      const int32_t* cell, const int64_t* id, int32_t precision) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_set_particles: set the mesh first");
  if (n < 0 || n >= ((int64_t)1 << 31) - 1) return fail(AMPS_E_RANGE, "particle count out of [0, 2^31-1)");
  if (n > 0 && (!x || !v || !spec || !blk || !cell)) return fail(AMPS_E_ARG, "amps_gpu_set_particles: null array");
  if (precision != AMPS_FP64 && precision != AMPS_FP32)
    return fail(AMPS_E_ARG, "amps_gpu_set_particles: precision must be AMPS_FP64 or AMPS_FP32");
  (void)hipSetDevice(c->device);
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int nb = c->mv.nb;
  const int64_t nkeys = (int64_t)nb * NCELL;
  std::vector<int> cnt((size_t)nkeys + 1, 0);
  for (int64_t p = 0; p < n; p++) {
    if (blk[p] < 0 || blk[p] >= nb || cell[p] < 0 || cell[p] >= NCELL)
      return fail(AMPS_E_ARG, "amps_gpu_set_particles: block or cell id out of range");
    if (spec[p] < 0 || spec[p] >= std::max(1, c->sp.n) || spec[p] > 255)
      return fail(AMPS_E_ARG, "amps_gpu_set_particles: species id out of range");
    cnt[(size_t)blk[p] * NCELL + cell[p] + 1]++;
  }
  for (int64_t k = 0; k < nkeys; k++) cnt[(size_t)k + 1] += cnt[(size_t)k];
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  std::vector<int64_t> perm((size_t)n);
  for (int64_t p = 0; p < n; p++) perm[(size_t)fill[(size_t)blk[p] * NCELL + cell[p]]++] = p;
  if (int r = alloc_particles(c, n)) return r;
  const int64_t cap = c->cap;
  std::vector<double> hx(3 * (size_t)cap, 0.0), hv(3 * (size_t)cap, 0.0);
  std::vector<uint8_t> hs((size_t)cap, 0);
  std::vector<int> hk((size_t)cap, 0);
  std::vector<int64_t> hid((size_t)cap, 0);
  for (int64_t q = 0; q < n; q++) {
    const int64_t p = perm[(size_t)q];
    for (int d = 0; d < 3; d++) {
      hx[(size_t)(d * cap + q)] = x[d * n + p];
      hv[(size_t)(d * cap + q)] = v[d * n + p];
    }
    hs[(size_t)q] = (uint8_t)spec[p];
    hk[(size_t)q] = blk[p] * NCELL + cell[p];
    hid[(size_t)q] = id ? id[p] : p;
  }
  c->cur = 0;
  HIP_TRY(hipMemcpy(c->d_x[0], hx.data(), sizeof(double) * 3 * cap, hipMemcpyHostToDevice));
  if (precision == AMPS_FP32) {
    std::vector<float> hf(3 * (size_t)cap);
    for (size_t i = 0; i < hf.size(); i++) hf[i] = (float)hv[i];
    HIP_TRY(hipMemcpy(c->d_v[0], hf.data(), sizeof(float) * 3 * cap, hipMemcpyHostToDevice));
  } else {
    HIP_TRY(hipMemcpy(c->d_v[0], hv.data(), sizeof(double) * 3 * cap, hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(c->d_spec[0], hs.data(), cap, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_key[0], hk.data(), sizeof(int) * cap, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_id[0], hid.data(), sizeof(int64_t) * cap, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(c->d_code, AMPS_CODE_FINISHED, cap));
  HIP_TRY(hipMemcpy(c->d_offset, cnt.data(), sizeof(int) * (nkeys + 1), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_off_layout, cnt.data(), sizeof(int) * (nkeys + 1), hipMemcpyHostToDevice));
  c->layout_valid = true;
  const int nn = (int)n;
  HIP_TRY(hipMemcpy(c->d_n, &nn, sizeof(int), hipMemcpyHostToDevice));
  c->n = n;
  c->precision = precision;
  c->offsets_valid = true;
  if (int r = build_tiles(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

extern "C" int64_t amps_gpu_particle_count(amps_ctx* c) {
  if (!c) return -1;
  int nn = 0;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return -1;
  if (hipMemcpy(&nn, c->d_n, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  c->n = nn;
  return nn;
}

extern "C" int amps_gpu_get_particles(amps_ctx* c, double* x, double* v, int32_t* spec,
                                      int32_t* blk, int32_t* cell, int64_t* id, int32_t* code) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  const int64_t n = amps_gpu_particle_count(c);
  if (n < 0) return fail(AMPS_E_HIP, "amps_gpu_get_particles: cannot read the particle count");
  if (n == 0) return AMPS_OK;
  const int64_t cap = c->cap;
  const int k = c->cur;
  if (x)
    for (int d = 0; d < 3; d++)
      HIP_TRY(hipMemcpy(x + d * n, c->d_x[k] + d * cap, sizeof(double) * n, hipMemcpyDeviceToHost));
  if (v && c->precision == AMPS_FP32) {
    std::vector<float> f((size_t)n);
    for (int d = 0; d < 3; d++) {
      HIP_TRY(hipMemcpy(f.data(), (const float*)c->d_v[k] + d * cap, sizeof(float) * n, hipMemcpyDeviceToHost));
      for (int64_t i = 0; i < n; i++) v[d * n + i] = f[(size_t)i];
    }
  } else if (v) {
    for (int d = 0; d < 3; d++)
      HIP_TRY(hipMemcpy(v + d * n, c->d_v[k] + d * cap, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  if (spec) {
    std::vector<uint8_t> s((size_t)n);
    HIP_TRY(hipMemcpy(s.data(), c->d_spec[k], n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) spec[i] = s[(size_t)i];
  }
  if (blk || cell) {
    std::vector<int> kk((size_t)n);
    HIP_TRY(hipMemcpy(kk.data(), c->d_key[k], sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      if (blk) blk[i] = kk[(size_t)i] / NCELL;
      if (cell) cell[i] = kk[(size_t)i] % NCELL;
    }
  }
  if (id) HIP_TRY(hipMemcpy(id, c->d_id[k], sizeof(int64_t) * n, hipMemcpyDeviceToHost));
  if (code) {
    std::vector<int8_t> s((size_t)n);
    HIP_TRY(hipMemcpy(s.data(), c->d_code, n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) code[i] = s[(size_t)i];
  }
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// migration
// ---------------------------------------------------------------------
// zeroed per-key counters (d_offset) and scatter cursors
static void clear_counts(amps_ctx* c) {
  const int64_t nkeys = (int64_t)c->mv.nb * NCELL;
  hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, c->stream, c->d_offset, nkeys + 1, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, c->stream, c->d_cursor, nkeys, 0);
}

// counted: the push has already counted the survivors into a cleared d_offset
// defer: leave the records where they are and write only each destination's
// source index (into d_tmpsrc, free once the movers are ordered) and key
// (pending_gather): the step's matrix-core cell pass reads the particles
// through them and the next push moves them as it updates them, so a step
// writes the particle state once instead of twice.
// classified: the push also classified the survivors (PushMig): d_offset
// holds the stayers per key, d_cursor the movers, d_wcnt and d_wmask the
// stayer counts and masks, the slab and overflow list the movers
static int compact_impl(amps_ctx* c, bool counted = false, bool defer = false, bool classified = false) {
  const int64_t nkeys = (int64_t)c->mv.nb * NCELL;
  const int64_t cap = c->cap;
  const unsigned gridp = (unsigned)((cap + 255) / 256);
  const int a = c->cur, b = 1 - c->cur;
  if (!counted) {
    clear_counts(c);
    hipLaunchKernelGGL(k_count_keys, dim3(gridp), dim3(256), 0, c->stream, c->d_n, c->d_key[a],
                       c->d_code, c->d_offset);
  }
  if (classified)
    hipLaunchKernelGGL(k_add_ints, dim3(1024), dim3(256), 0, c->stream, nkeys, c->d_offset, c->d_cursor);
  if (int r = scan_exclusive(c, c->d_offset, nkeys + 1, nullptr)) return r;
  const int vf32 = c->precision == AMPS_FP32 ? 1 : 0;
  if (c->deterministic) {
    // in every key, the particles that stayed in array order, then those that
    // moved in, in array order (every particle a mover when the arrays have no
    // layout of a previous compaction yet, e.g. a generated population)
    const unsigned gridk = (unsigned)((nkeys + 255) / 256);
    const int64_t nw = (int64_t)gridp * 4;
    if (!classified)
      hipLaunchKernelGGL(k_mig_classify, dim3(gridp), dim3(256), 0, c->stream, c->d_n, c->d_key[a], c->d_code,
                         c->d_off_layout, c->layout_valid ? 1 : 0, c->d_offset, c->d_cursor, c->d_tmpsrc,
                         c->d_wcnt, c->d_wmask);
    if (int r = scan_exclusive(c, c->d_wcnt, nw + 1, nullptr)) return r;
    if (classified)
      hipLaunchKernelGGL(k_mig_place, dim3(1024), dim3(256), 0, c->stream, c->d_mig_slab + nkeys * MIG_SLAB,
                         c->d_mig_ovf, c->d_key[a], c->d_offset, c->d_cursor, c->d_tmpsrc);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->d_big + nkeys, 1, 0);
    // the movers' destinations in dest[] only for the scatter; in a deferred
    // move their source indices are the sorted lists, in place in tmpsrc
    const int write_dest = defer ? 0 : 1;
    if (classified)
      hipLaunchKernelGGL(k_mig_order_slab, dim3(gridk), dim3(256), 0, c->stream, nkeys, c->mig_small, write_dest,
                         c->d_mig_slab, c->d_offset, c->d_cursor, c->d_tmpsrc, c->d_dest, c->d_big, c->d_off_layout,
                         c->d_wcnt, c->d_wmask, c->d_kbase);
    else
      hipLaunchKernelGGL(k_mig_order, dim3(gridk), dim3(256), 0, c->stream, nkeys, c->mig_small, write_dest,
                         c->d_offset, c->d_cursor, c->d_tmpsrc, c->d_dest, c->d_big, c->layout_valid ? 1 : 0,
                         c->d_off_layout, c->d_wcnt, c->d_wmask, c->d_kbase);
    // spare array of the merge sort: the other buffer's keys, written later
    hipLaunchKernelGGL(k_mig_order_big, dim3(512), dim3(256), 0, c->stream, nkeys, write_dest, c->d_offset,
                       c->d_cursor, c->d_tmpsrc, c->d_key[b], c->d_dest, c->d_big);
    if (defer) {
      hipLaunchKernelGGL(k_mig_dest, dim3((unsigned)((cap + 1023) / 1024)), dim3(256), 0, c->stream, c->d_n,
                         c->d_key[a], c->d_kbase, c->d_wcnt, c->d_wmask, cap, c->d_tmpsrc);
      c->pending_gather = true;
    } else {
      hipLaunchKernelGGL(k_mig_scatter, dim3(gridp), dim3(256), 0, c->stream, c->d_n, c->d_key[a], c->d_code,
                         c->d_kbase, c->d_wcnt, c->d_wmask, c->d_dest, cap, c->d_x[a], c->d_v[a],
                         vf32, c->d_spec[a], c->d_id[a], c->d_x[b], c->d_v[b], c->d_spec[b], c->d_key[b],
                         c->d_id[b]);
    }
  } else {
    hipLaunchKernelGGL(k_scatter, dim3(gridp), dim3(256), 0, c->stream, c->d_n, c->d_key[a], c->d_code, c->d_offset,
                       c->d_cursor, cap, c->d_x[a], c->d_v[a], vf32, c->d_spec[a], c->d_id[a], c->d_x[b], c->d_v[b],
                       c->d_spec[b], c->d_key[b], c->d_id[b]);
  }
  // the new layout, for the next compaction
  hipLaunchKernelGGL(k_copy_ints, dim3((unsigned)((nkeys + 1 + 255) / 256)), dim3(256), 0, c->stream, nkeys + 1,
                     c->d_offset, c->d_off_layout);
  c->layout_valid = true;
  hipLaunchKernelGGL(k_copy_total, dim3(1), dim3(1), 0, c->stream, c->d_offset + nkeys, c->d_n);
  c->cur = b;
  c->offsets_valid = true;
  // all survivors carry a committed code
  std::swap(c->d_code, c->d_code2);
  hipLaunchKernelGGL(k_fill_int, dim3(256), dim3(256), 0, c->stream, (int*)c->d_code, (cap + 3) / 4,
                     0x03030303);
  return build_tiles(c);
}

extern "C" int amps_gpu_compact(amps_ctx* c) {
  if (!c || !c->has_mesh || c->cap == 0) return fail(AMPS_E_STATE, "amps_gpu_compact: no particles");
  (void)hipSetDevice(c->device);
  tbegin(c);
  int r = compact_impl(c);
  tend(c, AMPS_K_COMPACT);
  return r;
}

// ---------------------------------------------------------------------
// mover
// ---------------------------------------------------------------------
static int check_mover(const amps_ctx* c, const amps_push_opts* o, const char* who) {
  if (o->mover != AMPS_MOVER_LAPENTA && o->mover != AMPS_MOVER_BORIS && o->mover != AMPS_MOVER_RELATIVISTIC)
    return fail(AMPS_E_ARG, std::string(who) + ": unknown mover");
  if (o->mover == AMPS_MOVER_RELATIVISTIC && c->precision == AMPS_FP32)
    return fail(AMPS_E_ARG, std::string(who) + ": the relativistic mover runs in FP64 only");
  if (c->b_center) {
    if (!c->has_center) return fail(AMPS_E_STATE, std::string(who) + ": b_center set but no cell-centred B given");
    if (c->precision == AMPS_FP32 || o->mover == AMPS_MOVER_RELATIVISTIC)
      return fail(AMPS_E_ARG, std::string(who) + ": cell-centred B runs with the FP64 Lapenta and Boris movers");
    if (c->dep_kernel == 0)
      return fail(AMPS_E_ARG, std::string(who) + ": cell-centred B runs with deposit_kernel 1, 2 or 3");
  }
  return 0;
}

// count: when non-null, the push also counts the surviving particles per key
// (the migration's counting pass), into a zeroed array; classify: it also
// classifies them for the deterministic migration (the stayers alone in
// count, see PushMig), into zeroed d_cursor, d_wcnt, d_wmask and overflow count
static int launch_push(amps_ctx* c, const amps_push_opts* o, int* count = nullptr, bool classify = false) {
  const int k = c->cur;
  PushMig pm = {0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (classify) {
    const int64_t nkeys = (int64_t)c->mv.nb * NCELL;
    pm = {nkeys, c->d_off_layout, c->d_cursor, c->d_mig_slab, c->d_mig_ovf, c->d_mig_slab + nkeys * MIG_SLAB,
          c->d_wmask, c->d_wcnt};
  }
  const bool w4 = c->push_waves == 4;
  // records the last compaction left in the other buffer: the push reads them
  // there and writes them to their places
  PushSrc ps = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  if (c->pending_gather) {
    const int r = 1 - k;
    ps = {c->d_tmpsrc, c->d_x[r], c->d_v[r], c->d_spec[r], c->d_id[r], c->d_spec[k], c->d_id[k]};
    c->pending_gather = false;
  }
#define LAUNCH_PUSH32(MV, W)                                                                              \
  hipLaunchKernelGGL((k_push_f32<MV, W>), dim3((unsigned)c->max_tiles), dim3(256), 0, c->stream, c->mv,     \
                     c->faces, c->sp, c->d_E, c->d_B, c->d_tiles, c->d_ntiles, o->bc_mode, o->r_sphere, c->cap, \
                     c->d_x[k], (float*)c->d_v[k], c->d_spec[k], c->d_key[k], c->d_code, count, ps, pm)
  if (c->precision == AMPS_FP32) {
    if (o->mover == AMPS_MOVER_LAPENTA) {
      if (w4) LAUNCH_PUSH32(AMPS_MOVER_LAPENTA, 4);
      else LAUNCH_PUSH32(AMPS_MOVER_LAPENTA, 1);
    } else {
      if (w4) LAUNCH_PUSH32(AMPS_MOVER_BORIS, 4);
      else LAUNCH_PUSH32(AMPS_MOVER_BORIS, 1);
    }
    HIP_TRY(hipGetLastError());
    return 0;
  }
#undef LAUNCH_PUSH32
#define LAUNCH_PUSH64(MV, BCV, W)                                                                       \
  hipLaunchKernelGGL((k_push_f64<MV, BCV, W>), dim3((unsigned)c->max_tiles), dim3(256), 0, c->stream, c->mv,  \
                     c->faces, c->sp, c->d_E, c->d_B, c->d_tiles, c->d_ntiles, o->bc_mode, o->r_sphere, c->cap, \
                     c->d_x[k], c->d_v[k], c->d_spec[k], c->d_key[k], c->d_code, c->d_Bc, count, ps, pm)
  const bool bc = c->b_center != 0;
  if (o->mover == AMPS_MOVER_LAPENTA) {
    if (bc) LAUNCH_PUSH64(AMPS_MOVER_LAPENTA, true, 1);
    else if (w4) LAUNCH_PUSH64(AMPS_MOVER_LAPENTA, false, 4);
    else LAUNCH_PUSH64(AMPS_MOVER_LAPENTA, false, 1);
  } else if (o->mover == AMPS_MOVER_RELATIVISTIC) {
    LAUNCH_PUSH64(AMPS_MOVER_RELATIVISTIC, false, 1);
  } else {
    if (bc) LAUNCH_PUSH64(AMPS_MOVER_BORIS, true, 1);
    else if (w4) LAUNCH_PUSH64(AMPS_MOVER_BORIS, false, 4);
    else LAUNCH_PUSH64(AMPS_MOVER_BORIS, false, 1);
  }
#undef LAUNCH_PUSH64
  HIP_TRY(hipGetLastError());
  return 0;
}

extern "C" int amps_gpu_move_particles(amps_ctx* c, const amps_push_opts* o) {
  if (!c || !o) return fail(AMPS_E_ARG, "amps_gpu_move_particles: null argument");
  if (!c->has_mesh || !c->has_fields || !c->has_species)
    return fail(AMPS_E_STATE, "amps_gpu_move_particles: mesh, fields and species are required");
  if (int r = check_mover(c, o, "amps_gpu_move_particles")) return r;
  if (o->bc_mode != AMPS_BC_DELETE && o->bc_mode != AMPS_BC_REFLECT && o->bc_mode != AMPS_BC_USER)
    return fail(AMPS_E_ARG, "amps_gpu_move_particles: unknown boundary mode");
  if (c->cap == 0) return AMPS_OK;
  (void)hipSetDevice(c->device);
  if (o->bc_mode == AMPS_BC_USER)
    if (int r = prepare_exits(c)) return r;
  if (o->tile > 0 && o->tile != c->tile) {
    if (int r = amps_gpu_set_param(c, "tile", o->tile)) return r;
  }
  if (!c->offsets_valid) {
    if (int r = amps_gpu_compact(c)) return r;
  } else if (!c->tiles_valid) {
    if (int r = build_tiles(c)) return r;
  }
  tbegin(c);
  if (int r = launch_push(c, o)) return r;
  tend(c, AMPS_K_PUSH);
  c->offsets_valid = false;
  c->tiles_valid = false;
  if (o->bc_mode == AMPS_BC_USER)
    if (int r = collect_exits(c)) return r;
  if (o->compact) return amps_gpu_compact(c);
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// ECSIM deposit
// ---------------------------------------------------------------------
template <int NCH, bool EX>
static void launch_dep(amps_ctx* c, double dt) {
  const int k = c->cur;
  hipLaunchKernelGGL((k_deposit_cell<NCH, EX>), dim3((unsigned)c->mv.nb * 8), dim3(64 * NCH), 0,
                     c->stream, c->mv, c->sp, dt, c->d_B, c->d_offset, c->cap, c->d_x[k],
                     c->d_v[k], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[k], c->d_cellpart);
}

static int launch_deposit(amps_ctx* c, double dt) {
  tbegin(c);
  const bool ex = c->dep_exact != 0;
  if (c->dep_kernel == 2 && !ex) {
    if (int r = launch_deposit_cell_mfma(c, dt)) return r;
  } else if (c->dep_kernel == 3 && !ex) {
    if (int r = launch_deposit_cell_lanes(c, dt)) return r;
  } else if (c->dep_kernel >= 1) {
    // kernels 2 and 3 have no bit-exact form: their exact mode is kernel 1's
    if (int r = launch_deposit_cell_rec(c, dt, c->dep_kernel == 3 ? 4 : c->dep_nch)) return r;
  } else {
    switch (c->dep_nch) {
      case 2: ex ? launch_dep<2, true>(c, dt) : launch_dep<2, false>(c, dt); break;
      case 3: ex ? launch_dep<3, true>(c, dt) : launch_dep<3, false>(c, dt); break;
      case 4: ex ? launch_dep<4, true>(c, dt) : launch_dep<4, false>(c, dt); break;
      case 6: ex ? launch_dep<6, true>(c, dt) : launch_dep<6, false>(c, dt); break;
      default: return fail(AMPS_E_ARG, "deposit_kernel 0 runs with deposit_chunks 2, 3, 4 or 6");
    }
  }
  HIP_TRY(hipGetLastError());
  tend(c, AMPS_K_DEPOSIT_CELL);
  tbegin(c);
  // a multiple of 8 workgroups (see xcd_tile); the surplus ones exit
  // node_batch 0 (staged gather only): a resident grid striding over the nodes
  const bool stride_grid = c->node_batch == 0 && c->dep_kernel >= 1 && !c->dep_exact && c->node_staged;
  const int nb_eff = (c->node_batch == 0 && !stride_grid) ? 8 : c->node_batch;
  const int64_t per_wg = 4 * (int64_t)(nb_eff > 0 ? nb_eff : 1);
  int64_t node_wgs = ((c->nnodes + per_wg - 1) / per_wg + 7) / 8 * 8;
  if (stride_grid) {
    // every workgroup resident at once, so that each XCD's band moves as one
    int per_cu = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k_deposit_node_staged),
                                                         256, 0));
    const int64_t resident = std::max<int64_t>(8, (int64_t)c->n_cu * std::max(per_cu, 1) / 8 * 8);
    node_wgs = std::min<int64_t>(node_wgs, resident);
  }
  // the record-batched and matrix-core cell passes write the moment form unless bit-exact
  if (c->dep_kernel >= 1 && !c->dep_exact && c->node_staged)
    hipLaunchKernelGGL(k_deposit_node_staged, dim3((unsigned)node_wgs), dim3(256), 0, c->stream,
                       c->nnodes, nb_eff, c->d_node_order, c->d_inc8_oct, c->d_cellpart, c->d_nodeout);
  else if (c->dep_kernel >= 1 && !c->dep_exact)
    hipLaunchKernelGGL(k_deposit_node<true>, dim3((unsigned)node_wgs), dim3(256), 0, c->stream,
                       c->nnodes, nb_eff, c->d_node_order, c->d_inc8_oct, c->d_cellpart, c->d_nodeout);
  else
    hipLaunchKernelGGL(k_deposit_node<false>, dim3((unsigned)node_wgs), dim3(256), 0, c->stream,
                       c->nnodes, nb_eff, c->d_node_order, c->d_inc8, c->d_cellpart, c->d_nodeout);
  HIP_TRY(hipGetLastError());
  tend(c, AMPS_K_DEPOSIT_NODE);
  return 0;
}

static int ensure_deposit_buffers(amps_ctx* c) {
  if (!c->d_cellpart) {
    // the cell records (either layout), then the zero moment record
    const size_t n = std::max<size_t>((size_t)c->mv.nb * NCELL * CELL_VALS,
                                      (size_t)(c->zero_rec + 1) * CELL_STRIDE_MOMENT);
    HIP_TRY(hipMalloc(&c->d_cellpart, sizeof(double) * n));
    HIP_TRY(hipMemset(c->d_cellpart + (size_t)c->zero_rec * CELL_STRIDE_MOMENT, 0,
                      sizeof(double) * CELL_STRIDE_MOMENT));
  }
  if (!c->d_nodeout) HIP_TRY(hipMalloc(&c->d_nodeout, sizeof(double) * (size_t)c->nnodes * NODE_STRIDE));
  return 0;
}

extern "C" int amps_gpu_update_j_mass_matrix(amps_ctx* c, double dt, double* J_out, double* M_out) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  if (!c->has_mesh || !c->has_fields || !c->has_species)
    return fail(AMPS_E_STATE, "amps_gpu_update_j_mass_matrix: mesh, fields and species are required");
  if (c->b_center && (!c->has_center || c->dep_kernel == 0))
    return fail(AMPS_E_STATE, "amps_gpu_update_j_mass_matrix: b_center needs the cell-centred B and deposit_kernel 1 or 2");
  (void)hipSetDevice(c->device);
  if (int r = ensure_deposit_buffers(c)) return r;
  if (c->cap == 0) {
    HIP_TRY(hipMemsetAsync(c->d_nodeout, 0, sizeof(double) * (size_t)c->nnodes * NODE_STRIDE, c->stream));
  } else {
    if (!c->offsets_valid)
      if (int r = amps_gpu_compact(c)) return r;
    if (int r = launch_deposit(c, dt)) return r;
  }
  // strided copies of the J and M columns of the resident [node][NODE_STRIDE] rows
  if (J_out || M_out) HIP_TRY(hipStreamSynchronize(c->stream));
  if (J_out && c->nnodes > 0)
    HIP_TRY(hipMemcpy2D(J_out, 3 * sizeof(double), c->d_nodeout, NODE_STRIDE * sizeof(double), 3 * sizeof(double),
                        (size_t)c->nnodes, hipMemcpyDeviceToHost));
  if (M_out && c->nnodes > 0)
    HIP_TRY(hipMemcpy2D(M_out, 243 * sizeof(double), c->d_nodeout + 3, NODE_STRIDE * sizeof(double),
                        243 * sizeof(double), (size_t)c->nnodes, hipMemcpyDeviceToHost));
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// timing
// ---------------------------------------------------------------------
extern "C" int amps_gpu_set_timing(amps_ctx* c, int32_t enable) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  c->timing = enable != 0;
  return AMPS_OK;
}
extern "C" int amps_gpu_get_timing(amps_ctx* c, double* ms, int64_t* launches) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  for (int k = 0; k < AMPS_K_COUNT; k++) {
    if (ms) ms[k] = c->ms[k];
    if (launches) launches[k] = c->launches[k];
  }
  return AMPS_OK;
}
extern "C" int amps_gpu_reset_timing(amps_ctx* c) {
  if (!c) return fail(AMPS_E_ARG, "null context");
  for (int k = 0; k < AMPS_K_COUNT; k++) c->ms[k] = 0, c->launches[k] = 0;
  return AMPS_OK;
}

// ---------------------------------------------------------------------
// synthetic population, full steps, copy probe
// ---------------------------------------------------------------------
extern "C" int amps_gpu_generate_particles(amps_ctx* c, int64_t n_per_species, int32_t nspec,
                                           const double* vth, const double* drift, uint64_t seed,
                                           int32_t precision) {
  if (precision != AMPS_FP64 && precision != AMPS_FP32)
    return fail(AMPS_E_ARG, "amps_gpu_generate_particles: bad precision");
  if (!c || !vth || !drift) return fail(AMPS_E_ARG, "amps_gpu_generate_particles: null argument");
  if (!c->has_mesh) return fail(AMPS_E_STATE, "amps_gpu_generate_particles: set the mesh first");
  if (nspec <= 0 || nspec > MAX_SPEC) return fail(AMPS_E_ARG, "amps_gpu_generate_particles: nspec out of [1,8]");
  if (c->has_species && nspec > c->sp.n) return fail(AMPS_E_ARG, "amps_gpu_generate_particles: more species than set");
  if (n_per_species <= 0 || n_per_species * nspec >= ((int64_t)1 << 31) - 1)
    return fail(AMPS_E_RANGE, "amps_gpu_generate_particles: total count out of [1, 2^31-1)");
  (void)hipSetDevice(c->device);
  HIP_TRY(hipStreamSynchronize(c->stream));
  const int64_t ntot = n_per_species * nspec;
  if (int r = alloc_particles(c, ntot)) return r;
  const int64_t ncells = (int64_t)c->mv.nb * NCELL;
  GenArgs ga;
  std::memset(&ga, 0, sizeof(ga));
  ga.ntotal = ntot;
  ga.base = n_per_species / ncells;
  ga.rem = n_per_species % ncells;
  ga.nspec = nspec;
  for (int s2 = 0; s2 < nspec; s2++) ga.vth[s2] = vth[s2];
  for (int d = 0; d < 3; d++) ga.drift[d] = drift[d];
  ga.seed = seed;
  ga.vf32 = precision == AMPS_FP32 ? 1 : 0;
  c->cur = 0;
  c->layout_valid = false;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)((ntot + 255) / 256)), dim3(256), 0, c->stream, c->mv, ga,
                     c->cap, c->d_x[0], c->d_v[0], c->d_spec[0], c->d_key[0], c->d_id[0], c->d_code);
  HIP_TRY(hipGetLastError());
  const int nn = (int)ntot;
  HIP_TRY(hipMemcpyAsync(c->d_n, &nn, sizeof(int), hipMemcpyHostToDevice, c->stream));
  c->n = ntot;
  c->precision = precision;
  if (int r = compact_impl(c)) return r;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return AMPS_OK;
}

// the record move a step's compaction left to the next push
static int settle_records(amps_ctx* c) {
  if (!c->pending_gather) return 0;
  const int k = c->cur, r = 1 - k;
  hipLaunchKernelGGL(k_mig_move, dim3((unsigned)((c->cap + 255) / 256)), dim3(256), 0, c->stream, c->d_n,
                     c->d_tmpsrc, c->cap, c->d_x[r], c->d_v[r], c->precision == AMPS_FP32 ? 1 : 0, c->d_spec[r],
                     c->d_id[r], c->d_key[r], c->d_x[k], c->d_v[k], c->d_spec[k], c->d_id[k], c->d_key[k]);
  c->pending_gather = false;
  HIP_TRY(hipGetLastError());
  return 0;
}

static int one_step(amps_ctx* c, const amps_push_opts* o, double dt) {
  // the push counts the survivors per key for the migration that follows,
  // and for the deterministic one (the arrays in a compaction's layout)
  // classifies them
  const bool cls = c->deterministic && c->layout_valid;
  clear_counts(c);
  if (cls) {
    const int64_t nkeys = (int64_t)c->mv.nb * NCELL;
    const int64_t nw = (c->cap + 255) / 256 * 4;
    hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, c->stream, c->d_wcnt, nw + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1024), dim3(256), 0, c->stream, (int*)c->d_wmask, 2 * nw, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->d_mig_slab + nkeys * MIG_SLAB, 1, 0);
  }
  tbegin(c);
  if (int r = launch_push(c, o, c->d_offset, cls)) return r;
  tend(c, AMPS_K_PUSH);
  if (o->bc_mode == AMPS_BC_USER)
    if (int r = collect_exits(c)) return r;
  // the deposit follows at once: with the matrix-core (single-phase) or the
  // lane cell pass, the migration leaves the records where they are (the
  // cell pass reads them there, the next push moves them)
  const bool defer = c->deterministic && !c->dep_exact &&
                     ((c->dep_kernel == 2 && !c->dep_paired) || c->dep_kernel == 3);
  tbegin(c);
  if (int r = compact_impl(c, true, defer, cls)) return r;
  tend(c, AMPS_K_COMPACT);
  return launch_deposit(c, dt);
}

static int run_steps_impl(amps_ctx* c, const amps_push_opts* o, double dt, int64_t nsteps, int32_t use_graph) {
  if (!c->has_mesh || !c->has_fields || !c->has_species || c->cap == 0)
    return fail(AMPS_E_STATE, "amps_gpu_run_steps: mesh, fields, species and particles are required");
  if (int r = check_mover(c, o, "amps_gpu_run_steps")) return r;
  if (o->bc_mode != AMPS_BC_DELETE && o->bc_mode != AMPS_BC_REFLECT && o->bc_mode != AMPS_BC_USER)
    return fail(AMPS_E_ARG, "amps_gpu_run_steps: unknown boundary mode");
  if (nsteps <= 0) return AMPS_OK;
  (void)hipSetDevice(c->device);
  if (o->bc_mode == AMPS_BC_USER)
    if (int r = prepare_exits(c)) return r;
  if (o->tile > 0 && o->tile != c->tile)
    if (int r = amps_gpu_set_param(c, "tile", o->tile)) return r;
  if (int r = ensure_deposit_buffers(c)) return r;
  if (!c->offsets_valid)
    if (int r = amps_gpu_compact(c)) return r;
  if (!c->tiles_valid)
    if (int r = build_tiles(c)) return r;
  // eager first step (allocates every lazily sized buffer)
  if (int r = one_step(c, o, dt)) return r;
  int64_t done = 1;
  if (use_graph && nsteps > 1) {
    if (c->step_exec) (void)hipGraphExecDestroy(c->step_exec);
    if (c->step_graph) (void)hipGraphDestroy(c->step_graph);
    c->step_exec = nullptr;
    c->step_graph = nullptr;
    const bool tim = c->timing;
    c->timing = false;
    // two steps per graph: the ping-pong particle buffers and code arrays
    // return to their captured roles after every replay
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int r = one_step(c, o, dt);
    if (!r) r = one_step(c, o, dt);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->timing = tim;
    if (r) return r;
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    c->step_graph = g;
    HIP_TRY(hipGraphInstantiate(&c->step_exec, g, nullptr, nullptr, 0));
    for (; done + 2 <= nsteps; done += 2) HIP_TRY(hipGraphLaunch(c->step_exec, c->stream));
  }
  for (; done < nsteps; done++)
    if (int r = one_step(c, o, dt)) return r;
  c->offsets_valid = true;
  c->tiles_valid = true;
  return AMPS_OK;
}

extern "C" int amps_gpu_run_steps(amps_ctx* c, const amps_push_opts* o, double dt, int64_t nsteps,
                                  int32_t use_graph) {
  if (!c || !o) return fail(AMPS_E_ARG, "amps_gpu_run_steps: null argument");
  const int r = run_steps_impl(c, o, dt, nsteps, use_graph);
  // the last step's records to their places, on every path out
  const int r2 = settle_records(c);
  return r ? r : r2;
}

extern "C" int amps_gpu_copy_probe(amps_ctx* c, int64_t nbytes, int32_t reps, double* ms) {
  if (!c || nbytes < 16 || reps <= 0) return fail(AMPS_E_ARG, "amps_gpu_copy_probe: bad arguments");
  (void)hipSetDevice(c->device);
  float4 *a = nullptr, *b = nullptr;
  const int64_t n4 = nbytes / 16;
  HIP_TRY(hipMalloc(&a, n4 * 16));
  HIP_TRY(hipMalloc(&b, n4 * 16));
  hipLaunchKernelGGL(k_fill_int, dim3(4096), dim3(256), 0, c->stream, (int*)a, n4 * 4, 1);
  // one 16 KB block per workgroup (at most 65,536 workgroups, then a grid-stride loop)
  const unsigned grid = (unsigned)std::min<int64_t>((n4 + 1023) / 1024, 65536);
  hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, a, b, n4);
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  for (int i = 0; i < reps; i++)
    hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(256), 0, c->stream, a, b, n4);
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  HIP_TRY(hipEventSynchronize(c->ev1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, c->ev0, c->ev1));
  if (ms) *ms = t / reps;
  (void)hipFree(a);
  (void)hipFree(b);
  return AMPS_OK;
}
