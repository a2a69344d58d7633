// amps_common.h — shared definitions of the gfx950 particle hot path library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ampsgpu.h"

namespace amps {

constexpr int NC = 8;                  // cells per block edge (_BLOCK_CELLS_X_)
constexpr int NN = 9;                  // corner nodes per block edge
constexpr int NCELL = NC * NC * NC;    // 512
constexpr int NCORNER = NN * NN * NN;  // 729
constexpr int CELL_VALS = 24 + 36 * 9; // per-cell J (8x3) + lower-triangular 8x8 pairs x 9
// per-cell J (8x3) + 27 weight-product moments x 9: the pair block of corners
// (i, j) depends only on the per-axis sums of their offset bits
constexpr int CELL_VALS_MOMENT = 24 + 27 * 9;
constexpr int NODE_VALS = 3 + 243;     // per-node J + mass-matrix row block
// Record strides padded to whole 128-byte lines: a line that a store only
// partly covers is first read from HBM, so records that start and end on line
// boundaries and are written by full-line stores cost no reads.
constexpr int CELL_STRIDE_MOMENT = 272;  // 17 lines
constexpr int NODE_STRIDE = 256;         // 16 lines
constexpr int MAX_SPEC = 8;

// Device view of the block mesh (leaf blocks of the AMR octree).
struct MeshView {
  double gxmin[3], gxmax[3], dx_root[3], dx_maxref[3], eps;
  int lmax, nb, lf, nf;      // finest block level present and its lattice size
  const double* blk_xmin;    // [nb*3]
  const double* blk_xmax;    // [nb*3]
  const double* blk_dxc;     // FindCellIndex cell size dxRootBlock/2^L/8, [nb*3]
  const int* blk_level;      // refinement level, [nb]
  const int* nbinfo;         // min/max neighbour level and face/edge/corner neighbours, [nb*28]
  const double* blk_dxs;     // InitStencil cell size (xmax-xmin)/8, [nb*3]
  const int* leaf_of;        // [nf^3] finest-level lattice -> block id
  const int* blk_lat;        // [nb*4] block origin on the level-lmax lattice, and its extent
};

struct Species {
  int n;
  double charge[MAX_SPEC], mass[MAX_SPEC], dt[MAX_SPEC], weight[MAX_SPEC];
};

struct Tile {
  int block;
  int start;
  int count;
  int pad;
};

// Boundary face of the root box: ExternalBoundaryFaceTable (pic_mover.cpp:24-28).
struct Face {
  double norm[3], e0[3], e1[3], x0[3], lE0, lE1;
};
struct Faces {
  Face f[6];
};

// Where a push reads its particles.  src null: in place.  Otherwise the
// particles are still where the last compaction found them: particle p of the
// new order is read at src[p] from x, v, spec, id (the other buffer), and its
// species and id are written to spec_out, id_out at p with its new state (the
// push makes the migration's record move).
struct PushSrc {
  const int* src;
  const double* x;
  const void* v;
  const uint8_t* spec;
  const int64_t* id;
  uint8_t* spec_out;
  int64_t* id_out;
};

// The migration's classification made by the push (a deterministic step
// whose arrays have the layout of a previous compaction; lay null: none).
// lay[] holds that layout's key offsets.  A survivor whose new key's old run
// holds it (a stayer) sets its bit of the 64-aligned stayer mask word wmask
// and counts in wcnt and in the push's count[]; a survivor that changes keys
// (a mover) counts in cursor[] and takes its rank among its new key's movers
// from that atomic, and is listed at slab[rank * nkeys + key] (rank-major:
// the first movers of neighbouring cells share cache lines), or, past the
// slab, as (position, rank) in ovf[] (novf entries).
constexpr int MIG_SLAB = 16;
struct PushMig {
  int64_t nkeys;
  const int* lay;
  int* cursor;
  int* slab;
  int2* ovf;
  int* novf;
  uint64_t* wmask;
  int* wcnt;
};

// Synthetic-population generator arguments (k_generate).
struct GenArgs {
  int64_t ntotal, base, rem;
  int nspec;
  double vth[MAX_SPEC], drift[3];
  uint64_t seed;
  int vf32;  // velocities stored as float
};

}  // namespace amps

struct amps_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // mesh
  bool has_mesh = false;
  amps::MeshView mv{};
  std::vector<int> h_level;
  std::vector<int64_t> h_ijk;
  std::vector<double> h_xmin, h_xmax;
  double *d_xmin = nullptr, *d_xmax = nullptr, *d_dxc = nullptr, *d_dxs = nullptr;
  int* d_leaf = nullptr;
  int* d_blk_lat = nullptr;
  int* d_level = nullptr;   // [nb] block refinement levels
  int* d_nbinfo = nullptr;  // [nb][28] min/max neighbour level, face[6], edge[12], corner[8] neighbours
  double* d_Bc = nullptr;   // [nb][10^3][3] cell-centred B with one ghost layer
  bool has_center = false;
  double* d_Bm = nullptr;   // [nb][512][3] Mode3D compact field of interior cells
  bool has_mode3d = false;
  amps::Faces faces{};
  int64_t nnodes = 0;
  std::vector<int32_t> h_node_of_corner;
  std::vector<int64_t> h_node_keys;
  int* d_inc8 = nullptr;  // [nnodes][8] in node_order: cell_gid*8 + corner, ascending cell order, -1 padded
  // [nnodes][8] in node_order, slot t = the octant (t & 1, t >> 1 & 1, t >> 2)
  // of the node that the cell lies in; an empty octant holds zero_rec*8 + the
  // corner a cell there would have (the moment-record gathers read zeros)
  int* d_inc8_oct = nullptr;
  int64_t zero_rec = 0;   // a moment record index past every cell record, kept zero
  int* d_node_order = nullptr; // node processing order of the corner flush
  // fields and species
  bool has_fields = false;
  double *d_E = nullptr, *d_B = nullptr;
  amps::Species sp{};
  bool has_species = false;
  // particles: ping-pong SoA buffers
  int64_t n = 0, cap = 0;
  int precision = AMPS_FP64;
  int cur = 0;
  double* d_x[2] = {nullptr, nullptr};
  double* d_v[2] = {nullptr, nullptr};
  uint8_t* d_spec[2] = {nullptr, nullptr};
  int* d_key[2] = {nullptr, nullptr};
  int64_t* d_id[2] = {nullptr, nullptr};
  int8_t* d_code = nullptr;    // per-particle mover code of the current order
  int8_t* d_code2 = nullptr;   // spare code buffer (swapped at compaction)
  int* d_n = nullptr;          // live particle count on the device
  // migration and tiling
  int* d_count = nullptr;      // [nkeys+1]
  int* d_offset = nullptr;     // [nkeys+1]
  int* d_cursor = nullptr;     // [nkeys]
  int* d_scan_tmp = nullptr;
  int64_t scan_tmp_len = 0;
  amps::Tile* d_tiles = nullptr;
  int* d_ntiles = nullptr;
  int* d_tcount = nullptr;     // per-block tile count / offsets [nb+1]
  int64_t max_tiles = 0;
  int tile = 8192;
  bool offsets_valid = false;  // d_offset describes the current order
  bool tiles_valid = false;    // d_tiles describes the current order
  int dep_nch = 12;            // deposit output chunks per cell (waves per workgroup)
  int node_batch = 8;          // nodes per wavefront in the deposit node gather (0: grid-stride)
  int n_cu = 256;              // compute units of the device
  int node_staged = 1;         // moment-record node gather staged through LDS
  int deterministic = 1;       // migration: in-key order independent of atomic scheduling
  int mig_small = 64;          // movers per key sorted by one lane (more: a workgroup merge sort)
  int dep_waves = 2;           // eight-lanes-per-cell pass: register budget in wavefronts per SIMD (2 or 3)
  int dep_paired = 0;          // matrix-core cell pass: two wavefronts per SIMD in opposite phases
  int push_waves = 1;          // 4: push kernels built for 4 waves per SIMD (128 VGPRs)
  int dep_exact = 0;           // 1: unfused mul+add in the deposit (bit-exact order)
  int dep_kernel = 2;          // cell pass: 0 per-lane particle loop, 1 LDS record batches, 2 matrix cores,
                               // 3 eight lanes per cell
  int b_center = 0;            // 1: B from the cell-centred field (mover and deposit)
  // deposit
  double* d_cellpart = nullptr;
  double* d_nodeout = nullptr;
  int* d_ecsim_nbr = nullptr;  // [nnodes][27] neighbour nodes of the ECSIM stencil, -1 outside
  // ECSIM field equation (ecsim_field.hip)
  uint8_t* d_ecsim_bnd = nullptr;  // [nnodes] 1: corner on the domain boundary
  int* d_ecsim_cells8 = nullptr;   // [nnodes][8] cells at node + (ci, cj, ck), ci, cj, ck in {-1, 0}
  int* d_ecsim_noc = nullptr;      // [nblocks][729] node of each block corner
  double* d_ecsim_work = nullptr;  // field vectors, the Krylov basis and reduction partials
  size_t ecsim_work_len = 0;       // doubles allocated at d_ecsim_work
  // deterministic migration: the offsets of the current array layout, keys
  // with many movers, the movers' destinations and source positions, and the
  // stayer count (scanned) and mask of every wavefront of the scatter grid
  int* d_off_layout = nullptr;  // [nkeys + 1]
  int* d_big = nullptr;         // [nkeys + 1]: keys with many movers, count last
  int* d_kbase = nullptr;       // [nkeys]: a key's new offset less the stayers before its old run
  int* d_dest = nullptr;        // [cap]
  int* d_tmpsrc = nullptr;      // [cap]
  int* d_wcnt = nullptr;        // [ceil(cap / 256) * 4 + 1]
  uint64_t* d_wmask = nullptr;  // [ceil(cap / 256) * 4]
  // the push's classification (PushMig): movers listed per key, the rest
  // after the slab
  int* d_mig_slab = nullptr;    // [nkeys * MIG_SLAB + 1], the overflow count last
  int2* d_mig_ovf = nullptr;    // [cap]
  bool layout_valid = false;
  // the last compaction left the particle records in the other buffer: the
  // destination's source indices are in d_tmpsrc.  The step's matrix-core
  // cell pass reads the particles through them, the next push moves them;
  // amps_gpu_run_steps, the only caller that sets it, moves any left over
  // before it returns.
  bool pending_gather = false;
  // exit list of the USER_FUNCTION boundary mode (exits.hip)
  int64_t exit_cap = (int64_t)1 << 20;  // records the list holds
  int64_t exit_alloc = 0;               // records allocated
  int64_t* d_ex_id = nullptr;           // [exit_alloc]
  double* d_ex_x = nullptr;             // [exit_alloc][3]
  double* d_ex_v = nullptr;             // [exit_alloc][3]
  int* d_ex_meta = nullptr;             // [exit_alloc][3]: species, face, block
  int* d_ex_wg = nullptr;               // per-workgroup exit counts, then offsets
  int64_t* d_ex_state = nullptr;        // [0] records in the list, [1] records dropped
  // timing
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double ms[AMPS_K_COUNT] = {0};
  int64_t launches[AMPS_K_COUNT] = {0};
  // captured full-step graph
  hipGraph_t step_graph = nullptr;
  hipGraphExec_t step_exec = nullptr;
  // tracer / surrogate state lives in their own files
  void* trace_state = nullptr;
  void* mlp_state = nullptr;
};

namespace amps {
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// timing helpers (no-ops when timing is off)
void tbegin(amps_ctx* c);
void tend(amps_ctx* c, int k);
// per-feature state release (trace.hip, mlp.hip)
void trace_free(amps_ctx* c);
void mlp_free(amps_ctx* c);
// device-wide exclusive scan of n ints in place; total written to total_out
int scan_exclusive(amps_ctx* c, int* data, int64_t n, int* total_out);
// ECSIM cell pass with LDS record batches (deposit.hip)
int launch_deposit_cell_rec(amps_ctx* c, double dt, int nch);
int launch_deposit_cell_mfma(amps_ctx* c, double dt);
int launch_deposit_cell_lanes(amps_ctx* c, double dt);
int alloc_particles(amps_ctx* c, int64_t n);  // both particle sets for at least n
int build_tiles(amps_ctx* c);                 // push tiles from the current offsets
// append the exits of the last push (codes AMPS_CODE_EXIT_FACE0..+5) to the
// exit list (exits.hip); no host synchronisation, capturable
int collect_exits(amps_ctx* c);
int prepare_exits(amps_ctx* c);  // allocates the list (outside any capture)
void exits_free(amps_ctx* c);
// looks up this device's address of the energy/CFL scratch array (diag.hip)
int energy_scratch_init(int device);
// ECSIM corner-stencil neighbour table [nnodes][27] (ecsim.hip) and the test
// that the mesh has one refinement level, as the field solver needs
int build_nbr(amps_ctx* c);
bool single_level(const amps_ctx* c);
// field-solver tables and work buffers (ecsim_field.hip)
void ecsim_field_free(amps_ctx* c);
}  // namespace amps

#define HIP_TRY(expr)                                    \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return amps::hip_fail(e_, #expr); \
  } while (0)
