/*
 * ampsgpu.h — C ABI of the MI355X (gfx950) particle hot path for AMPS.
 *
 * The library owns all device memory of one context; callers pass plain host
 * pointers and sizes.  Every function returns 0 on success and a nonzero
 * AMPS_E_* status otherwise; amps_gpu_last_error() returns the message of
 * the last failure on the calling thread.  One context per device; a context
 * is not re-entrant (calls on one context must not overlap).
 *
 * Each entry point names the reference interface it replaces.
 */
#ifndef AMPSGPU_H
#define AMPSGPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMPS_OK 0
#define AMPS_E_ARG 1      /* invalid argument / shape */
#define AMPS_E_STATE 2    /* call out of order (e.g. no mesh yet) */
#define AMPS_E_HIP 3      /* HIP runtime error */
#define AMPS_E_RANGE 4    /* size beyond a documented limit */
#define AMPS_E_NUMERIC 5  /* unresolved numerical failure (tracer) */

/* Mover return codes, pic.h:5955-5960, plus the two classifications the
 * reference folds into _PARTICLE_LEFT_THE_DOMAIN_. */
#define AMPS_CODE_REFLECTED (-1)  /* _PARTICLE_REJECTED_ON_THE_FACE_ (committed) */
#define AMPS_CODE_LEFT_DOMAIN 2   /* _PARTICLE_LEFT_THE_DOMAIN_: deleted on an outer face */
#define AMPS_CODE_FINISHED 3      /* _PARTICLE_MOTION_FINISHED_ */
#define AMPS_CODE_SPHERE 5        /* absorbed by the internal sphere (reference: 2) */

#define AMPS_MOVER_LAPENTA 0      /* PIC::Mover::Lapenta2017, pic_mover_boris.cpp:876-1380 */
#define AMPS_MOVER_BORIS 1        /* PIC::Mover::Boris, pic_mover_boris.cpp:126-553 */

#define AMPS_BC_DELETE 0          /* _PIC_PARTICLE_DOMAIN_BOUNDARY_INTERSECTION_PROCESSING_MODE__DELETE_ */
#define AMPS_BC_REFLECT 1         /* ..._SPECULAR_REFLECTION_ (pic_mover_boris.cpp:1236-1253) */

#define AMPS_FP64 0
#define AMPS_FP32 1               /* FP32 arithmetic, compensated FP64 positions */

typedef struct amps_ctx amps_ctx;

/* ---- lifecycle ------------------------------------------------------- */
/* Replaces the device set-up of LapentaMultiThreadedMoverGPU
 * (src/pic/pic_mover.cpp:324-575) and UpdateJMassMatrixGPU
 * (src/pic/pic_field_solver_ecsim.cpp:3998-4337). */
int amps_gpu_create(int device, amps_ctx** out);
int amps_gpu_destroy(amps_ctx* ctx);
const char* amps_gpu_last_error(void);
int amps_gpu_synchronize(amps_ctx* ctx);
/* Device properties: name (>= 64 bytes), CU count, HBM bytes. */
int amps_gpu_device_info(amps_ctx* ctx, char* name, int* n_cu, int64_t* hbm_bytes);

/* Tuning parameters (no reference counterpart; set by the autotuner):
 * "tile" (particles per push workgroup, 64..2^20), "deposit_chunks"
 * (2,3,4,6 waves per deposit workgroup), "deposit_exact" (0/1: unfused
 * mul+add in the deposit accumulation for a bit-exact summation). */
int amps_gpu_set_param(amps_ctx* ctx, const char* name, int64_t value);

/* ---- mesh ------------------------------------------------------------ */
/* The leaf blocks of cMeshAMR3d in block-id order (the reference orders
 * them along its Morton curve, meshAMRgeneric.h:11587-11620).  Block b has
 * refinement level level[b] and integer position ijk[3b..3b+2] on that
 * level's lattice; 8^3 cells per block (_BLOCK_CELLS_X_=8); lmax is
 * _MAX_REFINMENT_LEVEL_.  Replaces findTreeNode (meshAMRgeneric.h:2793-2882)
 * and FindCellIndex (:2256-2323). */
typedef struct {
  double xmin[3], xmax[3];
  int32_t lmax;
  int32_t nblocks;
} amps_mesh_desc;
int amps_gpu_set_mesh(amps_ctx* ctx, const amps_mesh_desc* desc, const int32_t* level,
                      const int64_t* ijk);
/* Unique corner nodes (corner objects shared between blocks).  Nodes are
 * ordered by lattice key ((gz*S+gy)*S+gx, S=2^lmax+1). */
int64_t amps_gpu_node_count(amps_ctx* ctx);
int amps_gpu_get_node_map(amps_ctx* ctx, int32_t* node_of_corner /* nb*729 */,
                          int64_t* node_keys /* nnodes, may be NULL */);

/* ---- fields and species ---------------------------------------------- */
/* Corner E^{n+theta} and B per block, [block][k][j][i][3] with local corner
 * i+9(j+9k) (SetBlock_E/SetBlock_B, pic_mover.cpp:84-186). */
int amps_gpu_set_fields(amps_ctx* ctx, const double* E_corner, const double* B_corner);
/* Normalized charge and mass per species, the mover time step per species
 * (_SPECIES_DEPENDENT_GLOBAL_TIME_STEP_) and the particle weight. */
int amps_gpu_set_species(amps_ctx* ctx, int32_t nspec, const double* charge,
                         const double* mass, const double* dt, const double* weight);

/* ---- particles (ParticleBuffer <-> SoA) ------------------------------ */
/* Host SoA input: x and v are three planes of n doubles (x[d*n+p]).  The
 * library sorts by (block, cell), keeping the input order within a cell. */
int amps_gpu_set_particles(amps_ctx* ctx, int64_t n, const double* x, const double* v,
                           const int32_t* spec, const int32_t* blk, const int32_t* cell,
                           const int64_t* id, int32_t precision);
int64_t amps_gpu_particle_count(amps_ctx* ctx);
/* Current device order (sorted by block, cell after a compaction). */
int amps_gpu_get_particles(amps_ctx* ctx, double* x, double* v, int32_t* spec, int32_t* blk,
                           int32_t* cell, int64_t* id, int32_t* code);

/* ---- mover ----------------------------------------------------------- */
typedef struct {
  int32_t mover;      /* AMPS_MOVER_* */
  int32_t bc_mode;    /* AMPS_BC_* */
  int32_t compact;    /* 1: migrate (sort by block/cell, drop deleted) after the push */
  int32_t tile;       /* particles per workgroup tile, 0 = autotuned/default */
  double r_sphere;    /* internal absorbing sphere radius, <= 0: none */
} amps_push_opts;
/* One PIC::Mover::MoveParticles step (pic_mover.cpp:580-1088) through the
 * per-particle Lapenta2017/Boris mover; codes stay readable until the next
 * compaction.  Exits (AMPS_CODE_LEFT_DOMAIN) are what the reference hands
 * to ProcessOutsideDomainParticles (pic.h:6043-6044). */
int amps_gpu_move_particles(amps_ctx* ctx, const amps_push_opts* opts);
/* Migration: stable-in-wave counting sort by (block, cell) with deleted
 * particles dropped (replaces the temp-list swap, pic_mover.cpp:982-992). */
int amps_gpu_compact(amps_ctx* ctx);

/* ---- ECSIM deposition ------------------------------------------------ */
/* ECSIM::UpdateJMassMatrix (pic_field_solver_ecsim.cpp:3244-3995) through
 * ProcessCell (:1882-2438).  J_out[nnodes*3] and M_out[nnodes*243] (slot
 * n = a(dx)+3a(dy)+9a(dz), a(0,-1,+1)=(0,1,2), entries 3*ii+jj,
 * MassMatrixOffsetTable :621-639) are copied to the host when non-NULL;
 * the dense [node][246] result stays resident either way. */
int amps_gpu_update_j_mass_matrix(amps_ctx* ctx, double dt, double* J_out, double* M_out);

/* ---- profiling ------------------------------------------------------- */
/* Per-kernel HIP-event timing on the library stream (ms, accumulated). */
#define AMPS_K_PUSH 0
#define AMPS_K_COMPACT 1
#define AMPS_K_DEPOSIT_CELL 2
#define AMPS_K_DEPOSIT_NODE 3
#define AMPS_K_TRACE 4
#define AMPS_K_MLP 5
#define AMPS_K_COUNT 6
int amps_gpu_set_timing(amps_ctx* ctx, int32_t enable);
int amps_gpu_get_timing(amps_ctx* ctx, double* ms /* AMPS_K_COUNT */, int64_t* launches);
int amps_gpu_reset_timing(amps_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
