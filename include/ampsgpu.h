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
/* AMPS_BC_USER: left the domain through external face f = code - 8 (0..5,
 * ExternalBoundaryFaceTable order), the state moved onto the face and the
 * particle queued for ProcessOutsideDomainParticles, then deleted. */
#define AMPS_CODE_EXIT_FACE0 8

#define AMPS_MOVER_LAPENTA 0      /* PIC::Mover::Lapenta2017, pic_mover_boris.cpp:876-1380 */
#define AMPS_MOVER_BORIS 1        /* PIC::Mover::Boris, pic_mover_boris.cpp:126-553 */
#define AMPS_MOVER_RELATIVISTIC 2 /* PIC::Mover::Relativistic::Boris, pic_mover_relativistic_boris.cpp:16-583 (FP64) */

#define AMPS_BC_DELETE 0          /* _PIC_PARTICLE_DOMAIN_BOUNDARY_INTERSECTION_PROCESSING_MODE__DELETE_ */
#define AMPS_BC_REFLECT 1         /* ..._SPECULAR_REFLECTION_ (pic_mover_boris.cpp:1236-1253) */
#define AMPS_BC_USER 2            /* ..._USER_FUNCTION_ (pic_mover_boris.cpp:1156-1231): exits go
                                     to the exit list (amps_gpu_get_exits) */

#define AMPS_FP64 0
#define AMPS_FP32 1               /* FP32 push arithmetic and velocities; FP64 positions,
                                     FP32 increments accumulated in FP64; FP64 deposit */

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
 * (waves per deposit workgroup: 2,3,4,6 for kernel 0, 4,6,12 for kernels 1
 * and 2; particles per cell per record batch, 4 or 8, for kernel 3),
 * "deposit_exact" (0/1: unfused mul+add in the deposit accumulation
 * for a bit-exact summation; kernels 2 and 3 then run kernel 1's exact form),
 * "deposit_kernel" (0: each lane evaluates its cell's particles itself,
 * 1: particles evaluated once into shared-memory record batches,
 * 2 (default): the per-cell moment sums as FP64 matrix-core products,
 * 3: eight lanes per cell on the vector ALUs), "deposit_waves" (kernel 3:
 * 2 (default) or 3 wavefronts per SIMD), "deterministic" (1 (default): the
 * migration keeps, in every cell, the particles that stayed in their
 * previous order followed by those that moved in, in array order, so that
 * repeated runs give the same layout and bit-identical deposits; 0: the
 * order in which the atomics of the scatter land), "migration_sort_small"
 * (1..64, default 64: cells receiving more movers than this are ordered by a
 * workgroup merge sort instead of one lane), "deposit_paired"
 * (kernel 2: 1 two wavefronts per SIMD in opposite VALU/MFMA phases,
 * 0 (default): one phase sequence per wavefront), "node_batch"
 * (1..64 consecutive nodes per wavefront in the deposit node gather,
 * default 8; 0: a resident grid whose wavefronts stride over the nodes, for
 * the staged gather — the others then take 8),
 * "node_staged" (1 (default): the node gather of the moment record copies
 * each node's inputs through shared memory with run-wise loads; 0: one
 * gathered load per output value),
 * "push_waves" (1: push kernels with the compiler's register allocation,
 * 4: held to 128 registers for four wavefronts per SIMD),
 * "b_center" (1: the FP64 Lapenta/Boris movers and the deposit take B from
 * the cell-centred field of amps_gpu_set_center_fields, the centre-based B
 * mode _PIC_FIELD_SOLVER_B_MODE_; 0 (default): B at the corners),
 * "exit_capacity" (records of the AMPS_BC_USER exit list, 1..2^30, default
 * 2^20; a new value empties the list at the next USER_FUNCTION push). */
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
/* Cell-centred B with one ghost layer per block, [block][10][10][10][3],
 * cell (i,j,k) in [-1,8]^3 at (i+1) + 10((j+1) + 10(k+1)): what SetBlock_B
 * stages for a centre-based B (pic_mover.cpp:84-186); the caller fills the
 * ghost cells (its halo exchange).  A ghost cell whose centre is outside the
 * domain is absent from every stencil. */
int amps_gpu_set_center_fields(amps_ctx* ctx, const double* B_center);
/* B at n points x (planes x[d*n+p]) of blocks blk by the cell-centred linear
 * stencil CellCentered::Linear::InitStencil (pic_interpolation_routines.cpp:
 * 224-1070): trilinear on cell centres, the multi-block stencil with 2x2x2
 * fine averages at coarse/fine interfaces and the coarse/fine blend. */
int amps_gpu_interp_center(amps_ctx* ctx, int64_t n, const double* x, const int32_t* blk, double* B);
/* Mode3D compact field (GlobalMagneticField.h:18-30): B of every block's
 * interior cells, [block][512][3] with cell i + 8(j + 8k); the block id is
 * the leaf's dense Temp_ID.  The tracer's mesh field (field_model 1) reads
 * it. */
int amps_gpu_set_mode3d_field(amps_ctx* ctx, const double* B_cells);
/* B at n points x[3n] from that field as cMode3DMeshFieldEval::GetB_T
 * (CutoffRigidityMode3D.cpp:702-800): zero outside the mesh, otherwise the
 * row stencil of CellCentered::Linear::InitStencil (pic_interpolation_
 * routines.cpp:224-1070 with a cRowStencil, pic.h:7295-7420), every support
 * cell taken to the leaf and interior cell holding its centre
 * (:33-105), summed as InterpolateField_ (GlobalMagneticField.cpp:292-330). */
int amps_gpu_mode3d_field(amps_ctx* ctx, int64_t n, const double* x, double* B);

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

/* ParticleBuffer records (picParticleDataMacro.h): ParticleDataLength bytes
 * per particle, next/prev list links at bytes 0 and 8, and the byte offsets
 * of the mandatory fields: _PIC_PARTICLE_DATA__SPECIES_ID_OFFSET_ (8 bytes in
 * the aligned state-vector mode, 1 otherwise), __VELOCITY_OFFSET_ and
 * __POSITION_OFFSET_ (3 doubles each) and __WEIGHT_CORRECTION_OFFSET_ (-1
 * when _INDIVIDUAL_PARTICLE_WEIGHT_MODE_ is off; the device carries no
 * per-particle weight: exported as 1, ignored on import).  Optional fields
 * are exported as zero bytes and ignored on import. */
typedef struct {
  int32_t record_bytes;
  int32_t species_offset;
  int32_t species_bytes;
  int32_t velocity_offset;
  int32_t position_offset;
  int32_t weight_offset;
} amps_aos_layout;
/* The particle part of a restart file as SaveParticleDataBlock writes it
 * (pic_restart.cpp:312-370), for all blocks at once: counts[512 b + 64 i +
 * 8 j + k] = ParticleNumberTable[i][j][k] of block b, and the records of
 * block 0, 1, ... each cell by cell in that i, j, k order, a cell's particles
 * from the last of its list to the first (the device order within a cell is
 * the list order from its head).  The caller writes AMRnodeID and the total
 * of each block with particles in front of its table.  Compacts first if the
 * particles are not sorted.  records may be NULL (counts only). */
int amps_gpu_get_particles_restart(amps_ctx* ctx, const amps_aos_layout* layout, int32_t* counts,
                                   void* records, int64_t cap_records, int64_t* n_records);
/* The inverse, ReadParticleDataBlock (pic_restart.cpp:430-512): counts and
 * records in the order above for all blocks (zero counts for blocks absent
 * from the file); every record is pushed onto the head of its cell, so the
 * device order within a cell is the file order reversed.  Particle ids are
 * the record indices. */
int amps_gpu_set_particles_restart(amps_ctx* ctx, const amps_aos_layout* layout, const int32_t* counts,
                                   int64_t n_records, const void* records, int32_t precision);

/* ---- mover ----------------------------------------------------------- */
typedef struct {
  int32_t mover;      /* AMPS_MOVER_* */
  int32_t bc_mode;    /* AMPS_BC_* */
  int32_t compact;    /* 1: migrate (sort by block/cell, drop deleted) after the push */
  int32_t tile;       /* particles per workgroup tile, 0 = autotuned/default */
  double r_sphere;    /* internal absorbing sphere radius, <= 0: none */
} amps_push_opts;
/* One PIC::Mover::MoveParticles step (pic_mover.cpp:580-1088) through the
 * per-particle Lapenta2017, Boris or relativistic Boris mover (the last in
 * FP64 only, with SpeedOfLight in the units of the mesh and fields); codes
 * stay readable until the next
 * compaction.  Exits (AMPS_CODE_LEFT_DOMAIN) are what the reference hands
 * to ProcessOutsideDomainParticles (pic.h:6043-6044). */
int amps_gpu_move_particles(amps_ctx* ctx, const amps_push_opts* opts);
/* The exit list of AMPS_BC_USER: every particle that left the domain through
 * an external face since the list was last cleared, in device order and step
 * order, with the arguments the reference passes to
 * PIC::Mover::ProcessOutsideDomainParticles(ptr, xInit, vInit,
 * nIntersectionFace, newNode) (pic.h:6043-6044; pic_mover_boris.cpp:1156-1231,
 * pic_mover_relativistic_boris.cpp:317-420): the particle id, the state moved
 * onto the exit face (x[3i..3i+2] at the face intersection minus EPS times
 * the face normal, v[3i..3i+2] advanced to the intersection time), its
 * species, the face index and the block holding x (-1 when none, where the
 * reference aborts).  The list is built on the device (per-workgroup counts,
 * a scan, ordered writes) after each push, so only the exits cross PCIe; its
 * capacity is the parameter "exit_capacity" (default 2^20 records).
 * *n_exits receives the count; min(count, cap) records are copied (NULL
 * arrays: count only); clear = 1 empties the list.  Returns AMPS_E_RANGE when
 * exits were dropped because the list was full. */
int amps_gpu_get_exits(amps_ctx* ctx, int64_t cap, int64_t* n_exits, int64_t* id, double* x, double* v,
                       int32_t* spec, int32_t* face, int32_t* blk, int32_t clear);

/* Migration: stable-in-wave counting sort by (block, cell) with deleted
 * particles dropped (replaces the temp-list swap, pic_mover.cpp:982-992). */
int amps_gpu_compact(amps_ctx* ctx);

/* Synthetic population generated on the device (no reference counterpart;
 * the reference's test apps fill cells on the host, srcFastWave/main.cpp:
 * 120-300): n_per_species particles of each species spread over all cells
 * (equal counts, the first cells get one extra), uniform in each cell,
 * Maxwellian with thermal speed vth[s] plus drift[3].  Ids are consecutive;
 * block/cell ids come from the same point location as the mover. */
int amps_gpu_generate_particles(amps_ctx* ctx, int64_t n_per_species, int32_t nspec,
                                const double* vth, const double* drift, uint64_t seed,
                                int32_t precision);

/* nsteps full particle steps: MoveParticles, migration (compaction) and
 * UpdateJMassMatrix at deposit_dt.  use_graph=1 runs one eager step, then
 * replays a hipGraph holding two captured steps (an odd remainder runs
 * eagerly).  On return the particle arrays (in array order) and the
 * resident J and M are those of the same steps made by separate
 * amps_gpu_move_particles (with compaction) and amps_gpu_update_j_mass_matrix
 * calls; inside the call a step may leave its particle records for the next
 * step's push to move. */
int amps_gpu_run_steps(amps_ctx* ctx, const amps_push_opts* opts, double deposit_dt,
                       int64_t nsteps, int32_t use_graph);

/* Device copy of nbytes (read + write) repeated reps times; ms receives the
 * average duration.  Used to calibrate HBM counters and bandwidth. */
int amps_gpu_copy_probe(amps_ctx* ctx, int64_t nbytes, int32_t reps, double* ms);

/* ---- ECSIM deposition ------------------------------------------------ */
/* ECSIM::UpdateJMassMatrix (pic_field_solver_ecsim.cpp:3244-3995) through
 * ProcessCell (:1882-2438).  J_out[nnodes*3] and M_out[nnodes*243] (slot
 * n = a(dx)+3a(dy)+9a(dz), a(0,-1,+1)=(0,1,2), entries 3*ii+jj,
 * MassMatrixOffsetTable :621-639) are copied to the host when non-NULL;
 * the dense [node][246] result stays resident either way. */
int amps_gpu_update_j_mass_matrix(amps_ctx* ctx, double dt, double* J_out, double* M_out);

/* The particle diagnostics of ECSIM::UpdateJMassMatrix (pic_field_solver_
 * ecsim.cpp:2229-2238, 2355-2358, 3925-3929, 3972-3987) over the resident
 * particles: *energy = the sum of 0.5 m w |v|^2 (ParticleEnergy), cfl[s] = the
 * maximum over cells of (sum over the cell's particles of species s of
 * |v| dt) / (count |dx|) (the per-species max CFL), cells summed in particle
 * order, blocks in block order.  cfl may be NULL. */
int amps_gpu_particle_energy_cfl(amps_ctx* ctx, double dt, double* energy, double* cfl);

/* No counterpart in the reference: amps_gpu_particle_energy_cfl for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * loops, harnesses that forbid hipMalloc).  It calls no hipMalloc/hipFree:
 * the block partials live in a static device array of the library sized for
 * at most 32768 blocks (2.25 MiB per GPU, held until the process ends; more
 * blocks give AMPS_E_RANGE), and the result goes to the caller's DEVICE array
 * out[1 + nspec] (out[0] the energy, out[1 + s] the CFL of species s),
 * asynchronously on the context's stream (amps_gpu_synchronize before
 * reading it).  The particles must be sorted (amps_gpu_compact, or any step
 * that compacts), else AMPS_E_STATE.  All calls on one GPU share the scratch
 * array: calls that overlap on one GPU race. */
int amps_gpu_particle_energy_cfl_noalloc(amps_ctx* ctx, double dt, double* out);

/* The mass-matrix term of the ECSIM field equation over the resident M of
 * the last amps_gpu_update_j_mass_matrix: for every node n and component c,
 *   out[3n+c] = sum over slots s of ((Ex M[9s+3c] + Ey M[9s+3c+1]) + Ez M[9s+3c+2]) * c1 * c2
 * with E of the slot's neighbour node (UpdateRhs_Legacy mass part,
 * pic_field_solver_ecsim.cpp:6116-6124, c1 = -4 pi dt theta, c2 = E_conv;
 * the neighbour order of GetStencil :1061-1074).  E_nodes and out are
 * [nnodes][3] in node-key order.  accumulate = 1: out holds on entry the
 * part of each row already summed (UpdateRhs_Legacy's res before its mass
 * loop, :6040-6115) and the 27 terms are added onto it one at a time, as the
 * reference does; 0: the sum starts from zero.  A neighbour is one cell away;
 * one outside the domain contributes 0.  The reference's ECSIM solver runs on
 * one resolution level (UpdateJMassMatrix takes BlockTable[0]'s level,
 * :3566): a mesh of several levels is refused with AMPS_E_STATE.  reps > 1
 * repeats the launch (for timing). */
int amps_gpu_ecsim_mass_apply(amps_ctx* ctx, const double* E_nodes, double c1, double c2, double* out,
                              int32_t accumulate, int32_t reps);

/* ---- ECSIM implicit field equation (single-level mesh) ---------------- */
/* Units and step of the field solve: dtTotal, theta (0.5 in the reference),
 * cDt = LightSpeed * dtTotal in the solver's units, and the conversions
 * E_conv, B_conv, length_conv of pic_field_solver_ecsim.cpp:202-212.  The
 * div-E correction (corrCoeff, the 375-point stencil) is not built: it is 0,
 * the reference's default (:99). */
typedef struct {
  double dt;
  double theta;
  double cDt;
  double E_conv;
  double B_conv;
  double length_conv;
} amps_ecsim_params;

/* y = A x, the matrix of GetStencil_Legacy (pic_field_solver_ecsim.cpp:
 * 740-1228) with the mass term of UpdateMatrixElement (:6004-6016) over the
 * resident M: identity - (c dt theta)^2 (Laplacian - grad div) + 4 pi dt
 * theta M on the 27 corner neighbours; rows of corners on the domain
 * boundary are the identity, and such corners are dropped from the other
 * rows (:782-799, :1032).  x, y: [nnodes][3] in node-key order.  Needs the
 * J/M of amps_gpu_update_j_mass_matrix and a single-level mesh (the
 * reference's solver runs on one level), else AMPS_E_STATE.  reps > 1
 * repeats the launch (timing). */
int amps_gpu_ecsim_operator(amps_ctx* ctx, const amps_ecsim_params* p, const double* x, double* y, int32_t reps);

/* The right-hand side of UpdateRhs_Legacy (:6040-6137) for every row: the
 * 27 neighbours' E^n times the Laplacian / grad-div coefficients, the mass
 * term and the node's J times -4 pi dt theta, and c dt theta curl B from the
 * 8 + 8 cell-centred B of :1234-1375.  E_n: [nnodes][3]; B_c: [nblocks][512]
 * [3] cell-centred B (cell i + 8j + 64k of each block); rhs: [nnodes][3], 0
 * on boundary rows. */
int amps_gpu_ecsim_rhs(amps_ctx* ctx, const amps_ecsim_params* p, const double* E_n, const double* B_c, double* rhs);

/* Solves A dE = rhs on the device with restarted GMRES (zero initial guess,
 * SetInitialGuess :6296-6301; stops when ||rhs - A dE|| <= tol ||rhs|| or
 * after max_iter Arnoldi steps).  The reference calls SWMF's
 * linear_solver_wrapper("GMRES") (srcInterface/LinearSystemCornerNode.h:3282);
 * its iterates are not reproduced, its converged solution is.  iters: the
 * Arnoldi steps taken; rel_residual: ||rhs - A dE|| / ||rhs|| of the result. */
int amps_gpu_ecsim_solve(amps_ctx* ctx, const amps_ecsim_params* p, const double* rhs, double tol, int32_t max_iter,
                         int32_t restart, double* dE, int32_t* iters, double* rel_residual);

/* UpdateB (:5160-5270) for every cell: B^{n+1} = B^n - c dt curl E^{n+theta}
 * from the 2x2 face averages of the cell's corners (the interior formula in
 * every cell; the reference's Dirichlet/Neumann treatment of domain-boundary
 * cells is left to the caller). */
int amps_gpu_ecsim_update_b(amps_ctx* ctx, const amps_ecsim_params* p, const double* E_theta, const double* B_c,
                            double* B_out);

/* One field step of ECSIM::TimeStep (:6179-6270) on the device: the
 * right-hand side, GMRES, ProcessFinalSolution (E^{n+theta} = dE / E_conv +
 * E^n, :6311-6318), UpdateE (E^{n+1} = (E^{n+theta} - (1 - theta) E^n) /
 * theta, :5909-5960) and UpdateB.  E: E^n in, E^{n+1} out; E_theta (may be
 * NULL): E^{n+theta} out; B_c: B^n in, B^{n+1} out. */
int amps_gpu_ecsim_time_step(amps_ctx* ctx, const amps_ecsim_params* p, double* E, double* E_theta, double* B_c,
                             double tol, int32_t max_iter, int32_t restart, int32_t* iters, double* rel_residual);

/* ---- gridless cutoff tracer ------------------------------------------ */
/* Parameters of one trace (SI).  The FP64 path is TraceTrajectoryImpl +
 * TraceTrajectoryWithSingleRetry (CutoffRigidityGridless.cpp:1246-1496) with
 * StepParticleChecked (GridlessParticleMovers.cpp:1035-1098) of the selected
 * full-orbit mover: BORIS (:781-801), HC4 (:723-779, four checked positive
 * macro-steps) or RK2/RK4/RK6 (RKStepGeneric :655-692, stage chords
 * checked); the legacy step selector (CutoffRigidityGridless.cpp:1086-1152)
 * whose gyro budget, distance limiter and floor are switchable, the analytic
 * dipole (DipoleInterface.h:99-125) plus a uniform field, and a box or
 * spherical outer boundary.  The FP32 path runs BORIS and HC4. */
typedef struct {
  double q_C;            /* signed backtrace charge [C] */
  double qabs_C;         /* |physical charge| for rigidity conversion [C] */
  double m0_kg;
  double dt_trace_s;     /* DT_TRACE */
  double max_time_s;
  int64_t max_steps;
  double max_distance_m; /* <= 0: off */
  double dphi_max;       /* gyro-phase budget per step [rad] (legacy 0.15) */
  int32_t adaptive;      /* adaptiveDt */
  int32_t dist_limiter;  /* 0.20 * d_near / v */
  int32_t dt_floor;      /* legacy max(1e-12, 1e-9*DT, 100 km / v) floor */
  int32_t outer_sphere;  /* 0: box, 1: sphere of radius r_outer_Re */
  int32_t retry;         /* 0 none; 1 single retry on a numerical failure
                            (TraceTrajectoryWithSingleRetry); 2 single retry of
                            every unresolved trace (DS_RETRY_UNRESOLVED) */
  int32_t mover;         /* MoverType (GridlessParticleMovers.h:300-310): 0 BORIS,
                            1 HC4, 2 RK2, 3 RK4, 4 RK6; GC2..HYBRID rejected */
  double box_min_Re[3], box_max_Re[3];
  double r_outer_Re, r_inner_Re;
  double earth_radius_m; /* length of one Re for all geometry conversions */
  double moment_Am2;     /* dipole |m| */
  double m_hat[3];
  double B_ext_T[3];     /* uniform field added to the dipole */
  int32_t field_model;   /* 0: analytic dipole + B_ext_T; 1: Mode3D mesh field
                            (amps_gpu_set_mode3d_field, cMode3DMeshFieldEval) */
  int32_t dt_selector;   /* 0: SelectLegacyCutoffTraceDt (gridless); 1: its Mode3D
                            form SelectLegacyCutoffTraceDt3D (CutoffRigidityMode3D.cpp:
                            881-936: box and inner distances in metres, limiter
                            and floor always on) */
  int32_t policy;        /* 0: the legacy cutoff policy (boundaries tested at the
                            step start); 1: StructuredAccurate, the density /
                            transmission policy of TraceTrajectoryShared(Ex)
                            (CutoffRigidityGridless.cpp:1154-1228, 1403-1440,
                            1563-1585): SelectStructuredTraceDt, the first
                            boundary event on every accepted chord
                            (util/TrajectoryBoundary.h:133-141), a normalised
                            start direction; box boundary, FP64 only */
  int32_t pad2_;
  double event_tol_m;    /* boundaryEventTolerance_m (default 1 m) */
  /* TrajectoryTrap::Detector (util/TrajectoryTrapDetector.h:23-153), the
   * MagneticallyTrappedForbidden termination of both policies: TRAP_DETECTION,
   * TRAP_MIN_MIRROR_POINTS (8), TRAP_MIN_BOUNCE_CYCLES (4), TRAP_OUTER_MARGIN
   * (1 Re), TRAP_RADIAL_GROWTH_TOLERANCE (0.05 Re), TRAP_ENERGY_RELATIVE_
   * TOLERANCE (1e-4), TRAP_PARALLEL_DEADBAND (1e-6); lengths in earth radii */
  int32_t trap;
  int32_t trap_min_mirror, trap_min_bounce, pad3_;
  double trap_outer_margin_Re, trap_radial_tol_Re, trap_energy_tol, trap_deadband;
} amps_trace_params;

/* TrajectoryExitState (CutoffRigidityGridless.h:258-269) of an allowed
 * structured trace (PopulateOuterExit, CutoffRigidityGridless.cpp:1323-1358);
 * valid = 0 for every other trajectory. */
typedef struct {
  double x_exit_m[3], v_exit_unit[3], cos_alpha;
  int32_t valid, pad_;
} amps_trace_exit;

/* TrajectoryResult (termination values of util/TrajectoryTermination.h:10-21) */
typedef struct {
  int32_t termination;
  int32_t retry_count;
  int64_t steps;
  double time_s, distance_m;
  double x[3], p[3];
} amps_trace_result;

/* Batched TraceTrajectoryShared (CutoffRigidityGridless.h:245-315): n
 * trajectories from x0[3n] (m) with arrival directions dir[3n] (the backtrace
 * starts along -dir) at rigidities R_GV[n].  One persistent kernel; lanes
 * refill from a device work queue.  steps_total (may be NULL) receives the
 * number of Boris steps taken. */
int amps_gpu_trace_batch(amps_ctx* ctx, const amps_trace_params* params, int32_t precision,
                         int64_t n, const double* x0, const double* dir, const double* R_GV,
                         amps_trace_result* out, int64_t* steps_total);

/* amps_gpu_trace_batch with the exit states (TraceTrajectorySharedEx,
 * CutoffRigidityGridless.cpp:1575-1585) of the allowed structured traces in
 * exits[n]. */
int amps_gpu_trace_batch_ex(amps_ctx* ctx, const amps_trace_params* params, int32_t precision,
                            int64_t n, const double* x0, const double* dir, const double* R_GV,
                            amps_trace_result* out, amps_trace_exit* exits, int64_t* steps_total);

/* Isotropic transmission of the density / spectrum solver
 * (ComputeWeightBlockAtEnergy / ComputeT_atEnergy, DensityGridless.cpp:
 * 880-925, through TraceDensityDirection :849-879): for every point x0[3p] and
 * rigidity R_GV[e], all ndir arrival directions dirs[3 ndir] are traced (one
 * batched launch) and, per (p, e), sampled, resolved (allowed or physically
 * forbidden), retried, weight_sum (allowed count) and the termination counts
 * term_counts[9 (p nR + e) + t] are reported; T = weight_sum / resolved.  Any
 * output array may be NULL. */
int amps_gpu_transmission(amps_ctx* ctx, const amps_trace_params* params, int32_t precision, int64_t npts,
                          const double* x0, int32_t nR, const double* R_GV, int32_t ndir, const double* dirs,
                          double* weight_sum, int32_t* sampled, int32_t* resolved, int32_t* retried,
                          int32_t* term_counts, int64_t* steps_total);

#define AMPS_SCAN_FULL 0   /* trace every bin of the bracket */
#define AMPS_SCAN_UPPER 1  /* downward from the bracket top to the first forbidden bin */
/* every bin of the bracket; a (location, direction) pair whose bracket holds
 * no transition has its other bins traced by the same launch, as soon as its
 * bracket is done, and its cutoff read off the whole grid (no reference
 * counterpart: the surrogate's bracket with the full-scan fallback) */
#define AMPS_SCAN_BRACKETED 2
#define AMPS_BIN_ABOVE (-2) /* bracket top forbidden below Rmax: cutoff above the bracket */
#define AMPS_BIN_BELOW (-3) /* bracket allowed down to a bottom above Rmin */
/* Cutoff map (CutoffForDirectionUpperScan_GV, CutoffRigidityGridless.cpp:
 * 2805-2834, on the grid of BuildCutoffSearchGrid_GV :2578-2617): for each of
 * nloc (location, direction) pairs, classify the rigidities grid[nR] inside
 * [bin_lo, bin_hi] (NULL: all bins).  states[nloc*nR]: -1 untraced,
 * 0 forbidden (physical or trace limit), 1 allowed, 2 numerical failure.
 * cutoff_bin[nloc]: index of the allowed side of the highest forbidden ->
 * allowed transition, 0 when all allowed down to Rmin, -1 when Rmax is
 * forbidden, or AMPS_BIN_ABOVE / AMPS_BIN_BELOW when the bracket does not
 * contain the transition (never in AMPS_SCAN_BRACKETED mode, which rescans
 * such pairs; their states outside the bracket are then traced, not -1). */
int amps_gpu_cutoff_scan(amps_ctx* ctx, const amps_trace_params* params, int32_t precision,
                         int32_t mode, int64_t nloc, const double* x0, const double* dir,
                         int32_t nR, const double* grid, const int32_t* bin_lo,
                         const int32_t* bin_hi, int8_t* states, int32_t* cutoff_bin,
                         int64_t* steps_total);
/* RefineCutoffTransition (srcEarth/util/CutoffBandSearch.h:212-242) for n
 * (location, direction) brackets at once, each [R_forbid, R_allow] with the
 * forbidden state at R_forbid: bisection to max(abs_tol, rel_tol max(|lo|,
 * |hi|)), one batched trace launch per round.  cutoff receives the allowed
 * end (-1 for an empty bracket), iterations the midpoints traced, status 0
 * resolved, 1 a trace limit ended it (trace_limit_forbidden = 0), 2 a
 * numerical failure after the retry ended it (the reference throws). */
int amps_gpu_refine_cutoff(amps_ctx* ctx, const amps_trace_params* params, int32_t precision,
                           int64_t n, const double* x0, const double* dir, const double* R_forbid,
                           const double* R_allow, double abs_tol, double rel_tol,
                           int32_t trace_limit_forbidden, double* cutoff, int32_t* iterations,
                           int32_t* status, int64_t* steps_total);
/* AnalyzeCutoffBandSamples (srcEarth/util/CutoffBandSearch.h:96-195) for n
 * state vectors of nR samples (host code): out[15 i ..] = lowerForbiddenIndex,
 * lowerAllowedIndex, upperForbiddenIndex, upperAllowedIndex, nTransitions,
 * nAllowedIntervals, nUnresolved, allAllowed, allForbidden, lowerBelowRange,
 * lowerAboveRange, upperBelowRange, upperAboveRange, lowerBracketUnresolved,
 * upperBracketUnresolved. */
int amps_gpu_band_topology(int64_t n, int32_t nR, const int8_t* states, int32_t* out);
/* IntegrateEffectiveCutoff (srcEarth/util/CutoffBandSearch.h:284-366), the
 * PENUMBRA_SCAN effective cutoff, for n (location, direction) pairs: every
 * coarse interval of grid[nR] inside [lower, upper] whose end states differ
 * is bisected (all pairs' bisections advance together, one batched trace
 * launch per round) and the allowed width of the penumbra summed.  states
 * [n*nR]: the scan's sample states (0 physically forbidden, 1 allowed,
 * 2 unresolved, as CutoffSampleState; -1 untraced counts as unresolved).
 * effective receives upper - allowed width clamped to [lower, upper] (-1
 * when not integrated), allowed_width and refined (midpoints traced) what the
 * reference accumulates, status 0 integrated, 1 unresolved, 2 a numerical
 * failure after the retry (the reference throws). */
int amps_gpu_effective_cutoff(amps_ctx* ctx, const amps_trace_params* params, int32_t precision,
                              int64_t n, const double* x0, const double* dir, int32_t nR,
                              const double* grid, const int8_t* states, const double* lower,
                              const double* upper, double abs_tol, double rel_tol,
                              int32_t trace_limit_forbidden, double* effective, double* allowed_width,
                              int32_t* refined, int32_t* status, int64_t* steps_total);
/* BuildUniformSphereDirs (srcEarth/gridless/CutoffRigidityGridless.cpp:
 * 1652-1674): n > 0 near-uniform unit vectors dirs[3k..3k+2] on the sphere
 * (z uniform in [-1, 1], azimuth advanced by the golden angle).  Host code. */
int amps_gpu_uniform_sphere_dirs(int32_t n, double* dirs);
/* BuildLocalENU_GSM + LocalLonLatToDir_GSM (CutoffRigidityGridless.cpp:
 * 1757-1800) for npts locations x0[3p] and ndir sky directions (lon_deg[k]
 * azimuth from local north toward east, lat_deg[k] elevation toward up):
 * dirs[3 (p ndir + k) + d]; frames[9p..9p+8] = E, N, U when non-NULL.  Host
 * code. */
int amps_gpu_local_dirs(int64_t npts, const double* x0, int32_t ndir, const double* lon_deg,
                        const double* lat_deg, double* dirs, double* frames);
/* The same for n (location, sky direction) pairs: dirs[3p..3p+2] for x0[3p],
 * lon_deg[p], lat_deg[p].  Host code. */
int amps_gpu_local_dirs_paired(int64_t n, const double* x0, const double* lon_deg, const double* lat_deg,
                               double* dirs);
/* Tracer tuning (autotuner): "workgroups" (persistent grid, 0 = 8 per CU),
 * "chunk" (steps between work-queue refills). */
int amps_gpu_trace_set_param(amps_ctx* ctx, const char* name, int64_t value);

/* ---- learned cutoff surrogate (no reference counterpart) -------------- */
/* MLP 6 -> 128 -> 128 -> 1 (GELU) on BF16 MFMA with FP32 accumulation,
 * trained with Adam on the mean-squared error of log10 cutoff rigidity.
 * Parameters (17,537 floats): W1[128][6] b1[128] W2[128][128] b2[128]
 * w3[128] b3.  Inputs per sample: sin/cos latitude, sin/cos longitude,
 * zenith and azimuth (scaled by the caller). */
#define AMPS_MLP_NPARAM 17537
int amps_gpu_mlp_init(amps_ctx* ctx, const float* theta);
int amps_gpu_mlp_get_params(amps_ctx* ctx, float* theta, int64_t* adam_steps);
/* Resident training set: X[n][6], targets T[n]. */
int amps_gpu_mlp_set_data(amps_ctx* ctx, int64_t n, const float* X, const float* T);
/* Gradient of the MSE over the first n (multiple of 64) resident rows. */
int amps_gpu_mlp_grad(amps_ctx* ctx, int64_t n, float* grad, float* loss);
/* nsteps Adam steps (batch a multiple of 64) over consecutive batches,
 * replayed from one captured hipGraph. */
int amps_gpu_mlp_train(amps_ctx* ctx, int64_t nsteps, int32_t batch, double lr, float* loss_out);
int amps_gpu_mlp_predict(amps_ctx* ctx, int64_t n, const float* X, float* y);

/* ---- profiling ------------------------------------------------------- */
/* Per-kernel HIP-event timing on the library stream (ms, accumulated). */
#define AMPS_K_PUSH 0
#define AMPS_K_COMPACT 1
#define AMPS_K_DEPOSIT_CELL 2
#define AMPS_K_DEPOSIT_NODE 3
#define AMPS_K_TRACE 4
#define AMPS_K_MLP 5
#define AMPS_K_ECSIM 6
#define AMPS_K_COUNT 7
int amps_gpu_set_timing(amps_ctx* ctx, int32_t enable);
int amps_gpu_get_timing(amps_ctx* ctx, double* ms /* AMPS_K_COUNT */, int64_t* launches);
int amps_gpu_reset_timing(amps_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
