/*
 * amps_oracle.h — CPU FP64 oracle for the AMPS particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the algorithms
 * of the reference's particle mover, ECSIM deposition, AMR point location and
 * gridless cutoff tracer so that the GPU library can be checked against it.
 * Only the test suite, the smoke entry point and the CPU-baseline leg of the
 * benchmark may load it; the product path never does.
 *
 * Every function cites the reference file:line it follows.  Units are the
 * reference's: normalized PIC units (c = 1) for the mover/deposit, SI for the
 * gridless tracer.  Where the reference has vectorised (AVX) and scalar
 * branches of one computation, the scalar branch is restated; the AVX
 * branches are parity unpinned.
 */
#ifndef AMPS_ORACLE_H
#define AMPS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NC 8      /* cells per block edge (_BLOCK_CELLS_X_) */
#define ORC_NN 9      /* corner nodes per block edge */
#define ORC_NCORNER (ORC_NN * ORC_NN * ORC_NN)
#define ORC_NCELL (ORC_NC * ORC_NC * ORC_NC)

/* Mover return codes (pic.h:5955-5960) plus the two classifications the
 * reference folds into _PARTICLE_LEFT_THE_DOMAIN_. */
#define ORC_CODE_REFLECTED (-1)   /* _PARTICLE_REJECTED_ON_THE_FACE_, committed */
#define ORC_CODE_LEFT_DOMAIN 2    /* deleted on an outer face */
#define ORC_CODE_FINISHED 3       /* _PARTICLE_MOTION_FINISHED_ */
#define ORC_CODE_SPHERE 5         /* absorbed by the internal sphere */
#define ORC_CODE_EXIT_FACE0 8     /* USER_FUNCTION mode: left through face (code - 8), state on the face */

#define ORC_MAX_SPEC 8
#define ORC_BC_DELETE 0
#define ORC_BC_REFLECT 1
#define ORC_BC_USER 2             /* ..._USER_FUNCTION_: ProcessOutsideDomainParticles gets the face state */

#define ORC_MOVER_LAPENTA 0   /* PIC::Mover::Lapenta2017 (alpha-matrix form) */
#define ORC_MOVER_BORIS 1     /* PIC::Mover::Boris (h/s rotation form) */
#define ORC_MOVER_RELATIVISTIC 2 /* PIC::Mover::Relativistic::Boris (gyro-limited sub-steps) */

#define ORC_SPEED_OF_LIGHT 299792458.0                  /* SpeedOfLight, constants.h:40 */
#define ORC_PI_TIMES_2 6.28318530717958647692528676655900576839433879875021 /* PiTimes2, constants.h:17 */

typedef struct orc_mesh orc_mesh;

orc_mesh* orc_mesh_create(const double gxmin[3], const double gxmax[3], int lmax,
                          int nblocks, const int* level, const int64_t* ijk);
void orc_mesh_destroy(orc_mesh* m);
double orc_mesh_eps(const orc_mesh* m);
/* block bounds (for tests) */
void orc_block_bounds(const orc_mesh* m, int b, double xmin[3], double xmax[3]);
/* point location: returns block id or -1 (findTreeNode(double*)) */
int orc_find_block(const orc_mesh* m, const double x[3], int start_block);
/* returns local cell id i+8(j+8k) or -1 (FindCellIndex) */
int orc_find_cell(const orc_mesh* m, const double x[3], int block);
/* trilinear corner stencil; x may be nudged (InitStencil). w_norm[8] in
 * stencil loop order, table[8] in the reference's table order, corner[8]
 * local corner index in stencil loop order. */
void orc_corner_stencil(const orc_mesh* m, double x[3], int block, double w_norm[8],
                        double table[8], int corner[8]);

/* One push of all particles; x,v,blk,cell updated in place, code written. */
int orc_push(const orc_mesh* m, const double* E_corner, const double* B_corner,
             const double* charge, const double* mass, const double* dt_spec,
             int mover, int bc_mode, double r_sphere, int64_t n, double* x, double* v,
             const int* spec, int* blk, int* cell, int* code, int nthreads);

/* Unique corner nodes: fills node_of_corner[nb*729], returns node count and,
 * if keys!=NULL, the lattice key of each node (sorted ascending).  Corners of
 * blocks that share a position on the finest lattice are one node, whatever
 * the blocks' levels; a hanging node (a fine corner on a coarse face) is a
 * node of its own.  The reference's deposit runs on one level, so this rule
 * at coarse/fine interfaces is decided here: parity unpinned there (the
 * count is checked against the distinct corner positions). */
int64_t orc_build_nodes(const orc_mesh* m, int* node_of_corner, int64_t* keys);

/* Particle energy and per-species CFL of UpdateJMassMatrix (see the .c) */
int orc_energy_cfl(const orc_mesh* m, const double* mass, const double* weight, int nspec, double dt,
                   int64_t n, const double* v, const int* spec, const int* blk, const int* cell,
                   double* energy, double* cfl, double* cell_energy, double* cell_cfl);

/* ECSIM J + mass matrix.  Particles may be in any order; within a cell they
 * are processed in array order.  J[nnodes*3], M[nnodes*243] are overwritten.
 * Pinned by the single-particle closed forms, symmetry and linearity; parity
 * unpinned at coarse/fine block interfaces, where the node rule of
 * orc_build_nodes applies. */
int orc_deposit(const orc_mesh* m, const double* B_corner, const double* charge,
                const double* mass, const double* weight, double dt, int64_t n,
                const double* x, const double* v, const int* spec, const int* blk,
                const int* cell, const int* node_of_corner, int64_t nnodes, double* J,
                double* M, int nthreads);

/* The deposit in the reference's parallel arrangement (eight parity sets of
 * cells per level, each cell added straight into its corners); equal to
 * orc_deposit up to rounding; only the rows of nodes around occupied cells
 * are written.  The CPU baseline's deposit. */
int orc_deposit_colored(const orc_mesh* m, const double* B_corner, const double* charge, const double* mass,
                        const double* weight, double dt, int64_t n, const double* x, const double* v,
                        const int* spec, const int* blk, const int* cell, const int* node_of_corner,
                        int64_t nnodes, double* J, double* M, int nthreads);

/* ---------------- gridless cutoff tracer (SI units) ---------------- */
typedef struct {
  double q_C;            /* signed backtrace charge [C] */
  double qabs_C;         /* |physical charge| for rigidity conversion [C] */
  double m0_kg;
  double dt_trace_s;     /* DT_TRACE */
  double max_time_s;
  int64_t max_steps;
  double max_distance_m; /* <= 0: off */
  double dphi_max;       /* gyro-phase budget per step [rad] */
  int32_t adaptive;      /* adaptiveDt */
  int32_t dist_limiter;  /* 0.20 * d_near / v */
  int32_t dt_floor;      /* legacy max(1e-12, 1e-9*DT, 100 km / v) floor */
  int32_t outer_sphere;  /* 0: box, 1: sphere of radius r_outer_Re */
  int32_t retry;         /* single retry on numerical failure */
  int32_t mover;         /* MoverType: 0 BORIS, 1 HC4, 2 RK2, 3 RK4, 4 RK6 */
  double box_min_Re[3], box_max_Re[3];
  double r_outer_Re, r_inner_Re;
  double earth_radius_m; /* length of one Re for all geometry conversions */
  double moment_Am2;     /* dipole |m| */
  double m_hat[3];
  double B_ext_T[3];     /* uniform field added to the dipole */
  int32_t field_model;   /* 0 dipole + uniform, 1 Mode3D mesh (orc_set_trace_mesh) */
  int32_t dt_selector;   /* 0 SelectLegacyCutoffTraceDt, 1 its Mode3D form (metres) */
  int32_t policy;        /* 0 legacy cutoff policy, 1 StructuredAccurate (density) */
  int32_t pad_;
  double event_tol_m;    /* boundaryEventTolerance_m */
  int32_t trap;          /* trapDetection */
  int32_t trap_min_mirror, trap_min_bounce, pad3_;
  double trap_outer_margin_Re, trap_radial_tol_Re, trap_energy_tol, trap_deadband;
} orc_trace_params;

/* TrajectoryTrap::Detector state (util/TrajectoryTrapDetector.h:23-153) */
typedef struct {
  int initialized, trapped, last_sign, mirror, bounce, stable, have_prev, pad_;
  double p_ref, p_min, p_max, r_min, r_max, prev_r_min, prev_r_max, min_margin;
} orc_trap;
void orc_trap_init(orc_trap* t);
/* Detector::Update with the trace parameters' trap configuration; returns trapped */
int orc_trap_update(const orc_trace_params* P, orc_trap* t, const double x[3], const double p[3], const double B[3]);

typedef struct {
  int32_t termination;   /* TrajectoryTermination value */
  int32_t retry_count;
  int64_t steps;
  double time_s, distance_m;
  double x[3], p[3];
} orc_trace_result;

/* TrajectoryExitState (CutoffRigidityGridless.h:258-269) of an allowed
 * structured trace; valid 0 otherwise. */
typedef struct {
  double x_exit_m[3], v_exit_unit[3], cos_alpha;
  int32_t valid, pad_;
} orc_trace_exit;

void orc_dipole_B(const orc_trace_params* P, const double x[3], double B[3]);
void orc_boris_step(const orc_trace_params* P, double x[3], double p[3], double dt);
/* HC4Step (unchecked fourth-order composition), in place. */
void orc_hc4_step(const orc_trace_params* P, double x[3], double p[3], double dt);
/* StepParticleChecked with mover P->mover: 1 when the step completed, 0 when
 * it met the inner sphere of radius r_inner_m (x, p then hold what the
 * reference leaves in them). */
int orc_step_checked(const orc_trace_params* P, double x[3], double p[3], double dt, double r_inner_m);
void orc_trace(const orc_trace_params* P, const double x0[3], const double dir_unit[3],
               double R_GV, orc_trace_result* res);
void orc_trace_ex(const orc_trace_params* P, const double x0[3], const double dir_unit[3],
                  double R_GV, orc_trace_result* res, orc_trace_exit* ex);
void orc_trace_batch_ex(const orc_trace_params* P, int64_t n, const double* x0, const double* dir,
                        const double* R_GV, orc_trace_result* res, orc_trace_exit* ex, int nthreads);
double orc_finalize_step(double dt, double time_remaining);
void orc_find_first_event(const double a[3], const double z[3], const double bmin[3], const double bmax[3],
                          double rInner, double tol, int* type, double* fraction, double pos[3], int* face);
void orc_trace_batch(const orc_trace_params* P, int64_t n, const double* x0,
                     const double* dir, const double* R_GV, orc_trace_result* res,
                     int nthreads);
/* Rigidity search grid (BuildCutoffSearchGrid_GV, log spacing). */
void orc_rigidity_grid(double rmin, double rmax, int n, double* grid);
/* Upper-cutoff bin from allowed flags (1 = allowed), UPPER_SCAN semantics:
 * -1 when the highest rigidity is forbidden, else the index of the allowed
 * side of the highest forbidden->allowed transition (0 when all allowed). */
int orc_upper_cutoff_bin(const int* allowed, int n);
/* AnalyzeCutoffBandSamples: states 0 forbidden, 1 allowed, 2 unresolved;
 * out[15] = lowerForbidden, lowerAllowed, upperForbidden, upperAllowed,
 * nTransitions, nAllowedIntervals, nUnresolved, allAllowed, allForbidden,
 * lowerBelow, lowerAbove, upperBelow, upperAbove, lowerUnres, upperUnres. */
void orc_band_topology(const int* states, int n, int* out);

/* ---- cell-centred linear interpolation (CellCentered::Linear::InitStencil,
 * pic_interpolation_routines.cpp:224-1070) ---------------------------------
 * Cell-centred data per block with one ghost layer: Bc[block][10][10][10][3],
 * cell (i,j,k) in [-1,8]^3 at ((i+1) + 10((j+1) + 10(k+1))).  A ghost cell
 * exists when its centre lies inside the domain.  The stencil is returned as
 * (block, ghost-layer cell index, weight) in the reference's order. */
#define ORC_GN 10
#define ORC_CC_MAX 80
int orc_center_stencil(const orc_mesh* m, const double x[3], int block, int* st_block,
                       int* st_cell, double* st_w);
void orc_interp_center(const orc_mesh* m, const double* Bc, int64_t n, const double* x,
                       const int* block, double* B, int nthreads);
/* Centre-based B for the mover and the deposit (_PIC_FIELD_SOLVER_B_MODE_ =
 * centre): when set (non-NULL, the caller keeps the array alive), orc_push's
 * Lapenta/Boris movers and orc_deposit interpolate B with the cell-centred
 * stencil instead of the corner values. */
void orc_set_center_field(orc_mesh* m, const double* Bc);
/* Mode3D compact-array field: row stencil at x (entries: block, interior
 * cell i + 8(j + 8k), weight; returns the length) and the field at n points
 * x[3n] from Bm[(block*512 + cell)*3 + d] into B[3n] (0 where a row is
 * empty). */
int orc_row_stencil(const orc_mesh* m, const double x[3], int* st_block, int* st_cell, double* st_w);
int orc_mode3d_field(const orc_mesh* m, const double* Bm, int64_t n, const double* x, double* B);
/* mesh and field of the tracer's field_model 1 (not thread-safe to change
 * during a trace batch) */
void orc_set_trace_mesh(const orc_mesh* m, const double* Bm);
/* per-block neighbour data used by the stencil: min and max neighbour level,
 * and the neighbour block across each face (6), edge (12) and corner (8) */
void orc_block_neighbours(const orc_mesh* m, int block, int* minmax, int* face, int* edge,
                          int* corner);
/* SelectLegacyCutoffTraceDt (CutoffRigidityGridless.cpp:1086-1152) and its
 * Mode3D form (CutoffRigidityMode3D.cpp:881-936) at one state */
double orc_select_dt(const orc_trace_params* P, const double x[3], const double p[3], double dtTrace,
                     double timeRemaining);
double orc_select_dt3d(const orc_trace_params* P, const double x[3], const double p[3], double dtTrace,
                       double timeRemaining, const double boxmin[3], const double boxmax[3], double rInner);

#ifdef __cplusplus
}
#endif
#endif
