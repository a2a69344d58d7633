// trace_common.h — kernel-side view of a tracing launch.
#pragma once
#include <hip/hip_runtime.h>

#include "amps_common.h"

namespace amps {

enum { TRACE_MODE_LIST = 0, TRACE_MODE_FULL = 1, TRACE_MODE_UPPER = 2, TRACE_MODE_BRACKET = 3 };

struct TraceView {
  amps_trace_params P;
  // FP64 geometry in metres
  double box_min_m[3], box_max_m[3], r_inner_m, r_outer_m;
  // FP32 path: length unit len = Re [m], time unit len / c
  double len, rin_n, rout_n, maxdist_n, floor_abs_n, floor_rel_n, floor_dist_n;
  float f_cb, f_mhat[3], f_bext[3], f_dphi;
  // work
  int mode, chunk, nR, nbin_item;
  int64_t nitems;
  unsigned long long* queue;
  unsigned long long* steps_total;
  const double* x0;   // [n or nloc][3] metres
  const double* dir;  // arrival directions (unit)
  const double* R;    // LIST mode rigidities [GV]
  const double* grid; // scan rigidities [GV]
  const int* lo;      // per-location bracket (may be null: full range)
  const int* hi;
  amps_trace_result* out;
  amps_trace_exit* exits;  // LIST mode exit states (may be null)
  int8_t* states;     // [nloc][nR]: -1 untraced, 0 forbidden, 1 allowed, 2 failed
  int* cutoff;        // UPPER mode raw result; BRACKET mode cutoff bin
  // BRACKET mode: locations, per-location outstanding items of the current
  // phase and phase (0 bracket running, 1 resolved, 2 whole grid rescanned)
  int64_t nloc;
  int* pend;
  int* status;
  // mesh field (field_model 1): the block mesh and the compact interior-cell B
  MeshView mesh;
  const double* Bm;
};

}  // namespace amps
