// center_device.h — device form of the cell-centred linear stencil
// (CellCentered::Linear::InitStencil, pic_interpolation_routines.cpp:224-1070)
// evaluated straight into B.  Expression order follows the reference; the
// library is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>

#include "amps_common.h"
#include "pic_device.h"

namespace amps {

constexpr int CC_GN = NC + 2;                    // cells per axis with the ghost layer
constexpr int CC_CELLS = CC_GN * CC_GN * CC_GN;  // 1000

__device__ __forceinline__ int cc_gidx(int i, int j, int k) { return (i + 1) + CC_GN * ((j + 1) + CC_GN * (k + 1)); }

// a cell of block b's ghost-layered array exists when it is interior or its
// centre lies inside the domain
__device__ __forceinline__ bool cc_exists(const MeshView& m, int b, int i, int j, int k) {
  const int ijk[3] = {i, j, k};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (ijk[d] < -1 || ijk[d] > NC) return false;
    if (ijk[d] >= 0 && ijk[d] < NC) continue;
    const double xmn = m.blk_xmin[3 * b + d];
    const double dx = (m.blk_xmax[3 * b + d] - xmn) / NC;
    const double c = xmn + (ijk[d] + 0.5) * dx;
    if (!(c >= m.gxmin[d] && c < m.gxmax[d])) return false;
  }
  return true;
}

__device__ __forceinline__ void cc_add(const double* __restrict__ Bc, int b, int cell, double w, double B[3]) {
  const double* v = Bc + ((size_t)b * CC_CELLS + (size_t)cell) * 3;
#pragma unroll
  for (int d = 0; d < 3; d++) B[d] += w * v[d];
}

// The 8-point stencil of GetTriliniarInterpolationStencil (:820-907) or the
// one-cell Constant::InitStencil (:168-219) it falls back to.
struct CcTri {
  int blk;
  int cell[8];
  double w[8];
  bool valid[8];
};

__device__ __forceinline__ void cc_constant(const MeshView& m, const double x[3], int b, CcTri& T) {
  T.blk = b;
#pragma unroll
  for (int l = 0; l < 8; l++) T.valid[l] = false, T.w[l] = 0.0, T.cell[l] = 0;
  if (b < 0) return;
  const int c = find_cell(m, x, b);
  if (c < 0) return;
  T.valid[0] = true;
  T.w[0] = 1.0;
  T.cell[0] = cc_gidx(c % NC, (c / NC) % NC, c / (NC * NC));
}

__device__ __forceinline__ double cc_weight(int l, const double w[3]) {
  switch (l) {
    case 0: return (1.0 - w[0]) * (1.0 - w[1]) * (1.0 - w[2]);
    case 1: return w[0] * (1.0 - w[1]) * (1.0 - w[2]);
    case 2: return (1.0 - w[0]) * w[1] * (1.0 - w[2]);
    case 3: return w[0] * w[1] * (1.0 - w[2]);
    case 4: return (1.0 - w[0]) * (1.0 - w[1]) * w[2];
    case 5: return w[0] * (1.0 - w[1]) * w[2];
    case 6: return (1.0 - w[0]) * w[1] * w[2];
    default: return w[0] * w[1] * w[2];
  }
}

// slots in the reference's loop order (i outer, k inner); the renormalisation
// test reads the stencil being built (the reference reads a shared table there)
__device__ __forceinline__ void cc_tri(const MeshView& m, double iLoc, double jLoc, double kLoc, const double x[3],
                                       int b, CcTri& T) {
  const int i0 = (iLoc < 0.5) ? -1 : (int)(iLoc - 0.50);
  const int j0 = (jLoc < 0.5) ? -1 : (int)(jLoc - 0.50);
  const int k0 = (kLoc < 0.5) ? -1 : (int)(kLoc - 0.50);
  double w[3];
  w[0] = iLoc - (i0 + 0.5);
  w[1] = jLoc - (j0 + 0.5);
  w[2] = kLoc - (k0 + 0.5);
  T.blk = b;
  int n = 0;
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const int i = s >> 2, j = (s >> 1) & 1, k = s & 1;
    T.cell[s] = cc_gidx(i0 + i, j0 + j, k0 + k);
    T.w[s] = cc_weight(i + 2 * j + 4 * k, w);
    T.valid[s] = cc_exists(m, b, i0 + i, j0 + j, k0 + k);
    n += T.valid[s] ? 1 : 0;
  }
  if (n == 0) {
    cc_constant(m, x, b, T);
    return;
  }
  if (n != 8) {
    double norm = 0.0;
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (T.valid[s]) norm += T.w[s];
    if (norm > 0.0) {
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (T.valid[s]) T.w[s] /= norm;
    }
  }
}

__device__ __forceinline__ void cc_accum_tri(const double* __restrict__ Bc, const CcTri& T, double B[3]) {
#pragma unroll
  for (int s = 0; s < 8; s++)
    if (T.valid[s]) cc_add(Bc, T.blk, T.cell[s], T.w[s], B);
}

// One coarse-stencil entry, blended with the fine stencil F when `blend`:
// weight * cw, plus the matching fine entry's weight * fw (cStencilGeneric::Add).
__device__ __forceinline__ void cc_emit(const double* __restrict__ Bc, int b, int cell, double w, bool blend,
                                        double cw, double fw, const CcTri& F, bool used[8], double B[3]) {
  double wt = w;
  if (blend) {
    wt = w * cw;
    if (b == F.blk) {
#pragma unroll
      for (int s = 0; s < 8; s++)
        if (F.valid[s] && !used[s] && F.cell[s] == cell) {
          wt += F.w[s] * fw;
          used[s] = true;
          break;
        }
    }
  }
  cc_add(Bc, b, cell, wt, B);
}

// GetTriliniarInterpolationMutiBlockStencil (:912-1070) on block b's cell
// frame, accumulated into B; with `blend` the fine stencil F joins it.
__device__ __forceinline__ void cc_multiblock(const MeshView& m, const double* __restrict__ Bc, const double x[3],
                                              int b, bool blend, double cw, double fw, const CcTri& F,
                                              double B[3]) {
  double dxCell[3], xmn[3], xLoc[3];
  int ijkMin[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xmn[d] = m.blk_xmin[3 * b + d];
    dxCell[d] = (m.blk_xmax[3 * b + d] - xmn[d]) / NC;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ijkMin[d] = (x[d] - xmn[d] < 0.5 * dxCell[d]) ? -1 : (int)((x[d] - xmn[d] - 0.5 * dxCell[d]) / dxCell[d]);
    const double xLower = xmn[d] + (ijkMin[d] + 0.5) * dxCell[d];
    xLoc[d] = (x[d] - xLower) / dxCell[d];
    if (xLoc[d] < 0.0) xLoc[d] = 0.0;
    if (xLoc[d] > 1.0) xLoc[d] = 1.0;
  }
  const int L = m.blk_level[b];
  // pass 1: the block of every logical support point and availability
  int S[8];
  bool ok = true;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int ijk[3] = {ijkMin[0] + (q >> 2), ijkMin[1] + ((q >> 1) & 1), ijkMin[2] + (q & 1)};
    double xl[3];
#pragma unroll
    for (int d = 0; d < 3; d++) xl[d] = xmn[d] + (ijk[d] + 0.5) * dxCell[d];
    S[q] = find_block(m, xl);
    if (S[q] < 0) {
      ok = false;
      continue;
    }
    if (m.blk_level[S[q]] == L) {
      if (!cc_exists(m, b, ijk[0], ijk[1], ijk[2])) ok = false;
    } else {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const int f = 2 * ((int)((xl[d] - m.blk_xmin[3 * S[q] + d]) / dxCell[d]));
        if (f < 0 || f + 1 > NC - 1) ok = false;  // fine cells f and f+1
      }
    }
  }
  bool used[8] = {false, false, false, false, false, false, false, false};
  if (!ok) {
    // the constant stencil of the block holding x
    CcTri K;
    cc_constant(m, x, find_block(m, x), K);
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (K.valid[s]) cc_emit(Bc, K.blk, K.cell[s], K.w[s], blend, cw, fw, F, used, B);
  } else {
    // pass 2: support points in the reference's order (di outer, dk inner)
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int di = q >> 2, dj = (q >> 1) & 1, dk = q & 1;
      const int ijk[3] = {ijkMin[0] + di, ijkMin[1] + dj, ijkMin[2] + dk};
      const double W = cc_weight(di + 2 * dj + 4 * dk, xLoc);
      if (m.blk_level[S[q]] == L) {
        cc_emit(Bc, b, cc_gidx(ijk[0], ijk[1], ijk[2]), W, blend, cw, fw, F, used, B);
      } else {
        // a logical coarse centre covered by a finer block: its 2x2x2 fine cells
        int iNeib[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          const double xl = xmn[d] + (ijk[d] + 0.5) * dxCell[d];
          iNeib[d] = 2 * ((int)((xl - m.blk_xmin[3 * S[q] + d]) / dxCell[d]));
        }
        const double FineCellWeight = (1.0 / 8.0) * W;
#pragma unroll
        for (int f = 0; f < 8; f++)
          cc_emit(Bc, S[q], cc_gidx(iNeib[0] + (f >> 2), iNeib[1] + ((f >> 1) & 1), iNeib[2] + (f & 1)),
                  FineCellWeight, blend, cw, fw, F, used, B);
      }
    }
  }
  if (blend) {
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (F.valid[s] && !used[s]) cc_add(Bc, F.blk, F.cell[s], F.w[s] * fw, B);
  }
}

// Edges: axis and the low/high side of the two other axes; corners c:
// x = c&1, y = (c>>1)&1, z = c>>2 (pic_interpolation_routines.cpp:421-613).
__constant__ static const signed char cc_edge_axis[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
__constant__ static const signed char cc_edge_side[12][3] = {{-1, 0, 0}, {-1, 0, 1}, {-1, 1, 1}, {-1, 1, 0},
                                                             {0, -1, 0}, {1, -1, 0}, {1, -1, 1}, {0, -1, 1},
                                                             {0, 0, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}};
__constant__ static const signed char cc_face_edges[6][4] = {{4, 11, 7, 8}, {5, 10, 6, 9}, {0, 9, 3, 8},
                                                             {1, 10, 2, 11}, {0, 5, 1, 4}, {3, 6, 2, 7}};
__constant__ static const signed char cc_face_nodes[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                                             {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};

// Branch of CellCentered::Linear::InitStencil (:224-816, AMPS branch of an
// AMR mesh) at x in block b: 0 tri-linear in b, 1 multi-block stencil of b,
// 2 multi-block stencil of the coarser block `coarser` (blended with b's
// tri-linear stencil when 0.5 < dmin <= 1).  xLoc receives (iLoc, jLoc, kLoc).
__device__ __forceinline__ int cc_branch(const MeshView& m, const double x[3], int b, double xLoc[3], int& coarser,
                                         double& dmin) {
  double lo[3], hi[3], dxCell[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    lo[d] = m.blk_xmin[3 * b + d];
    hi[d] = m.blk_xmax[3 * b + d];
    dxCell[d] = (hi[d] - lo[d]) / NC;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) xLoc[d] = (x[d] - lo[d]) / (hi[d] - lo[d]) * NC;
  const double N = NC;
  const int* info = m.nbinfo + 28 * b;
  const int L = m.blk_level[b], minL = info[0], maxL = info[1];
  const bool deep1 = (1.0 < xLoc[0]) && (xLoc[0] < N - 1) && (1.0 < xLoc[1]) && (xLoc[1] < N - 1) &&
                     (1.0 < xLoc[2]) && (xLoc[2] < N - 1);
  coarser = -1;
  dmin = 10.0 * NC * NC * NC;
  if ((L == minL && L == maxL) || deep1) return 0;
  if (L == minL) {
    if ((0.5 < xLoc[0]) && (xLoc[0] < N - 0.5) && (0.5 < xLoc[1]) && (xLoc[1] < N - 0.5) && (0.5 < xLoc[2]) &&
        (xLoc[2] < N - 0.5))
      return 0;
    return 1;
  }
  // a coarser block near x, through the faces, edges and corners it is near
  unsigned edgeTested = 0, cornerTested = 0;
  auto contains = [&](int nb) {
    int cnt = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
      if ((m.blk_xmin[3 * nb + d] - dxCell[d] <= x[d]) && (m.blk_xmax[3 * nb + d] + dxCell[d] >= x[d])) cnt++;
    return cnt == 3;
  };
  for (int idim = 0; idim < 3; idim++) {
    int iFace;
    if (xLoc[idim] <= 1.0) iFace = 2 * idim;
    else if (xLoc[idim] >= N - 1.0) iFace = 2 * idim + 1;
    else continue;
    int nb = info[2 + iFace];
    if (nb >= 0 && m.blk_level[nb] < L && contains(nb)) {
      if ((iFace & 1) == 0) {
        if (xLoc[idim] < 1.0 && xLoc[idim] < dmin) dmin = xLoc[idim], coarser = nb;
      } else {
        if (xLoc[idim] > N - 1 && N - xLoc[idim] < dmin) dmin = N - xLoc[idim], coarser = nb;
      }
    }
    for (int q = 0; q < 4; q++) {
      const int e = cc_face_edges[iFace][q];
      if (edgeTested & (1u << e)) continue;
      edgeTested |= 1u << e;
      nb = info[8 + e];
      if (nb < 0 || !(m.blk_level[nb] < L) || !contains(nb)) continue;
      bool ok = true;
      for (int d = 0; d < 3; d++) {
        if (d == cc_edge_axis[e]) continue;
        if (cc_edge_side[e][d] ? !(xLoc[d] > N - 1) : !(xLoc[d] < 1.0)) ok = false;
      }
      if (!ok) continue;
      for (int d = 0; d < 3; d++) {
        if (d == cc_edge_axis[e]) continue;
        const double dist = cc_edge_side[e][d] ? N - xLoc[d] : xLoc[d];
        if (dist < dmin) dmin = dist, coarser = nb;
      }
    }
    for (int q = 0; q < 4; q++) {
      const int cn = cc_face_nodes[iFace][q];
      if (cornerTested & (1u << cn)) continue;
      cornerTested |= 1u << cn;
      nb = info[20 + cn];
      if (nb < 0 || !(m.blk_level[nb] < L) || !contains(nb)) continue;
      bool ok = true;
      for (int d = 0; d < 3; d++)
        if (((cn >> d) & 1) ? !(xLoc[d] > N - 1) : !(xLoc[d] < 1.0)) ok = false;
      if (!ok) continue;
      for (int d = 0; d < 3; d++) {
        const double dist = ((cn >> d) & 1) ? N - xLoc[d] : xLoc[d];
        if (dist < dmin) dmin = dist, coarser = nb;
      }
    }
  }
  return coarser >= 0 ? 2 : 0;
}

// B at x in block b from the ghost-layered cell-centred field Bc.
static __device__ __noinline__ void center_B(const MeshView& m, const double* __restrict__ Bc, const double x[3], int b,
                                             double B[3]) {
  B[0] = B[1] = B[2] = 0.0;
  double xLoc[3], dmin;
  int coarser;
  const int br = cc_branch(m, x, b, xLoc, coarser, dmin);
  CcTri T;
  if (br == 0) {
    cc_tri(m, xLoc[0], xLoc[1], xLoc[2], x, b, T);
    cc_accum_tri(Bc, T, B);
    return;
  }
  if (br == 1) {
    cc_multiblock(m, Bc, x, b, false, 1.0, 0.0, T, B);
    return;
  }
  const bool blend = (0.5 < dmin) && (dmin <= 1.0);
  double cw = 1.0, fw = 0.0;
  if (blend) {
    cc_tri(m, xLoc[0], xLoc[1], xLoc[2], x, b, T);
    cw = 1.0 - (dmin - 0.5) / 0.5;
    fw = (dmin - 0.5) / 0.5;
  }
  cc_multiblock(m, Bc, x, coarser, blend, cw, fw, T, B);
}

// ---------------------------------------------------------------------
// Mode3D compact-array field.  The decomposition-independent row stencil
// (cRowStencil, pic.h:7295-7420) of the same InitStencil: every support
// cell, ghost indices included, is taken to the leaf holding its centre and
// the interior cell under it (AddRowStencilCell / ...ByCoordinate,
// pic_interpolation_routines.cpp:33-105), entries naming one cell merge, and
// the field is the weighted sum in row order over Bm[(block*512 + i +
// 8(j + 8k))*3 + d] (InterpolateField_, GlobalMagneticField.cpp:292-330).
// ---------------------------------------------------------------------
constexpr int CC_ROW_MAX = 72;  // 8 support points x 8 fine cells + 8 blended fine entries

template <int CAP>
struct RowList {
  int n;
  int key[CAP];  // block * 512 + interior cell
  double w[CAP];
};

// cRowStencil::AddCell (pic.h:7351-7375)
template <int CAP>
__device__ __forceinline__ void row_addkey(RowList<CAP>& r, double w, int key) {
  for (int e = 0; e < r.n; e++)
    if (r.key[e] == key) {
      r.w[e] += w;
      return;
    }
  if (r.n < CAP) {
    r.key[r.n] = key;
    r.w[r.n] = w;
    r.n++;
  }
}

// AddRowStencilCellByCoordinate (pic_interpolation_routines.cpp:33-82, non-periodic)
template <int CAP>
__device__ __forceinline__ bool row_add_coord(const MeshView& m, RowList<CAP>& r, double w, const double x[3]) {
#pragma unroll
  for (int d = 0; d < 3; d++)
    if ((x[d] < m.gxmin[d]) || (x[d] > m.gxmax[d])) return false;
  const int b = find_block(m, x);
  if (b < 0) return false;
  int ijk[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xmn = m.blk_xmin[3 * b + d];
    const double blockLength = m.blk_xmax[3 * b + d] - xmn;
    const double xl = (x[d] - xmn) / blockLength * NC;
    ijk[d] = (int)floor(xl);
    if ((ijk[d] < 0) && (xl > -1.0e-10)) ijk[d] = 0;
    if ((ijk[d] >= NC) && (xl < NC + 1.0e-10)) ijk[d] = NC - 1;
    if ((ijk[d] < 0) || (ijk[d] >= NC)) return false;
  }
  row_addkey(r, w, b * NCELL + ijk[0] + NC * (ijk[1] + NC * ijk[2]));
  return true;
}

// AddRowStencilCell (:91-105): indices of block s, ghosts included
template <int CAP>
__device__ __forceinline__ bool row_add_cell(const MeshView& m, RowList<CAP>& r, double w, int s, int i, int j,
                                             int k) {
  const int ijk[3] = {i, j, k};
  double x[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double dx = (m.blk_xmax[3 * s + d] - m.blk_xmin[3 * s + d]) / NC;
    x[d] = m.blk_xmin[3 * s + d] + (ijk[d] + 0.5) * dx;
  }
  return row_add_coord(m, r, w, x);
}

// GetTriliniarInterpolationStencil, row part (:820-896)
template <int CAP>
__device__ __forceinline__ void row_tri(const MeshView& m, const double loc[3], int b, RowList<CAP>& r) {
  r.n = 0;
  const int i0 = (loc[0] < 0.5) ? -1 : (int)(loc[0] - 0.50);
  const int j0 = (loc[1] < 0.5) ? -1 : (int)(loc[1] - 0.50);
  const int k0 = (loc[2] < 0.5) ? -1 : (int)(loc[2] - 0.50);
  double w[3];
  w[0] = loc[0] - (i0 + 0.5);
  w[1] = loc[1] - (j0 + 0.5);
  w[2] = loc[2] - (k0 + 0.5);
  for (int s = 0; s < 8; s++) {
    const int i = s >> 2, j = (s >> 1) & 1, k = s & 1;
    row_add_cell(m, r, cc_weight(i + 2 * j + 4 * k, w), b, i0 + i, j0 + j, k0 + k);
  }
  if ((r.n > 0) && (r.n != 8)) {
    double norm = 0.0;
    for (int e = 0; e < r.n; e++) norm += r.w[e];
    if (norm > 0.0)
      for (int e = 0; e < r.n; e++) r.w[e] /= norm;
  }
}

// GetTriliniarInterpolationMutiBlockStencil, row part (:912-1054)
__device__ __forceinline__ void row_multiblock(const MeshView& m, const double x[3], int b,
                                               RowList<CC_ROW_MAX>& r) {
  r.n = 0;
  double dxCell[3], xmn[3], xLoc[3];
  int ijkMin[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    xmn[d] = m.blk_xmin[3 * b + d];
    dxCell[d] = (m.blk_xmax[3 * b + d] - xmn[d]) / NC;
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    ijkMin[d] = (x[d] - xmn[d] < 0.5 * dxCell[d]) ? -1 : (int)((x[d] - xmn[d] - 0.5 * dxCell[d]) / dxCell[d]);
    const double xLower = xmn[d] + (ijkMin[d] + 0.5) * dxCell[d];
    xLoc[d] = (x[d] - xLower) / dxCell[d];
    if (xLoc[d] < 0.0) xLoc[d] = 0.0;
    if (xLoc[d] > 1.0) xLoc[d] = 1.0;
  }
  const int L = m.blk_level[b];
  bool geometry = true;
  for (int q = 0; q < 8; q++) {
    const int di = q >> 2, dj = (q >> 1) & 1, dk = q & 1;
    const int ijk[3] = {ijkMin[0] + di, ijkMin[1] + dj, ijkMin[2] + dk};
    double xl[3];
#pragma unroll
    for (int d = 0; d < 3; d++) xl[d] = xmn[d] + (ijk[d] + 0.5) * dxCell[d];
    const double W = cc_weight(di + 2 * dj + 4 * dk, xLoc);
    const int S = find_block(m, xl);
    if (S < 0) {
      geometry = false;
      continue;
    }
    if (m.blk_level[S] == L) {
      row_add_cell(m, r, W, b, ijk[0], ijk[1], ijk[2]);
    } else {
      // a logical coarse centre covered by a finer block: its 2x2x2 fine cells
      int iNeib[3];
#pragma unroll
      for (int d = 0; d < 3; d++) iNeib[d] = 2 * ((int)((xl[d] - m.blk_xmin[3 * S + d]) / dxCell[d]));
      const double FineCellWeight = (1.0 / 8.0) * W;
      for (int f = 0; f < 8; f++)
        row_add_cell(m, r, FineCellWeight, S, iNeib[0] + (f >> 2), iNeib[1] + ((f >> 1) & 1), iNeib[2] + (f & 1));
    }
  }
  if (!geometry) {
    // BuildConstantRowStencil (:109-115) at the leaf holding x
    r.n = 0;
    row_add_coord(m, r, 1.0, x);
  }
}

// cMode3DMeshFieldEval::GetB_T mesh path (CutoffRigidityMode3D.cpp:702-800):
// zero outside the mesh, else the row's weighted sum.  A stencil whose eight
// support cells are interior cells of the block holding x (the common case)
// is summed directly: each such cell's centre lies inside that block and maps
// back to itself, so the row is the same eight entries in the same order.
static __device__ __noinline__ void mode3d_B(const MeshView& m, const double* __restrict__ Bm, const double x[3],
                                             double B[3]) {
  B[0] = B[1] = B[2] = 0.0;
  const int b = find_block(m, x);
  if (b < 0) return;
  double loc[3], dmin;
  int coarser;
  const int br = cc_branch(m, x, b, loc, coarser, dmin);
  if (br == 0) {
    const int i0 = (loc[0] < 0.5) ? -1 : (int)(loc[0] - 0.50);
    const int j0 = (loc[1] < 0.5) ? -1 : (int)(loc[1] - 0.50);
    const int k0 = (loc[2] < 0.5) ? -1 : (int)(loc[2] - 0.50);
    if (i0 >= 0 && i0 <= NC - 2 && j0 >= 0 && j0 <= NC - 2 && k0 >= 0 && k0 <= NC - 2) {
      double w[3];
      w[0] = loc[0] - (i0 + 0.5);
      w[1] = loc[1] - (j0 + 0.5);
      w[2] = loc[2] - (k0 + 0.5);
      const double* base = Bm + (size_t)b * NCELL * 3;
#pragma unroll
      for (int s = 0; s < 8; s++) {
        const int i = s >> 2, j = (s >> 1) & 1, k = s & 1;
        const double W = cc_weight(i + 2 * j + 4 * k, w);
        const double* v = base + 3 * ((i0 + i) + NC * ((j0 + j) + NC * (k0 + k)));
        B[0] += W * v[0];
        B[1] += W * v[1];
        B[2] += W * v[2];
      }
      return;
    }
  }
  RowList<CC_ROW_MAX> r;
  if (br == 0) {
    row_tri(m, loc, b, r);
  } else if (br == 1) {
    row_multiblock(m, x, b, r);
  } else {
    row_multiblock(m, x, coarser, r);
    if ((0.5 < dmin) && (dmin <= 1.0)) {
      RowList<8> f;
      row_tri(m, loc, b, f);
      const double cw = 1.0 - (dmin - 0.5) / 0.5, fw = (dmin - 0.5) / 0.5;
      for (int e = 0; e < r.n; e++) r.w[e] *= cw;
      for (int e = 0; e < f.n; e++) f.w[e] *= fw;
      for (int e = 0; e < f.n; e++) row_addkey(r, f.w[e], f.key[e]);
    }
  }
  for (int e = 0; e < r.n; e++) {
    const double* v = Bm + (size_t)r.key[e] * 3;
    B[0] += r.w[e] * v[0];
    B[1] += r.w[e] * v[1];
    B[2] += r.w[e] * v[2];
  }
}

}  // namespace amps
