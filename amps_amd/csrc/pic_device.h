// pic_device.h — per-particle device routines of the FP64 push and deposit.
// Each routine cites the reference lines it follows; the expression order
// is the reference's (the library is compiled with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include "amps_common.h"

namespace amps {

struct BlockGeom {
  double xmin[3], xmax[3], dxs[3], dxc[3];
};

__device__ __forceinline__ void load_block_geom(const MeshView& m, int b, BlockGeom& g) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    g.xmin[d] = m.blk_xmin[3 * b + d];
    g.xmax[d] = m.blk_xmax[3 * b + d];
    g.dxs[d] = m.blk_dxs[3 * b + d];
    g.dxc[d] = m.blk_dxc[3 * b + d];
  }
}

__device__ __forceinline__ int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }

// Workgroups are dealt to the 8 XCDs round-robin (XCD = blockIdx.x % 8).
// This maps them so that each XCD works through one contiguous range of
// logical tiles, whose shared data then stays in that XCD's L2.  gridDim.x
// must be a multiple of 8.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_tile() {
  return (int)((blockIdx.x % N_XCD) * (gridDim.x / N_XCD) + blockIdx.x / N_XCD);
}

// CornerBased::InitStencil — pic_interpolation_routines.cpp:1074-1194.
// wn[l], l = 4i+2j+k in the reference's stencil loop order, normalized as
// cStencil::Normalize (pic.h:7255-7262); table[] in the reference's
// InterpolationCoefficientTable order 0,1,3,2,4,5,7,6.  x may be nudged off
// the block's upper face exactly as the reference does.
__device__ __forceinline__ void corner_stencil(const BlockGeom& g, double x[3], double wn[8],
                                               double table[8], int iX[3]) {
  double xLoc[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (fabs(x[d] - g.xmax[d]) < 1e-10 * g.dxs[d]) x[d] = g.xmax[d] - 1e-10 * g.dxs[d];
    xLoc[d] = (x[d] - g.xmin[d]) / g.dxs[d];
    iX[d] = (int)(xLoc[d]);
    xLoc[d] -= iX[d];
  }
  constexpr int tslot[8] = {0, 4, 3, 7, 1, 5, 2, 6};
  double norm = 0.0;
#pragma unroll
  for (int l = 0; l < 8; l++) {
    const int i = l >> 2, j = (l >> 1) & 1, k = l & 1;
    const double a = i ? xLoc[0] : (1.0 - xLoc[0]);
    const double b = j ? xLoc[1] : (1.0 - xLoc[1]);
    const double c = k ? xLoc[2] : (1.0 - xLoc[2]);
    const double w = a * b * c;
    table[tslot[l]] = w;
    wn[l] = w;
    norm += w;
  }
  if (norm > 0.0) {
#pragma unroll
    for (int l = 0; l < 8; l++) wn[l] /= norm;
  }
}

// findTreeNode(int*) as a lookup: the descent of meshAMRgeneric.h:2793-2848
// returns the leaf containing the lattice point, or NULL outside the root.
__device__ __forceinline__ int lookup_leaf(const MeshView& m, const int ix[3]) {
  const int lim = 1 << m.lmax;
#pragma unroll
  for (int d = 0; d < 3; d++)
    if (ix[d] < 0 || ix[d] >= lim) return -1;
  const int sh = m.lmax - m.lf;
  return m.leaf_of[((ix[2] >> sh) * m.nf + (ix[1] >> sh)) * m.nf + (ix[0] >> sh)];
}

// findTreeNode(double*) — meshAMRgeneric.h:2852-2882, including the
// half-open correction against the found block's bounds.
__device__ __forceinline__ int find_block(const MeshView& m, const double x[3]) {
  int ix[3];
#pragma unroll
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((x[d] - m.gxmin[d]) / m.dx_maxref[d]);
  int res = lookup_leaf(m, ix);
  if (res >= 0) {
    bool flag = false;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (x[d] < m.blk_xmin[3 * res + d]) ix[d]--, flag = true;
      if (x[d] >= m.blk_xmax[3 * res + d]) ix[d]++, flag = true;
    }
    if (flag) res = lookup_leaf(m, ix);
  }
  return res;
}

// The tile's block on the level-lmax lattice: find_block_in() returns it
// without reading leaf_of when the lattice point lies in its range (the
// lookup would return it), and find_cell_in() takes its bounds from the
// block geometry already in registers (the same values find_cell reads).
struct BlockLat {
  int b, lo[3], hi[3];
};

__device__ __forceinline__ void load_block_lat(const MeshView& m, int b, BlockLat& l) {
  l.b = b;
  const int ext = m.blk_lat[4 * b + 3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    l.lo[d] = m.blk_lat[4 * b + d];
    l.hi[d] = l.lo[d] + ext;
  }
}

__device__ __forceinline__ int find_block_in(const MeshView& m, const double x[3], const BlockGeom& g,
                                             const BlockLat& l) {
  int ix[3];
#pragma unroll
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((x[d] - m.gxmin[d]) / m.dx_maxref[d]);
  bool in = true;
#pragma unroll
  for (int d = 0; d < 3; d++) in = in && ix[d] >= l.lo[d] && ix[d] < l.hi[d];
  if (!in) return find_block(m, x);
  bool flag = false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (x[d] < g.xmin[d]) ix[d]--, flag = true;
    if (x[d] >= g.xmax[d]) ix[d]++, flag = true;
  }
  return flag ? lookup_leaf(m, ix) : l.b;
}

// FindCellIndex — meshAMRgeneric.h:2256-2323 (ExitFlag=false).
__device__ __forceinline__ int find_cell(const MeshView& m, const double x[3], int b) {
  int ijk[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xmn = m.blk_xmin[3 * b + d], xmx = m.blk_xmax[3 * b + d];
    if ((x[d] < xmn) || (xmx < x[d])) return -1;
    int i = (int)((x[d] - xmn) / m.blk_dxc[3 * b + d]);
    if (i == NC) i = NC - 1;
    ijk[d] = clampi(i, 0, NC - 1);
  }
  return ijk[0] + NC * (ijk[1] + NC * ijk[2]);
}

__device__ __forceinline__ int find_cell_in(const MeshView& m, const double x[3], int b, const BlockGeom& g,
                                            const BlockLat& l) {
  if (b != l.b) return find_cell(m, x, b);
  int ijk[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double xmn = g.xmin[d], xmx = g.xmax[d];
    if ((x[d] < xmn) || (xmx < x[d])) return -1;
    int i = (int)((x[d] - xmn) / g.dxc[d]);
    if (i == NC) i = NC - 1;
    ijk[d] = clampi(i, 0, NC - 1);
  }
  return ijk[0] + NC * (ijk[1] + NC * ijk[2]);
}

// Sphere, domain-boundary and block/cell relocation of a pushed particle
// (pic_mover_boris.cpp:1088-1375), shared by the FP64 and FP32 pushes.  On
// entry xInit/vInit hold the start state, xFinal/vFinal the pushed state;
// xp/vp receive the committed state.
__device__ __forceinline__ int relocate(const MeshView& m, const Faces& faces, const BlockGeom& g,
                                        const BlockLat& lat, double dtTotal,
                                        int bc_mode, double r_sphere, double xInit[3],
                                        double vInit[3], double xFinal[3], double vFinal[3],
                                        double xp[3], double vp[3], int* blk, int* cell) {
  // internal sphere (:1088-1117): absorbed
  double rFinal2;
  if (r_sphere > 0.0 &&
      (rFinal2 = xFinal[0] * xFinal[0] + xFinal[1] * xFinal[1] + xFinal[2] * xFinal[2]) <
          r_sphere * r_sphere) {
    const double r = sqrt(rFinal2);
#pragma unroll
    for (int d = 0; d < 3; d++) xFinal[d] *= r_sphere / r;
#pragma unroll
    for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
    return AMPS_CODE_SPHERE;
  }
  int newNode = find_block_in(m, xFinal, g, lat);
  int code = AMPS_CODE_FINISHED;
  if (newNode < 0) {
    if (bc_mode == AMPS_BC_DELETE) {
#pragma unroll
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return AMPS_CODE_LEFT_DOMAIN;
    }
    // exit face (:1156-1229); the face index starts at -1 here
    int nIntersectionFace = -1;
    const double vMiddle[3] = {0.5 * (vInit[0] + vFinal[0]), 0.5 * (vInit[1] + vFinal[1]),
                               0.5 * (vInit[2] + vFinal[2])};
    double dtIntersection = -1.0, r0[3];
    for (int nface = 0; nface < 6; nface++) {
      const Face& F = faces.f[nface];
      double cx = 0.0, cv = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) {
        r0[d] = xInit[d] - F.x0[d];
        cx += r0[d] * F.norm[d];
        cv += vMiddle[d] * F.norm[d];
      }
      if (cv > 0.0) {
        const double dt = -cx / cv;
        if ((dtIntersection < 0.0) || ((dt < dtIntersection) && (dt > 0.0))) {
          double cE0 = 0.0, cE1 = 0.0;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double c = r0[d] + dt * vMiddle[d];
            cE0 += c * F.e0[d], cE1 += c * F.e1[d];
          }
          if ((cE0 < -m.eps) || (cE0 > F.lE0 + m.eps) || (cE1 < -m.eps) || (cE1 > F.lE1 + m.eps))
            continue;
          nIntersectionFace = nface, dtIntersection = dt;
        }
      }
    }
    if (nIntersectionFace == -1) {
#pragma unroll
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return AMPS_CODE_LEFT_DOMAIN;
    }
    const Face& F = faces.f[nIntersectionFace];
    const double tVelocityIncrement =
        ((dtIntersection / dtTotal < 1) ? dtIntersection / dtTotal : 1);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      xInit[d] += dtIntersection * vMiddle[d] - F.norm[d] * m.eps;
      vInit[d] += tVelocityIncrement * (vFinal[d] - vInit[d]);
    }
    newNode = find_block(m, xInit);
    if (newNode < 0) {
#pragma unroll
      for (int d = 0; d < 3; d++) {
        if (m.gxmin[d] >= xInit[d]) xInit[d] = m.gxmin[d] + m.eps;
        if (m.gxmax[d] <= xInit[d]) xInit[d] = m.gxmax[d] - m.eps;
      }
      newNode = find_block(m, xInit);
      if (newNode < 0 && bc_mode != AMPS_BC_USER) {
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
        return AMPS_CODE_LEFT_DOMAIN;
      }
    }
    if (bc_mode == AMPS_BC_USER) {
      // ProcessOutsideDomainParticles(ptr, xInit, vInit, nIntersectionFace,
      // newNode) (:1231): the state moved onto the face is what the callback
      // receives; the particle is then deleted (:1262-1265)
#pragma unroll
      for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
      *blk = newNode;
      *cell = 0;
      return AMPS_CODE_EXIT_FACE0 + nIntersectionFace;
    }
    // specular reflection (:1236-1253), committed
    double cc = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) cc += F.norm[d] * vInit[d];
#pragma unroll
    for (int d = 0; d < 3; d++) vInit[d] -= 2.0 * cc * F.norm[d];
#pragma unroll
    for (int d = 0; d < 3; d++) vFinal[d] = vInit[d], xFinal[d] = xInit[d];
    code = AMPS_CODE_REFLECTED;
  }
  const int c_id = find_cell_in(m, xFinal, newNode, g, lat);
  if (c_id < 0) return AMPS_CODE_LEFT_DOMAIN;
  *blk = newNode;
  *cell = c_id;
#pragma unroll
  for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
  return code;
}

// cell-centred linear stencil (center_device.h)
static __device__ __noinline__ void center_B(const MeshView& m, const double* __restrict__ Bc, const double x[3],
                                             int b, double B[3]);

// One particle of PIC::Mover::Lapenta2017 (pic_mover_boris.cpp:876-1380) or
// PIC::Mover::Boris (:126-553, rotation :253-277) with corner E and B from
// LDS, or with BC the cell-centred B of Bc (the centre-based B mode,
// :979-1000).  Returns the AMPS_CODE_*; xp/vp hold the committed state.
template <int MOVER, bool BC = false>
__device__ __forceinline__ int push_particle(const MeshView& m, const Faces& faces,
                                             const BlockGeom& g, const BlockLat& lat, const double* sE,
                                             const double* sB, double q, double mass,
                                             double dtTotal, int bc_mode, double r_sphere,
                                             double xp[3], double vp[3], int* blk, int* cell,
                                             const double* Bc = nullptr) {
  double vInit[3] = {vp[0], vp[1], vp[2]}, xInit[3] = {xp[0], xp[1], xp[2]};
  double vFinal[3], xFinal[3];
  double wn[8], table[8];
  int iX[3];
  corner_stencil(g, xInit, wn, table, iX);
#pragma unroll
  for (int d = 0; d < 3; d++) iX[d] = clampi(iX[d], 0, NC - 1);
  double E[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int l = 0; l < 8; l++) {
    const int c = (iX[0] + (l >> 2)) + NN * ((iX[1] + ((l >> 1) & 1)) + NN * (iX[2] + (l & 1)));
#pragma unroll
    for (int d = 0; d < 3; d++) {
      E[d] += wn[l] * sE[3 * c + d];
      if constexpr (!BC) B[d] += wn[l] * sB[3 * c + d];
    }
  }
  if constexpr (BC) center_B(m, Bc, xInit, *blk, B);
  if (MOVER == AMPS_MOVER_LAPENTA) {
    // :1033-1081
    const double QdT_over_m = q * dtTotal / mass;
    const double QdT_over_2m = 0.5 * QdT_over_m;
    const double QdT_over_2m_squared = QdT_over_2m * QdT_over_2m;
    double BB[3][3], P[3], alpha[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      P[i] = -QdT_over_2m * B[i];
#pragma unroll
      for (int j = 0; j <= i; j++) {
        BB[i][j] = QdT_over_2m_squared * B[i] * B[j];
        BB[j][i] = BB[i][j];
      }
    }
    const double c0 =
        1.0 / (1.0 + QdT_over_2m_squared * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]));
    alpha[0][0] = c0 * (1.0 + BB[0][0]);
    alpha[0][1] = c0 * (-P[2] + BB[0][1]);
    alpha[0][2] = c0 * (P[1] + BB[0][2]);
    alpha[1][0] = c0 * (P[2] + BB[1][0]);
    alpha[1][1] = c0 * (1.0 + BB[1][1]);
    alpha[1][2] = c0 * (-P[0] + BB[1][2]);
    alpha[2][0] = c0 * (-P[1] + BB[2][0]);
    alpha[2][1] = c0 * (P[0] + BB[2][1]);
    alpha[2][2] = c0 * (1.0 + BB[2][2]);
#pragma unroll
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 3; j++) s += alpha[i][j] * (vInit[j] + QdT_over_2m * E[j]);
      vFinal[i] = 2.0 * s - vInit[i];
    }
#pragma unroll
    for (int i = 0; i < 3; i++) xFinal[i] = xInit[i] + dtTotal * vFinal[i];
  } else {
    // BorisSplitAcceleration_default :87-98, rotation :253-277
    const double Charge2Mass = q / mass;
    double accl[3], rot[3], u[3], h[3], U[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      accl[d] = 0.0 + Charge2Mass * E[d];
      rot[d] = 0.0 - Charge2Mass * B[d];
    }
    const double dtTemp = dtTotal, dtTempOverTwo = dtTotal / 2.0;
#pragma unroll
    for (int d = 0; d < 3; d++) u[d] = vInit[d] + dtTempOverTwo * accl[d];
#pragma unroll
    for (int d = 0; d < 3; d++) h[d] = -dtTempOverTwo * rot[d];
    const double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
    const double uh = u[0] * h[0] + u[1] * h[1] + u[2] * h[2];
    U[0] = ((1 - h2) * u[0] + 2 * (u[1] * h[2] - h[1] * u[2] + uh * h[0])) / (1 + h2);
    U[1] = ((1 - h2) * u[1] + 2 * (u[2] * h[0] - h[2] * u[0] + uh * h[1])) / (1 + h2);
    U[2] = ((1 - h2) * u[2] + 2 * (u[0] * h[1] - h[0] * u[1] + uh * h[2])) / (1 + h2);
#pragma unroll
    for (int d = 0; d < 3; d++) vFinal[d] = U[d] + dtTempOverTwo * accl[d];
#pragma unroll
    for (int d = 0; d < 3; d++) xFinal[d] = xInit[d] + dtTemp * vFinal[d];
  }

  return relocate(m, faces, g, lat, dtTotal, bc_mode, r_sphere, xInit, vInit, xFinal, vFinal, xp, vp, blk,
                  cell);
}

constexpr double SPEED_OF_LIGHT = 299792458.0;  // SpeedOfLight, constants.h:40
constexpr double PI_TIMES_2 = 6.28318530717958647692528676655900576839433879875021;  // constants.h:17

// One particle of PIC::Mover::Relativistic::Boris
// (pic_mover_relativistic_boris.cpp:16-583, scalar branch, forward time):
// sub-steps of at most one gyro period (:113-124) with the fields of each
// sub-step's block — the tile's block from LDS, any other block from HBM —
// momentum-form Boris rotation (:134-167), sphere absorption, and the exit
// face found along the chord xFinal - xInit with a fractional intersection
// time (:317-399; the found face's normal is used where the reference indexes
// its face table with the spent loop counter).  A reflected particle is
// committed at the face and the rest of its step dropped (the reference
// aborts there, :426-432).
__device__ __forceinline__ int push_particle_rel(const MeshView& m, const Faces& faces, int b0,
                                                 const BlockGeom& g0, const double* sE,
                                                 const double* sB, const double* __restrict__ Eg,
                                                 const double* __restrict__ Bg, double ElectricCharge,
                                                 double mass, double dtTotalIn, int bc_mode,
                                                 double r_sphere, double xp[3], double vp[3],
                                                 int* blk, int* cell) {
  const double c2 = SPEED_OF_LIGHT * SPEED_OF_LIGHT;
  double vInit[3] = {vp[0], vp[1], vp[2]}, xInit[3] = {xp[0], xp[1], xp[2]};
  double vFinal[3] = {vp[0], vp[1], vp[2]}, xFinal[3] = {xp[0], xp[1], xp[2]};
  int curb = b0;
  BlockGeom g = g0;
  while (dtTotalIn > 0.0) {
    double gamma = 1.0 / sqrt(1.0 - (vInit[0] * vInit[0] + vInit[1] * vInit[1] + vInit[2] * vInit[2]) / c2);
    double wn[8], table[8];
    int iX[3];
    corner_stencil(g, xInit, wn, table, iX);
#pragma unroll
    for (int d = 0; d < 3; d++) iX[d] = clampi(iX[d], 0, NC - 1);
    const double* fE = (curb == b0) ? sE : Eg + (size_t)curb * NCORNER * 3;
    const double* fB = (curb == b0) ? sB : Bg + (size_t)curb * NCORNER * 3;
    double E[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int l = 0; l < 8; l++) {
      const int c = (iX[0] + (l >> 2)) + NN * ((iX[1] + ((l >> 1) & 1)) + NN * (iX[2] + (l & 1)));
#pragma unroll
      for (int d = 0; d < 3; d++) {
        E[d] += wn[l] * fE[3 * c + d];
        B[d] += wn[l] * fB[3 * c + d];
      }
    }
    double dt;
    const double lB = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    if (lB > 1.0E-25) {
      const double GyroFreq = fabs(ElectricCharge) * lB / (PI_TIMES_2 * mass * gamma);
      const double dtMax = 1.0 / GyroFreq;
      dt = (dtMax < dtTotalIn) ? dtMax : dtTotalIn;
    } else {
      dt = dtTotalIn;
    }
    dtTotalIn -= dt;
    const double QdT_over_twoM = ElectricCharge * dt / (2.0 * mass);
    double uMinus[3], t[3], s[3], uPrime[3], uPlus[3], uFinal[3], l2 = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) uMinus[d] = gamma * vInit[d] + QdT_over_twoM * E[d];
    gamma = sqrt(1.0 + (uMinus[0] * uMinus[0] + uMinus[1] * uMinus[1] + uMinus[2] * uMinus[2]) / c2);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      t[d] = QdT_over_twoM / gamma * B[d];
      l2 += t[d] * t[d];
    }
    uPrime[0] = uMinus[1] * t[2] - uMinus[2] * t[1];
    uPrime[1] = uMinus[2] * t[0] - uMinus[0] * t[2];
    uPrime[2] = uMinus[0] * t[1] - uMinus[1] * t[0];
#pragma unroll
    for (int d = 0; d < 3; d++) uPrime[d] += uMinus[d];
#pragma unroll
    for (int d = 0; d < 3; d++) s[d] = 2.0 * t[d] / (1.0 + l2);
    uPlus[0] = uPrime[1] * s[2] - uPrime[2] * s[1];
    uPlus[1] = uPrime[2] * s[0] - uPrime[0] * s[2];
    uPlus[2] = uPrime[0] * s[1] - uPrime[1] * s[0];
#pragma unroll
    for (int d = 0; d < 3; d++) uPlus[d] += uMinus[d];
#pragma unroll
    for (int d = 0; d < 3; d++) uFinal[d] = uPlus[d] + QdT_over_twoM * E[d];
    gamma = sqrt(1.0 + (uFinal[0] * uFinal[0] + uFinal[1] * uFinal[1] + uFinal[2] * uFinal[2]) / c2);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      vFinal[d] = uFinal[d] / gamma;
      xFinal[d] = xInit[d] + vFinal[d] * dt;
    }
    double rFinal2;
    if (r_sphere > 0.0 &&
        (rFinal2 = xFinal[0] * xFinal[0] + xFinal[1] * xFinal[1] + xFinal[2] * xFinal[2]) <
            r_sphere * r_sphere) {
      const double r = sqrt(rFinal2);
#pragma unroll
      for (int d = 0; d < 3; d++) xFinal[d] *= r_sphere / r;
#pragma unroll
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return AMPS_CODE_SPHERE;
    }
    int newNode = find_block(m, xFinal);
    if (newNode < 0) {
      if (bc_mode == AMPS_BC_DELETE) {
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
        return AMPS_CODE_LEFT_DOMAIN;
      }
      int nIntersectionFace = -1;
      double vEffective[3], r0[3], dtIntersection = -1.0;
#pragma unroll
      for (int d = 0; d < 3; d++) vEffective[d] = xFinal[d] - xInit[d];
      for (int nface = 0; nface < 6; nface++) {
        const Face& F = faces.f[nface];
        double cx = 0.0, cv = 0.0;
#pragma unroll
        for (int d = 0; d < 3; d++) {
          r0[d] = xInit[d] - F.x0[d];
          cx += r0[d] * F.norm[d];
          cv += vEffective[d] * F.norm[d];
        }
        const double dtEffective = (cv > 0.0) ? -cx / cv : -1.0;
        if (dtEffective > 0.0) {
          if ((dtIntersection < 0.0) || ((dtEffective < dtIntersection) && (dtEffective > 0.0))) {
            double cE0 = 0.0, cE1 = 0.0;
#pragma unroll
            for (int d = 0; d < 3; d++) {
              const double c = r0[d] + dtEffective * vEffective[d];
              cE0 += c * F.e0[d], cE1 += c * F.e1[d];
            }
            if ((cE0 < -m.eps) || (cE0 > F.lE0 + m.eps) || (cE1 < -m.eps) || (cE1 > F.lE1 + m.eps))
              continue;
            nIntersectionFace = nface, dtIntersection = dtEffective;
          }
        }
      }
      if (nIntersectionFace == -1) {
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
        return AMPS_CODE_LEFT_DOMAIN;
      }
      const Face& F = faces.f[nIntersectionFace];
#pragma unroll
      for (int d = 0; d < 3; d++) {
        xInit[d] += dtIntersection * (xFinal[d] - xInit[d]) - F.norm[d] * m.eps;
        vInit[d] += dtIntersection * (vFinal[d] - vInit[d]);
      }
      newNode = find_block(m, xInit);
      if (newNode < 0) {
#pragma unroll
        for (int d = 0; d < 3; d++) {
          if (m.gxmin[d] >= xInit[d]) xInit[d] = m.gxmin[d] + m.eps;
          if (m.gxmax[d] <= xInit[d]) xInit[d] = m.gxmax[d] - m.eps;
        }
        newNode = find_block(m, xInit);
        if (newNode < 0 && bc_mode != AMPS_BC_USER) {
#pragma unroll
          for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
          return AMPS_CODE_LEFT_DOMAIN;
        }
      }
      if (bc_mode == AMPS_BC_USER) {
        // ProcessOutsideDomainParticles (pic_mover_relativistic_boris.cpp:418-420)
#pragma unroll
        for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
        *blk = newNode;
        *cell = 0;
        return AMPS_CODE_EXIT_FACE0 + nIntersectionFace;
      }
      double cc = 0.0;
#pragma unroll
      for (int d = 0; d < 3; d++) cc += F.norm[d] * vInit[d];
#pragma unroll
      for (int d = 0; d < 3; d++) vInit[d] -= 2.0 * cc * F.norm[d];
      const int c_id = find_cell(m, xInit, newNode);
      if (c_id < 0) return AMPS_CODE_LEFT_DOMAIN;
      *blk = newNode;
      *cell = c_id;
#pragma unroll
      for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
      return AMPS_CODE_REFLECTED;
    }
    if (newNode != curb) {
      curb = newNode;
      load_block_geom(m, curb, g);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) xInit[d] = xFinal[d], vInit[d] = vFinal[d];
  }
  const int c_id = find_cell(m, xFinal, curb);
  if (c_id < 0) return AMPS_CODE_LEFT_DOMAIN;
  *blk = curb;
  *cell = c_id;
#pragma unroll
  for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
  return AMPS_CODE_FINISHED;
}

// FP32 push: the same mover with FP32 stencil weights, field interpolation
// and velocity update; the position increment dt*v is formed in FP32 and
// accumulated into the FP64 position, and block/cell location, sphere and
// boundary handling run on the FP64 position (relocate()).
template <int MOVER>
__device__ __forceinline__ int push_particle_f32(const MeshView& m, const Faces& faces,
                                                 const BlockGeom& g, const BlockLat& lat, const float* sE,
                                                 const float* sB, float q, float mass,
                                                 double dtTotal, int bc_mode, double r_sphere,
                                                 double xp[3], float vf[3], int* blk, int* cell) {
#pragma clang fp contract(fast)
  double xInit[3] = {xp[0], xp[1], xp[2]};
  float xl[3];
  int iX[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (fabs(xInit[d] - g.xmax[d]) < 1e-10 * g.dxs[d]) xInit[d] = g.xmax[d] - 1e-10 * g.dxs[d];
    const double loc = (xInit[d] - g.xmin[d]) / g.dxs[d];
    iX[d] = (int)loc;
    xl[d] = (float)(loc - iX[d]);
    iX[d] = clampi(iX[d], 0, NC - 1);
  }
  float w[8], norm = 0.0f;
#pragma unroll
  for (int l = 0; l < 8; l++) {
    const float a = (l >> 2) ? xl[0] : (1.0f - xl[0]);
    const float b = ((l >> 1) & 1) ? xl[1] : (1.0f - xl[1]);
    const float c = (l & 1) ? xl[2] : (1.0f - xl[2]);
    w[l] = a * b * c;
    norm += w[l];
  }
  const float inorm = 1.0f / norm;
  float E[3] = {0.f, 0.f, 0.f}, B[3] = {0.f, 0.f, 0.f};
#pragma unroll
  for (int l = 0; l < 8; l++) {
    const int c = (iX[0] + (l >> 2)) + NN * ((iX[1] + ((l >> 1) & 1)) + NN * (iX[2] + (l & 1)));
    const float wl = w[l] * inorm;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      E[d] += wl * sE[3 * c + d];
      B[d] += wl * sB[3 * c + d];
    }
  }
  const float dt = (float)dtTotal;
  float vF[3];
  if (MOVER == AMPS_MOVER_LAPENTA) {
    const float h = 0.5f * q * dt / mass;
    const float hB[3] = {h * B[0], h * B[1], h * B[2]};
    const float c0 = 1.0f / (1.0f + hB[0] * hB[0] + hB[1] * hB[1] + hB[2] * hB[2]);
    const float u[3] = {vf[0] + h * E[0], vf[1] + h * E[1], vf[2] + h * E[2]};
    // alpha u = c0 (u + (hB.u) hB + u x hB)
    const float hu = hB[0] * u[0] + hB[1] * u[1] + hB[2] * u[2];
    const float au[3] = {c0 * (u[0] + hu * hB[0] + (u[1] * hB[2] - u[2] * hB[1])),
                         c0 * (u[1] + hu * hB[1] + (u[2] * hB[0] - u[0] * hB[2])),
                         c0 * (u[2] + hu * hB[2] + (u[0] * hB[1] - u[1] * hB[0]))};
#pragma unroll
    for (int d = 0; d < 3; d++) vF[d] = 2.0f * au[d] - vf[d];
  } else {
    const float qm = q / mass, ht = 0.5f * dt;
    const float u[3] = {vf[0] + ht * qm * E[0], vf[1] + ht * qm * E[1], vf[2] + ht * qm * E[2]};
    const float hh[3] = {ht * qm * B[0], ht * qm * B[1], ht * qm * B[2]};
    const float h2 = hh[0] * hh[0] + hh[1] * hh[1] + hh[2] * hh[2];
    const float uh = u[0] * hh[0] + u[1] * hh[1] + u[2] * hh[2];
    const float i1 = 1.0f / (1.0f + h2);
    const float U[3] = {((1 - h2) * u[0] + 2 * (u[1] * hh[2] - hh[1] * u[2] + uh * hh[0])) * i1,
                        ((1 - h2) * u[1] + 2 * (u[2] * hh[0] - hh[2] * u[0] + uh * hh[1])) * i1,
                        ((1 - h2) * u[2] + 2 * (u[0] * hh[1] - hh[0] * u[1] + uh * hh[2])) * i1};
#pragma unroll
    for (int d = 0; d < 3; d++) vF[d] = U[d] + ht * qm * E[d];
  }
  double vInit[3] = {(double)vf[0], (double)vf[1], (double)vf[2]};
  double vFinal[3] = {(double)vF[0], (double)vF[1], (double)vF[2]};
  double xFinal[3];
#pragma unroll
  for (int d = 0; d < 3; d++) xFinal[d] = xInit[d] + (double)(dt * vF[d]);
  double vo[3];
  const int code =
      relocate(m, faces, g, lat, dtTotal, bc_mode, r_sphere, xInit, vInit, xFinal, vFinal, xp, vo, blk, cell);
#pragma unroll
  for (int d = 0; d < 3; d++) vf[d] = (float)vo[d];
  return code;
}

// ---------------------------------------------------------------------
// ProcessCell per-particle body restricted to one chunk [O0,O1) of the
// 348 per-cell outputs: 0..23 J (corner c = o/3, dim o%3), then 36 lower
// triangular corner pairs (i >= j, i-major) x 9 alpha entries.
// ---------------------------------------------------------------------
constexpr int pair_i(int pr) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= pr) i++;
  return i;
}
constexpr int pair_j(int pr) { return pr - pair_i(pr) * (pair_i(pr) + 1) / 2; }

template <int O0, int O1, bool EXACT>
__device__ __forceinline__ void cell_chunk(const BlockGeom& g, const Species& sp, double dtTotal,
                                           const double* sB, int sx, int sy, int sz, int p0,
                                           int p1, int64_t cap, const double* __restrict__ x,
                                           const double* __restrict__ v, int vf32,
                                           const uint8_t* __restrict__ spec, double CellVolume,
                                           double* out) {
  constexpr int NO = O1 - O0;
  double acc[NO];
#pragma unroll
  for (int o = 0; o < NO; o++) acc[o] = 0.0;
  for (int p = p0; p < p1; ++p) {
    double xI[3] = {x[p], x[cap + p], x[2 * cap + p]};
    double vI[3];
    if (vf32) {
      const float* vf = (const float*)v;
      vI[0] = vf[p], vI[1] = vf[cap + p], vI[2] = vf[2 * cap + p];
    } else {
      vI[0] = v[p], vI[1] = v[cap + p], vI[2] = v[2 * cap + p];
    }
    const int s = spec[p];
    double wn[8], table[8];
    int iX[3];
    corner_stencil(g, xI, wn, table, iX);
    iX[0] = clampi(iX[0], sx, sx + 3);
    iX[1] = clampi(iX[1], sy, sy + 3);
    iX[2] = clampi(iX[2], sz, sz + 3);
    double Bv[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int l = 0; l < 8; l++) {
      const int c = (iX[0] - sx + (l >> 2)) + 5 * ((iX[1] - sy + ((l >> 1) & 1)) + 5 * (iX[2] - sz + (l & 1)));
#pragma unroll
      for (int d = 0; d < 3; d++) Bv[d] += wn[l] * sB[3 * c + d];
    }
    double chargeQ = sp.charge[s], mass = sp.mass[s];
    chargeQ *= sp.weight[s];
    mass *= sp.weight[s];
    const double QdT_over_m = chargeQ * dtTotal / mass;
    const double QdT_over_2m = 0.5 * QdT_over_m;
    const double QdT_over_2m_squared = QdT_over_2m * QdT_over_2m;
    double BB[3][3], P[3], alpha[9];
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
      P[ii] = -QdT_over_2m * Bv[ii];
#pragma unroll
      for (int jj = 0; jj <= ii; jj++) {
        BB[ii][jj] = QdT_over_2m_squared * Bv[ii] * Bv[jj];
        BB[jj][ii] = BB[ii][jj];
      }
    }
    const double c0 =
        1.0 / (1.0 + QdT_over_2m_squared * (Bv[0] * Bv[0] + Bv[1] * Bv[1] + Bv[2] * Bv[2]));
    alpha[0] = c0 * (1.0 + BB[0][0]);
    alpha[1] = c0 * (-P[2] + BB[0][1]);
    alpha[2] = c0 * (P[1] + BB[0][2]);
    alpha[3] = c0 * (P[2] + BB[1][0]);
    alpha[4] = c0 * (1.0 + BB[1][1]);
    alpha[5] = c0 * (-P[0] + BB[1][2]);
    alpha[6] = c0 * (-P[1] + BB[2][0]);
    alpha[7] = c0 * (P[0] + BB[2][1]);
    alpha[8] = c0 * (1.0 + BB[2][2]);
    double vRot[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int jj = 0; jj < 3; jj++) vRot[i] += alpha[3 * i + jj] * vI[jj];
    const double matrixConst = chargeQ * QdT_over_2m / CellVolume;
    double twc[8];
#pragma unroll
    for (int i = 0; i < 8; i++) twc[i] = matrixConst * table[i];
#pragma unroll
    for (int o = O0; o < O1; o++) {
      if (o < 24) {
        const double t = chargeQ * table[o / 3];
        if (EXACT) acc[o - O0] += t * vRot[o % 3];
        else acc[o - O0] = fma(t, vRot[o % 3], acc[o - O0]);
      } else {
        const int pr = (o - 24) / 9, k = (o - 24) % 9;
        const double twp = table[pair_j(pr)] * twc[pair_i(pr)];
        if (EXACT) acc[o - O0] += alpha[k] * twp;
        else acc[o - O0] = fma(alpha[k], twp, acc[o - O0]);
      }
    }
  }
#pragma unroll
  for (int o = O0; o < O1; o++) out[o] = (o < 24) ? acc[o - O0] / CellVolume : acc[o - O0];
}

template <int NCH, bool EXACT, int W = 0>
__device__ __forceinline__ void dispatch_chunk(int w, const BlockGeom& g, const Species& sp,
                                               double dtTotal, const double* sB, int sx, int sy,
                                               int sz, int p0, int p1, int64_t cap,
                                               const double* __restrict__ x,
                                               const double* __restrict__ v, int vf32,
                                               const uint8_t* __restrict__ spec,
                                               double CellVolume, double* out) {
  if constexpr (W < NCH) {
    constexpr int CH = (CELL_VALS + NCH - 1) / NCH;
    constexpr int O0 = W * CH;
    constexpr int O1 = (O0 + CH < CELL_VALS) ? O0 + CH : CELL_VALS;
    if (w == W)
      cell_chunk<O0, O1, EXACT>(g, sp, dtTotal, sB, sx, sy, sz, p0, p1, cap, x, v, vf32, spec,
                                CellVolume, out);
    else
      dispatch_chunk<NCH, EXACT, W + 1>(w, g, sp, dtTotal, sB, sx, sy, sz, p0, p1, cap, x, v,
                                        vf32, spec, CellVolume, out);
  }
}

}  // namespace amps
