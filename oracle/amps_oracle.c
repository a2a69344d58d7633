/*
 * amps_oracle.c — CPU FP64 restatement of the AMPS particle hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see amps_oracle.h).  Built with
 * -ffp-contract=off so that every expression rounds exactly as written,
 * in the order the reference writes it.
 */
#include "amps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* AMR octree: restates cTreeNodeAMR's search fields                      */
/* (src/meshAMR/meshAMRgeneric.h:2793-2882).                              */
/* ===================================================================== */
typedef struct {
  int64_t xmin_g[3]; /* xMinGlobalIndex */
  int64_t size_g;    /* NodeGeometricSizeIndex */
  int level;         /* RefinmentLevel */
  double xmin[3], xmax[3];
  int child[8];      /* downNode[i+2(j+2k)] */
  int up;            /* upNode */
  int block;         /* leaf block id, -1 for internal nodes */
} orc_tnode;

struct orc_mesh {
  double gxmin[3], gxmax[3], dx_root[3], dx_maxref[3], eps;
  int lmax, nblocks, nnodes, cap;
  orc_tnode* t;
  int* block_node;
  int* nb_info;  /* per block: minNeib, maxNeib, face[6], edge[12], corner[8] (28 ints) */
  const double* Bc; /* centre-based B with ghost layer, or NULL */
};

static int new_tnode(orc_mesh* m, int up, const int64_t xg[3], int64_t size, int level) {
  if (m->nnodes == m->cap) {
    m->cap = m->cap ? 2 * m->cap : 1024;
    m->t = (orc_tnode*)realloc(m->t, sizeof(orc_tnode) * (size_t)m->cap);
  }
  orc_tnode* n = &m->t[m->nnodes];
  for (int d = 0; d < 3; d++) {
    n->xmin_g[d] = xg[d];
    /* Block bounds on the max-refinement lattice; exact for the dyadic domains
     * used here, as any splitting rule of the reference is. */
    n->xmin[d] = m->gxmin[d] + (double)xg[d] * m->dx_maxref[d];
    n->xmax[d] = m->gxmin[d] + (double)(xg[d] + size) * m->dx_maxref[d];
  }
  n->size_g = size;
  n->level = level;
  for (int c = 0; c < 8; c++) n->child[c] = -1;
  n->up = up;
  n->block = -1;
  return m->nnodes++;
}

static void build_neighbour_info(orc_mesh* m);
static void center_B_at(const orc_mesh* m, const double x[3], int tn, double B[3]);

orc_mesh* orc_mesh_create(const double gxmin[3], const double gxmax[3], int lmax, int nblocks,
                          const int* level, const int64_t* ijk) {
  orc_mesh* m = (orc_mesh*)calloc(1, sizeof(orc_mesh));
  m->lmax = lmax;
  m->nblocks = nblocks;
  /* meshAMRgeneric.h:2338-2366 (init): dxRootBlock, EPS, dx_max_refinment */
  for (int d = 0; d < 3; d++) {
    m->gxmin[d] = gxmin[d];
    m->gxmax[d] = gxmax[d];
    m->dx_root[d] = gxmax[d] - gxmin[d];
    m->dx_maxref[d] = (gxmax[d] - gxmin[d]) / (1 << lmax);
  }
  m->eps = 0.0001 * m->dx_root[0] / (double)ORC_NC / (1 << lmax);
  for (int d = 1; d < 3; d++) {
    double e = 0.0001 * m->dx_root[d] / (double)ORC_NC / (1 << lmax);
    if (m->eps > e) m->eps = e;
  }
  int64_t zero[3] = {0, 0, 0};
  new_tnode(m, -1, zero, (int64_t)1 << lmax, 0);
  m->block_node = (int*)malloc(sizeof(int) * (size_t)nblocks);
  for (int b = 0; b < nblocks; b++) {
    int L = level[b];
    if (L < 0 || L > lmax - 3) goto fail;
    int64_t tg[3];
    for (int d = 0; d < 3; d++) tg[d] = ijk[3 * b + d] << (lmax - L);
    int cur = 0;
    while (m->t[cur].level < L) {
      if (m->t[cur].block >= 0) goto fail; /* leaf cannot be refined */
      int64_t half = m->t[cur].size_g / 2;
      int c = 0;
      int64_t cg[3];
      for (int d = 0; d < 3; d++) {
        int bit = (tg[d] - m->t[cur].xmin_g[d]) >= half;
        c |= bit << d;
        cg[d] = m->t[cur].xmin_g[d] + (bit ? half : 0);
      }
      if (m->t[cur].child[c] < 0) {
        int nn = new_tnode(m, cur, cg, half, m->t[cur].level + 1);
        m->t[cur].child[c] = nn;
      }
      cur = m->t[cur].child[c];
    }
    if (m->t[cur].block >= 0) goto fail;
    for (int c = 0; c < 8; c++)
      if (m->t[cur].child[c] >= 0) goto fail;
    m->t[cur].block = b;
    m->block_node[b] = cur;
  }
  /* every internal node must have all eight children (the leaves tile the root) */
  for (int i = 0; i < m->nnodes; i++) {
    if (m->t[i].block >= 0) continue;
    for (int c = 0; c < 8; c++)
      if (m->t[i].child[c] < 0) goto fail;
  }
  build_neighbour_info(m);
  return m;
fail:
  orc_mesh_destroy(m);
  return NULL;
}

void orc_mesh_destroy(orc_mesh* m) {
  if (!m) return;
  free(m->t);
  free(m->block_node);
  free(m->nb_info);
  free(m);
}

double orc_mesh_eps(const orc_mesh* m) { return m->eps; }

void orc_block_bounds(const orc_mesh* m, int b, double xmin[3], double xmax[3]) {
  const orc_tnode* n = &m->t[m->block_node[b]];
  for (int d = 0; d < 3; d++) {
    xmin[d] = n->xmin[d];
    xmax[d] = n->xmax[d];
  }
}

/* findTreeNode(int*) — meshAMRgeneric.h:2793-2848 */
static int find_tnode_int(const orc_mesh* m, const int ix[3], int start) {
  int cur = start;
  for (;;) {
    const orc_tnode* n = &m->t[cur];
    int inblock = 1;
    for (int d = 0; d < 3; d++) {
      if ((ix[d] < n->xmin_g[d]) || (ix[d] >= n->xmin_g[d] + n->size_g)) {
        inblock = 0;
        break;
      }
    }
    if (inblock) {
      int i = (ix[0] - n->xmin_g[0] < n->size_g / 2) ? 0 : 1;
      int j = (ix[1] - n->xmin_g[1] < n->size_g / 2) ? 0 : 1;
      int k = (ix[2] - n->xmin_g[2] < n->size_g / 2) ? 0 : 1;
      int t = n->child[i + 2 * (j + 2 * k)];
      if (t >= 0) {
        cur = t;
        continue;
      }
      return cur;
    }
    if (n->up >= 0) {
      cur = n->up;
      continue;
    }
    return -1;
  }
}

/* findTreeNode(double*) — meshAMRgeneric.h:2852-2882 */
static int find_tnode(const orc_mesh* m, const double x[3], int start) {
  int ix[3];
  for (int d = 0; d < 3; d++) ix[d] = (int)floor((x[d] - m->gxmin[d]) / m->dx_maxref[d]);
  int res = find_tnode_int(m, ix, start);
  int flag = 0;
  if (res >= 0) {
    for (int d = 0; d < 3; d++) {
      if (x[d] < m->t[res].xmin[d]) ix[d]--, flag = 1;
      if (x[d] >= m->t[res].xmax[d]) ix[d]++, flag = 1;
    }
  }
  if (flag) res = find_tnode_int(m, ix, res);
  return res;
}

int orc_find_block(const orc_mesh* m, const double x[3], int start_block) {
  int start = (start_block >= 0) ? m->block_node[start_block] : 0;
  int r = find_tnode(m, x, start);
  return (r >= 0) ? m->t[r].block : -1;
}

/* FindCellIndex — meshAMRgeneric.h:2256-2323 (ExitFlag=false) */
static int find_cell_node(const orc_mesh* m, const double x[3], const orc_tnode* n, int ijk[3]) {
  for (int d = 0; d < 3; d++) {
    if ((x[d] < n->xmin[d]) || (n->xmax[d] < x[d])) return -1;
    double dx = m->dx_root[d] / (1 << n->level) / (double)ORC_NC;
    int i = (int)((x[d] - n->xmin[d]) / dx);
    if (i == ORC_NC) i = ORC_NC - 1;
    ijk[d] = i;
  }
  return ijk[0] + ORC_NC * (ijk[1] + ORC_NC * ijk[2]);
}

int orc_find_cell(const orc_mesh* m, const double x[3], int block) {
  int ijk[3];
  return find_cell_node(m, x, &m->t[m->block_node[block]], ijk);
}

/* CornerBased::InitStencil — pic_interpolation_routines.cpp:1074-1194.
 * Stencil loop order: i outer, k inner (l = 4i+2j+k).  table[] follows the
 * reference's InterpolationCoefficientTable ordering 0,1,3,2,4,5,7,6.
 * Normalize(): pic.h:7255-7262. */
static void corner_stencil_node(const orc_tnode* n, double x[3], double w_norm[8],
                                double table[8], int corner[8]) {
  double dx[3], xLoc[3];
  int iX[3];
  for (int d = 0; d < 3; d++) dx[d] = (n->xmax[d] - n->xmin[d]) / ORC_NC;
  for (int d = 0; d < 3; d++) {
    if (fabs(x[d] - n->xmax[d]) < 1e-10 * dx[d]) x[d] = n->xmax[d] - 1e-10 * dx[d];
    xLoc[d] = (x[d] - n->xmin[d]) / dx[d];
    iX[d] = (int)(xLoc[d]);
    xLoc[d] -= iX[d];
  }
  static const int tslot[8] = {0, 4, 3, 7, 1, 5, 2, 6}; /* l = 4i+2j+k -> table index */
  double norm = 0.0;
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        int l = 4 * i + 2 * j + k;
        double a = i ? xLoc[0] : (1.0 - xLoc[0]);
        double b = j ? xLoc[1] : (1.0 - xLoc[1]);
        double c = k ? xLoc[2] : (1.0 - xLoc[2]);
        double w = a * b * c;
        table[tslot[l]] = w;
        w_norm[l] = w;
        corner[l] = (iX[0] + i) + ORC_NN * ((iX[1] + j) + ORC_NN * (iX[2] + k));
        norm += w;
      }
  if (norm > 0.0)
    for (int l = 0; l < 8; l++) w_norm[l] /= norm;
}

void orc_corner_stencil(const orc_mesh* m, double x[3], int block, double w_norm[8],
                        double table[8], int corner[8]) {
  corner_stencil_node(&m->t[m->block_node[block]], x, w_norm, table, corner);
}

/* ===================================================================== */
/* Mover                                                                  */
/* ===================================================================== */
typedef struct {
  double norm[3], e0[3], e1[3], x0[3], lE0, lE1;
} orc_face;

/* ExternalBoundaryFaceTable + Init — pic_mover.cpp:24-28, 60-74 */
static void build_faces(const orc_mesh* m, orc_face f[6]) {
  static const double nrm[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  static const int nX0[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 1}};
  static const double e0[6][3] = {{0, 1, 0}, {0, 1, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}, {1, 0, 0}};
  static const double e1[6][3] = {{0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 0, 1}, {0, 1, 0}, {0, 1, 0}};
  for (int nf = 0; nf < 6; nf++) {
    double cE0 = 0.0, cE1 = 0.0;
    for (int d = 0; d < 3; d++) {
      f[nf].norm[d] = nrm[nf][d];
      f[nf].e0[d] = e0[nf][d];
      f[nf].e1[d] = e1[nf][d];
      f[nf].x0[d] = (nX0[nf][d] == 0) ? m->gxmin[d] : m->gxmax[d];
      double a = ((e0[nf][d] + nX0[nf][d] < 0.5) ? m->gxmin[d] : m->gxmax[d]) - f[nf].x0[d];
      double b = ((e1[nf][d] + nX0[nf][d] < 0.5) ? m->gxmin[d] : m->gxmax[d]) - f[nf].x0[d];
      cE0 += pow(a, 2);
      cE1 += pow(b, 2);
    }
    f[nf].lE0 = sqrt(cE0);
    f[nf].lE1 = sqrt(cE1);
  }
}

/* One particle: PIC::Mover::Lapenta2017 (pic_mover_boris.cpp:876-1380) or
 * PIC::Mover::Boris (pic_mover_boris.cpp:126-553, rotation at :253-277), with
 * corner E and B.  Returns the code; x/v/blk/cell updated when committed. */
static int push_one(const orc_mesh* m, const orc_face* faces, const double* E_corner,
                    const double* B_corner, double q, double mass, double dtTotal, int mover,
                    int bc_mode, double r_sphere, double xp[3], double vp[3], int* blk, int* cell) {
  double vInit[3] = {vp[0], vp[1], vp[2]}, xInit[3] = {xp[0], xp[1], xp[2]};
  double vFinal[3], xFinal[3];
  int startNode = m->block_node[*blk];
  const orc_tnode* sn = &m->t[startNode];

  /* stencil + field accumulation (:945-977) */
  double w[8], table[8];
  int cr[8];
  corner_stencil_node(sn, xInit, w, table, cr);
  double E[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
  const double* Eb = E_corner + (size_t)(*blk) * ORC_NCORNER * 3;
  const double* Bb = B_corner + (size_t)(*blk) * ORC_NCORNER * 3;
  for (int l = 0; l < 8; l++) {
    const double* e = Eb + 3 * cr[l];
    const double* b = Bb + 3 * cr[l];
    for (int d = 0; d < 3; d++) {
      E[d] += w[l] * e[d];
      B[d] += w[l] * b[d];
    }
  }

  /* centre-based B: CellCentered::Linear::InitStencil at the (nudged) xInit
   * (pic_mover_boris.cpp:979-1000) */
  if (m->Bc) center_B_at(m, xInit, startNode, B);

  if (mover == ORC_MOVER_LAPENTA) {
    /* :1033-1081 */
    double QdT_over_m = q * dtTotal / mass;
    double QdT_over_2m = 0.5 * QdT_over_m;
    double QdT_over_2m_squared = QdT_over_2m * QdT_over_2m;
    double BB[3][3], P[3], alpha[3][3];
    for (int i = 0; i < 3; i++) {
      P[i] = -QdT_over_2m * B[i];
      for (int j = 0; j <= i; j++) {
        BB[i][j] = QdT_over_2m_squared * B[i] * B[j];
        BB[j][i] = BB[i][j];
      }
    }
    double c0 = 1.0 / (1.0 + QdT_over_2m_squared * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]));
    alpha[0][0] = c0 * (1.0 + BB[0][0]);
    alpha[0][1] = c0 * (-P[2] + BB[0][1]);
    alpha[0][2] = c0 * (P[1] + BB[0][2]);
    alpha[1][0] = c0 * (P[2] + BB[1][0]);
    alpha[1][1] = c0 * (1.0 + BB[1][1]);
    alpha[1][2] = c0 * (-P[0] + BB[1][2]);
    alpha[2][0] = c0 * (-P[1] + BB[2][0]);
    alpha[2][1] = c0 * (P[0] + BB[2][1]);
    alpha[2][2] = c0 * (1.0 + BB[2][2]);
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 3; j++) s += alpha[i][j] * (vInit[j] + QdT_over_2m * E[j]);
      vFinal[i] = 2.0 * s - vInit[i];
    }
    for (int i = 0; i < 3; i++) xFinal[i] = xInit[i] + dtTotal * vFinal[i];
  } else {
    /* BorisSplitAcceleration_default :87-98 and the rotation :253-277 */
    double Charge2Mass = q / mass, accl[3], rot[3], u[3], h[3], U[3];
    for (int d = 0; d < 3; d++) {
      accl[d] = 0.0 + Charge2Mass * E[d];
      rot[d] = 0.0 - Charge2Mass * B[d];
    }
    double dtTemp = dtTotal, dtTempOverTwo = dtTotal / 2.0;
    for (int d = 0; d < 3; d++) u[d] = vInit[d] + dtTempOverTwo * accl[d];
    for (int d = 0; d < 3; d++) h[d] = -dtTempOverTwo * rot[d];
    double h2 = h[0] * h[0] + h[1] * h[1] + h[2] * h[2];
    double uh = u[0] * h[0] + u[1] * h[1] + u[2] * h[2];
    U[0] = ((1 - h2) * u[0] + 2 * (u[1] * h[2] - h[1] * u[2] + uh * h[0])) / (1 + h2);
    U[1] = ((1 - h2) * u[1] + 2 * (u[2] * h[0] - h[2] * u[0] + uh * h[1])) / (1 + h2);
    U[2] = ((1 - h2) * u[2] + 2 * (u[0] * h[1] - h[0] * u[1] + uh * h[2])) / (1 + h2);
    for (int d = 0; d < 3; d++) vFinal[d] = U[d] + dtTempOverTwo * accl[d];
    for (int d = 0; d < 3; d++) xFinal[d] = xInit[d] + dtTemp * vFinal[d];
  }

  /* internal sphere (:1088-1117): absorb */
  int newNode;
  double rFinal2;
  if (r_sphere > 0.0 &&
      (rFinal2 = xFinal[0] * xFinal[0] + xFinal[1] * xFinal[1] + xFinal[2] * xFinal[2]) <
          r_sphere * r_sphere) {
    double r = sqrt(rFinal2);
    for (int d = 0; d < 3; d++) xFinal[d] *= r_sphere / r;
    for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
    return ORC_CODE_SPHERE;
  }
  newNode = find_tnode(m, xFinal, startNode);

  int code = ORC_CODE_FINISHED;
  if (newNode < 0) {
    if (bc_mode == ORC_BC_DELETE) {
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return ORC_CODE_LEFT_DOMAIN;
    }
    /* exit-face search and move to the face (:1156-1229); the reference reads
     * nIntersectionFace uninitialised, here it starts at -1 */
    int nIntersectionFace = -1;
    double vMiddle[3] = {0.5 * (vInit[0] + vFinal[0]), 0.5 * (vInit[1] + vFinal[1]),
                         0.5 * (vInit[2] + vFinal[2])};
    double dtIntersection = -1.0, r0[3], c;
    for (int nface = 0; nface < 6; nface++) {
      double cx = 0.0, cv = 0.0;
      for (int d = 0; d < 3; d++) {
        r0[d] = xInit[d] - faces[nface].x0[d];
        cx += r0[d] * faces[nface].norm[d];
        cv += vMiddle[d] * faces[nface].norm[d];
      }
      if (cv > 0.0) {
        double dt = -cx / cv;
        if ((dtIntersection < 0.0) || ((dt < dtIntersection) && (dt > 0.0))) {
          double cE0 = 0.0, cE1 = 0.0;
          for (int d = 0; d < 3; d++) {
            c = r0[d] + dt * vMiddle[d];
            cE0 += c * faces[nface].e0[d], cE1 += c * faces[nface].e1[d];
          }
          if ((cE0 < -m->eps) || (cE0 > faces[nface].lE0 + m->eps) || (cE1 < -m->eps) ||
              (cE1 > faces[nface].lE1 + m->eps))
            continue;
          nIntersectionFace = nface, dtIntersection = dt;
        }
      }
    }
    if (nIntersectionFace == -1) {
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return ORC_CODE_LEFT_DOMAIN; /* no face found: deleted */
    }
    double tVelocityIncrement = ((dtIntersection / dtTotal < 1) ? dtIntersection / dtTotal : 1);
    for (int d = 0; d < 3; d++) {
      xInit[d] += dtIntersection * vMiddle[d] - faces[nIntersectionFace].norm[d] * m->eps;
      vInit[d] += tVelocityIncrement * (vFinal[d] - vInit[d]);
    }
    newNode = find_tnode(m, xInit, startNode);
    if (newNode < 0) {
      for (int d = 0; d < 3; d++) {
        if (m->gxmin[d] >= xInit[d]) xInit[d] = m->gxmin[d] + m->eps;
        if (m->gxmax[d] <= xInit[d]) xInit[d] = m->gxmax[d] - m->eps;
      }
      newNode = find_tnode(m, xInit, startNode);
      if (newNode < 0 && bc_mode != ORC_BC_USER) {
        for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
        return ORC_CODE_LEFT_DOMAIN;
      }
    }
    /* USER_FUNCTION (:1230-1231): ProcessOutsideDomainParticles(ptr, xInit,
     * vInit, nIntersectionFace, newNode) receives the state on the face; the
     * particle is then deleted (:1262-1265).  newNode is -1 where the
     * reference aborts ("cannot find the node"). */
    if (bc_mode == ORC_BC_USER) {
      for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
      *blk = newNode >= 0 ? m->t[newNode].block : -1;
      *cell = 0;
      return ORC_CODE_EXIT_FACE0 + nIntersectionFace;
    }
    /* specular reflection (:1236-1253); the reference then aborts in its
     * switch(code) — here the reflected particle is committed. */
    {
      double cc = 0.0;
      for (int d = 0; d < 3; d++) cc += faces[nIntersectionFace].norm[d] * vInit[d];
      for (int d = 0; d < 3; d++) vInit[d] -= 2.0 * cc * faces[nIntersectionFace].norm[d];
    }
    for (int d = 0; d < 3; d++) vFinal[d] = vInit[d], xFinal[d] = xInit[d];
    code = ORC_CODE_REFLECTED;
  }

  /* FindCellIndex (:1286) and commit (:1374-1375) */
  int ijk[3];
  int c_id = find_cell_node(m, xFinal, &m->t[newNode], ijk);
  if (c_id < 0) return ORC_CODE_LEFT_DOMAIN;
  *blk = m->t[newNode].block;
  *cell = c_id;
  for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
  return code;
}

/* PIC::Mover::Relativistic::Boris, pic_mover_relativistic_boris.cpp:16-583,
 * scalar (non-AVX) branch, forward time integration.  Sub-steps limited to one
 * gyro period 1/f_gyro (:113-124, specfunc.h:1290-1292) with the fields
 * re-interpolated in the block of each sub-step; momentum form
 * u- = gamma v + (q dt/2m) E, rotation t = (q dt/2m/gamma) B,
 * s = 2t/(1+|t|^2), u = u+ + (q dt/2m) E, v = u/gamma(u) (:134-167).
 * Exit faces use the chord xFinal - xInit as the effective velocity and the
 * fractional intersection time (:317-399).  Where the reference reads the face
 * table at the loop counter after the loop (:391, :396 index nface == 6), the
 * found face nIntersectionFace is used.  A reflected particle is committed at
 * the face (the reference aborts in its final switch, :426-432) and the rest
 * of the step is dropped; sphere contact absorbs as in the other movers. */
static int push_rel_one(const orc_mesh* m, const orc_face* faces, const double* E_corner,
                        const double* B_corner, double ElectricCharge, double mass, double dtTotalIn,
                        int bc_mode, double r_sphere, double xp[3], double vp[3], int* blk,
                        int* cell) {
  const double c2 = ORC_SPEED_OF_LIGHT * ORC_SPEED_OF_LIGHT;
  double vInit[3] = {vp[0], vp[1], vp[2]}, xInit[3] = {xp[0], xp[1], xp[2]};
  double vFinal[3] = {vp[0], vp[1], vp[2]}, xFinal[3] = {xp[0], xp[1], xp[2]};
  int startNode = m->block_node[*blk];
  int newNode = startNode;
  while (dtTotalIn > 0.0) {
    double gamma = 1.0 / sqrt(1.0 - (vInit[0] * vInit[0] + vInit[1] * vInit[1] + vInit[2] * vInit[2]) / c2);
    const orc_tnode* sn = &m->t[startNode];
    double w[8], table[8];
    int cr[8];
    corner_stencil_node(sn, xInit, w, table, cr);
    double E[3] = {0.0, 0.0, 0.0}, B[3] = {0.0, 0.0, 0.0};
    const double* Eb = E_corner + (size_t)sn->block * ORC_NCORNER * 3;
    const double* Bb = B_corner + (size_t)sn->block * ORC_NCORNER * 3;
    for (int l = 0; l < 8; l++)
      for (int d = 0; d < 3; d++) {
        E[d] += w[l] * Eb[3 * cr[l] + d];
        B[d] += w[l] * Bb[3 * cr[l] + d];
      }
    double dt;
    const double lB = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
    if (lB > 1.0E-25) {
      /* GetGyroFrequency: |q| |B| / (2 pi m gamma(v)), gamma(v) as above */
      const double GyroFreq = fabs(ElectricCharge) * lB / (ORC_PI_TIMES_2 * mass * gamma);
      const double dtMax = 1.0 / GyroFreq;
      dt = (dtMax < dtTotalIn) ? dtMax : dtTotalIn;
    } else {
      dt = dtTotalIn;
    }
    dtTotalIn -= dt;
    const double QdT_over_twoM = ElectricCharge * dt / (2.0 * mass);
    double uMinus[3], t[3], s[3], uPrime[3], uPlus[3], uFinal[3], l = 0.0;
    for (int d = 0; d < 3; d++) uMinus[d] = gamma * vInit[d] + QdT_over_twoM * E[d];
    gamma = sqrt(1.0 + (uMinus[0] * uMinus[0] + uMinus[1] * uMinus[1] + uMinus[2] * uMinus[2]) / c2);
    for (int d = 0; d < 3; d++) {
      t[d] = QdT_over_twoM / gamma * B[d];
      l += t[d] * t[d];
    }
    uPrime[0] = uMinus[1] * t[2] - uMinus[2] * t[1];
    uPrime[1] = uMinus[2] * t[0] - uMinus[0] * t[2];
    uPrime[2] = uMinus[0] * t[1] - uMinus[1] * t[0];
    for (int d = 0; d < 3; d++) uPrime[d] += uMinus[d];
    for (int d = 0; d < 3; d++) s[d] = 2.0 * t[d] / (1.0 + l);
    uPlus[0] = uPrime[1] * s[2] - uPrime[2] * s[1];
    uPlus[1] = uPrime[2] * s[0] - uPrime[0] * s[2];
    uPlus[2] = uPrime[0] * s[1] - uPrime[1] * s[0];
    for (int d = 0; d < 3; d++) uPlus[d] += uMinus[d];
    for (int d = 0; d < 3; d++) uFinal[d] = uPlus[d] + QdT_over_twoM * E[d];
    gamma = sqrt(1.0 + (uFinal[0] * uFinal[0] + uFinal[1] * uFinal[1] + uFinal[2] * uFinal[2]) / c2);
    for (int d = 0; d < 3; d++) {
      vFinal[d] = uFinal[d] / gamma;
      xFinal[d] = xInit[d] + vFinal[d] * dt;
    }
    /* internal sphere (:270-292): absorbed */
    double rFinal2;
    if (r_sphere > 0.0 &&
        (rFinal2 = xFinal[0] * xFinal[0] + xFinal[1] * xFinal[1] + xFinal[2] * xFinal[2]) <
            r_sphere * r_sphere) {
      const double r = sqrt(rFinal2);
      for (int d = 0; d < 3; d++) xFinal[d] *= r_sphere / r;
      for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
      return ORC_CODE_SPHERE;
    }
    newNode = find_tnode(m, xFinal, startNode);
    if (newNode < 0) {
      if (bc_mode == ORC_BC_DELETE) {
        for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
        return ORC_CODE_LEFT_DOMAIN;
      }
      int nIntersectionFace = -1;
      double vEffective[3], r0[3], dtIntersection = -1.0;
      for (int d = 0; d < 3; d++) vEffective[d] = xFinal[d] - xInit[d];
      for (int nface = 0; nface < 6; nface++) {
        double cx = 0.0, cv = 0.0;
        for (int d = 0; d < 3; d++) {
          r0[d] = xInit[d] - faces[nface].x0[d];
          cx += r0[d] * faces[nface].norm[d];
          cv += vEffective[d] * faces[nface].norm[d];
        }
        const double dtEffective = (cv > 0.0) ? -cx / cv : -1.0;
        if (dtEffective > 0.0) {
          if ((dtIntersection < 0.0) || ((dtEffective < dtIntersection) && (dtEffective > 0.0))) {
            double cE0 = 0.0, cE1 = 0.0;
            for (int d = 0; d < 3; d++) {
              const double c = r0[d] + dtEffective * vEffective[d];
              cE0 += c * faces[nface].e0[d], cE1 += c * faces[nface].e1[d];
            }
            if ((cE0 < -m->eps) || (cE0 > faces[nface].lE0 + m->eps) || (cE1 < -m->eps) ||
                (cE1 > faces[nface].lE1 + m->eps))
              continue;
            nIntersectionFace = nface, dtIntersection = dtEffective;
          }
        }
      }
      if (nIntersectionFace == -1) {
        for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
        return ORC_CODE_LEFT_DOMAIN;
      }
      for (int d = 0; d < 3; d++) {
        xInit[d] += dtIntersection * (xFinal[d] - xInit[d]) - faces[nIntersectionFace].norm[d] * m->eps;
        vInit[d] += dtIntersection * (vFinal[d] - vInit[d]);
      }
      newNode = find_tnode(m, xInit, startNode);
      if (newNode < 0) {
        for (int d = 0; d < 3; d++) {
          if (m->gxmin[d] >= xInit[d]) xInit[d] = m->gxmin[d] + m->eps;
          if (m->gxmax[d] <= xInit[d]) xInit[d] = m->gxmax[d] - m->eps;
        }
        newNode = find_tnode(m, xInit, startNode);
        if (newNode < 0 && bc_mode != ORC_BC_USER) {
          for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
          return ORC_CODE_LEFT_DOMAIN;
        }
      }
      /* USER_FUNCTION (pic_mover_relativistic_boris.cpp:418-420) */
      if (bc_mode == ORC_BC_USER) {
        for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
        *blk = newNode >= 0 ? m->t[newNode].block : -1;
        *cell = 0;
        return ORC_CODE_EXIT_FACE0 + nIntersectionFace;
      }
      double cc = 0.0;
      for (int d = 0; d < 3; d++) cc += faces[nIntersectionFace].norm[d] * vInit[d];
      for (int d = 0; d < 3; d++) vInit[d] -= 2.0 * cc * faces[nIntersectionFace].norm[d];
      int ijk[3];
      const int c_id = find_cell_node(m, xInit, &m->t[newNode], ijk);
      if (c_id < 0) return ORC_CODE_LEFT_DOMAIN;
      *blk = m->t[newNode].block;
      *cell = c_id;
      for (int d = 0; d < 3; d++) xp[d] = xInit[d], vp[d] = vInit[d];
      return ORC_CODE_REFLECTED;
    }
    startNode = newNode;
    for (int d = 0; d < 3; d++) xInit[d] = xFinal[d], vInit[d] = vFinal[d];
  }
  /* FindCellIndex (:553) and commit (:561-562) */
  int ijk[3];
  const int c_id = find_cell_node(m, xFinal, &m->t[newNode], ijk);
  if (c_id < 0) return ORC_CODE_LEFT_DOMAIN;
  *blk = m->t[newNode].block;
  *cell = c_id;
  for (int d = 0; d < 3; d++) xp[d] = xFinal[d], vp[d] = vFinal[d];
  return ORC_CODE_FINISHED;
}

int orc_push(const orc_mesh* m, const double* E_corner, const double* B_corner,
             const double* charge, const double* mass, const double* dt_spec, int mover,
             int bc_mode, double r_sphere, int64_t n, double* x, double* v, const int* spec,
             int* blk, int* cell, int* code, int nthreads) {
  orc_face faces[6];
  build_faces(m, faces);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t p = 0; p < n; p++) {
    double xp[3] = {x[p], x[n + p], x[2 * n + p]};
    double vp[3] = {v[p], v[n + p], v[2 * n + p]};
    int s = spec[p];
    if (mover == ORC_MOVER_RELATIVISTIC)
      code[p] = push_rel_one(m, faces, E_corner, B_corner, charge[s], mass[s], dt_spec[s], bc_mode,
                             r_sphere, xp, vp, &blk[p], &cell[p]);
    else
      code[p] = push_one(m, faces, E_corner, B_corner, charge[s], mass[s], dt_spec[s], mover,
                         bc_mode, r_sphere, xp, vp, &blk[p], &cell[p]);
    for (int d = 0; d < 3; d++) {
      x[d * n + p] = xp[d];
      v[d * n + p] = vp[d];
    }
  }
  return 0;
}

/* ===================================================================== */
/* ECSIM deposition                                                       */
/* ===================================================================== */
static int cmp_i64(const void* a, const void* b) {
  int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
  return (x > y) - (x < y);
}

/* node lattice key of local corner (i,j,k) of block node n */
static int64_t corner_key(const orc_mesh* m, const orc_tnode* n, int i, int j, int k) {
  int64_t S = ((int64_t)1 << m->lmax) + 1;
  int64_t cs = n->size_g / ORC_NC;
  int64_t gx = n->xmin_g[0] + i * cs, gy = n->xmin_g[1] + j * cs, gz = n->xmin_g[2] + k * cs;
  return (gz * S + gy) * S + gx;
}

int64_t orc_build_nodes(const orc_mesh* m, int* node_of_corner, int64_t* keys_out) {
  int64_t total = (int64_t)m->nblocks * ORC_NCORNER;
  int64_t* keys = (int64_t*)malloc(sizeof(int64_t) * (size_t)total);
  for (int b = 0; b < m->nblocks; b++) {
    const orc_tnode* n = &m->t[m->block_node[b]];
    for (int k = 0; k < ORC_NN; k++)
      for (int j = 0; j < ORC_NN; j++)
        for (int i = 0; i < ORC_NN; i++)
          keys[(int64_t)b * ORC_NCORNER + i + ORC_NN * (j + ORC_NN * k)] = corner_key(m, n, i, j, k);
  }
  int64_t* sorted = (int64_t*)malloc(sizeof(int64_t) * (size_t)total);
  memcpy(sorted, keys, sizeof(int64_t) * (size_t)total);
  qsort(sorted, (size_t)total, sizeof(int64_t), cmp_i64);
  int64_t nu = 0;
  for (int64_t i = 0; i < total; i++)
    if (i == 0 || sorted[i] != sorted[i - 1]) sorted[nu++] = sorted[i];
  for (int64_t i = 0; i < total; i++) {
    int64_t lo = 0, hi = nu - 1;
    while (lo < hi) {
      int64_t mid = (lo + hi) / 2;
      if (sorted[mid] < keys[i]) lo = mid + 1;
      else hi = mid;
    }
    node_of_corner[i] = (int)lo;
  }
  if (keys_out) memcpy(keys_out, sorted, sizeof(int64_t) * (size_t)nu);
  free(keys);
  free(sorted);
  return nu;
}

/* IndexMatrix — pic_field_solver_ecsim.cpp:1377-1379 */
static const int IndexMatrix[8][8] = {{0, 2, 8, 6, 18, 20, 26, 24}, {1, 0, 6, 7, 19, 18, 24, 25},
                                      {4, 3, 0, 1, 22, 21, 18, 19}, {3, 5, 2, 0, 21, 23, 20, 18},
                                      {9, 11, 17, 15, 0, 2, 8, 6},  {10, 9, 15, 16, 1, 0, 6, 7},
                                      {13, 12, 9, 10, 4, 3, 0, 1},  {12, 14, 11, 9, 3, 5, 2, 0}};
/* cell corner c (ProcessCell offset[] order, :2033-2040) -> (di,dj,dk) */
static const int CornerOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                    {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

#define CELL_VALS (24 + 36 * 9)

/* Per-cell accumulation: ProcessCell scalar path, pic_field_solver_ecsim.cpp:1882-2438
 * (normalized units: B_conv = length_conv = LightSpeed = 1, :193-206). */
static void process_cell(const orc_mesh* m, int tn, const double* Bb, const double* charge,
                         const double* mass, const double* weight, double dtTotal,
                         int64_t np, const int64_t* idx, int64_t ntot, const double* x,
                         const double* v, const int* spec, double* out) {
  const orc_tnode* n = &m->t[tn];
  double Jg[8][3], MM[8][8][9];
  memset(Jg, 0, sizeof(Jg));
  memset(MM, 0, sizeof(MM));
  double dxc[3], CellVolume = 1;
  for (int d = 0; d < 3; d++) dxc[d] = (n->xmax[d] - n->xmin[d]) / ORC_NC * 1.0;
  for (int d = 0; d < 3; d++) CellVolume *= dxc[d];
  for (int64_t q = 0; q < np; q++) {
    int64_t p = idx[q];
    double xInit[3] = {x[p], x[ntot + p], x[2 * ntot + p]};
    double vInit[3] = {v[p], v[ntot + p], v[2 * ntot + p]};
    int s = spec[p];
    double w[8], table[8], B[3] = {0.0, 0.0, 0.0};
    int cr[8];
    corner_stencil_node(n, xInit, w, table, cr);
    if (m->Bc) {
      center_B_at(m, xInit, tn, B); /* :2097-2129, centre-based B */
    } else {
      for (int l = 0; l < 8; l++)
        for (int d = 0; d < 3; d++) B[d] += w[l] * Bb[3 * cr[l] + d];
    }
    double chargeQ = charge[s], m_ = mass[s];
    chargeQ *= weight[s];
    m_ *= weight[s];
    double QdT_over_m = chargeQ * dtTotal / m_;
    double QdT_over_2m = 0.5 * QdT_over_m;
    double QdT_over_2m_squared = QdT_over_2m * QdT_over_2m;
    double BB[3][3], P[3], alpha[9];
    for (int ii = 0; ii < 3; ii++) {
      P[ii] = -QdT_over_2m * B[ii];
      for (int jj = 0; jj <= ii; jj++) {
        BB[ii][jj] = QdT_over_2m_squared * B[ii] * B[jj];
        BB[jj][ii] = BB[ii][jj];
      }
    }
    double c0 = 1.0 / (1.0 + QdT_over_2m_squared * (B[0] * B[0] + B[1] * B[1] + B[2] * B[2]));
    alpha[0] = c0 * (1.0 + BB[0][0]);
    alpha[1] = c0 * (-P[2] + BB[0][1]);
    alpha[2] = c0 * (P[1] + BB[0][2]);
    alpha[3] = c0 * (P[2] + BB[1][0]);
    alpha[4] = c0 * (1.0 + BB[1][1]);
    alpha[5] = c0 * (-P[0] + BB[1][2]);
    alpha[6] = c0 * (-P[1] + BB[2][0]);
    alpha[7] = c0 * (P[0] + BB[2][1]);
    alpha[8] = c0 * (1.0 + BB[2][2]);
    /* the deposit weights are the raw table from a second InitStencil call */
    corner_stencil_node(n, xInit, w, table, cr);
    double vRot[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; i++)
      for (int jj = 0; jj < 3; jj++) vRot[i] += alpha[3 * i + jj] * vInit[jj];
    for (int c = 0; c < 8; c++) {
      double t = chargeQ * table[c];
      for (int d = 0; d < 3; d++) Jg[c][d] += t * vRot[d];
    }
    double matrixConst = chargeQ * QdT_over_2m / CellVolume;
    for (int i = 0; i < 8; i++) {
      double tempWeightConst = matrixConst * table[i];
      for (int j = 0; j <= i; j++) {
        double tempWeightProduct = table[j] * tempWeightConst;
        for (int k = 0; k < 9; k++) MM[i][j][k] += alpha[k] * tempWeightProduct;
      }
    }
  }
  /* out: 8x3 J already divided by the volume, then 36 pairs (i>=j) x 9 */
  for (int c = 0; c < 8; c++)
    for (int d = 0; d < 3; d++) out[3 * c + d] = Jg[c][d] / CellVolume;
  int pr = 0;
  for (int i = 0; i < 8; i++)
    for (int j = 0; j <= i; j++, pr++)
      for (int k = 0; k < 9; k++) out[24 + 9 * pr + k] = MM[i][j][k];
}

int orc_deposit(const orc_mesh* m, const double* B_corner, const double* charge,
                const double* mass, const double* weight, double dt, int64_t n, const double* x,
                const double* v, const int* spec, const int* blk, const int* cell,
                const int* node_of_corner, int64_t nnodes, double* J, double* M, int nthreads) {
  int64_t ncells = (int64_t)m->nblocks * ORC_NCELL;
  /* stable counting sort of particles by (block, cell) */
  int64_t* start = (int64_t*)calloc((size_t)ncells + 1, sizeof(int64_t));
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t p = 0; p < n; p++) start[(int64_t)blk[p] * ORC_NCELL + cell[p] + 1]++;
  for (int64_t c = 0; c < ncells; c++) start[c + 1] += start[c];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncells);
  memcpy(fill, start, sizeof(int64_t) * (size_t)ncells);
  for (int64_t p = 0; p < n; p++) idx[fill[(int64_t)blk[p] * ORC_NCELL + cell[p]]++] = p;
  free(fill);
  memset(J, 0, sizeof(double) * (size_t)nnodes * 3);
  memset(M, 0, sizeof(double) * (size_t)nnodes * 243);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const int64_t CH = 16384;
  double* buf = (double*)malloc(sizeof(double) * (size_t)CH * CELL_VALS);
  for (int64_t c0 = 0; c0 < ncells; c0 += CH) {
    int64_t c1 = (c0 + CH < ncells) ? c0 + CH : ncells;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 64)
#endif
    for (int64_t c = c0; c < c1; c++) {
      int64_t np = start[c + 1] - start[c];
      if (np == 0) continue;
      int b = (int)(c / ORC_NCELL);
      process_cell(m, m->block_node[b], B_corner + (size_t)b * ORC_NCORNER * 3, charge, mass, weight, dt, np,
                   idx + start[c], n, x, v, spec, buf + (c - c0) * CELL_VALS);
    }
    /* corner flush in cell order (:2362-2369, :2398-2426) */
    for (int64_t c = c0; c < c1; c++) {
      if (start[c + 1] == start[c]) continue;
      const double* o = buf + (c - c0) * CELL_VALS;
      int b = (int)(c / ORC_NCELL), lc = (int)(c % ORC_NCELL);
      int ci = lc % ORC_NC, cj = (lc / ORC_NC) % ORC_NC, ck = lc / (ORC_NC * ORC_NC);
      int nodeid[8];
      for (int q = 0; q < 8; q++) {
        int li = ci + CornerOff[q][0], lj = cj + CornerOff[q][1], lk = ck + CornerOff[q][2];
        nodeid[q] = node_of_corner[(int64_t)b * ORC_NCORNER + li + ORC_NN * (lj + ORC_NN * lk)];
      }
      for (int q = 0; q < 8; q++)
        for (int d = 0; d < 3; d++) J[(int64_t)nodeid[q] * 3 + d] += o[3 * q + d];
      int pr = 0;
      for (int i = 0; i < 8; i++)
        for (int j = 0; j <= i; j++, pr++) {
          const double* mm = o + 24 + 9 * pr;
          if (i == j) {
            double* cm = M + (int64_t)nodeid[i] * 243;
            for (int k = 0; k < 9; k++) cm[k] += mm[k];
          } else {
            double* mi = M + (int64_t)nodeid[i] * 243 + 9 * IndexMatrix[i][j];
            double* mj = M + (int64_t)nodeid[j] * 243 + 9 * IndexMatrix[j][i];
            for (int k = 0; k < 9; k++) {
              mi[k] += mm[k];
              mj[k] += mm[k];
            }
          }
        }
    }
  }
  free(buf);
  free(start);
  free(idx);
  return 0;
}

/* The same deposit arranged as the reference runs it in parallel
 * (UpdateJMassMatrix, pic_field_solver_ecsim.cpp:3585-3625 builds the sets,
 * the cell loop runs them one after the other): the cells fall into eight
 * sets by the parity of their global lattice index, and the cells of one set
 * share no corner node, so each thread adds a finished cell straight into
 * its corners.  The reference builds the sets on one resolution level; here
 * the sets are formed per level as well, so cells of different sizes that
 * share a node never run together.  Only cells holding particles are
 * visited, and only the rows of their corner nodes are zeroed and written:
 * J and M rows of other nodes are left as they are.  The per-node summation
 * order is the set order, not the cell order of orc_deposit, so the results
 * agree with it up to rounding.  This is the CPU baseline's deposit. */
int orc_deposit_colored(const orc_mesh* m, const double* B_corner, const double* charge, const double* mass,
                        const double* weight, double dt, int64_t n, const double* x, const double* v,
                        const int* spec, const int* blk, const int* cell, const int* node_of_corner,
                        int64_t nnodes, double* J, double* M, int nthreads) {
  (void)nnodes;
  const int64_t ncells = (int64_t)m->nblocks * ORC_NCELL;
  int64_t* start = (int64_t*)calloc((size_t)ncells + 1, sizeof(int64_t));
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t p = 0; p < n; p++) start[(int64_t)blk[p] * ORC_NCELL + cell[p] + 1]++;
  for (int64_t c = 0; c < ncells; c++) start[c + 1] += start[c];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncells);
  memcpy(fill, start, sizeof(int64_t) * (size_t)ncells);
  for (int64_t p = 0; p < n; p++) idx[fill[(int64_t)blk[p] * ORC_NCELL + cell[p]]++] = p;
  free(fill);
  /* the occupied cells, grouped by (level, parity set) */
  int lmin = 1 << 30, lmax = -1;
  for (int b = 0; b < m->nblocks; b++) {
    const int L = m->t[m->block_node[b]].level;
    lmin = L < lmin ? L : lmin;
    lmax = L > lmax ? L : lmax;
  }
  const int nlev = lmax >= lmin ? lmax - lmin + 1 : 0, nset = 8 * nlev;
  int64_t* sstart = (int64_t*)calloc((size_t)nset + 1, sizeof(int64_t));
  int64_t nocc = 0;
  int* cset = (int*)malloc(sizeof(int) * (size_t)(ncells > 0 ? ncells : 1));
  for (int64_t c = 0; c < ncells; c++) {
    cset[c] = -1;
    if (start[c + 1] == start[c]) continue;
    const int b = (int)(c / ORC_NCELL), lc = (int)(c % ORC_NCELL);
    const orc_tnode* tn = &m->t[m->block_node[b]];
    const int64_t cs = tn->size_g / ORC_NC;
    const int64_t gi = tn->xmin_g[0] / cs + lc % ORC_NC, gj = tn->xmin_g[1] / cs + (lc / ORC_NC) % ORC_NC,
                  gk = tn->xmin_g[2] / cs + lc / (ORC_NC * ORC_NC);
    cset[c] = 8 * (tn->level - lmin) + (int)((gi & 1) + 2 * (gj & 1) + 4 * (gk & 1));
    sstart[cset[c] + 1]++;
    nocc++;
  }
  for (int s2 = 0; s2 < nset; s2++) sstart[s2 + 1] += sstart[s2];
  int64_t* sfill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nset > 0 ? nset : 1));
  memcpy(sfill, sstart, sizeof(int64_t) * (size_t)nset);
  int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nocc > 0 ? nocc : 1));
  for (int64_t c = 0; c < ncells; c++)
    if (cset[c] >= 0) order[sfill[cset[c]]++] = c;
  free(sfill);
  free(cset);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  for (int s2 = 0; s2 < nset; s2++) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t e = sstart[s2]; e < sstart[s2 + 1]; e++) {
      const int64_t c = order[e];
      const int b = (int)(c / ORC_NCELL), lc = (int)(c % ORC_NCELL);
      const int ci = lc % ORC_NC, cj = (lc / ORC_NC) % ORC_NC, ck = lc / (ORC_NC * ORC_NC);
      for (int q = 0; q < 8; q++) {
        const int node = node_of_corner[(int64_t)b * ORC_NCORNER + (ci + CornerOff[q][0]) +
                                        ORC_NN * ((cj + CornerOff[q][1]) + ORC_NN * (ck + CornerOff[q][2]))];
        memset(J + (int64_t)node * 3, 0, sizeof(double) * 3);
        memset(M + (int64_t)node * 243, 0, sizeof(double) * 243);
      }
    }
  }
  for (int s2 = 0; s2 < nset; s2++) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int64_t e = sstart[s2]; e < sstart[s2 + 1]; e++) {
      double o[CELL_VALS];
      const int64_t c = order[e];
      const int b = (int)(c / ORC_NCELL), lc = (int)(c % ORC_NCELL);
      process_cell(m, m->block_node[b], B_corner + (size_t)b * ORC_NCORNER * 3, charge, mass, weight, dt,
                   start[c + 1] - start[c], idx + start[c], n, x, v, spec, o);
      const int ci = lc % ORC_NC, cj = (lc / ORC_NC) % ORC_NC, ck = lc / (ORC_NC * ORC_NC);
      int nodeid[8];
      for (int q = 0; q < 8; q++)
        nodeid[q] = node_of_corner[(int64_t)b * ORC_NCORNER + (ci + CornerOff[q][0]) +
                                   ORC_NN * ((cj + CornerOff[q][1]) + ORC_NN * (ck + CornerOff[q][2]))];
      for (int q = 0; q < 8; q++)
        for (int d = 0; d < 3; d++) J[(int64_t)nodeid[q] * 3 + d] += o[3 * q + d];
      int pr = 0;
      for (int i = 0; i < 8; i++)
        for (int j = 0; j <= i; j++, pr++) {
          const double* mm = o + 24 + 9 * pr;
          if (i == j) {
            double* cm = M + (int64_t)nodeid[i] * 243;
            for (int k = 0; k < 9; k++) cm[k] += mm[k];
          } else {
            double* mi = M + (int64_t)nodeid[i] * 243 + 9 * IndexMatrix[i][j];
            double* mj = M + (int64_t)nodeid[j] * 243 + 9 * IndexMatrix[j][i];
            for (int k = 0; k < 9; k++) {
              mi[k] += mm[k];
              mj[k] += mm[k];
            }
          }
        }
    }
  }
  free(order);
  free(sstart);
  free(start);
  free(idx);
  return 0;
}

/* Particle energy and CFL numbers of UpdateJMassMatrix (pic_field_solver_
 * ecsim.cpp:2229-2238 per particle, :2355-2358 per cell, :3925-3929 and
 * :3972-3987 the reductions).  Per cell, in the order the particles are given
 * (grouped by (block, cell)): ParticleEnergyCell += 0.5 mass vsqr with the
 * weighted mass, vmean[spec] += sqrt(vsqr) dt, and cflCell[spec] =
 * vmean[spec] / (count[spec] sqrt(dx^2+dy^2+dz^2)).  The totals are reduced
 * in a fixed order: each block's cells in cell order, then the blocks in
 * order (energy), and the maximum over cells of each species' cflCell
 * starting from 0, a cell without particles of the species (0/0) never
 * winning, as in the reference's comparison.  cell_energy[ncells] and
 * cell_cfl[ncells][nspec] may be NULL. */
int orc_energy_cfl(const orc_mesh* m, const double* mass, const double* weight, int nspec, double dt,
                   int64_t n, const double* v, const int* spec, const int* blk, const int* cell,
                   double* energy, double* cfl, double* cell_energy, double* cell_cfl) {
  const int64_t ncells = (int64_t)m->nblocks * ORC_NCELL;
  int64_t* start = (int64_t*)calloc((size_t)ncells + 1, sizeof(int64_t));
  int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  for (int64_t p = 0; p < n; p++) start[(int64_t)blk[p] * ORC_NCELL + cell[p] + 1]++;
  for (int64_t c = 0; c < ncells; c++) start[c + 1] += start[c];
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncells);
  memcpy(fill, start, sizeof(int64_t) * (size_t)ncells);
  for (int64_t p = 0; p < n; p++) idx[fill[(int64_t)blk[p] * ORC_NCELL + cell[p]]++] = p;
  free(fill);
  for (int s = 0; s < nspec; s++) cfl[s] = 0.0;
  double total = 0.0;
  for (int b = 0; b < m->nblocks; b++) {
    const orc_tnode* t = &m->t[m->block_node[b]];
    double dx[3];
    for (int d = 0; d < 3; d++) dx[d] = (t->xmax[d] - t->xmin[d]) / ORC_NC;
    const double ldx = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
    double eblk = 0.0;
    for (int lc = 0; lc < ORC_NCELL; lc++) {
      const int64_t c = (int64_t)b * ORC_NCELL + lc;
      double ecell = 0.0, vmean[ORC_MAX_SPEC];
      int cnt[ORC_MAX_SPEC];
      for (int s = 0; s < nspec; s++) vmean[s] = 0.0, cnt[s] = 0;
      for (int64_t q = start[c]; q < start[c + 1]; q++) {
        const int64_t p = idx[q];
        const int s = spec[p];
        const double vx = v[p], vy = v[n + p], vz = v[2 * n + p];
        const double vsqr = vx * vx + vy * vy + vz * vz;
        const double ms = mass[s] * weight[s];
        vmean[s] += sqrt(vsqr) * dt;
        ecell += 0.5 * ms * vsqr;
        cnt[s]++;
      }
      eblk += ecell;
      if (cell_energy) cell_energy[c] = ecell;
      for (int s = 0; s < nspec; s++) {
        const double cc = vmean[s] / (cnt[s] * ldx);
        if (cell_cfl) cell_cfl[c * nspec + s] = cc;
        if (cfl[s] < cc) cfl[s] = cc;
      }
    }
    total += eblk;
  }
  *energy = total;
  free(start);
  free(idx);
  return 0;
}

/* ===================================================================== */
/* Gridless cutoff tracer (SI)                                            */
/* ===================================================================== */
#define ORC_C 299792458.0 /* SpeedOfLight, constants.h:40 */

static inline double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static inline double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */
static inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* Dipole::GetB_Tesla — DipoleInterface.h:99-125, plus a uniform field */
void orc_dipole_B(const orc_trace_params* P, const double xm[3], double B[3]) {
  const double mu0_over_4pi = 1.0e-7;
  double x = xm[0], y = xm[1], z = xm[2];
  double r2 = x * x + y * y + z * z;
  if (r2 <= 0.0) {
    B[0] = B[1] = B[2] = 0.0;
  } else {
    double r = sqrt(r2), r3 = r2 * r, r5 = r3 * r2;
    double M = P->moment_Am2;
    double mx = M * P->m_hat[0], my = M * P->m_hat[1], mz = M * P->m_hat[2];
    double mdotr = mx * x + my * y + mz * z;
    B[0] = mu0_over_4pi * (3.0 * x * mdotr / r5 - mx / r3);
    B[1] = mu0_over_4pi * (3.0 * y * mdotr / r5 - my / r3);
    B[2] = mu0_over_4pi * (3.0 * z * mdotr / r5 - mz / r3);
  }
  B[0] = B[0] + P->B_ext_T[0];
  B[1] = B[1] + P->B_ext_T[1];
  B[2] = B[2] + P->B_ext_T[2];
}

/* GammaFromMomentum — GridlessParticleMovers.cpp:111-115 */
static inline double gamma_p(const double p[3], double m0) {
  double p2 = dot3(p, p);
  double mc = m0 * ORC_C;
  return sqrt(1.0 + p2 / (mc * mc));
}

static void trace_field(const orc_trace_params* P, const double x[3], double B[3]);

/* BorisStep — GridlessParticleMovers.cpp:781-801 */
void orc_boris_step(const orc_trace_params* P, double x[3], double p[3], double dt) {
  double B[3];
  trace_field(P, x, B);
  double m0 = P->m0_kg, q = P->q_C;
  double gam = gamma_p(p, m0);
  double iv = 1.0 / (gam * m0);
  double vv[3] = {iv * p[0], iv * p[1], iv * p[2]};
  double h = 0.5 * dt;
  for (int d = 0; d < 3; d++) x[d] = x[d] + h * vv[d];
  double tc = (q * dt) / (2.0 * gam * m0);
  double t[3] = {tc * B[0], tc * B[1], tc * B[2]};
  double t2 = dot3(t, t);
  double sc = 2.0 / (1.0 + t2);
  double s[3] = {sc * t[0], sc * t[1], sc * t[2]};
  double pm[3] = {p[0], p[1], p[2]};
  double pp[3] = {pm[0] + (pm[1] * t[2] - pm[2] * t[1]), pm[1] + (pm[2] * t[0] - pm[0] * t[2]),
                  pm[2] + (pm[0] * t[1] - pm[1] * t[0])};
  p[0] = pm[0] + (pp[1] * s[2] - pp[2] * s[1]);
  p[1] = pm[1] + (pp[2] * s[0] - pp[0] * s[2]);
  p[2] = pm[2] + (pp[0] * s[1] - pp[1] * s[0]);
  double gam2 = gamma_p(p, m0);
  double iv2 = 1.0 / (gam2 * m0);
  double v2[3] = {iv2 * p[0], iv2 * p[1], iv2 * p[2]};
  for (int d = 0; d < 3; d++) x[d] = x[d] + h * v2[d];
}

/* InsideInnerSphereM / SegmentHitsInnerSphere — GridlessParticleMovers.cpp:132-149 */
static inline int inside_inner_m(const double x[3], double r) { return (r > 0.0) && (dot3(x, x) <= r * r); }
static int segment_hits_inner(const double a[3], const double b[3], double r) {
  if (r <= 0.0) return 0;
  if (inside_inner_m(a, r) || inside_inner_m(b, r)) return 1;
  double d[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  double dd = dot3(d, d);
  if (dd <= 0.0) return 0;
  double t = -dot3(a, d) / dd;
  if (t < 0.0) t = 0.0;
  if (t > 1.0) t = 1.0;
  double q[3] = {a[0] + t * d[0], a[1] + t * d[1], a[2] + t * d[2]};
  return dot3(q, q) <= r * r;
}

/* HigueraCaryMagneticMidpointStep — GridlessParticleMovers.cpp:723-757: half
 * drift with v(p_n), B at the midpoint, Boris/Cayley rotation with gamma(p_n),
 * half drift with v(p_{n+1}). */
static void hc_midpoint_step(const orc_trace_params* P, double x[3], double p[3], double dt) {
  double m0 = P->m0_kg, q = P->q_C;
  double g0 = gamma_p(p, m0);
  double iv0 = 1.0 / (g0 * m0);
  double v0[3] = {iv0 * p[0], iv0 * p[1], iv0 * p[2]};
  double h = 0.5 * dt;
  double xm[3] = {x[0] + h * v0[0], x[1] + h * v0[1], x[2] + h * v0[2]};
  double B[3];
  trace_field(P, xm, B);
  double gamma0 = gamma_p(p, m0);
  double tc = (q * dt) / (2.0 * gamma0 * m0);
  double t[3] = {tc * B[0], tc * B[1], tc * B[2]};
  double t2 = dot3(t, t);
  double sc = 2.0 / (1.0 + t2);
  double s[3] = {sc * t[0], sc * t[1], sc * t[2]};
  double pm[3] = {p[0], p[1], p[2]};
  double pp[3] = {pm[0] + (pm[1] * t[2] - pm[2] * t[1]), pm[1] + (pm[2] * t[0] - pm[0] * t[2]),
                  pm[2] + (pm[0] * t[1] - pm[1] * t[0])};
  p[0] = pm[0] + (pp[1] * s[2] - pp[2] * s[1]);
  p[1] = pm[1] + (pp[2] * s[0] - pp[0] * s[2]);
  p[2] = pm[2] + (pp[0] * s[1] - pp[1] * s[0]);
  double g1 = gamma_p(p, m0);
  double iv1 = 1.0 / (g1 * m0);
  double v1[3] = {iv1 * p[0], iv1 * p[1], iv1 * p[2]};
  for (int d = 0; d < 3; d++) x[d] = xm[d] + h * v1[d];
}

/* HC4StepUnchecked — GridlessParticleMovers.cpp:759-779: Yoshida triple jump
 * H2(w1 dt) H2(w0 dt) H2(w1 dt), w1 = 1/(2 - 2^(1/3)), w0 = -2^(1/3)/(2 - 2^(1/3)). */
void orc_hc4_step(const orc_trace_params* P, double x[3], double p[3], double dt) {
  const double cbrt2 = 1.25992104989487316477;
  const double w1 = 1.0 / (2.0 - cbrt2);
  const double w0 = -cbrt2 / (2.0 - cbrt2);
  hc_midpoint_step(P, x, p, w1 * dt);
  hc_midpoint_step(P, x, p, w0 * dt);
  hc_midpoint_step(P, x, p, w1 * dt);
}

/* LorentzRhs — GridlessParticleMovers.cpp:151-157: (dx/dt, dp/dt) = (v, q v x B). */
static void lorentz_rhs(const orc_trace_params* P, const double x[3], const double p[3], double kx[3],
                        double kp[3]) {
  double B[3];
  trace_field(P, x, B);
  double iv = 1.0 / (gamma_p(p, P->m0_kg) * P->m0_kg);
  double v[3] = {iv * p[0], iv * p[1], iv * p[2]};
  double q = P->q_C;
  kx[0] = v[0], kx[1] = v[1], kx[2] = v[2];
  kp[0] = q * (v[1] * B[2] - v[2] * B[1]);
  kp[1] = q * (v[2] * B[0] - v[0] * B[2]);
  kp[2] = q * (v[0] * B[1] - v[1] * B[0]);
}

/* Butcher tables of StepParticleChecked's RK2 (midpoint), RK4 (classical)
 * and RK6 (Butcher's seven-stage sixth-order method),
 * GridlessParticleMovers.cpp:1061-1098.  a is ns x ns, row-major. */
static const double RK2_A[4] = {0.0, 0.0, 0.5, 0.0};
static const double RK2_B[2] = {0.0, 1.0};
static const double RK4_A[16] = {0.0, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 0.5, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
static const double RK4_B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
static const double RK6_A[49] = {
    0.0,         0.0,        0.0,          0.0,         0.0, 0.0,          0.0,
    1.0 / 3.0,   0.0,        0.0,          0.0,         0.0, 0.0,          0.0,
    0.0,         2.0 / 3.0,  0.0,          0.0,         0.0, 0.0,          0.0,
    1.0 / 12.0,  1.0 / 3.0,  -1.0 / 12.0,  0.0,         0.0, 0.0,          0.0,
    -1.0 / 16.0, 9.0 / 8.0,  -3.0 / 16.0,  -3.0 / 8.0,  0.0, 0.0,          0.0,
    0.0,         9.0 / 8.0,  -3.0 / 8.0,   -3.0 / 4.0,  0.5, 0.0,          0.0,
    9.0 / 44.0,  -9.0 / 11.0, 63.0 / 44.0, 18.0 / 11.0, 0.0, -16.0 / 11.0, 0.0};
static const double RK6_B[7] = {11.0 / 120.0, 0.0, 27.0 / 40.0, 27.0 / 40.0, -4.0 / 15.0, -4.0 / 15.0, 11.0 / 120.0};

/* RKStepGeneric — GridlessParticleMovers.cpp:655-692: explicit RK with the
 * chord between consecutive stage positions and the final chord checked
 * against the inner sphere.  On a stage chord hit x takes the stage position
 * and p is left at its start value; on the final chord both take y1. */
static int rk_step_checked(const orc_trace_params* P, int ns, const double* a, const double* b, double x[3],
                           double p[3], double dt, double rIn) {
  double kx[7][3], kp[7][3];
  double y0x[3] = {x[0], x[1], x[2]}, y0p[3] = {p[0], p[1], p[2]};
  double prev[3] = {x[0], x[1], x[2]};
  for (int i = 0; i < ns; i++) {
    double yx[3] = {y0x[0], y0x[1], y0x[2]}, yp[3] = {y0p[0], y0p[1], y0p[2]};
    for (int j = 0; j < i; j++) {
      double c = dt * a[i * ns + j];
      for (int d = 0; d < 3; d++) {
        yx[d] = yx[d] + c * kx[j][d];
        yp[d] = yp[d] + c * kp[j][d];
      }
    }
    if (segment_hits_inner(prev, yx, rIn)) {
      for (int d = 0; d < 3; d++) x[d] = yx[d];
      return 0;
    }
    if (inside_inner_m(yx, rIn)) {
      for (int d = 0; d < 3; d++) x[d] = yx[d], p[d] = yp[d];
      return 0;
    }
    for (int d = 0; d < 3; d++) prev[d] = yx[d];
    lorentz_rhs(P, yx, yp, kx[i], kp[i]);
  }
  double y1x[3] = {y0x[0], y0x[1], y0x[2]}, y1p[3] = {y0p[0], y0p[1], y0p[2]};
  for (int i = 0; i < ns; i++) {
    double c = dt * b[i];
    for (int d = 0; d < 3; d++) {
      y1x[d] = y1x[d] + c * kx[i][d];
      y1p[d] = y1p[d] + c * kp[i][d];
    }
  }
  for (int d = 0; d < 3; d++) x[d] = y1x[d], p[d] = y1p[d];
  if (segment_hits_inner(prev, y1x, rIn)) return 0;
  if (inside_inner_m(y1x, rIn)) return 0;
  return 1;
}

/* StepParticleChecked — GridlessParticleMovers.cpp:1035-1098 for the
 * full-orbit movers.  BORIS: one step, chord checked.  HC4: four positive
 * macro-steps of dt/4, each chord checked (the composition's negative inner
 * stage is never classified). */
int orc_step_checked(const orc_trace_params* P, double x[3], double p[3], double dt, double rIn) {
  switch (P->mover) {
    case 1: {
      const double dtSub = dt / 4.0;
      for (int i = 0; i < 4; i++) {
        double x0[3] = {x[0], x[1], x[2]};
        orc_hc4_step(P, x, p, dtSub);
        if (segment_hits_inner(x0, x, rIn) || inside_inner_m(x, rIn)) return 0;
      }
      return 1;
    }
    case 2: return rk_step_checked(P, 2, RK2_A, RK2_B, x, p, dt, rIn);
    case 3: return rk_step_checked(P, 4, RK4_A, RK4_B, x, p, dt, rIn);
    case 4: return rk_step_checked(P, 7, RK6_A, RK6_B, x, p, dt, rIn);
    default: {
      double x0[3] = {x[0], x[1], x[2]};
      orc_boris_step(P, x, p, dt);
      return !segment_hits_inner(x0, x, rIn) && !inside_inner_m(x, rIn);
    }
  }
}

/* SelectLegacyCutoffTraceDt — CutoffRigidityGridless.cpp:1086-1152.  The
 * gyro budget, limiter and floor are switchable; with outer_sphere the
 * boundary distance is the distance to the outer sphere. */
static double select_dt(const orc_trace_params* P, const double x[3], const double p[3],
                        double dtTrace, double timeRemaining) {
  double dt = dtTrace;
  if (timeRemaining < dt) dt = timeRemaining;
  if (!P->adaptive) return dt;
  double m0 = P->m0_kg;
  double p2 = dot3(p, p);
  double mc = m0 * ORC_C;
  double gamma = sqrt(1.0 + p2 / (mc * mc));
  double pMag = sqrt(dmax(0.0, p2));
  double vMag = (gamma > 0.0 && m0 > 0.0) ? pMag / (gamma * m0) : 0.0;
  double B[3];
  trace_field(P, x, B);
  double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    double omega = fabs(P->q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P->dphi_max / omega);
  }
  double vFor = vMag;
  if (P->dist_limiter) {
    double Re = P->earth_radius_m;
    double xRe[3] = {x[0] / Re, x[1] / Re, x[2] / Re};
    double rRe = sqrt(xRe[0] * xRe[0] + xRe[1] * xRe[1] + xRe[2] * xRe[2]);
    double dInner = 1.0e300;
    if (rRe > P->r_inner_Re) dInner = (rRe - P->r_inner_Re) * Re;
    double dBox = 1.0e300;
    if (P->outer_sphere) {
      if (rRe < P->r_outer_Re) dBox = (P->r_outer_Re - rRe) * Re;
    } else if (xRe[0] >= P->box_min_Re[0] && xRe[0] <= P->box_max_Re[0] &&
               xRe[1] >= P->box_min_Re[1] && xRe[1] <= P->box_max_Re[1] &&
               xRe[2] >= P->box_min_Re[2] && xRe[2] <= P->box_max_Re[2]) {
      double dxp = (P->box_max_Re[0] - xRe[0]) * Re, dxm = (xRe[0] - P->box_min_Re[0]) * Re;
      double dyp = (P->box_max_Re[1] - xRe[1]) * Re, dym = (xRe[1] - P->box_min_Re[1]) * Re;
      double dzp = (P->box_max_Re[2] - xRe[2]) * Re, dzm = (xRe[2] - P->box_min_Re[2]) * Re;
      dBox = dmin(dmin(dmin(dxp, dxm), dmin(dyp, dym)), dmin(dzp, dzm));
    }
    double dNear = dmin(dInner, dBox);
    if (vFor > 0.0 && dNear < 1.0e299) dt = dmin(dt, 0.20 * dNear / vFor);
  }
  if (P->dt_floor) {
    double safeSpeed = dmax(vFor, 1.0e-12);
    double dtFloor = dmax(dmax(1.0e-12, 1.0e-9 * dmax(dtTrace, 1.0)), 100.0e3 / safeSpeed);
    dt = dmax(dtFloor, dt);
  }
  if (timeRemaining > 0.0) dt = dmin(dt, timeRemaining);
  return dt;
}

/* SelectLegacyCutoffTraceDt3D — CutoffRigidityMode3D.cpp:881-936: the
 * gridless selector's rules with the box and inner-sphere distances taken in
 * metres (a point outside the box has no box limit; the inner distance is
 * max(0, |x| - r_inner)); the limiter and the floor are always on. */
static double select_dt3d(const orc_trace_params* P, const double x[3], const double p[3], double dtTrace,
                          double timeRemaining, const double boxmin[3], const double boxmax[3], double rInner) {
  double dt = dtTrace;
  if (timeRemaining < dt) dt = timeRemaining;
  if (!P->adaptive) return dt;
  double m0 = P->m0_kg;
  double p2 = dot3(p, p);
  double mc = m0 * ORC_C;
  double gamma = sqrt(1.0 + p2 / (mc * mc));
  double pMag = sqrt(dmax(0.0, p2));
  double vMag = (gamma > 0.0 && m0 > 0.0) ? pMag / (gamma * m0) : 0.0;
  double B[3];
  trace_field(P, x, B);
  double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    double omega = fabs(P->q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P->dphi_max / omega);
  }
  double dxp = boxmax[0] - x[0], dxm = x[0] - boxmin[0];
  double dyp = boxmax[1] - x[1], dym = x[1] - boxmin[1];
  double dzp = boxmax[2] - x[2], dzm = x[2] - boxmin[2];
  double dBox = dmin(dmin(dmin(dmin(dmin(dxp, dxm), dyp), dym), dzp), dzm);
  double dBoxSafe = (dBox < 0.0) ? 1.0e300 : dBox;
  double dInner = dmax(0.0, sqrt(dot3(x, x)) - rInner);
  double dNear = dmin(dBoxSafe, dInner);
  double vFor = vMag;
  if (vFor > 0.0 && dNear < 1.0e299) dt = dmin(dt, 0.20 * dNear / vFor);
  double safeSpeed = dmax(vFor, 1.0e-12);
  double dtFloor = dmax(dmax(1.0e-12, 1.0e-9 * dmax(dtTrace, 1.0)), 100.0e3 / safeSpeed);
  dt = dmax(dtFloor, dt);
  if (timeRemaining > 0.0) dt = dmin(dt, timeRemaining);
  return dt;
}

static int isfin3(const double a[3]) { return isfinite(a[0]) && isfinite(a[1]) && isfinite(a[2]); }

/* the two legacy step selectors, for their closed-form tests */
double orc_select_dt(const orc_trace_params* P, const double x[3], const double p[3], double dtTrace,
                     double timeRemaining) {
  return select_dt(P, x, p, dtTrace, timeRemaining);
}
double orc_select_dt3d(const orc_trace_params* P, const double x[3], const double p[3], double dtTrace,
                       double timeRemaining, const double boxmin[3], const double boxmax[3], double rInner) {
  return select_dt3d(P, x, p, dtTrace, timeRemaining, boxmin, boxmax, rInner);
}

/* FinalizeUpperBoundedStep — util/TrajectoryTimeStep.h:15-24 */
double orc_finalize_step(double dt, double timeRemaining) {
  if (!(dt > 0.0) || !isfinite(dt)) return NAN;
  if (timeRemaining > 0.0) dt = dmin(dt, timeRemaining);
  if (!(dt > 0.0) || !isfinite(dt)) return NAN;
  return dt;
}

/* SelectStructuredTraceDt — CutoffRigidityGridless.cpp:1154-1228: DT_TRACE
 * capped by the gyro-angle budget, no boundary limiter, no floor. */
static double select_dt_structured(const orc_trace_params* P, const double x[3], const double p[3],
                                   double dtTrace, double timeRemaining) {
  double dt = dtTrace;
  if (!P->adaptive) return orc_finalize_step(dt, timeRemaining);
  double m0 = P->m0_kg;
  double p2 = dot3(p, p);
  double mc = m0 * ORC_C;
  double gamma = sqrt(1.0 + p2 / (mc * mc));
  double B[3];
  trace_field(P, x, B);
  double Bmag = sqrt(dot3(B, B));
  if (Bmag > 0.0) {
    double omega = fabs(P->q_C) * Bmag / (gamma * m0);
    if (omega > 0.0) dt = dmin(dt, P->dphi_max / omega);
  }
  return orc_finalize_step(dt, timeRemaining);
}

/* TrajectoryBoundary::FindFirstEvent — util/TrajectoryBoundary.h:46-141:
 * the first inner-sphere or outer-box crossing on the chord start -> end
 * (type 0 none, 1 inner sphere, 2 outer box). */
void orc_find_first_event(const double a[3], const double z[3], const double bmin[3], const double bmax[3],
                          double rInner, double tol, int* type, double* fraction, double pos[3], int* face) {
  int bt = 0, bface = -1;
  double bf = INFINITY, bp[3] = {0.0, 0.0, 0.0};
  if (isfin3(a) && isfin3(z)) {
    const double d[3] = {z[0] - a[0], z[1] - a[1], z[2] - a[2]};
    /* ConsiderInnerSphere :50-85 */
    if (rInner > 0.0) {
      const double r0 = dmax(0.0, rInner + 0.0);
      if (a[0] * a[0] + a[1] * a[1] + a[2] * a[2] <= r0 * r0) {
        bt = 1, bf = 0.0, bface = -1;
        bp[0] = a[0], bp[1] = a[1], bp[2] = a[2];
      } else {
        const double aa = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (aa > 0.0) {
          const double segmentLength = sqrt(aa);
          const double fractionTol = (segmentLength > 0.0) ? dmax(0.0, tol) / segmentLength : 0.0;
          const double bb = 2.0 * (a[0] * d[0] + a[1] * d[1] + a[2] * d[2]);
          const double cc = a[0] * a[0] + a[1] * a[1] + a[2] * a[2] - rInner * rInner;
          double disc = bb * bb - 4.0 * aa * cc;
          if (!(disc < 0.0)) {
            disc = dmax(0.0, disc);
            const double root = sqrt(disc);
            const double cand[2] = {(-bb - root) / (2.0 * aa), (-bb + root) / (2.0 * aa)};
            for (int c = 0; c < 2; c++) {
              if (cand[c] < -fractionTol || cand[c] > 1.0 + fractionTol) continue;
              const double t = dmax(0.0, dmin(1.0, cand[c]));
              if (t >= bf) continue;
              bt = 1, bf = t, bface = -1;
              for (int i = 0; i < 3; i++) bp[i] = a[i] + bf * d[i];
            }
          }
        }
      }
    }
    /* ConsiderOuterBox :87-131 */
    const int ain = a[0] >= bmin[0] && a[0] <= bmax[0] && a[1] >= bmin[1] && a[1] <= bmax[1] &&
                    a[2] >= bmin[2] && a[2] <= bmax[2];
    if (!ain) {
      if (0.0 < bf) {
        bt = 2, bf = 0.0, bface = -1;
        bp[0] = a[0], bp[1] = a[1], bp[2] = a[2];
      }
    } else {
      const int zin = z[0] >= bmin[0] && z[0] <= bmax[0] && z[1] >= bmin[1] && z[1] <= bmax[1] &&
                      z[2] >= bmin[2] && z[2] <= bmax[2];
      if (!zin) {
        const double segmentLength = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double fractionTol = (segmentLength > 0.0) ? dmax(0.0, tol) / segmentLength : 0.0;
        for (int axis = 0; axis < 3; axis++) {
          if (d[axis] == 0.0) continue;
          const int side = (d[axis] > 0.0) ? 1 : 0;
          const double plane = side ? bmax[axis] : bmin[axis];
          const double cand = (plane - a[axis]) / d[axis];
          if (cand < -fractionTol || cand > 1.0 + fractionTol) continue;
          const double t = dmax(0.0, dmin(1.0, cand));
          if (t >= bf) continue;
          const double q[3] = {a[0] + t * d[0], a[1] + t * d[1], a[2] + t * d[2]};
          int onFace = 1;
          for (int j = 0; j < 3; j++) {
            if (j == axis) continue;
            if (q[j] < bmin[j] - tol || q[j] > bmax[j] + tol) {
              onFace = 0;
              break;
            }
          }
          if (!onFace) continue;
          bt = 2, bf = t, bface = 2 * axis + side;
          bp[0] = q[0], bp[1] = q[1], bp[2] = q[2];
        }
      }
    }
  }
  *type = bt, *fraction = bf, *face = bface;
  pos[0] = bp[0], pos[1] = bp[1], pos[2] = bp[2];
}

/* PopulateOuterExit — CutoffRigidityGridless.cpp:1323-1358: position of the
 * event, velocity direction from the momentum interpolated to the event
 * fraction, cos(pitch angle) against B a tolerance inside the box. */
static int populate_exit(const orc_trace_params* P, const double pos[3], double fraction, const double xb[3],
                         const double xa[3], const double pb[3], const double pa[3], orc_trace_exit* ex) {
  for (int d = 0; d < 3; d++) ex->x_exit_m[d] = pos[d];
  const double a = dmax(0.0, dmin(1.0, fraction));
  const double pc[3] = {pb[0] + a * (pa[0] - pb[0]), pb[1] + a * (pa[1] - pb[1]), pb[2] + a * (pa[2] - pb[2])};
  const double p2n = dot3(pc, pc);
  const double mc = P->m0_kg * ORC_C;
  const double g = sqrt(1.0 + p2n / (mc * mc));
  if (!(g > 0.0) || !isfinite(g)) return 0;
  const double iv = 1.0 / (g * P->m0_kg);
  double v[3] = {iv * pc[0], iv * pc[1], iv * pc[2]};
  const double n = sqrt(dot3(v, v));
  if (n > 0.0) {
    const double in = 1.0 / n;
    for (int d = 0; d < 3; d++) v[d] = in * v[d];
  } else {
    v[0] = v[1] = v[2] = 0.0;
  }
  for (int d = 0; d < 3; d++) ex->v_exit_unit[d] = v[d];
  const double ch[3] = {xa[0] - xb[0], xa[1] - xb[1], xa[2] - xb[2]};
  const double chl = sqrt(dot3(ch, ch));
  double inset = 0.0;
  if (chl > 0.0) inset = dmin(a, dmax(1.0, P->event_tol_m) / chl);
  const double f = dmax(0.0, a - inset);
  const double xe[3] = {xb[0] + f * ch[0], xb[1] + f * ch[1], xb[2] + f * ch[2]};
  double B[3];
  trace_field(P, xe, B);
  const double Bn = sqrt(dot3(B, B));
  if (!(Bn > 0.0) || !isfinite(Bn)) return 0;
  const double ib = 1.0 / Bn;
  const double bh[3] = {ib * B[0], ib * B[1], ib * B[2]};
  ex->cos_alpha = dot3(v, bh);
  return isfinite(ex->cos_alpha);
}

/* TrajectoryTrap::Detector — util/TrajectoryTrapDetector.h:23-153: counts
 * mirror points (sign changes of p.B beyond a deadband), two per bounce
 * cycle, and calls the orbit trapped once enough cycles have a stable radial
 * envelope, kept off the box by the margin, with |p| unchanged to the
 * tolerance. */
void orc_trap_init(orc_trap* t) {
  memset(t, 0, sizeof(*t));
  t->p_min = INFINITY;
  t->r_min = INFINITY;
  t->min_margin = INFINITY;
}

static double trap_margin(const orc_trace_params* P, const double x[3]) {
  double bmin[3], bmax[3];
  for (int d = 0; d < 3; d++) bmin[d] = P->box_min_Re[d] * P->earth_radius_m, bmax[d] = P->box_max_Re[d] * P->earth_radius_m;
  return dmin(dmin(dmin(x[0] - bmin[0], bmax[0] - x[0]), dmin(x[1] - bmin[1], bmax[1] - x[1])),
              dmin(x[2] - bmin[2], bmax[2] - x[2]));
}

int orc_trap_update(const orc_trace_params* P, orc_trap* t, const double x[3], const double p[3], const double B[3]) {
  if (!P->trap || t->trapped) return t->trapped;
  if (!isfin3(x) || !isfin3(p) || !isfin3(B)) return 0;
  const double pMag = sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  const double bMag = sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  const double r = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  if (!(pMag > 0.0) || !(bMag > 0.0) || !isfinite(r)) return 0;
  if (!t->initialized) {
    t->initialized = 1;
    t->p_ref = pMag, t->p_min = pMag, t->p_max = pMag;
    t->r_min = r, t->r_max = r;
    t->min_margin = dmin(t->min_margin, trap_margin(P, x));
  } else {
    t->p_min = dmin(t->p_min, pMag);
    t->p_max = dmax(t->p_max, pMag);
    t->r_min = dmin(t->r_min, r);
    t->r_max = dmax(t->r_max, r);
    t->min_margin = dmin(t->min_margin, trap_margin(P, x));
  }
  const double mu = (p[0] * B[0] + p[1] * B[1] + p[2] * B[2]) / (pMag * bMag);
  const int sign = (mu > P->trap_deadband) ? 1 : ((mu < -P->trap_deadband) ? -1 : 0);
  if (sign == 0) return t->trapped;
  if (t->last_sign == 0) {
    t->last_sign = sign;
    return t->trapped;
  }
  if (sign == t->last_sign) return t->trapped;
  t->last_sign = sign;
  ++t->mirror;
  if ((t->mirror % 2) != 0) return t->trapped;
  ++t->bounce;
  if (t->have_prev) {
    const double ch = dmax(fabs(t->r_min - t->prev_r_min), fabs(t->r_max - t->prev_r_max));
    if (ch <= P->trap_radial_tol_Re * P->earth_radius_m) ++t->stable;
    else t->stable = 0;
  }
  t->prev_r_min = t->r_min;
  t->prev_r_max = t->r_max;
  t->have_prev = 1;
  const double spread = (t->p_ref > 0.0) ? (t->p_max - t->p_min) / t->p_ref : INFINITY;
  t->trapped = (t->mirror >= P->trap_min_mirror && t->bounce >= P->trap_min_bounce && t->stable >= 2 &&
                t->min_margin >= P->trap_outer_margin_Re * P->earth_radius_m && spread <= P->trap_energy_tol);
  t->r_min = r, t->r_max = r;
  t->min_margin = dmin(t->min_margin, trap_margin(P, x));
  return t->trapped;
}

/* TraceTrajectoryImpl — CutoffRigidityGridless.cpp:1246-1457 (full-orbit
 * mover P->mover, trap detector off).  policy 0: the legacy cutoff policy
 * (SelectLegacyCutoffTraceDt, boundaries checked at the step start);
 * policy 1: StructuredAccurate (SelectStructuredTraceDt and the first
 * boundary event on each accepted chord, with the exit state when ex is
 * given).  Termination values follow util/TrajectoryTermination.h:10-21. */
static void trace_impl(const orc_trace_params* P, const double x0[3], const double dir[3],
                       double R_GV, double dtTrace, int64_t maxSteps, double maxTime,
                       orc_trace_result* res, orc_trace_exit* ex) {
  double Re = P->earth_radius_m;
  double boxmin[3], boxmax[3];
  for (int d = 0; d < 3; d++) boxmin[d] = P->box_min_Re[d] * Re, boxmax[d] = P->box_max_Re[d] * Re;
  double rInner = P->r_inner_Re * Re, rOuter = P->r_outer_Re * Re;
  double pMag = (R_GV * 1.0e9 * P->qabs_C) / ORC_C;
  double v0[3] = {-1.0 * dir[0], -1.0 * dir[1], -1.0 * dir[2]};
  if (P->policy == 1) {
    /* TraceTrajectoryShared(Ex) (:1563-1585) normalise the start direction */
    const double n = sqrt(dot3(v0, v0));
    if (n > 0.0) {
      const double in = 1.0 / n;
      for (int d = 0; d < 3; d++) v0[d] = in * v0[d];
    } else {
      v0[0] = v0[1] = v0[2] = 0.0;
    }
  }
  double p[3] = {pMag * v0[0], pMag * v0[1], pMag * v0[2]};
  double x[3] = {x0[0], x0[1], x0[2]};
  double t = 0.0, s = 0.0;
  int64_t nSteps = 0;
  int term = -1;
  if (ex) memset(ex, 0, sizeof(*ex));
  orc_trap trap;
  orc_trap_init(&trap);
  if (P->trap) {
    double B0[3];
    trace_field(P, x, B0);
    orc_trap_update(P, &trap, x, p, B0);
  }
  while (nSteps < maxSteps && t < maxTime && (P->max_distance_m <= 0.0 || s < P->max_distance_m)) {
    if (!isfin3(x) || !isfin3(p)) { term = 8; break; }
    if (rInner > 0.0 && x[0] * x[0] + x[1] * x[1] + x[2] * x[2] <= dmax(0.0, rInner + 0.0) * dmax(0.0, rInner + 0.0)) {
      term = 1;
      break;
    }
    int inside;
    if (P->outer_sphere) inside = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2] <= rOuter * rOuter);
    else
      inside = x[0] >= boxmin[0] - 0.0 && x[0] <= boxmax[0] + 0.0 && x[1] >= boxmin[1] - 0.0 &&
               x[1] <= boxmax[1] + 0.0 && x[2] >= boxmin[2] - 0.0 && x[2] <= boxmax[2] + 0.0;
    if (!inside) {
      term = 0;
      if (P->policy == 1 && ex) {
        /* an OuterBox event at the step start: fraction 0 on a zero chord */
        if (!populate_exit(P, x, 0.0, x, x, p, p, ex)) term = 7;
        else ex->valid = 1;
      }
      break;
    }
    double dt;
    if (P->policy == 1) dt = select_dt_structured(P, x, p, dtTrace, maxTime - t);
    else if (P->dt_selector == 1) dt = select_dt3d(P, x, p, dtTrace, maxTime - t, boxmin, boxmax, rInner);
    else dt = select_dt(P, x, p, dtTrace, maxTime - t);
    if (!(dt > 0.0) || !isfinite(dt)) { term = 6; break; }
    double xb[3] = {x[0], x[1], x[2]};
    double pb[3] = {p[0], p[1], p[2]};
    if (!orc_step_checked(P, x, p, dt, rInner)) { term = 1; break; }
    if (!isfin3(x) || !isfin3(p)) { term = 8; break; }
    int etype = 0, eface = -1;
    double efrac = 0.0, epos[3] = {0.0, 0.0, 0.0};
    if (P->policy == 1) orc_find_first_event(xb, x, boxmin, boxmax, rInner, P->event_tol_m, &etype, &efrac, epos, &eface);
    double dd[3] = {x[0] - xb[0], x[1] - xb[1], x[2] - xb[2]};
    double seg = sqrt(dot3(dd, dd));
    s += isfinite(seg) ? seg : 0.0;
    t += dt;
    ++nSteps;
    if (etype == 1) { term = 1; break; }
    if (etype == 2) {
      term = 0;
      if (ex) {
        if (!populate_exit(P, epos, efrac, xb, x, pb, p, ex)) term = 7;
        else ex->valid = 1;
      }
      break;
    }
    if (P->trap) {
      double Bt[3];
      trace_field(P, x, Bt);
      if (orc_trap_update(P, &trap, x, p, Bt)) { term = 2; break; }
    }
  }
  if (term < 0) {
    if (nSteps >= maxSteps) term = 4;
    else if (P->max_distance_m > 0.0 && s >= P->max_distance_m) term = 5;
    else if (t >= maxTime) term = 3;
    else term = 8;
  }
  if (ex && term != 0) ex->valid = 0;
  res->termination = term;
  res->retry_count = 0;
  res->steps = nSteps;
  res->time_s = t;
  res->distance_m = s;
  for (int d = 0; d < 3; d++) res->x[d] = x[d], res->p[d] = p[d];
}

/* TraceTrajectoryWithSingleRetry — CutoffRigidityGridless.cpp:1459-1496
 * (retry 1: numerical failures), or TraceDensityDirection's retry —
 * DensityGridless.cpp:849-879 (retry 2: every unresolved termination) */
void orc_trace_ex(const orc_trace_params* P, const double x0[3], const double dir[3], double R_GV,
                  orc_trace_result* res, orc_trace_exit* ex) {
  trace_impl(P, x0, dir, R_GV, P->dt_trace_s, P->max_steps, P->max_time_s, res, ex);
  int t = res->termination;
  if (!P->retry) return;
  if (P->retry == 2 ? t <= 2 : t <= 5) return;
  double dt2 = dmax(1.0e-12, 0.5 * P->dt_trace_s);
  int64_t ms2 = (P->max_steps <= (int64_t)2147483647 / 2) ? 2 * P->max_steps : (int64_t)2147483647;
  trace_impl(P, x0, dir, R_GV, dt2, ms2, 2.0 * P->max_time_s, res, ex);
  res->retry_count = 1;
}

void orc_trace(const orc_trace_params* P, const double x0[3], const double dir[3], double R_GV,
               orc_trace_result* res) {
  orc_trace_ex(P, x0, dir, R_GV, res, NULL);
}

void orc_trace_batch(const orc_trace_params* P, int64_t n, const double* x0, const double* dir,
                     const double* R_GV, orc_trace_result* res, int nthreads) {
  orc_trace_batch_ex(P, n, x0, dir, R_GV, res, NULL, nthreads);
}

void orc_trace_batch_ex(const orc_trace_params* P, int64_t n, const double* x0, const double* dir,
                        const double* R_GV, orc_trace_result* res, orc_trace_exit* ex, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int64_t i = 0; i < n; i++) orc_trace_ex(P, x0 + 3 * i, dir + 3 * i, R_GV[i], &res[i], ex ? &ex[i] : NULL);
}

/* BuildCutoffSearchGrid_GV (LOG spacing) — CutoffRigidityGridless.cpp:2578-2617 */
void orc_rigidity_grid(double Rmin, double Rmax, int nScan, double* grid) {
  if (nScan < 2) nScan = 2;
  double lmin = log(Rmin), lmax = log(Rmax);
  for (int i = 0; i < nScan; i++) {
    double a = (double)i / (double)(nScan - 1);
    grid[i] = exp((1.0 - a) * lmin + a * lmax);
  }
  grid[0] = Rmin;
  grid[nScan - 1] = Rmax;
}

/* CutoffForDirectionUpperScan_GV bin semantics — CutoffRigidityGridless.cpp:2805-2834 */
int orc_upper_cutoff_bin(const int* allowed, int n) {
  if (!allowed[n - 1]) return -1;
  for (int i = n - 2; i >= 0; --i)
    if (!allowed[i]) return i + 1;
  return 0;
}

/* AnalyzeCutoffBandSamples — util/CutoffBandSearch.h:96-194 */
void orc_band_topology(const int* st, int n, int* out) {
  int lF = -1, lA = -1, uF = -1, uA = -1, firstA = -1, lastF = -1, nT = 0, nAI = 0, nU = 0;
  int allA = 0, allF = 0, lBelow = 0, lAbove = 0, uBelow = 0, uAbove = 0, lUn = 0, uUn = 0;
  if (n > 0) {
    int inA = 0;
    for (int i = 0; i < n; i++) {
      int s = st[i];
      if (s == 1) {
        if (firstA < 0) firstA = i;
        if (!inA) { ++nAI; inA = 1; }
      } else {
        inA = 0;
        if (s == 0) lastF = i;
        else ++nU;
      }
      if (i > 0 && st[i - 1] != 2 && s != 2 && st[i - 1] != s) ++nT;
    }
    allA = (firstA == 0 && lastF < 0 && nU == 0);
    allF = (firstA < 0 && lastF == n - 1 && nU == 0);
    if (firstA == 0) lBelow = 1;
    else if (firstA > 0) {
      int clean = 1;
      for (int i = 0; i < firstA; i++)
        if (st[i] != 0) { clean = 0; break; }
      if (clean) { lF = firstA - 1; lA = firstA; }
      else lUn = 1;
    } else if (nU > 0) lUn = 1;
    else lAbove = 1;
    if (lastF < 0) {
      if (nU > 0) uUn = 1;
      else uBelow = 1;
    } else if (lastF == n - 1) uAbove = 1;
    else {
      int clean = 1;
      for (int i = lastF + 1; i < n; i++)
        if (st[i] != 1) { clean = 0; break; }
      if (clean) { uF = lastF; uA = lastF + 1; }
      else uUn = 1;
    }
  }
  int v[15] = {lF, lA, uF, uA, nT, nAI, nU, allA, allF, lBelow, lAbove, uBelow, uAbove, lUn, uUn};
  for (int i = 0; i < 15; i++) out[i] = v[i];
}


/* ===================================================================== */
/* Cell-centred linear interpolation                                      */
/* ===================================================================== */
/* Edges (pic_interpolation_routines.cpp:421-521): axis and the low/high side
 * of the two other axes; corners c: x = c&1, y = (c>>1)&1, z = c>>2. */
static const int EdgeAxis[12] = {0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2};
static const int EdgeSide[12][3] = {{-1, 0, 0}, {-1, 0, 1}, {-1, 1, 1}, {-1, 1, 0},
                                    {0, -1, 0}, {1, -1, 0}, {1, -1, 1}, {0, -1, 1},
                                    {0, 0, -1}, {1, 0, -1}, {1, 1, -1}, {0, 1, -1}};

static int block_at(const orc_mesh* m, const double x[3]) {
  int r = find_tnode(m, x, 0);
  return (r >= 0) ? m->t[r].block : -1;
}

/* Neighbours by position: a point just outside the block across each face,
 * edge and corner.  Sub-face and sub-edge points also sample finer
 * neighbours for the min/max levels (SetNeibRefinmentLevelLimits,
 * meshAMRgeneric.h:1018-1046). */
static void build_neighbour_info(orc_mesh* m) {
  m->nb_info = (int*)malloc(sizeof(int) * 28 * (size_t)m->nblocks);
  for (int b = 0; b < m->nblocks; b++) {
    const orc_tnode* n = &m->t[m->block_node[b]];
    int* info = m->nb_info + 28 * b;
    double lo[3], hi[3], dl[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = n->xmin[d], hi[d] = n->xmax[d];
      dl[d] = (hi[d] - lo[d]) / 64.0;
    }
    int mn = -1, mx = -1;
#define ORC_SEEN(bb)                                                        \
  do {                                                                      \
    if ((bb) >= 0) {                                                        \
      const int L_ = m->t[m->block_node[bb]].level;                         \
      if (mn == -1 || mn > L_) mn = L_;                                     \
      if (mx < L_) mx = L_;                                                 \
    }                                                                       \
  } while (0)
    for (int f = 0; f < 6; f++) {
      const int a = f / 2, side = f % 2;
      double p[3];
      for (int d = 0; d < 3; d++) p[d] = 0.5 * (lo[d] + hi[d]);
      p[a] = side ? hi[a] + dl[a] : lo[a] - dl[a];
      info[2 + f] = block_at(m, p);
      const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
      for (int q = 0; q < 4; q++) {
        p[a1] = lo[a1] + (hi[a1] - lo[a1]) * ((q & 1) ? 0.75 : 0.25);
        p[a2] = lo[a2] + (hi[a2] - lo[a2]) * ((q & 2) ? 0.75 : 0.25);
        ORC_SEEN(block_at(m, p));
      }
    }
    for (int e = 0; e < 12; e++) {
      double p[3];
      const int ax = EdgeAxis[e];
      for (int d = 0; d < 3; d++) {
        if (d == ax) p[d] = 0.5 * (lo[d] + hi[d]);
        else p[d] = EdgeSide[e][d] ? hi[d] + dl[d] : lo[d] - dl[d];
      }
      info[8 + e] = block_at(m, p);
      for (int q = 0; q < 2; q++) {
        p[ax] = lo[ax] + (hi[ax] - lo[ax]) * (q ? 0.75 : 0.25);
        ORC_SEEN(block_at(m, p));
      }
    }
    for (int c = 0; c < 8; c++) {
      double p[3];
      for (int d = 0; d < 3; d++) p[d] = ((c >> d) & 1) ? hi[d] + dl[d] : lo[d] - dl[d];
      info[20 + c] = block_at(m, p);
      ORC_SEEN(info[20 + c]);
    }
#undef ORC_SEEN
    info[0] = mn;
    info[1] = mx;
  }
}

void orc_block_neighbours(const orc_mesh* m, int block, int* minmax, int* face, int* edge,
                          int* corner) {
  const int* info = m->nb_info + 28 * block;
  minmax[0] = info[0], minmax[1] = info[1];
  for (int i = 0; i < 6; i++) face[i] = info[2 + i];
  for (int i = 0; i < 12; i++) edge[i] = info[8 + i];
  for (int i = 0; i < 8; i++) corner[i] = info[20 + i];
}

static inline int gidx(int i, int j, int k) { return (i + 1) + ORC_GN * ((j + 1) + ORC_GN * (k + 1)); }

typedef struct {
  int n;
  int blk[ORC_CC_MAX], cell[ORC_CC_MAX];
  double w[ORC_CC_MAX];
} orc_cst;

static void cst_add(orc_cst* s, double w, int blk, int cell) {
  s->blk[s->n] = blk, s->cell[s->n] = cell, s->w[s->n] = w;
  s->n++;
}

/* Normalize(), pic.h:7254-7261 */
static void cst_normalize(orc_cst* s) {
  double norm = 0.0;
  for (int i = 0; i < s->n; i++) norm += s->w[i];
  if (norm > 0.0)
    for (int i = 0; i < s->n; i++) s->w[i] /= norm;
}

/* a cell of the block's ghost-layered array exists when it is interior or
 * its centre lies inside the domain */
static int cell_exists(const orc_mesh* m, const orc_tnode* n, int i, int j, int k) {
  const int ijk[3] = {i, j, k};
  for (int d = 0; d < 3; d++) {
    if (ijk[d] < -1 || ijk[d] > ORC_NC) return 0;
    if (ijk[d] >= 0 && ijk[d] < ORC_NC) continue;
    const double dx = (n->xmax[d] - n->xmin[d]) / ORC_NC;
    const double c = n->xmin[d] + (ijk[d] + 0.5) * dx;
    if (!(c >= m->gxmin[d] && c < m->gxmax[d])) return 0;
  }
  return 1;
}

/* CellCentered::Constant::InitStencil, :168-219 */
static void cst_constant(const orc_mesh* m, const double x[3], int tn, orc_cst* s) {
  s->n = 0;
  int ijk[3];
  if (tn < 0 || find_cell_node(m, x, &m->t[tn], ijk) < 0) return;
  cst_add(s, 1.0, m->t[tn].block, gidx(ijk[0], ijk[1], ijk[2]));
}

/* GetTriliniarInterpolationStencil, :820-907 (the renormalisation test reads
 * the stencil being built; the reference reads a shared table there) */
static void cst_tri(const orc_mesh* m, double iLoc, double jLoc, double kLoc, const double x[3], int tn,
                    orc_cst* s) {
  const orc_tnode* n = &m->t[tn];
  s->n = 0;
  const int i0 = (iLoc < 0.5) ? -1 : (int)(iLoc - 0.50);
  const int j0 = (jLoc < 0.5) ? -1 : (int)(jLoc - 0.50);
  const int k0 = (kLoc < 0.5) ? -1 : (int)(kLoc - 0.50);
  double w[3];
  w[0] = iLoc - (i0 + 0.5);
  w[1] = jLoc - (j0 + 0.5);
  w[2] = kLoc - (k0 + 0.5);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        double W;
        switch (i + 2 * j + 4 * k) {
          case 0: W = (1.0 - w[0]) * (1.0 - w[1]) * (1.0 - w[2]); break;
          case 1: W = w[0] * (1.0 - w[1]) * (1.0 - w[2]); break;
          case 2: W = (1.0 - w[0]) * w[1] * (1.0 - w[2]); break;
          case 3: W = w[0] * w[1] * (1.0 - w[2]); break;
          case 4: W = (1.0 - w[0]) * (1.0 - w[1]) * w[2]; break;
          case 5: W = w[0] * (1.0 - w[1]) * w[2]; break;
          case 6: W = (1.0 - w[0]) * w[1] * w[2]; break;
          default: W = w[0] * w[1] * w[2]; break;
        }
        if (cell_exists(m, n, i0 + i, j0 + j, k0 + k)) cst_add(s, W, n->block, gidx(i0 + i, j0 + j, k0 + k));
      }
  if (s->n == 0) {
    cst_constant(m, x, tn, s);
    return;
  }
  if (s->n != 8) cst_normalize(s);
}

/* GetTriliniarInterpolationMutiBlockStencil, :912-1070 */
static void cst_multiblock(const orc_mesh* m, const double x[3], int tn, orc_cst* s) {
  const orc_tnode* n = &m->t[tn];
  s->n = 0;
  double dxCell[3], xLower[3], xLoc[3];
  int ijkMin[3];
  for (int d = 0; d < 3; d++) dxCell[d] = (n->xmax[d] - n->xmin[d]) / ORC_NC;
  for (int d = 0; d < 3; d++) {
    ijkMin[d] = (x[d] - n->xmin[d] < 0.5 * dxCell[d]) ? -1 : (int)((x[d] - n->xmin[d] - 0.5 * dxCell[d]) / dxCell[d]);
    xLower[d] = n->xmin[d] + (ijkMin[d] + 0.5) * dxCell[d];
    xLoc[d] = (x[d] - xLower[d]) / dxCell[d];
    if (xLoc[d] < 0.0) xLoc[d] = 0.0;
    if (xLoc[d] > 1.0) xLoc[d] = 1.0;
  }
  int geometry = 1, physical = 1;
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        const int ijk[3] = {ijkMin[0] + di, ijkMin[1] + dj, ijkMin[2] + dk};
        double xl[3];
        for (int d = 0; d < 3; d++) xl[d] = n->xmin[d] + (ijk[d] + 0.5) * dxCell[d];
        double W;
        switch (di + 2 * dj + 4 * dk) {
          case 0: W = (1.0 - xLoc[0]) * (1.0 - xLoc[1]) * (1.0 - xLoc[2]); break;
          case 1: W = xLoc[0] * (1.0 - xLoc[1]) * (1.0 - xLoc[2]); break;
          case 2: W = (1.0 - xLoc[0]) * xLoc[1] * (1.0 - xLoc[2]); break;
          case 3: W = xLoc[0] * xLoc[1] * (1.0 - xLoc[2]); break;
          case 4: W = (1.0 - xLoc[0]) * (1.0 - xLoc[1]) * xLoc[2]; break;
          case 5: W = xLoc[0] * (1.0 - xLoc[1]) * xLoc[2]; break;
          case 6: W = (1.0 - xLoc[0]) * xLoc[1] * xLoc[2]; break;
          default: W = xLoc[0] * xLoc[1] * xLoc[2]; break;
        }
        const int sn = find_tnode(m, xl, tn);
        if (sn < 0) {
          geometry = 0;
          continue;
        }
        const orc_tnode* S = &m->t[sn];
        if (S->level == n->level) {
          if (cell_exists(m, n, ijk[0], ijk[1], ijk[2])) cst_add(s, W, n->block, gidx(ijk[0], ijk[1], ijk[2]));
          else physical = 0;
        } else {
          /* a logical coarse centre covered by a finer block: the average of
           * its 2x2x2 fine cells */
          int iNeib[3];
          for (int d = 0; d < 3; d++) iNeib[d] = 2 * ((int)((xl[d] - S->xmin[d]) / dxCell[d]));
          for (int ii = 0; ii < 2; ii++)
            for (int jj = 0; jj < 2; jj++)
              for (int kk = 0; kk < 2; kk++) {
                const int f0 = iNeib[0] + ii, f1 = iNeib[1] + jj, f2 = iNeib[2] + kk;
                const double FineCellWeight = (1.0 / 8.0) * W;
                if (f0 < 0 || f0 >= ORC_NC || f1 < 0 || f1 >= ORC_NC || f2 < 0 || f2 >= ORC_NC) physical = 0;
                else cst_add(s, FineCellWeight, S->block, gidx(f0, f1, f2));
              }
        }
      }
  if (!geometry || !physical) cst_constant(m, x, find_tnode(m, x, tn), s);
}

/* Branch of CellCentered::Linear::InitStencil, :224-816 (AMPS branch of an
 * AMR mesh): 0 tri-linear in tn, 1 multi-block stencil of tn, 2 multi-block
 * stencil of the coarser block *coarser (blended with tn's tri-linear
 * stencil when 0.5 < *dmin <= 1).  loc receives (iLoc, jLoc, kLoc). */
static int cc_branch(const orc_mesh* m, const double x[3], int tn, double loc[3], int* coarser_out,
                     double* dmin_out) {
  const orc_tnode* n = &m->t[tn];
  const int* info = m->nb_info + 28 * n->block;
  const int minL = info[0], maxL = info[1];
  const double iLoc = (x[0] - n->xmin[0]) / (n->xmax[0] - n->xmin[0]) * ORC_NC;
  const double jLoc = (x[1] - n->xmin[1]) / (n->xmax[1] - n->xmin[1]) * ORC_NC;
  const double kLoc = (x[2] - n->xmin[2]) / (n->xmax[2] - n->xmin[2]) * ORC_NC;
  const double N = ORC_NC;
  loc[0] = iLoc, loc[1] = jLoc, loc[2] = kLoc;
  if (n->level == minL && n->level == maxL) return 0;
  if ((1.0 < iLoc) && (iLoc < N - 1) && (1.0 < jLoc) && (jLoc < N - 1) && (1.0 < kLoc) && (kLoc < N - 1)) return 0;
  if (n->level == minL) {
    if ((0.5 < iLoc) && (iLoc < N - 0.5) && (0.5 < jLoc) && (jLoc < N - 0.5) && (0.5 < kLoc) && (kLoc < N - 0.5))
      return 0;
    return 1;
  }
  /* search a coarser block through the faces, edges and corners near x */
  static const int faceEdges[6][4] = {{4, 11, 7, 8}, {5, 10, 6, 9}, {0, 9, 3, 8},
                                      {1, 10, 2, 11}, {0, 5, 1, 4}, {3, 6, 2, 7}};
  static const int FaceNodeMap[6][4] = {{0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5},
                                        {2, 3, 6, 7}, {0, 1, 2, 3}, {4, 5, 6, 7}};
  double dxCell[3];
  for (int d = 0; d < 3; d++) dxCell[d] = (n->xmax[d] - n->xmin[d]) / ORC_NC;
  const double xLoc[3] = {iLoc, jLoc, kLoc};
  double dmin = 10.0 * ORC_NC * ORC_NC * ORC_NC;
  int coarser = -1;
  int edgeTested[12] = {0}, cornerTested[8] = {0};
#define ORC_TRY(dist) \
  if ((dist) < dmin) dmin = (dist), coarser = nb
  for (int idim = 0; idim < 3; idim++) {
    int iFace;
    if (xLoc[idim] <= 1.0) iFace = 2 * idim;
    else if (xLoc[idim] >= N - 1.0) iFace = 2 * idim + 1;
    else continue;
    int nb = info[2 + iFace];
    if (nb >= 0 && m->t[m->block_node[nb]].level < n->level) {
      const orc_tnode* B = &m->t[m->block_node[nb]];
      int cnt = 0;
      for (int d = 0; d < 3; d++)
        if ((B->xmin[d] - dxCell[d] <= x[d]) && (B->xmax[d] + dxCell[d] >= x[d])) cnt++;
      if (cnt == 3) {
        const int a = iFace / 2;
        if ((iFace & 1) == 0) {
          if (xLoc[a] < 1.0) ORC_TRY(xLoc[a]);
        } else {
          if (xLoc[a] > N - 1) ORC_TRY(N - xLoc[a]);
        }
      }
    }
    for (int q = 0; q < 4; q++) {
      const int e = faceEdges[iFace][q];
      if (edgeTested[e]) continue;
      edgeTested[e] = 1;
      nb = info[8 + e];
      if (nb < 0 || !(m->t[m->block_node[nb]].level < n->level)) continue;
      const orc_tnode* B = &m->t[m->block_node[nb]];
      int cnt = 0;
      for (int d = 0; d < 3; d++)
        if ((B->xmin[d] - dxCell[d] <= x[d]) && (B->xmax[d] + dxCell[d] >= x[d])) cnt++;
      if (cnt != 3) continue;
      /* both non-axis coordinates within one cell of the edge's faces */
      int ok = 1;
      for (int d = 0; d < 3; d++) {
        if (d == EdgeAxis[e]) continue;
        if (EdgeSide[e][d] ? !(xLoc[d] > N - 1) : !(xLoc[d] < 1.0)) ok = 0;
      }
      if (!ok) continue;
      for (int d = 0; d < 3; d++) {
        if (d == EdgeAxis[e]) continue;
        ORC_TRY(EdgeSide[e][d] ? N - xLoc[d] : xLoc[d]);
      }
    }
    for (int q = 0; q < 4; q++) {
      const int c = FaceNodeMap[iFace][q];
      if (cornerTested[c]) continue;
      cornerTested[c] = 1;
      nb = info[20 + c];
      if (nb < 0 || !(m->t[m->block_node[nb]].level < n->level)) continue;
      const orc_tnode* B = &m->t[m->block_node[nb]];
      int cnt = 0;
      for (int d = 0; d < 3; d++)
        if ((B->xmin[d] - dxCell[d] <= x[d]) && (B->xmax[d] + dxCell[d] >= x[d])) cnt++;
      if (cnt != 3) continue;
      int ok = 1;
      for (int d = 0; d < 3; d++)
        if (((c >> d) & 1) ? !(xLoc[d] > N - 1) : !(xLoc[d] < 1.0)) ok = 0;
      if (!ok) continue;
      for (int d = 0; d < 3; d++) ORC_TRY(((c >> d) & 1) ? N - xLoc[d] : xLoc[d]);
    }
  }
#undef ORC_TRY
  *coarser_out = coarser;
  *dmin_out = dmin;
  return coarser >= 0 ? 2 : 0;
}

/* CellCentered::Linear::InitStencil, :224-816, AMPS branch of an AMR mesh */
static void cst_init(const orc_mesh* m, const double x[3], int tn, orc_cst* s) {
  double loc[3], dmin = 0.0;
  int coarser = -1;
  const int br = cc_branch(m, x, tn, loc, &coarser, &dmin);
  if (br == 0) {
    cst_tri(m, loc[0], loc[1], loc[2], x, tn, s);
    return;
  }
  if (br == 1) {
    cst_multiblock(m, x, tn, s);
    return;
  }
  cst_multiblock(m, x, m->block_node[coarser], s);
  if ((0.5 < dmin) && (dmin <= 1.0)) {
    orc_cst fine;
    cst_tri(m, loc[0], loc[1], loc[2], x, tn, &fine);
    const double cw = 1.0 - (dmin - 0.5) / 0.5, fw = (dmin - 0.5) / 0.5;
    for (int i = 0; i < s->n; i++) s->w[i] *= cw;
    for (int i = 0; i < fine.n; i++) fine.w[i] *= fw;
    /* cStencilGeneric::Add, pic.h:7273-7291 */
    for (int i = 0; i < fine.n; i++) {
      int found = 0;
      for (int j = 0; j < s->n; j++)
        if (s->blk[j] == fine.blk[i] && s->cell[j] == fine.cell[i]) {
          s->w[j] += fine.w[i];
          found = 1;
          break;
        }
      if (!found) cst_add(s, fine.w[i], fine.blk[i], fine.cell[i]);
    }
  }
}

int orc_center_stencil(const orc_mesh* m, const double x[3], int block, int* st_block,
                       int* st_cell, double* st_w) {
  orc_cst s;
  cst_init(m, x, m->block_node[block], &s);
  for (int i = 0; i < s.n; i++) st_block[i] = s.blk[i], st_cell[i] = s.cell[i], st_w[i] = s.w[i];
  return s.n;
}

void orc_interp_center(const orc_mesh* m, const double* Bc, int64_t n, const double* x,
                       const int* block, double* B, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int64_t p = 0; p < n; p++) {
    const double xp[3] = {x[p], x[n + p], x[2 * n + p]};
    orc_cst s;
    cst_init(m, xp, m->block_node[block[p]], &s);
    double b[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < s.n; i++) {
      const double* v = Bc + ((size_t)s.blk[i] * ORC_GN * ORC_GN * ORC_GN + (size_t)s.cell[i]) * 3;
      for (int d = 0; d < 3; d++) b[d] += s.w[i] * v[d];
    }
    for (int d = 0; d < 3; d++) B[d * n + p] = b[d];
  }
}

void orc_set_center_field(orc_mesh* m, const double* Bc) { m->Bc = Bc; }

static void center_B_at(const orc_mesh* m, const double x[3], int tn, double B[3]) {
  orc_cst s;
  cst_init(m, x, tn, &s);
  B[0] = B[1] = B[2] = 0.0;
  for (int i = 0; i < s.n; i++) {
    const double* v = m->Bc + ((size_t)s.blk[i] * ORC_GN * ORC_GN * ORC_GN + (size_t)s.cell[i]) * 3;
    for (int d = 0; d < 3; d++) B[d] += s.w[i] * v[d];
  }
}

/* ===================================================================== */
/* Mode3D compact-array field: the decomposition-independent row stencil  */
/* (cRowStencil, pic.h:7295-7420) built by CellCentered::Linear::        */
/* InitStencil with a row (pic_interpolation_routines.cpp:224-1070), and  */
/* GlobalMagneticField's InterpolateField_ (GlobalMagneticField.cpp:      */
/* 292-330) over B[(block*512 + i + 8(j + 8k))*3 + d] of interior cells.  */
/* ===================================================================== */
#define ORC_ROW_MAX 128 /* nMaxRowStencilLength = 2 nMaxStencilLength, pic.h:7185, 7199 */
typedef struct {
  int n;
  int blk[ORC_ROW_MAX], cell[ORC_ROW_MAX]; /* cell = i + 8(j + 8k), interior */
  double w[ORC_ROW_MAX];
} orc_row;

/* cRowStencil::AddCell, pic.h:7351-7375: an entry of the same block and cell
 * takes the weight; otherwise a new entry is appended */
static void row_addcell(orc_row* r, double w, int blk, int cell) {
  for (int e = 0; e < r->n; e++)
    if (r->blk[e] == blk && r->cell[e] == cell) {
      r->w[e] += w;
      return;
    }
  if (r->n == ORC_ROW_MAX) abort();
  r->blk[r->n] = blk, r->cell[r->n] = cell, r->w[r->n] = w;
  r->n++;
}

/* cRowStencil::Normalize, pic.h:7396-7403 */
static void row_normalize(orc_row* r) {
  double norm = 0.0;
  for (int e = 0; e < r->n; e++) norm += r->w[e];
  if (norm > 0.0)
    for (int e = 0; e < r->n; e++) r->w[e] /= norm;
}

/* AddRowStencilCellByCoordinate, pic_interpolation_routines.cpp:33-82
 * (non-periodic): a centre outside the global box is omitted; otherwise the
 * leaf holding it and the interior cell under it, with the 1e-10 face
 * clamps. */
static int row_add_coord(const orc_mesh* m, orc_row* r, double w, const double xin[3], int start) {
  const double x[3] = {xin[0], xin[1], xin[2]};
  for (int d = 0; d < 3; d++)
    if ((x[d] < m->gxmin[d]) || (x[d] > m->gxmax[d])) return 0;
  const int tn = find_tnode(m, x, start < 0 ? 0 : start);
  if (tn < 0) return 0;
  const orc_tnode* n = &m->t[tn];
  int ijk[3];
  for (int d = 0; d < 3; d++) {
    const double blockLength = n->xmax[d] - n->xmin[d];
    const double xLoc = (x[d] - n->xmin[d]) / blockLength * ORC_NC;
    ijk[d] = (int)floor(xLoc);
    if ((ijk[d] < 0) && (xLoc > -1.0e-10)) ijk[d] = 0;
    if ((ijk[d] >= ORC_NC) && (xLoc < ORC_NC + 1.0e-10)) ijk[d] = ORC_NC - 1;
    if ((ijk[d] < 0) || (ijk[d] >= ORC_NC)) return 0;
  }
  row_addcell(r, w, n->block, ijk[0] + ORC_NC * (ijk[1] + ORC_NC * ijk[2]));
  return 1;
}

/* AddRowStencilCell, :91-105: (possibly ghost) indices of tn to the centre */
static int row_add_cell(const orc_mesh* m, orc_row* r, double w, int tn, int i, int j, int k) {
  const orc_tnode* S = &m->t[tn];
  const int ijk[3] = {i, j, k};
  double x[3];
  for (int d = 0; d < 3; d++) {
    const double dx = (S->xmax[d] - S->xmin[d]) / ORC_NC;
    x[d] = S->xmin[d] + (ijk[d] + 0.5) * dx;
  }
  return row_add_coord(m, r, w, x, tn);
}

/* GetTriliniarInterpolationStencil, row part (:820-896) */
static void row_tri(const orc_mesh* m, double iLoc, double jLoc, double kLoc, int tn, orc_row* r) {
  r->n = 0;
  const int i0 = (iLoc < 0.5) ? -1 : (int)(iLoc - 0.50);
  const int j0 = (jLoc < 0.5) ? -1 : (int)(jLoc - 0.50);
  const int k0 = (kLoc < 0.5) ? -1 : (int)(kLoc - 0.50);
  double w[3];
  w[0] = iLoc - (i0 + 0.5);
  w[1] = jLoc - (j0 + 0.5);
  w[2] = kLoc - (k0 + 0.5);
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++)
      for (int k = 0; k < 2; k++) {
        double W;
        switch (i + 2 * j + 4 * k) {
          case 0: W = (1.0 - w[0]) * (1.0 - w[1]) * (1.0 - w[2]); break;
          case 1: W = w[0] * (1.0 - w[1]) * (1.0 - w[2]); break;
          case 2: W = (1.0 - w[0]) * w[1] * (1.0 - w[2]); break;
          case 3: W = w[0] * w[1] * (1.0 - w[2]); break;
          case 4: W = (1.0 - w[0]) * (1.0 - w[1]) * w[2]; break;
          case 5: W = w[0] * (1.0 - w[1]) * w[2]; break;
          case 6: W = (1.0 - w[0]) * w[1] * w[2]; break;
          default: W = w[0] * w[1] * w[2]; break;
        }
        row_add_cell(m, r, W, tn, i0 + i, j0 + j, k0 + k);
      }
  if ((r->n > 0) && (r->n != 8)) row_normalize(r);
}

/* GetTriliniarInterpolationMutiBlockStencil, row part (:912-1054) */
static void row_multiblock(const orc_mesh* m, const double x[3], int tn, orc_row* r) {
  const orc_tnode* n = &m->t[tn];
  r->n = 0;
  double dxCell[3], xLower[3], xLoc[3];
  int ijkMin[3];
  for (int d = 0; d < 3; d++) dxCell[d] = (n->xmax[d] - n->xmin[d]) / ORC_NC;
  for (int d = 0; d < 3; d++) {
    ijkMin[d] = (x[d] - n->xmin[d] < 0.5 * dxCell[d]) ? -1 : (int)((x[d] - n->xmin[d] - 0.5 * dxCell[d]) / dxCell[d]);
    xLower[d] = n->xmin[d] + (ijkMin[d] + 0.5) * dxCell[d];
    xLoc[d] = (x[d] - xLower[d]) / dxCell[d];
    if (xLoc[d] < 0.0) xLoc[d] = 0.0;
    if (xLoc[d] > 1.0) xLoc[d] = 1.0;
  }
  int geometry = 1;
  for (int di = 0; di < 2; di++)
    for (int dj = 0; dj < 2; dj++)
      for (int dk = 0; dk < 2; dk++) {
        const int ijk[3] = {ijkMin[0] + di, ijkMin[1] + dj, ijkMin[2] + dk};
        double xl[3];
        for (int d = 0; d < 3; d++) xl[d] = n->xmin[d] + (ijk[d] + 0.5) * dxCell[d];
        double W;
        switch (di + 2 * dj + 4 * dk) {
          case 0: W = (1.0 - xLoc[0]) * (1.0 - xLoc[1]) * (1.0 - xLoc[2]); break;
          case 1: W = xLoc[0] * (1.0 - xLoc[1]) * (1.0 - xLoc[2]); break;
          case 2: W = (1.0 - xLoc[0]) * xLoc[1] * (1.0 - xLoc[2]); break;
          case 3: W = xLoc[0] * xLoc[1] * (1.0 - xLoc[2]); break;
          case 4: W = (1.0 - xLoc[0]) * (1.0 - xLoc[1]) * xLoc[2]; break;
          case 5: W = xLoc[0] * (1.0 - xLoc[1]) * xLoc[2]; break;
          case 6: W = (1.0 - xLoc[0]) * xLoc[1] * xLoc[2]; break;
          default: W = xLoc[0] * xLoc[1] * xLoc[2]; break;
        }
        const int sn = find_tnode(m, xl, tn);
        if (sn < 0) {
          geometry = 0;
          continue;
        }
        const orc_tnode* S = &m->t[sn];
        if (S->level == n->level) {
          row_add_cell(m, r, W, tn, ijk[0], ijk[1], ijk[2]);
        } else {
          int iNeib[3];
          for (int d = 0; d < 3; d++) iNeib[d] = 2 * ((int)((xl[d] - S->xmin[d]) / dxCell[d]));
          for (int ii = 0; ii < 2; ii++)
            for (int jj = 0; jj < 2; jj++)
              for (int kk = 0; kk < 2; kk++) {
                const double FineCellWeight = (1.0 / 8.0) * W;
                row_add_cell(m, r, FineCellWeight, sn, iNeib[0] + ii, iNeib[1] + jj, iNeib[2] + kk);
              }
        }
      }
  if (!geometry) {
    /* BuildConstantRowStencil (:109-115) at the leaf holding x */
    const int xn = find_tnode(m, x, tn);
    r->n = 0;
    row_add_coord(m, r, 1.0, x, xn);
  }
}

/* CellCentered::Linear::InitStencil(x, node, row) (pic.h:7528-7532) */
static void row_init(const orc_mesh* m, const double x[3], int tn, orc_row* r) {
  double loc[3], dmin = 0.0;
  int coarser = -1;
  const int br = cc_branch(m, x, tn, loc, &coarser, &dmin);
  if (br == 0) {
    row_tri(m, loc[0], loc[1], loc[2], tn, r);
    return;
  }
  if (br == 1) {
    row_multiblock(m, x, tn, r);
    return;
  }
  row_multiblock(m, x, m->block_node[coarser], r);
  if ((0.5 < dmin) && (dmin <= 1.0)) {
    /* the coarse/fine blend of :660-681: MultiplyScalar on both rows, then
     * cRowStencil::Add (AddCell per fine entry, merging) */
    orc_row fine;
    row_tri(m, loc[0], loc[1], loc[2], tn, &fine);
    const double cw = 1.0 - (dmin - 0.5) / 0.5, fw = (dmin - 0.5) / 0.5;
    for (int e = 0; e < r->n; e++) r->w[e] *= cw;
    for (int e = 0; e < fine.n; e++) fine.w[e] *= fw;
    for (int e = 0; e < fine.n; e++) row_addcell(r, fine.w[e], fine.blk[e], fine.cell[e]);
  }
}

int orc_row_stencil(const orc_mesh* m, const double x[3], int* st_block, int* st_cell, double* st_w) {
  const int tn = find_tnode(m, x, 0);
  if (tn < 0) return 0;
  orc_row r;
  row_init(m, x, tn, &r);
  for (int e = 0; e < r.n; e++) st_block[e] = r.blk[e], st_cell[e] = r.cell[e], st_w[e] = r.w[e];
  return r.n;
}

/* cMode3DMeshFieldEval::GetB_T, CutoffRigidityMode3D.cpp:702-800, mesh path:
 * zero outside the mesh; otherwise InterpolateField_ (GlobalMagneticField.cpp:
 * 292-330), the row's weighted sum in row order.  Returns 0 for an empty
 * row (the reference exits there). */
static int mode3d_B(const orc_mesh* m, const double* Bm, const double x[3], double B[3]) {
  B[0] = B[1] = B[2] = 0.0;
  const int tn = find_tnode(m, x, 0);
  if (tn < 0) return 1;
  orc_row r;
  row_init(m, x, tn, &r);
  if (r.n <= 0) return 0;
  for (int e = 0; e < r.n; e++) {
    const double* v = Bm + ((size_t)r.blk[e] * ORC_NCELL + (size_t)r.cell[e]) * 3;
    B[0] += r.w[e] * v[0];
    B[1] += r.w[e] * v[1];
    B[2] += r.w[e] * v[2];
  }
  return 1;
}

int orc_mode3d_field(const orc_mesh* m, const double* Bm, int64_t n, const double* x, double* B) {
  int ok = 1;
  for (int64_t p = 0; p < n; p++) ok &= mode3d_B(m, Bm, x + 3 * p, B + 3 * p);
  return ok;
}

static const orc_mesh* g_trace_mesh = NULL;
static const double* g_trace_Bm = NULL;
void orc_set_trace_mesh(const orc_mesh* m, const double* Bm) {
  g_trace_mesh = m;
  g_trace_Bm = Bm;
}

/* the tracer's field: the analytic dipole plus the uniform field, or the
 * Mode3D mesh field of orc_set_trace_mesh */
static void trace_field(const orc_trace_params* P, const double x[3], double B[3]) {
  if (P->field_model == 1) mode3d_B(g_trace_mesh, g_trace_Bm, x, B);
  else orc_dipole_B(P, x, B);
}
