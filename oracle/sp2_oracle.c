/*
 * sp2_oracle.c -- CPU restatement of ModEM's SP2 3D-MT path (TEST INFRASTRUCTURE).
 *
 * The oracle the HIP library is checked against.  Only tests/, the smoke check
 * and the benchmark's CPU-baseline leg load it; the product never does.
 *
 * Each function names the reference routine it restates (paths relative to
 * the reference's f90/ tree).  Loop and summation order follow the reference
 * so that assembled operators agree bit for bit with the flang build of the
 * reference (tests/test_oracle_vs_ref.py).  Build: oracle/Makefile, with
 * -ffp-contract=off (the reference's flang build uses no FMA) and C99 complex
 * arithmetic (gcc and flang both divide through __divdc3).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   grid/metric/topology/index sets/sigma maps/AAii/CCii/CCib/GDii/L/U/VDsG:
 *     bit-exact against the flang reference build; plus T*G = 0 exactly and
 *     CC*G = 0 to rounding.
 *   Mult_Aii, PC_Lsolve/PC_Usolve: agreement with the reference to 1e-15 rel;
 *     L*U reproduces the block pattern of A (DILU invariant).
 *   BiCGStab (both semantics): converged solution vs the reference CPU BICG and
 *     vs a dense direct solve on small grids.
 *   PCG: vs the reference PCG and a dense solve.
 *   RHS/post-processing: closed forms against the reference's definitions.
 *   edge2model / pmultT: against the reference dEdgeToModelParam dump and
 *     the adjoint identity <Cell2Edge m, e> = <m, Edge2Cell e> in V-weights.
 */
#include "sp2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* math_constants.f90:12-19 (note the truncated PI the reference uses) */
#define O_PI 3.14159265357898
#define O_MU0 (O_PI * 4e-7)
#define O_ISIGN (-1)
/* SIGMA_AIR is a default-real literal widened to real(8) */
#define O_SIGMA_AIR ((double)1.0e-10f)


/* Complex division as the reference's flang build performs it: flang lowers
 * every complex quotient (complex/complex and complex/real) to __divdc3, whose
 * algorithm is the C99 Annex G example (G.5.1, _Cdivd): scale the divisor by
 * 2^-logb(max(|c|,|d|)), divide, scale back.  Restated here from the standard
 * so that quotients round exactly as in the reference. */
static zdouble zdiv(zdouble num, zdouble den) {
  double a = creal(num), b = cimag(num), c = creal(den), d = cimag(den);
  int ilogbw = 0;
  double logbw = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(logbw)) {
    ilogbw = (int)logbw;
    c = scalbn(c, -ilogbw);
    d = scalbn(d, -ilogbw);
  }
  double denom = c * c + d * d;
  double x = scalbn((a * c + b * d) / denom, -ilogbw);
  double y = scalbn((b * c - a * d) / denom, -ilogbw);
  if (isnan(x) && isnan(y)) {
    if (denom == 0.0 && (!isnan(a) || !isnan(b))) {
      x = copysign(INFINITY, c) * a;
      y = copysign(INFINITY, c) * b;
    } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
      a = copysign(isinf(a) ? 1.0 : 0.0, a);
      b = copysign(isinf(b) ? 1.0 : 0.0, b);
      x = INFINITY * (a * c + b * d);
      y = INFINITY * (b * c - a * d);
    } else if (isinf(logbw) && logbw > 0.0 && isfinite(a) && isfinite(b)) {
      c = copysign(isinf(c) ? 1.0 : 0.0, c);
      d = copysign(isinf(d) ? 1.0 : 0.0, d);
      x = 0.0 * (a * c + b * d);
      y = 0.0 * (b * c - a * d);
    }
  }
  return CMPLX(x, y);
}
static zdouble zdivr(zdouble num, double den) { return zdiv(num, CMPLX(den, 0.0)); }

#define XALLOC(T, n) ((T *)calloc((size_t)((n) > 0 ? (n) : 1), sizeof(T)))

/* ------------------------------------------------------------------ grid */
typedef struct {
  int nx, ny, nz, nzAir, nzEarth;
  double *dx, *dy, *dz, *delX, *delY, *delZ;
} o_grid;

/* GridDef.f90:301-399 setup_grid (delX etc.) */
static void grid_setup(o_grid *g) {
  int i;
  g->delX = XALLOC(double, g->nx + 1);
  g->delY = XALLOC(double, g->ny + 1);
  g->delZ = XALLOC(double, g->nz + 1);
  g->delX[0] = g->dx[0];
  for (i = 1; i < g->nx; i++) g->delX[i] = g->dx[i - 1] + g->dx[i];
  g->delX[g->nx] = g->dx[g->nx - 1];
  for (i = 0; i <= g->nx; i++) g->delX[i] = g->delX[i] / 2.0;
  g->delY[0] = g->dy[0];
  for (i = 1; i < g->ny; i++) g->delY[i] = g->dy[i - 1] + g->dy[i];
  g->delY[g->ny] = g->dy[g->ny - 1];
  for (i = 0; i <= g->ny; i++) g->delY[i] = g->delY[i] / 2.0;
  g->delZ[0] = g->dz[0];
  for (i = 1; i < g->nz; i++) g->delZ[i] = g->dz[i - 1] + g->dz[i];
  g->delZ[g->nz] = g->dz[g->nz - 1];
  for (i = 0; i <= g->nz; i++) g->delZ[i] = g->delZ[i] / 2.0;
}

/* component dims, vecTranslate.f90:275-325 setLimits */
enum { XEDGE, YEDGE, ZEDGE, XFACE, YFACE, ZFACE, CORNER, CENTER };
static void limits(const o_grid *g, int t, int *a, int *b, int *c) {
  int nx = g->nx, ny = g->ny, nz = g->nz;
  switch (t) {
    case XEDGE: *a = nx; *b = ny + 1; *c = nz + 1; break;
    case YEDGE: *a = nx + 1; *b = ny; *c = nz + 1; break;
    case ZEDGE: *a = nx + 1; *b = ny + 1; *c = nz; break;
    case XFACE: *a = nx + 1; *b = ny; *c = nz; break;
    case YFACE: *a = nx; *b = ny + 1; *c = nz; break;
    case ZFACE: *a = nx; *b = ny; *c = nz + 1; break;
    case CORNER: *a = nx + 1; *b = ny + 1; *c = nz + 1; break;
    default: *a = nx; *b = ny; *c = nz; break;
  }
}
static int ncomp(const o_grid *g, int t) {
  int a, b, c;
  limits(g, t, &a, &b, &c);
  return a * b * c;
}
/* vectorIndex (vecTranslate.f90:367-390), 0-based i,j,k -> 0-based index */
static int vidx(const o_grid *g, int t, int i, int j, int k) {
  int a, b, c;
  limits(g, t, &a, &b, &c);
  return (k * b + j) * a + i;
}

/* ------------------------------------------------------------ CSR toolkit */
static o_csr csr_new(int m, int n, int nz) {
  o_csr A;
  A.nrow = m;
  A.ncol = n;
  A.row = XALLOC(int, m + 1);
  A.col = XALLOC(int, nz);
  A.val = XALLOC(double, nz);
  return A;
}
static void csr_free(o_csr *A) {
  free(A->row); free(A->col); free(A->val);
  A->row = A->col = NULL; A->val = NULL;
}
static o_csrz csrz_new(int m, int n, int nz) {
  o_csrz A;
  A.nrow = m;
  A.ncol = n;
  A.row = XALLOC(int, m + 1);
  A.col = XALLOC(int, nz);
  A.val = XALLOC(zdouble, nz);
  return A;
}
static void csrz_free(o_csrz *A) {
  free(A->row); free(A->col); free(A->val);
  A->row = A->col = NULL; A->val = NULL;
}
static int nnz_of(const o_csr *A) { return A->row[A->nrow]; }

/* drop exact zeros keeping order (the compaction step shared by
 * DIAGxRMAT/RMATxDIAG/RMATxRMAT/RMATplusRMAT, spOpTools.f90:680-844, 477-596) */
static o_csr csr_compact(o_csr *T) {
  int m = T->nrow, i, j, k = 0, nz = 0;
  for (j = 0; j < nnz_of(T); j++) nz += (fabs(T->val[j]) > 0);
  o_csr B = csr_new(m, T->ncol, nz);
  for (i = 0; i < m; i++) {
    B.row[i] = k;
    for (j = T->row[i]; j < T->row[i + 1]; j++)
      if (fabs(T->val[j]) > 0) { B.col[k] = T->col[j]; B.val[k] = T->val[j]; k++; }
  }
  B.row[m] = k;
  csr_free(T);
  return B;
}

/* DIAGxRMAT spOpTools.f90:680-744: B = diag(D) A */
static o_csr diag_x_rmat(const double *D, const o_csr *A) {
  o_csr T = csr_new(A->nrow, A->ncol, nnz_of(A));
  memcpy(T.row, A->row, sizeof(int) * (A->nrow + 1));
  for (int i = 0; i < A->nrow; i++)
    for (int j = A->row[i]; j < A->row[i + 1]; j++) {
      T.val[j] = D[i] * A->val[j];
      T.col[j] = A->col[j];
    }
  return csr_compact(&T);
}
/* RMATxDIAG spOpTools.f90:779-844: B = A diag(D) */
static o_csr rmat_x_diag(const o_csr *A, const double *D) {
  o_csr T = csr_new(A->nrow, A->ncol, nnz_of(A));
  memcpy(T.row, A->row, sizeof(int) * (A->nrow + 1));
  for (int i = 0; i < A->nrow; i++)
    for (int j = A->row[i]; j < A->row[i + 1]; j++) {
      T.val[j] = A->val[j] * D[A->col[j]];
      T.col[j] = A->col[j];
    }
  return csr_compact(&T);
}
/* RMATtrans spOpTools.f90:917-948 via CSR2IJS/IJS2CSR (:285-361): a stable
 * counting sort, so transposed rows list their columns in ascending order. */
static o_csr rmat_trans(const o_csr *A) {
  int nz = nnz_of(A), i, j;
  o_csr B = csr_new(A->ncol, A->nrow, nz);
  int *cnt = XALLOC(int, A->ncol + 1);
  for (j = 0; j < nz; j++) cnt[A->col[j]]++;
  B.row[0] = 0;
  for (i = 0; i < A->ncol; i++) B.row[i + 1] = B.row[i] + cnt[i];
  memset(cnt, 0, sizeof(int) * (A->ncol + 1));
  for (i = 0; i < A->nrow; i++)
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      int r = A->col[j], pos = B.row[r] + cnt[r]++;
      B.col[pos] = i;
      B.val[pos] = A->val[j];
    }
  free(cnt);
  return B;
}
/* RMATxRMAT spOpTools.f90:477-596: C = A B; columns of a row in first-encounter
 * order, partial sums in encounter order, exact zeros dropped. */
static o_csr rmat_x_rmat(const o_csr *A, const o_csr *B) {
  int m = A->nrow, i, j, k, l;
  int maxA = 0, maxB = 0;
  for (i = 0; i < A->nrow; i++) maxA = (A->row[i + 1] - A->row[i]) > maxA ? (A->row[i + 1] - A->row[i]) : maxA;
  for (i = 0; i < B->nrow; i++) maxB = (B->row[i + 1] - B->row[i]) > maxB ? (B->row[i + 1] - B->row[i]) : maxB;
  int cap = maxA * maxB;
  int *colT = XALLOC(int, cap);
  int *rowT = XALLOC(int, m + 1);
  for (i = 0; i < m; i++) {
    int nc = 0;
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      int jj = A->col[j];
      for (k = B->row[jj]; k < B->row[jj + 1]; k++) {
        int nw = 1;
        for (l = 0; l < nc; l++) if (colT[l] == B->col[k]) { nw = 0; break; }
        if (nw) colT[nc++] = B->col[k];
      }
    }
    rowT[i + 1] = rowT[i] + nc;
  }
  o_csr T = csr_new(m, B->ncol, rowT[m]);
  memcpy(T.row, rowT, sizeof(int) * (m + 1));
  for (i = 0; i < m; i++) {
    int nc = 0, i1 = T.row[i];
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      int jj = A->col[j];
      for (k = B->row[jj]; k < B->row[jj + 1]; k++) {
        int pos = -1;
        for (l = 0; l < nc; l++) if (colT[l] == B->col[k]) { pos = l; break; }
        if (pos < 0) {
          colT[nc] = B->col[k];
          T.val[i1 + nc] = A->val[j] * B->val[k];
          T.col[i1 + nc] = B->col[k];
          nc++;
        } else {
          T.val[i1 + pos] = T.val[i1 + pos] + A->val[j] * B->val[k];
        }
      }
    }
  }
  free(colT);
  free(rowT);
  return csr_compact(&T);
}
/* subMatrix_Real spOpTools.f90:1334-1383 */
static o_csr submatrix(const o_csr *A, const int *r, int m, const int *c, int n) {
  int *colT = XALLOC(int, A->ncol);
  int i, j, nz = 0;
  for (i = 0; i < A->ncol; i++) colT[i] = -1;
  for (i = 0; i < n; i++) colT[c[i]] = i;
  for (i = 0; i < m; i++)
    for (j = A->row[r[i]]; j < A->row[r[i] + 1]; j++) nz += (colT[A->col[j]] >= 0);
  o_csr B = csr_new(m, n, nz);
  int k = 0;
  for (i = 0; i < m; i++) {
    B.row[i] = k;
    for (j = A->row[r[i]]; j < A->row[r[i] + 1]; j++)
      if (colT[A->col[j]] >= 0) { B.col[k] = colT[A->col[j]]; B.val[k] = A->val[j]; k++; }
  }
  B.row[m] = k;
  free(colT);
  return B;
}
/* RMATplusRMAT spOpTools.f90:2197-2297 (droptol = 0): A's entries first, then
 * B's new columns appended; sums that are exactly zero are dropped. */
static o_csr rmat_plus_rmat(const o_csr *A, const o_csr *B) {
  int m = A->nrow, i, j, k;
  o_csr T = csr_new(m, A->ncol, nnz_of(A) + nnz_of(B));
  int w = 0;
  for (i = 0; i < m; i++) {
    T.row[i] = w;
    int a0 = A->row[i], a1 = A->row[i + 1], base = w;
    for (j = a0; j < a1; j++) { T.col[w] = A->col[j]; T.val[w] = A->val[j]; w++; }
    for (k = B->row[i]; k < B->row[i + 1]; k++) {
      int found = -1;
      for (j = a0; j < a1; j++) if (B->col[k] == A->col[j]) { found = j; break; }
      if (found >= 0) {
        double t = B->val[k] + A->val[found];
        double test = (A->val[found] != 0.0) ? fabs(t) / fabs(A->val[found]) : fabs(t) / fabs(B->val[k]);
        T.val[base + (found - a0)] = (test > 0.0) ? t : 0.0;
      } else {
        T.col[w] = B->col[k];
        T.val[w] = B->val[k];
        w++;
      }
    }
  }
  T.row[m] = w;
  return csr_compact(&T);
}
/* RMATxCVEC spOpTools.f90:404-427 */
static void rmat_x_cvec(const o_csr *A, const zdouble *x, zdouble *y) {
  for (int i = 0; i < A->nrow; i++) {
    zdouble s = 0;
    for (int j = A->row[i]; j < A->row[i + 1]; j++) s = s + A->val[j] * x[A->col[j]];
    y[i] = s;
  }
}
/* RMATxRVEC spOpTools.f90:429-451 */
static void rmat_x_rvec(const o_csr *A, const double *x, double *y) {
  for (int i = 0; i < A->nrow; i++) {
    double s = 0.0;
    for (int j = A->row[i]; j < A->row[i + 1]; j++) s = s + A->val[j] * x[A->col[j]];
    y[i] = s;
  }
}

/* --------------------------------------------------------------- problem */
struct sp2o_problem {
  o_grid g;
  int nXe, nYe, nZe, nE, nNode;
  int nInt, nBdry, nNi, nNb;
  int *EDGEi, *EDGEb, *NODEi, *NODEb;
  int nxInt, nyInt, nzInt; /* interior edges per component */
  double *FaceA, *EdgeL, *DualFaceA, *DualEdgeL, *Vnode, *Vedge;
  double *sigEdge, *sigNode; /* as formed by UpdateCond (boundary edges zeroed) */
  double *logcond, airLog;
  o_csr T, G, CCii, CCib, AAii, GDii;
  double *vms; /* VomegaMuSig */
  double omega;
  o_csrz L, U;
  int have_lu;
  o_csr VDiv, VDs, Gai, VDsG, VDsG_L, VDsG_U;
  int have_divcor;
};

/* spOpTopology_SG.f90:24-171 setCurlTopology; faces x 4 edges */
static o_csr curl_topology(const o_grid *g) {
  int nXf = ncomp(g, XFACE), nYf = ncomp(g, YFACE), nZf = ncomp(g, ZFACE);
  int nXe = ncomp(g, XEDGE), nYe = ncomp(g, YEDGE);
  int m = nXf + nYf + nZf, n = nXe + nYe + ncomp(g, ZEDGE);
  o_csr T = csr_new(m, n, 4 * m);
  for (int r = 0; r <= m; r++) T.row[r] = 4 * r;
  int a, b, c, i, j, k, r;
  limits(g, XFACE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = (k * b + j) * a + i;
    T.col[4 * r + 0] = vidx(g, YEDGE, i, j, k) + nXe;       T.val[4 * r + 0] = 1;
    T.col[4 * r + 2] = vidx(g, ZEDGE, i, j, k) + nXe + nYe; T.val[4 * r + 2] = -1;
    T.col[4 * r + 1] = vidx(g, YEDGE, i, j, k + 1) + nXe;   T.val[4 * r + 1] = -1;
    T.col[4 * r + 3] = vidx(g, ZEDGE, i, j + 1, k) + nXe + nYe; T.val[4 * r + 3] = 1;
  }
  limits(g, YFACE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = nXf + (k * b + j) * a + i;
    T.col[4 * r + 0] = vidx(g, XEDGE, i, j, k);             T.val[4 * r + 0] = -1;
    T.col[4 * r + 2] = vidx(g, ZEDGE, i, j, k) + nXe + nYe; T.val[4 * r + 2] = 1;
    T.col[4 * r + 1] = vidx(g, XEDGE, i, j, k + 1);         T.val[4 * r + 1] = 1;
    T.col[4 * r + 3] = vidx(g, ZEDGE, i + 1, j, k) + nXe + nYe; T.val[4 * r + 3] = -1;
  }
  limits(g, ZFACE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = nXf + nYf + (k * b + j) * a + i;
    T.col[4 * r + 0] = vidx(g, XEDGE, i, j, k);             T.val[4 * r + 0] = 1;
    T.col[4 * r + 2] = vidx(g, YEDGE, i, j, k) + nXe;       T.val[4 * r + 2] = -1;
    T.col[4 * r + 1] = vidx(g, XEDGE, i, j + 1, k);         T.val[4 * r + 1] = -1;
    T.col[4 * r + 3] = vidx(g, YEDGE, i + 1, j, k) + nXe;   T.val[4 * r + 3] = 1;
  }
  return T;
}

/* spOpTopology_SG.f90:173-266 setGradTopology; edges x 2 nodes (-1, +1) */
static o_csr grad_topology(const o_grid *g) {
  int nXe = ncomp(g, XEDGE), nYe = ncomp(g, YEDGE), nZe = ncomp(g, ZEDGE);
  int m = nXe + nYe + nZe, n = ncomp(g, CORNER);
  o_csr G = csr_new(m, n, 2 * m);
  int a, b, c, i, j, k, r;
  for (r = 0; r <= m; r++) G.row[r] = 2 * r;
  for (r = 0; r < m; r++) { G.val[2 * r] = -1; G.val[2 * r + 1] = 1; }
  limits(g, XEDGE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = (k * b + j) * a + i;
    G.col[2 * r] = vidx(g, CORNER, i, j, k);
    G.col[2 * r + 1] = vidx(g, CORNER, i + 1, j, k);
  }
  limits(g, YEDGE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = nXe + (k * b + j) * a + i;
    G.col[2 * r] = vidx(g, CORNER, i, j, k);
    G.col[2 * r + 1] = vidx(g, CORNER, i, j + 1, k);
  }
  limits(g, ZEDGE, &a, &b, &c);
  for (k = 0; k < c; k++) for (j = 0; j < b; j++) for (i = 0; i < a; i++) {
    r = nXe + nYe + (k * b + j) * a + i;
    G.col[2 * r] = vidx(g, CORNER, i, j, k);
    G.col[2 * r + 1] = vidx(g, CORNER, i, j, k + 1);
  }
  return G;
}

/* vecTranslate.f90:436-527 boundaryIndex (EDGE and CORNER) */
static void boundary_index(sp2o_problem *p) {
  const o_grid *g = &p->g;
  int nx = g->nx, ny = g->ny, nz = g->nz, i, j, k, t = 0;
  char *flag = XALLOC(char, p->nE);
  for (k = 0; k <= nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i < nx; i++)
    flag[t++] = (j == 0 || j == ny || k == 0 || k == nz);
  for (k = 0; k <= nz; k++) for (j = 0; j < ny; j++) for (i = 0; i <= nx; i++)
    flag[t++] = (i == 0 || i == nx || k == 0 || k == nz);
  for (k = 0; k < nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i <= nx; i++)
    flag[t++] = (i == 0 || i == nx || j == 0 || j == ny);
  int nb = 0;
  for (i = 0; i < p->nE; i++) nb += flag[i];
  p->nBdry = nb;
  p->nInt = p->nE - nb;
  p->EDGEi = XALLOC(int, p->nInt);
  p->EDGEb = XALLOC(int, p->nBdry);
  int ib = 0, ii = 0;
  for (i = 0; i < p->nE; i++) { if (flag[i]) p->EDGEb[ib++] = i; else p->EDGEi[ii++] = i; }
  free(flag);
  flag = XALLOC(char, p->nNode);
  t = 0;
  for (k = 0; k <= nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i <= nx; i++)
    flag[t++] = (i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz);
  nb = 0;
  for (i = 0; i < p->nNode; i++) nb += flag[i];
  p->nNb = nb;
  p->nNi = p->nNode - nb;
  p->NODEi = XALLOC(int, p->nNi);
  p->NODEb = XALLOC(int, p->nNb);
  ib = ii = 0;
  for (i = 0; i < p->nNode; i++) { if (flag[i]) p->NODEb[ib++] = i; else p->NODEi[ii++] = i; }
  free(flag);
  p->nxInt = nx * (ny - 1) * (nz - 1);
  p->nyInt = (nx - 1) * ny * (nz - 1);
  p->nzInt = (nx - 1) * (ny - 1) * nz;
}

/* GridCalc.f90:230-418 EdgeLength, FaceLength, EdgeArea, FaceArea;
 * :152-186 NodeVolume; :44-90 EdgeVolume = EdgeLength * EdgeArea */
static void metrics(sp2o_problem *p) {
  const o_grid *g = &p->g;
  int nx = g->nx, ny = g->ny, nz = g->nz, i, j, k, t;
  int nF = ncomp(g, XFACE) + ncomp(g, YFACE) + ncomp(g, ZFACE);
  p->FaceA = XALLOC(double, nF);
  p->DualEdgeL = XALLOC(double, nF);
  p->EdgeL = XALLOC(double, p->nE);
  p->DualFaceA = XALLOC(double, p->nE);
  p->Vedge = XALLOC(double, p->nE);
  p->Vnode = XALLOC(double, p->nNode);
  /* edges */
  t = 0;
  for (k = 0; k <= nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i < nx; i++, t++) {
    p->EdgeL[t] = g->dx[i]; p->DualFaceA[t] = g->delY[j] * g->delZ[k];
  }
  for (k = 0; k <= nz; k++) for (j = 0; j < ny; j++) for (i = 0; i <= nx; i++, t++) {
    p->EdgeL[t] = g->dy[j]; p->DualFaceA[t] = g->delX[i] * g->delZ[k];
  }
  for (k = 0; k < nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i <= nx; i++, t++) {
    p->EdgeL[t] = g->dz[k]; p->DualFaceA[t] = g->delX[i] * g->delY[j];
  }
  for (t = 0; t < p->nE; t++) p->Vedge[t] = p->EdgeL[t] * p->DualFaceA[t];
  /* faces; FaceLength uses (d(i-1)+d(i))/2.0d0 with half-cells at the ends */
  t = 0;
  for (k = 0; k < nz; k++) for (j = 0; j < ny; j++) for (i = 0; i <= nx; i++, t++) {
    double l = (i == 0) ? g->dx[0] : (i == nx ? g->dx[nx - 1] : g->dx[i - 1] + g->dx[i]);
    p->DualEdgeL[t] = l / 2.0;
    p->FaceA[t] = g->dy[j] * g->dz[k];
  }
  for (k = 0; k < nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i < nx; i++, t++) {
    double l = (j == 0) ? g->dy[0] : (j == ny ? g->dy[ny - 1] : g->dy[j - 1] + g->dy[j]);
    p->DualEdgeL[t] = l / 2.0;
    p->FaceA[t] = g->dx[i] * g->dz[k];
  }
  for (k = 0; k <= nz; k++) for (j = 0; j < ny; j++) for (i = 0; i < nx; i++, t++) {
    double l = (k == 0) ? g->dz[0] : (k == nz ? g->dz[nz - 1] : g->dz[k - 1] + g->dz[k]);
    p->DualEdgeL[t] = l / 2.0;
    p->FaceA[t] = g->dx[i] * g->dy[j];
  }
  t = 0;
  for (k = 0; k <= nz; k++) for (j = 0; j <= ny; j++) for (i = 0; i <= nx; i++, t++)
    p->Vnode[t] = g->delX[i] * g->delY[j] * g->delZ[k];
}

/* ----------------------------------------------- cell <-> edge/node maps */
#define CI(i, j, k) ((((k) * g->ny) + (j)) * g->nx + (i))
/* GridCalc.f90:430-590 Cell2Edge (1-based loops restated 0-based) */
static void cell2edge(const o_grid *g, const double *C, double *E) {
  int nx = g->nx, ny = g->ny, nz = g->nz, ix, iy, iz;
  int nXe = nx * (ny + 1) * (nz + 1), nYe = (nx + 1) * ny * (nz + 1);
  double *Ex = E, *Ey = E + nXe, *Ez = E + nXe + nYe;
#define EX(i, j, k) Ex[((k) * (ny + 1) + (j)) * nx + (i)]
#define EY(i, j, k) Ey[((k) * ny + (j)) * (nx + 1) + (i)]
#define EZ(i, j, k) Ez[((k) * (ny + 1) + (j)) * (nx + 1) + (i)]
#define C1(i, j, k) C[CI((i) - 1, (j) - 1, (k) - 1)]
  memset(E, 0, sizeof(double) * (nXe + nYe + (nx + 1) * (ny + 1) * nz));
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++)
    EX(ix - 1, iy - 1, iz - 1) = (C1(ix, iy - 1, iz - 1) + C1(ix, iy, iz - 1) + C1(ix, iy - 1, iz) + C1(ix, iy, iz)) / 4.0;
  for (ix = 2; ix <= nx; ix++) for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++)
    EY(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy, iz - 1) + C1(ix, iy, iz - 1) + C1(ix - 1, iy, iz) + C1(ix, iy, iz)) / 4.0;
  for (ix = 2; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) for (iz = 1; iz <= nz; iz++)
    EZ(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy - 1, iz) + C1(ix - 1, iy, iz) + C1(ix, iy - 1, iz) + C1(ix, iy, iz)) / 4.0;
  iz = 1;
  for (iy = 1; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) EY(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy, iz) + C1(ix, iy, iz)) / 2.0;
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) EX(ix - 1, iy - 1, iz - 1) = (C1(ix, iy - 1, iz) + C1(ix, iy, iz)) / 2.0;
  iz = nz + 1;
  for (iy = 1; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) EY(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy, iz - 1) + C1(ix, iy, iz - 1)) / 2.0;
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) EX(ix - 1, iy - 1, iz - 1) = (C1(ix, iy - 1, iz - 1) + C1(ix, iy, iz - 1)) / 2.0;
  ix = 1;
  for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) EY(ix - 1, iy - 1, iz - 1) = (C1(ix, iy, iz - 1) + C1(ix, iy, iz)) / 2.0;
  for (iz = 1; iz <= nz; iz++) for (iy = 2; iy <= ny; iy++) EZ(ix - 1, iy - 1, iz - 1) = (C1(ix, iy - 1, iz) + C1(ix, iy, iz)) / 2.0;
  ix = nx + 1;
  for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) EY(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy, iz - 1) + C1(ix - 1, iy, iz)) / 2.0;
  for (iz = 1; iz <= nz; iz++) for (iy = 2; iy <= ny; iy++) EZ(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy - 1, iz) + C1(ix - 1, iy, iz)) / 2.0;
  iy = 1;
  for (ix = 1; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) EX(ix - 1, iy - 1, iz - 1) = (C1(ix, iy, iz - 1) + C1(ix, iy, iz)) / 2.0;
  for (iz = 1; iz <= nz; iz++) for (ix = 2; ix <= nx; ix++) EZ(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy, iz) + C1(ix, iy, iz)) / 2.0;
  iy = ny + 1;
  for (ix = 1; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) EX(ix - 1, iy - 1, iz - 1) = (C1(ix, iy - 1, iz - 1) + C1(ix, iy - 1, iz)) / 2.0;
  for (iz = 1; iz <= nz; iz++) for (ix = 2; ix <= nx; ix++) EZ(ix - 1, iy - 1, iz - 1) = (C1(ix - 1, iy - 1, iz) + C1(ix, iy - 1, iz)) / 2.0;
  for (iz = 1; iz <= nz; iz++) {
    EZ(0, 0, iz - 1) = C1(1, 1, iz);
    EZ(0, ny, iz - 1) = C1(1, ny, iz);
    EZ(nx, 0, iz - 1) = C1(nx, 1, iz);
    EZ(nx, ny, iz - 1) = C1(nx, ny, iz);
  }
  for (iy = 1; iy <= ny; iy++) {
    EY(0, iy - 1, 0) = C1(1, iy, 1);
    EY(0, iy - 1, nz) = C1(1, iy, nz);
    EY(nx, iy - 1, 0) = C1(nx, iy, 1);
    EY(nx, iy - 1, nz) = C1(nx, iy, nz);
  }
  for (ix = 1; ix <= nx; ix++) {
    EX(ix - 1, 0, 0) = C1(ix, 1, 1);
    EX(ix - 1, 0, nz) = C1(ix, 1, nz);
    EX(ix - 1, ny, 0) = C1(ix, ny, 1);
    EX(ix - 1, ny, nz) = C1(ix, ny, nz);
  }
}

/* GridCalc.f90:1047-1226 Edge2Cell: the scatter transpose of Cell2Edge */
static void edge2cell(const o_grid *g, const double *E, double *C) {
  int nx = g->nx, ny = g->ny, nz = g->nz, ix, iy, iz;
  int nXe = nx * (ny + 1) * (nz + 1), nYe = (nx + 1) * ny * (nz + 1);
  const double *Ex = E, *Ey = E + nXe, *Ez = E + nXe + nYe;
  memset(C, 0, sizeof(double) * nx * ny * nz);
#define CA(i, j, k, v) C1(i, j, k) = C1(i, j, k) + (v)
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) {
    double e = EX(ix - 1, iy - 1, iz - 1);
    CA(ix, iy - 1, iz - 1, e / 4.0); CA(ix, iy, iz - 1, e / 4.0); CA(ix, iy - 1, iz, e / 4.0); CA(ix, iy, iz, e / 4.0);
  }
  for (ix = 2; ix <= nx; ix++) for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) {
    double e = EY(ix - 1, iy - 1, iz - 1);
    CA(ix - 1, iy, iz - 1, e / 4.0); CA(ix, iy, iz - 1, e / 4.0); CA(ix - 1, iy, iz, e / 4.0); CA(ix, iy, iz, e / 4.0);
  }
  for (ix = 2; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) for (iz = 1; iz <= nz; iz++) {
    double e = EZ(ix - 1, iy - 1, iz - 1);
    CA(ix - 1, iy - 1, iz, e / 4.0); CA(ix - 1, iy, iz, e / 4.0); CA(ix, iy - 1, iz, e / 4.0); CA(ix, iy, iz, e / 4.0);
  }
  iz = 1;
  for (iy = 1; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) { double e = EY(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy, iz, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) { double e = EX(ix - 1, iy - 1, iz - 1); CA(ix, iy - 1, iz, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  iz = nz + 1;
  for (iy = 1; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) { double e = EY(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy, iz - 1, e / 2.0); CA(ix, iy, iz - 1, e / 2.0); }
  for (ix = 1; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) { double e = EX(ix - 1, iy - 1, iz - 1); CA(ix, iy - 1, iz - 1, e / 2.0); CA(ix, iy, iz - 1, e / 2.0); }
  ix = 1;
  for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) { double e = EY(ix - 1, iy - 1, iz - 1); CA(ix, iy, iz - 1, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  for (iz = 1; iz <= nz; iz++) for (iy = 2; iy <= ny; iy++) { double e = EZ(ix - 1, iy - 1, iz - 1); CA(ix, iy - 1, iz, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  ix = nx + 1;
  for (iy = 1; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) { double e = EY(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy, iz - 1, e / 2.0); CA(ix - 1, iy, iz, e / 2.0); }
  for (iz = 1; iz <= nz; iz++) for (iy = 2; iy <= ny; iy++) { double e = EZ(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy - 1, iz, e / 2.0); CA(ix - 1, iy, iz, e / 2.0); }
  iy = 1;
  for (ix = 1; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) { double e = EX(ix - 1, iy - 1, iz - 1); CA(ix, iy, iz - 1, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  for (iz = 1; iz <= nz; iz++) for (ix = 2; ix <= nx; ix++) { double e = EZ(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy, iz, e / 2.0); CA(ix, iy, iz, e / 2.0); }
  iy = ny + 1;
  for (ix = 1; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) { double e = EX(ix - 1, iy - 1, iz - 1); CA(ix, iy - 1, iz - 1, e / 2.0); CA(ix, iy - 1, iz, e / 2.0); }
  for (iz = 1; iz <= nz; iz++) for (ix = 2; ix <= nx; ix++) { double e = EZ(ix - 1, iy - 1, iz - 1); CA(ix - 1, iy - 1, iz, e / 2.0); CA(ix, iy - 1, iz, e / 2.0); }
  for (iz = 1; iz <= nz; iz++) {
    CA(1, 1, iz, EZ(0, 0, iz - 1));
    CA(1, ny, iz, EZ(0, ny, iz - 1));
    CA(nx, 1, iz, EZ(nx, 0, iz - 1));
    CA(nx, ny, iz, EZ(nx, ny, iz - 1));
  }
  for (iy = 1; iy <= ny; iy++) {
    CA(1, iy, 1, EY(0, iy - 1, 0));
    CA(1, iy, nz, EY(0, iy - 1, nz));
    CA(nx, iy, 1, EY(nx, iy - 1, 0));
    CA(nx, iy, nz, EY(nx, iy - 1, nz));
  }
  for (ix = 1; ix <= nx; ix++) {
    CA(ix, 1, 1, EX(ix - 1, 0, 0));
    CA(ix, 1, nz, EX(ix - 1, 0, nz));
    CA(ix, ny, 1, EX(ix - 1, ny, 0));
    CA(ix, ny, nz, EX(ix - 1, ny, nz));
  }
#undef CA
}

/* GridCalc.f90:864-1043 Cell2Node (accumulated term by term, /8, /4, /2) */
static void cell2node(const o_grid *g, const double *C, double *N) {
  int nx = g->nx, ny = g->ny, nz = g->nz, ix, iy, iz;
#define NV(i, j, k) N[(((k) - 1) * (ny + 1) + ((j) - 1)) * (nx + 1) + ((i) - 1)]
#define NA(i, j, k, v) NV(i, j, k) = NV(i, j, k) + (v)
  memset(N, 0, sizeof(double) * (nx + 1) * (ny + 1) * (nz + 1));
  for (ix = 2; ix <= nx; ix++) for (iy = 2; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) {
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 8.0); NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 8.0);
    NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 8.0);     NA(ix, iy, iz, C1(ix - 1, iy, iz) / 8.0);
    NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 8.0);     NA(ix, iy, iz, C1(ix, iy - 1, iz) / 8.0);
    NA(ix, iy, iz, C1(ix, iy, iz - 1) / 8.0);         NA(ix, iy, iz, C1(ix, iy, iz) / 8.0);
  }
  iy = 1;
  for (ix = 2; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) {
    NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 4.0); NA(ix, iy, iz, C1(ix, iy, iz - 1) / 4.0);
    NA(ix, iy, iz, C1(ix - 1, iy, iz) / 4.0);     NA(ix, iy, iz, C1(ix, iy, iz) / 4.0);
  }
  iy = ny + 1;
  for (ix = 2; ix <= nx; ix++) for (iz = 2; iz <= nz; iz++) {
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 4.0); NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 4.0);
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 4.0);     NA(ix, iy, iz, C1(ix, iy - 1, iz) / 4.0);
  }
  ix = 1;
  for (iy = 2; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) {
    NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 4.0); NA(ix, iy, iz, C1(ix, iy, iz - 1) / 4.0);
    NA(ix, iy, iz, C1(ix, iy - 1, iz) / 4.0);     NA(ix, iy, iz, C1(ix, iy, iz) / 4.0);
  }
  ix = nx + 1;
  for (iy = 2; iy <= ny; iy++) for (iz = 2; iz <= nz; iz++) {
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 4.0); NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 4.0);
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 4.0);     NA(ix, iy, iz, C1(ix - 1, iy, iz) / 4.0);
  }
  iz = 1;
  for (iy = 2; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) {
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 4.0); NA(ix, iy, iz, C1(ix, iy - 1, iz) / 4.0);
    NA(ix, iy, iz, C1(ix - 1, iy, iz) / 4.0);     NA(ix, iy, iz, C1(ix, iy, iz) / 4.0);
  }
  iz = nz + 1;
  for (iy = 2; iy <= ny; iy++) for (ix = 2; ix <= nx; ix++) {
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 4.0); NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 4.0);
    NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 4.0);     NA(ix, iy, iz, C1(ix, iy, iz - 1) / 4.0);
  }
  for (ix = 2; ix <= nx; ix++) {
    iy = 1; iz = 1;
    NA(ix, iy, iz, C1(ix - 1, iy, iz) / 2.0); NA(ix, iy, iz, C1(ix, iy, iz) / 2.0);
    iz = nz + 1;
    NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix, iy, iz - 1) / 2.0);
    iy = ny + 1; iz = 1;
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 2.0); NA(ix, iy, iz, C1(ix, iy - 1, iz) / 2.0);
    iz = nz + 1;
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 2.0);
  }
  for (iy = 2; iy <= ny; iy++) {
    ix = 1; iz = 1;
    NA(ix, iy, iz, C1(ix, iy - 1, iz) / 2.0); NA(ix, iy, iz, C1(ix, iy, iz) / 2.0);
    iz = nz + 1;
    NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix, iy, iz - 1) / 2.0);
    ix = nx + 1; iz = 1;
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 2.0); NA(ix, iy, iz, C1(ix - 1, iy, iz) / 2.0);
    iz = nz + 1;
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 2.0);
  }
  for (iz = 2; iz <= nz; iz++) {
    ix = 1; iy = 1;
    NA(ix, iy, iz, C1(ix, iy, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix, iy, iz) / 2.0);
    iy = ny + 1;
    NA(ix, iy, iz, C1(ix, iy - 1, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix, iy - 1, iz) / 2.0);
    ix = nx + 1; iy = 1;
    NA(ix, iy, iz, C1(ix - 1, iy, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix - 1, iy, iz) / 2.0);
    iy = ny + 1;
    NA(ix, iy, iz, C1(ix - 1, iy - 1, iz - 1) / 2.0); NA(ix, iy, iz, C1(ix - 1, iy - 1, iz) / 2.0);
  }
  NV(1, 1, 1) = C1(1, 1, 1);
  NV(nx + 1, 1, 1) = C1(nx, 1, 1);
  NV(1, ny + 1, 1) = C1(1, ny, 1);
  NV(nx + 1, ny + 1, 1) = C1(nx, ny, 1);
  NV(1, 1, nz + 1) = C1(1, 1, nz);
  NV(nx + 1, 1, nz + 1) = C1(nx, 1, nz);
  NV(1, ny + 1, nz + 1) = C1(1, ny, nz);
  NV(nx + 1, ny + 1, nz + 1) = C1(nx, ny, nz);
#undef NA
#undef NV
}

/* full-grid cell conductivity from LOGE parameters (ModelMap.inc:108-120) */
static double *cell_sigma(const sp2o_problem *p, const double *logcond) {
  const o_grid *g = &p->g;
  double *c = XALLOC(double, g->nx * g->ny * g->nz);
  double air = exp(p->airLog);
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++)
        c[CI(i, j, k)] = (k < g->nzAir) ? air
                          : exp(logcond[((k - g->nzAir) * g->ny + j) * g->nx + i]);
  return c;
}
static double *cell_volume(const o_grid *g) {
  double *v = XALLOC(double, g->nx * g->ny * g->nz);
  for (int k = 0; k < g->nz; k++) for (int j = 0; j < g->ny; j++) for (int i = 0; i < g->nx; i++)
    v[CI(i, j, k)] = g->dx[i] * g->dy[j] * g->dz[k];
  return v;
}
/* ModelParamToEdge ModelMap.inc:88-185: Econd = Cell2Edge(V_C*sigma)/Cell2Edge(V_C) */
static void model_to_edge(const sp2o_problem *p, const double *logcond, double *E) {
  const o_grid *g = &p->g;
  int nc = g->nx * g->ny * g->nz;
  double *s = cell_sigma(p, logcond), *vc = cell_volume(g), *ct = XALLOC(double, nc);
  double *ghost = XALLOC(double, p->nE), *et = XALLOC(double, p->nE);
  cell2edge(g, vc, ghost);
  for (int i = 0; i < nc; i++) ct[i] = vc[i] * s[i];
  cell2edge(g, ct, et);
  for (int i = 0; i < p->nE; i++) E[i] = et[i] / ghost[i];
  free(s); free(vc); free(ct); free(ghost); free(et);
}
/* ModelParamToNode ModelMap.inc:187-277 */
static void model_to_node(const sp2o_problem *p, const double *logcond, double *N) {
  const o_grid *g = &p->g;
  int nc = g->nx * g->ny * g->nz;
  double *s = cell_sigma(p, logcond), *vc = cell_volume(g), *ct = XALLOC(double, nc);
  double *ghost = XALLOC(double, p->nNode), *nt = XALLOC(double, p->nNode);
  cell2node(g, vc, ghost);
  for (int i = 0; i < nc; i++) ct[i] = vc[i] * s[i];
  cell2node(g, ct, nt);
  for (int i = 0; i < p->nNode; i++) N[i] = nt[i] / ghost[i];
  free(s); free(vc); free(ct); free(ghost); free(nt);
}
#undef C1
#undef EX
#undef EY
#undef EZ

/* ------------------------------------------------------ operator set-up */
/* modelOperator3D.f90:371-405 CurlCurlSetUp */
static void curlcurl_setup(sp2o_problem *p) {
  int nF = p->T.nrow;
  double *Dt = XALLOC(double, nF);
  o_csr Temp = rmat_x_diag(&p->T, p->EdgeL);
  for (int i = 0; i < nF; i++) Dt[i] = p->DualEdgeL[i] / p->FaceA[i];
  o_csr CC = diag_x_rmat(Dt, &Temp);
  csr_free(&Temp);
  o_csr Tt = rmat_trans(&p->T);
  Temp = rmat_x_rmat(&Tt, &CC);
  csr_free(&CC);
  CC = diag_x_rmat(p->EdgeL, &Temp);
  csr_free(&Temp);
  csr_free(&Tt);
  p->CCii = submatrix(&CC, p->EDGEi, p->nInt, p->EDGEi, p->nInt);
  p->CCib = submatrix(&CC, p->EDGEi, p->nInt, p->EDGEb, p->nBdry);
  csr_free(&CC);
  free(Dt);
}

/* modelOperator3D.f90:1086-1130 airNIndex */
static void air_nindex(const sp2o_problem *p, const double *sigEdge, double *Nair, double *Nearth) {
  o_csr Dt = rmat_trans(&p->G);
  int j, i;
  for (j = 0; j < nnz_of(&Dt); j++) Dt.val[j] = fabs(Dt.val[j]);
  double *Eearth = XALLOC(double, p->nE);
  const double thr = (double)1.1f * O_SIGMA_AIR; /* default-real 1.1 in the reference */
  for (i = 0; i < p->nE; i++) {
    double ea = 0.0, ee = 0.0;
    if (sigEdge[i] < thr) ea = 1.0; else ee = 1.0;
    if (sigEdge[i] == 0.0) { ea = 0.0; ee = 0.0; }
    (void)ea;
    Eearth[i] = ee;
  }
  rmat_x_rvec(&Dt, Eearth, Nair);
  for (i = 0; i < p->nNode; i++) {
    Nearth[i] = 0.0;
    if (Nair[i] > 0.0) { Nearth[i] = 1.0; Nair[i] = 0.0; } else { Nair[i] = 1.0; }
  }
  for (i = 0; i < p->nNb; i++) { Nearth[p->NODEb[i]] = 0.0; Nair[p->NODEb[i]] = 0.0; }
  free(Eearth);
  csr_free(&Dt);
}

/* modelOperator3D.f90:916-1008 GradDivSetup2 */
static void graddiv_setup2(sp2o_problem *p) {
  int Ne = p->nE, Nn = p->nNode, i;
  double *M1air = XALLOC(double, Ne), *M2air = XALLOC(double, Nn);
  double *M0e = XALLOC(double, Ne), *M1e = XALLOC(double, Ne), *M2e = XALLOC(double, Nn);
  double *M3e = XALLOC(double, Ne), *M4e = XALLOC(double, Ne);
  double *Nair = XALLOC(double, Nn), *Nearth = XALLOC(double, Nn);
  air_nindex(p, p->sigEdge, Nair, Nearth);
  for (i = 0; i < Ne; i++) {
    M1air[i] = p->DualFaceA[i];
    M0e[i] = p->DualFaceA[i];
    M1e[i] = p->sigEdge[i] * p->DualFaceA[i];
    M3e[i] = 1.0 / p->EdgeL[i];
    M4e[i] = p->DualFaceA[i];
  }
  for (i = 0; i < Nn; i++) {
    M2air[i] = Nair[i] / p->Vnode[i];
    M2e[i] = 1.0 / p->sigNode[i] * Nearth[i] / p->Vnode[i];
  }
  o_csr Dt = rmat_trans(&p->G);
  /* air part */
  o_csr GD = diag_x_rmat(M1air, &p->G);
  o_csr GDa = rmat_x_diag(&GD, M2air);
  csr_free(&GD);
  GD = rmat_x_rmat(&GDa, &Dt);
  csr_free(&GDa);
  GDa = rmat_x_diag(&GD, M1air);
  csr_free(&GD);
  /* earth part */
  GD = diag_x_rmat(M0e, &p->G);
  o_csr GDe = rmat_x_diag(&GD, M2e);
  csr_free(&GD);
  GD = rmat_x_rmat(&GDe, &Dt);
  csr_free(&GDe);
  GDe = rmat_x_diag(&GD, M1e);
  csr_free(&GD);
  GD = rmat_plus_rmat(&GDa, &GDe);
  csr_free(&GDa);
  csr_free(&GDe);
  o_csr GDii = submatrix(&GD, p->EDGEi, p->nInt, p->EDGEi, p->nInt);
  csr_free(&GD);
  p->AAii = rmat_plus_rmat(&p->CCii, &GDii);
  csr_free(&GDii);
  /* GD for the RHS correction terms */
  GDe = diag_x_rmat(M3e, &p->G);
  GD = rmat_x_diag(&GDe, M2e);
  csr_free(&GDe);
  GDe = rmat_x_rmat(&GD, &Dt);
  csr_free(&GD);
  GD = rmat_x_diag(&GDe, M4e);
  csr_free(&GDe);
  p->GDii = submatrix(&GD, p->EDGEi, p->nInt, p->EDGEi, p->nInt);
  csr_free(&GD);
  csr_free(&Dt);
  free(M1air); free(M2air); free(M0e); free(M1e); free(M2e); free(M3e); free(M4e);
  free(Nair); free(Nearth);
}

sp2o_problem *sp2o_create(int nx, int ny, int nzAir, int nzEarth, const double *dx,
                          const double *dy, const double *dz, const double *logcond,
                          double airLogCond) {
  sp2o_problem *p = XALLOC(sp2o_problem, 1);
  o_grid *g = &p->g;
  g->nx = nx; g->ny = ny; g->nzAir = nzAir; g->nzEarth = nzEarth; g->nz = nzAir + nzEarth;
  g->dx = XALLOC(double, nx); memcpy(g->dx, dx, sizeof(double) * nx);
  g->dy = XALLOC(double, ny); memcpy(g->dy, dy, sizeof(double) * ny);
  g->dz = XALLOC(double, g->nz); memcpy(g->dz, dz, sizeof(double) * g->nz);
  grid_setup(g);
  p->nXe = ncomp(g, XEDGE); p->nYe = ncomp(g, YEDGE); p->nZe = ncomp(g, ZEDGE);
  p->nE = p->nXe + p->nYe + p->nZe;
  p->nNode = ncomp(g, CORNER);
  p->logcond = XALLOC(double, nx * ny * nzEarth);
  memcpy(p->logcond, logcond, sizeof(double) * nx * ny * nzEarth);
  p->airLog = airLogCond;
  /* ModelDataInit modelOperator3D.f90:69-117 */
  p->T = curl_topology(g);
  p->G = grad_topology(g);
  boundary_index(p);
  metrics(p);
  curlcurl_setup(p);
  /* UpdateCond modelOperator3D.f90:167-213 */
  p->sigEdge = XALLOC(double, p->nE);
  p->sigNode = XALLOC(double, p->nNode);
  model_to_edge(p, logcond, p->sigEdge);
  model_to_node(p, logcond, p->sigNode);
  for (int i = 0; i < p->nBdry; i++) p->sigEdge[p->EDGEb[i]] = 0.0;
  graddiv_setup2(p);
  p->omega = 1.0;
  p->vms = XALLOC(double, p->nInt);
  for (int i = 0; i < p->nInt; i++)
    p->vms[i] = O_MU0 * p->omega * p->sigEdge[p->EDGEi[i]] * p->Vedge[p->EDGEi[i]];
  return p;
}

void sp2o_destroy(sp2o_problem *p) {
  if (!p) return;
  free(p->g.dx); free(p->g.dy); free(p->g.dz); free(p->g.delX); free(p->g.delY); free(p->g.delZ);
  free(p->EDGEi); free(p->EDGEb); free(p->NODEi); free(p->NODEb);
  free(p->FaceA); free(p->EdgeL); free(p->DualFaceA); free(p->DualEdgeL); free(p->Vnode); free(p->Vedge);
  free(p->sigEdge); free(p->sigNode); free(p->logcond); free(p->vms);
  csr_free(&p->T); csr_free(&p->G); csr_free(&p->CCii); csr_free(&p->CCib);
  csr_free(&p->AAii); csr_free(&p->GDii);
  if (p->have_lu) { csrz_free(&p->L); csrz_free(&p->U); }
  if (p->have_divcor) {
    csr_free(&p->VDiv); csr_free(&p->VDs); csr_free(&p->Gai);
    csr_free(&p->VDsG); csr_free(&p->VDsG_L); csr_free(&p->VDsG_U);
  }
  free(p);
}

void sp2o_sizes(const sp2o_problem *p, int *o) {
  o[0] = p->nE; o[1] = p->nInt; o[2] = p->nBdry; o[3] = p->nNode; o[4] = p->nNi; o[5] = p->nNb;
  o[6] = p->nxInt; o[7] = p->nyInt; o[8] = p->nzInt;
}
void sp2o_index(const sp2o_problem *p, int *ei, int *eb, int *ni, int *nb) {
  memcpy(ei, p->EDGEi, sizeof(int) * p->nInt);
  memcpy(eb, p->EDGEb, sizeof(int) * p->nBdry);
  memcpy(ni, p->NODEi, sizeof(int) * p->nNi);
  memcpy(nb, p->NODEb, sizeof(int) * p->nNb);
}
void sp2o_metric(const sp2o_problem *p, double *faceA, double *edgeL, double *dualFaceA,
                 double *dualEdgeL, double *vnode, double *vedge) {
  int nF = p->T.nrow;
  memcpy(faceA, p->FaceA, sizeof(double) * nF);
  memcpy(edgeL, p->EdgeL, sizeof(double) * p->nE);
  memcpy(dualFaceA, p->DualFaceA, sizeof(double) * p->nE);
  memcpy(dualEdgeL, p->DualEdgeL, sizeof(double) * nF);
  memcpy(vnode, p->Vnode, sizeof(double) * p->nNode);
  memcpy(vedge, p->Vedge, sizeof(double) * p->nE);
}
void sp2o_sigma(const sp2o_problem *p, double *sE, double *sN) {
  /* ModelParamToEdge itself, before UpdateCond zeroes the boundary */
  model_to_edge(p, p->logcond, sE);
  memcpy(sN, p->sigNode, sizeof(double) * p->nNode);
}
void sp2o_model2edge(const sp2o_problem *p, const double *logcond, double *sE) {
  model_to_edge(p, logcond, sE);
}

static const o_csr *pick(const sp2o_problem *p, int which) {
  switch (which) {
    case 0: return &p->AAii;
    case 1: return &p->CCii;
    case 2: return &p->CCib;
    case 3: return &p->GDii;
    case 4: return &p->T;
    case 5: return &p->G;
    case 6: return p->have_divcor ? &p->VDiv : NULL;
    case 7: return p->have_divcor ? &p->VDs : NULL;
    case 8: return p->have_divcor ? &p->Gai : NULL;
    case 9: return p->have_divcor ? &p->VDsG : NULL;
    case 10: return p->have_divcor ? &p->VDsG_L : NULL;
    case 11: return p->have_divcor ? &p->VDsG_U : NULL;
  }
  return NULL;
}
int sp2o_csr_nnz(const sp2o_problem *p, int which, int *nrow, int *ncol) {
  const o_csr *A = pick(p, which);
  if (!A) return -1;
  *nrow = A->nrow;
  *ncol = A->ncol;
  return nnz_of(A);
}
void sp2o_csr_get(const sp2o_problem *p, int which, int *row, int *col, double *val) {
  const o_csr *A = pick(p, which);
  if (!A) return;
  memcpy(row, A->row, sizeof(int) * (A->nrow + 1));
  memcpy(col, A->col, sizeof(int) * nnz_of(A));
  memcpy(val, A->val, sizeof(double) * nnz_of(A));
}

/* --------------------------------------------------- block DILU set-up */
/* CSR_R2Cdiag spOpTools.f90:1809-1825: B = A + ISIGN*i*diag(d) */
static o_csrz csr_r2cdiag(const o_csr *A, const double *d) {
  o_csrz B = csrz_new(A->nrow, A->ncol, nnz_of(A));
  memcpy(B.row, A->row, sizeof(int) * (A->nrow + 1));
  memcpy(B.col, A->col, sizeof(int) * nnz_of(A));
  for (int j = 0; j < nnz_of(A); j++) B.val[j] = A->val[j];
  for (int i = 0; i < A->nrow; i++)
    for (int j = B.row[i]; j < B.row[i + 1]; j++)
      if (B.col[j] == i) { B.val[j] = B.val[j] + (O_ISIGN * I) * d[i]; break; }
  return B;
}
/* lowerTri_Cmplx / upperTri_Cmplx spOpTools.f90:1188-1282 (diagonal kept) */
static o_csrz tri_part(const o_csrz *A, int lower) {
  int m = A->nrow, i, j, nz = 0;
  for (i = 0; i < m; i++)
    for (j = A->row[i]; j < A->row[i + 1]; j++)
      nz += lower ? (A->col[j] <= i) : (A->col[j] >= i);
  o_csrz B = csrz_new(m, A->ncol, nz);
  int k = 0;
  for (i = 0; i < m; i++) {
    B.row[i] = k;
    for (j = A->row[i]; j < A->row[i + 1]; j++)
      if (lower ? (A->col[j] <= i) : (A->col[j] >= i)) { B.col[k] = A->col[j]; B.val[k] = A->val[j]; k++; }
  }
  B.row[m] = k;
  return B;
}
/* CMATtrans spOpTools.f90:879-915 (conjugate = .true.) */
static o_csrz cmat_trans(const o_csrz *A) {
  int nz = A->row[A->nrow], i, j;
  o_csrz B = csrz_new(A->ncol, A->nrow, nz);
  int *cnt = XALLOC(int, A->ncol + 1);
  for (j = 0; j < nz; j++) cnt[A->col[j]]++;
  for (i = 0; i < A->ncol; i++) B.row[i + 1] = B.row[i] + cnt[i];
  memset(cnt, 0, sizeof(int) * (A->ncol + 1));
  for (i = 0; i < A->nrow; i++)
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      int r = A->col[j], pos = B.row[r] + cnt[r]++;
      B.col[pos] = i;
      B.val[pos] = conj(A->val[j]);
    }
  free(cnt);
  return B;
}
/* Dilu_Cmplx_AS spOpTools.f90:1996-2055 */
static void dilu_cmplx_as(const o_csrz *A, o_csrz *L, o_csrz *U) {
  *L = tri_part(A, 1);
  *U = tri_part(A, 0);
  o_csrz AT = cmat_trans(A);
  int m = A->nrow, i, j, k;
  zdouble *d = XALLOC(zdouble, m);
  for (i = 0; i < m; i++) {
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      if (A->col[j] == i) {
        d[i] = d[i] + A->val[j];
      } else if (A->col[j] < i) {
        for (k = AT.row[i]; k < AT.row[i + 1]; k++)
          if (AT.col[k] == A->col[j]) d[i] = d[i] - A->val[j] * AT.val[k] * d[A->col[j]];
      }
    }
    d[i] = zdiv(1.0, d[i]);
  }
  csrz_free(&AT);
  for (i = 0; i < m; i++)
    for (j = L->row[i]; j < L->row[i + 1]; j++)
      L->val[j] = (L->col[j] == i) ? (zdouble)1.0 : L->val[j] * d[L->col[j]];
  for (i = 0; i < m; i++)
    for (j = U->row[i]; j < U->row[i + 1]; j++)
      if (U->col[j] == i) { U->val[j] = zdiv(1.0, d[i]); break; }
  free(d);
}
/* BlkDiag_Cmplx spOpTools.f90:1474-1520 */
static o_csrz blkdiag(o_csrz *B, int nb) {
  int m = 0, n = 0, nz = 0, b, i, j;
  for (b = 0; b < nb; b++) { m += B[b].nrow; n += B[b].ncol; nz += B[b].row[B[b].nrow]; }
  o_csrz C = csrz_new(m, n, nz);
  int r0 = 0, c0 = 0, z0 = 0;
  for (b = 0; b < nb; b++) {
    for (i = 0; i < B[b].nrow; i++) C.row[r0 + i] = B[b].row[i] + z0;
    for (j = 0; j < B[b].row[B[b].nrow]; j++) { C.col[z0 + j] = B[b].col[j] + c0; C.val[z0 + j] = B[b].val[j]; }
    r0 += B[b].nrow; c0 += B[b].ncol; z0 += B[b].row[B[b].nrow];
  }
  C.row[m] = z0;
  return C;
}
/* PC_setup modelOperator3D.f90:512-611 */
static void pc_setup(sp2o_problem *p) {
  int starts[4] = {0, p->nxInt, p->nxInt + p->nyInt, p->nInt};
  o_csrz Lb[3], Ub[3];
  for (int b = 0; b < 3; b++) {
    int n = starts[b + 1] - starts[b];
    int *ix = XALLOC(int, n);
    double *d = XALLOC(double, n);
    for (int i = 0; i < n; i++) { ix[i] = starts[b] + i; d[i] = p->vms[starts[b] + i]; }
    o_csr Cxx = submatrix(&p->AAii, ix, n, ix, n);
    o_csrz Axx = csr_r2cdiag(&Cxx, d);
    dilu_cmplx_as(&Axx, &Lb[b], &Ub[b]);
    csr_free(&Cxx);
    csrz_free(&Axx);
    free(ix);
    free(d);
  }
  if (p->have_lu) { csrz_free(&p->L); csrz_free(&p->U); }
  p->L = blkdiag(Lb, 3);
  p->U = blkdiag(Ub, 3);
  p->have_lu = 1;
  for (int b = 0; b < 3; b++) { csrz_free(&Lb[b]); csrz_free(&Ub[b]); }
}

/* UpdateFreq modelOperator3D.f90:155-164 -> updateOmegaMuSig (:236-282) */
void sp2o_update_freq(sp2o_problem *p, double omega) {
  for (int i = 0; i < p->nInt; i++) {
    if (p->omega > 0) p->vms[i] = p->vms[i] / p->omega;
    p->vms[i] = p->vms[i] * omega;
  }
  p->omega = omega;
  pc_setup(p);
}
void sp2o_vms(const sp2o_problem *p, double *v) { memcpy(v, p->vms, sizeof(double) * p->nInt); }
int sp2o_lu_nnz(const sp2o_problem *p, int which) {
  if (!p->have_lu) return -1;
  return which ? p->U.row[p->U.nrow] : p->L.row[p->L.nrow];
}
void sp2o_lu_get(const sp2o_problem *p, int which, int *row, int *col, zdouble *val) {
  const o_csrz *A = which ? &p->U : &p->L;
  memcpy(row, A->row, sizeof(int) * (A->nrow + 1));
  memcpy(col, A->col, sizeof(int) * A->row[A->nrow]);
  memcpy(val, A->val, sizeof(zdouble) * A->row[A->nrow]);
}

/* -------------------------------------------------------- operator apply */
/* Mult_Aii modelOperator3D.f90:415-436 */
void sp2o_mult_aii(const sp2o_problem *p, int adjt, const zdouble *x, zdouble *y) {
  if (adjt) {
    o_csr At = rmat_trans(&p->AAii);
    rmat_x_cvec(&At, x, y);
    csr_free(&At);
    for (int i = 0; i < p->nInt; i++) y[i] = y[i] - (I * O_ISIGN) * p->vms[i] * x[i];
  } else {
    rmat_x_cvec(&p->AAii, x, y);
    for (int i = 0; i < p->nInt; i++) y[i] = y[i] + (I * O_ISIGN) * p->vms[i] * x[i];
  }
}
/* LTsolve_Cmplx / UTsolve_Cmplx spOpTools.f90:1664-1730 */
static void ltsolve_z(const o_csrz *L, const zdouble *b, zdouble *x) {
  for (int i = 0; i < L->nrow; i++) {
    zdouble s = b[i], d = 1.0;
    for (int j = L->row[i]; j < L->row[i + 1]; j++) {
      if (L->col[j] < i) s = s - L->val[j] * x[L->col[j]];
      else d = L->val[j];
    }
    x[i] = zdiv(s, d);
  }
}
static void utsolve_z(const o_csrz *U, const zdouble *b, zdouble *x) {
  for (int i = U->nrow - 1; i >= 0; i--) {
    zdouble s = b[i], d = 1.0;
    for (int j = U->row[i]; j < U->row[i + 1]; j++) {
      if (U->col[j] > i) s = s - U->val[j] * x[U->col[j]];
      else d = U->val[j];
    }
    x[i] = zdiv(s, d);
  }
}
void sp2o_pc_lsolve(const sp2o_problem *p, const zdouble *x, zdouble *y) { ltsolve_z(&p->L, x, y); }
void sp2o_pc_usolve(const sp2o_problem *p, const zdouble *x, zdouble *y) { utsolve_z(&p->U, x, y); }

/* ------------------------------------------------------------- BiCGStab */
static zdouble zdot(int n, const zdouble *a, const zdouble *b) { /* Fortran dot_product */
  zdouble s = 0;
  for (int i = 0; i < n; i++) s = s + conj(a[i]) * b[i];
  return s;
}
static double znrm(int n, const zdouble *a) { return sqrt(creal(zdot(n, a, a))); }

static void precond(const sp2o_problem *p, int kind, const zdouble *r, zdouble *tmp, zdouble *z) {
  int n = p->nInt;
  if (kind == 0) {
    ltsolve_z(&p->L, r, tmp);
    utsolve_z(&p->U, tmp, z);
  } else if (kind == 1) {
    for (int i = 0; i < n; i++) {
      double a = 0.0;
      for (int j = p->AAii.row[i]; j < p->AAii.row[i + 1]; j++)
        if (p->AAii.col[j] == i) a = p->AAii.val[j];
      z[i] = zdiv(r[i], a + (O_ISIGN * I) * p->vms[i]);
    }
  } else {
    memcpy(z, r, sizeof(zdouble) * n);
  }
}

/* BICG solver.f90:720-926 (gpu=0) and cuBiCG solver.f90:931-2009 (gpu=1) */
int sp2o_bicg(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
              int gpu, int pc, int *niter, double *rerr) {
  int n = p->nInt, iter, i, status = 1;
  for (i = 0; i < maxit; i++) rerr[i] = 0.0;
  double bnorm = znrm(n, b);
  if (isnan(bnorm)) { *niter = 0; return 3; }
  if (bnorm == 0.0) { for (i = 0; i < n; i++) x[i] = 0; *niter = 1; return 0; }
  zdouble *R = XALLOC(zdouble, n), *RT = XALLOC(zdouble, n), *P = XALLOC(zdouble, n);
  zdouble *PT = XALLOC(zdouble, n), *PH = XALLOC(zdouble, n), *S = XALLOC(zdouble, n);
  zdouble *SH = XALLOC(zdouble, n), *V = XALLOC(zdouble, n), *T = XALLOC(zdouble, n);
  zdouble *X0 = XALLOC(zdouble, n);
  sp2o_mult_aii(p, 0, x, R);
  for (i = 0; i < n; i++) R[i] = b[i] - R[i];
  double rnorm = znrm(n, R), btol = tol * bnorm, rnorm0 = rnorm;
  rerr[0] = (float)(rnorm / bnorm);
  if (rnorm <= btol) { *niter = 1; status = 0; goto done; }
  memcpy(X0, x, sizeof(zdouble) * n);
  zdouble RHO = 1.0, RHO1 = 1.0, ALPHA = 1.0, OMEGA = 1.0, BETA;
  if (!gpu) memcpy(RT, R, sizeof(zdouble) * n);
  *niter = maxit;
  for (iter = 1; iter <= maxit; iter++) {
    int restart = gpu ? ((iter % 120) == 1) : (iter == 1);
    if (gpu && restart) {
      memcpy(RT, R, sizeof(zdouble) * n);
      RHO1 = 1.0;
      RHO = zdot(n, RT, R);
      memcpy(P, R, sizeof(zdouble) * n);
      OMEGA = 1.0;
    } else {
      RHO1 = RHO;
      RHO = zdot(n, RT, R);
      if (!gpu && RHO == 0.0) { status = 2; break; }
      if (iter == 1) {
        memcpy(P, R, sizeof(zdouble) * n);
      } else {
        BETA = zdiv(RHO, RHO1) * zdiv(ALPHA, OMEGA);
        if (BETA == 0.0) { status = 2; break; }
        for (i = 0; i < n; i++) P[i] = R[i] + BETA * (P[i] - OMEGA * V[i]);
      }
    }
    precond(p, pc, P, PT, PH);
    sp2o_mult_aii(p, 0, PH, V);
    zdouble RTV = zdot(n, RT, V);
    if (!gpu && RTV == 0.0) { status = 2; break; }
    ALPHA = zdiv(RHO, RTV);
    if (ALPHA == 0.0) { status = 2; break; }
    for (i = 0; i < n; i++) x[i] = x[i] + ALPHA * PH[i];
    for (i = 0; i < n; i++) S[i] = R[i] - ALPHA * V[i];
    precond(p, pc, S, PT, SH);
    sp2o_mult_aii(p, 0, SH, T);
    zdouble TT = zdot(n, T, T);
    if (!gpu && TT == 0.0) { status = 2; break; }
    OMEGA = zdiv(zdot(n, T, S), TT);
    if (OMEGA == 0.0) { status = 2; break; }
    for (i = 0; i < n; i++) x[i] = x[i] + OMEGA * SH[i];
    for (i = 0; i < n; i++) R[i] = S[i] - OMEGA * T[i];
    rnorm = znrm(n, R);
    /* the CPU solver stores real(rnorm/bnorm), a default (single) real */
    rerr[iter - 1] = gpu ? rnorm / bnorm : (float)(rnorm / bnorm);
    if (rnorm < btol) {
      if (!gpu) { *niter = iter; status = 0; break; }
      /* cuBiCG re-checks with the true residual and keeps it (solver.f90:1855-1895) */
      sp2o_mult_aii(p, 0, x, R);
      for (i = 0; i < n; i++) R[i] = b[i] - R[i];
      rnorm = znrm(n, R);
      if (rnorm < btol) { rerr[iter - 1] = (float)(rnorm / bnorm); *niter = iter; status = 0; break; }
    }
    if (rnorm < rnorm0) { rnorm0 = rnorm; if (gpu) memcpy(X0, x, sizeof(zdouble) * n); }
  }
  if (gpu && status != 0) memcpy(x, X0, sizeof(zdouble) * n); /* best iterate */
  if (status == 2) *niter = maxit;
done:
  free(R); free(RT); free(P); free(PT); free(PH); free(S); free(SH); free(V); free(T); free(X0);
  return status;
}

/* -------------------------------------------------------------------- QMR */
/* PC_Lsolve / PC_Usolve with adjt = .true. (modelOperator3D.f90:622-644):
 * UTsolve_Cmplx(LH) and LTsolve_Cmplx(UH), LH = L^H and UH = U^H as PC_setup
 * forms them with CMATtrans (conjugate). */
static void pc_lsolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y) {
  o_csrz LH = cmat_trans(&p->L);
  utsolve_z(&LH, x, y);
  csrz_free(&LH);
}
static void pc_usolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y) {
  o_csrz UH = cmat_trans(&p->U);
  ltsolve_z(&UH, x, y);
  csrz_free(&UH);
}
void sp2o_pc_lsolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y) { pc_lsolve_h(p, x, y); }
void sp2o_pc_usolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y) { pc_usolve_h(p, x, y); }

/* QMR solver.f90:140-370: QMR without look-ahead (Barrett et al., Templates,
 * p. 24) with M1 = L, M2 = U of the block DILU; returns the last iterate,
 * rerr(niter) = the smallest relative residual seen.  Status: 0 converged
 * (niter < maxit), 1 not converged, 2 breakdown (the reference's `failed`;
 * rho or psi = 0, where the reference stops the program, also returns 2). */
int sp2o_qmr(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
             int *niter, double *rerr) {
  int n = p->nInt, i, iter, failed = 0;
  for (i = 0; i < maxit; i++) rerr[i] = 0.0;
  zdouble *R = XALLOC(zdouble, n), *VT = XALLOC(zdouble, n), *Y = XALLOC(zdouble, n);
  zdouble *Z = XALLOC(zdouble, n), *WT = XALLOC(zdouble, n), *V = XALLOC(zdouble, n);
  zdouble *W = XALLOC(zdouble, n), *YT = XALLOC(zdouble, n), *ZT = XALLOC(zdouble, n);
  zdouble *P = XALLOC(zdouble, n), *Q = XALLOC(zdouble, n), *PT = XALLOC(zdouble, n);
  zdouble *D = XALLOC(zdouble, n), *S = XALLOC(zdouble, n);
  sp2o_mult_aii(p, 0, x, R);
  for (i = 0; i < n; i++) R[i] = b[i] - R[i];
  double bnorm = creal(csqrt(zdot(n, b, b))), rnorm = creal(csqrt(zdot(n, R, R)));
  double rnormin = rnorm;
  if (isnan(fabs(bnorm))) { *niter = 0; failed = 3; goto done; }
  iter = 1;
  rerr[iter - 1] = (float)(rnorm / bnorm);
  memcpy(VT, R, sizeof(zdouble) * n);
  ltsolve_z(&p->L, VT, Y);
  zdouble RHO = csqrt(zdot(n, Y, Y));
  memcpy(WT, R, sizeof(zdouble) * n);
  pc_usolve_h(p, WT, Z);
  zdouble PSI = csqrt(zdot(n, Z, Z));
  zdouble GAMM = 1.0, ETA = -1.0, EPSIL = 0.0, THET = 0.0, THET1 = 0.0;
  while (rerr[iter - 1] > tol && iter < maxit) {
    if (RHO == 0.0 || PSI == 0.0) { failed = 2; break; }
    zdouble rhoInv = zdiv(1.0, RHO) * CMPLX(1.0, 0.0);
    zdouble psiInv = zdiv(1.0, PSI) * CMPLX(1.0, 0.0);
    for (i = 0; i < n; i++) V[i] = VT[i] * rhoInv;
    for (i = 0; i < n; i++) W[i] = WT[i] * psiInv;
    for (i = 0; i < n; i++) Y[i] = Y[i] * rhoInv;
    for (i = 0; i < n; i++) Z[i] = Z[i] * psiInv;
    zdouble DELTA = zdot(n, Z, Y);
    if (DELTA == 0.0) { failed = 2; break; }
    utsolve_z(&p->U, Y, YT);
    pc_lsolve_h(p, Z, ZT);
    if (iter == 1) {
      memcpy(P, YT, sizeof(zdouble) * n);
      memcpy(Q, ZT, sizeof(zdouble) * n);
    } else {
      zdouble PDE = zdiv(PSI * DELTA, EPSIL);
      zdouble RDE = RHO * conj(zdiv(DELTA, EPSIL));
      for (i = 0; i < n; i++) P[i] = YT[i] - PDE * P[i];
      for (i = 0; i < n; i++) Q[i] = ZT[i] - RDE * Q[i];
    }
    sp2o_mult_aii(p, 0, P, PT);
    EPSIL = zdot(n, Q, PT);
    if (EPSIL == 0.0) { failed = 2; break; }
    zdouble BETA = zdiv(EPSIL, DELTA);
    if (BETA == 0.0) { failed = 2; break; }
    for (i = 0; i < n; i++) VT[i] = PT[i] - BETA * V[i];
    zdouble RHO1 = RHO;
    ltsolve_z(&p->L, VT, Y);
    RHO = csqrt(zdot(n, Y, Y));
    sp2o_mult_aii(p, 1, Q, WT);
    for (i = 0; i < n; i++) WT[i] = WT[i] - conj(BETA) * W[i];
    pc_usolve_h(p, WT, Z);
    PSI = csqrt(zdot(n, Z, Z));
    if (iter > 1) THET1 = THET;
    THET = zdiv(RHO, GAMM * cabs(BETA));
    zdouble GAMM1 = GAMM;
    GAMM = zdiv(1.0, csqrt(1.0 + THET * THET));
    if (GAMM == 0.0) break;
    ETA = -zdiv(ETA * RHO1 * GAMM * GAMM, BETA * GAMM1 * GAMM1);
    if (iter == 1) {
      for (i = 0; i < n; i++) D[i] = ETA * P[i];
      for (i = 0; i < n; i++) S[i] = ETA * PT[i];
    } else {
      zdouble TM2 = THET1 * THET1 * GAMM * GAMM;
      for (i = 0; i < n; i++) D[i] = ETA * P[i] + TM2 * D[i];
      for (i = 0; i < n; i++) S[i] = ETA * PT[i] + TM2 * S[i];
    }
    for (i = 0; i < n; i++) x[i] = x[i] + D[i];
    for (i = 0; i < n; i++) R[i] = R[i] - S[i];
    rnorm = creal(csqrt(zdot(n, R, R)));
    if (rnorm < rnormin) rnormin = rnorm;
    iter = iter + 1;
    rerr[iter - 1] = (float)(rnorm / bnorm);
  }
  *niter = iter;
  rerr[iter - 1] = (float)(rnormin / bnorm);
done:
  free(R); free(VT); free(Y); free(Z); free(WT); free(V); free(W); free(YT); free(ZT);
  free(P); free(Q); free(PT); free(D); free(S);
  if (failed) return failed;
  return *niter < maxit ? 0 : 1;
}

/* ------------------------------------------------- divergence correction */
/* lowerTri_Real/upperTri_Real + Dilu_Real spOpTools.f90:1092-1186, 1903-1948 */
static o_csr tri_real(const o_csr *A, int lower) {
  int m = A->nrow, i, j, nz = 0;
  for (i = 0; i < m; i++) for (j = A->row[i]; j < A->row[i + 1]; j++) nz += lower ? (A->col[j] <= i) : (A->col[j] >= i);
  o_csr B = csr_new(m, A->ncol, nz);
  int k = 0;
  for (i = 0; i < m; i++) {
    B.row[i] = k;
    for (j = A->row[i]; j < A->row[i + 1]; j++)
      if (lower ? (A->col[j] <= i) : (A->col[j] >= i)) { B.col[k] = A->col[j]; B.val[k] = A->val[j]; k++; }
  }
  B.row[m] = k;
  return B;
}
static void dilu_real(const o_csr *A, o_csr *L, o_csr *U) {
  *L = tri_real(A, 1);
  *U = tri_real(A, 0);
  int m = A->nrow, i, j;
  double *d = XALLOC(double, m);
  for (i = 0; i < m; i++) {
    d[i] = 0.0;
    for (j = A->row[i]; j < A->row[i + 1]; j++) {
      if (A->col[j] == i) d[i] = d[i] + A->val[j];
      else if (A->col[j] < i) d[i] = d[i] - A->val[j] * A->val[j] * d[A->col[j]];
    }
    d[i] = 1.0 / d[i];
  }
  for (i = 0; i < m; i++)
    for (j = L->row[i]; j < L->row[i + 1]; j++)
      L->val[j] = (L->col[j] == i) ? 1.0 : L->val[j] * d[L->col[j]];
  for (i = 0; i < m; i++)
    for (j = U->row[i]; j < U->row[i + 1]; j++)
      if (U->col[j] == i) { U->val[j] = 1.0 / d[i]; break; }
  free(d);
}
/* DivCorInit + DivCorSetup modelOperator3D.f90:648-730 */
void sp2o_divcor_setup(sp2o_problem *p) {
  if (p->have_divcor) {
    csr_free(&p->VDiv); csr_free(&p->VDs); csr_free(&p->Gai);
    csr_free(&p->VDsG); csr_free(&p->VDsG_L); csr_free(&p->VDsG_U);
  }
  int i;
  o_csr Temp = rmat_trans(&p->G);
  o_csr Temp2 = rmat_x_diag(&Temp, p->DualFaceA);
  p->VDiv = submatrix(&Temp2, p->NODEi, p->nNi, p->EDGEi, p->nInt);
  csr_free(&Temp); csr_free(&Temp2);
  double *d = XALLOC(double, p->nE);
  for (i = 0; i < p->nE; i++) d[i] = 1. / p->EdgeL[i];
  int *all = XALLOC(int, p->nNode);
  for (i = 0; i < p->nNode; i++) all[i] = i;
  Temp = diag_x_rmat(d, &p->G);
  p->Gai = submatrix(&Temp, p->EDGEi, p->nInt, all, p->nNode);
  csr_free(&Temp);
  free(d);
  free(all);
  double *dd = XALLOC(double, p->nInt);
  for (i = 0; i < p->nInt; i++) {
    dd[i] = p->vms[i] / (O_MU0 * p->omega);
    dd[i] = -dd[i] / p->Vedge[p->EDGEi[i]];
  }
  p->VDs = rmat_x_diag(&p->VDiv, dd);
  free(dd);
  int *allr = XALLOC(int, p->Gai.nrow);
  for (i = 0; i < p->Gai.nrow; i++) allr[i] = i;
  Temp = submatrix(&p->Gai, allr, p->Gai.nrow, p->NODEi, p->nNi);
  p->VDsG = rmat_x_rmat(&p->VDs, &Temp);
  dilu_real(&p->VDsG, &p->VDsG_L, &p->VDsG_U);
  csr_free(&Temp);
  free(allr);
  p->have_divcor = 1;
}
/* LTsolve_Real / UTsolve_Real spOpTools.f90:1732-1798 */
static void ltsolve_r(const o_csr *L, const zdouble *b, zdouble *x) {
  for (int i = 0; i < L->nrow; i++) {
    zdouble s = b[i];
    double d = 1.0;
    for (int j = L->row[i]; j < L->row[i + 1]; j++) {
      if (L->col[j] < i) s = s - L->val[j] * x[L->col[j]];
      else d = L->val[j];
    }
    x[i] = zdivr(s, d);
  }
}
static void utsolve_r(const o_csr *U, const zdouble *b, zdouble *x) {
  for (int i = U->nrow - 1; i >= 0; i--) {
    zdouble s = b[i];
    double d = 1.0;
    for (int j = U->row[i]; j < U->row[i + 1]; j++) {
      if (U->col[j] > i) s = s - U->val[j] * x[U->col[j]];
      else d = U->val[j];
    }
    x[i] = zdivr(s, d);
  }
}
/* PCG solver.f90:57-136 with A = DivCgrad, Minv = DivCgradILU */
int sp2o_pcg(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
             int *niter, double *rerr) {
  int n = p->nNi, i, it;
  zdouble *r = XALLOC(zdouble, n), *s = XALLOC(zdouble, n), *pp = XALLOC(zdouble, n);
  zdouble *q = XALLOC(zdouble, n), *tmp = XALLOC(zdouble, n);
  zdouble delta, deltaOld = 1, beta, alpha;
  for (i = 0; i < maxit; i++) rerr[i] = 0.0;
  rmat_x_cvec(&p->VDsG, x, r);
  for (i = 0; i < n; i++) r[i] = b[i] - r[i];
  double bnorm = znrm(n, b), rnorm = znrm(n, r);
  it = 1;
  rerr[0] = (float)(rnorm / bnorm);
  while (rerr[it - 1] > tol && it < maxit) {
    ltsolve_r(&p->VDsG_L, r, tmp);
    utsolve_r(&p->VDsG_U, tmp, s);
    delta = zdot(n, r, s);
    if (it == 1) memcpy(pp, s, sizeof(zdouble) * n);
    else {
      beta = zdiv(delta, deltaOld);
      for (i = 0; i < n; i++) pp[i] = s[i] + beta * pp[i];
    }
    rmat_x_cvec(&p->VDsG, pp, q);
    alpha = zdiv(delta, zdot(n, pp, q));
    for (i = 0; i < n; i++) x[i] = x[i] + pp[i] * alpha;
    for (i = 0; i < n; i++) r[i] = r[i] - q[i] * alpha;
    deltaOld = delta;
    it++;
    rnorm = znrm(n, r);
    rerr[it - 1] = (float)(rnorm / bnorm);
  }
  *niter = it;
  free(r); free(s); free(pp); free(q); free(tmp);
  return rerr[it - 1] > tol;
}
void sp2o_divc(const sp2o_problem *p, const zdouble *e, zdouble *phi) { rmat_x_cvec(&p->VDs, e, phi); }
void sp2o_grad(const sp2o_problem *p, const zdouble *phi, zdouble *e) { rmat_x_cvec(&p->Gai, phi, e); }

/* --------------------------------------------------- RHS / post-process */
/* FWDsolve3D, forward case with boundary data and no source
 * (EMsolve3D.f90:261-274, 318): b = -CCib * e(EDGEb) */
void sp2o_rhs_fwd(const sp2o_problem *p, const zdouble *e_all, zdouble *b) {
  zdouble *eb = XALLOC(zdouble, p->nBdry);
  for (int i = 0; i < p->nBdry; i++) eb[i] = e_all[p->EDGEb[i]];
  rmat_x_cvec(&p->CCib, eb, b);
  for (int i = 0; i < p->nInt; i++) b[i] = -b[i];
  free(eb);
}
/* FWDsolve3D, transposed case with a dense source (EMsolve3D.f90:217-256):
 * b = s_i + V_E * GDii (s_i * iOmegaMuInv / V_E) */
void sp2o_rhs_trn(const sp2o_problem *p, double omega, const zdouble *s_all, zdouble *b) {
  int n = p->nInt;
  zdouble iOmegaMuInv = zdiv(1.0, CMPLX(0.0, 1.0 * O_ISIGN * omega * O_MU0));
  zdouble *si = XALLOC(zdouble, n), *st = XALLOC(zdouble, n);
  for (int i = 0; i < n; i++) si[i] = s_all[p->EDGEi[i]];
  for (int i = 0; i < n; i++) si[i] = zdivr(si[i] * iOmegaMuInv, p->Vedge[p->EDGEi[i]]);
  rmat_x_cvec(&p->GDii, si, st);
  for (int i = 0; i < n; i++) b[i] = s_all[p->EDGEi[i]] + st[i] * p->Vedge[p->EDGEi[i]];
  free(si); free(st);
}
/* FWDsolve3D post-processing for TRN without boundary data (EMsolve3D.f90:463-478):
 * e(EDGEi) = ei; s(EDGEb) = CCib^T ei; e = V_E e; e(EDGEb) = -s(EDGEb) */
void sp2o_post_trn(const sp2o_problem *p, const zdouble *ei, zdouble *e) {
  o_csr Ct = rmat_trans(&p->CCib);
  zdouble *yb = XALLOC(zdouble, p->nBdry);
  rmat_x_cvec(&Ct, ei, yb);
  for (int i = 0; i < p->nE; i++) e[i] = 0;
  for (int i = 0; i < p->nInt; i++) e[p->EDGEi[i]] = ei[i];
  for (int i = 0; i < p->nE; i++) e[i] = p->Vedge[i] * e[i];
  for (int i = 0; i < p->nBdry; i++) e[p->EDGEb[i]] = -yb[i];
  free(yb);
  csr_free(&Ct);
}

/* EdgeToModelParam ModelMap.inc:302-372: m = V_C * Edge2Cell(e / V_E_ghost) * exp(m0) */
void sp2o_edge2model(const sp2o_problem *p, const double *e, double *dm) {
  const o_grid *g = &p->g;
  int nc = g->nx * g->ny * g->nz;
  double *vc = cell_volume(g), *ghost = XALLOC(double, p->nE), *et = XALLOC(double, p->nE);
  double *ct = XALLOC(double, nc);
  cell2edge(g, vc, ghost);
  for (int i = 0; i < p->nE; i++) et[i] = e[i] / ghost[i];
  edge2cell(g, et, ct);
  for (int i = 0; i < nc; i++) ct[i] = vc[i] * ct[i];
  int ne = g->nx * g->ny * g->nzEarth, off = g->nx * g->ny * g->nzAir;
  for (int i = 0; i < ne; i++) dm[i] = ct[off + i] * exp(p->logcond[i]);
  free(vc); free(ghost); free(et); free(ct);
}
/* PmultT SolverSens.f90:63-116 (real part): dm = dEdgeToModelParam(Re(-i*ISIGN*mu0*omega * sum_pol e0.*e)) */
void sp2o_pmultT(const sp2o_problem *p, double omega, int npol, const zdouble *e0, const zdouble *e,
                 double *dm) {
  int n = p->nE;
  zdouble c = 0.0 + I * (-O_ISIGN * O_MU0 * omega);
  zdouble *acc = XALLOC(zdouble, n);
  double *re = XALLOC(double, n);
  for (int i = 0; i < n; i++) acc[i] = e0[i] * e[i];
  for (int k = 1; k < npol; k++)
    for (int i = 0; i < n; i++) acc[i] = acc[i] + e0[(size_t)k * n + i] * e[(size_t)k * n + i];
  for (int i = 0; i < n; i++) re[i] = creal(c * acc[i]);
  sp2o_edge2model(p, re, dm);
  free(acc);
  free(re);
}

/* ======================================================================
 * Data functionals: impedance responses (dataResp), their linearisations
 * (Lrows) and the adjoint source LmultT.
 * ====================================================================== */

/* math_constants.f90:32 D2R = PI/180 (truncated PI); :43 LARGE_REAL = 1.0e13
 * is written as a default-kind (single precision) literal, so its value is the
 * float nearest 1e13, 9999998976. */
#define O_D2R (O_PI / 180.0)
static const double O_LARGE_REAL = (double)1.0e13f;

/* utilities.f90:100-114 clean(x) = dnint(x*LARGE_REAL)/LARGE_REAL */
static double o_clean(double x) { return round(x * O_LARGE_REAL) / O_LARGE_REAL; }

/* utilities.f90:196-225 minNode: 1-based ix with xNode(ix) <= x < xNode(ix+1);
 * 0 below the first node, n beyond the last. */
static int o_min_node(double x, const double *node, int n) {
  for (int i = 1; i <= n; i++)
    if (o_clean(node[i - 1]) > o_clean(x)) return i - 1;
  return n;
}

/* node and centre coordinates, GridDef.f90:301-399 setup_grid (arrays are
 * 0-based here: xE[i-1] is the reference's xEdge(i)) */
typedef struct {
  double *xE, *yE, *zE, *xC, *yC, *zC;
} o_coords;

static void o_coords_make(const o_grid *g, const double org[3], o_coords *c) {
  int i;
  double cum;
  c->xE = XALLOC(double, g->nx + 1); c->yE = XALLOC(double, g->ny + 1);
  c->zE = XALLOC(double, g->nz + 1); c->xC = XALLOC(double, g->nx);
  c->yC = XALLOC(double, g->ny); c->zC = XALLOC(double, g->nz);
  c->xE[0] = org[0]; c->yE[0] = org[1]; c->zE[0] = org[2];
  cum = 0.0;
  for (i = 0; i < g->nx; i++) { cum = cum + g->dx[i]; c->xE[i + 1] = cum + org[0]; }
  cum = 0.0;
  for (i = 0; i < g->ny; i++) { cum = cum + g->dy[i]; c->yE[i + 1] = cum + org[1]; }
  cum = 0.0;
  for (i = 0; i < g->nz; i++) { cum = cum + g->dz[i]; c->zE[i + 1] = cum; }
  double airThick = c->zE[g->nzAir];
  cum = 0.0;
  for (i = 0; i < g->nx; i++) { cum = cum + g->delX[i]; c->xC[i] = cum + org[0]; }
  cum = 0.0;
  for (i = 0; i < g->ny; i++) { cum = cum + g->delY[i]; c->yC[i] = cum + org[1]; }
  cum = 0.0;
  for (i = 0; i < g->nz; i++) { cum = cum + g->delZ[i]; c->zC[i] = cum; }
  /* the top edge keeps its initial oz before the shift, as in the reference */
  for (i = 0; i < g->nz; i++) {
    c->zC[i] = c->zC[i] - airThick + org[2];
    c->zE[i] = c->zE[i] - airThick + org[2];
  }
  c->zE[g->nz] = c->zE[g->nz] - airThick + org[2];
}
static void o_coords_free(o_coords *c) {
  free(c->xE); free(c->yE); free(c->zE); free(c->xC); free(c->yC); free(c->zC);
}

/* sparsevecc (FIELDS/FiniteDiff3D/sg_sparse_vector.f90:78-102): coefficients
 * c on grid components (xyz, i, j, k), 1-based indices as in the reference. */
#define SV_MAX 1024
typedef struct {
  int n, face;
  int i[SV_MAX], j[SV_MAX], k[SV_MAX], xyz[SV_MAX];
  zdouble c[SV_MAX];
} o_svec;

static o_svec *sv_new(void) { return XALLOC(o_svec, 1); }

/* sg_sparse_vector.f90:430-501 linComb_sparsevecc: out = c1*a (in a's order),
 * then each entry of b is added to the FIRST matching entry of a, or appended. */
static void sv_lincomb(const o_svec *a, zdouble c1, const o_svec *b, zdouble c2, o_svec *out) {
  o_svec *t = sv_new();
  int old[SV_MAX];
  for (int m = 0; m < b->n; m++) {
    old[m] = -1;
    for (int n = 0; n < a->n; n++)
      if (a->xyz[n] == b->xyz[m] && a->i[n] == b->i[m] && a->j[n] == b->j[m] && a->k[n] == b->k[m]) {
        old[m] = n;
        break;
      }
  }
  t->face = a->face;
  for (int n = 0; n < a->n; n++) {
    t->i[n] = a->i[n]; t->j[n] = a->j[n]; t->k[n] = a->k[n]; t->xyz[n] = a->xyz[n];
    t->c[n] = c1 * a->c[n];
  }
  int nm = a->n;
  for (int m = 0; m < b->n; m++) {
    if (old[m] < 0) {
      if (nm >= SV_MAX) abort();
      t->i[nm] = b->i[m]; t->j[nm] = b->j[m]; t->k[nm] = b->k[m]; t->xyz[nm] = b->xyz[m];
      t->c[nm] = c2 * b->c[m];
      nm++;
    } else {
      t->c[old[m]] = t->c[old[m]] + c2 * b->c[m];
    }
  }
  t->n = nm;
  memcpy(out, t, sizeof(o_svec));
  free(t);
}

/* EMfieldInterp.f90:141-324 EinterpSetUp (tri-linear; no conductivity weights) */
static void o_einterp(const o_grid *g, const o_coords *co, const double x[3], int xyz, o_svec *L) {
  int nxMax = g->nx + 1, nyMax = g->ny + 1, nzMax = g->nz + 1;
  int i0 = o_min_node(x[0], co->xE, g->nx + 1);
  int j0 = o_min_node(x[1], co->yE, g->ny + 1);
  int k0 = o_min_node(x[2], co->zE, g->nz + 1);
  double w[3][2];
  if (xyz == 1) { i0 = o_min_node(x[0], co->xC, g->nx); nxMax = nxMax - 1; }
  else if (xyz == 2) { j0 = o_min_node(x[1], co->yC, g->ny); nyMax = nyMax - 1; }
  else { k0 = o_min_node(x[2], co->zC, g->nz); nzMax = nzMax - 1; }
  if (i0 > 0 && i0 < nxMax)
    w[0][1] = xyz == 1 ? (x[0] - co->xC[i0 - 1]) / g->delX[i0] : (x[0] - co->xE[i0 - 1]) / g->dx[i0 - 1];
  else w[0][1] = i0 <= 0 ? 1.0 : 0.0;
  if (j0 > 0 && j0 < nyMax)
    w[1][1] = xyz == 2 ? (x[1] - co->yC[j0 - 1]) / g->delY[j0] : (x[1] - co->yE[j0 - 1]) / g->dy[j0 - 1];
  else w[1][1] = j0 <= 0 ? 1.0 : 0.0;
  if (k0 > 0 && k0 < nzMax)
    w[2][1] = xyz == 3 ? (x[2] - co->zC[k0 - 1]) / g->delZ[k0] : (x[2] - co->zE[k0 - 1]) / g->dz[k0 - 1];
  else w[2][1] = k0 <= 0 ? 1.0 : 0.0;
  for (int d = 0; d < 3; d++) w[d][0] = 1.0 - w[d][1];
  L->n = 0;
  L->face = 0;
  for (int n = 0; n < 2; n++)
    for (int m = 0; m < 2; m++)
      for (int p = 0; p < 2; p++) {
        double wadd = w[0][n] * w[1][m] * w[2][p];
        if (wadd > 0) {
          int q = L->n++;
          L->i[q] = i0 + n; L->j[q] = j0 + m; L->k[q] = k0 + p; L->xyz[q] = xyz;
          L->c[q] = CMPLX(wadd, 0.0);
        }
      }
}

/* EMfieldInterp.f90:329-465 BinterpSetUp (face-centred H, tri-linear) */
static void o_binterp(const o_grid *g, const o_coords *co, const double x[3], int xyz, o_svec *L) {
  int nxMax = g->nx, nyMax = g->ny, nzMax = g->nz, i0, j0, k0;
  double w[3][2];
  if (xyz == 1) {
    i0 = o_min_node(x[0], co->xE, g->nx + 1); j0 = o_min_node(x[1], co->yC, g->ny);
    k0 = o_min_node(x[2], co->zC, g->nz); nxMax = nxMax + 1;
  } else if (xyz == 2) {
    i0 = o_min_node(x[0], co->xC, g->nx); j0 = o_min_node(x[1], co->yE, g->ny + 1);
    k0 = o_min_node(x[2], co->zC, g->nz); nyMax = nyMax + 1;
  } else {
    i0 = o_min_node(x[0], co->xC, g->nx); j0 = o_min_node(x[1], co->yC, g->ny);
    k0 = o_min_node(x[2], co->zE, g->nz + 1); nzMax = nzMax + 1;
  }
  if (i0 > 0 && i0 < nxMax)
    w[0][1] = xyz == 1 ? (x[0] - co->xE[i0 - 1]) / g->dx[i0 - 1] : (x[0] - co->xC[i0 - 1]) / g->delX[i0];
  else w[0][1] = i0 <= 0 ? 1.0 : 0.0;
  if (j0 > 0 && j0 < nyMax)
    w[1][1] = xyz == 2 ? (x[1] - co->yE[j0 - 1]) / g->dy[j0 - 1] : (x[1] - co->yC[j0 - 1]) / g->delY[j0];
  else w[1][1] = j0 <= 0 ? 1.0 : 0.0;
  if (k0 > 0 && k0 < nzMax)
    w[2][1] = xyz == 3 ? (x[2] - co->zE[k0 - 1]) / g->dz[k0 - 1] : (x[2] - co->zC[k0 - 1]) / g->delZ[k0];
  else w[2][1] = k0 <= 0 ? 1.0 : 0.0;
  for (int d = 0; d < 3; d++) w[d][0] = 1.0 - w[d][1];
  L->n = 0;
  L->face = 1;
  for (int n = 0; n < 2; n++)
    for (int m = 0; m < 2; m++)
      for (int p = 0; p < 2; p++) {
        double wadd = w[0][n] * w[1][m] * w[2][p];
        if (wadd > 0) {
          int q = L->n++;
          L->i[q] = i0 + n; L->j[q] = j0 + m; L->k[q] = k0 + p; L->xyz[q] = xyz;
          L->c[q] = CMPLX(wadd, 0.0);
        }
      }
}

/* EMfieldInterp.f90:470-647 BfromESetUp: H at x from the curl of E on the four
 * edges around each face of the B interpolation, h = e*sum(l)/S, times
 * i_omega_inv = ISIGN*i/omega.  l_E and S_F are GridCalc.f90:230-254 EdgeLength
 * and :373-418 FaceArea (dx, dy, dz and their products). */
static void o_bfrome(const o_grid *g, const o_coords *co, double omega, const double x[3], int xyz,
                     o_svec *L) {
  zdouble iw = O_ISIGN * CMPLX(0.0, 1.0 / omega);
  o_svec *LCH = sv_new(), *one = sv_new(), *tmp = sv_new();
  o_binterp(g, co, x, xyz, LCH);
  one->face = 0;
  for (int ii = 0; ii < LCH->n; ii++) {
    int EI[4], EJ[4], EK[4], AX[4];
    zdouble C[4];
    const int i = LCH->i[ii], j = LCH->j[ii], k = LCH->k[ii];
    /* l_E%x(i,.,.) = dx(i), %y = dy(j), %z = dz(k); S_F%x = dy*dz, %y = dx*dz, %z = dx*dy */
    if (LCH->xyz[ii] == 1) {
      int ax[4] = {2, 2, 3, 3}, jj[4] = {j, j, j, j + 1}, kk[4] = {k, k + 1, k, k};
      double S = g->dy[j - 1] * g->dz[k - 1];
      for (int q = 0; q < 4; q++) { AX[q] = ax[q]; EI[q] = i; EJ[q] = jj[q]; EK[q] = kk[q]; }
      C[0] = CMPLX(1 * g->dy[EJ[0] - 1] / S, 0.0);
      C[1] = CMPLX(-1 * g->dy[EJ[1] - 1] / S, 0.0);
      C[2] = CMPLX(-1 * g->dz[EK[2] - 1] / S, 0.0);
      C[3] = CMPLX(1 * g->dz[EK[3] - 1] / S, 0.0);
    } else if (LCH->xyz[ii] == 2) {
      int ax[4] = {3, 3, 1, 1}, ii4[4] = {i, i + 1, i, i}, kk[4] = {k, k, k, k + 1};
      double S = g->dx[i - 1] * g->dz[k - 1];
      for (int q = 0; q < 4; q++) { AX[q] = ax[q]; EI[q] = ii4[q]; EJ[q] = j; EK[q] = kk[q]; }
      C[0] = CMPLX(1 * g->dz[EK[0] - 1] / S, 0.0);
      C[1] = CMPLX(-1 * g->dz[EK[1] - 1] / S, 0.0);
      C[2] = CMPLX(-1 * g->dx[EI[2] - 1] / S, 0.0);
      C[3] = CMPLX(1 * g->dx[EI[3] - 1] / S, 0.0);
    } else {
      int ax[4] = {1, 1, 2, 2}, ii4[4] = {i, i, i, i + 1}, jj[4] = {j, j + 1, j, j};
      double S = g->dx[i - 1] * g->dy[j - 1];
      for (int q = 0; q < 4; q++) { AX[q] = ax[q]; EI[q] = ii4[q]; EJ[q] = jj[q]; EK[q] = k; }
      C[0] = CMPLX(1 * g->dx[EI[0] - 1] / S, 0.0);
      C[1] = CMPLX(-1 * g->dx[EI[1] - 1] / S, 0.0);
      C[2] = CMPLX(-1 * g->dy[EJ[2] - 1] / S, 0.0);
      C[3] = CMPLX(1 * g->dy[EJ[3] - 1] / S, 0.0);
    }
    if (ii == 0) {
      L->n = 4;
      L->face = 0;
      for (int q = 0; q < 4; q++) {
        L->i[q] = EI[q]; L->j[q] = EJ[q]; L->k[q] = EK[q]; L->xyz[q] = AX[q];
        L->c[q] = C[q] * LCH->c[0] * iw;
      }
    } else {
      memcpy(tmp, L, sizeof(o_svec));
      one->n = 4;
      for (int q = 0; q < 4; q++) {
        one->i[q] = EI[q]; one->j[q] = EJ[q]; one->k[q] = EK[q]; one->xyz[q] = AX[q]; one->c[q] = C[q];
      }
      sv_lincomb(tmp, CMPLX(1.0, 0.0), one, LCH->c[ii] * iw, L);
    }
  }
  free(LCH); free(one); free(tmp);
}

/* all-edge 0-based index of (xyz, i, j, k) (1-based), [x ; y ; z] ordering */
static int o_edge_index(const sp2o_problem *p, int xyz, int i, int j, int k) {
  int off = xyz == 1 ? 0 : (xyz == 2 ? p->nXe : p->nXe + p->nYe);
  return off + vidx(&p->g, XEDGE + xyz - 1, i - 1, j - 1, k - 1);
}

/* sg_sparse_vector.f90:669-728 dotProd_noConj_scvector_f */
static zdouble o_sdot(const sp2o_problem *p, const o_svec *L, const zdouble *e) {
  zdouble c = 0.0;
  for (int q = 0; q < L->n; q++) c = c + L->c[q] * e[o_edge_index(p, L->xyz[q], L->i[q], L->j[q], L->k[q])];
  return c;
}

/* sg_sparse_vector.f90:803-860 add_scvector: v = cs*L + v entry by entry */
static void o_sadd(const sp2o_problem *p, zdouble cs, const o_svec *L, zdouble *v) {
  for (int q = 0; q < L->n; q++) {
    int e = o_edge_index(p, L->xyz[q], L->i[q], L->j[q], L->k[q]);
    v[e] = cs * L->c[q] + v[e];
  }
}

/* DataFunc.f90:931-957 rotate_Model2Data */
static void o_rotate(zdouble *ch1, zdouble *ch2, double a1, double a2, double a) {
  zdouble t1 = *ch1, t2 = *ch2;
  double tmpv = 1.0 / (sin(a2 * O_D2R - a1 * O_D2R));
  zdouble r1 = sin(a2 * O_D2R - a * O_D2R) * t1 - cos(a2 * O_D2R - a * O_D2R) * t2;
  *ch1 = zdivr(r1, tmpv);
  zdouble r2 = -sin(a1 * O_D2R - a * O_D2R) * t1 + cos(a1 * O_D2R - a * O_D2R) * t2;
  *ch2 = zdivr(r2, tmpv);
}

/* inverse of the 2x2 matrix M (column j = mode) by Kramer's rule, in place,
 * DataFunc.f90:165-172 */
static void o_kramer(zdouble M[3][2]) {
  zdouble det = M[0][0] * M[1][1] - M[0][1] * M[1][0];
  zdouble t = M[0][0];
  M[0][0] = zdiv(M[1][1], det);
  M[1][1] = zdiv(t, det);
  M[0][1] = zdiv(-M[0][1], det);
  M[1][0] = zdiv(-M[1][0], det);
}

/* dataTypes.f90:87-222: number of real components and complex flag per type */
static int o_dt_ncomp(int dt, int *cplx) {
  static const int nc[12] = {0, 8, 4, 4, 8, 4, 4, 2, 2, 2, 2, 2};
  static const int cx[12] = {0, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 1};
  if (dt < 1 || dt > 11) return 0;
  if (cplx) *cplx = cx[dt];
  return nc[dt];
}

/* one receiver: location x, reference location r, azimuths (degrees):
 * Hx, Hy, Hz, Ex, Ey, Hx_ref, Hy_ref (fields_orientation.f90:23-58) */
typedef struct {
  double x[3], r[3], az[7];
} o_site;

/* DataFunc.f90:43-462 dataResp.  Z[6] is the module state the reference keeps
 * for Lrows (:37); Binv (2x2, [row][mode]) is its optional output. */
static void o_data_resp(const sp2o_problem *p, const o_coords *co, int dt, double omega,
                        const o_site *s, const zdouble *e1, const zdouble *e2, double *resp,
                        zdouble Z[6], zdouble Binv[2][2]) {
  const o_grid *g = &p->g;
  const zdouble *pol[2] = {e1, e2};
  zdouble EE[3][2], BB[3][2], RR[3][2], tz[4];
  double Hx = s->az[0], Hy = s->az[1], Ex = s->az[3], Ey = s->az[4], Hxr = s->az[5], Hyr = s->az[6];
  o_svec *Lex = sv_new(), *Ley = sv_new(), *Lbx = sv_new(), *Lby = sv_new(), *Lbz = sv_new(),
         *Lrx = sv_new(), *Lry = sv_new();
  int cplx = 0, ncomp = o_dt_ncomp(dt, &cplx), nfunc = cplx ? ncomp / 2 : ncomp;
  memset(EE, 0, sizeof(EE)); memset(BB, 0, sizeof(BB)); memset(RR, 0, sizeof(RR));
  switch (dt) {
    case 7: o_einterp(g, co, s->x, 1, Lex); for (int q = 0; q < 6; q++) Z[q] = o_sdot(p, Lex, e1); break;
    case 8: o_einterp(g, co, s->x, 2, Ley); for (int q = 0; q < 6; q++) Z[q] = o_sdot(p, Ley, e1); break;
    case 9: o_bfrome(g, co, omega, s->x, 1, Lbx); for (int q = 0; q < 6; q++) Z[q] = o_sdot(p, Lbx, e1); break;
    case 10: o_bfrome(g, co, omega, s->x, 2, Lby); for (int q = 0; q < 6; q++) Z[q] = o_sdot(p, Lby, e1); break;
    case 11: o_bfrome(g, co, omega, s->x, 3, Lbz); for (int q = 0; q < 6; q++) Z[q] = o_sdot(p, Lbz, e1); break;
    case 1: case 2: case 5: case 6:
      o_einterp(g, co, s->x, 1, Lex);
      o_einterp(g, co, s->x, 2, Ley);
      o_bfrome(g, co, omega, s->x, 1, Lbx);
      o_bfrome(g, co, omega, s->x, 2, Lby);
      for (int m = 0; m < 2; m++) {
        EE[0][m] = o_sdot(p, Lex, pol[m]);
        EE[1][m] = o_sdot(p, Ley, pol[m]);
        o_rotate(&EE[0][m], &EE[1][m], Ex, Ey, 0.0);
        BB[0][m] = o_sdot(p, Lbx, pol[m]);
        BB[1][m] = o_sdot(p, Lby, pol[m]);
        o_rotate(&BB[0][m], &BB[1][m], Hx, Hy, 0.0);
      }
      o_kramer(BB);
      if (dt == 1) {
        for (int j = 0; j < 2; j++)
          for (int i = 0; i < 2; i++) Z[2 * i + j] = EE[i][0] * BB[0][j] + EE[i][1] * BB[1][j];
      } else if (dt == 2) {
        Z[0] = EE[0][0] * BB[0][1] + EE[0][1] * BB[1][1];
        Z[1] = EE[1][0] * BB[0][0] + EE[1][1] * BB[1][0];
      } else if (dt == 5) {
        tz[0] = EE[0][0] * BB[0][1] + EE[0][1] * BB[1][1];
        tz[1] = EE[1][0] * BB[0][0] + EE[1][1] * BB[1][0];
        Z[0] = log10(pow(cabs(tz[0]), 2) * O_MU0 / omega);
        Z[1] = atan2(O_ISIGN * cimag(tz[0]), creal(tz[0]));
        Z[2] = log10(pow(cabs(tz[1]), 2) * O_MU0 / omega);
        Z[3] = atan2(O_ISIGN * cimag(tz[1]), creal(tz[1]));
      } else {
        for (int j = 0; j < 2; j++)
          for (int i = 0; i < 2; i++) tz[2 * i + j] = EE[i][0] * BB[0][j] + EE[i][1] * BB[1][j];
        double detX = creal(tz[0]) * creal(tz[3]) - creal(tz[1]) * creal(tz[2]);
        Z[0] = O_ISIGN * (creal(tz[3]) * cimag(tz[0]) - creal(tz[1]) * cimag(tz[2])) / detX;
        Z[1] = O_ISIGN * (creal(tz[3]) * cimag(tz[1]) - creal(tz[1]) * cimag(tz[3])) / detX;
        Z[2] = O_ISIGN * (creal(tz[0]) * cimag(tz[2]) - creal(tz[2]) * cimag(tz[0])) / detX;
        Z[3] = O_ISIGN * (creal(tz[0]) * cimag(tz[3]) - creal(tz[2]) * cimag(tz[1])) / detX;
      }
      break;
    case 3:
      o_bfrome(g, co, omega, s->x, 1, Lbx);
      o_bfrome(g, co, omega, s->x, 2, Lby);
      o_bfrome(g, co, omega, s->x, 3, Lbz);
      for (int m = 0; m < 2; m++) {
        BB[0][m] = o_sdot(p, Lbx, pol[m]);
        BB[1][m] = o_sdot(p, Lby, pol[m]);
        BB[2][m] = o_sdot(p, Lbz, pol[m]);
        o_rotate(&BB[0][m], &BB[1][m], Hx, Hy, 0.0);
      }
      o_kramer(BB);
      Z[0] = BB[2][0] * BB[0][0] + BB[2][1] * BB[1][0];
      Z[1] = BB[2][0] * BB[0][1] + BB[2][1] * BB[1][1];
      break;
    case 4:
      o_bfrome(g, co, omega, s->x, 1, Lbx);
      o_bfrome(g, co, omega, s->x, 2, Lby);
      o_bfrome(g, co, omega, s->r, 1, Lrx);
      o_bfrome(g, co, omega, s->r, 2, Lry);
      for (int m = 0; m < 2; m++) {
        BB[0][m] = o_sdot(p, Lbx, pol[m]);
        BB[1][m] = o_sdot(p, Lby, pol[m]);
        o_rotate(&BB[0][m], &BB[1][m], Hx, Hy, 0.0);
        RR[0][m] = o_sdot(p, Lrx, pol[m]);
        RR[1][m] = o_sdot(p, Lry, pol[m]);
        o_rotate(&RR[0][m], &RR[1][m], Hxr, Hyr, 0.0);
      }
      o_kramer(RR);
      for (int j = 0; j < 2; j++)
        for (int i = 0; i < 2; i++) Z[2 * i + j] = BB[i][0] * RR[0][j] + BB[i][1] * RR[1][j];
      Z[0] = Z[0] - 1.0;
      Z[3] = Z[3] - 1.0;
      break;
    default: break;
  }
  if (resp) {
    int ic = 0;
    for (int f = 0; f < nfunc; f++) {
      resp[ic++] = creal(Z[f]);
      if (cplx) resp[ic++] = cimag(Z[f]);
    }
  }
  if (Binv)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++) Binv[a][b] = dt == 4 ? RR[a][b] : BB[a][b];
  free(Lex); free(Ley); free(Lbx); free(Lby); free(Lbz); free(Lrx); free(Lry);
}

/* DataFunc.f90:465-869 Lrows: sparse rows of the linearised functionals of
 * one site, L[n][k] for functional n applied to mode k.  Returns nfunc. */
static int o_lrows(const sp2o_problem *p, const o_coords *co, int dt, double omega, const o_site *s,
                   const zdouble *e1, const zdouble *e2, o_svec *L[4][2]) {
  const o_grid *g = &p->g;
  zdouble Z[6] = {0}, Binv[2][2];
  int IJ[3][4] = {{0}}, ncomp = 0, hz = 0;
  switch (dt) {
    case 1: case 6:
      ncomp = 4;
      for (int j = 1; j <= 2; j++)
        for (int i = 1; i <= 2; i++) {
          IJ[0][2 * (i - 1) + j - 1] = i; IJ[1][2 * (i - 1) + j - 1] = j; IJ[2][2 * (i - 1) + j - 1] = i;
        }
      o_data_resp(p, co, 1, omega, s, e1, e2, NULL, Z, Binv);
      break;
    case 2: case 5:
      ncomp = 2;
      IJ[0][0] = 1; IJ[1][0] = 2; IJ[0][1] = 2; IJ[1][1] = 1; IJ[2][0] = 1; IJ[2][1] = 2;
      o_data_resp(p, co, 1, omega, s, e1, e2, NULL, Z, Binv);
      break;
    case 3:
      ncomp = 2; hz = 1;
      IJ[0][0] = 1; IJ[0][1] = 1; IJ[1][0] = 1; IJ[1][1] = 2; IJ[2][0] = 3; IJ[2][1] = 3;
      o_data_resp(p, co, 3, omega, s, e1, e2, NULL, Z, Binv);
      break;
    case 4:
      ncomp = 4;
      for (int j = 1; j <= 2; j++)
        for (int i = 1; i <= 2; i++) {
          IJ[0][2 * (i - 1) + j - 1] = i; IJ[1][2 * (i - 1) + j - 1] = j; IJ[2][2 * (i - 1) + j - 1] = i + 3;
        }
      o_data_resp(p, co, 4, omega, s, e1, e2, NULL, Z, Binv);
      break;
    case 11: hz = 1; ncomp = 1; break;
    default: ncomp = 1; break;
  }
  o_svec *Lex = sv_new(), *Ley = sv_new(), *Lbx = sv_new(), *Lby = sv_new(), *Lbz = sv_new(),
         *Lrx = sv_new(), *Lry = sv_new(), *Lexr = sv_new(), *Leyr = sv_new(), *Lbxr = sv_new(),
         *Lbyr = sv_new(), *Lrxr = sv_new(), *Lryr = sv_new(), *L1 = sv_new(), *L2 = sv_new(),
         *L3 = sv_new(), *P[4];
  for (int q = 0; q < 4; q++) P[q] = sv_new();
  o_einterp(g, co, s->x, 1, Lex);
  o_einterp(g, co, s->x, 2, Ley);
  o_bfrome(g, co, omega, s->x, 1, Lbx);
  o_bfrome(g, co, omega, s->x, 2, Lby);
  if (hz) o_bfrome(g, co, omega, s->x, 3, Lbz);
  o_bfrome(g, co, omega, s->r, 1, Lrx);
  o_bfrome(g, co, omega, s->r, 2, Lry);
  double Hx = s->az[0] * O_D2R, Hy = s->az[1] * O_D2R, Ex = s->az[3] * O_D2R, Ey = s->az[4] * O_D2R;
  double Hxr = s->az[5] * O_D2R, Hyr = s->az[6] * O_D2R;
  sv_lincomb(Lex, sin(Ey) / sin(Ey - Ex), Ley, -cos(Ey) / sin(Ey - Ex), Lexr);
  sv_lincomb(Lex, -sin(Ex) / sin(Ey - Ex), Ley, cos(Ex) / sin(Ey - Ex), Leyr);
  sv_lincomb(Lbx, sin(Hy) / sin(Hy - Hx), Lby, -cos(Hy) / sin(Hy - Hx), Lbxr);
  sv_lincomb(Lbx, -sin(Hx) / sin(Hy - Hx), Lby, cos(Hx) / sin(Hy - Hx), Lbyr);
  sv_lincomb(Lrx, sin(Hyr) / sin(Hyr - Hxr), Lry, -cos(Hyr) / sin(Hyr - Hxr), Lrxr);
  sv_lincomb(Lrx, -sin(Hxr) / sin(Hyr - Hxr), Lry, cos(Hxr) / sin(Hyr - Hxr), Lryr);
  const zdouble one = CMPLX(1.0, 0.0), zero = CMPLX(0.0, 0.0);
  if (dt >= 7) {
    const o_svec *src = dt == 7 ? Lex : dt == 8 ? Ley : dt == 9 ? Lbx : dt == 10 ? Lby : Lbz;
    memcpy(L[0][0], src, sizeof(o_svec));
    L[0][1]->n = 0;   /* created empty, never assigned */
  } else {
    for (int n = 0; n < ncomp; n++) {
      int i = IJ[0][n], j = IJ[1][n], pc = IJ[2][n];
      zdouble c1 = Z[2 * (i - 1)], c2 = Z[2 * (i - 1) + 1];
      if (dt == 4) sv_lincomb(Lrxr, c1, Lryr, c2, L1);
      else sv_lincomb(Lbxr, c1, Lbyr, c2, L1);
      for (int k = 0; k < 2; k++) {
        zdouble b = Binv[k][j - 1];
        const o_svec *row = pc == 1 ? Lexr : pc == 2 ? Leyr : pc == 3 ? Lbz : pc == 4 ? Lbxr : Lbyr;
        sv_lincomb(row, b, L1, -b, L[n][k]);
      }
    }
  }
  if (dt == 5) {
    for (int k = 0; k < 2; k++) {
      double a3 = cabs(Z[2]), a2 = cabs(Z[1]);
      /* complex/real quotients go through __divdc3 as in the reference */
      zdouble c = zdivr(CMPLX(0.0, 1.0) * conj(Z[2]), a3 * a3);
      sv_lincomb(L[1][k], c, L[1][k], zero, L[3][k]);
      c = zdivr(zdivr(2.0 * conj(Z[2]), a3 * a3), log(10.0));
      sv_lincomb(L[1][k], c, L[1][k], zero, L[2][k]);
      c = zdivr(CMPLX(0.0, 1.0) * conj(Z[1]), a2 * a2);
      sv_lincomb(L[0][k], c, L[0][k], zero, L[1][k]);
      c = zdivr(zdivr(2.0 * conj(Z[1]), a2 * a2), log(10.0));
      sv_lincomb(L[0][k], c, L[0][k], zero, L1);
      memcpy(L[0][k], L1, sizeof(o_svec));
    }
  }
  if (dt == 6) {
    double r1 = creal(Z[0]), r2 = creal(Z[1]), r3 = creal(Z[2]), r4 = creal(Z[3]);
    double m1 = cimag(Z[0]), m2 = cimag(Z[1]), m3 = cimag(Z[2]), m4 = cimag(Z[3]);
    for (int k = 0; k < 2; k++) {
      double detX = r1 * r4 - r2 * r3, PT[2][2];
      PT[0][0] = O_ISIGN * (r4 * m1 - r2 * m3) / detX;
      PT[0][1] = O_ISIGN * (r4 * m2 - r2 * m4) / detX;
      PT[1][0] = O_ISIGN * (r1 * m3 - r3 * m1) / detX;
      PT[1][1] = O_ISIGN * (r1 * m4 - r3 * m2) / detX;
      /* PTXX */
      sv_lincomb(L[0][k], CMPLX(-1.0 * PT[0][0] * r4 / detX, 0.0), L[1][k],
                 CMPLX((PT[0][0] * r3 - O_ISIGN * m3) / detX, 0.0), L1);
      sv_lincomb(L[2][k], CMPLX(PT[0][0] * r2 / detX, 0.0), L[3][k],
                 CMPLX((-1.0 * PT[0][0] * r1 + O_ISIGN * m1) / detX, 0.0), L2);
      sv_lincomb(L1, one, L2, one, L3);
      sv_lincomb(L[0][k], CMPLX(0.0, r4 / detX), L[2][k], CMPLX(0.0, -1.0 * r2 / detX), L1);
      sv_lincomb(L3, one, L1, one, P[0]);
      /* PTXY */
      sv_lincomb(L[0][k], CMPLX(-1.0 * PT[0][1] * r4 / detX, 0.0), L[1][k],
                 CMPLX((PT[0][1] * r3 - O_ISIGN * m4) / detX, 0.0), L1);
      sv_lincomb(L[2][k], CMPLX(PT[0][1] * r2 / detX, 0.0), L[3][k],
                 CMPLX((-1.0 * PT[0][1] * r1 + O_ISIGN * m2) / detX, 0.0), L2);
      sv_lincomb(L1, one, L2, one, L3);
      sv_lincomb(L[1][k], CMPLX(0.0, r4 / detX), L[3][k], CMPLX(0.0, -1.0 * r2 / detX), L1);
      sv_lincomb(L3, one, L1, one, P[1]);
      /* PTYX */
      sv_lincomb(L[0][k], CMPLX((-1.0 * PT[1][0] * r4 + O_ISIGN * m3) / detX, 0.0), L[1][k],
                 CMPLX(PT[1][0] * r3 / detX, 0.0), L1);
      sv_lincomb(L[2][k], CMPLX((PT[1][0] * r2 - O_ISIGN * m1) / detX, 0.0), L[3][k],
                 CMPLX(-1.0 * PT[1][0] * r1 / detX, 0.0), L2);
      sv_lincomb(L1, one, L2, one, L3);
      sv_lincomb(L[0][k], CMPLX(0.0, -1.0 * r3 / detX), L[2][k], CMPLX(0.0, r1 / detX), L1);
      sv_lincomb(L3, one, L1, one, P[2]);
      /* PTYY */
      sv_lincomb(L[0][k], CMPLX((-1.0 * PT[1][1] * r4 + O_ISIGN * m4) / detX, 0.0), L[1][k],
                 CMPLX(PT[1][1] * r3 / detX, 0.0), L1);
      sv_lincomb(L[2][k], CMPLX((PT[1][1] * r2 - O_ISIGN * m2) / detX, 0.0), L[3][k],
                 CMPLX(-1.0 * PT[1][1] * r1 / detX, 0.0), L2);
      sv_lincomb(L1, one, L2, one, L3);
      sv_lincomb(L[1][k], CMPLX(0.0, -1.0 * r3 / detX), L[3][k], CMPLX(0.0, r1 / detX), L1);
      sv_lincomb(L3, one, L1, one, P[3]);
      for (int q = 0; q < 4; q++) sv_lincomb(P[q], one, P[q], zero, L[q][k]);
    }
  }
  free(Lex); free(Ley); free(Lbx); free(Lby); free(Lbz); free(Lrx); free(Lry);
  free(Lexr); free(Leyr); free(Lbxr); free(Lbyr); free(Lrxr); free(Lryr);
  free(L1); free(L2); free(L3);
  for (int q = 0; q < 4; q++) free(P[q]);
  int cplx = 0, nc = o_dt_ncomp(dt, &cplx);
  return cplx ? nc / 2 : nc;
}

static void o_sites(int nsite, const double *x, const double *r, const double *az, o_site *S) {
  static const double az0[7] = {0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0};
  for (int s = 0; s < nsite; s++) {
    for (int d = 0; d < 3; d++) {
      S[s].x[d] = x[3 * s + d];
      S[s].r[d] = r ? r[3 * s + d] : 0.0;
    }
    for (int d = 0; d < 7; d++) S[s].az[d] = az ? az[7 * s + d] : az0[d];
  }
}

int sp2o_data_resp(const sp2o_problem *p, int dt, double omega, const double *origin, int nsite,
                   const double *x, const double *r, const double *az, const zdouble *e1,
                   const zdouble *e2, double *resp) {
  int cplx = 0, nc = o_dt_ncomp(dt, &cplx);
  if (!nc) return -1;
  o_coords co;
  o_coords_make(&p->g, origin, &co);
  o_site *S = XALLOC(o_site, nsite);
  o_sites(nsite, x, r, az, S);
  for (int s = 0; s < nsite; s++) {
    zdouble Z[6] = {0}, B[2][2];
    o_data_resp(p, &co, dt, omega, &S[s], e1, e2, resp + (size_t)s * nc, Z, B);
  }
  free(S);
  o_coords_free(&co);
  return nc;
}

/* SENS/DataSens.f90:130-225 LmultT for one transmitter and one data type:
 * comb_k = sum over sites and functionals of conj-data * L(n)%L(k), added
 * entry by entry (add_sparseVrhsV, SolnSpace.f90:1026-1067).  Complex data
 * enter as cmplx(re, -im), real data as cmplx(v, 0); a functional counts
 * when exist() of its first component is set. */
int sp2o_lmultT(const sp2o_problem *p, int dt, double omega, const double *origin, int nsite,
                const double *x, const double *r, const double *az, const double *d,
                const unsigned char *exist, const zdouble *e1, const zdouble *e2, zdouble *comb1,
                zdouble *comb2) {
  int cplx = 0, nc = o_dt_ncomp(dt, &cplx);
  if (!nc) return -1;
  o_coords co;
  o_coords_make(&p->g, origin, &co);
  o_site *S = XALLOC(o_site, nsite);
  o_sites(nsite, x, r, az, S);
  o_svec *L[4][2];
  for (int n = 0; n < 4; n++)
    for (int k = 0; k < 2; k++) L[n][k] = sv_new();
  for (int e = 0; e < p->nE; e++) comb1[e] = comb2[e] = 0.0;
  zdouble *comb[2] = {comb1, comb2};
  for (int s = 0; s < nsite; s++) {
    int nfunc = o_lrows(p, &co, dt, omega, &S[s], e1, e2, L);
    int ic = 0;
    for (int f = 0; f < nfunc; f++) {
      int ex = exist ? exist[(size_t)s * nc + ic] : 1;
      zdouble z;
      if (cplx) {
        z = CMPLX(d[(size_t)s * nc + ic], -d[(size_t)s * nc + ic + 1]);
        ic += 2;
      } else {
        z = CMPLX(d[(size_t)s * nc + ic], 0.0);
        ic += 1;
      }
      if (ex)
        for (int k = 0; k < 2; k++) o_sadd(p, z, L[f][k], comb[k]);
    }
  }
  for (int n = 0; n < 4; n++)
    for (int k = 0; k < 2; k++) free(L[n][k]);
  free(S);
  o_coords_free(&co);
  return nc;
}

/* ================================================================ 2-D TE boundary fields
 * BC_x0_WS (3D_MT/FWD/boundary_ws.f90:21-171): the boundary values and the
 * initial guess of a forward solve from one 2-D TE problem per grid slice,
 * solved by the Siripunvaraporn finite-difference code (3D_MT/FWD/Mod2d/).
 * Mode 1 (Ey polarisation): one problem per y-slice over (x, z), written to
 * the y-edges; mode 2 (Ex): one per x-slice over (y, z), written to the
 * x-edges.  All other edges are zero.
 *
 * The dense banded algebra is LAPACK's ZGBTRF/ZGBTRS, which the reference
 * links as a library; the vendored copy it ships (LAPACK/wsLAPACK.f90,
 * LAPACK/wsBLAS.f90) takes the unblocked ZGBTF2 path when KU <= 64 (ILAENV,
 * wsLAPACK.f90:5298-5311), i.e. for every 1-D solve and for 2-D slices with
 * nz <= 65.  The restatement below follows the published unblocked
 * algorithm (partial pivoting on |re|+|im|, first maximum) and the reference
 * BLAS operation order, so on those sizes it reproduces the vendored LAPACK
 * bit for bit; on larger nz the blocked factorisation rounds differently. */

static inline int o_imin(int a, int b) { return a < b ? a : b; }
static inline int o_imax(int a, int b) { return a > b ? a : b; }
static inline zdouble o_zmul(zdouble a, zdouble b) { /* textbook product, no contraction */
  return CMPLX(creal(a) * creal(b) - cimag(a) * cimag(b), creal(a) * cimag(b) + cimag(a) * creal(b));
}
static inline double o_cabs1(zdouble z) { return fabs(creal(z)) + fabs(cimag(z)); }

/* Banded LU with partial pivoting, unblocked (the ZGBTF2 algorithm).  ab is
 * column-major band storage with leading dimension ldab >= 2*kl+ku+1:
 * A(i,j) (1-based) sits at row kl+ku+1+i-j of column j.  Returns the first
 * zero pivot (0 if none); *nswap counts row interchanges. */
static int o_gbtf2(int m, int n, int kl, int ku, zdouble *ab, int ldab, int *ipiv, int *nswap) {
#define AB(i, j) ab[(size_t)((j) - 1) * ldab + ((i) - 1)]
  const int kv = ku + kl;
  int info = 0;
  for (int j = ku + 2; j <= o_imin(kv, n); ++j)
    for (int i = kv - j + 2; i <= kl; ++i) AB(i, j) = 0;
  int ju = 1;
  for (int j = 1; j <= o_imin(m, n); ++j) {
    if (j + kv <= n)
      for (int i = 1; i <= kl; ++i) AB(i, j + kv) = 0;
    const int km = o_imin(kl, m - j);
    int jp = 1;
    double smax = o_cabs1(AB(kv + 1, j));
    for (int i = 2; i <= km + 1; ++i) {
      const double v = o_cabs1(AB(kv + i, j));
      if (!(v <= smax)) { jp = i; smax = v; }
    }
    ipiv[j - 1] = jp + j - 1;
    if (AB(kv + jp, j) != 0) {
      ju = o_imax(ju, o_imin(j + ku + jp - 1, n));
      if (jp != 1) {
        /* swap matrix rows j and j+jp-1 over columns j..ju (band stride ldab-1) */
        zdouble *a = &AB(kv + jp, j), *b = &AB(kv + 1, j);
        for (int q = 0; q < ju - j + 1; ++q) {
          const zdouble t = a[(size_t)q * (ldab - 1)];
          a[(size_t)q * (ldab - 1)] = b[(size_t)q * (ldab - 1)];
          b[(size_t)q * (ldab - 1)] = t;
        }
        ++*nswap;
      }
      if (km > 0) {
        const zdouble r = zdiv(CMPLX(1.0, 0.0), AB(kv + 1, j));
        for (int i = 1; i <= km; ++i) AB(kv + 1 + i, j) = o_zmul(r, AB(kv + 1 + i, j));
        if (ju > j) {
          /* rank-1 update of the trailing band: A(j+i, j+c) -= l(i) u(c) */
          const zdouble *x = &AB(kv + 2, j), *y = &AB(kv, j + 1);
          zdouble *A = &AB(kv + 1, j + 1);
          for (int c = 0; c < ju - j; ++c) {
            const zdouble yc = y[(size_t)c * (ldab - 1)];
            if (yc != 0) {
              const zdouble t = o_zmul(CMPLX(-1.0, 0.0), yc);
              for (int i = 0; i < km; ++i)
                A[(size_t)c * (ldab - 1) + i] = A[(size_t)c * (ldab - 1) + i] + o_zmul(x[i], t);
            }
          }
        }
      }
    } else if (info == 0) {
      info = j;
    }
  }
  return info;
}

/* Solve A x = b with the factors of o_gbtf2 (the ZGBTRS 'N' algorithm with
 * one right-hand side: L with the interchanges, then the banded triangular
 * solve of U, ZTBSV order). */
static void o_gbtrs(int n, int kl, int ku, const zdouble *ab, int ldab, const int *ipiv, zdouble *b) {
  const int kd = ku + kl + 1;
  if (kl > 0)
    for (int j = 1; j <= n - 1; ++j) {
      const int lm = o_imin(kl, n - j), l = ipiv[j - 1];
      if (l != j) { const zdouble t = b[l - 1]; b[l - 1] = b[j - 1]; b[j - 1] = t; }
      if (b[j - 1] != 0) {
        const zdouble t = o_zmul(CMPLX(-1.0, 0.0), b[j - 1]);
        for (int i = 1; i <= lm; ++i) b[j - 1 + i] = b[j - 1 + i] + o_zmul(AB(kd + i, j), t);
      }
    }
  const int k = kl + ku, kp1 = k + 1;
  for (int j = n; j >= 1; --j)
    if (b[j - 1] != 0) {
      const int l = kp1 - j;
      b[j - 1] = zdiv(b[j - 1], AB(kp1, j));
      const zdouble t = b[j - 1];
      for (int i = j - 1; i >= o_imax(1, j - k); --i) b[i - 1] = b[i - 1] - o_zmul(t, AB(l + i, j));
    }
#undef AB
}

/* DistanceBetweenBlocks (WSutils.f90:182-193): c(1) = d(1), c(i) = d(i) + d(i-1),
 * c(n+1) = d(n).  c has n+1 entries. */
static void ws_cdist(int n, const double *d, double *c) {
  for (int i = 2; i <= n; ++i) c[i - 1] = d[i - 1] + d[i - 2];
  c[0] = d[0];
  c[n] = d[n - 1];
}

/* Fwd1D_TE (WSfwd1Dmod.f90:79-141) with Solve1D (:145-166): the 1-D TE field
 * of a layered column, 1 at the top node, over nz layers plus one layer of
 * the bottom skin depth.  s1: the nz-nza earth conductivities; x1d: nz+1. */
static void ws_fwd1d_te(double per, int nza, int nz, const double *dz, const double *s1, zdouble *x1d,
                        int *nswap) {
  const double omega = (2.0 * O_PI) / per, omue = omega * O_MU0;
  double *s1d = XALLOC(double, nz + 1), *dz1 = XALLOC(double, nz + 1);
  for (int iz = 1; iz <= nza; ++iz) s1d[iz - 1] = O_SIGMA_AIR;
  for (int iz = nza + 1; iz <= nz; ++iz) s1d[iz - 1] = s1[iz - nza - 1];
  s1d[nz] = s1d[nz - 1];
  const double skd = sqrt(2.0 / (s1d[nz] * omue));
  for (int iz = 1; iz <= nz; ++iz) dz1[iz - 1] = dz[iz - 1];
  dz1[nz] = skd;
  /* tridiagonal in band storage (kl = ku = 1, ldab = 4) */
  zdouble *a = XALLOC(zdouble, 4 * (size_t)nz), *xb = XALLOC(zdouble, nz);
  double *ad = XALLOC(double, nz);
  double au1 = 0.0;
  for (int jj = 1; jj <= nz; ++jj) {
    const int iz = jj + 1;
    const double au = 2.0 / dz1[iz - 2], adv = 2.0 / dz1[iz - 1];
    const double ss = s1d[iz - 1] * dz1[iz - 1] + s1d[iz - 2] * dz1[iz - 2];
    ad[jj - 1] = adv;
    a[(size_t)(jj - 1) * 4 + 2] = CMPLX((0.0 - au) - adv, omue * ss);
    if (jj == 1) au1 = au;
  }
  for (int jj = 2; jj <= nz; ++jj) a[(size_t)(jj - 1) * 4 + 1] = ad[jj - 2];
  for (int jj = 1; jj <= nz - 1; ++jj) a[(size_t)(jj - 1) * 4 + 3] = ad[jj - 1];
  xb[0] = CMPLX(-au1, 0.0);
  int *ipiv = XALLOC(int, nz);
  o_gbtf2(nz, nz, 1, 1, a, 4, ipiv, nswap);
  o_gbtrs(nz, 1, 1, a, 4, ipiv, xb);
  x1d[0] = 1.0;
  for (int i = 0; i < nz; ++i) x1d[i + 1] = xb[i];
  free(s1d); free(dz1); free(a); free(xb); free(ad); free(ipiv);
}

/* One 2-D TE problem (FwdTEmod.f90 Fwd2DsetupTE / UpdateCondTE / UpdateFreqTE /
 * SetBoundTE / Fwd2DsolveTE, with SetupA_TE, FormAII, SetBound2D_TE and
 * MulAibWithXb of WSfwd2Dmod.f90).  ny horizontal cells dy, nz vertical cells
 * dz of which nza are air; cond(ih, iz) at cond[(ih-1) + ny*(iz-1)] (the air
 * rows are not read: the 2-D code uses SIGMA_AIR there).  esol((ny+1), (nz+1))
 * at esol[(iy-1) + (ny+1)*(iz-1)]. */
static void ws_te2d(double per, int ny, int nz, int nza, const double *dy, const double *dz,
                    const double *cond, zdouble *esol, int *nswap) {
#define CC(iz, iy) ccon[(size_t)((iy) - 1) * nz + ((iz) - 1)]
  double *cz = XALLOC(double, nz + 1), *cy = XALLOC(double, ny + 1);
  ws_cdist(nz, dz, cz);
  ws_cdist(ny, dy, cy);
  /* SetupA_TE (WSfwd2Dmod.f90:345-438) */
  double *ccon = XALLOC(double, (size_t)nz * ny);
  for (int iy = 1; iy <= ny; ++iy)
    for (int iz = 1; iz <= nz; ++iz)
      CC(iz, iy) = iz <= nza ? O_SIGMA_AIR : cond[(size_t)(iy - 1) + (size_t)ny * (iz - 1)];
  const int mmi = (ny - 1) * (nz - 1), mmb = 2 * ny + 2 * nz;
  double *a1 = XALLOC(double, mmi), *a2 = XALLOC(double, mmi), *a3 = XALLOC(double, mmi),
         *a4 = XALLOC(double, mmi), *bte = XALLOC(double, mmb + 1);
  int jj = 0;
  for (int iy = 2; iy <= ny; ++iy)
    for (int iz = 2; iz <= nz; ++iz, ++jj) {
      const double s00 = CC(iz, iy), s10 = CC(iz - 1, iy), s01 = CC(iz, iy - 1), s11 = CC(iz - 1, iy - 1);
      const double s2d = (s00 * dz[iz - 1] * dy[iy - 1] + s01 * dz[iz - 1] * dy[iy - 2] +
                          s10 * dz[iz - 2] * dy[iy - 1] + s11 * dz[iz - 2] * dy[iy - 2]) /
                         (dz[iz - 1] * dy[iy - 1] + dz[iz - 1] * dy[iy - 2] + dz[iz - 2] * dy[iy - 1] +
                          dz[iz - 2] * dy[iy - 2]);
      const double ar = 2.0 * cz[iz - 1] / dy[iy - 1], al = 2.0 * cz[iz - 1] / dy[iy - 2];
      const double adv = 2.0 * cy[iy - 1] / dz[iz - 1], au = 2.0 * cy[iy - 1] / dz[iz - 2];
      a1[jj] = -ar - al - adv - au;
      a4[jj] = cy[iy - 1] * cz[iz - 1] * s2d;
      a2[jj] = iz == nz ? 0.0 : adv;
      a3[jj] = iy < ny ? ar : 0.0;
    }
  /* boundary couplings, 1-based positions as in the reference */
  for (int iz = 2; iz <= nz; ++iz) bte[iz] = 2.0 * cz[iz - 1] / dy[0];
  for (int iy = 2; iy <= ny; ++iy) bte[nz + iy] = 2.0 * cy[iy - 1] / dz[0];
  for (int iy = 2; iy <= ny; ++iy) bte[nz + ny + iy - 1] = 2.0 * cy[iy - 1] / dz[nz - 1];
  for (int iz = 2; iz <= nz; ++iz) bte[nz + 2 * ny - 1 + iz] = 2.0 * cz[iz - 1] / dy[ny - 1];
  /* FormAII (WSfwd2Dmod.f90:31-85): band storage and factorisation */
  const double omega = (2.0 * O_PI) / per, omue = omega * O_MU0;
  const int kl = nz - 1, ku = nz - 1, kc = kl + ku + 1, ldab = 2 * kl + ku + 1;
  zdouble *aii = XALLOC(zdouble, (size_t)ldab * mmi);
#define AII(i, j) aii[(size_t)((j) - 1) * ldab + ((i) - 1)]
  for (int j = 1; j <= mmi; ++j) AII(kc, j) = CMPLX(a1[j - 1], omue * a4[j - 1]);
  for (int j = 1; j <= mmi - 1; ++j) { AII(kc - 1, j + 1) = a2[j - 1]; AII(kc + 1, j) = a2[j - 1]; }
  for (int j = 1; j <= mmi - (nz - 1); ++j) {
    AII(kc - (nz - 1), j + (nz - 1)) = a3[j - 1];
    AII(kc + (nz - 1), j) = a3[j - 1];
  }
#undef AII
  int *ipiv = XALLOC(int, mmi);
  o_gbtf2(mmi, mmi, kl, ku, aii, ldab, ipiv, nswap);
  /* SetBound2D_TE (WSfwd2Dmod.f90:262-341): 1-D columns at the sides and,
   * dy-averaged, under every interior node of the bottom */
  zdouble *exb = XALLOC(zdouble, mmb + 1), *x1d = XALLOC(zdouble, nz + 1);
  double *s1 = XALLOC(double, nz);
  const int ne = nz - nza;
  for (int iz = 1; iz <= ne; ++iz) s1[iz - 1] = CC(iz + nza, 1);
  ws_fwd1d_te(per, nza, nz, dz, s1, x1d, nswap);
  for (int i = 1; i <= nz + 1; ++i) exb[i] = x1d[i - 1];
  for (int i = nz + 2; i <= nz + ny; ++i) exb[i] = 1.0;
  for (int iy = 2; iy <= ny; ++iy) {
    for (int iz = 1; iz <= ne; ++iz)
      s1[iz - 1] = (CC(iz + nza, iy - 1) * dy[iy - 2] + CC(iz + nza, iy) * dy[iy - 1]) / (dy[iy - 2] + dy[iy - 1]);
    ws_fwd1d_te(per, nza, nz, dz, s1, x1d, nswap);
    exb[nz + ny - 1 + iy] = x1d[nz];
  }
  for (int iz = 1; iz <= ne; ++iz) s1[iz - 1] = CC(iz + nza, ny);
  ws_fwd1d_te(per, nza, nz, dz, s1, x1d, nswap);
  for (int i = 0; i <= nz; ++i) exb[nz + 2 * ny + i] = x1d[i];
  /* MulAibWithXb (WSfwd2Dmod.f90:89-140): b = -A_ib x_b (the right side
   * overwrites, the top and bottom accumulate) */
  zdouble *exi = XALLOC(zdouble, mmi);
#define RL(a, x) CMPLX((a) * creal(x), (a) * cimag(x))
  for (int iz = 2; iz <= nz; ++iz) exi[iz - 2] = -RL(bte[iz], exb[iz]);
  for (int iz = 2, izz = 2 * ny + nz + 1; iz <= nz; ++iz, ++izz)
    exi[(ny - 2) * (nz - 1) + iz - 2] = -RL(bte[izz], exb[izz]);
  for (int iy = 2, izz = nz + 2; iy <= ny; ++iy, ++izz) {
    const int q = (iy - 2) * (nz - 1);
    exi[q] = exi[q] - RL(bte[izz], exb[izz]);
  }
  for (int iy = 2, izz = nz + ny + 1; iy <= ny; ++iy, ++izz) {
    const int q = (iy - 2) * (nz - 1) + nz - 2;
    exi[q] = exi[q] - RL(bte[izz], exb[izz]);
  }
#undef RL
  o_gbtrs(mmi, kl, ku, aii, ldab, ipiv, exi);
  /* Fwd2DsolveTE (FwdTEmod.f90:110-169): interior, then the four sides */
#define ES(iy, iz) esol[(size_t)((iy) - 1) + (size_t)(ny + 1) * ((iz) - 1)]
  int is = 0;
  for (int iy = 2; iy <= ny; ++iy)
    for (int iz = 2; iz <= nz; ++iz) ES(iy, iz) = exi[is++];
  for (int iz = 1; iz <= nz + 1; ++iz) ES(1, iz) = exb[iz];
  for (int iy = 2; iy <= ny; ++iy) ES(iy, 1) = exb[iy + nz];
  for (int iy = 2; iy <= ny; ++iy) ES(iy, nz + 1) = exb[iy + nz + ny - 1];
  for (int iz = 1; iz <= nz + 1; ++iz) ES(ny + 1, iz) = exb[nz + 2 * ny - 1 + iz];
#undef ES
#undef CC
  free(cz); free(cy); free(ccon); free(a1); free(a2); free(a3); free(a4); free(bte);
  free(aii); free(ipiv); free(exb); free(x1d); free(s1); free(exi);
}

int sp2o_bc_x0(int nx, int ny, int nzAir, int nzEarth, const double *dx, const double *dy,
               const double *dz, const double *logcond, double omega, int imode, zdouble *e0,
               int *nswap) {
  const int nz = nzAir + nzEarth;
  const int nXe = nx * (ny + 1) * (nz + 1), nYe = (nx + 1) * ny * (nz + 1), nZe = (nx + 1) * (ny + 1) * nz;
  memset(e0, 0, sizeof(zdouble) * (size_t)(nXe + nYe + nZe));
  *nswap = 0;
  if (imode != 1 && imode != 2) return -1;
  /* ComputeBC (FWD_SP2/modelOperator3D.f90:292-317) passes the period 2*PI/omega */
  const double per = (2.0 * O_PI) / omega;
  const int nh = imode == 2 ? ny : nx, nslice = imode == 2 ? nx : ny;
  const double *dh = imode == 2 ? dy : dx;
  double *cond = XALLOC(double, (size_t)nh * nz);
  zdouble *es = XALLOC(zdouble, (size_t)(nh + 1) * (nz + 1));
  for (int s = 1; s <= nslice; ++s) {
    /* sigma_C of the earth cells (ModelParamToCell, ModelMap.inc:26-66) */
    for (int ih = 1; ih <= nh; ++ih)
      for (int iz = nzAir + 1; iz <= nz; ++iz) {
        const int i = imode == 2 ? s : ih, j = imode == 2 ? ih : s;
        cond[(size_t)(ih - 1) + (size_t)nh * (iz - 1)] =
            exp(logcond[(size_t)(i - 1) + (size_t)nx * ((j - 1) + (size_t)ny * (iz - nzAir - 1))]);
      }
    ws_te2d(per, nh, nz, nzAir, dh, dz, cond, es, nswap);
    for (int ih = 1; ih <= nh + 1; ++ih)
      for (int iz = 1; iz <= nz + 1; ++iz) {
        const zdouble v = es[(size_t)(ih - 1) + (size_t)(nh + 1) * (iz - 1)];
        if (imode == 2) /* E0%x(iSlice, ih, iz) */
          e0[(size_t)(s - 1) + (size_t)nx * ((ih - 1) + (size_t)(ny + 1) * (iz - 1))] = v;
        else /* E0%y(ih, iSlice, iz) */
          e0[(size_t)nXe + (size_t)(ih - 1) + (size_t)(nx + 1) * ((s - 1) + (size_t)ny * (iz - 1))] = v;
      }
  }
  free(cond);
  free(es);
  return 0;
}
