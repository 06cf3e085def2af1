/*
 * sp2_oracle.h -- CPU restatement of ModEM's SP2 3D-MT forward/adjoint path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg may load it.  The product (libmtsp2) never links or calls it.
 *
 * Every routine follows a reference routine (cited in sp2_oracle.c) and keeps
 * its loop and summation order, so that operator assembly reproduces the
 * reference's FP64 values bit for bit.  Indices are 0-based here; the
 * reference is 1-based.  Vectors over edges use the reference ordering
 * [x-edges ; y-edges ; z-edges], each Fortran-ordered (i fastest).
 */
#ifndef SP2_ORACLE_H
#define SP2_ORACLE_H
#include <complex.h>
#include <stdint.h>

typedef double _Complex zdouble;

typedef struct {
  int nrow, ncol;
  int *row;    /* nrow+1, 0-based offsets */
  int *col;    /* 0-based column indices, reference order within a row */
  double *val;
} o_csr;

typedef struct {
  int nrow, ncol;
  int *row, *col;
  zdouble *val;
} o_csrz;

typedef struct sp2o_problem sp2o_problem;

/* grid + model set-up: ModelDataInit + UpdateCond */
sp2o_problem *sp2o_create(int nx, int ny, int nzAir, int nzEarth,
                          const double *dx, const double *dy, const double *dz,
                          const double *logcond /* nx*ny*nzEarth, Fortran order */,
                          double airLogCond);
void sp2o_destroy(sp2o_problem *p);

/* sizes: [nEdgeAll, nInt, nBdry, nNodeAll, nNodeInt, nNodeBdry, nxEdgeInt, nyEdgeInt, nzEdgeInt] */
void sp2o_sizes(const sp2o_problem *p, int *out9);
void sp2o_index(const sp2o_problem *p, int *edgei, int *edgeb, int *nodei, int *nodeb);
/* metric vectors: FaceA, EdgeL, DualFaceA, DualEdgeL, Vnode, Vedge */
void sp2o_metric(const sp2o_problem *p, double *faceA, double *edgeL, double *dualFaceA,
                 double *dualEdgeL, double *vnode, double *vedge);
void sp2o_sigma(const sp2o_problem *p, double *sigEdge, double *sigNode);

/* CSR export (0-based). which: 0 AAii, 1 CCii, 2 CCib, 3 GDii(RHS form), 4 T, 5 G,
 * 6 VDiv, 7 VDs, 8 Gai, 9 VDsG, 10 VDsG_L, 11 VDsG_U */
int sp2o_csr_nnz(const sp2o_problem *p, int which, int *nrow, int *ncol);
void sp2o_csr_get(const sp2o_problem *p, int which, int *row, int *col, double *val);

/* UpdateFreq: VomegaMuSig rescale + block-DILU set-up */
void sp2o_update_freq(sp2o_problem *p, double omega);
void sp2o_vms(const sp2o_problem *p, double *vms);
/* which: 0 L, 1 U */
int sp2o_lu_nnz(const sp2o_problem *p, int which);
void sp2o_lu_get(const sp2o_problem *p, int which, int *row, int *col, zdouble *val);

/* Mult_Aii (adjt=0) / transpose variant (adjt=1), PC_Lsolve, PC_Usolve */
void sp2o_mult_aii(const sp2o_problem *p, int adjt, const zdouble *x, zdouble *y);
void sp2o_pc_lsolve(const sp2o_problem *p, const zdouble *x, zdouble *y);
void sp2o_pc_usolve(const sp2o_problem *p, const zdouble *x, zdouble *y);

/* BiCGStab.  gpu_semantics=0: CPU BICG (solver.f90:720-926).
 * gpu_semantics=1: cuBiCG (restart every 120, true-residual recheck, best iterate).
 * precond: 0 = block DILU (reference), 1 = Jacobi (diag of A), 2 = none.
 * Returns 0 converged, 1 not converged, 2 breakdown. */
int sp2o_bicg(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
              int gpu_semantics, int precond, int *niter, double *rerr);

/* PC_Lsolve / PC_Usolve with adjt = .true.: y = L^-H x, y = U^-H x */
void sp2o_pc_lsolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y);
void sp2o_pc_usolve_h(const sp2o_problem *p, const zdouble *x, zdouble *y);
/* QMR (solver.f90:140-370), preconditioned by the block DILU's L and U.
 * Returns 0 converged (niter < maxit), 1 not converged, 2 breakdown, 3 NaN
 * right-hand side; x holds the last iterate, rerr[niter-1] the smallest
 * relative residual. */
int sp2o_qmr(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
             int *niter, double *rerr);

/* Divergence correction: DivCorInit + DivCorSetup, PCG (solver.f90:57-136) */
void sp2o_divcor_setup(sp2o_problem *p);
int sp2o_pcg(const sp2o_problem *p, const zdouble *b, zdouble *x, int maxit, double tol,
             int *niter, double *rerr);
void sp2o_divc(const sp2o_problem *p, const zdouble *e, zdouble *phi);   /* VDs * e  */
void sp2o_grad(const sp2o_problem *p, const zdouble *phi_all, zdouble *e); /* Gai * phi */

/* FWDsolve3D right-hand sides and post-processing (EMsolve3D.f90:217-320, 463-485) */
void sp2o_rhs_fwd(const sp2o_problem *p, const zdouble *e_all /* BC on boundary edges */,
                  zdouble *b);
void sp2o_rhs_trn(const sp2o_problem *p, double omega, const zdouble *s_all, zdouble *b);
void sp2o_post_trn(const sp2o_problem *p, const zdouble *ei, zdouble *e_all);

/* dEdgeToModelParam (ModelMap.inc:281-372) and PmultT (SolverSens.f90:63-116) */
void sp2o_edge2model(const sp2o_problem *p, const double *e_all, double *dm /* nx*ny*nzEarth */);
void sp2o_pmultT(const sp2o_problem *p, double omega, int npol, const zdouble *e0_all,
                 const zdouble *e_all, double *dm);

/* cell-conductivity -> edge/node maps alone (LOGE), for a fresh model vector */
void sp2o_model2edge(const sp2o_problem *p, const double *logcond, double *sigEdge);

/* Data functionals (DataFunc.f90:43-869, EMfieldInterp.f90:141-647) and the
 * adjoint source LmultT (SENS/DataSens.f90:130-225) for one transmitter.
 * dt: data type numbering of dataTypes.f90:65-77 (1..11).  Sites: x, r
 * (3*nsite, location and reference location in the grid frame), az (7*nsite,
 * degrees Hx,Hy,Hz,Ex,Ey,Hx_ref,Hy_ref; NULL = 0/90 defaults); origin = the
 * grid's (ox, oy, oz).  e1/e2: all-edge fields of the two polarisations.
 * resp / d / exist: nsite x ncomp(dt) (site-major).  Return ncomp, or -1. */
int sp2o_data_resp(const sp2o_problem *p, int dt, double omega, const double *origin, int nsite,
                   const double *x, const double *r, const double *az, const zdouble *e1,
                   const zdouble *e2, double *resp);
int sp2o_lmultT(const sp2o_problem *p, int dt, double omega, const double *origin, int nsite,
                const double *x, const double *r, const double *az, const double *d,
                const unsigned char *exist, const zdouble *e1, const zdouble *e2, zdouble *comb1,
                zdouble *comb2);

/* BC_x0_WS (3D_MT/FWD/boundary_ws.f90:21-171): the all-edge field E0 of the
 * 2-D TE slice solves for mode 1 (Ey polarisation) or 2 (Ex) at angular
 * frequency omega; the cell model is LOGE (air cells use SIGMA_AIR, as the
 * 2-D code does).  nswap receives the number of row interchanges the banded
 * LU made over all slices.  The forward problem's boundary data are E0 on
 * the boundary edges with the bottom-face x and y edges set to zero
 * (BC%xZMax = BC%yZMax = 0). */
int sp2o_bc_x0(int nx, int ny, int nzAir, int nzEarth, const double *dx, const double *dy,
               const double *dz, const double *logcond, double omega, int imode, zdouble *e0,
               int *nswap);

#endif
