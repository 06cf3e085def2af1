/*
 * mtsp2.h -- C ABI of libmtsp2, the MI355X (gfx950) SP2 3D-MT solver library.
 *
 * Drop-in for the GPU path of ModEM-GPU's SP2 forward/adjoint solver.  The
 * reference's Fortran host (f90/3D_MT/FWD_SP2) reaches its GPU through
 * ISO_C_BINDING interfaces (f90/3D_MT/FWD_SP2/cudaFortMap.f90:93-1408) inside
 * cuBiCG (f90/3D_MT/FWD_SP2/solver.f90:931-2009).  This header is what a
 * Fortran `bind(C)` interface block binds instead; the matching Fortran module
 * is modem-gpu_amd/fortran/mtsp2_fortran.f90.
 *
 * Conventions
 *  - Plain pointers and sizes only.  Complex numbers are C99 `double _Complex`
 *    (Fortran `complex(c_double_complex)`), stored interleaved (re, im).
 *  - Host pointers unless the name ends in `_dev` (device pointers).
 *  - Interior-edge vectors are in the reference's EDGEi order (x-, y-, then
 *    z-edges, each Fortran-ordered), length n_int.
 *  - Every function returns an int status (MT_OK = 0); mt_last_error() gives
 *    the message.  Nothing in the library calls exit()/stop, unlike the
 *    reference, whose host code stops on any CUDA error (solver.f90:1140-1143).
 *  - A context is bound to one device and one stream and is not thread-safe.
 */
#ifndef MTSP2_H
#define MTSP2_H

#include <stddef.h>
#ifdef __cplusplus
#include <complex>
typedef std::complex<double> mt_zdouble;
extern "C" {
#else
#include <complex.h>
typedef double _Complex mt_zdouble;
#endif

/* ---- status codes ---------------------------------------------------- */
#define MT_OK 0
#define MT_E_ARG 1        /* bad argument / size mismatch                     */
#define MT_E_DEVICE 2     /* HIP runtime error                                 */
#define MT_E_OOM 3        /* device allocation failed                          */
#define MT_E_NAN_RHS 4    /* right-hand side contains NaN (solver.f90:1500)    */
#define MT_E_BREAKDOWN 5  /* BiCGStab breakdown (alpha/beta/omega == 0)        */
#define MT_E_MAXIT 6      /* not converged within maxit (best iterate returned)*/
#define MT_E_STATE 7      /* operator / preconditioner not set                 */
#define MT_E_STENCIL 8    /* CSR does not fit the staggered-grid stencil       */
#define MT_E_PIVOT 9      /* a 2-D boundary slice where the reference's banded LU
                             would interchange rows (mt_bc_x0_dev)              */

/* ---- preconditioners and precisions ---------------------------------- */
#define MT_PC_DILU 0      /* block diagonal-ILU, the reference's PC_setup      */
#define MT_PC_JACOBI 1
#define MT_PC_NONE 2
#define MT_PC_DILU_RB 3   /* block DILU in red-black tile order: 2 launches per
                             sweep instead of one per tile level; changes the
                             iterates (not in the reference), meant for the
                             inner solves of MT_PREC_MIXED                    */
#define MT_PREC_FP64 0
#define MT_PREC_MIXED 1   /* FP32 inner BiCGStab + FP64 defect correction      */

typedef struct mt_ctx mt_ctx;

/* Solver controls.  Mirrors solverControl_t (solver.f90:25-36) plus the
 * options the reference hard-codes (restart interval 120, solver.f90:1545). */
typedef struct {
  int maxit;            /* max BiCGStab iterations (total, FP64 or inner)      */
  double tol;           /* relative residual ||b-Ax||/||b|| target             */
  int precond;          /* MT_PC_*                                             */
  int precision;        /* MT_PREC_*                                           */
  int restart;          /* restart interval; 0 -> 120 (the reference's)        */
  int best_iterate;     /* 1: return the smallest-residual iterate on failure  */
  double inner_tol;     /* mixed: inner relative reduction per refinement step (3e-4) */
  int max_outer;        /* mixed: max refinement steps                         */
  int op_transpose;     /* 1: solve with A^T (the adjoint systems); 0: A      */
  int chunk;            /* iterations per host status check / graph replay     */
  int use_graph;        /* 1: capture the iteration chunk in a hipGraph        */
} mt_solve_opts;

typedef struct {
  int niter;            /* total BiCGStab iterations (inner for mixed)         */
  int nouter;           /* refinement steps (mixed) / true-residual rechecks   */
  int status;           /* MT_OK, MT_E_BREAKDOWN, MT_E_MAXIT                   */
  double relres;        /* final FP64 true relative residual                   */
} mt_sys_stats;

void mt_default_opts(mt_solve_opts *opts);

/* ---- context ---------------------------------------------------------- */
/* Replaces the per-call stream/handle creation of cuBiCG (solver.f90:1118-
 * 1160) and kernelc_hookCtx (kernel_c.cu:341-460): one context, reused. */
int mt_ctx_create(int device, mt_ctx **ctx);
int mt_ctx_destroy(mt_ctx *ctx);
const char *mt_last_error(const mt_ctx *ctx);
/* Number of visible devices; replaces kernelc_getDevNum (kernel_c.cu:463-500). */
int mt_device_count(int *count);

/* ---- operator --------------------------------------------------------- */
/* Grid in cells (nz includes the air layers), as grid_t (GridDef.f90:34-91). */
int mt_set_grid(mt_ctx *ctx, int nx, int ny, int nz);

/* Uploads AAii (real CSR, n_int x n_int, as built by GradDivSetup2,
 * modelOperator3D.f90:916-1008) and the frequency-free diagonal
 * vms1 = mu0*sigma_E*V_E on interior edges (VomegaMuSig / omega,
 * modelOperator3D.f90:271).  index_base is 1 for the reference's Fortran CSR
 * (spOpTools.f90:11-28), 0 for C.  Replaces the per-solve H2D copies of
 * cuBiCG (solver.f90:1236-1263).  The CSR is converted to the device stencil
 * format; MT_E_STENCIL if an entry does not fit the staggered-grid stencil. */
int mt_operator_set_csr(mt_ctx *ctx, int n_int, int nnz, const int *rowptr, const int *col,
                        int index_base, const double *val, const double *vms1);

/* Builds AAii, vms1, V_E, CCib and the RHS-form GDii on the library side from
 * the grid and the LOGE cell conductivities (nx*ny*nzEarth, Fortran order),
 * restating ModelDataInit + UpdateCond (modelOperator3D.f90:69-117, 167-213).
 * Sets the grid too.  The operator is then device-resident. */
int mt_assemble(mt_ctx *ctx, int nx, int ny, int nzAir, int nzEarth, const double *dx,
                const double *dy, const double *dz, const double *logcond, double air_logcond);

/* UpdateCond (modelOperator3D.f90:167-213) on the device: a new LOGE model
 * (logcond, nx*ny*nzEarth, Fortran order, and the air value) on the grid of
 * the last mt_assemble.  sigma_E / sigma_N (cell averages, ModelMap.inc:88-277),
 * AAii = CCii + GDii(sigma) (GradDivSetup2 with airNIndex,
 * modelOperator3D.f90:916-1130) and vms1 are recomputed on the GPU; the
 * curl-curl part is the assembly's.  The device operator equals what
 * mt_assemble builds for the same model, bit for bit; host copies used by the
 * host-side entry points are rebuilt on their first use.  Replaces the
 * updateCond(sigma) call that runs on the CPU before every transmitter's
 * forward solve (ForwardSolver.f90:261). */
int mt_update_cond(mt_ctx *ctx, const double *logcond, double air_logcond);

/* Host-only assembly (no device needed): the same construction as
 * mt_assemble, returned as interior CSR (0-based, row-sorted) plus vms1.
 * Call with rowptr == NULL to get *n_int and *nnz, then with arrays. */
int mt_host_assemble(int nx, int ny, int nzAir, int nzEarth, const double *dx, const double *dy,
                     const double *dz, const double *logcond, double air_logcond, int *n_int,
                     int *nnz, int *rowptr, int *col, double *val, double *vms1);

/* Host-only block-DILU set-up (PC_setup, modelOperator3D.f90:512-611): the
 * inverse pivots 1/p (interior order, complex) of AAii + i*ISIGN*omega*vms1
 * given as interior CSR on grid (nx, ny, nz cells). */
int mt_host_dilu_pivots(int nx, int ny, int nz, int n_int, int nnz, const int *rowptr,
                        const int *col, int index_base, const double *val, const double *vms1,
                        double omega, mt_zdouble *dinv);

/* Copies of what mt_assemble built (host arrays sized by the caller):
 * n_int, nnz of AAii as CSR (0-based, row-sorted), and the arrays. */
int mt_operator_size(const mt_ctx *ctx, int *n_int, int *n_edge, int *nnz);
int mt_operator_get_csr(mt_ctx *ctx, int *rowptr, int *col, double *val, double *vms1);

/* ---- drop-in solve ------------------------------------------------------ */
/* cuBiCG(b, x, KSPiter, device_idx) (solver.f90:931-965): right-preconditioned
 * BiCGStab with conjugated inner products, restart every `restart` iterations,
 * true-residual recheck at convergence, best iterate returned on failure.
 * x holds x0 on entry.  rerr (length maxit, may be NULL) receives the relative
 * residual per iteration (KSPiter%rerr).  failed mirrors KSPiter%failed. */
int mt_bicg(mt_ctx *ctx, double omega, const mt_zdouble *b, mt_zdouble *x, int maxit,
            double tol, int precond, int *niter, double *rerr, int *failed);

/* QMR (solver.f90:140-370), the reference's default solver_name
 * (EMsolve3D.f90:70), which its CUDA build runs on the host
 * (EMsolve3D.f90:375-377): QMR without look-ahead, M1 = L and M2 = U of the
 * block DILU in the reference ordering (precond MT_PC_DILU; MT_PC_NONE for
 * none), A^H and the L^-H / U^-H solves for the dual recurrence.  Same
 * arguments as mt_bicg; as the reference, x returns the last iterate, rerr
 * (length maxit) the relative residual per iteration stored as a default
 * real, rerr[niter-1] the smallest one seen, and *niter < maxit means
 * converged.  Status MT_OK, MT_E_MAXIT, MT_E_BREAKDOWN (*failed = 1: rho,
 * psi, delta, epsilon or beta = 0, where the reference stops or exits) or
 * MT_E_NAN_RHS. */
int mt_qmr(mt_ctx *ctx, double omega, const mt_zdouble *b, mt_zdouble *x, int maxit, double tol,
           int precond, int *niter, double *rerr, int *failed);

/* ---- batched solve -------------------------------------------------------- */
/* nsys systems A(omega_s) x_s = b_s sharing AAii, solved together: both
 * polarizations and many frequencies in one launch sequence.  b and x are
 * nsys*n_int, system-major; x holds x0 on entry.  stats may be NULL. */
int mt_solve_batch(mt_ctx *ctx, int nsys, const double *omegas, const mt_zdouble *b,
                   mt_zdouble *x, const mt_solve_opts *opts, mt_sys_stats *stats);
/* Same with device-resident b/x (interior-edge order, system-major). */
int mt_solve_batch_dev(mt_ctx *ctx, int nsys, const double *omegas, const void *b_dev,
                       void *x_dev, const mt_solve_opts *opts, mt_sys_stats *stats);

/* Device buffers for callers without their own allocator (e.g. the
 * benchmark): allocate, free, and copy on the context's stream (kind: 0 H2D,
 * 1 D2H, 2 D2D; synchronous). */
int mt_dev_alloc(mt_ctx *ctx, size_t bytes, void **ptr);
int mt_dev_free(mt_ctx *ctx, void *ptr);
int mt_dev_copy(mt_ctx *ctx, void *dst, const void *src, size_t bytes, int kind);

/* ---- operator applications (Mult_Aii, PC_Lsolve/PC_Usolve) -------------- */
/* y = A(omega) x (Mult_Aii, modelOperator3D.f90:415-436).  op=0: AAii x +
 * i*ISIGN*omega*vms1 x; op=1: the transpose A^T; op=2: the conjugate transpose
 * A^H, which is what the reference's Mult_Aii(adjt=.true.) computes. */
int mt_apply(mt_ctx *ctx, double omega, int op, const mt_zdouble *x, mt_zdouble *y);
/* y = M^{-1} x for the preconditioner of frequency omega (L then U solves,
 * modelOperator3D.f90:622-644). */
int mt_precond_apply(mt_ctx *ctx, double omega, int precond, const mt_zdouble *x,
                     mt_zdouble *y);

/* ---- right-hand sides, post-processing (FWDsolve3D, EMsolve3D.f90:129-501) */
/* Forward RHS b = -CCib e_b from boundary values (all-edge vector e_all, only
 * its boundary entries are read).  Requires mt_assemble. */
int mt_rhs_fwd(mt_ctx *ctx, const mt_zdouble *e_all, mt_zdouble *b);
/* The same for a batch on the device, with the initial guess: for each of
 * nsys all-edge boundary fields (e_dev, [nsys][n_edge] complex), b = -CCib e_b
 * (Mult_Aib, modelOperator3D.f90:438-459) and, when x0_dev is not NULL,
 * x0 = e(EDGEi), the initial guess FWDsolve3D takes from the boundary field
 * (EMsolve3D.f90:217-274); b_dev, x0_dev: [nsys][n_int] in EDGEi order, ready
 * for mt_solve_batch_dev.  Requires mt_assemble. */
int mt_rhs_fwd_dev(mt_ctx *ctx, int nsys, const void *e_dev, void *b_dev, void *x0_dev);

/* BC_x0_WS (boundary_ws.f90:21-171, called through ComputeBC,
 * FWD_SP2/modelOperator3D.f90:292-317, from fwdSetup, ForwardSolver.f90:
 * 333-462) on the device, for a batch: the boundary values and initial guess
 * of nsys forward systems (omega_s, mode_s; mode 1 = Ey polarisation, one
 * 2-D TE problem per y-slice, mode 2 = Ex, one per x-slice) from the model
 * of the last mt_assemble / mt_update_cond.  e_dev ([nsys][n_edge] complex,
 * device) receives the field FWDsolve3D starts from: E0 on the interior
 * edges and the boundary conditions on the boundary edges (E0 there, with
 * the bottom-face x and y edges zero), ready for mt_rhs_fwd_dev.  Equal to
 * the reference's fields bit for bit while its partial pivoting keeps the
 * diagonal; MT_E_PIVOT (fields still written) where it would not.
 * Needs nz <= about 100 (the slice window lives in LDS).  Scratch, kept by
 * the context: the rows of U of the slices in flight, up to
 * 2*CUs*(nh-1)*(nz-1)*nz complex doubles (3.9 GB at 100x100x70 on 256 CUs). */
int mt_bc_x0_dev(mt_ctx *ctx, int nsys, const double *omegas, const int *modes, void *e_dev);
/* Transposed (adjoint) RHS of FWDsolve3D's TRN branch with a dense source
 * (EMsolve3D.f90:217-256): b = s_i + V_E GDii(s_i (i*ISIGN*omega*mu0)^-1 / V_E),
 * s_all on all edges, b on interior edges.  Host buffers; the work runs in
 * the kernels of mt_rhs_trn_dev.  Requires mt_assemble. */
int mt_rhs_trn(mt_ctx *ctx, double omega, const mt_zdouble *s_all, mt_zdouble *b);
/* TRN post-processing of FWDsolve3D without boundary data
 * (EMsolve3D.f90:463-478; the reference's nonzero_BC branch, :473-474, is
 * not covered: sensSolve never passes boundary data): e(EDGEi) = ei,
 * e = V_E e, e(EDGEb) = -CCib^T ei.  Host buffers; kernels of
 * mt_post_trn_dev. */
int mt_post_trn(mt_ctx *ctx, const mt_zdouble *ei, mt_zdouble *e_all);

/* Device forms for batches (nsys systems, system-major, device pointers):
 *  mt_post_fwd_dev: FWDsolve3D's forward post-processing (EMsolve3D.f90:463,
 *    481-485): e(EDGEi) = x on top of the boundary data already in e
 *    (x [nsys][n_int], e [nsys][n_edge], e.g. from mt_bc_x0_dev).
 *  mt_rhs_trn_dev: the TRN right-hand side above for each system (omegas[s],
 *    s_dev [nsys][n_edge], b_dev [nsys][n_int]); gd_correct = 0 gives
 *    b = s(EDGEi), the right-hand side of the exact transposed system
 *    A^T lam = s (no GDii term, no omega needed).
 *  mt_post_trn_dev: the TRN post-processing above (ei_dev [nsys][n_int],
 *    e_dev [nsys][n_edge]).
 * All require mt_assemble (sigma from the last mt_update_cond, if any). */
int mt_post_fwd_dev(mt_ctx *ctx, int nsys, const void *x_dev, void *e_dev);
int mt_rhs_trn_dev(mt_ctx *ctx, int nsys, const double *omegas, const void *s_dev, void *b_dev,
                   int gd_correct);
int mt_post_trn_dev(mt_ctx *ctx, int nsys, const void *ei_dev, void *e_dev);

/* ---- divergence correction (SdivCorr, EMsolve3D.f90:507-613) -------------- */
/* Node-Laplacian PCG (solver.f90:57-136) on VDsG = VDs*Gai with its DILU
 * (modelOperator3D.f90:648-763), FP64.  Corrects ei in place:
 * phi = VDsG^-1 (VDs ei - phi0), ei -= Gai phi.  phi0 (interior nodes) may be
 * NULL.  niter/relres may be NULL. */
int mt_divcor(mt_ctx *ctx, double omega, mt_zdouble *ei, const mt_zdouble *phi0, int maxit,
              double tol, int *niter, double *relres);

/* ---- gradient assembly (PmultT + dEdgeToModelParam) ----------------------- */
/* dm = Re(-i*ISIGN*mu0*omega * sum_pol e0_pol .* lam_pol) mapped to cells by
 * dEdgeToModelParam (SolverSens.f90:63-116, ModelMap.inc:281-372): all-edge
 * complex inputs (npol*n_edge), output nx*ny*nzEarth (Fortran order), added
 * into dm when accumulate != 0.  Requires mt_assemble. */
int mt_pmultT(mt_ctx *ctx, double omega, int npol, const mt_zdouble *e0_all,
              const mt_zdouble *lam_all, double *dm, int accumulate);
/* The same on the device for nfreq frequencies (JmultT summed over the
 * transmitters, SensComp.f90:408-485 per transmitter): e0_dev, lam_dev
 * [nfreq][npol][n_edge] complex (forward fields and TRN solutions as
 * FWDsolve3D returns them, all edges), dm_dev nx*ny*nzEarth doubles on the
 * device, overwritten (accumulate = 0) or added to.  The host form above
 * stages its arrays through the same kernels. */
int mt_pmultT_dev(mt_ctx *ctx, int nfreq, const double *omegas, int npol, const void *e0_dev,
                  const void *lam_dev, void *dm_dev, int accumulate);

/* Gradient of Re(sum_pol lam_pol^T A(omega) e_pol) w.r.t. the LOGE cell
 * parameters (nx*ny*nzEarth, Fortran order), accumulated into dm.  For a
 * misfit phi = 1/2 ||W (d - P e)||^2 with lam solving A^T lam = P^T conj(W^2
 * (d - P e)) this is d phi / d m.  e_all: forward fields on all edges
 * (npol*n_edge); lam: adjoint fields on interior edges (npol*n_int).
 * mode MT_GRAD_REFERENCE: only the i*ISIGN*omega*mu0*sigma_E*V_E diagonal, the
 *   term the reference's PmultT differentiates (SolverSens.f90:63-116);
 * mode MT_GRAD_EXACT: plus the grad-div part of AAii through sigma_E and
 *   sigma_N (GradDivSetup2, modelOperator3D.f90:916-1008): the exact discrete
 *   adjoint of the CC-GD system, which central differences confirm.
 * Has no counterpart entry point in the reference (its JmultT_TX,
 * SensComp.f90:408-485, chains PmultT and dEdgeToModelParam in Fortran). */
#define MT_GRAD_REFERENCE 0
#define MT_GRAD_EXACT 1
int mt_gradient(mt_ctx *ctx, double omega, int npol, const mt_zdouble *e_all,
                const mt_zdouble *lam, double *dm, int mode);
/* The same on the device for nfreq frequencies: e_dev [nfreq][npol][n_edge]
 * (forward fields, all edges), lam_dev [nfreq][npol][n_int] (adjoint
 * solutions of A^T lam = s, interior edges), dm_dev nx*ny*nzEarth doubles on
 * the device, overwritten (accumulate = 0) or added to.  The host form
 * mt_gradient stages its arrays through the same kernels; mt_host_gradient
 * is the host-only restatement the kernels are checked against. */
int mt_gradient_dev(mt_ctx *ctx, int nfreq, const double *omegas, int npol, const void *e_dev,
                    const void *lam_dev, void *dm_dev, int mode, int accumulate);
/* Host-only variant (assembles the model itself; for small problems/tests). */
int mt_host_gradient(int nx, int ny, int nzAir, int nzEarth, const double *dx, const double *dy,
                     const double *dz, const double *logcond, double air_logcond, double omega,
                     int npol, const mt_zdouble *e_all, const mt_zdouble *lam, double *dm,
                     int mode);

/* ---- tuning / introspection ------------------------------------------------ */
/* ---- data functionals (DataFunc.f90, EMfieldInterp.f90, SENS/DataSens.f90) */
/* Data types, numbered as dataTypes.f90:65-77; the number of real components
 * per site (ncomp) is 8, 4, 4, 8, 4, 4 for types 1-6 and 2 for types 7-11. */
#define MT_DT_FULL_IMPEDANCE 1
#define MT_DT_OFF_DIAGONAL_IMPEDANCE 2
#define MT_DT_FULL_VERTICAL_COMPONENTS 3
#define MT_DT_FULL_INTERSTATION_TF 4
#define MT_DT_OFF_DIAGONAL_RHO_PHASE 5
#define MT_DT_PHASE_TENSOR 6
#define MT_DT_EX_FIELD 7
#define MT_DT_EY_FIELD 8
#define MT_DT_BX_FIELD 9
#define MT_DT_BY_FIELD 10
#define MT_DT_BZ_FIELD 11
/* Receivers (the reference's rxDict, 3D_MT/DICT/receivers.f90, plus the
 * per-site orientations of dataBlock_t, SENS/DataSpace.f90:104): the grid
 * (cells, spacings, origin (ox, oy, oz) as in GridDef.f90:301-399), site
 * locations x (3*nsite), reference-site locations r (3*nsite, NULL = x) and
 * channel azimuths (7*nsite degrees Hx,Hy,Hz,Ex,Ey,Hx_ref,Hy_ref, NULL = the
 * defaults of fields_orientation.f90:38-58).  Builds the E and H
 * interpolation rows of EinterpSetUp / BfromESetUp (EMfieldInterp.f90:141-647)
 * on the host and keeps them on the device.  MT_E_ARG if a site's stencil
 * leaves the grid. */
int mt_receivers_set(mt_ctx *ctx, int nx, int ny, int nzAir, int nzEarth, const double *dx,
                     const double *dy, const double *dz, const double *origin, int nsite,
                     const double *x, const double *r, const double *azimuth);
/* Responses of data type dtype at every receiver for nfreq transmitters
 * (dataResp, DataFunc.f90:43-462): e_dev = device fields, all edges,
 * [freq][polarisation 1,2][n_edge] complex; resp = host
 * [freq][site][ncomp] (real and imaginary parts interleaved for complex
 * types, as dataResp's Resp). */
int mt_data_resp(mt_ctx *ctx, int dtype, int nfreq, const double *omegas, const void *e_dev,
                 double *resp);
/* Adjoint source of data d (LmultT, SENS/DataSens.f90:130-225, one dataBlock
 * per transmitter): comb_dev (device, [freq][mode 1,2][n_edge] complex,
 * overwritten) = sum over sites and functionals of z * L, with L the
 * linearised functionals of Lrows (DataFunc.f90:465-869) at the fields e_dev
 * and z = cmplx(d_re, -d_im) for complex types, cmplx(d, 0) for real ones.
 * d, exist: host [freq][site][ncomp]; exist (NULL = all present) is read at
 * each functional's first component. */
int mt_lmultT(mt_ctx *ctx, int dtype, int nfreq, const double *omegas, const void *e_dev,
              const double *d, const unsigned char *exist, void *comb_dev);
/* mt_lmultT without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops such as an inversion's
 * gradient loop).  No counterpart in the reference.  Same arguments and
 * results as mt_lmultT; its scratch (the per-(frequency, site) functional
 * coefficients, the data, the flags and the frequencies) is a file-scope
 * device array of 4.8 MB held by every GPU the library runs on until the
 * process ends: nfreq * nsite <= 16384 and nfreq <= 1024, else MT_E_ARG.
 * Every call on one GPU shares that array (which keeps the last call's
 * contents), so calls that overlap on one GPU race; comb_dev is cleared by a
 * kernel, not a memset.  Receivers come from mt_receivers_set as for
 * mt_lmultT (set-up, which allocates). */
int mt_lmultT_noalloc(mt_ctx *ctx, int dtype, int nfreq, const double *omegas, const void *e_dev,
                      const double *d, const unsigned char *exist, void *comb_dev);

/* Tunes, for nsys systems, what the mixed solver's inner iterations run, by
 * timing and keeping the fastest: the systems per workgroup of the brick SpMV
 * (all, 8 or 4) and the row chunks of the vector updates, as a total over the
 * active systems (1024 ... 32768); *variant receives the chosen total.  The
 * choices follow measured times, so they can differ between runs (and with
 * them the reduction order of the dot products).  mt_set_tuning's fixed
 * count per system, when set, takes precedence over the total. */
int mt_autotune(mt_ctx *ctx, int nsys, int *variant);
/* Times one block-DILU application (MT_PC_DILU or MT_PC_DILU_RB) on nsys
 * systems, two per frequency, at MT_PREC_FP64 or MT_PREC_MIXED (FP32), with
 * HIP events on the context stream: average ms per application and its
 * algorithmic bytes (6 coupling slots + per system right-hand side, result
 * and pivots).  Has no counterpart in the reference. */
int mt_bench_precond(mt_ctx *ctx, int nsys, int precision, int precond, int reps,
                     double *ms_per_apply, double *bytes_per_apply);
/* Chooses the SpMV of the solvers: 0 row-parallel (every row gathers its 14
 * neighbours from memory); 2 k-marching (a block keeps three k layers of x of
 * an (i, j) tile in LDS); 1 k-marching in FP32 solves and row-parallel in
 * FP64 ones; 3 (default) rows of several systems per block in FP32 solves
 * (each row's couplings loaded once for the block's systems) and
 * row-parallel in FP64 ones.  layers_per_segment (> 0) sets the k extent of
 * one k-marching block, or for variant 3 the systems per block (4 or 8); 0
 * keeps it.  Has no counterpart in the reference (cuSPARSE's SpMV is fixed
 * there). */
int mt_set_spmv_variant(mt_ctx *ctx, int variant, int layers_per_segment);
/* Launch shape of the red-black DILU (MT_PC_DILU_RB) sweeps: 1 = one system
 * per single-wave persistent workgroup with the next item prefetched into
 * registers; 4 = four systems of one tile per workgroup sharing the tile's
 * couplings in LDS; 5 = the same with the lean LDS layout (three workgroups
 * per CU); 6 (default) = for the inner solves on bricks, persistent
 * four-system workgroups (three per CU) that load their next item while
 * they sweep the current one (elsewhere as 5).  The same operator; 5 and 6
 * round the black tiles' backward sweep differently from 1 and 4 (6 is
 * bitwise 5).  No counterpart in the reference. */
int mt_set_dilu_variant(mt_ctx *ctx, int variant);
/* Layout of the mixed solver's inner FP32 vectors when they are
 * preconditioned by the red-black DILU with the default SpMV: 2 (default) =
 * bricks (each 8x8x8 tile of the red-black plan contiguous) with the
 * preconditioned vectors p-hat and s-hat stored pair-interleaved (systems
 * 2m and 2m+1 side by side, so the SpMV reads both with one 16-byte gather;
 * bitwise equal to 1); 1 = bricks, every vector system-major; 0 = the
 * all-edge layout of the outer solve.  systems_per_block: systems per
 * workgroup of the brick SpMV (0, the default: all active systems).  The same
 * operator and preconditioner; bricks and the all-edge layout differ in the
 * rounding of the inner sums.  No counterpart in the reference. */
int mt_set_inner_layout(mt_ctx *ctx, int bricks, int systems_per_block);
/* The inner solves' brick kernels on one vector, for checking them: y = A x
 * (what = 0) or y = M^-1 x (what = 1, the red-black DILU of A(omega)), of A
 * (transpose = 0) or A^T (1), computed in FP32 on bricks; x, y: n_int
 * complex in EDGEi order.  No counterpart in the reference. */
int mt_inner_apply(mt_ctx *ctx, double omega, int what, int transpose, const mt_zdouble *x,
                   mt_zdouble *y);
/* Fixes the row-tile group count per system (0: automatic). */
int mt_set_tuning(mt_ctx *ctx, int groups_per_system);
/* Times the solver's fused SpMV+dot kernel (v = A x, <rt,v>) for nsys systems
 * at MT_PREC_FP64 (FP64) or MT_PREC_MIXED (FP32) with HIP events on the
 * context stream: average ms per launch and algorithmic bytes per launch
 * (nnz*s_v + n*s_d + nsys*3*n*s_c). */
int mt_bench_spmv(mt_ctx *ctx, int nsys, int precision, int reps, double *ms_per_launch,
                  double *bytes_per_launch);
/* Kernel timing of the BiCGStab engine with HIP events on the solver stream,
 * summed over every solve since the last mt_timing_enable (which resets it;
 * on = 0 turns the events off).  family MT_KT_SPMV: one event pair around each
 * fused SpMV+dot launch; MT_KT_PRECOND: one pair around each preconditioner
 * application (block DILU: one launch per tile level and sweep; red-black
 * tile DILU: three launches), so its time includes the gaps between those
 * launches.  launches: kernel launches covered; ms: summed event time; bytes:
 * algorithmic bytes of those launches (SpMV: stored couplings + diagonal per
 * launch, x, y and the dot partner per system; preconditioner: the six
 * lattice couplings per application, right-hand side(s), result and pivots
 * per system).  While timing is on, solves launch their kernels eagerly
 * (use_graph is ignored): HIP does not time events recorded by graph nodes. */
#define MT_KT_SPMV 0
#define MT_KT_PRECOND 1
int mt_timing_enable(mt_ctx *ctx, int on);
int mt_kernel_timing(const mt_ctx *ctx, int family, long long *launches, double *ms,
                     double *bytes);

#ifdef __cplusplus
}
#endif
#endif
