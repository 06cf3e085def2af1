// Host-side operator construction for libmtsp2: the staggered-grid SP2 system
// in the library's stencil layout, the CSR -> stencil conversion for operators
// handed over by the reference's Fortran host, and the per-frequency block
// DILU pivots.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

#include "mt_internal.h"

namespace mt {

using zc = std::complex<double>;

struct HostOperator {
  Grid g;
  bool have_model = false;             // true after assemble()
  std::vector<double> coef;            // [15][nE] AAii, interior rows only
  std::vector<double> coefCC;          // [15][nE] its curl-curl part (sigma-free, assemble only)
  std::vector<double> vms1;            // [nE] mu0*sigma_E*V_E on interior edges
  std::vector<double> Vedge;           // [nE] (assemble only)
  std::vector<int> imap;               // [nInt] interior (EDGEi order) -> all-edge index
  std::vector<int> bmap;               // [nBdry] boundary edges (EDGEb order)
  // RHS / post-processing operators (assemble only), CSR over interior rows
  // with all-edge column indices:
  std::vector<int> ccib_row, ccib_col; // CC restricted to (interior, boundary)
  std::vector<double> ccib_val;
  std::vector<int> gd_row, gd_col;     // RHS-form GDii (interior, interior)
  std::vector<double> gd_val;
  // model description kept for the gradient map
  std::vector<double> dx, dy, dz;
  int nzAir = 0, nzEarth = 0;
  std::vector<double> logcond;         // nx*ny*nzEarth, Fortran order
  double airLog = 0.0;
  std::vector<double> Vcell_ghost_edge;  // Cell2Edge(V_C) on all edges
  std::vector<double> sigE;            // [nE] edge conductivity (0 on boundary edges)
  std::vector<double> dualA;           // [nE] DualFaceA
  std::vector<double> edgeL;           // [nE] EdgeL
  std::vector<double> sigN;            // [nN] node conductivity
  std::vector<double> vnode;           // [nN] Vnode
  std::vector<char> nodeEarth;         // [nN] earth class of interior nodes (airNIndex)
  std::vector<double> Vcell_ghost_node;  // [nN] sum of adjacent cell volumes
};

// Slots of the grad-div couplings of a row of component c through the node
// at its first (side 0) or second (side 1) end: slot[c][side][q] for the
// node's six edges q in the order of assemble()'s grad-div loop
// (x(i-1), x(i), y(j-1), y(j), z(k-1), z(k) around node (i, j, k)).
void graddiv_slots(int slot[3][2][6]);

// Gradient of Re(sum_pol lam_pol^T A(sigma) e_pol) w.r.t. the LOGE cell
// parameters, for one frequency: e_all = forward fields on all edges
// (npol x nE), lam = adjoint fields on interior edges (npol x nInt).
// mode 0: the reference's PmultT term only (SolverSens.f90:63-116), i.e. the
//         derivative of the i*ISIGN*omega*mu0*sigma_E*V_E diagonal;
// mode 1: plus the derivative of the grad-div part of AAii (GradDivSetup2,
//         modelOperator3D.f90:916-1008) through sigma_E and sigma_N -- the
//         exact discrete adjoint of the CC-GD system.
// dm (nx*ny*nzEarth, Fortran order) is accumulated into.
void grad_assemble(const HostOperator& op, double omega, int npol, const zc* e_all,
                   const zc* lam, int mode, double* dm);

// Divergence-correction operators on the padded node lattice
// ((nx+1)(ny+1)(nz+1) nodes, boundary entries zero):
//   VDs e   (DivC, modelOperator3D.f90:784-795) = -sum_e G_en DualFaceA_e sigma_e e_e
//   Gai phi (Grad, :797-805)                    = (phi_end - phi_start) / EdgeL on interior edges
//   VDsG = VDs Gai restricted to interior nodes (DivCorSetup, :670-705): a
//   symmetric 7-point node stencil, stored [7][nN] (self, i-1, i+1, j-1, j+1, k-1, k+1).
void divcor_node_stencil(const HostOperator& op, std::vector<double>& coefN,
                         std::vector<double>& wE);
// Dilu_Real (spOpTools.f90:1903-1948) of VDsG on the interior-node lattice:
// returns 1/p on all nodes (zero on boundary nodes).
void divcor_inv_pivots(const HostOperator& op, const std::vector<double>& coefN,
                       std::vector<double>& dinv);

void build_index_maps(HostOperator& op);

// ModelDataInit + UpdateCond (modelOperator3D.f90:69-117, 167-213,
// GradDivSetup2 :916-1008, CurlCurlSetUp :371-405) restated directly as
// stencil coefficients: CC by scattering each face's 4x4 curl-curl block,
// GD by scattering each interior node's 6x6 grad-div block.
Status assemble(HostOperator& op, int nx, int ny, int nzAir, int nzEarth, const double* dx,
                const double* dy, const double* dz, const double* logcond, double airLog);

// Converts an interior CSR (the reference's AAii) to the stencil layout.
Status csr_to_stencil(HostOperator& op, int n, int nnz, const int* rowptr, const int* col,
                      int base, const double* val, const double* vms1);

// Block DILU pivots of A(omega) = AAii + i*ISIGN*omega*diag(vms1), restricted
// to the x-x, y-y and z-z blocks (PC_setup, modelOperator3D.f90:512-611;
// Dilu_Cmplx_AS, spOpTools.f90:1996-2055): returns 1/p on all edges (zero on
// boundary edges).
void dilu_inv_pivots(const HostOperator& op, double omega, std::vector<zc>& dinv);

// Interior CSR export of the stencil (0-based, columns sorted).
void stencil_to_csr(const HostOperator& op, std::vector<int>& row, std::vector<int>& col,
                    std::vector<double>& val);

// Transposed stencil (A^T) coefficients.
void transpose_stencil(const HostOperator& op, std::vector<double>& coefT);

// dEdgeToModelParam (ModelMap.inc:281-372) for a real all-edge vector.
void edge_to_model(const HostOperator& op, const double* e, double* dm);

}  // namespace mt
