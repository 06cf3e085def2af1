// Receiver set-up for the data functionals: per site, the sparse rows that
// interpolate the E field (tri-linear over edges, EMfieldInterp.f90:141-324)
// and the H field from the curl of E (face interpolation of the edge
// circulation, EMfieldInterp.f90:329-647), as rows over all-edge indices.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mt_internal.h"

namespace mt {

// base rows of one site, in this order
enum : int { kSiteEx = 0, kSiteEy, kSiteHx, kSiteHy, kSiteHz, kSiteHrx, kSiteHry, kSiteRows };

struct SiteRows {
  int nsite = 0;
  // row (s*kSiteRows + b) holds entries ptr[row] .. ptr[row+1]-1: all-edge
  // index and real coefficient.  The H rows (b >= kSiteHx) leave out the
  // common factor ISIGN*i/omega of the reference's BfromESetUp.
  std::vector<int> ptr, idx;
  std::vector<double> val;
  // per site 3 channel frames (E, H at the site, H at the reference site),
  // each two 2x2 matrices: the rotation the responses use
  // (rotate_Model2Data, DataFunc.f90:931-957) and the one the linearised
  // functionals use (Lrows, DataFunc.f90:626-656), row-major
  std::vector<double> rot;   // [nsite][3][2][4]
  // transpose of the rows for the deterministic adjoint gather: distinct
  // edges uedge[u], with items gptr[u] .. gptr[u+1]-1 = (row, coefficient)
  std::vector<int> uedge, gptr, grow;
  std::vector<double> gval;
};

// x, r: 3*nsite (location and reference location, the grid frame of the
// reference: x north, y east, z down, origin at the Earth's surface shifted by
// `origin`); az: 7*nsite degrees (Hx, Hy, Hz, Ex, Ey, Hx_ref, Hy_ref) or null
// for the defaults 0/90.  dz holds all layers (air first).  Returns MT_OK or
// MT_E_ARG with `err` set (a site whose stencil leaves the grid).
int build_site_rows(const Grid& g, const double* dx, const double* dy, const double* dz, int nzAir,
                    const double origin[3], int nsite, const double* x, const double* r,
                    const double* az, SiteRows& out, std::string& err);

}  // namespace mt
