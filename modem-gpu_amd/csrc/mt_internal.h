// Internal definitions of libmtsp2 (not installed; the public C ABI is include/mtsp2.h).
//
// Data layout in HBM
// ------------------
// Vectors over edges are kept PADDED to all grid edges in the reference's edge
// order [x-edges ; y-edges ; z-edges], each Fortran-ordered (i fastest), with
// the boundary entries held at zero.  The interior-edge vectors the reference
// passes (EDGEi order) are scattered into / gathered out of this layout at the
// boundary of the library.  Padding makes every stencil neighbour of an
// interior edge an in-bounds address, so the operator kernels need no
// predicates: a missing neighbour has a zero coefficient and a boundary
// neighbour has a zero vector entry.
//
// The operator AAii (real, shared by every frequency) is stored as a
// 15-coefficient stencil, structure-of-arrays: coef[slot * nE + edge].  Slot
// order, for an edge of component c at edge-lattice position (i,j,k):
//   0 self; 1..6 same component at a-1, a+1, b-1, b+1, c-1, c+1 where (a,b,c)
//   are the (i,j,k) axes; 7..10 first other component; 11..14 second other
//   component (x rows: y then z; y rows: x then z; z rows: x then y), each in
//   the order listed in stencil_neighbours() below.
// Only interior-interior couplings are stored (exactly AAii's entries); every
// other slot is zero.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_complex.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mtsp2.h"

namespace mt {

constexpr int kSlots = 15;
constexpr double kPi = 3.14159265357898;  // the reference's constant (math_constants.f90:12)
constexpr double kMu0 = kPi * 4e-7;
// SIGMA_AIR (math_constants.f90:19) is a default-real literal widened to real(8)
constexpr double kSigmaAir = double(1.0e-10f);
constexpr int kIsign = -1;

struct Grid {
  int nx = 0, ny = 0, nz = 0;  // cells (nz includes air layers)
  int64_t nXe = 0, nYe = 0, nZe = 0, nE = 0;
  int64_t nXi = 0, nYi = 0, nZi = 0, nInt = 0;
  void set(int x, int y, int z) {
    nx = x; ny = y; nz = z;
    nXe = int64_t(nx) * (ny + 1) * (nz + 1);
    nYe = int64_t(nx + 1) * ny * (nz + 1);
    nZe = int64_t(nx + 1) * (ny + 1) * nz;
    nE = nXe + nYe + nZe;
    nXi = int64_t(nx) * (ny - 1) * (nz - 1);
    nYi = int64_t(nx - 1) * ny * (nz - 1);
    nZi = int64_t(nx - 1) * (ny - 1) * nz;
    nInt = nXi + nYi + nZi;
  }
  // Zero tail after the last system's vector: the z-edge stencil of the top
  // layer reaches one z-plane past the end of the z block (its coefficient is
  // zero), so every padded vector buffer is allocated with this many extra
  // entries to keep that read in bounds.
  int64_t tail() const { return int64_t(nx + 1) * (ny + 1) + 256; }
  int64_t vlen(int nsys) const { return int64_t(nsys) * nE + tail(); }
  // all-edge index of component c at edge-lattice (i,j,k)
  int64_t eidx(int c, int i, int j, int k) const {
    if (c == 0) return i + int64_t(nx) * (j + int64_t(ny + 1) * k);
    if (c == 1) return nXe + i + int64_t(nx + 1) * (j + int64_t(ny) * k);
    return nXe + nYe + i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k);
  }
  void dims(int c, int& a, int& b, int& d) const {
    if (c == 0) { a = nx; b = ny + 1; d = nz + 1; }
    else if (c == 1) { a = nx + 1; b = ny; d = nz + 1; }
    else { a = nx + 1; b = ny + 1; d = nz; }
  }
  bool interior(int c, int i, int j, int k) const {
    if (c == 0) return j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 1 && i >= 0 && i < nx;
    if (c == 1) return i >= 1 && i <= nx - 1 && k >= 1 && k <= nz - 1 && j >= 0 && j < ny;
    return i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1 && k >= 0 && k < nz;
  }
};

// Relative position (component, di, dj, dk) of each stencil slot for a row of
// component c.  Used by the host (CSR conversion, assembly) and mirrored in the
// kernels' address arithmetic.
struct Nb { int comp, di, dj, dk; };
inline void stencil_neighbours(int c, Nb out[kSlots]) {
  out[0] = {c, 0, 0, 0};
  out[1] = {c, -1, 0, 0}; out[2] = {c, 1, 0, 0};
  out[3] = {c, 0, -1, 0}; out[4] = {c, 0, 1, 0};
  out[5] = {c, 0, 0, -1}; out[6] = {c, 0, 0, 1};
  if (c == 0) {
    out[7] = {1, 0, -1, 0}; out[8] = {1, 1, -1, 0}; out[9] = {1, 0, 0, 0}; out[10] = {1, 1, 0, 0};
    out[11] = {2, 0, 0, -1}; out[12] = {2, 1, 0, -1}; out[13] = {2, 0, 0, 0}; out[14] = {2, 1, 0, 0};
  } else if (c == 1) {
    out[7] = {0, -1, 0, 0}; out[8] = {0, 0, 0, 0}; out[9] = {0, -1, 1, 0}; out[10] = {0, 0, 1, 0};
    out[11] = {2, 0, 0, -1}; out[12] = {2, 0, 1, -1}; out[13] = {2, 0, 0, 0}; out[14] = {2, 0, 1, 0};
  } else {
    out[7] = {0, -1, 0, 0}; out[8] = {0, 0, 0, 0}; out[9] = {0, -1, 0, 1}; out[10] = {0, 0, 0, 1};
    out[11] = {1, 0, -1, 0}; out[12] = {1, 0, 0, 0}; out[13] = {1, 0, -1, 1}; out[14] = {1, 0, 0, 1};
  }
}
// slot of neighbour (cb, di, dj, dk) for a row of component ca, or -1
inline int slot_of(int ca, int cb, int di, int dj, int dk) {
  Nb nb[kSlots];
  stencil_neighbours(ca, nb);
  for (int s = 0; s < kSlots; ++s)
    if (nb[s].comp == cb && nb[s].di == di && nb[s].dj == dj && nb[s].dk == dk) return s;
  return -1;
}

struct Status {
  int code = MT_OK;
  std::string msg;
};

// Per-device-buffer RAII-free helpers: the context owns everything and frees
// it in mt_ctx_destroy.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace mt

#define MT_HIP_CHECK(ctx, call)                                                  \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      (void)hipGetLastError(); /* clear it: the next call must not report it */  \
      return mt_fail((ctx), MT_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    }                                                                            \
  } while (0)
