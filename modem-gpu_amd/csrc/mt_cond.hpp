// UpdateCond on the device (modelOperator3D.f90:167-213): a new conductivity
// model on the grid of the last assembly.  sigma_E and sigma_N are the
// volume-weighted averages of the cells around each edge / node
// (ModelParamToEdge / ModelParamToNode, ModelMap.inc:88-277), sigma_E is zero
// on boundary edges, and AAii = CCii + GDii(sigma_E, sigma_N) (GradDivSetup2
// with airNIndex, modelOperator3D.f90:916-1130) with vms1 = mu0 sigma_E V_E.
// The curl-curl part does not depend on sigma and is kept from the assembly.
//
// Every sum is formed in the order of the host assembly (cells in (k, j, i)
// order, grad-div nodes in (k, j, i) order, a node's edges in their listed
// order) with contraction off, so the device operator equals the host one
// bit for bit.
#pragma once

#include "mt_device.hpp"
#include "mt_internal.h"

namespace mt {

struct CondArgs {
  int nx, ny, nz;
  int nzAir;                  // air layers (the sensitivity kernels' earth cells start below)
  int64_t nXe, nYe, nE, nInt;
  const double *dx, *dy, *dz, *delX, *delY, *delZ;
  const double* csig;         // [nz][ny][nx] cell conductivity (air layers included)
  const int* imap;            // interior row -> all-edge index
  const double* coefCC;       // [15][nE] curl-curl part
  const int* slot;            // [3][2][6] grad-div slots (graddiv_slots)
  double* sigE;               // [nE]
  unsigned char* earthE;      // [nE] sigma_E counts as earth
  double* sigN;               // [(nx+1)(ny+1)(nz+1)]
  double* m2;                 // [nodes] 1/(sigma_N V_node) (earth) or 1/V_node (air)
  double* coef;               // [15][nE] AAii out
  double* vms1;               // [nE] out
};

__device__ __forceinline__ void cond_edge_pos(const CondArgs& a, int64_t e, int& c, int& i, int& j,
                                              int& k) {
  int64_t q;
  int A, B;
  if (e < a.nXe) { c = 0; q = e; A = a.nx; B = a.ny + 1; }
  else if (e < a.nXe + a.nYe) { c = 1; q = e - a.nXe; A = a.nx + 1; B = a.ny; }
  else { c = 2; q = e - a.nXe - a.nYe; A = a.nx + 1; B = a.ny + 1; }
  i = int(q % A);
  q /= A;
  j = int(q % B);
  k = int(q / B);
}

__device__ __forceinline__ int64_t cond_eidx(const CondArgs& a, int c, int i, int j, int k) {
  if (c == 0) return i + int64_t(a.nx) * (j + int64_t(a.ny + 1) * k);
  if (c == 1) return a.nXe + i + int64_t(a.nx + 1) * (j + int64_t(a.ny) * k);
  return a.nXe + a.nYe + i + int64_t(a.nx + 1) * (j + int64_t(a.ny + 1) * k);
}

__device__ __forceinline__ bool cond_interior(const CondArgs& a, int c, int i, int j, int k) {
  if (c == 0) return j >= 1 && j <= a.ny - 1 && k >= 1 && k <= a.nz - 1;
  if (c == 1) return i >= 1 && i <= a.nx - 1 && k >= 1 && k <= a.nz - 1;
  return i >= 1 && i <= a.nx - 1 && j >= 1 && j <= a.ny - 1;
}

// sigma_E, earth class, vms1 of every edge
__global__ void k_cond_edges(CondArgs a) {
#pragma clang fp contract(off)
  const double thr = double(1.1f) * kSigmaAir;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < a.nE;
       e += (int64_t)gridDim.x * blockDim.x) {
    int c, i, j, k;
    cond_edge_pos(a, e, c, i, j, k);
    double num = 0.0, den = 0.0;
    // the cells around the edge, in the host's (k, j, i) cell order
    for (int k2 = (c == 2 ? k : k - 1); k2 <= k; ++k2)
      for (int j2 = (c == 1 ? j : j - 1); j2 <= j; ++j2)
        for (int i2 = (c == 0 ? i : i - 1); i2 <= i; ++i2) {
          if (i2 < 0 || j2 < 0 || k2 < 0 || i2 >= a.nx || j2 >= a.ny || k2 >= a.nz) continue;
          const double vc = a.dx[i2] * a.dy[j2] * a.dz[k2];
          const double vs = vc * a.csig[i2 + int64_t(a.nx) * (j2 + int64_t(a.ny) * k2)];
          num += vs;
          den += vc;
        }
    const bool in = cond_interior(a, c, i, j, k);
    const double s = in ? num / den : 0.0;
    a.sigE[e] = s;
    a.earthE[e] = (s != 0.0) && !(s < thr);
    const double len = c == 0 ? a.dx[i] : (c == 1 ? a.dy[j] : a.dz[k]);
    const double dual = c == 0 ? a.delY[j] * a.delZ[k] : (c == 1 ? a.delX[i] * a.delZ[k] : a.delX[i] * a.delY[j]);
    a.vms1[e] = in ? kMu0 * s * (len * dual) : 0.0;
  }
}

// sigma_N and the grad-div node weight of every node
__global__ void k_cond_nodes(CondArgs a) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nN;
       n += (int64_t)gridDim.x * blockDim.x) {
    int64_t q = n;
    const int i = int(q % (a.nx + 1));
    q /= (a.nx + 1);
    const int j = int(q % (a.ny + 1));
    const int k = int(q / (a.ny + 1));
    double num = 0.0, den = 0.0;
    for (int k2 = k - 1; k2 <= k; ++k2)
      for (int j2 = j - 1; j2 <= j; ++j2)
        for (int i2 = i - 1; i2 <= i; ++i2) {
          if (i2 < 0 || j2 < 0 || k2 < 0 || i2 >= a.nx || j2 >= a.ny || k2 >= a.nz) continue;
          const double vc = a.dx[i2] * a.dy[j2] * a.dz[k2];
          const double vs = vc * a.csig[i2 + int64_t(a.nx) * (j2 + int64_t(a.ny) * k2)];
          num += vs;
          den += vc;
        }
    const double sn = num / den;
    a.sigN[n] = sn;
    const double vnode = a.delX[i] * a.delY[j] * a.delZ[k];
    double w = 0.0;
    if (i > 0 && i < a.nx && j > 0 && j < a.ny && k > 0 && k < a.nz) {
      const bool earth = a.earthE[cond_eidx(a, 0, i - 1, j, k)] || a.earthE[cond_eidx(a, 0, i, j, k)] ||
                         a.earthE[cond_eidx(a, 1, i, j - 1, k)] || a.earthE[cond_eidx(a, 1, i, j, k)] ||
                         a.earthE[cond_eidx(a, 2, i, j, k - 1)] || a.earthE[cond_eidx(a, 2, i, j, k)];
      w = earth ? 1.0 / sn / vnode : 1.0 / vnode;
      if (!earth) w = -w;   // sign marks an air node (the weight itself is positive)
    }
    a.m2[n] = w;
  }
}

// AAii rows: the curl-curl part plus the grad-div terms of the row's two end
// nodes, first node then second, the node's six edges in order
__global__ void k_cond_rows(CondArgs a) {
#pragma clang fp contract(off)
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < a.nInt;
       r += (int64_t)gridDim.x * blockDim.x) {
    const int64_t e = a.imap[r];
    int c, i, j, k;
    cond_edge_pos(a, e, c, i, j, k);
    double cf[kSlots];
#pragma unroll
    for (int s = 0; s < kSlots; ++s) cf[s] = a.coefCC[(int64_t)s * a.nE + e];
    const double dualE = c == 0 ? a.delY[j] * a.delZ[k] : (c == 1 ? a.delX[i] * a.delZ[k] : a.delX[i] * a.delY[j]);
    for (int side = 0; side < 2; ++side) {
      const int ni = i + (c == 0 ? side : 0), nj = j + (c == 1 ? side : 0), nk = k + (c == 2 ? side : 0);
      if (!(ni > 0 && ni < a.nx && nj > 0 && nj < a.ny && nk > 0 && nk < a.nz)) continue;
      const double wn = a.m2[ni + int64_t(a.nx + 1) * (nj + int64_t(a.ny + 1) * nk)];
      const bool earth = wn > 0.0;
      const double m2 = earth ? wn : -wn;
      const double se = side == 0 ? -1.0 : 1.0;   // G = -1 at the first node, +1 at the second
      const int qc[6] = {0, 0, 1, 1, 2, 2};
      const int qi[6] = {ni - 1, ni, ni, ni, ni, ni}, qj[6] = {nj, nj, nj - 1, nj, nj, nj};
      const int qk[6] = {nk, nk, nk, nk, nk - 1, nk};
#pragma unroll
      for (int q = 0; q < 6; ++q) {
        if (!cond_interior(a, qc[q], qi[q], qj[q], qk[q])) continue;
        const int64_t eq = cond_eidx(a, qc[q], qi[q], qj[q], qk[q]);
        const double sq = (q & 1) ? -1.0 : 1.0;
        const double dualQ = qc[q] == 0 ? a.delY[qj[q]] * a.delZ[qk[q]]
                                        : (qc[q] == 1 ? a.delX[qi[q]] * a.delZ[qk[q]] : a.delX[qi[q]] * a.delY[qj[q]]);
        const double colw = earth ? a.sigE[eq] * dualQ : dualQ;
        const double v = dualE * se * m2 * sq * colw;
        const int s = a.slot[(c * 2 + side) * 6 + q];
        cf[s] = cf[s] + v;
      }
    }
#pragma unroll
    for (int s = 0; s < kSlots; ++s) a.coef[(int64_t)s * a.nE + e] = cf[s];
  }
}

template <class U>
__global__ void k_cast_array(const double* __restrict__ in, U* __restrict__ out, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = U(in[q]);
}

}  // namespace mt
