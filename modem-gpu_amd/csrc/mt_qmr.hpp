// Vector kernels of the QMR solver (the reference's CPU default,
// solver.f90:140-370) on padded all-edge FP64 vectors.  The scalars of the
// recurrence are formed on the host from deterministic per-block partial
// sums; every vector operation here is elementwise over [0, nE), boundary
// entries stay zero.
#pragma once

#include "mt_device.hpp"
#include "mt_internal.h"

namespace mt {

constexpr int QMR_BS = 256;

__device__ __forceinline__ Z zscale(Z a, Z s) { return zmul(a, s); }

// Per-block partial sums of up to three inner products sum conj(a_k) b_k
// (Fortran dot_product) over [0, n); partial k of block g at part[k*G + g].
__global__ __launch_bounds__(QMR_BS) void k_q_dots(int64_t n, int nd, const Z* __restrict__ a0,
                                                   const Z* __restrict__ b0, const Z* __restrict__ a1,
                                                   const Z* __restrict__ b1, const Z* __restrict__ a2,
                                                   const Z* __restrict__ b2, Z* __restrict__ part) {
  __shared__ Z red[3][QMR_BS];
  Z acc[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  const Z* A[3] = {a0, a1, a2};
  const Z* B[3] = {b0, b1, b2};
  for (int64_t q = (int64_t)blockIdx.x * QMR_BS + threadIdx.x; q < n; q += (int64_t)gridDim.x * QMR_BS) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k >= nd) break;
      const Z x = A[k][q], y = B[k][q];
      acc[k].x += x.x * y.x + x.y * y.y;
      acc[k].y += x.x * y.y - x.y * y.x;
    }
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int s = QMR_BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x].x += red[k][threadIdx.x + s].x;
        red[k][threadIdx.x].y += red[k][threadIdx.x + s].y;
      }
    __syncthreads();
  }
  if (threadIdx.x < nd) part[threadIdx.x * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}

// V = VT s1, W = WT s2, Y = Y s1, Z = Z s2 (the normalisation at the top of
// an iteration)
__global__ void k_q_scale(int64_t n, Z s1, Z s2, const Z* __restrict__ vt, const Z* __restrict__ wt,
                          Z* __restrict__ v, Z* __restrict__ w, Z* __restrict__ y, Z* __restrict__ z) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    v[q] = zmul(vt[q], s1);
    w[q] = zmul(wt[q], s2);
    y[q] = zmul(y[q], s1);
    z[q] = zmul(z[q], s2);
  }
}

// out = in (mode 0), out = in / d (mode 1: the diagonal Q = conj(P) applied
// as a division by its stored inverse), out = in * d (mode 2); d = 0 (a
// boundary entry) gives 0
__global__ void k_q_diag(int64_t n, int mode, const Z* __restrict__ in, const Z* __restrict__ d,
                         Z* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const Z x = in[q];
    if (mode == 0) { out[q] = x; continue; }
    const Z dd = d[q];
    if (dd.x == 0.0 && dd.y == 0.0) { out[q] = {0.0, 0.0}; continue; }
    out[q] = mode == 1 ? zdiv(x, dd) : zmul(x, dd);
  }
}

// out = a - s * b
__global__ void k_q_axpy(int64_t n, Z s, const Z* __restrict__ a, const Z* __restrict__ b, Z* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    const Z t = zmul(s, b[q]);
    out[q] = {a[q].x - t.x, a[q].y - t.y};
  }
}

// D = eta P + tm2 D, S = eta PT + tm2 S (first: D = eta P, S = eta PT),
// x += D, R -= S
__global__ void k_q_update(int64_t n, int first, Z eta, Z tm2, const Z* __restrict__ p,
                           const Z* __restrict__ pt, Z* __restrict__ d, Z* __restrict__ s,
                           Z* __restrict__ x, Z* __restrict__ r) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    Z dq = zmul(eta, p[q]), sq = zmul(eta, pt[q]);
    if (!first) {
      const Z a = zmul(tm2, d[q]), b = zmul(tm2, s[q]);
      dq = {dq.x + a.x, dq.y + a.y};
      sq = {sq.x + b.x, sq.y + b.y};
    }
    d[q] = dq;
    s[q] = sq;
    x[q] = {x[q].x + dq.x, x[q].y + dq.y};
    r[q] = {r[q].x - sq.x, r[q].y - sq.y};
  }
}

// conjugated copy (the pivots of A^H from those of A)
__global__ void k_q_conj(int64_t n, const Z* __restrict__ in, Z* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = {in[q].x, -in[q].y};
}

}  // namespace mt
