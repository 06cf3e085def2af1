// Device kernels of libmtsp2 (gfx950, wave64).
//
// Everything here is memory-bound integer-indexed sparse work; there is no
// GEMM and no MFMA.  Design points:
//  * vectors are padded to all edges (boundary entries zero) so stencil
//    addresses are affine and always in bounds (see mt_internal.h);
//  * batched kernels take a list of active systems and map blocks to
//    (system, row tile) XCD-aware: the 8 XCDs work on different row tiles and
//    consecutive blocks on one XCD cover the systems of one tile, so the
//    shared stencil coefficients are fetched once per XCD L2;
//  * dot products: each block reduces its lanes (wave64 shuffles or DPP
//    moves, then LDS) and stores a per-block partial; a finalize kernel
//    with one block per system, next on the stream, sums the partials in a
//    fixed order (no float atomics, deterministic) and updates the solver
//    scalars on the device: no host round trip per dot (the reference reads
//    every dot back to the host, solver.f90:1098-1104);
//  * the inner solves of the mixed solver keep their vectors on bricks
//    (mt_brick.hpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mt {

constexpr double kIsignD = -1.0;  // ISIGN of the reference (math_constants.f90:16)

template <class T>
struct alignas(2 * sizeof(T)) C2 {
  T x, y;
};
using Z = C2<double>;

template <class T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <class T>
__device__ __forceinline__ C2<T> cdotm(C2<T> a, C2<T> b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ C2<T> cscale(T s, C2<T> a) { return {s * a.x, s * a.y}; }
__host__ __device__ __forceinline__ Z zdiv(Z a, Z b) {
  double d = b.x * b.x + b.y * b.y;
  return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__host__ __device__ __forceinline__ Z zmul(Z a, Z b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <class T, class U>
__device__ __forceinline__ C2<T> ccast(C2<U> a) { return {T(a.x), T(a.y)}; }

// ---------------------------------------------------------------- operator
// Stencil operator view.  coef is [15][nE] (slot-major), vms1 is [nE], both
// zero outside interior rows.  off/base select the neighbour addresses.
template <class V>
struct StencilView {
  const V* coef;
  const V* vms1;
  const V* rowc;   // [nE][16]: a row's 15 couplings and vms1, contiguous (may be null)
  int64_t nE;
  int64_t nXi, nYi, nInt;
  int nx, ny, nz;
  int64_t nXe, nYe;
  int off[3][15];  // neighbour offset relative to its base, per component/slot
};

struct RowInfo {
  int64_t e;   // all-edge index of the row
  int64_t b1;  // base (same lattice position) in the first other component
  int64_t b2;  // base in the second other component
  int c;       // component 0/1/2
};

template <class V>
__device__ __forceinline__ RowInfo decode_row(const StencilView<V>& A, int64_t r) {
  RowInfo ri;
  int i, j, k;
  const int nx = A.nx, ny = A.ny;
  if (r < A.nXi) {
    ri.c = 0;
    uint32_t q = uint32_t(r), ai = uint32_t(nx), bi = uint32_t(ny - 1);
    i = q % ai; q /= ai; j = 1 + q % bi; k = 1 + q / bi;
  } else if (r < A.nXi + A.nYi) {
    ri.c = 1;
    uint32_t q = uint32_t(r - A.nXi), ai = uint32_t(nx - 1), bi = uint32_t(ny);
    i = 1 + q % ai; q /= ai; j = q % bi; k = 1 + q / bi;
  } else {
    ri.c = 2;
    uint32_t q = uint32_t(r - A.nXi - A.nYi), ai = uint32_t(nx - 1), bi = uint32_t(ny - 1);
    i = 1 + q % ai; q /= ai; j = 1 + q % bi; k = q / bi;
  }
  const int64_t ex = i + int64_t(nx) * (j + int64_t(ny + 1) * k);
  const int64_t ey = A.nXe + i + int64_t(nx + 1) * (j + int64_t(ny) * k);
  const int64_t ez = A.nXe + A.nYe + i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k);
  if (ri.c == 0) { ri.e = ex; ri.b1 = ey; ri.b2 = ez; }
  else if (ri.c == 1) { ri.e = ey; ri.b1 = ex; ri.b2 = ez; }
  else { ri.e = ez; ri.b1 = ex; ri.b2 = ey; }
  return ri;
}

// Compile-time component tag for the row loops below.
template <int C>
struct Comp {
  static constexpr int value = C;
};

// Row r, counted from the first interior row of component C, -> RowInfo.
template <int C, class V>
__device__ __forceinline__ RowInfo decode_row_c(const StencilView<V>& A, int64_t r) {
  RowInfo ri;
  int i, j, k;
  const int nx = A.nx, ny = A.ny;
  uint32_t q = uint32_t(r);
  if (C == 0) {
    const uint32_t ai = uint32_t(nx), bi = uint32_t(ny - 1);
    i = q % ai; q /= ai; j = 1 + q % bi; k = 1 + q / bi;
  } else if (C == 1) {
    const uint32_t ai = uint32_t(nx - 1), bi = uint32_t(ny);
    i = 1 + q % ai; q /= ai; j = q % bi; k = 1 + q / bi;
  } else {
    const uint32_t ai = uint32_t(nx - 1), bi = uint32_t(ny - 1);
    i = 1 + q % ai; q /= ai; j = 1 + q % bi; k = q / bi;
  }
  const int64_t ex = i + int64_t(nx) * (j + int64_t(ny + 1) * k);
  const int64_t ey = A.nXe + i + int64_t(nx + 1) * (j + int64_t(ny) * k);
  const int64_t ez = A.nXe + A.nYe + i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k);
  ri.c = C;
  if (C == 0) { ri.e = ex; ri.b1 = ey; ri.b2 = ez; }
  else if (C == 1) { ri.e = ey; ri.b1 = ex; ri.b2 = ez; }
  else { ri.e = ez; ri.b1 = ex; ri.b2 = ey; }
  return ri;
}

// body(ri, Comp<C>()) for the interior rows r = s + t0 + n*step of each
// component's part [s, e) of [lo, hi): inside the body the component, and so
// the neighbour offsets, are compile-time (uniform kernel-argument reads).
template <class V, class F>
__device__ __forceinline__ void for_rows(const StencilView<V>& A, int64_t lo, int64_t hi,
                                         int64_t t0, int64_t step, F&& body) {
  const int64_t b1 = A.nXi, b2 = A.nXi + A.nYi;
  for (int64_t r = lo + t0; r < min(hi, b1); r += step) body(decode_row_c<0>(A, r), Comp<0>());
  for (int64_t r = max(lo, b1) + t0; r < min(hi, b2); r += step)
    body(decode_row_c<1>(A, r - b1), Comp<1>());
  for (int64_t r = max(lo, b2) + t0; r < hi; r += step) body(decode_row_c<2>(A, r - b2), Comp<2>());
}

// y_row = sum_s coef[s][e] x[nb_s] for a row of component C
template <int C, class V, class T>
__device__ __forceinline__ C2<T> stencil_row_c(const StencilView<V>& A, const RowInfo& ri,
                                               const C2<T>* __restrict__ x) {
  const V* __restrict__ cf = A.coef + ri.e;
  const int64_t nE = A.nE;
  C2<T> acc = cscale(T(cf[0]), x[ri.e]);
#pragma unroll
  for (int s = 1; s < 7; ++s) {
    const C2<T> xv = x[ri.e + A.off[C][s]];
    const T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
#pragma unroll
  for (int s = 7; s < 11; ++s) {
    const C2<T> xv = x[ri.b1 + A.off[C][s]];
    const T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
#pragma unroll
  for (int s = 11; s < 15; ++s) {
    const C2<T> xv = x[ri.b2 + A.off[C][s]];
    const T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
  return acc;
}

// y_row = sum_s coef[s][e] x[nb_s]  (real stencil part only)
template <class V, class T>
__device__ __forceinline__ C2<T> stencil_row(const StencilView<V>& A, const RowInfo& ri,
                                             const C2<T>* __restrict__ x) {
  const int* o = A.off[ri.c];
  const V* cf = A.coef + ri.e;
  const int64_t nE = A.nE;
  C2<T> acc = cscale(T(cf[0]), x[ri.e]);
#pragma unroll
  for (int s = 1; s < 7; ++s) {
    C2<T> xv = x[ri.e + o[s]];
    T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
#pragma unroll
  for (int s = 7; s < 11; ++s) {
    C2<T> xv = x[ri.b1 + o[s]];
    T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
#pragma unroll
  for (int s = 11; s < 15; ++s) {
    C2<T> xv = x[ri.b2 + o[s]];
    T c = T(cf[s * nE]);
    acc.x += c * xv.x; acc.y += c * xv.y;
  }
  return acc;
}

// ---------------------------------------------------------- reductions
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// Sum of a double over the wave's 64 lanes, valid in lane 63, by DPP lane
// moves inside the VALU (no LDS crossbar round trips): quad [1,0,3,2],
// quad [2,3,0,1], half-row mirror and row mirror leave each 16-lane row's
// sum in all of its lanes; row broadcast 15 adds row 0 into row 1 and row 2
// into row 3, row broadcast 31 adds lane 31 into rows 2 and 3.  A fixed
// order, so the result is deterministic.  Every lane of the wave must be
// active.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, int(b), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(b >> 32), CTRL, ROWS, 0xf, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ __forceinline__ double wave_sum63(double v) {
  v += dpp_d<0xb1, 0xf>(v);    // quad_perm [1,0,3,2]
  v += dpp_d<0x4e, 0xf>(v);    // quad_perm [2,3,0,1]
  v += dpp_d<0x141, 0xf>(v);   // row_half_mirror
  v += dpp_d<0x140, 0xf>(v);   // row_mirror
  v += dpp_d<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v += dpp_d<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
  return v;
}

// Sums NV complex values over the block; result valid in thread 0.
template <int NV, int BS>
__device__ __forceinline__ void block_sum(Z (&v)[NV], Z* lds /* >= NV*BS/64 */) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    v[q].x = wave_sum(v[q].x);
    v[q].y = wave_sum(v[q].y);
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) lds[w * NV + q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      Z s = lds[q];
      for (int ww = 1; ww < BS / 64; ++ww) s = cadd(s, lds[ww * NV + q]);
      v[q] = s;
    }
  }
}

// Per-system solver scalars, FP64 whatever the vector precision.
enum : int { ST_RUN = 0, ST_RECHECK = 1, ST_CONV = 2, ST_BREAK = 3, ST_MAXIT = 4 };

struct SysScal {
  Z rho, rho1, alpha, omg, beta;  // omg: BiCGStab omega
  double bnorm, btol, rnorm, rnorm_best, rnorm_start;
  double freq_omega;              // angular frequency of the system
  double scale;                   // mixed: residual scaling into FP32
  int iter, maxit, status, restart_due, best_due, restart, inner, fidx;
  int tot_inner;                  // mixed (outer record): inner iterations so far
  int pad[3];
};

// Block -> (system, row chunk) mapping, XCD-aware.  nblk = nact * G with G a
// multiple of 8.  Block b is dispatched to XCD b % 8.  Each XCD owns a
// contiguous eighth of the rows (chunks xcd*G/8 .. (xcd+1)*G/8-1) and
// consecutive blocks on one XCD cover all systems of one chunk, so the
// stencil couplings, shared by the systems, are fetched once per XCD.  (The
// alternative, an XCD sweeping whole systems, keeps a system's k-1/k+1 planes
// in its L2 but re-reads the couplings per system: measured slower, with
// twice the fetched bytes.)
__device__ __forceinline__ void map_sys_block(int nact, int G, int& si, int& g) {
  const int b = blockIdx.x;
  const int xcd = b & 7, w = b >> 3;
  si = w % nact;
  g = xcd * (G >> 3) + w / nact;
}

// Contiguous row range [lo, hi) of chunk g out of G over n rows (256-aligned).
__device__ __forceinline__ void chunk_range(int64_t n, int G, int g, int64_t& lo, int64_t& hi) {
  const int64_t per = ((n + G - 1) / G + 255) & ~int64_t(255);
  lo = per * g;
  hi = lo + per < n ? lo + per : n;
}

// Row range [lo, hi) of chunk g of the interior rows (EDGEi order: x rows,
// then y, then z) when the G chunks are dealt to the three components in
// turn: chunk g is part g/3 of component g%3.  An XCD owns G/8 consecutive
// chunks, hence the same fraction -- the same k-slab -- of all three
// components, whose rows read each other's x: the cross-component gathers
// then hit that XCD's L2 instead of the other XCDs' share of x being fetched
// again from HBM.
__device__ __forceinline__ void chunk_range3(int64_t nXi, int64_t nYi, int64_t nInt, int G, int g,
                                             int64_t& lo, int64_t& hi) {
  if (G < 3) {
    chunk_range(nInt, G, g, lo, hi);
    return;
  }
  const int c = g % 3, j = g / 3;
  const int cnt = (G - c + 2) / 3;   // chunks of component c
  const int64_t base = c == 0 ? 0 : (c == 1 ? nXi : nXi + nYi);
  const int64_t n = c == 0 ? nXi : (c == 1 ? nYi : nInt - nXi - nYi);
  const int64_t per = ((n + cnt - 1) / cnt + 63) & ~int64_t(63);
  lo = per * j < n ? per * j : n;
  hi = lo + per < n ? lo + per : n;
  lo += base;
  hi += base;
}

// Cross-block reductions never communicate inside a kernel: each block
// stores its partial sums, and a one-block finalize kernel launched next on
// the stream reduces them in fixed order.  (An in-kernel last-block scheme
// needs a device-scope fence per block, which on gfx950 writes back the
// whole L2 of the XCD each time.)

// Sum over g of partials[(si * G + g) * NV + q] for one system, by the whole
// block in a fixed order (deterministic); result valid in thread 0.
template <int NV, int BS>
__device__ __forceinline__ void reduce_sys(const Z* partials, int si, int G, Z (&v)[NV],
                                           Z* lds /* >= NV*BS/64 */) {
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = {0.0, 0.0};
  const Z* p = partials + (int64_t)si * G * NV;
  for (int g = threadIdx.x; g < G; g += BS) {
#pragma unroll
    for (int q = 0; q < NV; ++q) v[q] = cadd(v[q], p[(int64_t)g * NV + q]);
  }
  block_sum<NV, BS>(v, lds);
}

}  // namespace mt
