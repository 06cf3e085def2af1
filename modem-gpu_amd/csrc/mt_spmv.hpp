// k-marching SpMV for the BiCGStab engine: y = A x for one system per block
// over an (i, j) tile of lattice positions, marching up through the k layers.
//
// Every row of the staggered-grid operator at lattice position (i, j, k)
// reads x only at positions (i+-1, j+-1, k+-1) of the three edge components,
// so a block keeps three k layers of x (all three components, the tile plus a
// one-position halo) in LDS and loads one new layer per step.  Each x value
// is then fetched about (TI+2)(TJ+2)/(TI*TJ) times per segment instead of the
// ~3 times a row-parallel SpMV costs when 16 systems' k-1 / k+1 planes do not
// fit an XCD's L2, and the 14 neighbour gathers per row become LDS reads.
// The next layer's x values are in flight while the current layer is
// computed.
//
// Blocks: nact systems x Gkm work items (tile x k segment), Gkm padded to a
// multiple of 8; consecutive blocks of one XCD take the same work item for
// all systems (shared couplings in L2), the XCDs own contiguous item ranges.
#pragma once

#include <type_traits>

#include "mt_kernels.hpp"

namespace mt {

constexpr int KM_TI = 32, KM_TJ = 8, KM_BS = KM_TI * KM_TJ;
constexpr int KM_PI = KM_TI + 2, KM_PJ = KM_TJ + 2, KM_PL = KM_PI * KM_PJ;  // padded layer

struct KmGeom {
  int nti, ntj, nseg, ks;  // tiles in i, j; k segments of ks layers
  int Greal, Gpad;         // work items per system, padded to a multiple of 8
};

// Neighbour (component, di, dj, dk) of slot s of a row of component c; the
// same table as stencil_neighbours() on the host.
__host__ __device__ constexpr int km_nb(int c, int s, int q) {
  // rows: c*15 + s; columns: comp, di, dj, dk
  constexpr int t[3][15][4] = {
      {{0, 0, 0, 0}, {0, -1, 0, 0}, {0, 1, 0, 0}, {0, 0, -1, 0}, {0, 0, 1, 0}, {0, 0, 0, -1},
       {0, 0, 0, 1}, {1, 0, -1, 0}, {1, 1, -1, 0}, {1, 0, 0, 0}, {1, 1, 0, 0}, {2, 0, 0, -1},
       {2, 1, 0, -1}, {2, 0, 0, 0}, {2, 1, 0, 0}},
      {{1, 0, 0, 0}, {1, -1, 0, 0}, {1, 1, 0, 0}, {1, 0, -1, 0}, {1, 0, 1, 0}, {1, 0, 0, -1},
       {1, 0, 0, 1}, {0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 1, 0}, {0, 0, 1, 0}, {2, 0, 0, -1},
       {2, 0, 1, -1}, {2, 0, 0, 0}, {2, 0, 1, 0}},
      {{2, 0, 0, 0}, {2, -1, 0, 0}, {2, 1, 0, 0}, {2, 0, -1, 0}, {2, 0, 1, 0}, {2, 0, 0, -1},
       {2, 0, 0, 1}, {0, -1, 0, 0}, {0, 0, 0, 0}, {0, -1, 0, 1}, {0, 0, 0, 1}, {1, 0, -1, 0},
       {1, 0, 0, 0}, {1, 0, -1, 1}, {1, 0, 0, 1}}};
  return t[c][s][q];
}

// all-edge address of component c at lattice position (i, j, k), or -1 when
// the position is outside that component's lattice
template <class V>
__device__ __forceinline__ int64_t km_addr(const StencilView<V>& A, int c, int i, int j, int k) {
  const int nx = A.nx, ny = A.ny, nz = A.nz;
  if (i < 0 || j < 0 || k < 0) return -1;
  if (c == 0) {
    if (i >= nx || j > ny || k > nz) return -1;
    return i + int64_t(nx) * (j + int64_t(ny + 1) * k);
  }
  if (c == 1) {
    if (i > nx || j >= ny || k > nz) return -1;
    return A.nXe + i + int64_t(nx + 1) * (j + int64_t(ny) * k);
  }
  if (i > nx || j > ny || k >= nz) return -1;
  return A.nXe + A.nYe + i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k);
}

// interior row of component c at (i, j, k)?
__device__ __forceinline__ bool km_interior(int c, int i, int j, int k, int nx, int ny, int nz) {
  if (c == 0) return i < nx && j >= 1 && j <= ny - 1 && k >= 1 && k <= nz - 1;
  if (c == 1) return i >= 1 && i <= nx - 1 && j < ny && k >= 1 && k <= nz - 1;
  return i >= 1 && i <= nx - 1 && j >= 1 && j <= ny - 1 && k < nz;
}

// Row-major copy of the operator for the k-marching SpMV: row e holds its 15
// couplings and vms1 in 16 consecutive values (4 x 16 B loads in FP32).
template <class V>
__global__ void k_rows_from_slots(const V* __restrict__ coef, const V* __restrict__ vms1,
                                  int64_t nE, V* __restrict__ rows) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nE;
       e += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int s = 0; s < 15; ++s) rows[e * 16 + s] = coef[s * nE + e];
    rows[e * 16 + 15] = vms1[e];
  }
}

// WHICH=1: v = A ph, partials <rt,v>.  WHICH=2: t = A sh, partials <t,t>, <t,s>
// (s = r - alpha v when fused).
template <class T, class V, int WHICH>
__global__ __launch_bounds__(KM_BS) void k_spmv_km(IterArgs<T, V> a, KmGeom gm) {
  __shared__ C2<T> X[3][3][KM_PL];  // [component][layer slot][padded position]
  __shared__ Z lds[2 * KM_BS / 64];
  const int b = blockIdx.x;
  const int xcd = b & 7, w = b >> 3;
  const int si = w % a.nact;
  const int g = xcd * (gm.Gpad >> 3) + w / a.nact;
  const int sys = a.act[si];
  const SysScal& sc0 = a.sc[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  const StencilView<V>& A = a.A;
  if (g < gm.Greal && sc0.status == ST_RUN) {
    const int seg = g % gm.nseg, tile = g / gm.nseg;
    const int i0 = (tile % gm.nti) * KM_TI, j0 = (tile / gm.nti) * KM_TJ;
    const int k0 = seg * gm.ks, k1 = min(k0 + gm.ks, A.nz + 1);
    const int64_t o = (int64_t)sys * a.nE;
    const T dsc = T(kIsignD * sc0.freq_omega);
    const C2<T>* __restrict__ in = (WHICH == 1 ? a.ph : a.sh) + o;
    C2<T>* __restrict__ out = (WHICH == 1 ? a.v : a.t) + o;
    const bool fz = WHICH == 2 && a.fused;
    const C2<T>* __restrict__ pw = (WHICH == 1 ? a.rt : (fz ? a.r : a.s)) + o;
    const C2<T>* __restrict__ pv = a.v + o;
    const C2<T> al = ccast<T>(sc0.alpha);
    const int t = threadIdx.x;
    const int ti = t % KM_TI, tj = t / KM_TI;
    const int i = i0 + ti, j = j0 + tj;
    const int nx = A.nx, ny = A.ny, nz = A.nz;
    // lattice plane sizes (k strides) and depths of the three components
    const int64_t pls[3] = {int64_t(nx) * (ny + 1), int64_t(nx + 1) * ny,
                            int64_t(nx + 1) * (ny + 1)};
    const int dep[3] = {nz + 1, nz + 1, nz};
    // this thread's share of a padded layer (all components): address at
    // k = 0, component, and whether (i, j) lies in that component's lattice
    constexpr int NL = (3 * KM_PL + KM_BS - 1) / KM_BS;
    int64_t lbase[NL];
    int lcomp[NL];
    bool lok[NL];
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int q = t + KM_BS * u;
      const int c = q / KM_PL, pq = q - c * KM_PL;
      const int ii = i0 - 1 + pq % KM_PI, jj = j0 - 1 + pq / KM_PI;
      const int64_t e = km_addr(A, c < 3 ? c : 0, ii, jj, 0);
      lok[u] = q < 3 * KM_PL && e >= 0;
      lbase[u] = lok[u] ? e : 0;
      lcomp[u] = c < 3 ? c : 0;
    }
    // this thread's three rows: address at k = 0 and (i, j) interior-ness
    int64_t rbase[3];
    bool rok[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rok[c] = km_interior(c, i, j, 1, nx, ny, nz);  // the (i, j) part (k = 1 is interior)
      const int64_t e = km_addr(A, c, i, j, 0);
      rbase[c] = e >= 0 ? e : 0;
    }
    auto row_at = [&](int c, int k) {
      const bool kin = (c == 2) ? (k >= 0 && k < nz) : (k >= 1 && k <= nz - 1);
      return rok[c] && kin;
    };
    using V4 = typename std::conditional<sizeof(V) == 4, float4, double2>::type;
    constexpr int NV4 = 16 * sizeof(V) / 16;
    struct Buf {
      C2<T> x[NL];      // the next layer of x
    };
    struct Rows {
      V4 cv[3][NV4];    // a row's couplings 0..14 and vms1
      C2<T> w[3], v[3];
    };
    // loads of layer kx of x.  Every load is issued unconditionally (from a
    // valid address, the value masked afterwards): a fixed number of loads
    // in flight lets the compiler wait for exactly the older ones instead of
    // draining the memory counter.
    auto issue = [&](int kx, Buf& bf) {
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int c = lcomp[u];
        const bool ok = lok[u] && kx >= 0 && kx < dep[c];
        const C2<T> xv = in[ok ? lbase[u] + kx * pls[c] : 0];
        bf.x[u] = ok ? xv : C2<T>{T(0), T(0)};
      }
    };
    // the rows of layer kr (issued at the start of the layer's step)
    auto load_rows = [&](int kr, Rows& rw) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const bool ok = row_at(c, kr);
        const int64_t e = ok ? rbase[c] + kr * pls[c] : 0;
        const V4* __restrict__ rp = reinterpret_cast<const V4*>(A.rowc + e * 16);
#pragma unroll
        for (int q = 0; q < NV4; ++q) rw.cv[c][q] = rp[q];
        rw.w[c] = pw[e];
        rw.v[c] = WHICH == 2 ? pv[e] : C2<T>{T(0), T(0)};
      }
    };
    auto put_layer = [&](int kx, const Buf& bf) {
      const int sl = (kx + 3) % 3;
#pragma unroll
      for (int u = 0; u < NL; ++u) {
        const int q = t + KM_BS * u;
        if (q < 3 * KM_PL) X[lcomp[u]][sl][q - lcomp[u] * KM_PL] = bf.x[u];
      }
    };
    auto compute = [&](int k, const Rows& bf) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (!row_at(c, k)) continue;
        const int64_t e = rbase[c] + k * pls[c];
        const V* cvs = reinterpret_cast<const V*>(&bf.cv[c][0]);
        const T d = dsc * T(cvs[15]);
        C2<T> wv = bf.w[c];
        if (fz) wv = csub(wv, cmul(al, bf.v[c]));
        // the row SpMV's order: slots 0..14, then the diagonal term
        C2<T> y = {T(0), T(0)}, xs = {T(0), T(0)};
#pragma unroll
        for (int sl = 0; sl < 15; ++sl) {
          const int cb = km_nb(c, sl, 0), di = km_nb(c, sl, 1), dj = km_nb(c, sl, 2),
                    dk = km_nb(c, sl, 3);
          const C2<T> xv = X[cb][(k + dk + 3) % 3][(ti + 1 + di) + KM_PI * (tj + 1 + dj)];
          const T cs = T(cvs[sl]);
          if (sl == 0) {
            xs = xv;
            y = cscale(cs, xv);
          } else {
            y.x += cs * xv.x;
            y.y += cs * xv.y;
          }
        }
        y.x += -d * xs.y;  // i*ISIGN*omega*vms1*x
        y.y += d * xs.x;
        out[e] = y;
        if (WHICH == 1) {
          acc[0].x += double(wv.x) * y.x + double(wv.y) * y.y;
          acc[0].y += double(wv.x) * y.y - double(wv.y) * y.x;
        } else {
          acc[0].x += double(y.x) * y.x + double(y.y) * y.y;
          acc[1].x += double(y.x) * wv.x + double(y.y) * wv.y;
          acc[1].y += double(y.x) * wv.y - double(y.y) * wv.x;
        }
        // one component's 15 LDS reads in flight at a time (register pressure)
        __builtin_amdgcn_sched_barrier(0);
      }
    };
    // prologue: layers k0-1 and k0 in LDS, layer k0+1 in buffer A.  Step k:
    // put layer k+1 into LDS; issue layer k+2 into the other buffer and the
    // rows of layer k; compute layer k.
    Buf bA, bB;
    Rows rw;
    issue(k0 - 1, bA);
    put_layer(k0 - 1, bA);
    issue(k0, bA);
    put_layer(k0, bA);
    issue(k0 + 1, bA);
    for (int k = k0; k < k1; k += 2) {
      put_layer(k + 1, bA);
      __syncthreads();
      issue(k + 2, bB);
      load_rows(k, rw);
      compute(k, rw);
      __syncthreads();
      if (k + 1 >= k1) break;
      put_layer(k + 2, bB);
      __syncthreads();
      issue(k + 3, bA);
      load_rows(k + 1, rw);
      compute(k + 1, rw);
      __syncthreads();
    }
  }
  block_sum<2, KM_BS>(acc, lds);
  if (threadIdx.x == 0) {
    a.partials[((int64_t)si * gm.Gpad + g) * 2 + 0] = acc[0];
    a.partials[((int64_t)si * gm.Gpad + g) * 2 + 1] = acc[1];
  }
}

// ------------------------------------------------ several systems per row
// y_s = A x_s for S systems per block over a chunk of rows: each row's 15
// couplings and vms1 are loaded once (the row-major copy, 16 consecutive
// values) and applied to the S systems' x.  With 16 systems the couplings are
// 130 MB of the 920 MB an FP32 SpMV launch must move; loaded per system they
// are read 16 times.  Blocks: (chunk, group of S active systems), groups
// fastest inside an XCD; chunks dealt to the three components in turn
// (chunk_range3) so an XCD's cross-component gathers stay in its L2.  Partial
// dots per (system, chunk) as k_spmv_dot stores them.
template <class T, class V, int WHICH, int S>
__global__ __launch_bounds__(BS) void k_spmv_ms(IterArgs<T, V> a) {
  __shared__ Z lds[S][2 * BS / 64];
  const int ngrp = (a.nact + S - 1) / S;
  const int b = blockIdx.x, xcd = b & 7, w = b >> 3;
  const int grp = w % ngrp, g = xcd * (a.G >> 3) + w / ngrp;
  const StencilView<V>& A = a.A;
  const C2<T>* __restrict__ in[S];
  C2<T>* __restrict__ out[S];
  const C2<T>* __restrict__ pw[S];
  const C2<T>* __restrict__ pv[S];
  T dsc[S];
  C2<T> al[S];
  bool on[S];
  const bool fz = WHICH == 2 && a.fused;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int si = grp * S + s;
    on[s] = false;
    in[s] = out[s] = nullptr;
    pw[s] = pv[s] = nullptr;
    dsc[s] = T(0);
    al[s] = {T(0), T(0)};
    if (si < a.nact) {
      const int sys = a.act[si];
      const SysScal& sc0 = a.sc[sys];
      if (sc0.status == ST_RUN) {
        on[s] = true;
        const int64_t o = (int64_t)sys * a.nE;
        dsc[s] = T(kIsignD * sc0.freq_omega);
        in[s] = (WHICH == 1 ? a.ph : a.sh) + o;
        out[s] = (WHICH == 1 ? a.v : a.t) + o;
        pw[s] = (WHICH == 1 ? a.rt : (fz ? a.r : a.s)) + o;
        pv[s] = a.v + o;
        al[s] = ccast<T>(sc0.alpha);
      }
    }
  }
  Z acc[S][2];
#pragma unroll
  for (int s = 0; s < S; ++s) acc[s][0] = acc[s][1] = {0.0, 0.0};
  int64_t lo, hi;
  chunk_range3(A.nXi, A.nYi, A.nInt, a.G, g, lo, hi);
  for_rows(A, lo, hi, threadIdx.x, BS, [&](const RowInfo& ri, auto cc) {
    constexpr int C = decltype(cc)::value;
    V cf[16];
    {
      using V4 = typename std::conditional<std::is_same<V, float>::value, float4, double2>::type;
      constexpr int PER = sizeof(V4) / sizeof(V);
      const V4* __restrict__ rp = reinterpret_cast<const V4*>(A.rowc + ri.e * 16);
#pragma unroll
      for (int q = 0; q < 16 / PER; ++q) {
        const V4 v4 = rp[q];
        if constexpr (PER == 4) {
          cf[4 * q] = v4.x; cf[4 * q + 1] = v4.y; cf[4 * q + 2] = v4.z; cf[4 * q + 3] = v4.w;
        } else {
          cf[2 * q] = v4.x; cf[2 * q + 1] = v4.y;
        }
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      if (!on[s]) continue;
      const C2<T>* __restrict__ x = in[s];
      const C2<T> xe = x[ri.e];
      C2<T> y = cscale(T(cf[0]), xe);
#pragma unroll
      for (int k = 1; k < 15; ++k) {
        const int64_t base = k < 7 ? ri.e : (k < 11 ? ri.b1 : ri.b2);
        const C2<T> xv = x[base + A.off[C][k]];
        const T c = T(cf[k]);
        y.x += c * xv.x;
        y.y += c * xv.y;
      }
      const T d = dsc[s] * T(cf[15]);
      y.x += -d * xe.y;
      y.y += d * xe.x;
      out[s][ri.e] = y;
      C2<T> wv = pw[s][ri.e];
      if (fz) wv = csub(wv, cmul(al[s], pv[s][ri.e]));
      if (WHICH == 1) {
        acc[s][0].x += double(wv.x) * y.x + double(wv.y) * y.y;
        acc[s][0].y += double(wv.x) * y.y - double(wv.y) * y.x;
      } else {
        acc[s][0].x += double(y.x) * y.x + double(y.y) * y.y;
        acc[s][1].x += double(y.x) * wv.x + double(y.y) * wv.y;
        acc[s][1].y += double(y.x) * wv.y - double(y.y) * wv.x;
      }
    }
  });
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int si = grp * S + s;
    if (si >= a.nact) continue;   // uniform over the block
    block_sum<2, BS>(acc[s], lds[s]);
    if (threadIdx.x == 0) {
      a.partials[((int64_t)si * a.G + g) * 2 + 0] = acc[s][0];
      a.partials[((int64_t)si * a.G + g) * 2 + 1] = acc[s][1];
    }
  }
}

}  // namespace mt
