// Brick layout of the inner solves of the mixed-precision solver.
//
// The FP32 inner BiCGStab (red-black tile DILU, multi-system SpMV) keeps its
// vectors, couplings and pivots tile by tile: the 8x8x8 tiles of the three
// interior edge lattices of the red-black tile plan, numbered as the plan
// numbers them (lattice tbase + ti + ntx*(tj + nty*tk)), each tile's 512
// points contiguous (i fastest, then j, then k), points past a partial tile's
// extent held at zero.  In the lattice order a tile row is 8 points (64 bytes
// of FP32 complex), half of a 128-byte line whose other half belongs to the
// neighbouring tile of the other colour: the red-black sweeps fetched about
// twice the bytes they used, in short pieces.  Bricked, a tile is 4 KB
// contiguous per system, and the SpMV's k-1 and k+1 neighbours of 7 of a
// tile's 8 planes lie in the tile itself.
//
// The preconditioned vectors p-hat and s-hat, which the SpMV gathers, can be
// kept with the systems of a pair (2m, 2m+1) side by side, entry
// (s/2)*2nB + 2*point + s%2: the SpMV then reads both systems of a pair with
// one 16-byte gather per neighbour (k_spmv_btp), the DILU writes them so
// (DiluArgs::zout), and every other inner vector stays system-major.
//
// The outer FP64 defect correction stays in the all-edge layout: its scaled
// residual is permuted into bricks before an inner solve and the inner
// correction permuted back after it (k_bk_to_brick / k_bk_from_brick, once
// per outer iteration).
#pragma once

#include "mt_dilu.hpp"

namespace mt {

constexpr int kBkPts = 512;
// chunks of tiles of the brick SpMV (at most; a multiple of the 8 XCDs)
constexpr int kBtChunks = 2048;

// Tile geometry of the bricks (the red-black tile plan's lattices).
struct BrickGeom {
  int A[3], B[3], C[3];        // interior lattice dims per edge component
  int oi[3], oj[3], ok[3];     // offset of the interior lattice in the full lattice
  int ntx[3], nty[3], ntz[3];  // tiles per direction
  int tbase[4];                // first tile of each lattice; tbase[3] = all tiles
  int64_t nB;                  // points per system: tiles * 512
};

// (component, di, dj, dk) of stencil slot s for a row of component c, in
// full-lattice index steps (the order of stencil_neighbours)
struct NbS { int cb, di, dj, dk; };
__host__ __device__ constexpr NbS bk_nb(int c, int s) {
  if (s == 0) return {c, 0, 0, 0};
  if (s <= 6) {
    const int d = (s & 1) ? -1 : 1, ax = (s - 1) >> 1;
    return {c, ax == 0 ? d : 0, ax == 1 ? d : 0, ax == 2 ? d : 0};
  }
  if (c == 0) {
    switch (s) {
      case 7: return {1, 0, -1, 0};  case 8: return {1, 1, -1, 0};
      case 9: return {1, 0, 0, 0};   case 10: return {1, 1, 0, 0};
      case 11: return {2, 0, 0, -1}; case 12: return {2, 1, 0, -1};
      case 13: return {2, 0, 0, 0};  default: return {2, 1, 0, 0};
    }
  }
  if (c == 1) {
    switch (s) {
      case 7: return {0, -1, 0, 0};  case 8: return {0, 0, 0, 0};
      case 9: return {0, -1, 1, 0};  case 10: return {0, 0, 1, 0};
      case 11: return {2, 0, 0, -1}; case 12: return {2, 0, 1, -1};
      case 13: return {2, 0, 0, 0};  default: return {2, 0, 1, 0};
    }
  }
  switch (s) {
    case 7: return {0, -1, 0, 0};  case 8: return {0, 0, 0, 0};
    case 9: return {0, -1, 0, 1};  case 10: return {0, 0, 0, 1};
    case 11: return {1, 0, -1, 0}; case 12: return {1, 0, 0, 0};
    case 13: return {1, 0, -1, 1}; default: return {1, 0, 0, 1};
  }
}

// Brick index of the interior edge of component C at full-lattice (I, J, K),
// or -1 when that edge is not interior (a boundary edge, whose coupling is
// zero, or past the lattice).
template <int C>
__device__ __forceinline__ int bk_at(const BrickGeom& g, int I, int J, int K) {
  const int i = I - g.oi[C], j = J - g.oj[C], k = K - g.ok[C];
  const bool in = (unsigned)i < (unsigned)g.A[C] && (unsigned)j < (unsigned)g.B[C] &&
                  (unsigned)k < (unsigned)g.C[C];
  const int t = g.tbase[C] + (i >> 3) + g.ntx[C] * ((j >> 3) + g.nty[C] * (k >> 3));
  const int q = t * kBkPts + (i & 7) + 8 * (j & 7) + 64 * (k & 7);
  return in ? q : -1;
}

// Lattice and tile position of tile t.
__device__ __forceinline__ void bk_tile(const BrickGeom& g, int t, int& c, int& ti, int& tj,
                                        int& tk) {
  c = t >= g.tbase[2] ? 2 : (t >= g.tbase[1] ? 1 : 0);
  int q = t - g.tbase[c];
  ti = q % g.ntx[c];
  q /= g.ntx[c];
  tj = q % g.nty[c];
  tk = q / g.nty[c];
}

// Tiles [t0, t1) of chunk g of G, the chunks dealt to the three lattices in
// turn (chunk g is part g/3 of lattice g%3): an XCD's G/8 consecutive chunks
// are then the same k-slab fraction of all three lattices, whose rows gather
// each other's values.
__device__ __forceinline__ void bk_tile_range3(const BrickGeom& g, int G, int gi, int& t0,
                                               int& t1) {
  const int c = gi % 3, j = gi / 3;
  const int cnt = (G - c + 2) / 3;
  const int64_t n = g.tbase[c + 1] - g.tbase[c];
  // balanced: every chunk gets floor or ceil of n / cnt tiles
  t0 = g.tbase[c] + int(n * j / cnt);
  t1 = g.tbase[c] + int(n * (j + 1) / cnt);
}

// ------------------------------------------------------------ permutations
// dst[q][b] = src[q][bmap[b]] (0 where bmap[b] < 0), q < nq
template <class T>
__global__ __launch_bounds__(BS) void k_bk_gather(const int* __restrict__ bmap, int64_t nB, int nq,
                                                  int64_t sstride, int64_t dstride,
                                                  const C2<T>* __restrict__ src,
                                                  C2<T>* __restrict__ dst) {
  for (int64_t idx = blockIdx.x * int64_t(BS) + threadIdx.x; idx < nB * nq;
       idx += int64_t(gridDim.x) * BS) {
    const int q = int(idx / nB);
    const int64_t b = idx - q * nB;
    const int e = bmap[b];
    dst[q * dstride + b] = e >= 0 ? src[q * sstride + e] : C2<T>{T(0), T(0)};
  }
}

// Row-major couplings [n][16] (15 slots and vms1) into brick order.
template <class V>
__global__ __launch_bounds__(BS) void k_bk_rows(const int* __restrict__ bmap, int64_t nB,
                                                const V* __restrict__ src, V* __restrict__ dst) {
  for (int64_t idx = blockIdx.x * int64_t(BS) + threadIdx.x; idx < nB * 16;
       idx += int64_t(gridDim.x) * BS) {
    const int64_t b = idx >> 4;
    const int e = bmap[b];
    dst[idx] = e >= 0 ? src[int64_t(e) * 16 + (idx & 15)] : V(0);
  }
}

// Slots 1..6 (the same-component couplings the DILU sweeps use) of the
// slot-major stencil [15][nE] into [6][nB].
template <class V>
__global__ __launch_bounds__(BS) void k_bk_slots(const int* __restrict__ bmap, int64_t nB,
                                                 int64_t nE, const V* __restrict__ coef,
                                                 V* __restrict__ dst) {
  for (int64_t idx = blockIdx.x * int64_t(BS) + threadIdx.x; idx < nB * 6;
       idx += int64_t(gridDim.x) * BS) {
    const int s = int(idx / nB);
    const int64_t b = idx - s * nB;
    const int e = bmap[b];
    dst[idx] = e >= 0 ? coef[(1 + s) * nE + e] : V(0);
  }
}

// Inner right-hand sides: r (bricks) = r32 (all-edge layout) of the active
// systems, the inner correction d (bricks) = 0.
__global__ __launch_bounds__(BS) void k_bk_to_brick(const int* __restrict__ act, int nact,
                                                    const int* __restrict__ bmap, int64_t nB,
                                                    int64_t nE, const C2<float>* __restrict__ r32,
                                                    C2<float>* __restrict__ rb,
                                                    C2<float>* __restrict__ db) {
  for (int64_t idx = blockIdx.x * int64_t(BS) + threadIdx.x; idx < nB * nact;
       idx += int64_t(gridDim.x) * BS) {
    const int q = int(idx / nB);
    const int64_t b = idx - q * nB;
    const int64_t s = act[q];
    const int e = bmap[b];
    rb[s * nB + b] = e >= 0 ? r32[s * nE + e] : C2<float>{0.f, 0.f};
    db[s * nB + b] = {0.f, 0.f};
  }
}

// The inner corrections back into the all-edge layout.
__global__ __launch_bounds__(BS) void k_bk_from_brick(const int* __restrict__ act, int nact,
                                                      const int* __restrict__ bmap, int64_t nB,
                                                      int64_t nE, const C2<float>* __restrict__ db,
                                                      C2<float>* __restrict__ d32) {
  for (int64_t idx = blockIdx.x * int64_t(BS) + threadIdx.x; idx < nB * nact;
       idx += int64_t(gridDim.x) * BS) {
    const int q = int(idx / nB);
    const int64_t b = idx - q * nB;
    const int64_t s = act[q];
    const int e = bmap[b];
    if (e >= 0) d32[s * nE + e] = db[s * nB + b];
  }
}

// FP64 <-> FP32 copies of complex vectors (n entries)
__global__ __launch_bounds__(BS) void k_bk_narrow(const Z* __restrict__ x, C2<float>* __restrict__ y,
                                                  int64_t n) {
  for (int64_t q = blockIdx.x * int64_t(BS) + threadIdx.x; q < n; q += int64_t(gridDim.x) * BS)
    y[q] = {float(x[q].x), float(x[q].y)};
}
__global__ __launch_bounds__(BS) void k_bk_widen(const C2<float>* __restrict__ x, Z* __restrict__ y,
                                                 int64_t n) {
  for (int64_t q = blockIdx.x * int64_t(BS) + threadIdx.x; q < n; q += int64_t(gridDim.x) * BS)
    y[q] = {double(x[q].x), double(x[q].y)};
}

// k_pupdate on the tiles of one colour of the red-black plan (tiles[0,
// ntiles): {lattice, ti, tj, tk}): the black tiles' p, when the red forward
// sweep forms the red tiles' itself.  Block (system, chunk g of the tiles),
// as k_pupdate's; the same arithmetic per point.
template <class T, class V>
__global__ __launch_bounds__(BS) void k_pupdate_tiles(IterArgs<T, V> a, const int4* __restrict__ tiles,
                                                      int ntiles, LatticeInfo l0, LatticeInfo l1,
                                                      LatticeInfo l2) {
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc = a.sc[sys];
  if (sc.status != ST_RUN) return;
  const int64_t o = (int64_t)sys * a.nE;
  const bool rs = sc.restart_due;
  const C2<T> beta = ccast<T>(sc.beta), om = ccast<T>(sc.omg);
  const int t0 = int(int64_t(ntiles) * g / a.G), t1 = int(int64_t(ntiles) * (g + 1) / a.G);
  for (int t = t0; t < t1; ++t) {
    const int4 tl = tiles[t];
    const LatticeInfo L = tl.x == 0 ? l0 : (tl.x == 1 ? l1 : l2);
    const int64_t b0 = o + (int64_t)(L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w)) * kBkPts;
    for (int q = threadIdx.x; q < kBkPts; q += BS) {
      const int64_t e = b0 + q;
      const C2<T> r = a.r[e];
      C2<T> p;
      if (rs) {
        a.rt[e] = r;
        p = r;
      } else {
        p = cadd(r, cmul(beta, csub(a.p[e], cmul(om, a.v[e]))));
      }
      a.p[e] = p;
    }
  }
}

// Point p of tile t for the brick SpMV: whether it is an interior edge, its
// brick index e, its 14 neighbours' brick indices (a neighbour that is not an
// interior edge has a zero coupling and the row's own index) and its 16
// couplings (the brick-ordered row-major copy: 15 slots and the omega-free
// diagonal).
template <class V>
__device__ __forceinline__ void bt_point(const BrickGeom& bg, const V* __restrict__ rowc, int t,
                                         int p, bool& valid, int& e, V (&cf)[16], int (&nb)[15]) {
  int c, ti, tj, tk;
  bk_tile(bg, t, c, ti, tj, tk);
  const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
  const int i = ti * 8 + li, j = tj * 8 + lj, k = tk * 8 + lk;
  valid = i < bg.A[c] && j < bg.B[c] && k < bg.C[c];
  e = t * kBkPts + p;
  const int ee = e;
  auto setup = [&](auto cc) {
    constexpr int C = decltype(cc)::value;
    const int I = i + bg.oi[C], J = j + bg.oj[C], K = k + bg.ok[C];
#pragma unroll
    for (int q = 1; q < 15; ++q) {
      const NbS d = bk_nb(C, q);
      int r;
      if (d.cb == 0) r = bk_at<0>(bg, I + d.di, J + d.dj, K + d.dk);
      else if (d.cb == 1) r = bk_at<1>(bg, I + d.di, J + d.dj, K + d.dk);
      else r = bk_at<2>(bg, I + d.di, J + d.dj, K + d.dk);
      nb[q] = r < 0 ? ee : r;
    }
  };
  if (c == 0) setup(Comp<0>());
  else if (c == 1) setup(Comp<1>());
  else setup(Comp<2>());
  using V4 = typename std::conditional<std::is_same<V, float>::value, float4, double2>::type;
  constexpr int PER = sizeof(V4) / sizeof(V);
  const V4* __restrict__ rp = reinterpret_cast<const V4*>(rowc + int64_t(e) * 16);
#pragma unroll
  for (int q = 0; q < 16 / PER; ++q) {
    const V4 v4 = rp[q];   // pads hold zero couplings
    if constexpr (PER == 4) {
      cf[4 * q] = v4.x; cf[4 * q + 1] = v4.y; cf[4 * q + 2] = v4.z; cf[4 * q + 3] = v4.w;
    } else {
      cf[2 * q] = v4.x; cf[2 * q + 1] = v4.y;
    }
  }
}

// ------------------------------------------------------------------ SpMV
// y_s = A x_s on bricks with the dot products of k_spmv_ms (WHICH=1:
// v = A ph, <rt,v>; WHICH=2: t = A sh, <t,t>, <t,s>), one block per chunk of
// tiles and group of sg active systems (four systems' sums per thread, as
// k_spmv_ms keeps them, cap a brick kernel at two or three waves per SIMD by
// registers): a thread owns one point of the tile, forms its 14 neighbours'
// brick indices and loads its 16 couplings (the brick-ordered row-major copy)
// once, then runs the group's systems one after another, two at a time.  A
// neighbour that is not an interior edge has a zero coupling and the row's
// own index.  The row sums run in k_spmv_ms's order.  Each wave reduces its
// lanes' dot contributions per system (wave_sum63) into LDS; the block then
// sums its waves in order into the per-(system, chunk) partials.
constexpr int kBtThreads = 512, kBtWaves = kBtThreads / 64, kBtMaxSys = 64;
template <class T, class V, int WHICH>
__global__ __launch_bounds__(kBtThreads) void k_spmv_bt(IterArgs<T, V> a, BrickGeom bg,
                                                        const V* __restrict__ rowc, int sg) {
  constexpr int NQ = WHICH == 1 ? 2 : 3;   // doubles per system: <rt,v>; <t,t>, <t,s>
  __shared__ double red[kBtWaves][kBtMaxSys][NQ];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  // blocks (chunk g, group of sg active systems), the groups of a chunk
  // consecutive inside an XCD (they read the same couplings)
  const int ngrp = (a.nact + sg - 1) / sg;
  const int w = blockIdx.x >> 3, grp = w % ngrp;
  const int g = (blockIdx.x & 7) * (a.G >> 3) + w / ngrp;
  const int s0 = grp * sg, nact = min(a.nact, s0 + sg);
  for (int q = tid; q < kBtWaves * kBtMaxSys * NQ; q += kBtThreads) (&red[0][0][0])[q] = 0.0;
  __syncthreads();
  // the <t,s> partner: s stored by the preconditioner, else r - alpha v
  const bool fz = WHICH == 2 && a.fused && !a.s_ok;
  int t0, t1;
  bk_tile_range3(bg, a.G, g, t0, t1);
  for (int t = t0; t < t1; ++t) {
    const int p = tid;
    bool valid;
    int e;
    V cf[16];
    int nb[15];
    bt_point(bg, rowc, t, p, valid, e, cf, nb);
    // one system: y = A x at this point, its dot contributions
    auto one = [&](int si, double* dq) {
      const int sys = a.act[si];
      const SysScal& sc0 = a.sc[sys];
#pragma unroll
      for (int q = 0; q < NQ; ++q) dq[q] = 0.0;
      if (sc0.status != ST_RUN) return false;
      const int64_t o = (int64_t)sys * a.nE;
      const C2<T>* __restrict__ x = (WHICH == 1 ? a.ph : a.sh) + o;
      const C2<T> xe = x[e];
      C2<T> y = cscale(T(cf[0]), xe);
#pragma unroll
      for (int q = 1; q < 15; ++q) {
        const C2<T> xv = x[nb[q]];
        const T cq = T(cf[q]);
        y.x += cq * xv.x;
        y.y += cq * xv.y;
      }
      const T d = T(kIsignD * sc0.freq_omega) * T(cf[15]);
      y.x += -d * xe.y;
      y.y += d * xe.x;
      if (valid) (WHICH == 1 ? a.v : a.t)[o + e] = y;
      if (!valid) y = {T(0), T(0)};
      C2<T> wv = (WHICH == 1 ? a.rt : (fz ? a.r : a.s))[o + e];
      if (fz) wv = csub(wv, cmul(ccast<T>(sc0.alpha), a.v[o + e]));
      if (WHICH == 1) {
        dq[0] = double(wv.x) * y.x + double(wv.y) * y.y;
        dq[1] = double(wv.x) * y.y - double(wv.y) * y.x;
      } else {
        dq[0] = double(y.x) * y.x + double(y.y) * y.y;
        dq[1] = double(y.x) * wv.x + double(y.y) * wv.y;
        dq[2] = double(y.x) * wv.y - double(y.y) * wv.x;
      }
      return true;
    };
    for (int si = s0; si < nact; si += 2) {
      double d0[NQ], d1[NQ];
      const bool r0 = one(si, d0);
      const bool r1 = si + 1 < nact ? one(si + 1, d1) : false;
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        d0[q] = wave_sum63(d0[q]);
        if (si + 1 < nact) d1[q] = wave_sum63(d1[q]);
      }
      if (lane == 63) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (r0) red[wv][si][q] += d0[q];
          if (r1) red[wv][si + 1][q] += d1[q];
        }
      }
    }
  }
  __syncthreads();
  for (int si = s0 + tid; si < nact; si += kBtThreads) {
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      v[q] = 0.0;
      for (int w = 0; w < kBtWaves; ++w) v[q] += red[w][si][q];
    }
    Z* out = a.partials + ((int64_t)si * a.G + g) * 2;
    if (WHICH == 1) {
      out[0] = {v[0], v[1]};
      out[1] = {0.0, 0.0};
    } else {
      out[0] = {v[0], 0.0};
      out[1] = {v[1], v[2]};
    }
  }
}

// k_spmv_bt for the FP32 inner solves with paired x: ph (WHICH=1) or sh
// (WHICH=2) holds the two systems of a pair interleaved, so one 16-byte
// gather per neighbour serves both (the addressers, which bound the
// single-system kernel, process half the gather instructions).  A block
// takes a chunk of tiles and a group of spg entries of the pair list
// {system 2m, system 2m+1, active index of each or -1}; a pair with one
// system stopped reads both halves and keeps only the running one.  Per
// system the arithmetic, its order and the partial sums (per active index,
// over the block's tiles and waves in order) are k_spmv_bt's.
template <int WHICH>
__global__ __launch_bounds__(kBtThreads, 4) void k_spmv_btp(IterArgs<float, float> a, BrickGeom bg,
                                                         const float* __restrict__ rowc,
                                                         const int4* __restrict__ pairs, int npair,
                                                         int spg) {
  constexpr int NQ = WHICH == 1 ? 2 : 3;
  __shared__ double red[kBtWaves][kBtMaxSys][NQ];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int ngrp = (npair + spg - 1) / spg;
  const int w = blockIdx.x >> 3, grp = w % ngrp;
  const int g = (blockIdx.x & 7) * (a.G >> 3) + w / ngrp;
  const int p0 = grp * spg, p1 = min(npair, p0 + spg);
  for (int q = tid; q < kBtWaves * kBtMaxSys * NQ; q += kBtThreads) (&red[0][0][0])[q] = 0.0;
  __syncthreads();
  // the <t,s> partner: s stored by the preconditioner, else r - alpha v
  const bool fz = WHICH == 2 && a.fused && !a.s_ok;
  const float4* __restrict__ xi = reinterpret_cast<const float4*>(WHICH == 1 ? a.phi : a.shi);
  const int64_t xpair = a.ppair >> 1;   // float4 entries per pair
  int t0, t1;
  bk_tile_range3(bg, a.G, g, t0, t1);
  for (int t = t0; t < t1; ++t) {
    bool valid;
    int e;
    float cf[16];
    int nb[15];
    bt_point(bg, rowc, t, tid, valid, e, cf, nb);
    for (int pi = p0; pi < p1; ++pi) {
      const int4 pr = pairs[pi];   // {2m, 2m+1, active index of 2m, of 2m+1}
      // both systems' records up front (status, omega, alpha)
      bool ra = false, rb = false;
      float wa = 0.f, wb = 0.f;
      C2<float> ala = {0.f, 0.f}, alb = {0.f, 0.f};
      if (pr.z >= 0) {
        const SysScal& sc0 = a.sc[pr.x];
        ra = sc0.status == ST_RUN;
        wa = float(kIsignD * sc0.freq_omega);
        ala = ccast<float>(sc0.alpha);
      }
      if (pr.w >= 0) {
        const SysScal& sc1 = a.sc[pr.y];
        rb = sc1.status == ST_RUN;
        wb = float(kIsignD * sc1.freq_omega);
        alb = ccast<float>(sc1.alpha);
      }
      if (!ra && !rb) continue;   // uniform over the block
      const float4* __restrict__ x = xi + (int64_t)(pr.x >> 1) * xpair;
      const float4 xe = x[e];
      C2<float> y0 = cscale(cf[0], C2<float>{xe.x, xe.y});
      C2<float> y1 = cscale(cf[0], C2<float>{xe.z, xe.w});
#pragma unroll
      for (int q = 1; q < 15; ++q) {
        const float4 xv = x[nb[q]];
        const float cq = cf[q];
        y0.x += cq * xv.x;
        y0.y += cq * xv.y;
        y1.x += cq * xv.z;
        y1.y += cq * xv.w;
      }
      // one system: the diagonal term, y stored (on: running), its dot
      // contributions.  Both systems of the pair are finished whether running
      // or not (a stopped or absent second one reads the first's partner):
      // were the second one's work in a branch, the compiler would sink its
      // half of every gather into it, splitting each 16-byte gather in two.
      auto fin = [&](int sys, bool on, float wsc, C2<float> al, C2<float> y, float xr, float xim,
                     double* dq) {
        const float d = wsc * cf[15];
        y.x += -d * xim;
        y.y += d * xr;
        const int64_t o = (int64_t)sys * a.nE;
        if (valid && on) (WHICH == 1 ? a.v : a.t)[o + e] = y;
        if (!valid) y = {0.f, 0.f};
        C2<float> pw = (WHICH == 1 ? a.rt : (fz ? a.r : a.s))[o + e];
        if (fz) pw = csub(pw, cmul(al, a.v[o + e]));
        if (WHICH == 1) {
          dq[0] = double(pw.x) * y.x + double(pw.y) * y.y;
          dq[1] = double(pw.x) * y.y - double(pw.y) * y.x;
        } else {
          dq[0] = double(y.x) * y.x + double(y.y) * y.y;
          dq[1] = double(y.x) * pw.x + double(y.y) * pw.y;
          dq[2] = double(y.x) * pw.y - double(y.y) * pw.x;
        }
      };
      double d0[NQ], d1[NQ];
      fin(pr.x, ra, wa, ala, y0, xe.x, xe.y, d0);
      fin(rb ? pr.y : pr.x, rb, wb, alb, y1, xe.z, xe.w, d1);
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        d0[q] = wave_sum63(d0[q]);
        d1[q] = wave_sum63(d1[q]);
      }
      if (lane == 63) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          if (ra) red[wv][pr.z][q] += d0[q];
          if (rb) red[wv][pr.w][q] += d1[q];
        }
      }
    }
  }
  __syncthreads();
  // the per-(system, chunk) partials of every active index of the group
  for (int k = tid; k < 2 * (p1 - p0); k += kBtThreads) {
    const int4 pr = pairs[p0 + (k >> 1)];
    const int si = (k & 1) ? pr.w : pr.z;
    if (si < 0) continue;
    double v[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      v[q] = 0.0;
      for (int w2 = 0; w2 < kBtWaves; ++w2) v[q] += red[w2][si][q];
    }
    Z* out = a.partials + ((int64_t)si * a.G + g) * 2;
    if (WHICH == 1) {
      out[0] = {v[0], v[1]};
      out[1] = {0.0, 0.0};
    } else {
      out[0] = {v[0], 0.0};
      out[1] = {v[1], v[2]};
    }
  }
}

// ------------------------------------------------- red-black DILU on bricks
// k_dilu_rbl's sweeps (the same LDS layout and level order) on bricked data:
// a tile's right-hand sides, pivots and the six couplings it uses are
// contiguous runs.  The red forward sweep leaves w = dinv*y of the red tiles
// in the output (k_dilu_rbl leaves y): that is what the black tiles take from
// their red neighbours, and the red backward sweep forms z = w - dinv*(sum of
// the upper terms), as the black backward sweep does (the same operator as
// dinv*(y - sum), rounded differently).  A tile reads its neighbours' k- and
// j-faces from their bricks (a k-face is 512 contiguous bytes, a j-face eight
// runs of 64) and their i-faces, one value per 64 bytes there, from the
// packed i-faces (DiluArgs::xf) the other colour wrote: w of the red tiles,
// z of the black tiles.
// DiluArgs on bricks: coef = the [6][nB] couplings (slots 1..6), cstride =
// nB, vstride = nB, dinv = [nfreq][nB].
template <class T, class V, int MODE, int NW, bool ZP = false>
__global__ __launch_bounds__(64 * NW) void k_dilu_bk(DiluArgs<T, V> a, const TileShape* shp) {
  constexpr int TX = 8, NP = kRbNP, NU = 8, NLEV = 22, NB = kRbNB;
  constexpr bool FWDP = MODE != RB_RED_BWD;   // has a forward sweep
  constexpr bool BWDP = MODE != RB_RED_FWD;   // has a backward sweep
  constexpr bool HALO = MODE != RB_RED_FWD;   // reads face neighbours
  struct Cpl { V c[3]; };
  __shared__ C2<T> Wall[NW][NB];                 // per system: the box
  // per-point data by level slot (TileShape::rslot)
  __shared__ C2<T> DVall[NW][NP];                // per system: pivots (red forward: then y)
  __shared__ Cpl CL[FWDP ? NP : 1];              // raw lower couplings (slots 1, 3, 5)
  __shared__ Cpl CH[BWDP ? NP : 1];              // raw upper couplings (slots 2, 4, 6)
  __shared__ int s_sys[NW], s_ok[NW], s_fidx[NW];
  __shared__ C2<T> s_al[NW];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int ngrp = (a.nact + NW - 1) / NW;
  const int nwork = a.ntiles * ngrp;
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const int chunk = (nwork + 7) / 8;
  const int w = xcd * chunk + lb;
  if (lb >= chunk || w >= nwork) return;   // uniform over the workgroup
  const int tix = w / ngrp, grp = w - tix * ngrp;
  if (tid < NW) {
    const int si = grp * NW + tid;
    int sy = 0, ok = 0, fi = 0;
    C2<T> al = {T(0), T(0)};
    if (si < a.nact) {
      sy = a.act[si];
      const SysScal& sc = a.sc[sy];
      ok = sc.status == a.status_req;
      fi = sc.fidx;
      al = ccast<T>(sc.alpha);
    }
    s_sys[tid] = sy;
    s_ok[tid] = ok;
    s_fidx[tid] = fi;
    s_al[tid] = al;
  }
  const int4 tl = a.tiles[tix];
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
  const int ex = min(TX, L.A - gi0), ey = min(TX, L.B - gj0), ez = min(TX, L.C - gk0);
  const int tself = L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w);
  const int tb = tself * kBkPts;   // the tile's first brick index
  const V* __restrict__ cfs = a.coef;
  const int64_t cst = a.cstride;
  const int st[3] = {1, kRbBX, kRbPK};
  // ---- the tile's couplings, once for the NW systems
  for (int p = tid; p < NP; p += 64 * NW) {
    const int sl = shp->rslot[p];
    Cpl c, h;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      c.c[s] = FWDP ? cfs[(int64_t)(2 * s) * cst + tb + p] : V(0);
      h.c[s] = BWDP ? cfs[(int64_t)(2 * s + 1) * cst + tb + p] : V(0);
    }
    if (FWDP) CL[sl] = c;
    if (BWDP) CH[sl] = h;
  }
  __syncthreads();   // s_ tables
  // this lane's point on every hyperplane of the tile (at most 48 per plane),
  // formed before the sweeps: box index | level slot << 10, or ~0u (no point,
  // or past a partial tile's extent)
  unsigned int pk[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) pk[h] = shp->lpts[h][lane];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) {
    const unsigned int v = pk[h];
    const int li = v & 7, lj = (v >> 3) & 7, lk = (v >> 6) & 7;
    const bool ok = v != 0xffffffffu && li < ex && lj < ey && lk < ez;
    pk[h] = ok ? (unsigned(rb_box(li, lj, lk)) | ((v >> 9) << 10)) : 0xffffffffu;
    asm volatile("" : "+v"(pk[h]));   // formed here, not per sweep step
  }
  const bool mine = __builtin_amdgcn_readfirstlane(s_ok[wv]) != 0;   // 0 past nact
  const int sys = __builtin_amdgcn_readfirstlane(s_sys[wv]);
  const int fidx = __builtin_amdgcn_readfirstlane(s_fidx[wv]);
  C2<T>* __restrict__ W = Wall[wv];
  C2<T>* __restrict__ DV = DVall[wv];
  const int64_t o = (int64_t)sys * a.vstride;
  // the other colour's results (red backward with zout: the black tiles' z,
  // pair-interleaved)
  // (ZP: the launch has a.zout)
  constexpr bool zp = ZP && MODE != RB_RED_FWD;
  constexpr int zs = zp ? 2 : 1;
  const C2<T>* __restrict__ pout =
      (MODE == RB_RED_BWD && zp) ? a.zout + (int64_t)(sys >> 1) * a.zpair + (sys & 1) : a.out + o;
  const C2<T>* __restrict__ xfi =
      a.xf + (MODE == RB_BLACK_FB ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys;   // faces read
  C2<T>* __restrict__ xfw =
      a.xf + (MODE == RB_RED_FWD ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys;    // faces written
  // face values kept for the fold (black: the high faces; red backward: the
  // low faces and the lower couplings of the tile points next to them)
  C2<T> fold_v[3];
  T fold_c[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    fold_v[s] = {T(0), T(0)};
    fold_c[s] = T(0);
  }
  if (mine) {
    for (int q = lane; q < NB; q += 64) W[q] = {T(0), T(0)};
    const C2<T>* __restrict__ pin = (FWDP ? a.in : a.out) + o + tb;
    const C2<T>* __restrict__ pin2 = a.in2 ? a.in2 + o + tb : nullptr;
    const C2<T>* __restrict__ dinv = a.dinv + (int64_t)fidx * cst + tb;
    const C2<T> al = s_al[wv];
    C2<T> rv[NU], v2[NU], dd[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      // points past a partial tile's extent hold zeros
      rv[u] = pin[p];
      v2[u] = (FWDP && pin2) ? pin2[p] : C2<T>{T(0), T(0)};
      dd[u] = dinv[p];
    }
    C2<T> fv[6];
    if (HALO) {
      const int fa = lane & 7, fb = lane >> 3;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1;
        const bool hi = f & 1;
        int ti, tj, tk;   // tile point next to the face
        if (dir == 0) { ti = hi ? 7 : 0; tj = fa; tk = fb; }
        else if (dir == 1) { ti = fa; tj = hi ? 7 : 0; tk = fb; }
        else { ti = fa; tj = fb; tk = hi ? 7 : 0; }
        const int ni = gi0 + ti + (dir == 0 ? (hi ? 1 : -1) : 0);
        const int nj = gj0 + tj + (dir == 1 ? (hi ? 1 : -1) : 0);
        const int nk = gk0 + tk + (dir == 2 ? (hi ? 1 : -1) : 0);
        fv[f] = {T(0), T(0)};
        const bool pin_ok = ti < ex && tj < ey && tk < ez;
        const bool ok = pin_ok && ni >= 0 && nj >= 0 && nk >= 0 && ni < L.A && nj < L.B && nk < L.C;
        if (ok) {
          // the neighbour tile's facing face: the i-faces packed (entry
          // tj + 8*tk = lane), the j- and k-faces from the neighbour's brick
          const int d = hi ? 1 : -1;
          const int tn = L.tbase + (tl.y + (dir == 0 ? d : 0)) +
                         L.ntx * ((tl.z + (dir == 1 ? d : 0)) + L.nty * (tl.w + (dir == 2 ? d : 0)));
          if (dir == 0) {
            fv[f] = xfi[(int64_t)tn * kRbFaces + (hi ? 0 : 1) * 64 + lane];
          } else {
            const int fl = hi ? 0 : TX - 1;   // the neighbour's facing layer
            const int pn = dir == 1 ? ti + TX * (fl + TX * tk) : ti + TX * (tj + TX * fl);
            fv[f] = pout[(MODE == RB_RED_BWD ? zs : 1) * ((int64_t)tn * kBkPts + pn)];
          }
        }
        if (MODE == RB_RED_BWD && !hi && pin_ok)
          fold_c[dir] = T(cfs[(int64_t)(2 * dir) * cst + tb + ti + TX * (tj + TX * tk)]);
      }
    }
    // ---- stage into LDS: the right-hand side in the box, the pivots
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      C2<T> r = rv[u];
      if (FWDP && pin2) {
        r = csub(r, cmul(al, v2[u]));
        if (a.sout) a.sout[o + tb + p] = r;
      }
      W[rb_box(li, lj, lk)] = r;
      DV[shp->rslot[p]] = dd[u];
    }
    if (HALO) {
      const int fa = lane & 7, fb = lane >> 3;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1;
        const bool hi = f & 1;
        int ti, tj, tk;
        if (dir == 0) { ti = hi ? 7 : 0; tj = fa; tk = fb; }
        else if (dir == 1) { ti = fa; tj = hi ? 7 : 0; tk = fb; }
        else { ti = fa; tj = fb; tk = hi ? 7 : 0; }
        const C2<T> v = fv[f];
        const bool inbox = (MODE == RB_BLACK_FB) ? !hi : hi;
        if (!inbox) {
          fold_v[dir] = v;
          continue;
        }
        if (!(ti < ex && tj < ey && tk < ez)) continue;
        const int qp = rb_box(ti, tj, tk);
        int qs = hi ? qp + st[dir] : qp - st[dir];
        if (qs >= NB) qs -= NB;
        W[qs] = v;
      }
    }
  }
  __syncthreads();   // the shared couplings; the rest is per wave
  if (!mine) return;
  // fold the face terms the sweeps do not read, one direction after another:
  // black forward, the high faces (red neighbours above, upper couplings);
  // red backward, the low faces (black neighbours below, lower couplings)
  if (HALO) {
    const int fa = lane & 7, fb = lane >> 3;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const int t0 = (MODE == RB_BLACK_FB) ? TX - 1 : 0;
      int ti, tj, tk;
      if (s == 0) { ti = t0; tj = fa; tk = fb; }
      else if (s == 1) { ti = fa; tj = t0; tk = fb; }
      else { ti = fa; tj = fb; tk = t0; }
      if (ti < ex && tj < ey && tk < ez) {
        const int q = rb_box(ti, tj, tk);
        const int sl = shp->rslot[ti + TX * (tj + TX * tk)];
        const T c = (MODE == RB_BLACK_FB) ? T(CH[sl].c[s]) : fold_c[s];
        const C2<T> v = fold_v[s];
        C2<T> r = W[q];
        if (MODE == RB_BLACK_FB) {
          r.x -= c * v.x;
          r.y -= c * v.y;
        } else {
          // the box holds w = dinv*y: subtract dinv*(c v)
          const C2<T> d = cmul(DV[sl], C2<T>{c * v.x, c * v.y});
          r.x -= d.x;
          r.y -= d.y;
        }
        W[q] = r;
      }
      wave_sync();
    }
  }
  // The sweeps take no masks: a red forward step's low neighbours outside the
  // tile are zero cells of the box (red forward stages no faces); a black
  // point uses every lower coupling; before the black backward sweep the
  // low-face cells are cleared, so that every upper neighbour outside the
  // tile (an i-high cell is the next row's i-low cell, a j-high cell the next
  // plane's j-low cells, a k-high index wraps to the k-low plane) reads zero;
  // red backward uses every upper coupling with the high faces in the box.
  // A neighbour past a partial tile's extent is a pad point: zero.
  if (FWDP) {
#pragma unroll
    for (int h = 0; h < NLEV; ++h) {
      const unsigned int pt = pk[h];
      if (pt != 0xffffffffu) {
        const int q = pt & 0x3ff, p = pt >> 10;
        const C2<T> dvp = DV[p];
        const Cpl c4 = CL[p];
        C2<T> y = W[q];
        C2<T> v[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) v[s] = W[q - st[s]];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = T(c4.c[s]);
          y.x -= c * v[s].x;
          y.y -= c * v[s].y;
        }
        W[q] = cmul(dvp, y);
      }
      wave_sync();
    }
  }
  if (MODE == RB_BLACK_FB) {
    const int fa = lane & 7, fb = lane >> 3;
    W[rb_box(0, fa, fb) - 1] = {T(0), T(0)};
    W[rb_box(fa, 0, fb) - kRbBX] = {T(0), T(0)};
    W[rb_box(fa, fb, 0) - kRbPK] = {T(0), T(0)};
    wave_sync();
  }
  if (BWDP) {
#pragma unroll
    for (int h = NLEV - 1; h >= 0; --h) {
      const unsigned int pt = pk[h];
      if (pt != 0xffffffffu) {
        const int q = pt & 0x3ff, p = pt >> 10;
        const C2<T> dvp = DV[p];
        const Cpl h4 = CH[p];
        const C2<T> b = W[q];
        C2<T> z[3];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          // q + st < 2 NB: the index past the box wraps to its start
          const unsigned int qn = unsigned(q + st[s]);
          z[s] = W[s == 0 ? qn : min(qn, qn - unsigned(NB))];
        }
        // y = -(sum of the upper terms); the box holds w = dv*y_forward
        C2<T> y = {T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = T(h4.c[s]);
          y.x -= c * z[s].x;
          y.y -= c * z[s].y;
        }
        const C2<T> d = cmul(dvp, y);
        W[q] = {b.x + d.x, b.y + d.y};
      }
      wave_sync();
    }
  }
  // ---- write the tile back (red forward: y; otherwise z) and the packed
  // faces the other colour reads next (red forward: dinv*y; black: z)
  C2<T>* __restrict__ dst =
      zp ? a.zout + (int64_t)(sys >> 1) * a.zpair + (sys & 1) + 2 * (int64_t)tb : a.out + o + tb;
  C2<T>* __restrict__ xfo = xfw + (int64_t)tself * kRbFaces;
  for (int p = lane; p < NP; p += 64) {
    const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
    if (!(li < ex && lj < ey && lk < ez)) continue;
    const int q = rb_box(li, lj, lk);
    const C2<T> zv = W[q];
    dst[zs * p] = zv;
    if (MODE != RB_RED_BWD && (li == 0 || li == TX - 1)) xfo[(li == 0 ? 0 : 64) + lj + 8 * lk] = zv;
  }
}

// The same three launches with persistent workgroups: a workgroup walks its
// XCD's items (tile, group of NW systems) with a stride, and while the waves
// sweep one item out of LDS, the loads of the next item (right-hand sides,
// pivots, couplings, faces; the system records and the tile before them) are
// in flight into registers.  The sweeps leave the loads little to overlap
// with otherwise: a workgroup's tile data arrive in one burst and the memory
// pipe idles while its waves sweep.  The arithmetic is k_dilu_bk's.
template <class T, class V, int MODE, int NW, bool ZP = false, bool PU = false>
__global__ __launch_bounds__(64 * NW) void k_dilu_bkp(DiluArgs<T, V> a, const TileShape* shp) {
  constexpr int TX = 8, NP = kRbNP, NU = 8, NLEV = 22, NB = kRbNB;
  static_assert(!PU || MODE == RB_RED_FWD, "the p-update is folded into the red forward sweep");
  constexpr bool FWDP = MODE != RB_RED_BWD;   // has a forward sweep
  constexpr bool BWDP = MODE != RB_RED_FWD;   // has a backward sweep
  constexpr bool HALO = MODE != RB_RED_FWD;   // reads face neighbours
  constexpr int CPT = NP / (64 * NW);         // tile points per thread for the couplings
  static_assert(CPT * 64 * NW == NP, "couplings: whole points per thread");
  struct Cpl { V c[3]; };
  __shared__ C2<T> Wall[NW][NB];
  __shared__ C2<T> DVall[NW][NP];
  __shared__ Cpl CL[FWDP ? NP : 1];
  __shared__ Cpl CH[BWDP ? NP : 1];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int ngrp = (a.nact + NW - 1) / NW;
  const int nwork = a.ntiles * ngrp;
  const int xcd = blockIdx.x & 7, nb8 = gridDim.x >> 3;
  const int chunk = (nwork + 7) / 8;
  const int wend = min(nwork, (xcd + 1) * chunk);
  int w = xcd * chunk + (blockIdx.x >> 3);
  if (w >= wend) return;   // uniform over the workgroup
  C2<T>* __restrict__ W = Wall[wv];
  C2<T>* __restrict__ DV = DVall[wv];
  const V* __restrict__ cfs = a.coef;
  const int64_t cst = a.cstride;
  const int st[3] = {1, kRbBX, kRbPK};
  const int fa = lane & 7, fb = lane >> 3;
  // constant over the items: the lane's point of every hyperplane (box index
  // | level slot << 10 | li << 19 | lj << 22 | lk << 25, or ~0u), the level
  // slots of its staging points and of its coupling points
  unsigned int pk[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) {
    const unsigned int v = shp->lpts[h][lane];
    const unsigned int li = v & 7, lj = (v >> 3) & 7, lk = (v >> 6) & 7;
    pk[h] = v == 0xffffffffu ? v
                             : (unsigned(rb_box(li, lj, lk)) | ((v >> 9) << 10) | (li << 19) | (lj << 22) |
                                (lk << 25));
  }
  int sl[NU], csl[CPT];
#pragma unroll
  for (int u = 0; u < NU; ++u) sl[u] = shp->rslot[lane + 64 * u];
#pragma unroll
  for (int c = 0; c < CPT; ++c) csl[c] = shp->rslot[tid + 64 * NW * c];
  // ---- the prefetched item
  struct Geo {
    int ok, sys, fidx, tself, ex, ey, ez;
    C2<T> al;
    int rs;           // PU: restart due
    C2<T> beta, om;   // PU: beta, omega
  };
  Geo nx;
  C2<T> rv[NU], v2[NU], dd[NU], fv[6];
  C2<T> r3[PU ? NU : 1];   // PU: r
  V cp[CPT][6];
  T fc[3];
  auto load = [&](int wi) {
    const int tix = wi / ngrp, grp = wi - tix * ngrp;
    const int si = grp * NW + wv;
    nx.ok = 0; nx.sys = 0; nx.fidx = 0;
    nx.al = {T(0), T(0)};
    if (si < a.nact) {
      nx.sys = a.act[si];
      const SysScal& sc = a.sc[nx.sys];
      nx.ok = sc.status == a.status_req;
      nx.fidx = sc.fidx;
      nx.al = ccast<T>(sc.alpha);
      if constexpr (PU) {
        nx.rs = sc.restart_due;
        nx.beta = ccast<T>(sc.beta);
        nx.om = ccast<T>(sc.omg);
      }
    }
    const int4 tl = a.tiles[tix];
    const LatticeInfo L = a.lat[tl.x];
    const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
    nx.ex = min(TX, L.A - gi0);
    nx.ey = min(TX, L.B - gj0);
    nx.ez = min(TX, L.C - gk0);
    nx.tself = L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w);
    const int tb = nx.tself * kBkPts;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int p = tid + 64 * NW * c;
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        cp[c][s] = FWDP ? cfs[(int64_t)(2 * s) * cst + tb + p] : V(0);
        cp[c][3 + s] = BWDP ? cfs[(int64_t)(2 * s + 1) * cst + tb + p] : V(0);
      }
    }
    if (!nx.ok) return;
    const int64_t o = (int64_t)nx.sys * a.vstride;
    const C2<T>* __restrict__ pin = (FWDP ? a.in : a.out) + o + tb;
    const C2<T>* __restrict__ pin2 = a.in2 ? a.in2 + o + tb : nullptr;
    const C2<T>* __restrict__ dinv = a.dinv + (int64_t)nx.fidx * cst + tb;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      rv[u] = pin[p];   // points past a partial tile's extent hold zeros
      if constexpr (PU) {
        v2[u] = a.pu_v[o + tb + p];
        r3[u] = a.pu_r[o + tb + p];
      } else {
        v2[u] = (FWDP && pin2) ? pin2[p] : C2<T>{T(0), T(0)};
      }
      dd[u] = dinv[p];
    }
    if (HALO) {
      // the other colour's results (red backward with zout: the black
      // tiles' z, pair-interleaved)
      constexpr bool zr = ZP && MODE == RB_RED_BWD;   // (ZP: the launch has a.zout)
      const C2<T>* __restrict__ pout =
          zr ? a.zout + (int64_t)(nx.sys >> 1) * a.zpair + (nx.sys & 1) : a.out + o;
      constexpr int zs = zr ? 2 : 1;
      const C2<T>* __restrict__ xfi =
          a.xf + (MODE == RB_BLACK_FB ? 0 : a.xfhalf) + (int64_t)nx.sys * a.xfsys;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1;
        const bool hi = f & 1;
        int ti, tj, tk;   // tile point next to the face
        if (dir == 0) { ti = hi ? 7 : 0; tj = fa; tk = fb; }
        else if (dir == 1) { ti = fa; tj = hi ? 7 : 0; tk = fb; }
        else { ti = fa; tj = fb; tk = hi ? 7 : 0; }
        const int ni = gi0 + ti + (dir == 0 ? (hi ? 1 : -1) : 0);
        const int nj = gj0 + tj + (dir == 1 ? (hi ? 1 : -1) : 0);
        const int nk = gk0 + tk + (dir == 2 ? (hi ? 1 : -1) : 0);
        fv[f] = {T(0), T(0)};
        const bool pin_ok = ti < nx.ex && tj < nx.ey && tk < nx.ez;
        const bool ok = pin_ok && ni >= 0 && nj >= 0 && nk >= 0 && ni < L.A && nj < L.B && nk < L.C;
        if (ok) {
          const int d = hi ? 1 : -1;
          const int tn = L.tbase + (tl.y + (dir == 0 ? d : 0)) +
                         L.ntx * ((tl.z + (dir == 1 ? d : 0)) + L.nty * (tl.w + (dir == 2 ? d : 0)));
          if (dir == 0) {
            fv[f] = xfi[(int64_t)tn * kRbFaces + (hi ? 0 : 1) * 64 + lane];
          } else {
            const int fl = hi ? 0 : TX - 1;   // the neighbour's facing layer
            const int pn = dir == 1 ? ti + TX * (fl + TX * tk) : ti + TX * (tj + TX * fl);
            fv[f] = pout[zs * ((int64_t)tn * kBkPts + pn)];
          }
        }
        if (MODE == RB_RED_BWD && !hi) {
          fc[dir] = T(0);
          if (pin_ok) fc[dir] = T(cfs[(int64_t)(2 * dir) * cst + tb + ti + TX * (tj + TX * tk)]);
        }
      }
    }
  };
  load(w);
  while (true) {
    // the point table is loop-invariant: kept opaque, so that what the sweeps
    // derive from it (addresses of 22 levels) is not hoisted out of the loop
    // into registers
#pragma unroll
    for (int h = 0; h < NLEV; ++h) asm volatile("" : "+v"(pk[h]));
    const Geo cu = nx;
    const int ex = cu.ex, ey = cu.ey, ez = cu.ez;
    const bool mine = __builtin_amdgcn_readfirstlane(cu.ok) != 0;
    __syncthreads();   // the previous item's sweeps and write-back are done
    // ---- stage: the shared couplings, per wave the box and the pivots
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      if (FWDP) CL[csl[c]] = Cpl{{cp[c][0], cp[c][1], cp[c][2]}};
      if (BWDP) CH[csl[c]] = Cpl{{cp[c][3], cp[c][4], cp[c][5]}};
    }
    C2<T> fold_v[3];
    T fold_c[3];
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      fold_v[s] = {T(0), T(0)};
      fold_c[s] = MODE == RB_RED_BWD ? fc[s] : T(0);
    }
    if (mine) {
      for (int q = lane; q < NB; q += 64) W[q] = {T(0), T(0)};
      wave_sync();
      C2<T>* __restrict__ so =
          a.sout ? a.sout + (int64_t)cu.sys * a.vstride + (int64_t)cu.tself * kBkPts : nullptr;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        C2<T> r = rv[u];
        if constexpr (PU) {
          // k_pupdate's p, stored for the red tiles
          const int64_t pe = (int64_t)cu.sys * a.vstride + (int64_t)cu.tself * kBkPts + lane + 64 * u;
          if (cu.rs) {
            a.pu_rt[pe] = r3[u];
            r = r3[u];
          } else {
            r = cadd(r3[u], cmul(cu.beta, csub(r, cmul(cu.om, v2[u]))));
          }
          a.pu_p[pe] = r;
        } else if (FWDP && a.in2) {
          r = csub(r, cmul(cu.al, v2[u]));
          if (so) so[lane + 64 * u] = r;
        }
        W[rb_box(fa, fb & 7, u)] = r;
        DV[sl[u]] = dd[u];
      }
      if (HALO) {
#pragma unroll
        for (int f = 0; f < 6; ++f) {
          const int dir = f >> 1;
          const bool hi = f & 1;
          int ti, tj, tk;
          if (dir == 0) { ti = hi ? 7 : 0; tj = fa; tk = fb; }
          else if (dir == 1) { ti = fa; tj = hi ? 7 : 0; tk = fb; }
          else { ti = fa; tj = fb; tk = hi ? 7 : 0; }
          const bool inbox = (MODE == RB_BLACK_FB) ? !hi : hi;
          if (!inbox) {
            fold_v[dir] = fv[f];
            continue;
          }
          if (!(ti < ex && tj < ey && tk < ez)) continue;
          const int qp = rb_box(ti, tj, tk);
          int qs = hi ? qp + st[dir] : qp - st[dir];
          if (qs >= NB) qs -= NB;
          W[qs] = fv[f];
        }
      }
    }
    // ---- the next item's loads, in flight during this item's sweeps
    const int wn = w + nb8;
    const bool more = wn < wend;
    if (more) load(wn);
    __syncthreads();   // the shared couplings
    if (mine) {
      const int64_t o = (int64_t)cu.sys * a.vstride;
      // fold the face terms the sweeps do not read (k_dilu_bk)
      if (HALO) {
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const int t0 = (MODE == RB_BLACK_FB) ? TX - 1 : 0;
          int ti, tj, tk;
          if (s == 0) { ti = t0; tj = fa; tk = fb; }
          else if (s == 1) { ti = fa; tj = t0; tk = fb; }
          else { ti = fa; tj = fb; tk = t0; }
          if (ti < ex && tj < ey && tk < ez) {
            const int q = rb_box(ti, tj, tk);
            const int sq = shp->rslot[ti + TX * (tj + TX * tk)];
            const T c = (MODE == RB_BLACK_FB) ? T(CH[sq].c[s]) : fold_c[s];
            const C2<T> v = fold_v[s];
            C2<T> r = W[q];
            if (MODE == RB_BLACK_FB) {
              r.x -= c * v.x;
              r.y -= c * v.y;
            } else {
              const C2<T> d = cmul(DV[sq], C2<T>{c * v.x, c * v.y});
              r.x -= d.x;
              r.y -= d.y;
            }
            W[q] = r;
          }
          wave_sync();
        }
      }
      const bool full = ex == TX && ey == TX && ez == TX;
      auto live = [&](unsigned int pt) {
        return pt != 0xffffffffu &&
               (full || (int((pt >> 19) & 7) < ex && int((pt >> 22) & 7) < ey && int((pt >> 25) & 7) < ez));
      };
      if (FWDP) {
#pragma unroll
        for (int h = 0; h < NLEV; ++h) {
          const unsigned int pt = pk[h];
          if (live(pt)) {
            const int q = pt & 0x3ff, p = (pt >> 10) & 0x1ff;
            const C2<T> dvp = DV[p];
            const Cpl c4 = CL[p];
            C2<T> y = W[q];
            C2<T> v[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) v[s] = W[q - st[s]];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
              const T c = T(c4.c[s]);
              y.x -= c * v[s].x;
              y.y -= c * v[s].y;
            }
            W[q] = cmul(dvp, y);
          }
          wave_sync();
        }
      }
      if (MODE == RB_BLACK_FB) {
        W[rb_box(0, fa, fb) - 1] = {T(0), T(0)};
        W[rb_box(fa, 0, fb) - kRbBX] = {T(0), T(0)};
        W[rb_box(fa, fb, 0) - kRbPK] = {T(0), T(0)};
        wave_sync();
        // nothing the forward sweep derived from the table is kept for the
        // backward one (registers for the prefetched item)
#pragma unroll
        for (int h = 0; h < NLEV; ++h) asm volatile("" : "+v"(pk[h]));
      }
      if (BWDP) {
#pragma unroll
        for (int h = NLEV - 1; h >= 0; --h) {
          const unsigned int pt = pk[h];
          if (live(pt)) {
            const int q = pt & 0x3ff, p = (pt >> 10) & 0x1ff;
            const C2<T> dvp = DV[p];
            const Cpl h4 = CH[p];
            const C2<T> b = W[q];
            C2<T> z[3];
#pragma unroll
            for (int s = 0; s < 3; ++s) {
              const unsigned int qn = unsigned(q + st[s]);
              z[s] = W[s == 0 ? qn : min(qn, qn - unsigned(NB))];
            }
            C2<T> y = {T(0), T(0)};
#pragma unroll
            for (int s = 0; s < 3; ++s) {
              const T c = T(h4.c[s]);
              y.x -= c * z[s].x;
              y.y -= c * z[s].y;
            }
            const C2<T> d = cmul(dvp, y);
            W[q] = {b.x + d.x, b.y + d.y};
          }
          wave_sync();
        }
      }
      // ---- write the item back (red forward: w; otherwise z) and the
      // packed i-faces the other colour reads next
      constexpr bool zp = ZP && MODE != RB_RED_FWD;
      constexpr int zs = zp ? 2 : 1;
      C2<T>* __restrict__ dst =
          zp ? a.zout + (int64_t)(cu.sys >> 1) * a.zpair + (cu.sys & 1) + 2 * (int64_t)cu.tself * kBkPts
             : a.out + o + (int64_t)cu.tself * kBkPts;
      C2<T>* __restrict__ xfo = a.xf + (MODE == RB_RED_FWD ? 0 : a.xfhalf) +
                                (int64_t)cu.sys * a.xfsys + (int64_t)cu.tself * kRbFaces;
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int p = lane + 64 * u;
        if (!(fa < ex && fb < ey && u < ez)) continue;
        const C2<T> zv = W[rb_box(fa, fb, u)];
        dst[zs * p] = zv;
        if (MODE != RB_RED_BWD && (fa == 0 || fa == TX - 1)) xfo[(fa == 0 ? 0 : 64) + fb + 8 * u] = zv;
      }
    }
    if (!more) break;
    w = wn;
  }
}

}  // namespace mt
