// Block-DILU preconditioner apply on gfx950: level-scheduled triangular
// solves on 3-D tiles.
//
// The reference preconditioner (PC_setup, modelOperator3D.f90:512-611) keeps
// only the x-x, y-y and z-z diagonal blocks of A.  Each block couples an edge
// to its 6 lattice neighbours of the same component, so in the reference's
// lexicographic order (i fastest) the lower neighbours of (i,j,k) are i-1, j-1
// and k-1, and L = I + E P^-1, U = P + F (Dilu_Cmplx_AS, spOpTools.f90:1996-
// 2055) are solved by wavefronts over the hyperplanes i+j+k = const.
//
// Instead of one launch per hyperplane (266 at 100x100x60+10), the lattice is
// cut into TX*TY*TZ tiles.  One launch handles every tile of one tile-level
// ti+tj+tk = L (all three blocks and all active systems at once); inside a
// tile one wave64 sweeps the tile's own hyperplanes out of LDS, with the
// upstream faces (final after the previous launch) staged into an LDS halo.
// Launch boundaries are the only inter-tile synchronisation, so nothing ever
// spins on a flag.
//   forward  y = L^-1 r :  y_e = r_e - sum_{l lower} a_el dinv_l y_l
//   backward z = U^-1 y :  z_e = dinv_e (y_e - sum_{u upper} a_eu z_u)
// with dinv = 1/pivot computed per frequency.  The backward sweep runs in
// place on the forward result.
#pragma once

#include "mt_kernels.hpp"

namespace mt {

constexpr int kMaxTilePts = 1024;
struct TileShape {
  int tx, ty, tz;
  int npts;
  int nlev;
  int lvl[64];                  // level start offsets into pts, nlev+1 entries
  unsigned int pts[kMaxTilePts];  // packed (li | lj<<8 | lk<<16) sorted by li+lj+lk
};

struct LatticeInfo {
  int A, B, C;        // interior lattice dims
  int oi, oj, ok;     // offset of the interior lattice in its full lattice
  int a, b;           // full-lattice dims (i, j) for addressing
  int64_t base;       // address of the full lattice's first point in the vector
};

template <class T, class V>
struct DiluArgs {
  const V* coef;           // [7 or 15][cstride]: slots 1..6 = a-1, a+1, b-1, b+1, c-1, c+1
  int64_t cstride;         // coefficient slot stride (= per-frequency pivot stride)
  int64_t vstride;         // per-system vector stride
  LatticeInfo lat[3];      // the independent blocks (3 edge components, or 1 node lattice)
  const int4* tiles;       // tiles of this level: {lattice, ti, tj, tk}
  int ntiles;
  const int* act;
  int nact;
  const SysScal* sc;
  const C2<T>* in;         // forward: right-hand side
  C2<T>* out;              // forward: y; backward: in place y -> z
  const C2<T>* dinv;       // [nfreq][cstride]
  int status_req;          // only systems with this status are processed
};

__device__ __forceinline__ int64_t lat_e(const LatticeInfo& L, int i, int j, int k) {
  return L.base + (i + L.oi) + int64_t(L.a) * ((j + L.oj) + int64_t(L.b) * (k + L.ok));
}

// FWD=true: lower solve with halo on the low side; false: upper solve.
template <class T, class V, int TX, int TY, int TZ, bool FWD>
__global__ __launch_bounds__(64) void k_dilu_tile(DiluArgs<T, V> a, const TileShape* shp) {
  constexpr int BX = TX + 1, BY = TY + 1, BZ = TZ + 1, NB = BX * BY * BZ, NP = TX * TY * TZ;
  __shared__ C2<T> yb[NB];
  __shared__ C2<T> dv[FWD ? NB : NP];
  __shared__ V cf[3][NP];
  __shared__ unsigned int spts[NP];
  __shared__ int slvl[TX + TY + TZ];
  const int tix = blockIdx.x / a.nact, si = blockIdx.x - tix * a.nact;
  const int sys = a.act[si];
  const SysScal& sc = a.sc[sys];
  if (sc.status != a.status_req) return;
  const int4 tl = a.tiles[tix];
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TY, gk0 = tl.w * TZ;
  const int ex = min(TX, L.A - gi0), ey = min(TY, L.B - gj0), ez = min(TZ, L.C - gk0);
  const int64_t o = (int64_t)sys * a.vstride;
  const C2<T>* dinv = a.dinv + (int64_t)sc.fidx * a.cstride;
  const int s0 = FWD ? 1 : 2;  // lower slots 1,3,5 / upper slots 2,4,6
  // box coordinate b <-> local tile coordinate l: FWD: l = b - 1, BWD: l = b
  for (int q = threadIdx.x; q < NB; q += 64) {
    const int bi = q % BX, bj = (q / BX) % BY, bk = q / (BX * BY);
    const int li = FWD ? bi - 1 : bi, lj = FWD ? bj - 1 : bj, lk = FWD ? bk - 1 : bk;
    const bool inTile = li >= 0 && lj >= 0 && lk >= 0 && li < TX && lj < TY && lk < TZ;
    const int gi = gi0 + li, gj = gj0 + lj, gk = gk0 + lk;
    const bool inLat = gi >= 0 && gj >= 0 && gk >= 0 && gi < L.A && gj < L.B && gk < L.C;
    C2<T> yv = {T(0), T(0)}, dd = {T(0), T(0)};
    if (inLat) {
      const int64_t e = lat_e(L, gi, gj, gk);
      yv = inTile ? (FWD ? a.in[o + e] : a.out[o + e]) : a.out[o + e];
      dd = dinv[e];
      if (inTile) {
        const int p = li + TX * (lj + TY * lk);
#pragma unroll
        for (int s = 0; s < 3; ++s) cf[s][p] = a.coef[(int64_t)(s0 + 2 * s) * a.cstride + e];
        if (!FWD) dv[p] = dd;
      }
    } else if (inTile) {
      const int p = li + TX * (lj + TY * lk);
#pragma unroll
      for (int s = 0; s < 3; ++s) cf[s][p] = V(0);
      if (!FWD) dv[p] = dd;
    }
    yb[q] = yv;
    if (FWD) dv[q] = dd;
  }
  for (int p = threadIdx.x; p < NP; p += 64) spts[p] = shp->pts[p];
  for (int p = threadIdx.x; p < TX + TY + TZ - 1; p += 64) slvl[p] = shp->lvl[p];
  __syncthreads();
  const int nlev = TX + TY + TZ - 2;
  for (int hh = 0; hh < nlev; ++hh) {
    const int h = FWD ? hh : nlev - 1 - hh;
    for (int pp = slvl[h] + threadIdx.x; pp < slvl[h + 1]; pp += 64) {
      const unsigned int pk = spts[pp];
      const int li = pk & 0xff, lj = (pk >> 8) & 0xff, lk = pk >> 16;
      if (li >= ex || lj >= ey || lk >= ez) continue;
      const int p = li + TX * (lj + TY * lk);
      if (FWD) {
        const int q = (li + 1) + BX * ((lj + 1) + BY * (lk + 1));
        C2<T> y = yb[q];
        const int qn[3] = {q - 1, q - BX, q - BX * BY};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = T(cf[s][p]);
          const C2<T> w = cmul(dv[qn[s]], yb[qn[s]]);
          y.x -= c * w.x;
          y.y -= c * w.y;
        }
        yb[q] = y;
      } else {
        const int q = li + BX * (lj + BY * lk);
        C2<T> y = yb[q];
        const int qn[3] = {q + 1, q + BX, q + BX * BY};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = T(cf[s][p]);
          const C2<T> z = yb[qn[s]];
          y.x -= c * z.x;
          y.y -= c * z.y;
        }
        yb[q] = cmul(dv[p], y);
      }
    }
    __syncthreads();
  }
  for (int p = threadIdx.x; p < NP; p += 64) {
    const int li = p % TX, lj = (p / TX) % TY, lk = p / (TX * TY);
    if (li >= ex || lj >= ey || lk >= ez) continue;
    const int q = FWD ? (li + 1) + BX * ((lj + 1) + BY * (lk + 1)) : li + BX * (lj + BY * lk);
    a.out[o + lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = yb[q];
  }
}

}  // namespace mt
