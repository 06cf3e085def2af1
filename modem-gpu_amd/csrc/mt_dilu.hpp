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

// The tile kernels run one wave64 per block.  A wave's LDS instructions are
// executed in order, so a read issued after a write sees it whichever lane
// wrote it: between the hyperplanes of a sweep only the compiler must be kept
// from reordering, no s_waitcnt for the writes is needed (a block barrier
// would wait for them to complete on every hyperplane).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kMaxTilePts = 1024;
// LDS box of the lean red-black sweeps (k_dilu_rbl): the 8x8x8 tile and one
// pad layer on the low side, i fastest, row pitch 9 and plane pitch 82.  With
// a plane pitch of 81 the points of a hyperplane i+j+k = h sit 8*dj + 80*dk
// apart, which is 0 or 8 modulo 16: every 8-byte access of a hyperplane step
// fell into two of the sixteen bank pairs.  Pitch 82 spreads them over all of
// them (one spare slot per plane).  An upper neighbour past the end of a row,
// plane or box lands in a low pad slot (index mod kRbNB), disjoint from the
// tile and from the other high faces.
constexpr int kRbBX = 9, kRbPK = 82, kRbNB = 9 * kRbPK, kRbNP = 512;
// packed i-faces per tile and system in DiluArgs::xf: two faces of 64
constexpr int kRbFaces = 128;
__host__ __device__ __forceinline__ int rb_box(int li, int lj, int lk) {
  return (li + 1) + kRbBX * (lj + 1) + kRbPK * (lk + 1);
}
struct TileShape {
  int tx, ty, tz;
  int npts;
  int nlev;
  int lvl[64];                  // level start offsets into pts, nlev+1 entries
  unsigned int pts[kMaxTilePts];  // packed (li | lj<<8 | lk<<16) sorted by li+lj+lk
  // 8x8x8 tiles only (k_dilu_rbl): the same levels with the points of a level
  // ordered so that lanes of a 16-lane group hit distinct box banks, packed
  // li | lj<<3 | lk<<6 | slot<<9 where slot is the position in this list, and
  // slot of tile point li + 8*(lj + 8*lk).  The per-point LDS data (pivots,
  // couplings) are stored by slot, so a level's lanes read consecutive words.
  unsigned int rpts[kRbNP];
  unsigned short rslot[kRbNP];
  // rpts by (hyperplane, lane): lane l's point of hyperplane h, or ~0u
  unsigned int lpts[22][64];
};

struct LatticeInfo {
  int A, B, C;        // interior lattice dims
  int oi, oj, ok;     // offset of the interior lattice in its full lattice
  int a, b;           // full-lattice dims (i, j) for addressing
  int64_t base;       // address of the full lattice's first point in the vector
  // tiles of a tile plan over this lattice: ntx x nty x ntz, numbered
  // tbase + ti + ntx*(tj + nty*tk) across all lattices of the plan
  int ntx, nty, tbase;
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
  const C2<T>* in2;        // forward, when set: right-hand side in - alpha*in2
  // brick red-black sweeps with in2: in - alpha*in2 of every point also stored
  // here (the BiCGStab s, which the SpMV's <t,s> and the r update then read)
  C2<T>* sout = nullptr;
  // brick red-black sweeps, when set: the result z (black and red backward
  // launches) goes to zout with the two systems of a pair interleaved, entry
  // (sys/2)*zpair + 2*point + sys%2 (zpair = twice the points per system);
  // the red forward sweep's w stays in out, and the red backward sweep reads
  // the black tiles' j- and k-faces from zout
  C2<T>* zout = nullptr;
  int64_t zpair = 0;
  // persistent red forward sweep of the first application, when pu_r is set:
  // its right-hand side is the BiCGStab p = r + beta (p - omega v) (on a
  // restart p = r and rt = r), formed from r (pu_r), the previous p (in) and
  // v (pu_v), and stored (pu_p, pu_rt) for the red tiles (the black tiles'
  // by k_pupdate_tiles before the launch)
  const C2<T>* pu_r = nullptr;
  const C2<T>* pu_v = nullptr;
  C2<T>* pu_p = nullptr;
  C2<T>* pu_rt = nullptr;
  C2<T>* out;              // forward: y; backward: in place y -> z
  const C2<T>* dinv;       // [nfreq][cstride]
  int status_req;          // only systems with this status are processed
  // red-black sweeps: the i-faces (li = 0 and li = 7) of every tile, packed
  // [half][system][tile][face][64], entry lj + 8*lk (kRbFaces per tile).
  // Half 0 holds dinv*y of the red tiles (written by the red forward sweep,
  // read by the black one), half 1 holds z of the black tiles (written by the
  // black sweep, read by the red backward one).  Packed, a neighbour face is
  // one coalesced read instead of 64 reads of 8 bytes, one per cache line.
  C2<T>* xf = nullptr;
  int64_t xfsys = 0, xfhalf = 0;
};

__device__ __forceinline__ int64_t lat_e(const LatticeInfo& L, int i, int j, int k) {
  return L.base + (i + L.oi) + int64_t(L.a) * ((j + L.oj) + int64_t(L.b) * (k + L.ok));
}

// ------------------------------------------ reference ordering, 8x8x8 tiles
// One tile level per launch (grid = tiles of the level x active systems), one
// wave per (tile, system): the tile's inputs and the upstream neighbours'
// facing layer (final after the previous launch) are staged into LDS, the 22
// hyperplanes swept, the tile written back.  A launch holds a few dozen to a
// few hundred tiles of one system, so a tile's time is the launch's time, and
// it is latency: the chain of dependent loads before the data arrive, the
// hyperplane steps, the stores.
//  * Loads: the system-independent tables (tile, lane's point of every
//    hyperplane, level slots) and the system lookups are requested together,
//    then every data load at once; the system's status is checked after (a
//    skipped system's loads are wasted, never out of bounds).
//  * A hyperplane step is one round of LDS reads: the lane's points are in
//    registers (box index and level slot, extent folded in), the pivots and
//    couplings sit by level slot (consecutive lanes, consecutive entries), and
//    the box is rb_box's pitch-82 layout, which spreads a hyperplane over the
//    LDS banks.
//   forward: the box holds r on the tile and w = dinv*y of the low faces;
//            a step forms y = r - sum a w, keeps y in the pivot's slot and
//            leaves w = dinv*y in the box; y is written back.
//   backward: the box holds y on the tile and z of the high faces (past the
//             box's last plane, kRbNB + kRbPK entries in all); a step forms
//             z = dinv*(y - sum a z) in place; z is written back.
// Per point the three terms are taken in the order i, j, k.
template <class T, class V, bool FWD>
__global__ __launch_bounds__(64) void k_dilu_ro(DiluArgs<T, V> a, const TileShape* __restrict__ shp) {
  constexpr int TX = 8, NP = kRbNP, NU = 8, NLEV = 22;
  constexpr int NB = FWD ? kRbNB : kRbNB + kRbPK;
  struct Cpl { V c[3]; };
  __shared__ C2<T> W[NB];
  __shared__ C2<T> DV[NP];   // pivots by level slot (forward: then y)
  __shared__ Cpl CF[NP];     // the sweep's couplings by level slot
  const int lane = threadIdx.x;
  const int tix = blockIdx.x / a.nact, si = blockIdx.x - tix * a.nact;
  // ---- system-independent tables and the system lookups, together
  const int4 tl = a.tiles[tix];
  const int sys = a.act[si];
  int sl[NU];
#pragma unroll
  for (int u = 0; u < NU; ++u) sl[u] = shp->rslot[lane + 64 * u];
  unsigned int pv[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) pv[h] = shp->lpts[h][lane];
  const SysScal& sc = a.sc[sys];
  const int status = sc.status, fidx = sc.fidx;
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
  const int ex = min(TX, L.A - gi0), ey = min(TX, L.B - gj0), ez = min(TX, L.C - gk0);
  const int64_t o = (int64_t)sys * a.vstride;
  const int64_t cst = a.cstride;
  const C2<T>* __restrict__ pout = a.out + o;
  const C2<T>* __restrict__ dinv = a.dinv + (int64_t)fidx * cst;
  const V* __restrict__ cf0 = a.coef + (FWD ? 1 : 2) * cst;   // slots 1,3,5 / 2,4,6
  const int st[3] = {1, kRbBX, kRbPK};
  // ---- the tile's own data (p = lane + 64u: li = lane&7, lj = lane>>3,
  // lk = u): forward r (- alpha*in2), backward y; pivots; couplings
  const int li0 = lane & 7, lj0 = lane >> 3;
  const bool col_ok = li0 < ex && lj0 < ey;
  const int64_t e0 = lat_e(L, gi0 + li0, gj0 + lj0, gk0);
  const int64_t kst = int64_t(L.a) * L.b;
  C2<T> rv[NU], dd[NU], r2[NU];
  V cc[NU][3];
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    rv[u] = dd[u] = r2[u] = {T(0), T(0)};
    cc[u][0] = cc[u][1] = cc[u][2] = V(0);
    if (col_ok && u < ez) {
      const int64_t e = e0 + u * kst;
      if (FWD) {
        rv[u] = a.in[o + e];
        if (a.in2) r2[u] = a.in2[o + e];
      } else {
        rv[u] = pout[e];   // in place
      }
      dd[u] = dinv[e];
#pragma unroll
      for (int s = 0; s < 3; ++s) cc[u][s] = cf0[(int64_t)(2 * s) * cst + e];
    }
  }
  // the facing layer of the upstream neighbours (forward: low faces, y and
  // pivot; backward: high faces, z)
  int fq[3];
  C2<T> fv[3], fd[3];
#pragma unroll
  for (int dir = 0; dir < 3; ++dir) {
    const int t0 = FWD ? 0 : TX - 1;
    const int ti = dir == 0 ? t0 : li0;
    const int tj = dir == 0 ? li0 : (dir == 1 ? t0 : lj0);
    const int tk = dir == 2 ? t0 : lj0;
    const int d = FWD ? -1 : 1;
    const int ni = gi0 + ti + (dir == 0 ? d : 0);
    const int nj = gj0 + tj + (dir == 1 ? d : 0);
    const int nk = gk0 + tk + (dir == 2 ? d : 0);
    fv[dir] = fd[dir] = {T(0), T(0)};
    fq[dir] = -1;
    const bool ok = ti < ex && tj < ey && tk < ez && ni >= 0 && nj >= 0 && nk >= 0 &&
                    ni < L.A && nj < L.B && nk < L.C;
    if (ok) {
      const int64_t l = lat_e(L, ni, nj, nk);
      fv[dir] = pout[l];
      if (FWD) fd[dir] = dinv[l];
      fq[dir] = FWD ? rb_box(ti, tj, tk) - st[dir] : rb_box(ti, tj, tk) + st[dir];
    }
  }
  if (status != a.status_req) return;   // uniform
  // this lane's point on every hyperplane: box index | level slot << 10
  unsigned int pk[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) {
    const unsigned int v = pv[h];
    const int li = v & 7, lj = (v >> 3) & 7, lk = (v >> 6) & 7;
    const bool ok = v != 0xffffffffu && li < ex && lj < ey && lk < ez;
    pk[h] = ok ? (unsigned(rb_box(li, lj, lk)) | ((v >> 9) << 10)) : 0xffffffffu;
  }
  // ---- stage: zero box (partial tiles and absent faces read zeros), then
  // the tile, the faces, the per-slot data
  for (int q = lane; q < NB; q += 64) W[q] = {T(0), T(0)};
  wave_sync();
  const C2<T> al = ccast<T>(sc.alpha);
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    C2<T> r = rv[u];
    if (FWD && a.in2) r = csub(r, cmul(al, r2[u]));
    W[rb_box(li0, lj0, u)] = r;
    DV[sl[u]] = dd[u];
    Cpl c;
#pragma unroll
    for (int s = 0; s < 3; ++s) c.c[s] = cc[u][s];
    CF[sl[u]] = c;
  }
#pragma unroll
  for (int dir = 0; dir < 3; ++dir)
    if (fq[dir] >= 0) W[fq[dir]] = FWD ? cmul(fd[dir], fv[dir]) : fv[dir];
  wave_sync();
  if (FWD) {
#pragma unroll
    for (int h = 0; h < NLEV; ++h) {
      const unsigned int pt = pk[h];
      if (pt != 0xffffffffu) {
        const int q = pt & 0x3ff, p = pt >> 10;
        const C2<T> dvp = DV[p];
        const Cpl c = CF[p];
        C2<T> y = W[q];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T cs = T(c.c[s]);
          const C2<T> w = W[q - st[s]];
          y.x -= cs * w.x;
          y.y -= cs * w.y;
        }
        DV[p] = y;
        W[q] = cmul(dvp, y);
      }
      wave_sync();
    }
  } else {
#pragma unroll
    for (int h = NLEV - 1; h >= 0; --h) {
      const unsigned int pt = pk[h];
      if (pt != 0xffffffffu) {
        const int q = pt & 0x3ff, p = pt >> 10;
        const C2<T> dvp = DV[p];
        const Cpl c = CF[p];
        C2<T> y = W[q];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T cs = T(c.c[s]);
          const C2<T> z = W[q + st[s]];
          y.x -= cs * z.x;
          y.y -= cs * z.y;
        }
        W[q] = cmul(dvp, y);
      }
      wave_sync();
    }
  }
  // ---- write the tile back: forward y, backward z
  C2<T>* __restrict__ dst = a.out + o;
#pragma unroll
  for (int u = 0; u < NU; ++u) {
    if (!(col_ok && u < ez)) continue;
    dst[e0 + u * kst] = FWD ? DV[sl[u]] : W[rb_box(li0, lj0, u)];
  }
}

// ------------------------------------------------- red-black tile ordering
// Block DILU in a second ordering, for the inner FP32 solves of the mixed
// solver (whose FP64 defect correction makes any fixed preconditioner
// admissible): the tiles of each lattice are coloured red/black by the parity
// of ti+tj+tk, all red tiles come first, then all black tiles, and inside a
// tile the points keep the lexicographic order.  Red tiles only couple to
// black tiles, so each colour is one launch per sweep instead of one per
// tile level.  With this ordering
//   lower(e) = in-tile lower neighbours, plus (black e) its face neighbours in
//              the adjacent red tiles;
//   upper(e) = in-tile upper neighbours, plus (red e) its face neighbours in
//              the adjacent black tiles,
// and the pivots, the forward and the backward sweep are those of DILU with
// these sets.  The colour whose cross-face neighbours are already final
// (black in the forward sweep and in the pivots, red in the backward sweep)
// takes their contributions into the tile's right-hand side (or diagonal)
// before it sweeps the tile like an isolated one.
//
// ------------------------------------------- red-black sweeps, persistent
// The red-black application as three launches over 8x8x8 tiles:
//   RB_RED_FWD   red tiles, forward:  y = L^-1 r on each tile alone;
//   RB_BLACK_FB  black tiles: forward with the red face neighbours (all six
//                faces are lower for a black point) and, since a black
//                point's upper neighbours all lie in its own tile, the
//                backward sweep of the same tile right after, in LDS;
//   RB_RED_BWD   red tiles, backward with the black face neighbours (upper).
// Work items are (tile of the colour, active system) pairs, systems fastest.
// The items are cut into 8 contiguous ranges, one per XCD (workgroups are
// dispatched round-robin over the XCDs), and the blocks of an XCD walk its
// range, so concurrent blocks of one XCD work on the same tiles for different
// systems and share the tile's couplings in that XCD's L2.  A block is
// persistent: while it sweeps one item out of LDS, the loads of its next item
// are in flight.  The box holds the tile and a one-point halo on all six
// faces (box stride 1, 10, 100).
enum : int { RB_RED_FWD = 0, RB_BLACK_FB = 1, RB_RED_BWD = 2 };

template <class T, class V, int MODE>
__global__ __launch_bounds__(64) void k_dilu_rb2(DiluArgs<T, V> a, const TileShape* shp) {
  constexpr int TX = 8, NP = 512, NU = 8, NLEV = 22, BX = 10, NB = 1000;
  constexpr bool FWDP = MODE != RB_RED_BWD;   // has a forward sweep
  constexpr bool BWDP = MODE != RB_RED_FWD;   // has a backward sweep
  constexpr bool HALO = MODE != RB_RED_FWD;   // reads face neighbours
  __shared__ C2<T> W[NB];                     // box: w = dinv*y (forward), y -> z (backward)
  // per tile point, packed so that a sweep step reads each with one LDS
  // instruction: rhs -> y with the pivot, and the couplings by four
  struct alignas(16) PtA { C2<T> rb, dv; };
  struct alignas(16) Cpl { V c[4]; };
  __shared__ PtA PD[FWDP ? NP : 1];           // forward sweeps: rhs -> y, pivot
  __shared__ C2<T> DV[FWDP ? 1 : NP];         // red backward: pivot
  __shared__ Cpl CL[FWDP ? NP : 1];           // lower couplings used by the forward sweep
  __shared__ Cpl CH[BWDP ? NP : 1];           // upper couplings used by the backward sweep
  __shared__ int s_sys[64], s_ok[64], s_fidx[64];
  __shared__ C2<T> s_al[64];
  const int lane = threadIdx.x;
  const int nwork = a.ntiles * a.nact;
  // the active systems (at most 64): index, status, frequency slot, alpha
  if (lane < a.nact) {
    const int sy = a.act[lane];
    const SysScal& sc = a.sc[sy];
    s_sys[lane] = sy;
    s_ok[lane] = sc.status == a.status_req;
    s_fidx[lane] = sc.fidx;
    s_al[lane] = ccast<T>(sc.alpha);
  }
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3, gx = gridDim.x >> 3;
  const int chunk = (nwork + 7) / 8;
  const int w0 = xcd * chunk + lb, wend = min(nwork, (xcd + 1) * chunk);
  // this lane's point on every hyperplane of the tile (at most 48 per plane)
  unsigned int pk[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) {
    const int b = shp->lvl[h], e = shp->lvl[h + 1];
    pk[h] = (b + lane < e) ? shp->pts[b + lane] : 0xffffffffu;
  }
  for (int q = lane; q < NB; q += 64) W[q] = {T(0), T(0)};
  wave_sync();
  const V* __restrict__ cfs = a.coef;
  const int64_t cst = a.cstride;
  // ---- prefetch registers (one item)
  C2<T> rv[NU], v2[NU], dd[NU], fv[6], fd[6];
  V cl[NU][3], ch[NU][3];
  auto item_ok = [&](int w) { return w < wend && s_ok[w % a.nact]; };
  int4 tl_next = make_int4(0, 0, 0, 0);
  auto issue = [&](int w) {
    const int tix = w / a.nact, si = w - tix * a.nact, sys = s_sys[si];
    const int4 tl = a.tiles[tix];
    tl_next = tl;
    const LatticeInfo L = a.lat[tl.x];
    const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
    const int ex = min(TX, L.A - gi0), ey = min(TX, L.B - gj0), ez = min(TX, L.C - gk0);
    const int64_t o = (int64_t)sys * a.vstride;
    const C2<T>* __restrict__ pin = a.in + o;
    const C2<T>* __restrict__ pin2 = a.in2 ? a.in2 + o : nullptr;
    const C2<T>* __restrict__ pout = a.out + o;
    const C2<T>* __restrict__ dinv = a.dinv + (int64_t)s_fidx[si] * cst;
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      rv[u] = v2[u] = dd[u] = {T(0), T(0)};
#pragma unroll
      for (int s = 0; s < 3; ++s) cl[u][s] = ch[u][s] = V(0);
      if (li < ex && lj < ey && lk < ez) {
        const int64_t e = lat_e(L, gi0 + li, gj0 + lj, gk0 + lk);
        rv[u] = FWDP ? pin[e] : pout[e];
        if (FWDP && pin2) v2[u] = pin2[e];
        dd[u] = dinv[e];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          if (FWDP || MODE == RB_RED_BWD) cl[u][s] = cfs[(int64_t)(1 + 2 * s) * cst + e];
          if (BWDP) ch[u][s] = cfs[(int64_t)(2 + 2 * s) * cst + e];
        }
      }
    }
    if (HALO) {
      // face f: lane (fa, fb) of the 8x8 face; the tile point next to the
      // face must be in the lattice, and its neighbour across the face too
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
        fv[f] = fd[f] = {T(0), T(0)};
        const bool ok = ti < ex && tj < ey && tk < ez && ni >= 0 && nj >= 0 && nk >= 0 &&
                        ni < L.A && nj < L.B && nk < L.C;
        if (ok) {
          if (dir == 0) {
            // the neighbour tile's facing i-face, packed (entry tj + 8*tk = lane)
            const int tid = L.tbase + (tl.y + (hi ? 1 : -1)) + L.ntx * (tl.z + L.nty * tl.w);
            fv[f] = a.xf[(MODE == RB_BLACK_FB ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
                         (int64_t)tid * 128 + (hi ? 0 : 64) + lane];
          } else {
            const int64_t l = lat_e(L, ni, nj, nk);
            fv[f] = pout[l];
            if (MODE == RB_BLACK_FB) fd[f] = dinv[l];
          }
        }
      }
    }
  };
  int w = w0;
  while (w < wend && !item_ok(w)) w += gx;
  if (w < wend) issue(w);
  for (; w < wend;) {
    const int tix = w / a.nact, si = w - tix * a.nact, sys = s_sys[si];
    const int4 tl = tl_next;
    const LatticeInfo L = a.lat[tl.x];
    const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
    const int ex = min(TX, L.A - gi0), ey = min(TX, L.B - gj0), ez = min(TX, L.C - gk0);
    const int64_t o = (int64_t)sys * a.vstride;
    const C2<T> al = s_al[si];
    // ---- stage the item into LDS
    if (HALO) {
      const int fa = lane & 7, fb = lane >> 3;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1;
        const bool hi = f & 1;
        int q;
        if (dir == 0) q = (hi ? 9 : 0) + BX * ((fa + 1) + BX * (fb + 1));
        else if (dir == 1) q = (fa + 1) + BX * ((hi ? 9 : 0) + BX * (fb + 1));
        else q = (fa + 1) + BX * ((fb + 1) + BX * (hi ? 9 : 0));
        // black: red neighbours enter as dinv*y (the packed i-faces already are)
        W[q] = (MODE == RB_BLACK_FB && dir != 0) ? cmul(fd[f], fv[f]) : fv[f];
      }
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
      const bool in0 = li > 0, in1 = lj > 0, in2 = lk > 0;                      // lower in tile
      const bool up0 = li + 1 < ex, up1 = lj + 1 < ey, up2 = lk + 1 < ez;      // upper in tile
      C2<T> r = rv[u];
      if (FWDP && a.in2) r = csub(r, cmul(al, v2[u]));
      if (FWDP) {
        PD[p] = PtA{r, dd[u]};
        // red: face neighbours are upper, so only in-tile lower couplings;
        // black: every lower-side coupling (in tile or across the face)
        Cpl c;
        c.c[0] = (MODE == RB_BLACK_FB || in0) ? cl[u][0] : V(0);
        c.c[1] = (MODE == RB_BLACK_FB || in1) ? cl[u][1] : V(0);
        c.c[2] = (MODE == RB_BLACK_FB || in2) ? cl[u][2] : V(0);
        c.c[3] = V(0);
        CL[p] = c;
      } else {
        DV[p] = dd[u];
        W[q] = r;
      }
      if (BWDP) {
        // black: upper neighbours in the tile only; red: in tile or across
        Cpl c;
        c.c[0] = (MODE == RB_RED_BWD || up0) ? ch[u][0] : V(0);
        c.c[1] = (MODE == RB_RED_BWD || up1) ? ch[u][1] : V(0);
        c.c[2] = (MODE == RB_RED_BWD || up2) ? ch[u][2] : V(0);
        c.c[3] = V(0);
        CH[p] = c;
      }
    }
    wave_sync();
    // fold the face terms the sweeps do not reach: black forward, the high
    // faces (red neighbours above); red backward, the low faces (black below)
    if (HALO) {
#pragma unroll
      for (int u = 0; u < NU; ++u) {
        const int p = lane + 64 * u;
        const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
        if (!(li < ex && lj < ey && lk < ez)) continue;
        const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
        if (MODE == RB_BLACK_FB) {
          C2<T> r = PD[p].rb;
          const bool f0 = li == TX - 1, f1 = lj == TX - 1, f2 = lk == TX - 1;
          const int st[3] = {1, BX, BX * BX};
          const bool fl[3] = {f0, f1, f2};
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            if (!fl[s]) continue;
            const T c = T(ch[u][s]);
            const C2<T> v = W[q + st[s]];
            r.x -= c * v.x;
            r.y -= c * v.y;
          }
          PD[p].rb = r;
        } else {
          C2<T> y = W[q];
          const int st[3] = {1, BX, BX * BX};
          const bool fl[3] = {li == 0, lj == 0, lk == 0};
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            if (!fl[s]) continue;
            const T c = T(cl[u][s]);
            const C2<T> v = W[q - st[s]];
            y.x -= c * v.x;
            y.y -= c * v.y;
          }
          W[q] = y;
        }
      }
      wave_sync();
    }
    // ---- the next item's loads go out now and land during the sweeps
    int wn = w + gx;
    while (wn < wend && !item_ok(wn)) wn += gx;
    if (wn < wend) issue(wn);
    // ---- sweeps
    if (FWDP) {
#pragma unroll
      for (int h = 0; h < NLEV; ++h) {
        const unsigned int pt = pk[h];
        const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = (pt >> 16) & 0xff;
        if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
          const int p = li + TX * (lj + TX * lk);
          const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
          const PtA pt = PD[p];
          const Cpl cl4 = CL[p];
          C2<T> y = pt.rb;
          const int qn[3] = {q - 1, q - BX, q - BX * BX};
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            const T c = T(cl4.c[s]);
            const C2<T> v = W[qn[s]];
            y.x -= c * v.x;
            y.y -= c * v.y;
          }
          PD[p].rb = y;
          W[q] = cmul(pt.dv, y);
        }
        wave_sync();
      }
    }
    if (BWDP) {
#pragma unroll
      for (int h = NLEV - 1; h >= 0; --h) {
        const unsigned int pt = pk[h];
        const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = (pt >> 16) & 0xff;
        if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
          const int p = li + TX * (lj + TX * lk);
          const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
          C2<T> y, dvp;
          if constexpr (FWDP) {
            const PtA pt = PD[p];
            y = pt.rb;
            dvp = pt.dv;
          } else {
            y = W[q];
            dvp = DV[p];
          }
          const Cpl ch4 = CH[p];
          const int qn[3] = {q + 1, q + BX, q + BX * BX};
#pragma unroll
          for (int s = 0; s < 3; ++s) {
            const T c = T(ch4.c[s]);
            const C2<T> z = W[qn[s]];
            y.x -= c * z.x;
            y.y -= c * z.y;
          }
          W[q] = cmul(dvp, y);
        }
        wave_sync();
      }
    }
    // ---- write the tile back (red forward: y; otherwise z), and the packed
    // i-faces the other colour reads next (red forward: dinv*y; black: z)
    C2<T>* __restrict__ dst = a.out + o;
    C2<T>* __restrict__ xfo =
        a.xf + (MODE == RB_RED_FWD ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
        (int64_t)(L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w)) * 128;
    for (int p = lane; p < NP; p += 64) {
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      if (!(li < ex && lj < ey && lk < ez)) continue;
      const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
      C2<T> zv;
      if constexpr (MODE == RB_RED_FWD) zv = PD[p].rb;
      else zv = W[q];
      dst[lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = zv;
      if (MODE != RB_RED_BWD && (li == 0 || li == TX - 1)) {
        C2<T> fvv = zv;
        if constexpr (MODE == RB_RED_FWD) fvv = cmul(PD[p].dv, zv);
        xfo[(li == 0 ? 0 : 64) + lj + 8 * lk] = fvv;
      }
    }
    wave_sync();
    w = wn;
  }
}

// ------------------------------------ red-black sweeps, systems per workgroup
// The same three launches with NW systems of one tile per workgroup, one wave
// per system.  The tile's couplings are loaded once, raw, into LDS shared by
// the NW waves (the sweeps mask them by colour and position); each wave then
// stages its own system's right-hand side, pivots and halo and sweeps the
// tile alone, with wave-level synchronisation only.  Per-system LDS is the
// box and the packed point data (16 KB in FP32), and nothing is held in
// registers across items, so two workgroups share a CU and one's loads run
// under the other's sweeps.  Work items are (tile, group of NW active
// systems), groups fastest, cut into 8 contiguous per-XCD ranges, so the
// groups of a tile run on one XCD and share its pivots and couplings in L2.
template <class T, class V, int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void k_dilu_rbm(DiluArgs<T, V> a, const TileShape* shp) {
  constexpr int TX = 8, NP = 512, NU = 8, NLEV = 22, BX = 10, NB = 1000;
  constexpr bool FWDP = MODE != RB_RED_BWD;   // has a forward sweep
  constexpr bool BWDP = MODE != RB_RED_FWD;   // has a backward sweep
  constexpr bool HALO = MODE != RB_RED_FWD;   // reads face neighbours
  struct alignas(16) PtA { C2<T> rb, dv; };
  struct Cpl { V c[3]; };
  __shared__ C2<T> Wall[NW][NB];                 // per system: the box
  __shared__ PtA PDall[NW][FWDP ? NP : 1];       // per system: rhs -> y, pivot
  __shared__ C2<T> DVall[NW][FWDP ? 1 : NP];     // per system (red backward): pivot
  __shared__ Cpl CL[NP];                         // raw lower couplings (slots 1, 3, 5)
  __shared__ Cpl CH[BWDP ? NP : 1];              // raw upper couplings (slots 2, 4, 6)
  __shared__ int s_sys[64], s_ok[64], s_fidx[64];
  __shared__ C2<T> s_al[64];
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int ngrp = (a.nact + NW - 1) / NW;
  const int nwork = a.ntiles * ngrp;
  const int xcd = blockIdx.x & 7, lb = blockIdx.x >> 3;
  const int chunk = (nwork + 7) / 8;
  const int w = xcd * chunk + lb;
  if (lb >= chunk || w >= nwork) return;   // uniform over the workgroup
  if (tid < a.nact) {
    const int sy = a.act[tid];
    const SysScal& sc = a.sc[sy];
    s_sys[tid] = sy;
    s_ok[tid] = sc.status == a.status_req;
    s_fidx[tid] = sc.fidx;
    s_al[tid] = ccast<T>(sc.alpha);
  }
  const int tix = w / ngrp, grp = w - tix * ngrp;
  const int4 tl = a.tiles[tix];
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TX, gk0 = tl.w * TX;
  const int ex = min(TX, L.A - gi0), ey = min(TX, L.B - gj0), ez = min(TX, L.C - gk0);
  const V* __restrict__ cfs = a.coef;
  const int64_t cst = a.cstride;
  // ---- the tile's couplings, once for the NW systems
  for (int p = tid; p < NP; p += 64 * NW) {
    const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
    Cpl c, h;
#pragma unroll
    for (int s = 0; s < 3; ++s) c.c[s] = h.c[s] = V(0);
    if (li < ex && lj < ey && lk < ez) {
      const int64_t e = lat_e(L, gi0 + li, gj0 + lj, gk0 + lk);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        c.c[s] = cfs[(int64_t)(1 + 2 * s) * cst + e];
        if (BWDP) h.c[s] = cfs[(int64_t)(2 + 2 * s) * cst + e];
      }
    }
    CL[p] = c;
    if (BWDP) CH[p] = h;
  }
  __syncthreads();   // s_ tables
  const int si = grp * NW + wv;
  const bool mine = si < a.nact && s_ok[si];
  const int sys = mine ? s_sys[si] : 0;
  C2<T>* __restrict__ W = Wall[wv];
  PtA* __restrict__ PD = PDall[wv];
  C2<T>* __restrict__ DV = DVall[wv];
  const int64_t o = (int64_t)sys * a.vstride;
  if (mine) {
    for (int q = lane; q < NB; q += 64) W[q] = {T(0), T(0)};
    const C2<T>* __restrict__ pin = a.in + o;
    const C2<T>* __restrict__ pin2 = a.in2 ? a.in2 + o : nullptr;
    const C2<T>* __restrict__ pout = a.out + o;
    const C2<T>* __restrict__ dinv = a.dinv + (int64_t)s_fidx[si] * cst;
    const C2<T> al = s_al[si];
    C2<T> rv[NU], v2[NU], dd[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      rv[u] = v2[u] = dd[u] = {T(0), T(0)};
      if (li < ex && lj < ey && lk < ez) {
        const int64_t e = lat_e(L, gi0 + li, gj0 + lj, gk0 + lk);
        rv[u] = FWDP ? pin[e] : pout[e];
        if (FWDP && pin2) v2[u] = pin2[e];
        dd[u] = dinv[e];
      }
    }
    C2<T> fv[6], fd[6];
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
        fv[f] = fd[f] = {T(0), T(0)};
        const bool ok = ti < ex && tj < ey && tk < ez && ni >= 0 && nj >= 0 && nk >= 0 &&
                        ni < L.A && nj < L.B && nk < L.C;
        if (ok) {
          if (dir == 0) {
            // the neighbour tile's facing i-face, packed (entry tj + 8*tk = lane)
            const int tn = L.tbase + (tl.y + (hi ? 1 : -1)) + L.ntx * (tl.z + L.nty * tl.w);
            fv[f] = a.xf[(MODE == RB_BLACK_FB ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
                         (int64_t)tn * 128 + (hi ? 0 : 64) + lane];
          } else {
            const int64_t l = lat_e(L, ni, nj, nk);
            fv[f] = pout[l];
            if (MODE == RB_BLACK_FB) fd[f] = dinv[l];
          }
        }
      }
    }
    // ---- stage into LDS
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
      C2<T> r = rv[u];
      if (FWDP && pin2) r = csub(r, cmul(al, v2[u]));
      if (FWDP) PD[p] = PtA{r, dd[u]};
      else { DV[p] = dd[u]; W[q] = r; }
    }
    if (HALO) {
      const int fa = lane & 7, fb = lane >> 3;
#pragma unroll
      for (int f = 0; f < 6; ++f) {
        const int dir = f >> 1;
        const bool hi = f & 1;
        int q;
        if (dir == 0) q = (hi ? 9 : 0) + BX * ((fa + 1) + BX * (fb + 1));
        else if (dir == 1) q = (fa + 1) + BX * ((hi ? 9 : 0) + BX * (fb + 1));
        else q = (fa + 1) + BX * ((fb + 1) + BX * (hi ? 9 : 0));
        // black: red neighbours enter as dinv*y (the packed i-faces already are)
        W[q] = (MODE == RB_BLACK_FB && dir != 0) ? cmul(fd[f], fv[f]) : fv[f];
      }
    }
  }
  __syncthreads();   // the shared couplings; the rest is per wave
  if (!mine) return;
  // fold the face terms the sweeps do not reach: black forward, the high
  // faces (red neighbours above); red backward, the low faces (black below)
  if (HALO) {
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      if (!(li < ex && lj < ey && lk < ez)) continue;
      const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
      const int st[3] = {1, BX, BX * BX};
      if (MODE == RB_BLACK_FB) {
        const bool fl[3] = {li == TX - 1, lj == TX - 1, lk == TX - 1};
        const Cpl h = CH[p];
        C2<T> r = PD[p].rb;
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          if (!fl[s]) continue;
          const T c = T(h.c[s]);
          const C2<T> v = W[q + st[s]];
          r.x -= c * v.x;
          r.y -= c * v.y;
        }
        PD[p].rb = r;
      } else {
        const bool fl[3] = {li == 0, lj == 0, lk == 0};
        const Cpl c4 = CL[p];
        C2<T> y = W[q];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          if (!fl[s]) continue;
          const T c = T(c4.c[s]);
          const C2<T> v = W[q - st[s]];
          y.x -= c * v.x;
          y.y -= c * v.y;
        }
        W[q] = y;
      }
    }
    wave_sync();
  }
  // this lane's point on every hyperplane of the tile (at most 48 per plane)
  unsigned int pk[NLEV];
#pragma unroll
  for (int h = 0; h < NLEV; ++h) {
    const int b = shp->lvl[h], e = shp->lvl[h + 1];
    pk[h] = (b + lane < e) ? shp->pts[b + lane] : 0xffffffffu;
  }
  if (FWDP) {
#pragma unroll
    for (int h = 0; h < NLEV; ++h) {
      const unsigned int pt = pk[h];
      const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = (pt >> 16) & 0xff;
      if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
        const int p = li + TX * (lj + TX * lk);
        const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
        const PtA pa = PD[p];
        const Cpl c4 = CL[p];
        // red: face neighbours are upper, so only in-tile lower couplings;
        // black: every lower-side coupling (in tile or across the face)
        const bool in[3] = {li > 0, lj > 0, lk > 0};
        C2<T> y = pa.rb;
        const int qn[3] = {q - 1, q - BX, q - BX * BX};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = (MODE == RB_BLACK_FB || in[s]) ? T(c4.c[s]) : T(0);
          const C2<T> v = W[qn[s]];
          y.x -= c * v.x;
          y.y -= c * v.y;
        }
        PD[p].rb = y;
        W[q] = cmul(pa.dv, y);
      }
      wave_sync();
    }
  }
  if (BWDP) {
#pragma unroll
    for (int h = NLEV - 1; h >= 0; --h) {
      const unsigned int pt = pk[h];
      const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = (pt >> 16) & 0xff;
      if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
        const int p = li + TX * (lj + TX * lk);
        const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
        C2<T> y, dvp;
        if constexpr (FWDP) {
          const PtA pa = PD[p];
          y = pa.rb;
          dvp = pa.dv;
        } else {
          y = W[q];
          dvp = DV[p];
        }
        const Cpl h4 = CH[p];
        // black: upper neighbours in the tile only; red: in tile or across
        const bool up[3] = {li + 1 < ex, lj + 1 < ey, lk + 1 < ez};
        const int qn[3] = {q + 1, q + BX, q + BX * BX};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = (MODE == RB_RED_BWD || up[s]) ? T(h4.c[s]) : T(0);
          const C2<T> z = W[qn[s]];
          y.x -= c * z.x;
          y.y -= c * z.y;
        }
        W[q] = cmul(dvp, y);
      }
      wave_sync();
    }
  }
  // ---- write the tile back (red forward: y; otherwise z), and the packed
  // i-faces the other colour reads next (red forward: dinv*y; black: z)
  C2<T>* __restrict__ dst = a.out + o;
  C2<T>* __restrict__ xfo =
      a.xf + (MODE == RB_RED_FWD ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
      (int64_t)(L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w)) * 128;
  for (int p = lane; p < NP; p += 64) {
    const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
    if (!(li < ex && lj < ey && lk < ez)) continue;
    const int q = (li + 1) + BX * ((lj + 1) + BX * (lk + 1));
    C2<T> zv;
    if constexpr (MODE == RB_RED_FWD) zv = PD[p].rb;
    else zv = W[q];
    dst[lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = zv;
    if (MODE != RB_RED_BWD && (li == 0 || li == TX - 1)) {
      C2<T> fvv = zv;
      if constexpr (MODE == RB_RED_FWD) fvv = cmul(PD[p].dv, zv);
      xfo[(li == 0 ? 0 : 64) + lj + 8 * lk] = fvv;
    }
  }
}

// ------------------------------ red-black sweeps, four systems, lean LDS
// The launches of k_dilu_rbm with the LDS of a workgroup cut below a third of
// a CU's 160 KB, so that three workgroups (12 waves) share a CU: the sweeps
// are bound by LDS latency, which resident waves hide.
//  * The right-hand side lives in the box in place: a point's forward step
//    reads its own box entry, so the per-point data shrink to the pivots
//    (red forward then keeps y in the pivot's slot, which it no longer needs).
//  * The box (rb_box) is the tile and one pad layer on the low side, with a
//    plane pitch that spreads a hyperplane over the LDS banks.  An
//    upper-neighbour index past the end wraps to the start, which puts the
//    high faces in low-side pad slots, disjoint from one another and from the
//    tile.  Pivots and couplings are stored by level slot, so the lanes of a
//    sweep step read consecutive entries.  A mode reads one side's faces from the box (red forward none,
//    black the low ones, red backward the high ones); the other side's terms
//    are folded into the right-hand side before the sweeps by the lanes that
//    loaded the face values, one direction after another, which is the
//    per-point order of the k_dilu_rbm folds.
//  * Black backward: z = (dv y) - dv (sum of the upper terms), from the box's
//    dv y (the same operator as dv (y - sum), rounded differently).  Red
//    forward and red backward round as k_dilu_rbm does.
//  * System tables per workgroup.
template <class T, class V, int MODE, int NW>
__global__ __launch_bounds__(64 * NW) void k_dilu_rbl(DiluArgs<T, V> a, const TileShape* shp) {
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
  const V* __restrict__ cfs = a.coef;
  const int64_t cst = a.cstride;
  const int st[3] = {1, kRbBX, kRbPK};
  // ---- the tile's couplings, once for the NW systems
  for (int p = tid; p < NP; p += 64 * NW) {
    const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
    const int sl = shp->rslot[p];
    Cpl c, h;
#pragma unroll
    for (int s = 0; s < 3; ++s) c.c[s] = h.c[s] = V(0);
    if (li < ex && lj < ey && lk < ez) {
      const int64_t e = lat_e(L, gi0 + li, gj0 + lj, gk0 + lk);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        if (FWDP) c.c[s] = cfs[(int64_t)(1 + 2 * s) * cst + e];
        if (BWDP) h.c[s] = cfs[(int64_t)(2 + 2 * s) * cst + e];
      }
    }
    if (FWDP) CL[sl] = c;
    if (BWDP) CH[sl] = h;
  }
  __syncthreads();   // s_ tables
  // this lane's point on every hyperplane of the tile (at most 48 per plane),
  // requested before the staging loads so that their latencies overlap: the
  // level bounds in one batch, then unconditional loads (index clamped into
  // the table) and a select
  unsigned int pk[NLEV];
  {
    int lv[NLEV + 1];
#pragma unroll
    for (int h = 0; h <= NLEV; ++h) lv[h] = shp->lvl[h];
#pragma unroll
    for (int h = 0; h < NLEV; ++h) {
      const int idx = lv[h] + lane;
      const unsigned int v = shp->rpts[min(idx, NP - 1)];
      pk[h] = idx < lv[h + 1] ? v : 0xffffffffu;
    }
  }
  const bool mine = s_ok[wv] != 0;   // 0 for the slots past nact
  const int sys = s_sys[wv];
  C2<T>* __restrict__ W = Wall[wv];
  C2<T>* __restrict__ DV = DVall[wv];
  const int64_t o = (int64_t)sys * a.vstride;
  const C2<T>* __restrict__ pout = a.out + o;
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
    const C2<T>* __restrict__ pin = a.in + o;
    const C2<T>* __restrict__ pin2 = a.in2 ? a.in2 + o : nullptr;
    const C2<T>* __restrict__ dinv = a.dinv + (int64_t)s_fidx[wv] * cst;
    const C2<T> al = s_al[wv];
    C2<T> rv[NU], v2[NU], dd[NU];
    int rs[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      rs[u] = shp->rslot[p];
      rv[u] = v2[u] = dd[u] = {T(0), T(0)};
      if (li < ex && lj < ey && lk < ez) {
        const int64_t e = lat_e(L, gi0 + li, gj0 + lj, gk0 + lk);
        rv[u] = FWDP ? pin[e] : pout[e];
        if (FWDP && pin2) v2[u] = pin2[e];
        dd[u] = dinv[e];
      }
    }
    C2<T> fv[6], fd[6];
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
        fv[f] = fd[f] = {T(0), T(0)};
        const bool pin_ok = ti < ex && tj < ey && tk < ez;
        const bool ok = pin_ok && ni >= 0 && nj >= 0 && nk >= 0 && ni < L.A && nj < L.B && nk < L.C;
        if (ok) {
          if (dir == 0) {
            // the neighbour tile's facing i-face, packed (entry tj + 8*tk = lane)
            const int tn = L.tbase + (tl.y + (hi ? 1 : -1)) + L.ntx * (tl.z + L.nty * tl.w);
            fv[f] = a.xf[(MODE == RB_BLACK_FB ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
                         (int64_t)tn * 128 + (hi ? 0 : 64) + lane];
          } else {
            const int64_t l = lat_e(L, ni, nj, nk);
            fv[f] = pout[l];
            if (MODE == RB_BLACK_FB) fd[f] = dinv[l];
          }
        }
        if (MODE == RB_RED_BWD && !hi && pin_ok)
          fold_c[dir] = T(cfs[(int64_t)(1 + 2 * dir) * cst +
                              lat_e(L, gi0 + ti, gj0 + tj, gk0 + tk)]);
      }
    }
    // ---- stage into LDS: the right-hand side in the box, the pivots
#pragma unroll
    for (int u = 0; u < NU; ++u) {
      const int p = lane + 64 * u;
      const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
      C2<T> r = rv[u];
      if (FWDP && pin2) r = csub(r, cmul(al, v2[u]));
      W[rb_box(li, lj, lk)] = r;
      DV[rs[u]] = dd[u];
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
        // black: red neighbours enter as dinv*y (the packed i-faces already are)
        const C2<T> v = (MODE == RB_BLACK_FB && dir != 0) ? cmul(fd[f], fv[f]) : fv[f];
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
        const T c = (MODE == RB_BLACK_FB) ? T(CH[shp->rslot[ti + TX * (tj + TX * tk)]].c[s]) : fold_c[s];
        const C2<T> v = fold_v[s];
        C2<T> r = W[q];
        r.x -= c * v.x;
        r.y -= c * v.y;
        W[q] = r;
      }
      wave_sync();
    }
  }
  if (FWDP) {
#pragma unroll
    for (int h = 0; h < NLEV; ++h) {
      const unsigned int pt = pk[h];
      const int li = pt & 7, lj = (pt >> 3) & 7, lk = (pt >> 6) & 7;
      if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
        const int p = pt >> 9;   // level slot
        const int q = rb_box(li, lj, lk);
        const C2<T> dvp = DV[p];
        const Cpl c4 = CL[p];
        // red: face neighbours are upper, so only in-tile lower couplings;
        // black: every lower-side coupling (in tile or across the face)
        const bool in[3] = {li > 0, lj > 0, lk > 0};
        C2<T> y = W[q];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = (MODE == RB_BLACK_FB || in[s]) ? T(c4.c[s]) : T(0);
          const C2<T> v = W[q - st[s]];
          y.x -= c * v.x;
          y.y -= c * v.y;
        }
        W[q] = cmul(dvp, y);
        if (MODE == RB_RED_FWD) DV[p] = y;
      }
      wave_sync();
    }
  }
  if (BWDP) {
#pragma unroll
    for (int h = NLEV - 1; h >= 0; --h) {
      const unsigned int pt = pk[h];
      const int li = pt & 7, lj = (pt >> 3) & 7, lk = (pt >> 6) & 7;
      if (pt != 0xffffffffu && li < ex && lj < ey && lk < ez) {
        const int p = pt >> 9;   // level slot
        const int q = rb_box(li, lj, lk);
        const C2<T> dvp = DV[p];
        const Cpl h4 = CH[p];
        // black: upper neighbours in the tile only; red: in tile or across
        const bool up[3] = {li + 1 < ex, lj + 1 < ey, lk + 1 < ez};
        C2<T> y = (MODE == RB_RED_BWD) ? W[q] : C2<T>{T(0), T(0)};
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          const T c = (MODE == RB_RED_BWD || up[s]) ? T(h4.c[s]) : T(0);
          int qn = q + st[s];
          if (qn >= NB) qn -= NB;
          const C2<T> z = W[qn];
          y.x -= c * z.x;
          y.y -= c * z.y;
        }
        if constexpr (MODE == RB_RED_BWD) {
          W[q] = cmul(dvp, y);
        } else {
          // y holds -(sum of the upper terms); the box holds dv*y_forward
          const C2<T> d = cmul(dvp, y);
          const C2<T> b = W[q];
          W[q] = {b.x + d.x, b.y + d.y};
        }
      }
      wave_sync();
    }
  }
  // ---- write the tile back (red forward: y; otherwise z), and the packed
  // i-faces the other colour reads next (red forward: dinv*y; black: z)
  C2<T>* __restrict__ dst = a.out + o;
  C2<T>* __restrict__ xfo =
      a.xf + (MODE == RB_RED_FWD ? 0 : a.xfhalf) + (int64_t)sys * a.xfsys +
      (int64_t)(L.tbase + tl.y + L.ntx * (tl.z + L.nty * tl.w)) * 128;
  for (int p = lane; p < NP; p += 64) {
    const int li = p & 7, lj = (p >> 3) & 7, lk = p >> 6;
    if (!(li < ex && lj < ey && lk < ez)) continue;
    const int q = rb_box(li, lj, lk);
    const C2<T> zv = (MODE == RB_RED_FWD) ? DV[shp->rslot[p]] : W[q];
    dst[lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = zv;
    if (MODE != RB_RED_BWD && (li == 0 || li == TX - 1))
      xfo[(li == 0 ? 0 : 64) + lj + 8 * lk] = W[q];   // red forward: dinv*y; black: z
  }
}

// ------------------------------------------------------------- pivots
// PC_setup's block-DILU pivots on the device (Dilu_Cmplx_AS, spOpTools.f90:
// 1996-2055, on the x-x, y-y, z-z blocks of A(omega)):
//   p_e = a_ee + i*ISIGN*omega*vms1_e - sum_{l lower} a_el a_le / p_l,  dinv_e = 1/p_e
// The recurrence has the dependencies of the forward sweep, so it runs on the
// same tile levels (one launch per level, all frequencies at once).  Every
// point combines its three lower neighbours in the host routine's order.
struct PivotArgs {
  const double* coef;      // [15][cstride] stencil slots (1,3,5 lower; 2,4,6 upper)
  int64_t cstride;
  const double* vms1;      // [cstride]
  LatticeInfo lat[3];
  const int4* tiles;
  int ntiles;
  const double* omegas;    // [nf]
  int nf;
  Z* dinv;                 // [nf][cstride]
};

template <int TX, int TY, int TZ>
__global__ __launch_bounds__(64) void k_dilu_pivots(PivotArgs a, const TileShape* shp) {
#pragma clang fp contract(off)
  constexpr int BX = TX + 1, BY = TY + 1, BZ = TZ + 1, NB = BX * BY * BZ, NP = TX * TY * TZ;
  constexpr int NQ = (NB + 63) / 64;
  constexpr int NLEV = TX + TY + TZ - 2;
  __shared__ Z pv[NB];          // dinv of every box point (halo: final)
  __shared__ double cu[3][NB];  // upper couplings a_le of every box point
  __shared__ double cl[3][NP];  // lower couplings a_el of the tile points
  __shared__ Z dg[NP];          // diagonal of the tile points
  __shared__ unsigned int spts[NP];
  __shared__ int slvl[NLEV + 1];
  const int lane = threadIdx.x;
  const int tix = blockIdx.x / a.nf, f = blockIdx.x - tix * a.nf;
  const int4 tl = a.tiles[tix];
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TY, gk0 = tl.w * TZ;
  const int ex = min(TX, L.A - gi0), ey = min(TY, L.B - gj0), ez = min(TZ, L.C - gk0);
  const double wsg = kIsignD * a.omegas[f];
  Z* __restrict__ dv = a.dinv + (int64_t)f * a.cstride;
  const double* __restrict__ cf = a.coef;
  const double* __restrict__ vm = a.vms1;
  for (int u = 0; u < NQ; ++u) {
    const int q = lane + 64 * u;
    if (q >= NB) break;
    const int bi = q % BX, bj = (q / BX) % BY, bk = q / (BX * BY);
    const int li = bi - 1, lj = bj - 1, lk = bk - 1;
    const bool inBoxTile = li >= 0 && lj >= 0 && lk >= 0 && li < TX && lj < TY && lk < TZ;
    const bool inTile = inBoxTile && li < ex && lj < ey && lk < ez;
    const int gi = gi0 + li, gj = gj0 + lj, gk = gk0 + lk;
    const bool inLat = gi >= 0 && gj >= 0 && gk >= 0 && gi < L.A && gj < L.B && gk < L.C;
    Z pvq = {0.0, 0.0};
    double up[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
    Z d = {0.0, 0.0};
    if (inLat) {
      const int64_t e = lat_e(L, gi, gj, gk);
      if (!inTile) pvq = dv[e];
#pragma unroll
      for (int s = 0; s < 3; ++s) up[s] = cf[(int64_t)(2 + 2 * s) * a.cstride + e];
      if (inTile) {
#pragma unroll
        for (int s = 0; s < 3; ++s) lo[s] = cf[(int64_t)(1 + 2 * s) * a.cstride + e];
        d = {cf[e], wsg * vm[e]};
      }
    }
    pv[q] = pvq;
#pragma unroll
    for (int s = 0; s < 3; ++s) cu[s][q] = up[s];
    if (inBoxTile) {
      const int p = li + TX * (lj + TY * lk);
#pragma unroll
      for (int s = 0; s < 3; ++s) cl[s][p] = lo[s];
      dg[p] = d;
    }
  }
  for (int p = lane; p < NP; p += 64) spts[p] = shp->pts[p];
  if (lane <= NLEV) slvl[lane] = shp->lvl[lane];
  wave_sync();
  for (int h = 0; h < NLEV; ++h) {
    for (int pp = slvl[h] + lane; pp < slvl[h + 1]; pp += 64) {
      const unsigned int pt = spts[pp];
      const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = pt >> 16;
      if (li >= ex || lj >= ey || lk >= ez) continue;
      const int p = li + TX * (lj + TY * lk);
      const int q = (li + 1) + BX * ((lj + 1) + BY * (lk + 1));
      const int qn[3] = {q - 1, q - BX, q - BX * BY};
      Z piv = dg[p];
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const double alo = cl[s][p];
        if (alo == 0.0) continue;
        const double t = alo * cu[s][qn[s]];
        const Z w = pv[qn[s]];
        piv.x = piv.x - t * w.x;
        piv.y = piv.y - t * w.y;
      }
      pv[q] = zdiv({1.0, 0.0}, piv);
    }
    wave_sync();
  }
  for (int p = lane; p < NP; p += 64) {
    const int li = p % TX, lj = (p / TX) % TY, lk = p / (TX * TY);
    if (li >= ex || lj >= ey || lk >= ez) continue;
    dv[lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = pv[(li + 1) + BX * ((lj + 1) + BY * (lk + 1))];
  }
}

// Pivots in the red-black tile ordering: red tiles (HALO=false) are
// factorised on their own; a black tile (HALO=true) first takes the
// contributions a_el a_le / p_l of its face neighbours in the (final) red
// tiles off its diagonal, then runs the in-tile recurrence.
template <int TX, int TY, int TZ, bool HALO>
__global__ __launch_bounds__(64) void k_dilu_pivots_rb(PivotArgs a, const TileShape* shp) {
#pragma clang fp contract(off)
  constexpr int NP = TX * TY * TZ;
  constexpr int NLEV = TX + TY + TZ - 2;
  __shared__ Z pv[NP];
  __shared__ double cu[3][NP];  // upper couplings of the tile points
  __shared__ double cl[3][NP];  // in-tile lower couplings
  __shared__ Z dg[NP];
  __shared__ unsigned int spts[NP];
  __shared__ int slvl[NLEV + 1];
  const int lane = threadIdx.x;
  const int tix = blockIdx.x / a.nf, f = blockIdx.x - tix * a.nf;
  const int4 tl = a.tiles[tix];
  const LatticeInfo L = a.lat[tl.x];
  const int gi0 = tl.y * TX, gj0 = tl.z * TY, gk0 = tl.w * TZ;
  const int ex = min(TX, L.A - gi0), ey = min(TY, L.B - gj0), ez = min(TZ, L.C - gk0);
  const double wsg = kIsignD * a.omegas[f];
  Z* __restrict__ dvo = a.dinv + (int64_t)f * a.cstride;
  const double* __restrict__ cf = a.coef;
  const int64_t cst = a.cstride;
  const int64_t sj = L.a, sk = int64_t(L.a) * L.b;
  for (int p = lane; p < NP; p += 64) {
    const int li = p % TX, lj = (p / TX) % TY, lk = p / (TX * TY);
    const bool inTile = li < ex && lj < ey && lk < ez;
    Z d = {0.0, 0.0};
    double up[3] = {0.0, 0.0, 0.0}, lo[3] = {0.0, 0.0, 0.0};
    if (inTile) {
      const int gi = gi0 + li, gj = gj0 + lj, gk = gk0 + lk;
      const int64_t e = lat_e(L, gi, gj, gk);
      d = {cf[e], wsg * a.vms1[e]};
#pragma unroll
      for (int s = 0; s < 3; ++s) up[s] = cf[(int64_t)(2 + 2 * s) * cst + e];
      if (li > 0) lo[0] = cf[1 * cst + e];
      if (lj > 0) lo[1] = cf[3 * cst + e];
      if (lk > 0) lo[2] = cf[5 * cst + e];
      if (HALO) {
        const bool fl[6] = {li == 0 && gi > 0, li + 1 == ex && gi + 1 < L.A,
                            lj == 0 && gj > 0, lj + 1 == ey && gj + 1 < L.B,
                            lk == 0 && gk > 0, lk + 1 == ez && gk + 1 < L.C};
        const int64_t st[6] = {-1, 1, -sj, sj, -sk, sk};
#pragma unroll
        for (int q = 0; q < 6; ++q) {
          if (!fl[q]) continue;
          const double ael = cf[(int64_t)(1 + q) * cst + e];
          if (ael == 0.0) continue;
          // the coupling back from l to e is l's slot in the opposite direction
          const int back = (q & 1) ? q : q + 2;  // 1-based slot: low<->high
          const double ale = cf[(int64_t)back * cst + e + st[q]];
          const double t = ael * ale;
          const Z w = dvo[e + st[q]];
          d.x = d.x - t * w.x;
          d.y = d.y - t * w.y;
        }
      }
    }
    dg[p] = d;
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      cu[s][p] = up[s];
      cl[s][p] = lo[s];
    }
    pv[p] = {0.0, 0.0};
  }
  for (int p = lane; p < NP; p += 64) spts[p] = shp->pts[p];
  if (lane <= NLEV) slvl[lane] = shp->lvl[lane];
  wave_sync();
  const int stp[3] = {1, TX, TX * TY};
  for (int h = 0; h < NLEV; ++h) {
    for (int pp = slvl[h] + lane; pp < slvl[h + 1]; pp += 64) {
      const unsigned int pt = spts[pp];
      const int li = pt & 0xff, lj = (pt >> 8) & 0xff, lk = pt >> 16;
      if (li >= ex || lj >= ey || lk >= ez) continue;
      const int p = li + TX * (lj + TY * lk);
      Z piv = dg[p];
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const double alo = cl[s][p];
        if (alo == 0.0) continue;
        const double t = alo * cu[s][p - stp[s]];
        const Z w = pv[p - stp[s]];
        piv.x = piv.x - t * w.x;
        piv.y = piv.y - t * w.y;
      }
      pv[p] = zdiv({1.0, 0.0}, piv);
    }
    wave_sync();
  }
  for (int p = lane; p < NP; p += 64) {
    const int li = p % TX, lj = (p / TX) % TY, lk = p / (TX * TY);
    if (li >= ex || lj >= ey || lk >= ez) continue;
    dvo[lat_e(L, gi0 + li, gj0 + lj, gk0 + lk)] = pv[p];
  }
}

template <class T>
__global__ void k_cast_pivots(const Z* __restrict__ in, C2<T>* __restrict__ out, int64_t n) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n;
       q += (int64_t)gridDim.x * blockDim.x)
    out[q] = ccast<T>(in[q]);
}

}  // namespace mt
