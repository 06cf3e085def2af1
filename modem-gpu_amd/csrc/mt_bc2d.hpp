// The boundary values and initial guess of a forward solve from 2-D TE slice
// solves (BC_x0_WS, 3D_MT/FWD/boundary_ws.f90:21-171), on the device.
//
// For mode 1 (Ey polarisation) the reference solves one 2-D TE problem per
// y-slice over the (x, z) cells of the slice and writes it to the y-edges;
// for mode 2 (Ex) one per x-slice over (y, z), written to the x-edges.  Each
// problem (Siripunvaraporn's code, 3D_MT/FWD/Mod2d/) has (nh-1)(nz-1)
// interior nodes, 1 at the top of the air, and on the sides and bottom the
// values of 1-D TE columns (Fwd1D_TE, WSfwd1Dmod.f90:79-141).  Its 5-point
// system is band-factored by LAPACK (FormAII -> ZGBTRF, WSfwd2Dmod.f90:31-85)
// and solved (Fwd2DsolveTE -> ZGBTRS, FwdTEmod.f90:110-169), bandwidth nz-1
// on both sides of the diagonal.
//
// k_bc_te1d solves every 1-D column, one thread each.  k_bc_te2d solves one
// slice per workgroup: the band's active (nz x nz) window stays in LDS as a
// circular buffer over rows and columns, the right-hand side is eliminated
// as the factorisation proceeds, the rows of U go to a per-workgroup scratch
// in HBM, and one wave back-substitutes from staged chunks of U.
//
// Every complex product and quotient is formed as the reference's BLAS and
// LAPACK form it (textbook products, C99 Annex G division as __divdc3 does
// it, no contraction) and every entry receives its updates in the same
// order, so the fields equal the reference's bit for bit while its partial
// pivoting keeps the diagonal.  (ZGERU's a + x*(-1*y) is formed as
// a - x*y: negation is exact and round-to-nearest symmetric, so the two
// agree in value and differ at most in the sign of a zero.)  A column where it would interchange rows (or
// a zero pivot) is flagged, and the entry point reports it.
#pragma once

#include "mt_device.hpp"
#include "mt_dilu.hpp"
#include "mt_internal.h"

namespace mt {

constexpr int kBcChunk = 32;   // factorisation steps per staged chunk of entering coefficients

struct BcArgs {
  int nx, ny, nz, nza;          // 3-D grid; nz counts the air layers
  int64_t nXe, nE;              // x-edges, all edges
  const double *dx, *dy, *dz;   // cell sizes
  const double *cx, *cy, *cz;   // DistanceBetweenBlocks of dx, dy, dz (n+1 entries each)
  const double* sig;            // earth cell conductivity [nzEarth][ny][nx]
  const double* omega;          // [nsys]
  const int* mode;              // [nsys] 1 or 2
  const int2* item;             // [nitem] (system, slice), slice 1-based
  int nitem;
  int ncolMax;                  // max(nx, ny) + 1: 1-D columns per item, padded
  int mmbMax;                   // boundary vector length per item (1-based, index 0 unused)
  double2* exb;                 // [nitem][mmbMax] boundary values of each slice (EXB)
  double2* side;                // [nitem][2][nz][2] factors of the two side columns
  double2* ucol;                // per workgroup [nmax][nz]: ucol[j-1][o] = U(j-o, j)
  double2* yx;                  // per workgroup [nmax]: eliminated rhs, then the solution
  int64_t uStride, yStride;
  int* flag;                    // [nitem] 1 if the reference's LU would pivot (or a zero pivot)
  double2* e;                   // [nsys][nE] out: E0 inside, the boundary conditions outside
};

struct BcSlice {
  int sys, s, mode, nh;
  const double *dh, *ch;
  double omue;
};

__device__ __forceinline__ BcSlice bc_slice(const BcArgs& a, int it) {
#pragma clang fp contract(off)
  const int2 q = a.item[it];
  BcSlice S;
  S.sys = q.x;
  S.s = q.y;
  S.mode = a.mode[q.x];
  S.nh = S.mode == 2 ? a.ny : a.nx;
  S.dh = S.mode == 2 ? a.dy : a.dx;
  S.ch = S.mode == 2 ? a.cy : a.cx;
  // ComputeBC passes period = 2 PI / omega; the 2-D and 1-D codes take
  // omega = 2 PI / period back and omega mu0 from it
  const double per = (2.0 * kPi) / a.omega[q.x];
  S.omue = ((2.0 * kPi) / per) * kMu0;
  return S;
}

__device__ __forceinline__ double2 bz_mul(double2 a, double2 b) {
#pragma clang fp contract(off)
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// complex quotient, C99 Annex G (G.5.1) for finite operands: the divisor
// scaled by 2^-logb(max(|c|, |d|))
__device__ __forceinline__ double2 bz_div(double2 n, double2 d) {
#pragma clang fp contract(off)
  const double a = n.x, b = n.y;
  double c = d.x, e = d.y;
  int il = 0;
  const double lw = logb(fmax(fabs(c), fabs(e)));
  if (isfinite(lw)) {
    il = int(lw);
    c = scalbn(c, -il);
    e = scalbn(e, -il);
  }
  const double den = c * c + e * e;
  return make_double2(scalbn((a * c + b * e) / den, -il), scalbn((b * c - a * e) / den, -il));
}

__device__ __forceinline__ bool bz_nz(double2 v) { return v.x != 0.0 || v.y != 0.0; }
__device__ __forceinline__ double bz_abs1(double2 v) { return fabs(v.x) + fabs(v.y); }

// cell conductivity of the slice, 1-based (ih, iz); the air is SIGMA_AIR
__device__ __forceinline__ double bc_sig(const BcArgs& a, const BcSlice& S, int ih, int iz) {
  if (iz <= a.nza) return kSigmaAir;
  const int i = S.mode == 2 ? S.s : ih, j = S.mode == 2 ? ih : S.s;
  return a.sig[(i - 1) + int64_t(a.nx) * ((j - 1) + int64_t(a.ny) * (iz - a.nza - 1))];
}

// node q (1-based) of the slice's interior, iz fastest
__device__ __forceinline__ void bc_node(int b, int q, int& iy, int& iz) {
  iy = 2 + (q - 1) / b;
  iz = 2 + (q - 1) % b;
}

// SetupA_TE (WSfwd2Dmod.f90:345-438) + FormAII's diagonal
__device__ __forceinline__ double2 bc_diag(const BcArgs& a, const BcSlice& S, int iy, int iz) {
#pragma clang fp contract(off)
  const double s00 = bc_sig(a, S, iy, iz), s10 = bc_sig(a, S, iy, iz - 1);
  const double s01 = bc_sig(a, S, iy - 1, iz), s11 = bc_sig(a, S, iy - 1, iz - 1);
  const double dzi = a.dz[iz - 1], dzm = a.dz[iz - 2], dyi = S.dh[iy - 1], dym = S.dh[iy - 2];
  const double s2d = (s00 * dzi * dyi + s01 * dzi * dym + s10 * dzm * dyi + s11 * dzm * dym) /
                     (dzi * dyi + dzi * dym + dzm * dyi + dzm * dym);
  const double cz = a.cz[iz - 1], cy = S.ch[iy - 1];
  const double ar = 2.0 * cz / dyi, al = 2.0 * cz / dym, ad = 2.0 * cy / dzi, au = 2.0 * cy / dzm;
  const double a1 = -ar - al - ad - au;
  const double a4 = cy * cz * s2d;
  return make_double2(a1, S.omue * a4);
}
// coupling to the next node in z (0 at the bottom row)
__device__ __forceinline__ double bc_a2(const BcArgs& a, const BcSlice& S, int iy, int iz) {
#pragma clang fp contract(off)
  return iz == a.nz ? 0.0 : 2.0 * S.ch[iy - 1] / a.dz[iz - 1];
}
// coupling to the next node in the horizontal (0 in the last column)
__device__ __forceinline__ double bc_a3(const BcArgs& a, const BcSlice& S, int iy, int iz) {
#pragma clang fp contract(off)
  return iy < S.nh ? 2.0 * a.cz[iz - 1] / S.dh[iy - 1] : 0.0;
}

// MulAibWithXb (WSfwd2Dmod.f90:89-140) for one node: the sides assign (the
// right after the left), then the top and the bottom subtract
__device__ __forceinline__ double2 bc_rhs(const BcArgs& a, const BcSlice& S, const double2* exb,
                                          int iy, int iz) {
#pragma clang fp contract(off)
  const int Nz = a.nz, Ny = S.nh;
  double2 r = make_double2(0.0, 0.0);
  if (iy == 2) {
    const double c = 2.0 * a.cz[iz - 1] / S.dh[0];
    const double2 x = exb[iz];
    r = make_double2(-(c * x.x), -(c * x.y));
  }
  if (iy == Ny) {
    const double c = 2.0 * a.cz[iz - 1] / S.dh[Ny - 1];
    const double2 x = exb[Nz + 2 * Ny - 1 + iz];
    r = make_double2(-(c * x.x), -(c * x.y));
  }
  if (iz == 2) {
    const double c = 2.0 * S.ch[iy - 1] / a.dz[0];
    const double2 x = exb[Nz + iy];
    r = make_double2(r.x - c * x.x, r.y - c * x.y);
  }
  if (iz == Nz) {
    const double c = 2.0 * S.ch[iy - 1] / a.dz[Nz - 1];
    const double2 x = exb[Nz + Ny + iy - 1];
    r = make_double2(r.x - c * x.x, r.y - c * x.y);
  }
  return r;
}

// conductivity of the 1-D column col at layer iz (1..nz+1; nz+1 repeats nz):
// the side columns of the slice, or the dh-weighted mean of the two cells
// over an interior bottom node (SetBound2D_TE, WSfwd2Dmod.f90:262-341)
__device__ __forceinline__ double bc_col_sig(const BcArgs& a, const BcSlice& S, int col, int iz) {
#pragma clang fp contract(off)
  const int z = iz > a.nz ? a.nz : iz;
  if (z <= a.nza) return kSigmaAir;
  if (col == 0) return bc_sig(a, S, 1, z);
  if (col == S.nh) return bc_sig(a, S, S.nh, z);
  const int iy = col + 1;
  return (bc_sig(a, S, iy - 1, z) * S.dh[iy - 2] + bc_sig(a, S, iy, z) * S.dh[iy - 1]) /
         (S.dh[iy - 2] + S.dh[iy - 1]);
}

// Fwd1D_TE + Solve1D for every column of every slice: the tridiagonal LU
// (ZGBTF2 with kl = ku = 1, which never pivots here: every pivot exceeds its
// real subdiagonal in modulus) and ZGBTRS.  Interior columns keep only the
// bottom value; the two side columns keep their whole profile.
__global__ void k_bc_te1d(BcArgs a) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int it = int(t / a.ncolMax), col = int(t % a.ncolMax);
  if (it >= a.nitem) return;
  const BcSlice S = bc_slice(a, it);
  if (col > S.nh) return;
  const int Nz = a.nz;
  const bool sidecol = col == 0 || col == S.nh;
  double2* exb = a.exb + (int64_t)it * a.mmbMax;
  double2* sd = a.side + ((int64_t)it * 2 + (col == 0 ? 0 : 1)) * Nz * 2;
  const double skd = sqrt(2.0 / (bc_col_sig(a, S, col, Nz + 1) * S.omue));
  // ac(jj) = i omue (s(jj+1) dz1(jj+1) + s(jj) dz1(jj)) - 2/dz1(jj) - 2/dz1(jj+1)
  double dprev = a.dz[0], sprev = bc_col_sig(a, S, col, 1);
  const double au1 = 2.0 / dprev;
  double dcur = Nz >= 2 ? a.dz[1] : skd, scur = bc_col_sig(a, S, col, 2);
  double adv = 2.0 / dcur;
  double2 p = make_double2((0.0 - au1) - adv, S.omue * (scur * dcur + sprev * dprev));
  double2 bj = make_double2(-au1, 0.0);
  for (int j = 1; j < Nz; ++j) {
    if (sidecol) {
      sd[2 * (j - 1)] = p;
      sd[2 * (j - 1) + 1] = bj;
    }
    const double2 r = bz_div(make_double2(1.0, 0.0), p);
    const double2 l = bz_mul(r, make_double2(adv, 0.0));
    const double2 tt = bz_mul(make_double2(-1.0, 0.0), make_double2(adv, 0.0));
    // next row: au(j+1) = 2/dz1(j+1) is this adv
    dprev = dcur;
    sprev = scur;
    dcur = j + 2 <= Nz ? a.dz[j + 1] : skd;
    scur = bc_col_sig(a, S, col, j + 2);
    const double au = adv;
    adv = 2.0 / dcur;
    const double2 ac = make_double2((0.0 - au) - adv, S.omue * (scur * dcur + sprev * dprev));
    const double2 lt = bz_mul(l, tt);
    p = make_double2(ac.x + lt.x, ac.y + lt.y);
    if (bz_nz(bj)) {
      const double2 lb = bz_mul(l, bz_mul(make_double2(-1.0, 0.0), bj));
      bj = make_double2(0.0 + lb.x, 0.0 + lb.y);
    } else {
      bj = make_double2(0.0, 0.0);
    }
  }
  double2 x = bz_nz(bj) ? bz_div(bj, p) : bj;
  if (!sidecol) {
    exb[Nz + S.nh - 1 + (col + 1)] = x;   // X1D(nz+1) under interior node column col+1
    return;
  }
  // back substitution of the side column: x(j-1) from b(j-1) - x(j) ad(j-1)
  const int base = col == 0 ? 1 : Nz + 2 * S.nh;   // EXB position of X1D(1)
  exb[base] = make_double2(1.0, 0.0);
  exb[base + Nz] = x;
  for (int j = Nz - 1; j >= 1; --j) {
    const double2 pj = sd[2 * (j - 1)];
    double2 b = sd[2 * (j - 1) + 1];
    if (bz_nz(x)) {
      const double ad = 2.0 / a.dz[j];          // A(j, j+1) = 2/dz1(j+1)
      const double2 u = bz_mul(x, make_double2(ad, 0.0));
      b = make_double2(b.x - u.x, b.y - u.y);
    }
    x = bz_nz(b) ? bz_div(b, pj) : b;
    exb[base + j] = x;
  }
  if (col == 0)
    for (int ii = Nz + 2; ii <= Nz + S.nh; ++ii) exb[ii] = make_double2(1.0, 0.0);   // the surface
}

// A workgroup barrier that orders LDS only: the factorisation's global
// stores (rows of U, the eliminated rhs) are not read back until after a full
// barrier, so the steps need not wait for them (__syncthreads would).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__host__ __device__ inline size_t bc_lds_bytes(int nz) {
  const size_t B1 = size_t(nz);
  return (B1 * B1 + B1 + 2 * kBcChunk + 1) * sizeof(double2) + 2 * kBcChunk * sizeof(double);
}

// One slice per workgroup, persistent over the items.  Window slot of node r
// is r mod nz (nz = bandwidth + 1); W[cs*nz + rs] holds A(r, c).
__global__ __launch_bounds__(256) void k_bc_te2d(BcArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double2 lds[];
  __shared__ int sflag;
  const int b = a.nz - 1, B1 = a.nz;
  double2* W = lds;
  double2* Bv = W + B1 * B1;
  double2* Ed = Bv + B1;
  double2* Er = Ed + kBcChunk;
  double* Ea2 = reinterpret_cast<double*>(Er + kBcChunk);
  double* Ea3 = Ea2 + kBcChunk;
  double2* Rp = reinterpret_cast<double2*>(Ea3 + kBcChunk);   // 1 / next pivot
  const int tid = threadIdx.x, nt = blockDim.x;
  double2* U = a.ucol + (int64_t)blockIdx.x * a.uStride;
  double2* Y = a.yx + (int64_t)blockIdx.x * a.yStride;
  // (selects below are written as assignments: a ?: between a local double2
  // and a memory operand becomes a select of addresses, which keeps the
  // local in scratch)

  for (int it = blockIdx.x; it < a.nitem; it += gridDim.x) {
    const BcSlice S = bc_slice(a, it);
    const int Ny = S.nh, n = (Ny - 1) * b;
    const double2* exb = a.exb + (int64_t)it * a.mmbMax;
    if (tid == 0) sflag = 0;
    int myflag = 0;
    for (int q = tid; q < B1 * B1; q += nt) W[q] = make_double2(0.0, 0.0);
    __syncthreads();
    // the window of step 1: nodes 1..min(b, n)
    const int n0 = min(b, n);
    for (int r = 1 + tid; r <= n0; r += nt) {
      int iy, iz;
      bc_node(b, r, iy, iz);
      W[(r % B1) * B1 + r % B1] = bc_diag(a, S, iy, iz);
      if (r + 1 <= n0) {
        const double2 v = make_double2(bc_a2(a, S, iy, iz), 0.0);
        W[((r + 1) % B1) * B1 + r % B1] = v;
        W[(r % B1) * B1 + (r + 1) % B1] = v;
      }
      Bv[r % B1] = bc_rhs(a, S, exb, iy, iz);
    }
    for (int j = 1; j <= n; ++j) {
      const int cs = (j - 1) % kBcChunk;
      if (cs == 0) {
        // coefficients entering at steps j .. j+kBcChunk-1: node q = step+b
        // (diagonal, rhs, z-coupling to q-1) and the horizontal coupling of the step's own node
        for (int s = tid; s < kBcChunk; s += nt) {
          const int jj = j + s, q = jj + b;
          int iy, iz;
          if (q <= n) {
            bc_node(b, q, iy, iz);
            Ed[s] = bc_diag(a, S, iy, iz);
            Er[s] = bc_rhs(a, S, exb, iy, iz);
            bc_node(b, q - 1, iy, iz);
            Ea2[s] = bc_a2(a, S, iy, iz);
          }
          if (jj <= n) {
            bc_node(b, jj, iy, iz);
            Ea3[s] = bc_a3(a, S, iy, iz);
          }
        }
      }
      lds_barrier();
      const int m = min(b, n - j), sj = j % B1, q = j + b, sq = q % B1;
      const double2 p = W[sj * B1 + sj];
      const double2 yv = Bv[sj];
      const double2 a3j = make_double2(Ea3[cs], 0.0);
      if (tid >= 1 && tid <= m) {
        // column j of L: l_i = A(j+i, j) / p, the rhs update, the pivot test
        const int i = tid;
        int si = sj + i;
        if (si >= B1) si -= B1;
        double2 ai = a3j;
        if (i != b) ai = W[sj * B1 + si];
        if (!(bz_abs1(ai) <= bz_abs1(p))) myflag = 1;
        double2 r;
        if (j == 1) r = bz_div(make_double2(1.0, 0.0), p);
        else r = Rp[0];
        const double2 l = bz_mul(r, ai);
        W[sj * B1 + si] = l;
        double2 base;
        if (i == b) base = Er[cs];
        else base = Bv[si];
        if (bz_nz(yv)) {
          const double2 u = bz_mul(l, yv);
          base = make_double2(base.x - u.x, base.y - u.y);
        }
        Bv[si] = base;
      }
      if (tid <= m) {
        // row j of U, by column into the scratch
        const int c = tid;
        int sc = sj + c;
        if (sc >= B1) sc -= B1;
        double2 u;
        if (c == 0) u = p;
        else if (c == b) u = a3j;
        else u = W[sc * B1 + sj];
        U[(int64_t)(j + c - 1) * B1 + c] = u;
      }
      if (tid == 0) {
        Y[j - 1] = yv;
        if (!bz_nz(p)) myflag = 1;
      }
      if (q <= n) {
        if (tid <= b) {
          // the entering column q: rows j..q
          const int r = j + tid;
          double2 v = make_double2(0.0, 0.0);
          if (r == j) v = a3j;
          else if (r == q - 1) v = make_double2(Ea2[cs], 0.0);
          else if (r == q) v = Ed[cs];
          W[sq * B1 + r % B1] = v;
        }
        if (tid >= 1 && tid <= b - 1) {
          // the entering row q: columns j+1..q-1
          const int c = j + tid;
          W[(c % B1) * B1 + sq] = make_double2(c == q - 1 ? Ea2[cs] : 0.0, 0.0);
        }
      }
      lds_barrier();
      // rank-1 update of the trailing m x m block (ZGERU order per entry):
      // thread (g, i) owns rows i and i + mh and every G-th column from 1+g,
      // so each U-row value it reads serves two entries and its multipliers
      // are read once; four columns are in flight at a time.  The owner of
      // the next pivot (row 1, column 1: thread 0) also forms its reciprocal
      // for the next step.
      if (m > 0) {
        const int mh = (m + 1) / 2, G = nt / mh, g = tid / mh, i = 1 + tid % mh, i2 = i + mh;
        if (g < G) {
          const bool two = i2 <= m;
          int si = sj + i, si2 = sj + i2;
          if (si >= B1) si -= B1;
          if (si2 >= B1) si2 -= B1;
          const double2 l = W[sj * B1 + si];
          double2 l2 = make_double2(0.0, 0.0);
          if (two) l2 = W[sj * B1 + si2];
          int c = 1 + g;
          for (; c + 3 * G <= m; c += 4 * G) {
            int sc[4];
            double2 y[4], w[4], w2[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              sc[k] = sj + c + k * G;
              if (sc[k] >= B1) sc[k] -= B1;
              y[k] = W[sc[k] * B1 + sj];
              w[k] = W[sc[k] * B1 + si];
              if (two) w2[k] = W[sc[k] * B1 + si2];
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const double2 u = bz_mul(l, y[k]);
              w[k] = make_double2(w[k].x - u.x, w[k].y - u.y);
              W[sc[k] * B1 + si] = w[k];
              if (two) {
                const double2 u2 = bz_mul(l2, y[k]);
                W[sc[k] * B1 + si2] = make_double2(w2[k].x - u2.x, w2[k].y - u2.y);
              }
            }
            if (tid == 0 && c == 1) Rp[0] = bz_div(make_double2(1.0, 0.0), w[0]);
          }
          for (; c <= m; c += G) {
            int sc = sj + c;
            if (sc >= B1) sc -= B1;
            const double2 yv2 = W[sc * B1 + sj];
            const double2 u = bz_mul(l, yv2);
            const double2 w0 = W[sc * B1 + si];
            const double2 wn = make_double2(w0.x - u.x, w0.y - u.y);
            W[sc * B1 + si] = wn;
            if (two) {
              const double2 u2 = bz_mul(l2, yv2);
              const double2 w1 = W[sc * B1 + si2];
              W[sc * B1 + si2] = make_double2(w1.x - u2.x, w1.y - u2.y);
            }
            if (tid == 0 && c == 1) Rp[0] = bz_div(make_double2(1.0, 0.0), wn);
          }
        }
      }
    }
    if (myflag) atomicOr(&sflag, 1);
    __syncthreads();
    if (tid == 0) a.flag[it] = sflag;
    // back substitution with U (ZTBSV order): x_j = y_j / U(j,j), then
    // y_{j-o} -= x_j U(j-o, j) for o = 1..b; the y window is Bv
    for (int r = max(1, n - b) + tid; r <= n; r += nt) Bv[r % B1] = Y[r - 1];
    for (int j1 = n; j1 >= 1; j1 -= b) {
      const int j0 = max(1, j1 - b + 1), cnt = j1 - j0 + 1;
      __syncthreads();
      // stage U columns j0..j1 into W[(jj-j0)*B1 + o] and the entering y_{jj-b} into W[b*B1 + jj-j0]
      for (int e = tid; e < cnt * B1; e += nt) {
        const int jj = j0 + e / B1, o = e % B1;
        double2 v = make_double2(0.0, 0.0);
        if (o <= jj - 1) v = U[(int64_t)(jj - 1) * B1 + o];
        W[(jj - j0) * B1 + o] = v;
      }
      for (int e = tid; e < cnt; e += nt) {
        const int jj = j0 + e;
        double2 v = make_double2(0.0, 0.0);
        if (jj - b >= 1) v = Y[jj - b - 1];
        W[b * B1 + e] = v;
      }
      __syncthreads();
      if (tid < 64) {
        for (int j = j1; j >= j0; --j) {
          const double2* Uc = W + (j - j0) * B1;
          const double2 yj = Bv[j % B1];
          const bool go = bz_nz(yj);
          double2 xj = yj;
          if (go) xj = bz_div(yj, Uc[0]);
          const int om = min(b, j - 1);
          for (int o = 1 + tid; o <= om; o += 64) {
            const int sr = (j - o) % B1;
            double2 v;
            if (o == b) v = W[b * B1 + (j - j0)];
            else v = Bv[sr];
            if (go) {
              const double2 u = bz_mul(xj, Uc[o]);
              v = make_double2(v.x - u.x, v.y - u.y);
            }
            Bv[sr] = v;
          }
          if (tid == 0) Y[j - 1] = xj;
          wave_sync();
        }
      }
    }
    __syncthreads();
    // Esol into the edge field (mode 2: x-edges of x-slice s; mode 1: y-edges
    // of y-slice s); the bottom face carries the zeroed xZMax / yZMax
    const int nzp = a.nz + 1, nhp = Ny + 1;
    double2* E = a.e + (int64_t)S.sys * a.nE;
    for (int e = tid; e < nhp * nzp; e += nt) {
      const int iz = 1 + e % nzp, ih = 1 + e / nzp;
      double2 v;
      if (iz == nzp) v = make_double2(0.0, 0.0);
      else if (ih == 1) v = exb[iz];
      else if (ih == nhp) v = exb[a.nz + 2 * Ny - 1 + iz];
      else if (iz == 1) v = exb[a.nz + ih];
      else v = Y[(ih - 2) * b + (iz - 2)];
      const int64_t idx =
          S.mode == 2 ? (S.s - 1) + int64_t(a.nx) * ((ih - 1) + int64_t(a.ny + 1) * (iz - 1))
                      : a.nXe + (ih - 1) + int64_t(a.nx + 1) * ((S.s - 1) + int64_t(a.ny) * (iz - 1));
      E[idx] = v;
    }
    __syncthreads();
  }
}

}  // namespace mt
