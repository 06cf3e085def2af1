// Data functionals on the device: impedance responses (dataResp,
// DataFunc.f90:43-462) and the adjoint source of a data residual (LmultT,
// SENS/DataSens.f90:130-225, with Lrows DataFunc.f90:465-869) for all
// (frequency, site) pairs of a batch at once.
//
// The site rows (E and H interpolation over edges, mt_sites.hpp) are fixed per
// receiver set.  Every linearised functional of the reference is a linear
// combination of the seven base rows of its site (Ex, Ey, Hx, Hy, Hz at the
// site, Hx, Hy at the reference site), so a functional is carried here as its
// 7 complex coefficients over those rows instead of a merged sparse vector:
//   k_site_eval     one thread per (frequency, site): the 14 base values
//                   (7 rows x 2 polarisations), the responses, and for LmultT
//                   the per-mode coefficient vector sum_n d_n L_n(k);
//   k_lmultT_gather one thread per (frequency, distinct edge): the adjoint
//                   source on that edge, summed over the (site, row) items in
//                   a fixed order -- deterministic, no atomics.
#pragma once

#include "mt_device.hpp"
#include "mt_internal.h"
#include "mt_sites.hpp"

namespace mt {

using Cd = C2<double>;

__device__ __forceinline__ Cd cdv(Cd a, Cd b) {
  // a / b, scaled as in Smith's algorithm
  if (fabs(b.x) >= fabs(b.y)) {
    const double r = b.y / b.x, den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  }
  const double r = b.x / b.y, den = b.x * r + b.y;
  return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}
__device__ __forceinline__ Cd cscl(double s, Cd a) { return {s * a.x, s * a.y}; }
__device__ __forceinline__ Cd cconj(Cd a) { return {a.x, -a.y}; }
__device__ __forceinline__ double cabs2(Cd a) { return a.x * a.x + a.y * a.y; }

struct SiteArgs {
  int nsite, nfreq, dtype, ncomp, want_coef;
  const double* omegas;          // [nfreq]
  const Cd* e;                   // [nfreq][2][nE] all-edge fields
  int64_t nE;
  const int* ptr;                // [nsite*7+1]
  const int* idx;                // all-edge indices
  const double* val;             // coefficients (H rows without ISIGN*i/omega)
  const double* rot;             // [nsite][3][8]
  double* resp;                  // [nfreq][nsite][ncomp] or null
  const double* d;               // [nfreq][nsite][ncomp] data residual or null
  const unsigned char* exist;    // same shape, or null (all present)
  Cd* coef;                      // [nfreq][nsite][2][7] or null
};

// 7-vector over the base rows
struct V7 {
  Cd c[7];
};
__device__ __forceinline__ V7 v7_zero() {
  V7 v;
#pragma unroll
  for (int b = 0; b < 7; ++b) v.c[b] = {0.0, 0.0};
  return v;
}
__device__ __forceinline__ V7 v7_unit(int b) {
  V7 v = v7_zero();
  v.c[b] = {1.0, 0.0};
  return v;
}
// a*x + b*y
__device__ __forceinline__ V7 v7_lin(Cd a, const V7& x, Cd b, const V7& y) {
  V7 v;
#pragma unroll
  for (int q = 0; q < 7; ++q) v.c[q] = cadd(cmul(a, x.c[q]), cmul(b, y.c[q]));
  return v;
}

__device__ __forceinline__ void rot2(const double* R, Cd& c1, Cd& c2) {
  const Cd a = c1, b = c2;
  c1 = cadd(cscl(R[0], a), cscl(R[1], b));
  c2 = cadd(cscl(R[2], a), cscl(R[3], b));
}
// 2x2 inverse by Kramer's rule, M[row][mode]
__device__ __forceinline__ void inv2(Cd M[2][2]) {
  const Cd det = csub(cmul(M[0][0], M[1][1]), cmul(M[0][1], M[1][0]));
  const Cd t = M[0][0];
  M[0][0] = cdv(M[1][1], det);
  M[1][1] = cdv(t, det);
  M[0][1] = cdv({-M[0][1].x, -M[0][1].y}, det);
  M[1][0] = cdv({-M[1][0].x, -M[1][0].y}, det);
}

__global__ __launch_bounds__(64) void k_site_eval(SiteArgs a) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= a.nfreq * a.nsite) return;
  const int f = t / a.nsite, s = t - f * a.nsite, dt = a.dtype;
  const double om = a.omegas[f];
  const Cd iw = {0.0, kIsignD / om};   // ISIGN*i/omega
  // which rows this data type reads
  const bool needE = dt == 1 || dt == 2 || dt == 5 || dt == 6 || dt == 7 || dt == 8;
  const bool needH = dt <= 6 || dt == 9 || dt == 10;
  const bool needZ = dt == 3 || dt == 11;
  const bool needR = dt == 4;
  Cd v[7][2];
#pragma unroll
  for (int b = 0; b < 7; ++b) {
    v[b][0] = v[b][1] = {0.0, 0.0};
    const bool need = b <= kSiteEy ? needE : b <= kSiteHy ? needH : b == kSiteHz ? needZ : needR;
    if (!need) continue;
    const int row = s * kSiteRows + b;
    for (int m = 0; m < 2; ++m) {
      const Cd* __restrict__ e = a.e + ((int64_t)f * 2 + m) * a.nE;
      Cd acc = {0.0, 0.0};
      for (int q = a.ptr[row]; q < a.ptr[row + 1]; ++q) {
        const Cd x = e[a.idx[q]];
        acc.x += a.val[q] * x.x;
        acc.y += a.val[q] * x.y;
      }
      v[b][m] = b >= kSiteHx ? cmul(iw, acc) : acc;
    }
  }
  const double* rot = a.rot + (int64_t)s * 24;   // E, H, Href frames
  Cd Z[4] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  Cd Zl[4];            // the Z the linearisation uses
  Cd B[2][2];          // inverse of the horizontal H (or reference H) matrix
  double resp[8];
  int nf = 0;
  if (dt >= 7) {
    const int b = dt == 7 ? kSiteEx : dt == 8 ? kSiteEy : dt == 9 ? kSiteHx : dt == 10 ? kSiteHy : kSiteHz;
    Z[0] = v[b][0];
    nf = 1;
  } else {
    Cd EE[2][2], HH[2][2], HZ[2], RR[2][2];
    for (int m = 0; m < 2; ++m) {
      EE[0][m] = v[kSiteEx][m]; EE[1][m] = v[kSiteEy][m];
      rot2(rot + 0, EE[0][m], EE[1][m]);
      HH[0][m] = v[kSiteHx][m]; HH[1][m] = v[kSiteHy][m];
      rot2(rot + 8, HH[0][m], HH[1][m]);
      HZ[m] = v[kSiteHz][m];
      RR[0][m] = v[kSiteHrx][m]; RR[1][m] = v[kSiteHry][m];
      rot2(rot + 16, RR[0][m], RR[1][m]);
    }
    if (dt == 4) {
      inv2(RR);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) Z[2 * i + j] = cadd(cmul(HH[i][0], RR[0][j]), cmul(HH[i][1], RR[1][j]));
      Z[0].x -= 1.0;
      Z[3].x -= 1.0;
      for (int q = 0; q < 4; ++q) Zl[q] = Z[q];
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) B[i][j] = RR[i][j];
      nf = 4;
    } else {
      inv2(HH);
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) B[i][j] = HH[i][j];
      if (dt == 3) {
        Z[0] = cadd(cmul(HZ[0], HH[0][0]), cmul(HZ[1], HH[1][0]));
        Z[1] = cadd(cmul(HZ[0], HH[0][1]), cmul(HZ[1], HH[1][1]));
        Zl[0] = Z[0]; Zl[1] = Z[1];
        nf = 2;
      } else {
        Cd Zf[4];   // full impedance tensor
        for (int i = 0; i < 2; ++i)
          for (int j = 0; j < 2; ++j) Zf[2 * i + j] = cadd(cmul(EE[i][0], HH[0][j]), cmul(EE[i][1], HH[1][j]));
        for (int q = 0; q < 4; ++q) Zl[q] = Zf[q];
        if (dt == 1) {
          for (int q = 0; q < 4; ++q) Z[q] = Zf[q];
          nf = 4;
        } else if (dt == 2) {
          Z[0] = Zf[1]; Z[1] = Zf[2];
          nf = 2;
        } else if (dt == 5) {
          Z[0] = {log10(cabs2(Zf[1]) * kMu0 / om), 0.0};
          Z[1] = {atan2(kIsignD * Zf[1].y, Zf[1].x), 0.0};
          Z[2] = {log10(cabs2(Zf[2]) * kMu0 / om), 0.0};
          Z[3] = {atan2(kIsignD * Zf[2].y, Zf[2].x), 0.0};
          nf = 4;
        } else {
          const double r1 = Zf[0].x, r2 = Zf[1].x, r3 = Zf[2].x, r4 = Zf[3].x;
          const double m1 = Zf[0].y, m2 = Zf[1].y, m3 = Zf[2].y, m4 = Zf[3].y;
          const double det = r1 * r4 - r2 * r3;
          Z[0] = {kIsignD * (r4 * m1 - r2 * m3) / det, 0.0};
          Z[1] = {kIsignD * (r4 * m2 - r2 * m4) / det, 0.0};
          Z[2] = {kIsignD * (r1 * m3 - r3 * m1) / det, 0.0};
          Z[3] = {kIsignD * (r1 * m4 - r3 * m2) / det, 0.0};
          nf = 4;
        }
      }
    }
  }
  const bool cplx = dt != 5 && dt != 6;
  {
    int ic = 0;
    for (int q = 0; q < nf; ++q) {
      resp[ic++] = Z[q].x;
      if (cplx) resp[ic++] = Z[q].y;
    }
    if (a.resp)
      for (int q = 0; q < a.ncomp; ++q) a.resp[((int64_t)f * a.nsite + s) * a.ncomp + q] = resp[q];
  }
  if (!a.want_coef) return;
  // ---- coefficient vector of sum_n z_n L_n(k), per mode k
  const int64_t dbase = ((int64_t)f * a.nsite + s) * a.ncomp;
  Cd zc[4];
  bool ex[4];
  {
    int ic = 0;
    for (int n = 0; n < nf; ++n) {
      ex[n] = a.exist ? a.exist[dbase + ic] != 0 : true;
      if (cplx) {
        zc[n] = {a.d[dbase + ic], -a.d[dbase + ic + 1]};
        ic += 2;
      } else {
        zc[n] = {a.d[dbase + ic], 0.0};
        ic += 1;
      }
    }
  }
  // rotated base rows (Lrows, DataFunc.f90:626-656)
  const double* Re = rot + 4;
  const double* Rh = rot + 12;
  const double* Rr = rot + 20;
  const V7 uEx = v7_unit(kSiteEx), uEy = v7_unit(kSiteEy), uHx = v7_unit(kSiteHx),
           uHy = v7_unit(kSiteHy), uHz = v7_unit(kSiteHz), uRx = v7_unit(kSiteHrx),
           uRy = v7_unit(kSiteHry);
  const V7 Lexr = v7_lin({Re[0], 0.0}, uEx, {Re[1], 0.0}, uEy);
  const V7 Leyr = v7_lin({Re[2], 0.0}, uEx, {Re[3], 0.0}, uEy);
  const V7 Lbxr = v7_lin({Rh[0], 0.0}, uHx, {Rh[1], 0.0}, uHy);
  const V7 Lbyr = v7_lin({Rh[2], 0.0}, uHx, {Rh[3], 0.0}, uHy);
  const V7 Lrxr = v7_lin({Rr[0], 0.0}, uRx, {Rr[1], 0.0}, uRy);
  const V7 Lryr = v7_lin({Rr[2], 0.0}, uRx, {Rr[3], 0.0}, uRy);
  const Cd one = {1.0, 0.0}, zero = {0.0, 0.0};
  for (int k = 0; k < 2; ++k) {
    V7 acc = v7_zero();
    if (dt >= 7) {
      if (k == 0 && ex[0]) {
        const V7& u = dt == 7 ? uEx : dt == 8 ? uEy : dt == 9 ? uHx : dt == 10 ? uHy : uHz;
        acc = v7_lin(zc[0], u, zero, u);
      }
    } else {
      // rows of the impedance-type functionals: L_n(k) = B(k,j) (row_i - L1_i)
      const int nbase = (dt == 1 || dt == 4 || dt == 6) ? 4 : 2;
      V7 L[4];
      for (int n = 0; n < nbase; ++n) {
        int i, j, pc;
        if (nbase == 4) { i = n / 2 + 1; j = n % 2 + 1; pc = dt == 4 ? i + 3 : i; }
        else if (dt == 3) { i = 1; j = n + 1; pc = 3; }
        else { i = n + 1; j = 2 - n; pc = n + 1; }
        const Cd c1 = Zl[2 * (i - 1)], c2 = Zl[2 * (i - 1) + 1];
        const V7 L1 = dt == 4 ? v7_lin(c1, Lrxr, c2, Lryr) : v7_lin(c1, Lbxr, c2, Lbyr);
        const V7& row = pc == 1 ? Lexr : pc == 2 ? Leyr : pc == 3 ? uHz : pc == 4 ? Lbxr : Lbyr;
        const Cd b = B[k][j - 1];
        L[n] = v7_lin(b, row, {-b.x, -b.y}, L1);
      }
      if (dt == 5) {
        // log10 rho and phase of Zxy, Zyx (Lrows, DataFunc.f90:717-741)
        const double a2 = cabs2(Zl[1]), a3 = cabs2(Zl[2]), ln10 = log(10.0);
        const V7 Lxy = L[0], Lyx = L[1];
        L[3] = v7_lin(cscl(1.0 / a3, cmul({0.0, 1.0}, cconj(Zl[2]))), Lyx, zero, Lyx);
        L[2] = v7_lin(cscl(2.0 / a3 / ln10, cconj(Zl[2])), Lyx, zero, Lyx);
        L[1] = v7_lin(cscl(1.0 / a2, cmul({0.0, 1.0}, cconj(Zl[1]))), Lxy, zero, Lxy);
        L[0] = v7_lin(cscl(2.0 / a2 / ln10, cconj(Zl[1])), Lxy, zero, Lxy);
      } else if (dt == 6) {
        // phase tensor (Lrows, DataFunc.f90:743-840)
        const double r1 = Zl[0].x, r2 = Zl[1].x, r3 = Zl[2].x, r4 = Zl[3].x;
        const double m1 = Zl[0].y, m2 = Zl[1].y, m3 = Zl[2].y, m4 = Zl[3].y;
        const double det = r1 * r4 - r2 * r3, sg = kIsignD;
        const double P11 = sg * (r4 * m1 - r2 * m3) / det, P12 = sg * (r4 * m2 - r2 * m4) / det;
        const double P21 = sg * (r1 * m3 - r3 * m1) / det, P22 = sg * (r1 * m4 - r3 * m2) / det;
        auto re = [](double x) { return Cd{x, 0.0}; };
        auto im = [](double x) { return Cd{0.0, x}; };
        auto pt = [&](double x11, double x12, double x21, double x22, double y1, double y2, int ya,
                      int yb) {
          V7 t = v7_lin(re(x11), L[0], re(x12), L[1]);
          t = v7_lin(one, t, one, v7_lin(re(x21), L[2], re(x22), L[3]));
          return v7_lin(one, t, one, v7_lin(im(y1), L[ya], im(y2), L[yb]));
        };
        const V7 p11 = pt(-P11 * r4 / det, (P11 * r3 - sg * m3) / det, P11 * r2 / det,
                          (-P11 * r1 + sg * m1) / det, r4 / det, -r2 / det, 0, 2);
        const V7 p12 = pt(-P12 * r4 / det, (P12 * r3 - sg * m4) / det, P12 * r2 / det,
                          (-P12 * r1 + sg * m2) / det, r4 / det, -r2 / det, 1, 3);
        const V7 p21 = pt((-P21 * r4 + sg * m3) / det, P21 * r3 / det, (P21 * r2 - sg * m1) / det,
                          -P21 * r1 / det, -r3 / det, r1 / det, 0, 2);
        const V7 p22 = pt((-P22 * r4 + sg * m4) / det, P22 * r3 / det, (P22 * r2 - sg * m2) / det,
                          -P22 * r1 / det, -r3 / det, r1 / det, 1, 3);
        L[0] = p11; L[1] = p12; L[2] = p21; L[3] = p22;
      }
      for (int n = 0; n < nf; ++n)
        if (ex[n]) acc = v7_lin(one, acc, zc[n], L[n]);
    }
    Cd* out = a.coef + (((int64_t)f * a.nsite + s) * 2 + k) * 7;
#pragma unroll
    for (int b = 0; b < 7; ++b) out[b] = acc.c[b];
  }
}

// comb[f][k][uedge[u]] = sum over items (row, val): val * w_b * coef[f][s][k][b],
// w_b = 1 for E rows and ISIGN*i/omega_f for H rows.  Other edges stay as
// filled before.
__global__ __launch_bounds__(256) void k_lmultT_gather(int nfreq, int nu, int nsite, int64_t nE,
                                                      const double* __restrict__ omegas,
                                                      const int* __restrict__ uedge,
                                                      const int* __restrict__ gptr,
                                                      const int* __restrict__ grow,
                                                      const double* __restrict__ gval,
                                                      const Cd* __restrict__ coef, Cd* comb) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= (int64_t)nfreq * nu) return;
  const int f = int(t / nu), u = int(t - (int64_t)f * nu);
  const Cd iw = {0.0, kIsignD / omegas[f]};
  Cd acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  for (int q = gptr[u]; q < gptr[u + 1]; ++q) {
    const int row = grow[q], s = row / kSiteRows, b = row - s * kSiteRows;
    const Cd w = b >= kSiteHx ? cscl(gval[q], iw) : Cd{gval[q], 0.0};
    for (int k = 0; k < 2; ++k) acc[k] = cadd(acc[k], cmul(w, coef[(((int64_t)f * nsite + s) * 2 + k) * 7 + b]));
  }
  for (int k = 0; k < 2; ++k) comb[((int64_t)f * 2 + k) * nE + uedge[u]] = acc[k];
}

}  // namespace mt
