// BiCGStab, operator and preconditioner kernels of libmtsp2.
//
// One BiCGStab iteration (right-preconditioned, conjugated inner products:
// the algorithm of cuBiCG, solver.f90:1548-1902) is 5 fused streaming kernels,
// 3 one-block finalize kernels and the preconditioner:
//   k_pupdate     restart ? (rt = r, p = r) : p = r + beta (p - omega v)   [+ Jacobi]
//   (precond)     ph = M^-1 p
//   k_spmv_dot<1> v = A ph, partials of <rt,v>      -> k_finalize<FIN_ALPHA>
//   k_supdate     x += alpha ph, s = r - alpha v  [+ best-iterate copy, Jacobi]
//   (precond)     sh = M^-1 s
//   k_spmv_dot<2> t = A sh, partials of <t,t>, <t,s> -> k_finalize<FIN_OMEGA>
//   k_rupdate     x += omega sh, r = s - omega t, partials of <rt,r>, <r,r>
//                                                    -> k_finalize<FIN_RUPDATE>
//   (mixed inner solves with DILU: k_supdate is folded into the
//    preconditioner input, the <t,s> dot and k_rupdate_fused)
//   k_recheck     only systems whose recursive residual met tol:
//                 r = b - A x (solver.f90:1855-1895)  -> k_finalize<FIN_RECHECK>
// The finalize kernels hold all solver scalar logic (convergence, breakdown,
// best iterate, restart, next rho/beta) on the device, so nothing is read back
// per dot.  Every kernel takes the active-system list and skips finished
// systems, so a fixed chunk of iterations is captured once in a hipGraph.
#pragma once

#include "mt_device.hpp"

namespace mt {

constexpr int BS = 256;
enum : int { PCK_DILU = 0, PCK_JACOBI = 1, PCK_NONE = 2, PCK_DILU_RB = 3 };
enum : int { FIN_INIT = 0, FIN_ALPHA = 1, FIN_OMEGA = 2, FIN_RUPDATE = 3, FIN_RECHECK = 4 };

template <class T, class V>
struct IterArgs {
  StencilView<V> A;
  SysScal* sc;
  const int* act;
  int nact, G;
  int pc;
  int64_t nE;
  C2<T>* x; C2<T>* r; C2<T>* rt; C2<T>* p; C2<T>* ph; C2<T>* v;
  C2<T>* s; C2<T>* sh; C2<T>* t; C2<T>* xbest;
  const C2<T>* b;
  const C2<T>* dinv;     // Jacobi inverse diagonal per frequency [nfreq][nE]
  Z* partials;           // [nact][G][2]
  double* rerr;          // [nsys][rerr_len]
  int rerr_len;
  double tol;
  // fused = 1 (inner solves of the mixed solver with a DILU preconditioner, no
  // best iterate): s = r - alpha v is never stored; the preconditioner, the
  // <t,s> dot and k_rupdate_fused recompute it
  int fused = 0;
  // s_ok = 1 (fused, brick red-black DILU): the preconditioner's first touch
  // stores s = r - alpha v in s, which the <t,s> dot and k_rupdate_fused read
  int s_ok = 0;
  // inner solves on bricks with paired x: ph and sh stored with the two
  // systems of a pair interleaved, entry (sys/2)*ppair + 2*point + sys%2
  // (ppair = twice the points per system; 0: ph and sh system-major)
  C2<T>* phi = nullptr;
  C2<T>* shi = nullptr;
  int64_t ppair = 0;
};

template <class T, class V>
__device__ __forceinline__ C2<T> apply_row(const StencilView<V>& A, const RowInfo& ri,
                                           const C2<T>* __restrict__ x, T dscale) {
  C2<T> y = stencil_row<V, T>(A, ri, x);
  // i*ISIGN*omega*vms1*x == (ISIGN*omega*vms1) * (i x)
  const T d = dscale * T(A.vms1[ri.e]);
  const C2<T> xe = x[ri.e];
  y.x += -d * xe.y;
  y.y += d * xe.x;
  return y;
}

// Same with the component C known at compile time.
template <int C, class T, class V>
__device__ __forceinline__ C2<T> apply_row_c(const StencilView<V>& A, const RowInfo& ri,
                                             const C2<T>* __restrict__ x, T dscale) {
  C2<T> y = stencil_row_c<C, V, T>(A, ri, x);
  const T d = dscale * T(A.vms1[ri.e]);
  const C2<T> xe = x[ri.e];
  y.x += -d * xe.y;
  y.y += d * xe.x;
  return y;
}

__device__ __forceinline__ void store_partials(Z* partials, int si, int G, int g, Z (&acc)[2],
                                               Z* lds) {
  block_sum<2, BS>(acc, lds);
  if (threadIdx.x == 0) {
    partials[((int64_t)si * G + g) * 2 + 0] = acc[0];
    partials[((int64_t)si * G + g) * 2 + 1] = acc[1];
  }
}

// ------------------------------------------------------------ scalar logic
__device__ inline void prep_next(SysScal& sc, Z rt_r, Z r_r) {
  sc.restart_due = (sc.iter % sc.restart) == 0;
  if (sc.restart_due) {
    sc.rho1 = {1.0, 0.0};
    sc.rho = r_r;
    sc.omg = {1.0, 0.0};
  } else {
    sc.rho1 = sc.rho;
    sc.rho = rt_r;
    sc.beta = zmul(zdiv(sc.rho, sc.rho1), zdiv(sc.alpha, sc.omg));
    if (sc.beta.x == 0.0 && sc.beta.y == 0.0) sc.status = ST_BREAK;
  }
}

__device__ inline void fin_rupdate(SysScal& sc, Z rt_r, Z r_r, double* rerr) {
  const double rn = sqrt(r_r.x);
  const int it = ++sc.iter;
  if (rerr) rerr[it - 1] = rn / sc.bnorm;
  sc.rnorm = rn;
  if (rn < sc.btol) {
    sc.status = sc.inner ? ST_CONV : ST_RECHECK;
    return;
  }
  if (rn < sc.rnorm_best) { sc.rnorm_best = rn; sc.best_due = 1; }
  if (it >= sc.maxit) { sc.status = ST_MAXIT; return; }
  prep_next(sc, rt_r, r_r);
}

__device__ inline void fin_recheck(SysScal& sc, Z rt_r, Z r_r, double* rerr) {
  const double rn = sqrt(r_r.x);
  sc.rnorm = rn;
  if (rn < sc.btol) {
    if (rerr) rerr[sc.iter - 1] = rn / sc.bnorm;
    sc.status = ST_CONV;
    return;
  }
  sc.status = ST_RUN;
  if (rn < sc.rnorm_best) { sc.rnorm_best = rn; sc.best_due = 1; }
  if (sc.iter >= sc.maxit) { sc.status = ST_MAXIT; return; }
  prep_next(sc, rt_r, r_r);
}

__device__ inline void fin_init(SysScal& sc, Z rr, Z bb, double tol, double* rerr) {
  const double rn = sqrt(rr.x), bn = sqrt(bb.x);
  sc.bnorm = bn;
  sc.btol = tol * bn;
  sc.rnorm = rn;
  sc.rnorm_start = rn;
  sc.rnorm_best = rn;
  sc.iter = 0;
  sc.best_due = 0;
  sc.alpha = {1.0, 0.0};
  sc.omg = {1.0, 0.0};
  sc.rho1 = {1.0, 0.0};
  sc.rho = rr;
  sc.restart_due = 1;
  if (rerr) rerr[0] = (bn > 0.0) ? rn / bn : 0.0;
  if (isnan(bn)) { sc.status = ST_BREAK; sc.iter = -1; return; }
  if (bn == 0.0) { sc.status = ST_CONV; sc.iter = 1; return; }
  if (rn <= sc.btol) { sc.status = ST_CONV; sc.iter = 1; return; }
  sc.status = ST_RUN;
}

// One block per active system (grid = nact): reduce that system's partials of
// the kernel just run, in fixed order, and update its solver scalars.
template <int KIND>
__global__ __launch_bounds__(BS) void k_finalize(SysScal* scs, const int* act, int nact, int G,
                                                 const Z* partials, double* rerr_base,
                                                 int rerr_len, double tol) {
  __shared__ Z lds[2 * BS / 64];
  const int q = blockIdx.x;
  Z red[2];
  reduce_sys<2, BS>(partials, q, G, red, lds);
  if (threadIdx.x != 0) return;
  const int s = act[q];
  SysScal& sc = scs[s];
  double* rerr = rerr_base ? rerr_base + (int64_t)s * rerr_len : nullptr;
  const Z a0 = red[0], a1 = red[1];
  if (KIND == FIN_INIT) {
    fin_init(sc, a0, a1, tol, rerr);
  } else if (KIND == FIN_ALPHA) {
    if (sc.status != ST_RUN) return;
    sc.alpha = zdiv(sc.rho, a0);
    // <rt, v> = 0 (the CPU BICG's check, solver.f90:720-926) or an alpha
    // that is zero or not finite: breakdown
    if ((a0.x == 0.0 && a0.y == 0.0) || (sc.alpha.x == 0.0 && sc.alpha.y == 0.0) ||
        !isfinite(sc.alpha.x) || !isfinite(sc.alpha.y))
      sc.status = ST_BREAK;
  } else if (KIND == FIN_OMEGA) {
    if (sc.status != ST_RUN) return;
    sc.omg = zdiv(a1, a0);
    sc.best_due = 0;  // consumed by k_supdate of this iteration
    if ((a0.x == 0.0 && a0.y == 0.0) || (sc.omg.x == 0.0 && sc.omg.y == 0.0) ||
        !isfinite(sc.omg.x) || !isfinite(sc.omg.y))
      sc.status = ST_BREAK;
  } else if (KIND == FIN_RUPDATE) {
    if (sc.status != ST_RUN) return;
    fin_rupdate(sc, a0, a1, rerr);
  } else if (KIND == FIN_RECHECK) {
    if (sc.status != ST_RECHECK) return;
    fin_recheck(sc, a0, a1, rerr);
  }
}

// y = d .* x over n complex entries (the Jacobi preconditioner apply)
__global__ __launch_bounds__(BS) void k_diag_mul(const Z* __restrict__ d, const Z* __restrict__ x,
                                                 Z* __restrict__ y, int64_t n) {
  for (int64_t q = blockIdx.x * int64_t(BS) + threadIdx.x; q < n; q += int64_t(gridDim.x) * BS)
    y[q] = zmul(d[q], x[q]);
}

// ------------------------------------------------------------------ kernels
// r = b - A x, partials <r,r>, <b,b> (then FIN_INIT); xbest = x0.
template <class T, class V>
__global__ __launch_bounds__(BS) void k_init(IterArgs<T, V> a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const int64_t o = (int64_t)sys * a.nE;
  const T dsc = T(kIsignD * a.sc[sys].freq_omega);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  const C2<T>* __restrict__ x = a.x + o;
  const C2<T>* __restrict__ b = a.b + o;
  C2<T>* __restrict__ r = a.r + o;
  C2<T>* __restrict__ xb = a.xbest ? a.xbest + o : nullptr;
  int64_t lo, hi;
  chunk_range3(a.A.nXi, a.A.nYi, a.A.nInt, a.G, g, lo, hi);
  for_rows(a.A, lo, hi, threadIdx.x, BS, [&](const RowInfo& ri, auto cc) {
    constexpr int C = decltype(cc)::value;
    const C2<T> bv = b[ri.e];
    const C2<T> y = csub(bv, apply_row_c<C, T, V>(a.A, ri, x, dsc));
    r[ri.e] = y;
    if (xb) xb[ri.e] = x[ri.e];
    acc[0].x += double(y.x) * y.x + double(y.y) * y.y;
    acc[1].x += double(bv.x) * bv.x + double(bv.y) * bv.y;
  });
  store_partials(a.partials, si, a.G, g, acc, lds);
}

template <class T, class V>
__global__ __launch_bounds__(BS) void k_pupdate(IterArgs<T, V> a) {
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc = a.sc[sys];
  if (sc.status != ST_RUN) return;
  const int64_t o = (int64_t)sys * a.nE;
  const bool rs = sc.restart_due;
  const C2<T> beta = ccast<T>(sc.beta), om = ccast<T>(sc.omg);
  const C2<T>* dv = (a.pc == PCK_JACOBI) ? a.dinv + (int64_t)sc.fidx * a.nE : nullptr;
  int64_t lo, hi;
  chunk_range(a.nE, a.G, g, lo, hi);
  for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
    const C2<T> r = a.r[o + e];
    C2<T> p;
    if (rs) {
      a.rt[o + e] = r;
      p = r;
    } else {
      p = cadd(r, cmul(beta, csub(a.p[o + e], cmul(om, a.v[o + e]))));
    }
    a.p[o + e] = p;
    if (dv) a.ph[o + e] = cmul(dv[e], p);
    else if (a.pc == PCK_NONE) a.ph[o + e] = p;
  }
}

// WHICH=1: v = A ph, <rt,v>.   WHICH=2: t = A sh, <t,t>, <t,s>.
template <class T, class V, int WHICH>
__global__ __launch_bounds__(BS) void k_spmv_dot(IterArgs<T, V> a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc0 = a.sc[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (sc0.status == ST_RUN) {
    const int64_t o = (int64_t)sys * a.nE;
    const T dsc = T(kIsignD * sc0.freq_omega);
    const C2<T>* __restrict__ in = (WHICH == 1 ? a.ph : a.sh) + o;
    C2<T>* __restrict__ out = (WHICH == 1 ? a.v : a.t) + o;
    // dot partner: rt, or s = r - alpha v (stored, or recomputed when fused)
    const bool fz = WHICH == 2 && a.fused;
    const C2<T>* __restrict__ pw = (WHICH == 1 ? a.rt : (fz ? a.r : a.s)) + o;
    const C2<T>* __restrict__ pv = a.v + o;
    const C2<T> al = ccast<T>(sc0.alpha);
    int64_t lo, hi;
    chunk_range3(a.A.nXi, a.A.nYi, a.A.nInt, a.G, g, lo, hi);
    for_rows(a.A, lo, hi, threadIdx.x, BS, [&](const RowInfo& ri, auto cc) {
      constexpr int C = decltype(cc)::value;
      C2<T> w = pw[ri.e];
      if (fz) w = csub(w, cmul(al, pv[ri.e]));
      const C2<T> y = apply_row_c<C, T, V>(a.A, ri, in, dsc);
      out[ri.e] = y;
      if (WHICH == 1) {
        acc[0].x += double(w.x) * y.x + double(w.y) * y.y;
        acc[0].y += double(w.x) * y.y - double(w.y) * y.x;
      } else {
        acc[0].x += double(y.x) * y.x + double(y.y) * y.y;
        acc[1].x += double(y.x) * w.x + double(y.y) * w.y;
        acc[1].y += double(y.x) * w.y - double(y.y) * w.x;
      }
    });
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

template <class T, class V>
__global__ __launch_bounds__(BS) void k_supdate(IterArgs<T, V> a) {
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc = a.sc[sys];
  if (sc.status != ST_RUN) return;
  const int64_t o = (int64_t)sys * a.nE;
  const C2<T> al = ccast<T>(sc.alpha);
  const bool best = sc.best_due && a.xbest;
  const C2<T>* dv = (a.pc == PCK_JACOBI) ? a.dinv + (int64_t)sc.fidx * a.nE : nullptr;
  int64_t lo, hi;
  chunk_range(a.nE, a.G, g, lo, hi);
  for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
    const C2<T> xo = a.x[o + e];
    if (best) a.xbest[o + e] = xo;
    a.x[o + e] = cadd(xo, cmul(al, a.ph[o + e]));
    const C2<T> sv = csub(a.r[o + e], cmul(al, a.v[o + e]));
    a.s[o + e] = sv;
    if (dv) a.sh[o + e] = cmul(dv[e], sv);
    else if (a.pc == PCK_NONE) a.sh[o + e] = sv;
  }
}

template <class T, class V>
__global__ __launch_bounds__(BS) void k_rupdate(IterArgs<T, V> a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc0 = a.sc[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (sc0.status == ST_RUN) {
    const int64_t o = (int64_t)sys * a.nE;
    const C2<T> om = ccast<T>(sc0.omg);
    int64_t lo, hi;
    chunk_range(a.nE, a.G, g, lo, hi);
    for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
      a.x[o + e] = cadd(a.x[o + e], cmul(om, a.sh[o + e]));
      const C2<T> r = csub(a.s[o + e], cmul(om, a.t[o + e]));
      a.r[o + e] = r;
      const C2<T> w = a.rt[o + e];
      acc[0].x += double(w.x) * r.x + double(w.y) * r.y;
      acc[0].y += double(w.x) * r.y - double(w.y) * r.x;
      acc[1].x += double(r.x) * r.x + double(r.y) * r.y;
    }
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// Fused form for the mixed inner solves: x += alpha ph + omega sh,
// r = (r - alpha v) - omega t, partials of <rt,r>, <r,r>.  The same
// operations, in the same order, as k_supdate followed by k_rupdate; with
// s_ok, s = r - alpha v is the one the preconditioner stored.
template <class T, class V>
__global__ __launch_bounds__(BS) void k_rupdate_fused(IterArgs<T, V> a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc0 = a.sc[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (sc0.status == ST_RUN) {
    const int64_t o = (int64_t)sys * a.nE;
    const C2<T> al = ccast<T>(sc0.alpha), om = ccast<T>(sc0.omg);
    C2<T>* __restrict__ x = a.x + o;
    C2<T>* __restrict__ r = a.r + o;
    // ph, sh system-major, or pair-interleaved (every second entry)
    const int64_t po = a.ppair ? (int64_t)(sys >> 1) * a.ppair + (sys & 1) : o;
    const int ps = a.ppair ? 2 : 1;
    const C2<T>* __restrict__ ph = (a.ppair ? a.phi : a.ph) + po;
    const C2<T>* __restrict__ sh = (a.ppair ? a.shi : a.sh) + po;
    const C2<T>* __restrict__ v = a.v + o;
    const C2<T>* __restrict__ t = a.t + o;
    const C2<T>* __restrict__ rt = a.rt + o;
    const C2<T>* __restrict__ sv = a.s + o;
    const bool sk = a.s_ok;
    int64_t lo, hi;
    chunk_range(a.nE, a.G, g, lo, hi);
#pragma unroll 2
    for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
      x[e] = cadd(cadd(x[e], cmul(al, ph[ps * e])), cmul(om, sh[ps * e]));
      const C2<T> se = sk ? sv[e] : csub(r[e], cmul(al, v[e]));
      const C2<T> rn = csub(se, cmul(om, t[e]));
      r[e] = rn;
      const C2<T> w = rt[e];
      acc[0].x += double(w.x) * rn.x + double(w.y) * rn.y;
      acc[0].y += double(w.x) * rn.y - double(w.y) * rn.x;
      acc[1].x += double(rn.x) * rn.x + double(rn.y) * rn.y;
    }
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// True-residual recheck for systems in ST_RECHECK (no-op for the others).
template <class T, class V>
__global__ __launch_bounds__(BS) void k_recheck(IterArgs<T, V> a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& sc0 = a.sc[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (sc0.status == ST_RECHECK) {
    const int64_t o = (int64_t)sys * a.nE;
    const T dsc = T(kIsignD * sc0.freq_omega);
    const C2<T>* __restrict__ x = a.x + o;
    const C2<T>* __restrict__ b = a.b + o;
    const C2<T>* __restrict__ rt = a.rt + o;
    C2<T>* __restrict__ r = a.r + o;
    int64_t lo, hi;
    chunk_range3(a.A.nXi, a.A.nYi, a.A.nInt, a.G, g, lo, hi);
    for_rows(a.A, lo, hi, threadIdx.x, BS, [&](const RowInfo& ri, auto cc) {
      constexpr int C = decltype(cc)::value;
      const C2<T> w = rt[ri.e];
      const C2<T> y = csub(b[ri.e], apply_row_c<C, T, V>(a.A, ri, x, dsc));
      r[ri.e] = y;
      acc[0].x += double(w.x) * y.x + double(w.y) * y.y;
      acc[0].y += double(w.x) * y.y - double(w.y) * y.x;
      acc[1].x += double(y.x) * y.x + double(y.y) * y.y;
    });
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// ---------------------------------------------------------- plain operator
// y = A(omega) x for one system (mt_apply, tests).
template <class T, class V>
__global__ __launch_bounds__(BS) void k_apply(StencilView<V> A, const C2<T>* x, C2<T>* y,
                                              double omega) {
  const T dsc = T(kIsignD * omega);
  const C2<T>* __restrict__ xr = x;
  C2<T>* __restrict__ yr = y;
  for_rows(A, 0, A.nInt, (int64_t)blockIdx.x * BS + threadIdx.x, (int64_t)gridDim.x * BS,
           [&](const RowInfo& ri, auto cc) {
             constexpr int C = decltype(cc)::value;
             yr[ri.e] = apply_row_c<C, T, V>(A, ri, xr, dsc);
           });
}

// --------------------------------------------------- interior <-> padded
template <class T, class U>
__global__ void k_scatter(const int* __restrict__ map, int64_t nInt, int64_t nE, int nsys,
                          const C2<U>* __restrict__ in, C2<T>* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nInt * nsys;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = q / nInt, r = q - s * nInt;
    out[s * nE + map[r]] = ccast<T>(in[q]);
  }
}
template <class T, class U>
__global__ void k_gather(const int* __restrict__ map, int64_t nInt, int64_t nE, int nsys,
                         const C2<U>* __restrict__ in, C2<T>* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nInt * nsys;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = q / nInt, r = q - s * nInt;
    out[q] = ccast<T>(in[s * nE + map[r]]);
  }
}

// Jacobi: dinv_f[e] = 1 / (a_ee + i*ISIGN*omega_f*vms1_e) on interior rows.
template <class T, class V>
__global__ void k_jacobi_dinv(StencilView<V> A, const double* omegas, int nfreq, C2<T>* dinv) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < A.nInt * nfreq;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t f = q / A.nInt, rr = q - f * A.nInt;
    const RowInfo ri = decode_row(A, rr);
    const Z d = {double(A.coef[ri.e]), kIsignD * omegas[f] * double(A.vms1[ri.e])};
    const Z iv = zdiv({1.0, 0.0}, d);
    dinv[f * A.nE + ri.e] = {T(iv.x), T(iv.y)};
  }
}

// ------------------------------------------------ mixed precision (outer)
// FP64 defect correction around the FP32 BiCGStab: the outer record `so`
// holds the FP64 state of each system, the inner record `si` the FP32 solve
// of A d = r_scaled.
struct OuterArgs {
  StencilView<double> A;
  SysScal* so;
  SysScal* si;
  const int* act;
  int nact, G;
  int64_t nE;
  const Z* b;
  Z* x;
  C2<float>* r32;
  C2<float>* d32;
  Z* partials;
  double tol, inner_tol;
  int max_outer, maxit_total;
  double* rerr;           // outer relative residual history [nsys][rerr_len]
  int rerr_len;
};

// partials of <b,b> (then k_outer_fin<0>)
__global__ __launch_bounds__(BS) void k_outer_bnorm(OuterArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const int64_t o = (int64_t)sys * a.nE;
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  int64_t lo, hi;
  chunk_range(a.nE, a.G, g, lo, hi);
  for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
    const Z bv = a.b[o + e];
    acc[0].x += bv.x * bv.x + bv.y * bv.y;
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// r = b - A x (FP64) -> r32 = scale*r, d32 = 0; partials of <r,r> (then k_outer_fin<1>)
__global__ __launch_bounds__(BS) void k_outer_resid(OuterArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& so0 = a.so[sys];
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (so0.status == ST_RUN) {
    const int64_t o = (int64_t)sys * a.nE;
    const double dsc = kIsignD * so0.freq_omega, sc = so0.scale;
    const Z* __restrict__ x = a.x + o;
    const Z* __restrict__ b = a.b + o;
    C2<float>* __restrict__ r32 = a.r32 + o;
    C2<float>* __restrict__ d32 = a.d32 + o;
    int64_t lo, hi;
    chunk_range3(a.A.nXi, a.A.nYi, a.A.nInt, a.G, g, lo, hi);
    for_rows(a.A, lo, hi, threadIdx.x, BS, [&](const RowInfo& ri, auto cc) {
      constexpr int C = decltype(cc)::value;
      const Z y = csub(b[ri.e], apply_row_c<C, double, double>(a.A, ri, x, dsc));
      r32[ri.e] = {float(sc * y.x), float(sc * y.y)};
      d32[ri.e] = {0.0f, 0.0f};
      acc[0].x += y.x * y.x + y.y * y.y;
    });
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// One block per active system (grid = nact).
// STAGE 0: bnorm, btol, FP32 scaling (a power of two near 1/||b||).
// STAGE 1: outer convergence and, for systems still refining, the inner set-up.
template <int STAGE>
__global__ __launch_bounds__(BS) void k_outer_fin(OuterArgs a) {
  __shared__ Z lds[2 * BS / 64];
  const int q = blockIdx.x;
  Z red[2];
  reduce_sys<2, BS>(a.partials, q, a.G, red, lds);
  if (threadIdx.x != 0) return;
  {
    const int s = a.act[q];
    SysScal& so = a.so[s];
    if (STAGE == 0) {
      const double bn = sqrt(red[0].x);
      so.bnorm = bn;
      so.btol = a.tol * bn;
      so.scale = (bn > 0.0 && isfinite(bn)) ? ldexp(1.0, -ilogb(bn)) : 1.0;
      so.iter = 0;
      so.tot_inner = 0;
      so.status = isnan(bn) ? ST_BREAK : (bn == 0.0 ? ST_CONV : ST_RUN);
      if (isnan(bn)) so.iter = -1;
      return;
    }
    SysScal& sn = a.si[s];
    if (so.status != ST_RUN) { sn.status = ST_CONV; return; }
    const double rn = sqrt(red[0].x);
    so.rnorm = rn;
    if (a.rerr && so.iter < a.rerr_len) a.rerr[(int64_t)s * a.rerr_len + so.iter] = rn / so.bnorm;
    so.iter += 1;
    if (rn < so.btol) { so.status = ST_CONV; sn.status = ST_CONV; return; }
    if (so.iter > a.max_outer || so.tot_inner >= a.maxit_total) {
      so.status = ST_MAXIT; sn.status = ST_CONV; return;
    }
    const double bi = rn * so.scale;
    sn.bnorm = bi;
    // ask the inner solve for the reduction still needed, at least inner_tol
    const double need = 0.5 * so.btol / rn;
    sn.btol = bi * fmax(a.inner_tol, need);
    sn.rnorm = bi;
    sn.rnorm_start = bi;
    sn.rnorm_best = bi;
    sn.iter = 0;
    sn.maxit = a.maxit_total - so.tot_inner;
    sn.best_due = 0;
    sn.restart_due = 1;
    sn.restart = so.restart;
    sn.rho = {bi * bi, 0.0};
    sn.rho1 = {1.0, 0.0};
    sn.alpha = {1.0, 0.0};
    sn.omg = {1.0, 0.0};
    sn.inner = 1;
    sn.fidx = so.fidx;
    sn.freq_omega = so.freq_omega;
    sn.status = ST_RUN;
  }
}

// x += d / scale for systems still refining; count inner iterations.
__global__ __launch_bounds__(BS) void k_outer_update(OuterArgs a) {
  int si, g;
  map_sys_block(a.nact, a.G, si, g);
  const int sys = a.act[si];
  const SysScal& so = a.so[sys];
  if (so.status != ST_RUN) return;
  const int64_t o = (int64_t)sys * a.nE;
  const double inv = 1.0 / so.scale;
  int64_t lo, hi;
  chunk_range(a.nE, a.G, g, lo, hi);
  for (int64_t e = lo + threadIdx.x; e < hi; e += BS) {
    const C2<float> d = a.d32[o + e];
    Z x = a.x[o + e];
    x.x += inv * double(d.x);
    x.y += inv * double(d.y);
    a.x[o + e] = x;
  }
  if (g == 0 && threadIdx.x == 0) a.so[sys].tot_inner += a.si[sys].iter;
}

}  // namespace mt
