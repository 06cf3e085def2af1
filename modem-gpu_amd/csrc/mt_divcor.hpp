// Divergence-correction kernels (SdivCorr, EMsolve3D.f90:507-613): DivC on
// the padded node lattice, the node-Laplacian PCG (solver.f90:57-136) with
// its block... single-lattice DILU (the tile sweeps of mt_dilu.hpp), and the
// gradient update of the edge field.  FP64 throughout, one system.
#pragma once

#include "mt_kernels.hpp"

namespace mt {

enum : int { PF_INIT = 0, PF_DELTA = 1, PF_ALPHA = 2, PF_R = 3 };

struct NodeArgs {
  const double* coefN;  // [7][nN] VDsG stencil
  const double* wE;     // [nE] DualFaceA*sigma_E on interior edges, 0 elsewhere
  const double* invL;   // [nE] 1/EdgeL
  int64_t nN, nE, nXe, nYe, nXi, nYi, nInt;
  int nx, ny, nz;
  Z* e;                 // padded edge field
  Z *b, *x, *r, *s, *p, *q;  // padded node vectors
  const Z* phi0;        // padded node source term or null
  SysScal* sc;
  Z* partials;
  int G;
  double* rerr;
  double tol;
};

__device__ __forceinline__ int64_t node_addr(const NodeArgs& a, int64_t r, int& i, int& j, int& k) {
  const uint32_t ai = uint32_t(a.nx - 1), bi = uint32_t(a.ny - 1);
  uint32_t q = uint32_t(r);
  i = 1 + q % ai; q /= ai;
  j = 1 + q % bi;
  k = 1 + q / bi;
  return i + int64_t(a.nx + 1) * (j + int64_t(a.ny + 1) * k);
}

// VDs e at node (i,j,k): -sum_e G_en w_e e_e over its 6 edges
__device__ __forceinline__ Z divc_node(const NodeArgs& a, int i, int j, int k) {
  const int64_t nx = a.nx, ny = a.ny;
  const int64_t xe = i + nx * (j + (ny + 1) * k);
  const int64_t ye = a.nXe + i + (nx + 1) * (j + ny * k);
  const int64_t ze = a.nXe + a.nYe + i + (nx + 1) * (j + (ny + 1) * k);
  const int64_t sx = 1, sy = nx + 1, sz = (nx + 1) * (ny + 1);
  Z acc = {0.0, 0.0};
  auto add = [&](int64_t ed, double sgn) {
    const double w = sgn * a.wE[ed];
    acc.x += w * a.e[ed].x;
    acc.y += w * a.e[ed].y;
  };
  add(xe - sx, -1.0); add(xe, 1.0);   // -(+1) for the edge ending here, -(-1) starting here
  add(ye - sy, -1.0); add(ye, 1.0);
  add(ze - sz, -1.0); add(ze, 1.0);
  return acc;
}

// b = VDs e (- phi0); partials <b,b>
__global__ __launch_bounds__(BS) void k_divc(NodeArgs a, Z* out) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(1, a.G, si, g);
  const int64_t nI = int64_t(a.nx - 1) * (a.ny - 1) * (a.nz - 1);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  int64_t lo, hi;
  chunk_range(nI, a.G, g, lo, hi);
  for (int64_t rr = lo + threadIdx.x; rr < hi; rr += BS) {
    int i, j, k;
    const int64_t n = node_addr(a, rr, i, j, k);
    Z v = divc_node(a, i, j, k);
    if (a.phi0) v = csub(v, a.phi0[n]);
    out[n] = v;
    acc[0].x += v.x * v.x + v.y * v.y;
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

__device__ __forceinline__ Z node_apply(const NodeArgs& a, const Z* v, int64_t n) {
  const int64_t sx = 1, sy = a.nx + 1, sz = int64_t(a.nx + 1) * (a.ny + 1);
  const double* c = a.coefN + n;
  const int64_t N = a.nN;
  Z y = cscale(c[0], v[n]);
  const int64_t off[6] = {-sx, sx, -sy, sy, -sz, sz};
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const double cc = c[(s + 1) * N];
    const Z w = v[n + off[s]];
    y.x += cc * w.x;
    y.y += cc * w.y;
  }
  return y;
}

// PCG start (x = 0): r = b; partials <r,r>, <b,b>
__global__ __launch_bounds__(BS) void k_pcg_init(NodeArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(1, a.G, si, g);
  const int64_t nI = int64_t(a.nx - 1) * (a.ny - 1) * (a.nz - 1);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  int64_t lo, hi;
  chunk_range(nI, a.G, g, lo, hi);
  for (int64_t rr = lo + threadIdx.x; rr < hi; rr += BS) {
    int i, j, k;
    const int64_t n = node_addr(a, rr, i, j, k);
    const Z bv = a.b[n];
    a.r[n] = bv;
    a.x[n] = {0.0, 0.0};
    a.p[n] = {0.0, 0.0};
    a.q[n] = {0.0, 0.0};
    acc[0].x += bv.x * bv.x + bv.y * bv.y;
    acc[1].x += bv.x * bv.x + bv.y * bv.y;
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// partials <r, s> (conjugated, dot_product(r, s))
__global__ __launch_bounds__(BS) void k_pcg_rs(NodeArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(1, a.G, si, g);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (a.sc->status == ST_RUN) {
    const int64_t nI = int64_t(a.nx - 1) * (a.ny - 1) * (a.nz - 1);
    int64_t lo, hi;
    chunk_range(nI, a.G, g, lo, hi);
    for (int64_t rr = lo + threadIdx.x; rr < hi; rr += BS) {
      int i, j, k;
      const int64_t n = node_addr(a, rr, i, j, k);
      const Z w = cdotm(a.r[n], a.s[n]);
      acc[0] = cadd(acc[0], w);
    }
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// p = s + beta p ; q = A p = A s + beta q (linearity: neighbours of the new p
// are not needed); partials <p, q>
__global__ __launch_bounds__(BS) void k_pcg_pq(NodeArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(1, a.G, si, g);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (a.sc->status == ST_RUN) {
    const Z beta = a.sc->beta;
    const int64_t nI = int64_t(a.nx - 1) * (a.ny - 1) * (a.nz - 1);
    int64_t lo, hi;
    chunk_range(nI, a.G, g, lo, hi);
    for (int64_t rr = lo + threadIdx.x; rr < hi; rr += BS) {
      int i, j, k;
      const int64_t n = node_addr(a, rr, i, j, k);
      const Z pn = cadd(a.s[n], cmul(beta, a.p[n]));
      const Z qn = cadd(node_apply(a, a.s, n), cmul(beta, a.q[n]));
      a.p[n] = pn;
      a.q[n] = qn;
      acc[0] = cadd(acc[0], cdotm(pn, qn));
    }
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// x += alpha p ; r -= alpha q ; partials <r, r>
__global__ __launch_bounds__(BS) void k_pcg_xr(NodeArgs a) {
  __shared__ Z lds[2 * BS / 64];
  int si, g;
  map_sys_block(1, a.G, si, g);
  Z acc[2] = {{0.0, 0.0}, {0.0, 0.0}};
  if (a.sc->status == ST_RUN) {
    const Z al = a.sc->alpha;
    const int64_t nI = int64_t(a.nx - 1) * (a.ny - 1) * (a.nz - 1);
    int64_t lo, hi;
    chunk_range(nI, a.G, g, lo, hi);
    for (int64_t rr = lo + threadIdx.x; rr < hi; rr += BS) {
      int i, j, k;
      const int64_t n = node_addr(a, rr, i, j, k);
      a.x[n] = cadd(a.x[n], cmul(al, a.p[n]));
      const Z rv = csub(a.r[n], cmul(al, a.q[n]));
      a.r[n] = rv;
      acc[0].x += rv.x * rv.x + rv.y * rv.y;
    }
  }
  store_partials(a.partials, si, a.G, g, acc, lds);
}

// PCG scalar logic of solver.f90:57-136 (loop while rerr(i) > tol and i < maxit)
template <int KIND>
__global__ __launch_bounds__(BS) void k_pcg_fin(NodeArgs a) {
  __shared__ Z lds[2 * BS / 64];
  Z red[2];
  reduce_sys<2, BS>(a.partials, 0, a.G, red, lds);
  if (threadIdx.x != 0) return;
  SysScal& sc = *a.sc;
  if (KIND == PF_INIT) {
    sc.bnorm = sqrt(red[1].x);
    const double rn = sqrt(red[0].x);
    sc.iter = 1;
    sc.rnorm = rn;
    const double re = rn / sc.bnorm;
    if (a.rerr) a.rerr[0] = re;
    sc.status = (re > a.tol && sc.iter < sc.maxit) ? ST_RUN : ST_CONV;
    return;
  }
  if (sc.status != ST_RUN) return;
  if (KIND == PF_DELTA) {
    sc.rho = red[0];  // delta
    sc.beta = (sc.iter == 1) ? Z{0.0, 0.0} : zdiv(sc.rho, sc.rho1);
  } else if (KIND == PF_ALPHA) {
    sc.alpha = zdiv(sc.rho, red[0]);
  } else {
    sc.rho1 = sc.rho;
    sc.iter += 1;
    const double rn = sqrt(red[0].x);
    sc.rnorm = rn;
    const double re = rn / sc.bnorm;
    if (a.rerr) a.rerr[sc.iter - 1] = re;
    if (!(re > a.tol && sc.iter < sc.maxit)) sc.status = (re > a.tol) ? ST_MAXIT : ST_CONV;
  }
}

// e -= Gai phi on interior edges (Grad, modelOperator3D.f90:797-805)
__global__ __launch_bounds__(BS) void k_grad_sub(NodeArgs a, StencilView<double> A) {
  for (int64_t rr = (int64_t)blockIdx.x * BS + threadIdx.x; rr < A.nInt;
       rr += (int64_t)gridDim.x * BS) {
    const RowInfo ri = decode_row(A, rr);
    // start node of the edge: its own (i,j,k); end node one step along its axis
    int64_t q = ri.e;
    int i, j, k;
    if (ri.c == 0) { i = int(q % a.nx); q /= a.nx; j = int(q % (a.ny + 1)); k = int(q / (a.ny + 1)); }
    else if (ri.c == 1) { q -= a.nXe; i = int(q % (a.nx + 1)); q /= (a.nx + 1); j = int(q % a.ny); k = int(q / a.ny); }
    else { q -= a.nXe + a.nYe; i = int(q % (a.nx + 1)); q /= (a.nx + 1); j = int(q % (a.ny + 1)); k = int(q / (a.ny + 1)); }
    const int64_t sy = a.nx + 1, sz = int64_t(a.nx + 1) * (a.ny + 1);
    const int64_t n0 = i + sy * j + sz * k;
    const int64_t n1 = n0 + (ri.c == 0 ? 1 : (ri.c == 1 ? sy : sz));
    const Z d = csub(a.x[n1], a.x[n0]);
    const double il = a.invL[ri.e];
    a.e[ri.e] = {a.e[ri.e].x - il * d.x, a.e[ri.e].y - il * d.y};
  }
}

}  // namespace mt
