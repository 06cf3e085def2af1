// Sensitivity kernels: the transposed-solve right-hand side and
// post-processing of FWDsolve3D, the forward post-processing, and the
// gradient assembly (PmultT + dEdgeToModelParam, and the exact discrete
// adjoint including the grad-div part), all on device-resident fields.
//
// Geometry and the model's sigma_E, sigma_N and grad-div node weights come
// from the UpdateCond kernels (mt_cond.hpp, CondArgs): m2[n] > 0 marks an
// interior earth node with weight 1/(sigma_N V_node), m2[n] < 0 an interior
// air node, 0 a boundary node.  Every reduction is formed in the order of the
// host restatement (mt_host_ops.cpp) with contraction off: the cell gradients
// equal the host's bit for bit; the TRN right-hand side sums its grad-div
// term node by node instead of over a merged matrix row (equal to rounding).
#pragma once

#include "mt_cond.hpp"
#include "mt_device.hpp"
#include "mt_internal.h"

namespace mt {

__device__ __forceinline__ double sens_len(const CondArgs& a, int c, int i, int j, int k) {
  return c == 0 ? a.dx[i] : (c == 1 ? a.dy[j] : a.dz[k]);
}
__device__ __forceinline__ double sens_dual(const CondArgs& a, int c, int i, int j, int k) {
  return c == 0 ? a.delY[j] * a.delZ[k] : (c == 1 ? a.delX[i] * a.delZ[k] : a.delX[i] * a.delY[j]);
}
__device__ __forceinline__ int64_t sens_nidx(const CondArgs& a, int i, int j, int k) {
  return i + int64_t(a.nx + 1) * (j + int64_t(a.ny + 1) * k);
}
// complex product without contraction (the host restatement's rounding)
__device__ __forceinline__ Z zmul_nc(Z p, Z q) {
#pragma clang fp contract(off)
  return {p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x};
}

// The six edges around node (i, j, k) in the grad-div order of the assembly:
// x(i-1), x(i), y(j-1), y(j), z(k-1), z(k); G = +1 where the node is the
// edge's second end (the first of each pair), -1 where it is its first end.
struct NodeEdges {
  int c[6], i[6], j[6], k[6];
  __device__ NodeEdges(int ni, int nj, int nk) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      c[q] = q / 2;
      i[q] = ni - (q == 0);
      j[q] = nj - (q == 2);
      k[q] = nk - (q == 4);
    }
  }
};

// Ghost volumes: per edge the summed volume of its adjacent cells (air cells
// included) and their count, per node the same over its cells, summed in
// (k, j, i) cell order (Cell2Edge / Cell2Node of V_C, ModelMap.inc:281-372).
__global__ void k_sens_ghost(CondArgs a, double* __restrict__ gE, unsigned char* __restrict__ cntE,
                             double* __restrict__ gN) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.nE + nN;
       t += (int64_t)gridDim.x * blockDim.x) {
    int c = 3, i, j, k;
    if (t < a.nE) {
      cond_edge_pos(a, t, c, i, j, k);
    } else {
      int64_t q = t - a.nE;
      i = int(q % (a.nx + 1));
      q /= (a.nx + 1);
      j = int(q % (a.ny + 1));
      k = int(q / (a.ny + 1));
    }
    double den = 0.0;
    int cnt = 0;
    for (int k2 = (c == 2 ? k : k - 1); k2 <= k; ++k2)
      for (int j2 = (c == 1 ? j : j - 1); j2 <= j; ++j2)
        for (int i2 = (c == 0 ? i : i - 1); i2 <= i; ++i2) {
          if (i2 < 0 || j2 < 0 || k2 < 0 || i2 >= a.nx || j2 >= a.ny || k2 >= a.nz) continue;
          den += a.dx[i2] * a.dy[j2] * a.dz[k2];
          ++cnt;
        }
    if (t < a.nE) {
      gE[t] = den;
      cntE[t] = (unsigned char)cnt;
    } else {
      gN[t - a.nE] = den;
    }
  }
}

// ------------------------------------------------ FWDsolve3D post-processing
// Forward case (EMsolve3D.f90:463, 481-485): e(EDGEi) = x on top of the
// boundary data already in e.
__global__ void k_post_fwd(int nsys, int64_t nInt, int64_t nE, const int* __restrict__ imap,
                           const Z* __restrict__ x, Z* __restrict__ e) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nInt * nsys;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / nInt, r = t - s * nInt;
    e[s * nE + imap[r]] = x[t];
  }
}

// Transposed case without boundary data (EMsolve3D.f90:463-478):
// e = V_E (ei on the interior edges), then on the boundary edges
// e(EDGEb) = -CCib^T ei, gathered per boundary edge over its interior rows in
// increasing row order (bt_row/bt_col/bt_val: CCib transposed, rows in EDGEb
// order).
__global__ void k_post_trn(CondArgs a, int nsys, const int* __restrict__ bmap, int64_t nBdry,
                           const int* __restrict__ bt_row, const int* __restrict__ bt_col,
                           const double* __restrict__ bt_val, const Z* __restrict__ ei,
                           Z* __restrict__ e) {
#pragma clang fp contract(off)
  const int64_t per = a.nInt + nBdry;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < per * nsys;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = t / per, r = t - s * per;
    const Z* __restrict__ es = ei + s * a.nInt;
    if (r < a.nInt) {
      const int64_t ed = a.imap[r];
      int c, i, j, k;
      cond_edge_pos(a, ed, c, i, j, k);
      const double v = sens_len(a, c, i, j, k) * sens_dual(a, c, i, j, k);
      const Z x = es[r];
      e[s * a.nE + ed] = {v * x.x, v * x.y};
    } else {
      const int64_t b = r - a.nInt;
      Z acc = {0.0, 0.0};
      for (int q = bt_row[b]; q < bt_row[b + 1]; ++q) {
        const Z x = es[bt_col[q]];
        acc.x = acc.x + bt_val[q] * x.x;
        acc.y = acc.y + bt_val[q] * x.y;
      }
      e[s * a.nE + bmap[b]] = {-acc.x, -acc.y};
    }
  }
}

// ------------------------------------------------ FWDsolve3D TRN right-hand side
// Node divergence of the source scaled as the reference's TRN branch does
// (EMsolve3D.f90:245-252): si = s(EDGEi) * iOmegaMuInv / V_E (iOmegaMuInv
// per system in inv), then
// div_n = sum over the node's interior edges of G * DualFaceA * si, on the
// interior earth nodes (the nodes of the RHS-form GDii); zero elsewhere.
__global__ void k_trn_div(CondArgs a, int nsys, const Z* __restrict__ inv,
                          const Z* __restrict__ s, Z* __restrict__ div) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nN * nsys;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sy = t / nN, n = t - sy * nN;
    Z acc = {0.0, 0.0};
    if (a.m2[n] > 0.0) {
      int64_t q = n;
      const int ni = int(q % (a.nx + 1));
      q /= (a.nx + 1);
      const int nj = int(q % (a.ny + 1));
      const int nk = int(q / (a.ny + 1));
      // inv[sy] = iOmegaMuInv = 1 / (i ISIGN omega mu0) of the system
      const Z iw = inv[sy];
      const NodeEdges E(ni, nj, nk);
      const Z* __restrict__ ss = s + sy * a.nE;
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        if (!cond_interior(a, E.c[p], E.i[p], E.j[p], E.k[p])) continue;
        const int64_t ed = cond_eidx(a, E.c[p], E.i[p], E.j[p], E.k[p]);
        const double ve = sens_len(a, E.c[p], E.i[p], E.j[p], E.k[p]) *
                          sens_dual(a, E.c[p], E.i[p], E.j[p], E.k[p]);
        const Z xi = zmul_nc(ss[ed], iw);
        const Z si = {xi.x / ve, xi.y / ve};
        const double g = (p & 1) ? -1.0 : 1.0;
        const double da = sens_dual(a, E.c[p], E.i[p], E.j[p], E.k[p]);
        acc.x = acc.x + g * da * si.x;
        acc.y = acc.y + g * da * si.y;
      }
    }
    div[t] = acc;
  }
}

// b = s(EDGEi) + V_E * GDii(si) with GDii(si)_e = (1/EdgeL_e) * sum over the
// edge's earth end nodes of G_en * m2_n * div_n  (gd != 0), or b = s(EDGEi)
// (gd == 0: the right-hand side of the exact transposed system A^T lam = s).
__global__ void k_trn_rhs(CondArgs a, int nsys, int gd, const Z* __restrict__ s,
                          const Z* __restrict__ div, Z* __restrict__ b) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.nInt * nsys;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t sy = t / a.nInt, r = t - sy * a.nInt;
    const int64_t ed = a.imap[r];
    const Z x = s[sy * a.nE + ed];
    if (!gd) { b[t] = x; continue; }
    int c, i, j, k;
    cond_edge_pos(a, ed, c, i, j, k);
    Z st = {0.0, 0.0};
    for (int side = 0; side < 2; ++side) {
      const int ni = i + (c == 0 ? side : 0), nj = j + (c == 1 ? side : 0), nk = k + (c == 2 ? side : 0);
      const int64_t n = sens_nidx(a, ni, nj, nk);
      const double m2 = a.m2[n];
      if (!(m2 > 0.0)) continue;
      const double g = side == 0 ? -1.0 : 1.0;
      const Z d = div[sy * nN + n];
      st.x = st.x + g * m2 * d.x;
      st.y = st.y + g * m2 * d.y;
    }
    const double len = sens_len(a, c, i, j, k), dual = sens_dual(a, c, i, j, k);
    const double v = len * dual, il = 1.0 / len;
    b[t] = {x.x + (il * st.x) * v, x.y + (il * st.y) * v};
  }
}

// ------------------------------------------------ gradient assembly
// PmultT (SolverSens.f90:63-116) for one frequency, mapped to the cells by
// dEdgeToModelParam (ModelMap.inc:281-372): the edge value is
// Re(-i ISIGN mu0 omega * sum_pol e0_pol lam_pol) on every edge, and cell q
// receives V_C exp(m_q) * sum over its 12 edges of (value / ghost) / count
// with ghost = (summed adjacent cell volume) / count.
__device__ __forceinline__ double pmult_edge(const Z* __restrict__ e0, const Z* __restrict__ lam,
                                             int npol, int64_t nE, int64_t ed, double ci) {
#pragma clang fp contract(off)
  Z acc = zmul_nc(e0[ed], lam[ed]);
  for (int p = 1; p < npol; ++p) {
    const Z t = zmul_nc(e0[p * nE + ed], lam[p * nE + ed]);
    acc.x = acc.x + t.x;
    acc.y = acc.y + t.y;
  }
  return 0.0 * acc.x - ci * acc.y;   // Re((0, ci) * acc)
}

__global__ void k_pmult_cells(CondArgs a, int npol, double omega, const Z* __restrict__ e0,
                              const Z* __restrict__ lam, const double* __restrict__ gE,
                              const unsigned char* __restrict__ cntE, double* __restrict__ dm,
                              int accumulate) {
#pragma clang fp contract(off)
  const int nze = a.nz - a.nzAir;
  const int64_t ncell = int64_t(a.nx) * a.ny * nze;
  const double ci = -double(kIsign) * kMu0 * omega;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ncell;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    const int i = int(r % a.nx);
    r /= a.nx;
    const int j = int(r % a.ny);
    const int k = int(r / a.ny) + a.nzAir;
    double acc = 0.0;
    auto add = [&](int64_t ed) {
      const double cnt = double(cntE[ed]);
      const double ghost = gE[ed] / cnt;
      acc += (pmult_edge(e0, lam, npol, a.nE, ed, ci) / ghost) / cnt;
    };
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk) add(cond_eidx(a, 0, i, j + dj, k + dk));
    for (int di = 0; di < 2; ++di)
      for (int dk = 0; dk < 2; ++dk) add(cond_eidx(a, 1, i + di, j, k + dk));
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj) add(cond_eidx(a, 2, i + di, j + dj, k));
    const double vc = a.dx[i] * a.dy[j] * a.dz[k];
    const double v = vc * acc * a.csig[i + int64_t(a.nx) * (j + int64_t(a.ny) * k)];
    dm[q] = accumulate ? dm[q] + v : v;
  }
}

// Exact gradient, node stage for one (frequency, polarisation): on every
// interior earth node D_n = sum_a G_an DualFaceA_a lam_a and
// S_n = sum_b G_bn sigma_b DualFaceA_b e_b over the node's interior edges
// (lam on interior edges in EDGEi order, reached through the row index of
// each edge; e on all edges).
__global__ void k_grad_nodes(CondArgs a, const int* __restrict__ rowOf, const Z* __restrict__ lam,
                             const Z* __restrict__ e, Z* __restrict__ D, Z* __restrict__ S) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < nN;
       n += (int64_t)gridDim.x * blockDim.x) {
    Z dd = {0.0, 0.0}, sd = {0.0, 0.0};
    if (a.m2[n] > 0.0) {
      int64_t q = n;
      const int ni = int(q % (a.nx + 1));
      q /= (a.nx + 1);
      const int nj = int(q % (a.ny + 1));
      const int nk = int(q / (a.ny + 1));
      const NodeEdges E(ni, nj, nk);
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        if (!cond_interior(a, E.c[p], E.i[p], E.j[p], E.k[p])) continue;
        const int64_t ed = cond_eidx(a, E.c[p], E.i[p], E.j[p], E.k[p]);
        const double da = sens_dual(a, E.c[p], E.i[p], E.j[p], E.k[p]);
        const Z l = lam[rowOf[ed]], x = e[ed];
        const Z dl = {da * l.x, da * l.y};
        const double sd_w = a.sigE[ed] * da;
        const Z ds = {sd_w * x.x, sd_w * x.y};
        if (p & 1) {   // the node is the edge's first end
          dd.x = dd.x - dl.x; dd.y = dd.y - dl.y;
          sd.x = sd.x - ds.x; sd.y = sd.y - ds.y;
        } else {
          dd.x = dd.x + dl.x; dd.y = dd.y + dl.y;
          sd.x = sd.x + ds.x; sd.y = sd.y + ds.y;
        }
      }
    }
    D[n] = dd;
    S[n] = sd;
  }
}

// Exact gradient, edge and node accumulation for one (frequency,
// polarisation): tE[e] += Re(i ISIGN omega mu0 V_e lam_e e_e) (the
// derivative of the diagonal), and with the grad-div part (gd != 0)
// tE[e] += Re(DualFaceA_e e_e (-D_n0/(sigma_N V_N)|earth + D_n1/(sigma_N V_N)|earth)),
// tN[n] += Re(-D_n S_n / (sigma_N^2 V_N)) on earth nodes.
__global__ void k_grad_accum(CondArgs a, double omega, int gd, const Z* __restrict__ lam,
                             const Z* __restrict__ e, const Z* __restrict__ D,
                             const Z* __restrict__ S, double* __restrict__ tE,
                             double* __restrict__ tN) {
#pragma clang fp contract(off)
  const int64_t nN = int64_t(a.nx + 1) * (a.ny + 1) * (a.nz + 1);
  const double ci = double(kIsign) * omega * kMu0;   // dcoef = (0, ci)
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < a.nInt + nN;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (t < a.nInt) {
      const int64_t ed = a.imap[t];
      int c, i, j, k;
      cond_edge_pos(a, ed, c, i, j, k);
      const double len = sens_len(a, c, i, j, k), dual = sens_dual(a, c, i, j, k);
      const double v = len * dual;
      const Z l = lam[t], x = e[ed];
      // ((dcoef * V) * L) * E
      const Z dv = {0.0 * v, ci * v};
      const Z t2 = zmul_nc(dv, l);
      const Z t3 = zmul_nc(t2, x);
      double te = tE[ed] + t3.x;
      if (gd) {
        Z acc = {0.0, 0.0};
        for (int side = 0; side < 2; ++side) {
          const int ni = i + (c == 0 ? side : 0), nj = j + (c == 1 ? side : 0), nk = k + (c == 2 ? side : 0);
          const int64_t n = sens_nidx(a, ni, nj, nk);
          if (!(a.m2[n] > 0.0)) continue;
          const double den = a.sigN[n] * (a.delX[ni] * a.delY[nj] * a.delZ[nk]);
          const Z q = {D[n].x / den, D[n].y / den};
          if (side == 0) { acc.x = acc.x - q.x; acc.y = acc.y - q.y; }
          else { acc.x = acc.x + q.x; acc.y = acc.y + q.y; }
        }
        const Z w = {dual * x.x, dual * x.y};
        te = te + zmul_nc(w, acc).x;
      }
      tE[ed] = te;
    } else if (gd) {
      const int64_t n = t - a.nInt;
      if (!(a.m2[n] > 0.0)) continue;
      int64_t q = n;
      const int ni = int(q % (a.nx + 1));
      q /= (a.nx + 1);
      const int nj = int(q % (a.ny + 1));
      const int nk = int(q / (a.ny + 1));
      const double sn = a.sigN[n];
      const double den = sn * sn * (a.delX[ni] * a.delY[nj] * a.delZ[nk]);
      const Z md = {-D[n].x, -D[n].y};
      const Z p = zmul_nc(md, S[n]);
      tN[n] = tN[n] + p.x / den;
    }
  }
}

// Exact gradient, chain rule to the cells: sigma_E and sigma_N are
// volume-weighted averages of exp(m) over the adjacent cells
// (ModelParamToEdge / ModelParamToNode, ModelMap.inc:88-277), so
// dm_q += V_C exp(m_q) * (sum over the 12 edges of tE / ghost_E
//                        + sum over the 8 nodes of tN / ghost_N).
__global__ void k_grad_cells(CondArgs a, int gd, const double* __restrict__ tE,
                             const double* __restrict__ tN, const double* __restrict__ gE,
                             const double* __restrict__ gN, double* __restrict__ dm) {
#pragma clang fp contract(off)
  const int nze = a.nz - a.nzAir;
  const int64_t ncell = int64_t(a.nx) * a.ny * nze;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ncell;
       q += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = q;
    const int i = int(r % a.nx);
    r /= a.nx;
    const int j = int(r % a.ny);
    const int k = int(r / a.ny) + a.nzAir;
    const double w = a.dx[i] * a.dy[j] * a.dz[k] * a.csig[i + int64_t(a.nx) * (j + int64_t(a.ny) * k)];
    double acc = 0.0;
    for (int dj = 0; dj < 2; ++dj)
      for (int dk = 0; dk < 2; ++dk) {
        const int64_t ed = cond_eidx(a, 0, i, j + dj, k + dk);
        acc += tE[ed] / gE[ed];
      }
    for (int di = 0; di < 2; ++di)
      for (int dk = 0; dk < 2; ++dk) {
        const int64_t ed = cond_eidx(a, 1, i + di, j, k + dk);
        acc += tE[ed] / gE[ed];
      }
    for (int di = 0; di < 2; ++di)
      for (int dj = 0; dj < 2; ++dj) {
        const int64_t ed = cond_eidx(a, 2, i + di, j + dj, k);
        acc += tE[ed] / gE[ed];
      }
    if (gd)
      for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj)
          for (int dk = 0; dk < 2; ++dk) {
            const int64_t n = sens_nidx(a, i + di, j + dj, k + dk);
            acc += tN[n] / gN[n];
          }
    dm[q] = dm[q] + w * acc;
  }
}

}  // namespace mt
