// Host-side operator construction (see mt_host_ops.hpp).
#include "mt_host_ops.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace mt {

namespace {

// (component, i, j, k) of an all-edge index
inline void edge_lattice(const Grid& g, int64_t e, int& c, int& i, int& j, int& k) {
  int64_t q;
  int a, b, d;
  if (e < g.nXe) { c = 0; q = e; }
  else if (e < g.nXe + g.nYe) { c = 1; q = e - g.nXe; }
  else { c = 2; q = e - g.nXe - g.nYe; }
  g.dims(c, a, b, d);
  i = int(q % a);
  q /= a;
  j = int(q % b);
  k = int(q / b);
  (void)d;
}

// slot lookup table: [ca][cb][di+1][dj+1][dk+1]
struct SlotTable {
  signed char t[3][3][3][3][3];
  SlotTable() {
    std::memset(t, -1, sizeof(t));
    for (int ca = 0; ca < 3; ++ca) {
      Nb nb[kSlots];
      stencil_neighbours(ca, nb);
      for (int s = 0; s < kSlots; ++s)
        t[ca][nb[s].comp][nb[s].di + 1][nb[s].dj + 1][nb[s].dk + 1] = (signed char)s;
    }
  }
  int get(int ca, int cb, int di, int dj, int dk) const {
    if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) return -1;
    return t[ca][cb][di + 1][dj + 1][dk + 1];
  }
};
const SlotTable& slots() {
  static SlotTable st;
  return st;
}

}  // namespace

void graddiv_slots(int slot[3][2][6]) {
  const SlotTable& ST = slots();
  // lattice offset of the node's edge q from the node
  const int qc[6] = {0, 0, 1, 1, 2, 2};
  const int qd[6][3] = {{-1, 0, 0}, {0, 0, 0}, {0, -1, 0}, {0, 0, 0}, {0, 0, -1}, {0, 0, 0}};
  for (int c = 0; c < 3; ++c)
    for (int side = 0; side < 2; ++side)
      for (int q = 0; q < 6; ++q) {
        int d[3] = {qd[q][0], qd[q][1], qd[q][2]};
        d[c] += side;   // the second node is one step along the edge
        slot[c][side][q] = ST.get(c, qc[q], d[0], d[1], d[2]);
      }
}

void build_index_maps(HostOperator& op) {
  const Grid& g = op.g;
  op.imap.resize(g.nInt);
  op.bmap.resize(g.nE - g.nInt);
  int64_t ni = 0, nb = 0;
  for (int c = 0; c < 3; ++c) {
    int a, b, d;
    g.dims(c, a, b, d);
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < a; ++i) {
          const int64_t e = g.eidx(c, i, j, k);
          if (g.interior(c, i, j, k)) op.imap[ni++] = int(e);
          else op.bmap[nb++] = int(e);
        }
  }
}

Status assemble(HostOperator& op, int nx, int ny, int nzAir, int nzEarth, const double* dx,
                const double* dy, const double* dz, const double* logcond, double airLog) {
  Status st;
  const int nz = nzAir + nzEarth;
  if (nx < 2 || ny < 2 || nz < 2 || nzEarth < 1 || nzAir < 0) {
    st.code = MT_E_ARG;
    st.msg = "assemble: grid needs nx, ny, nz >= 2 and nzEarth >= 1";
    return st;
  }
  Grid& g = op.g;
  g.set(nx, ny, nz);
  op.dx.assign(dx, dx + nx);
  op.dy.assign(dy, dy + ny);
  op.dz.assign(dz, dz + nz);
  op.nzAir = nzAir;
  op.nzEarth = nzEarth;
  op.logcond.assign(logcond, logcond + int64_t(nx) * ny * nzEarth);
  op.airLog = airLog;
  build_index_maps(op);

  // --- grid metrics (GridDef.f90:301-399, GridCalc.f90:44-418)
  auto half_sum = [](const std::vector<double>& d, int n, std::vector<double>& out) {
    out.assign(n + 1, 0.0);
    out[0] = d[0];
    for (int i = 1; i < n; ++i) out[i] = d[i - 1] + d[i];
    out[n] = d[n - 1];
    for (auto& v : out) v /= 2.0;
  };
  std::vector<double> delX, delY, delZ;
  half_sum(op.dx, nx, delX);
  half_sum(op.dy, ny, delY);
  half_sum(op.dz, nz, delZ);

  const int64_t nE = g.nE;
  std::vector<double> edgeL(nE), dualA(nE);
  for (int c = 0; c < 3; ++c) {
    int a, b, d;
    g.dims(c, a, b, d);
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < a; ++i) {
          const int64_t e = g.eidx(c, i, j, k);
          if (c == 0) { edgeL[e] = op.dx[i]; dualA[e] = delY[j] * delZ[k]; }
          else if (c == 1) { edgeL[e] = op.dy[j]; dualA[e] = delX[i] * delZ[k]; }
          else { edgeL[e] = op.dz[k]; dualA[e] = delX[i] * delY[j]; }
        }
  }
  op.Vedge.resize(nE);
  for (int64_t e = 0; e < nE; ++e) op.Vedge[e] = edgeL[e] * dualA[e];
  const int64_t nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
  auto nidx = [&](int i, int j, int k) { return i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k); };
  std::vector<double> vnode(nN);
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) vnode[nidx(i, j, k)] = delX[i] * delY[j] * delZ[k];

  // --- cell -> edge / node conductivity (ModelMap.inc:88-277): volume-weighted
  // average over the cells adjacent to each edge / node.
  const double airC = std::exp(airLog);
  auto csig = [&](int i, int j, int k) {
    return k < nzAir ? airC : std::exp(logcond[i + int64_t(nx) * (j + int64_t(ny) * (k - nzAir))]);
  };
  std::vector<double> numE(nE, 0.0), denE(nE, 0.0), numN(nN, 0.0), denN(nN, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double vc = op.dx[i] * op.dy[j] * op.dz[k];
        const double vs = vc * csig(i, j, k);
        for (int dj = 0; dj < 2; ++dj)
          for (int dk = 0; dk < 2; ++dk) {
            int64_t e = g.eidx(0, i, j + dj, k + dk);
            numE[e] += vs; denE[e] += vc;
          }
        for (int di = 0; di < 2; ++di)
          for (int dk = 0; dk < 2; ++dk) {
            int64_t e = g.eidx(1, i + di, j, k + dk);
            numE[e] += vs; denE[e] += vc;
          }
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj) {
            int64_t e = g.eidx(2, i + di, j + dj, k);
            numE[e] += vs; denE[e] += vc;
          }
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj)
            for (int dk = 0; dk < 2; ++dk) {
              int64_t n = nidx(i + di, j + dj, k + dk);
              numN[n] += vs; denN[n] += vc;
            }
      }
  std::vector<double> sigE(nE), sigN(nN);
  for (int64_t e = 0; e < nE; ++e) sigE[e] = numE[e] / denE[e];
  for (int64_t n = 0; n < nN; ++n) sigN[n] = numN[n] / denN[n];
  op.Vcell_ghost_edge = denE;
  op.Vcell_ghost_node = denN;
  op.sigN = sigN;
  op.vnode = vnode;
  // UpdateCond zeroes sigma on boundary edges (modelOperator3D.f90:192)
  for (int e : op.bmap) sigE[e] = 0.0;

  // --- air / earth node classes (airNIndex, modelOperator3D.f90:1086-1130)
  const double thr = double(1.1f) * kSigmaAir;
  std::vector<char> eearth(nE, 0);
  for (int64_t e = 0; e < nE; ++e) eearth[e] = (sigE[e] != 0.0) && !(sigE[e] < thr);
  auto node_inner = [&](int i, int j, int k) {
    return i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
  };

  op.vms1.assign(nE, 0.0);
  for (int e : op.imap) op.vms1[e] = kMu0 * sigE[e] * op.Vedge[e];
  op.sigE = sigE;
  op.dualA = dualA;
  op.edgeL = edgeL;

  op.coef.assign(size_t(kSlots) * nE, 0.0);
  const SlotTable& ST = slots();
  std::vector<char> isInt(nE, 0);
  for (int e : op.imap) isInt[e] = 1;
  struct EL { int64_t e; int c, i, j, k; double s; };
  auto add_pair = [&](const EL& ea, const EL& eb, double v) {
    if (!isInt[ea.e] || !isInt[eb.e]) return;
    const int s = ST.get(ea.c, eb.c, eb.i - ea.i, eb.j - ea.j, eb.k - ea.k);
    op.coef[size_t(s) * nE + ea.e] += v;
  };
  // boundary couplings of CC (for the forward RHS) accumulate into a map
  std::vector<std::vector<std::pair<int, double>>> ccib(g.nInt);
  std::vector<int> rowOfEdge(nE, -1);
  for (int64_t r = 0; r < g.nInt; ++r) rowOfEdge[op.imap[r]] = int(r);
  auto add_ccib = [&](const EL& ea, const EL& eb, double v) {
    if (!isInt[ea.e] || isInt[eb.e]) return;
    auto& row = ccib[rowOfEdge[ea.e]];
    for (auto& pr : row)
      if (pr.first == int(eb.e)) { pr.second += v; return; }
    row.push_back({int(eb.e), v});
  };

  // --- curl-curl: each face's 4 edges with the curl signs of setCurlTopology
  // (spOpTopology_SG.f90:24-171); weight DualEdgeL/FaceA
  auto mk = [&](int c, int i, int j, int k, double s) {
    return EL{g.eidx(c, i, j, k), c, i, j, k, s};
  };
  for (int fc = 0; fc < 3; ++fc) {
    int a = (fc == 0) ? nx + 1 : nx, b = (fc == 1) ? ny + 1 : ny, d = (fc == 2) ? nz + 1 : nz;
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < a; ++i) {
          EL ed[4];
          double len, area;
          if (fc == 0) {
            ed[0] = mk(1, i, j, k, 1); ed[1] = mk(1, i, j, k + 1, -1);
            ed[2] = mk(2, i, j, k, -1); ed[3] = mk(2, i, j + 1, k, 1);
            len = (i == 0) ? op.dx[0] : (i == nx ? op.dx[nx - 1] : op.dx[i - 1] + op.dx[i]);
            area = op.dy[j] * op.dz[k];
          } else if (fc == 1) {
            ed[0] = mk(0, i, j, k, -1); ed[1] = mk(0, i, j, k + 1, 1);
            ed[2] = mk(2, i, j, k, 1); ed[3] = mk(2, i + 1, j, k, -1);
            len = (j == 0) ? op.dy[0] : (j == ny ? op.dy[ny - 1] : op.dy[j - 1] + op.dy[j]);
            area = op.dx[i] * op.dz[k];
          } else {
            ed[0] = mk(0, i, j, k, 1); ed[1] = mk(0, i, j + 1, k, -1);
            ed[2] = mk(1, i, j, k, -1); ed[3] = mk(1, i + 1, j, k, 1);
            len = (k == 0) ? op.dz[0] : (k == nz ? op.dz[nz - 1] : op.dz[k - 1] + op.dz[k]);
            area = op.dx[i] * op.dy[j];
          }
          const double w = (len / 2.0) / area;
          for (int p = 0; p < 4; ++p)
            for (int q = 0; q < 4; ++q) {
              const double v = edgeL[ed[p].e] * ed[p].s * w * ed[q].s * edgeL[ed[q].e];
              add_pair(ed[p], ed[q], v);
              add_ccib(ed[p], ed[q], v);
            }
        }
  }

  op.coefCC = op.coef;   // the sigma-free part, for updates of sigma on the device

  // --- grad-div: each interior node's 6 edges (G = -1 at an edge's first
  // node, +1 at its second, setGradTopology spOpTopology_SG.f90:173-266)
  std::vector<std::vector<std::pair<int, double>>> gdr(g.nInt);
  for (int k = 1; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        EL ed[6] = {mk(0, i - 1, j, k, 1), mk(0, i, j, k, -1), mk(1, i, j - 1, k, 1),
                    mk(1, i, j, k, -1),    mk(2, i, j, k - 1, 1), mk(2, i, j, k, -1)};
        bool earth = false;
        for (auto& x : ed) earth = earth || eearth[x.e];
        const int64_t n = nidx(i, j, k);
        if (op.nodeEarth.empty()) op.nodeEarth.assign(nN, 0);
        op.nodeEarth[n] = earth;
        // air part: DualFaceA G diag(1/Vnode) G^T DualFaceA
        // earth part: DualFaceA G diag(1/(sigN Vnode)) G^T diag(sigE DualFaceA)
        const double m2 = earth ? 1.0 / sigN[n] / vnode[n] : 1.0 / vnode[n];
        for (int p = 0; p < 6; ++p)
          for (int q = 0; q < 6; ++q) {
            const double colw = earth ? sigE[ed[q].e] * dualA[ed[q].e] : dualA[ed[q].e];
            add_pair(ed[p], ed[q], dualA[ed[p].e] * ed[p].s * m2 * ed[q].s * colw);
          }
        if (earth) {
          // RHS-form GDii: diag(1/EdgeL) G diag(1/(sigN Vnode)) G^T diag(DualFaceA)
          for (int p = 0; p < 6; ++p) {
            if (!isInt[ed[p].e]) continue;
            auto& row = gdr[rowOfEdge[ed[p].e]];
            for (int q = 0; q < 6; ++q) {
              if (!isInt[ed[q].e]) continue;
              const double v = (1.0 / edgeL[ed[p].e]) * ed[p].s * m2 * ed[q].s * dualA[ed[q].e];
              bool f = false;
              for (auto& pr : row)
                if (pr.first == int(ed[q].e)) { pr.second += v; f = true; break; }
              if (!f) row.push_back({int(ed[q].e), v});
            }
          }
        }
      }
  (void)node_inner;
  auto to_csr = [&](std::vector<std::vector<std::pair<int, double>>>& rows, std::vector<int>& rp,
                    std::vector<int>& cl, std::vector<double>& vl) {
    rp.assign(rows.size() + 1, 0);
    for (size_t r = 0; r < rows.size(); ++r) {
      std::sort(rows[r].begin(), rows[r].end());
      rp[r + 1] = rp[r] + int(rows[r].size());
    }
    cl.resize(rp.back());
    vl.resize(rp.back());
    for (size_t r = 0; r < rows.size(); ++r)
      for (size_t q = 0; q < rows[r].size(); ++q) {
        cl[rp[r] + q] = rows[r][q].first;
        vl[rp[r] + q] = rows[r][q].second;
      }
  };
  to_csr(ccib, op.ccib_row, op.ccib_col, op.ccib_val);
  to_csr(gdr, op.gd_row, op.gd_col, op.gd_val);
  op.have_model = true;
  return st;
}

Status csr_to_stencil(HostOperator& op, int n, int nnz, const int* rowptr, const int* col,
                      int base, const double* val, const double* vms1) {
  Status st;
  const Grid& g = op.g;
  if (g.nx == 0) { st.code = MT_E_STATE; st.msg = "grid not set (mt_set_grid)"; return st; }
  if (int64_t(n) != g.nInt) {
    st.code = MT_E_ARG;
    st.msg = "operator size " + std::to_string(n) + " != interior edges " + std::to_string(g.nInt);
    return st;
  }
  if (rowptr[n] - base != nnz) { st.code = MT_E_ARG; st.msg = "rowptr[n] does not match nnz"; return st; }
  build_index_maps(op);
  op.coef.assign(size_t(kSlots) * g.nE, 0.0);
  op.vms1.assign(g.nE, 0.0);
  const SlotTable& ST = slots();
  for (int r = 0; r < n; ++r) {
    const int64_t e = op.imap[r];
    int ca, ia, ja, ka;
    edge_lattice(g, e, ca, ia, ja, ka);
    op.vms1[e] = vms1[r];
    for (int q = rowptr[r] - base; q < rowptr[r + 1] - base; ++q) {
      const int cidx = col[q] - base;
      if (cidx < 0 || cidx >= n) { st.code = MT_E_ARG; st.msg = "column index out of range"; return st; }
      int cb, ib, jb, kb;
      edge_lattice(g, op.imap[cidx], cb, ib, jb, kb);
      const int s = ST.get(ca, cb, ib - ia, jb - ja, kb - ka);
      if (s < 0) {
        st.code = MT_E_STENCIL;
        st.msg = "entry (" + std::to_string(r) + "," + std::to_string(cidx) +
                 ") is not a staggered-grid stencil neighbour";
        return st;
      }
      op.coef[size_t(s) * g.nE + e] += val[q];
    }
  }
  op.have_model = false;
  return st;
}

void dilu_inv_pivots(const HostOperator& op, double omega, std::vector<zc>& dinv) {
  const Grid& g = op.g;
  const int64_t nE = g.nE;
  dinv.assign(nE, zc(0.0, 0.0));
  const double* cf = op.coef.data();
  // lower slots 1 (a-1), 3 (b-1), 5 (c-1); the coupling back is slot+1 of the
  // neighbour.  Interior rows in EDGEi order visit each block lexicographically.
  for (int c = 0; c < 3; ++c) {
    int a, b, d;
    g.dims(c, a, b, d);
    const int64_t sa = 1, sb = a, sc = int64_t(a) * b;
    for (int k = 0; k < d; ++k)
      for (int j = 0; j < b; ++j)
        for (int i = 0; i < a; ++i) {
          if (!g.interior(c, i, j, k)) continue;
          const int64_t e = g.eidx(c, i, j, k);
          zc p(cf[e], kIsign * omega * op.vms1[e]);
          const int64_t st[3] = {sa, sb, sc};
          for (int q = 0; q < 3; ++q) {
            const double alo = cf[size_t(1 + 2 * q) * nE + e];
            if (alo == 0.0) continue;
            const int64_t l = e - st[q];
            const double aup = cf[size_t(2 + 2 * q) * nE + l];
            p -= alo * aup * dinv[l];
          }
          dinv[e] = 1.0 / p;
        }
  }
}

void stencil_to_csr(const HostOperator& op, std::vector<int>& row, std::vector<int>& col,
                    std::vector<double>& val) {
  const Grid& g = op.g;
  std::vector<int> rowOf(g.nE, -1);
  for (int64_t r = 0; r < g.nInt; ++r) rowOf[op.imap[r]] = int(r);
  row.assign(g.nInt + 1, 0);
  col.clear();
  val.clear();
  for (int64_t r = 0; r < g.nInt; ++r) {
    const int64_t e = op.imap[r];
    int c, i, j, k;
    edge_lattice(g, e, c, i, j, k);
    Nb nb[kSlots];
    stencil_neighbours(c, nb);
    std::vector<std::pair<int, double>> ent;
    for (int s = 0; s < kSlots; ++s) {
      const double v = op.coef[size_t(s) * g.nE + e];
      if (v == 0.0) continue;
      const int64_t f = g.eidx(nb[s].comp, i + nb[s].di, j + nb[s].dj, k + nb[s].dk);
      ent.push_back({rowOf[f], v});
    }
    std::sort(ent.begin(), ent.end());
    for (auto& pr : ent) { col.push_back(pr.first); val.push_back(pr.second); }
    row[r + 1] = int(col.size());
  }
}

void transpose_stencil(const HostOperator& op, std::vector<double>& coefT) {
  const Grid& g = op.g;
  const int64_t nE = g.nE;
  coefT.assign(size_t(kSlots) * nE, 0.0);
  const SlotTable& ST = slots();
  for (int64_t r = 0; r < g.nInt; ++r) {
    const int64_t e = op.imap[r];
    int c, i, j, k;
    edge_lattice(g, e, c, i, j, k);
    Nb nb[kSlots];
    stencil_neighbours(c, nb);
    for (int s = 0; s < kSlots; ++s) {
      const double v = op.coef[size_t(s) * nE + e];
      if (v == 0.0) continue;
      const int ci = i + nb[s].di, cj = j + nb[s].dj, ck = k + nb[s].dk;
      const int64_t f = g.eidx(nb[s].comp, ci, cj, ck);
      // A^T[f][e] = A[e][f]: slot of e as seen from row f
      const int t = ST.get(nb[s].comp, c, i - ci, j - cj, k - ck);
      coefT[size_t(t) * nE + f] += v;
    }
  }
}

void divcor_node_stencil(const HostOperator& op, std::vector<double>& coefN,
                         std::vector<double>& wE) {
  const Grid& g = op.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
  auto nidx = [&](int i, int j, int k) { return i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k); };
  std::vector<char> isInt(g.nE, 0);
  for (int e : op.imap) isInt[e] = 1;
  // weight of each interior edge in DivC: DualFaceA * sigma_E (VDs = VDiv diag(-sigma_E))
  wE.assign(g.nE, 0.0);
  for (int e : op.imap) wE[e] = op.dualA[e] * op.sigE[e];
  coefN.assign(size_t(7) * nN, 0.0);
  // edge e from node a to node b: Gai phi = (phi_b - phi_a)/l, VDs adds
  // -G_en w_e (.) at both ends: row a gets +w/l (phi_b - phi_a), row b -w/l (phi_b - phi_a)
  auto add_edge = [&](int64_t e, int64_t a, int64_t b, int sa, int sb, bool aInt, bool bInt) {
    if (!isInt[e]) return;
    const double c = wE[e] / op.edgeL[e];
    if (aInt) {
      coefN[a] -= c;                       // phi_a on row a
      if (bInt) coefN[size_t(sa) * nN + a] += c;  // phi_b on row a
    }
    if (bInt) {
      coefN[b] -= c;
      if (aInt) coefN[size_t(sb) * nN + b] += c;
    }
  };
  auto inner = [&](int i, int j, int k) {
    return i > 0 && i < nx && j > 0 && j < ny && k > 0 && k < nz;
  };
  for (int k = 0; k <= nz; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        if (i < nx) add_edge(g.eidx(0, i, j, k), nidx(i, j, k), nidx(i + 1, j, k), 2, 1, inner(i, j, k), inner(i + 1, j, k));
        if (j < ny) add_edge(g.eidx(1, i, j, k), nidx(i, j, k), nidx(i, j + 1, k), 4, 3, inner(i, j, k), inner(i, j + 1, k));
        if (k < nz) add_edge(g.eidx(2, i, j, k), nidx(i, j, k), nidx(i, j, k + 1), 6, 5, inner(i, j, k), inner(i, j, k + 1));
      }
}

void divcor_inv_pivots(const HostOperator& op, const std::vector<double>& coefN,
                       std::vector<double>& dinv) {
  const Grid& g = op.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
  dinv.assign(nN, 0.0);
  const int64_t st[3] = {1, nx + 1, int64_t(nx + 1) * (ny + 1)};
  for (int k = 1; k < nz; ++k)
    for (int j = 1; j < ny; ++j)
      for (int i = 1; i < nx; ++i) {
        const int64_t n = i + st[1] * j + st[2] * k;
        double p = coefN[n];
        for (int q = 0; q < 3; ++q) {
          const double alo = coefN[size_t(1 + 2 * q) * nN + n];
          if (alo == 0.0) continue;
          p -= alo * alo * dinv[n - st[q]];
        }
        dinv[n] = 1.0 / p;
      }
}

void grad_assemble(const HostOperator& op, double omega, int npol, const zc* e_all,
                   const zc* lam, int mode, double* dm) {
  const Grid& g = op.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int64_t nE = g.nE, nN = int64_t(nx + 1) * (ny + 1) * (nz + 1);
  auto nidx = [&](int i, int j, int k) { return i + int64_t(nx + 1) * (j + int64_t(ny + 1) * k); };
  std::vector<double> tE(nE, 0.0), tN(nN, 0.0);
  std::vector<zc> L(nE), E(nE), D(nN), S(nN);
  const zc dcoef(0.0, kIsign * omega * kMu0);  // d A_ee / d sigma_e = i ISIGN w mu0 V_e
  for (int pol = 0; pol < npol; ++pol) {
    std::fill(L.begin(), L.end(), zc(0.0));
    std::fill(E.begin(), E.end(), zc(0.0));
    for (int64_t r = 0; r < g.nInt; ++r) {
      const int e = op.imap[r];
      L[e] = lam[size_t(pol) * g.nInt + r];
      E[e] = e_all[size_t(pol) * nE + e];  // interior columns of AAii only
    }
    for (int e : op.imap) tE[e] += (dcoef * op.Vedge[e] * L[e] * E[e]).real();
    if (mode == 0) continue;
    // D_n = sum_a G_an A_a lam_a, S_n = sum_b G_bn sigma_b A_b e_b (interior edges)
    std::fill(D.begin(), D.end(), zc(0.0));
    std::fill(S.begin(), S.end(), zc(0.0));
    for (int c = 0; c < 3; ++c) {
      int a, b, d;
      g.dims(c, a, b, d);
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < b; ++j)
          for (int i = 0; i < a; ++i) {
            if (!g.interior(c, i, j, k)) continue;
            const int64_t e = g.eidx(c, i, j, k);
            const int64_t n0 = nidx(i, j, k);
            const int64_t n1 = nidx(i + (c == 0), j + (c == 1), k + (c == 2));
            const zc dl = op.dualA[e] * L[e], ds = op.sigE[e] * op.dualA[e] * E[e];
            D[n0] -= dl; D[n1] += dl;
            S[n0] -= ds; S[n1] += ds;
          }
    }
    // node weights M2e = Nearth / (sigma_N Vnode) on interior earth nodes
    for (int c = 0; c < 3; ++c) {
      int a, b, d;
      g.dims(c, a, b, d);
      for (int k = 0; k < d; ++k)
        for (int j = 0; j < b; ++j)
          for (int i = 0; i < a; ++i) {
            if (!g.interior(c, i, j, k)) continue;
            const int64_t e = g.eidx(c, i, j, k);
            const int64_t n0 = nidx(i, j, k);
            const int64_t n1 = nidx(i + (c == 0), j + (c == 1), k + (c == 2));
            zc acc(0.0);
            if (op.nodeEarth[n0]) acc -= D[n0] / (op.sigN[n0] * op.vnode[n0]);
            if (op.nodeEarth[n1]) acc += D[n1] / (op.sigN[n1] * op.vnode[n1]);
            tE[e] += (op.dualA[e] * E[e] * acc).real();
          }
    }
    for (int64_t n = 0; n < nN; ++n)
      if (op.nodeEarth[n])
        tN[n] += (-D[n] * S[n] / (op.sigN[n] * op.sigN[n] * op.vnode[n])).real();
  }
  // chain rule to the cells: sigma_E and sigma_N are volume-weighted averages
  // of exp(m) over adjacent cells (ModelParamToEdge/Node, ModelMap.inc:88-277)
  for (int k = op.nzAir; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int64_t q = i + int64_t(nx) * (j + int64_t(ny) * (k - op.nzAir));
        const double w = op.dx[i] * op.dy[j] * op.dz[k] * std::exp(op.logcond[q]);
        double acc = 0.0;
        for (int dj = 0; dj < 2; ++dj)
          for (int dk = 0; dk < 2; ++dk) {
            const int64_t e = g.eidx(0, i, j + dj, k + dk);
            acc += tE[e] / op.Vcell_ghost_edge[e];
          }
        for (int di = 0; di < 2; ++di)
          for (int dk = 0; dk < 2; ++dk) {
            const int64_t e = g.eidx(1, i + di, j, k + dk);
            acc += tE[e] / op.Vcell_ghost_edge[e];
          }
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj) {
            const int64_t e = g.eidx(2, i + di, j + dj, k);
            acc += tE[e] / op.Vcell_ghost_edge[e];
          }
        if (mode != 0)
          for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
              for (int dk = 0; dk < 2; ++dk) {
                const int64_t n = nidx(i + di, j + dj, k + dk);
                acc += tN[n] / op.Vcell_ghost_node[n];
              }
        dm[q] += w * acc;
      }
}

void edge_to_model(const HostOperator& op, const double* e, double* dm) {
  // EdgeToModelParam: m = V_C * Edge2Cell(e / Cell2Edge(V_C)) * exp(m0) on
  // earth cells; Edge2Cell is the transpose of Cell2Edge's averaging, i.e.
  // each edge spreads e/ghost over its adjacent cells with weight 1/ncells,
  // which here equals V_C-weighting through ghost = sum of V_C.
  const Grid& g = op.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  std::vector<double> cnt(g.nE, 0.0);
  // number of cells adjacent to each edge (4 inside, 2 on faces, 1 on lines)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        for (int dj = 0; dj < 2; ++dj)
          for (int dk = 0; dk < 2; ++dk) cnt[g.eidx(0, i, j + dj, k + dk)] += 1;
        for (int di = 0; di < 2; ++di)
          for (int dk = 0; dk < 2; ++dk) cnt[g.eidx(1, i + di, j, k + dk)] += 1;
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj) cnt[g.eidx(2, i + di, j + dj, k)] += 1;
      }
  // ghost(e) in the reference = (sum of adjacent V_C) / cnt(e)
  for (int k = op.nzAir; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double vc = op.dx[i] * op.dy[j] * op.dz[k];
        double acc = 0.0;
        auto addE = [&](int64_t ed) {
          const double ghost = op.Vcell_ghost_edge[ed] / cnt[ed];
          acc += (e[ed] / ghost) / cnt[ed];
        };
        for (int dj = 0; dj < 2; ++dj)
          for (int dk = 0; dk < 2; ++dk) addE(g.eidx(0, i, j + dj, k + dk));
        for (int di = 0; di < 2; ++di)
          for (int dk = 0; dk < 2; ++dk) addE(g.eidx(1, i + di, j, k + dk));
        for (int di = 0; di < 2; ++di)
          for (int dj = 0; dj < 2; ++dj) addE(g.eidx(2, i + di, j + dj, k));
        const int64_t q = i + int64_t(nx) * (j + int64_t(ny) * (k - op.nzAir));
        dm[q] = vc * acc * std::exp(op.logcond[q]);
      }
}

}  // namespace mt
