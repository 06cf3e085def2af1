// Receiver set-up for the data functionals (host side of the LmultT path).
#include "mt_sites.hpp"

#include <algorithm>
#include <cmath>
#include <map>

namespace mt {

namespace {

// Node coordinates as the reference's setup_grid forms them
// (GridDef.f90:301-399): cumulative sums, centres from the half-sums delX,
// z measured down from the Earth's surface (the top of the air layers is at
// -airThickness), all shifted by the origin.
struct Coords {
  std::vector<double> xE, yE, zE, xC, yC, zC, delX, delY, delZ;
};

std::vector<double> half_sums(const double* d, int n) {
  std::vector<double> h(n + 1);
  h[0] = d[0];
  for (int i = 1; i < n; ++i) h[i] = d[i - 1] + d[i];
  h[n] = d[n - 1];
  for (double& v : h) v = v / 2.0;
  return h;
}

Coords make_coords(const Grid& g, const double* dx, const double* dy, const double* dz, int nzAir,
                   const double org[3]) {
  Coords c;
  c.delX = half_sums(dx, g.nx);
  c.delY = half_sums(dy, g.ny);
  c.delZ = half_sums(dz, g.nz);
  auto cum = [](const double* d, int n, double o, std::vector<double>& out) {
    out.assign(n + 1, o);
    double s = 0.0;
    for (int i = 0; i < n; ++i) { s = s + d[i]; out[i + 1] = s + o; }
  };
  cum(dx, g.nx, org[0], c.xE);
  cum(dy, g.ny, org[1], c.yE);
  cum(dz, g.nz, 0.0, c.zE);
  c.zE[0] = org[2];
  const double air = c.zE[nzAir];
  auto cen = [](const std::vector<double>& h, int n, double o, std::vector<double>& out) {
    out.assign(n, 0.0);
    double s = 0.0;
    for (int i = 0; i < n; ++i) { s = s + h[i]; out[i] = s + o; }
  };
  cen(c.delX, g.nx, org[0], c.xC);
  cen(c.delY, g.ny, org[1], c.yC);
  cen(c.delZ, g.nz, 0.0, c.zC);
  for (int i = 0; i < g.nz; ++i) {
    c.zC[i] = c.zC[i] - air + org[2];
    c.zE[i] = c.zE[i] - air + org[2];
  }
  c.zE[g.nz] = c.zE[g.nz] - air + org[2];
  return c;
}

// The reference compares coordinates after rounding them to 1e-13 relative to
// LARGE_REAL (utilities.f90:100-114 clean; LARGE_REAL is the single-precision
// literal 1.0e13, math_constants.f90:43), so points within that of a node
// count as on it.
double clean(double v) {
  const double big = double(1.0e13f);
  return std::round(v * big) / big;
}
// 1-based ix with node[ix-1] <= v < node[ix] (utilities.f90:196-225 minNode)
int min_node(double v, const std::vector<double>& node) {
  const double cv = clean(v);
  for (size_t i = 0; i < node.size(); ++i)
    if (clean(node[i]) > cv) return int(i);
  return int(node.size());
}

struct Entry { int comp, i, j, k; double w; };   // 1-based lattice indices

// Tri-linear interpolation of an edge field (centre axis along `comp`) or a
// face field (edge axis along `comp`) at v: 8 weighted lattice points with
// weights > 0.  Edge fields: EinterpSetUp; face fields: BinterpSetUp.
std::vector<Entry> trilinear(const Grid& g, const Coords& c, const double* v, int comp, bool face,
                             const double* dx, const double* dy, const double* dz) {
  const std::vector<double>* nodes[3];
  const double* len[3] = {dx, dy, dz};
  const std::vector<double>* half[3] = {&c.delX, &c.delY, &c.delZ};
  const std::vector<double>* E[3] = {&c.xE, &c.yE, &c.zE};
  const std::vector<double>* C[3] = {&c.xC, &c.yC, &c.zC};
  const int ncell[3] = {g.nx, g.ny, g.nz};
  int i0[3], nmax[3];
  bool centre[3];
  for (int d = 0; d < 3; ++d) {
    // edge field: centre nodes along its own axis; face field: along the others
    centre[d] = face ? (d != comp) : (d == comp);
    nodes[d] = centre[d] ? C[d] : E[d];
    i0[d] = min_node(v[d], *nodes[d]);
    nmax[d] = centre[d] ? ncell[d] : ncell[d] + 1;
  }
  double w[3][2];
  for (int d = 0; d < 3; ++d) {
    if (i0[d] > 0 && i0[d] < nmax[d]) {
      const int q = i0[d];
      w[d][1] = centre[d] ? (v[d] - (*C[d])[q - 1]) / (*half[d])[q]
                          : (v[d] - (*E[d])[q - 1]) / len[d][q - 1];
    } else {
      w[d][1] = i0[d] <= 0 ? 1.0 : 0.0;
    }
    w[d][0] = 1.0 - w[d][1];
  }
  std::vector<Entry> out;
  for (int n = 0; n < 2; ++n)
    for (int m = 0; m < 2; ++m)
      for (int p = 0; p < 2; ++p) {
        const double wa = w[0][n] * w[1][m] * w[2][p];
        if (wa > 0) out.push_back({comp, i0[0] + n, i0[1] + m, i0[2] + p, wa});
      }
  return out;
}

bool edge_ok(const Grid& g, int comp, int i, int j, int k) {
  int a, b, d;
  g.dims(comp, a, b, d);
  return i >= 1 && i <= a && j >= 1 && j <= b && k >= 1 && k <= d;
}

}  // namespace

int build_site_rows(const Grid& g, const double* dx, const double* dy, const double* dz, int nzAir,
                    const double origin[3], int nsite, const double* x, const double* r,
                    const double* az, SiteRows& out, std::string& err) {
  static const double az0[7] = {0.0, 90.0, 0.0, 0.0, 90.0, 0.0, 90.0};
  const Coords c = make_coords(g, dx, dy, dz, nzAir, origin);
  out = SiteRows();
  out.nsite = nsite;
  out.ptr.push_back(0);
  out.rot.assign(size_t(nsite) * 24, 0.0);
  const double d2r = kPi / 180.0;
  for (int s = 0; s < nsite; ++s) {
    const double* xs = x + 3 * s;
    const double rz[3] = {0.0, 0.0, 0.0};
    const double* rs = r ? r + 3 * s : rz;
    for (int b = 0; b < kSiteRows; ++b) {
      std::map<int64_t, double> row;   // merged by edge, ascending index
      const double* at = (b >= kSiteHrx) ? rs : xs;
      if (b <= kSiteEy) {
        for (const Entry& e : trilinear(g, c, at, b, false, dx, dy, dz)) {
          if (!edge_ok(g, e.comp, e.i, e.j, e.k)) {
            err = "receiver " + std::to_string(s) + ": E interpolation leaves the grid";
            return MT_E_ARG;
          }
          row[g.eidx(e.comp, e.i - 1, e.j - 1, e.k - 1)] += e.w;
        }
      } else {
        const int comp = (b == kSiteHrx) ? 0 : (b == kSiteHry) ? 1 : b - kSiteHx;
        // face (comp; i, j, k) -> the four edges around it, circulation / area
        // (BfromESetUp: h = e*sum(l)/S with l = dx|dy|dz, S the face area)
        for (const Entry& f : trilinear(g, c, at, comp, true, dx, dy, dz)) {
          const int i = f.i, j = f.j, k = f.k;
          struct Q { int comp, i, j, k; double len, sgn; } q[4];
          double S;
          if (comp == 0) {
            S = dy[j - 1] * dz[k - 1];
            q[0] = {1, i, j, k, dy[j - 1], 1.0};   q[1] = {1, i, j, k + 1, dy[j - 1], -1.0};
            q[2] = {2, i, j, k, dz[k - 1], -1.0};  q[3] = {2, i, j + 1, k, dz[k - 1], 1.0};
          } else if (comp == 1) {
            S = dx[i - 1] * dz[k - 1];
            q[0] = {2, i, j, k, dz[k - 1], 1.0};   q[1] = {2, i + 1, j, k, dz[k - 1], -1.0};
            q[2] = {0, i, j, k, dx[i - 1], -1.0};  q[3] = {0, i, j, k + 1, dx[i - 1], 1.0};
          } else {
            S = dx[i - 1] * dy[j - 1];
            q[0] = {0, i, j, k, dx[i - 1], 1.0};   q[1] = {0, i, j + 1, k, dx[i - 1], -1.0};
            q[2] = {1, i, j, k, dy[j - 1], -1.0};  q[3] = {1, i + 1, j, k, dy[j - 1], 1.0};
          }
          for (const Q& e : q) {
            if (!edge_ok(g, e.comp, e.i, e.j, e.k)) {
              err = "receiver " + std::to_string(s) + ": H interpolation leaves the grid";
              return MT_E_ARG;
            }
            row[g.eidx(e.comp, e.i - 1, e.j - 1, e.k - 1)] += e.sgn * e.len / S * f.w;
          }
        }
      }
      for (const auto& kv : row) {
        out.idx.push_back(int(kv.first));
        out.val.push_back(kv.second);
      }
      out.ptr.push_back(int(out.idx.size()));
    }
    const double* a = az ? az + 7 * s : az0;
    // frames: (first, second) channel azimuths
    const double fr[3][2] = {{a[3], a[4]}, {a[0], a[1]}, {a[5], a[6]}};
    for (int f = 0; f < 3; ++f) {
      const double a1 = fr[f][0] * d2r, a2 = fr[f][1] * d2r;
      const double s12 = std::sin(a2 - a1);
      double* R = &out.rot[size_t(s) * 24 + f * 8];
      // responses: s12 * [[sin a2, -cos a2], [-sin a1, cos a1]]
      R[0] = s12 * std::sin(a2); R[1] = -s12 * std::cos(a2);
      R[2] = -s12 * std::sin(a1); R[3] = s12 * std::cos(a1);
      // linearised functionals: the same matrix divided by s12
      R[4] = std::sin(a2) / s12; R[5] = -std::cos(a2) / s12;
      R[6] = -std::sin(a1) / s12; R[7] = std::cos(a1) / s12;
    }
  }
  // transpose for the adjoint gather
  std::map<int, std::vector<std::pair<int, double>>> by_edge;
  const int nrow = nsite * kSiteRows;
  for (int row = 0; row < nrow; ++row)
    for (int q = out.ptr[row]; q < out.ptr[row + 1]; ++q) by_edge[out.idx[q]].push_back({row, out.val[q]});
  out.gptr.push_back(0);
  for (const auto& kv : by_edge) {
    out.uedge.push_back(kv.first);
    for (const auto& it : kv.second) {
      out.grow.push_back(it.first);
      out.gval.push_back(it.second);
    }
    out.gptr.push_back(int(out.grow.size()));
  }
  return MT_OK;
}

}  // namespace mt
