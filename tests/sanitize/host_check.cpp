// Host-code check under AddressSanitizer and UndefinedBehaviorSanitizer: the
// library's host-side operator construction (mt_host_ops.cpp), receiver
// set-up (mt_sites.cpp) and the oracle (oracle/sp2_oracle.c) on a small grid
// with partial-tile dimensions.  Built and run by tests/test_host_sanitizers.py
// with -fsanitize=address,undefined -fno-sanitize-recover=all: any invalid
// access or undefined operation aborts with a nonzero status.  The numeric
// checks here are consistency checks between the two codes (the parity of
// each against the reference is pinned by the other tests).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mt_host_ops.hpp"
#include "mt_sites.hpp"

extern "C" {
#include "sp2_oracle.h"
}

namespace {

int g_fail = 0;
void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAIL: %s\n", what);
    ++g_fail;
  }
}

using zc = std::complex<double>;

// the oracle's C complex type, built with the GNU real/imag part extensions
zdouble zmake(double re, double im) {
  zdouble v;
  __real__ v = re;
  __imag__ v = im;
  return v;
}

}  // namespace

int main() {
  const int nx = 11, ny = 9, nzAir = 3, nzEarth = 6, nz = nzAir + nzEarth;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  std::vector<double> dx(nx), dy(ny), dz(nz), lc(size_t(nx) * ny * nzEarth);
  for (auto& v : dx) v = 200.0 + 300.0 * U(rng);
  for (auto& v : dy) v = 200.0 + 300.0 * U(rng);
  for (int k = 0; k < nz; ++k) dz[k] = 50.0 * std::pow(1.4, std::abs(k - nzAir));
  for (auto& v : lc) v = std::log(0.001 + 0.5 * U(rng));
  const double air = std::log(1e-10);

  // ---- library host operator
  mt::HostOperator op;
  mt::Status st = mt::assemble(op, nx, ny, nzAir, nzEarth, dx.data(), dy.data(), dz.data(),
                               lc.data(), air);
  check(st.code == MT_OK, "assemble");
  const mt::Grid& g = op.g;
  std::vector<int> row, col;
  std::vector<double> val;
  mt::stencil_to_csr(op, row, col, val);
  check(int64_t(row.size()) == g.nInt + 1, "csr rows");

  // CSR -> stencil round trip, bit for bit
  mt::HostOperator op2;
  op2.g = g;
  std::vector<double> vms_int(size_t(g.nInt));
  for (int64_t r = 0; r < g.nInt; ++r) vms_int[r] = op.vms1[op.imap[r]];
  st = mt::csr_to_stencil(op2, int(g.nInt), int(val.size()), row.data(), col.data(), 0, val.data(),
                          vms_int.data());
  check(st.code == MT_OK, "csr_to_stencil");
  check(op2.coef.size() == op.coef.size() &&
            std::memcmp(op2.coef.data(), op.coef.data(), op.coef.size() * sizeof(double)) == 0,
        "stencil round trip");
  // a CSR with an out-of-range column is rejected, not read
  {
    std::vector<int> bad = col;
    bad[bad.size() / 2] = int(g.nInt) + 5;
    mt::HostOperator op3;
    op3.g = g;
    st = mt::csr_to_stencil(op3, int(g.nInt), int(val.size()), row.data(), bad.data(), 0,
                            val.data(), vms_int.data());
    check(st.code != MT_OK, "csr_to_stencil rejects a bad column");
  }

  std::vector<double> coefT;
  mt::transpose_stencil(op, coefT);
  check(coefT.size() == op.coef.size(), "transpose size");
  std::vector<zc> dinv;
  mt::dilu_inv_pivots(op, 2 * M_PI * 0.3, dinv);
  check(int64_t(dinv.size()) == g.nE, "pivots size");
  std::vector<double> coefN, wE, dinvN;
  mt::divcor_node_stencil(op, coefN, wE);
  mt::divcor_inv_pivots(op, coefN, dinvN);
  std::vector<double> ereal(size_t(g.nE)), dm(lc.size(), 0.0);
  for (auto& v : ereal) v = U(rng) - 0.5;
  mt::edge_to_model(op, ereal.data(), dm.data());
  std::vector<zc> e_all(size_t(2 * g.nE)), lam(size_t(2 * g.nInt));
  for (auto& v : e_all) v = zc(U(rng) - 0.5, U(rng) - 0.5);
  for (auto& v : lam) v = zc(U(rng) - 0.5, U(rng) - 0.5);
  for (int mode = 0; mode < 2; ++mode)
    mt::grad_assemble(op, 2 * M_PI * 0.3, 2, e_all.data(), lam.data(), mode, dm.data());
  for (double v : dm) check(std::isfinite(v), "gradient finite");

  // ---- receivers: valid sites, then one off the grid (error path)
  {
    const double origin[3] = {0.0, 0.0, 0.0};
    double x0 = 0, y0 = 0;
    for (int i = 0; i < 3; ++i) x0 += dx[i];
    for (int j = 0; j < 3; ++j) y0 += dy[j];
    std::vector<double> xs = {x0 + 17.0, y0 + 29.0, 0.0, x0 + dx[3] + 40.0, y0 + dy[3] + 3.0, 0.0};
    mt::SiteRows rows;
    std::string err;
    int rc = mt::build_site_rows(g, dx.data(), dy.data(), dz.data(), nzAir, origin, 2, xs.data(),
                                 nullptr, nullptr, rows, err);
    check(rc == MT_OK, "site rows");
    check(int(rows.ptr.size()) == 2 * mt::kSiteRows + 1, "site row count");
    // far outside the grid: the interpolation takes the nearest cells
    std::vector<double> off = {-1e7, 1e7, 0.0};
    mt::SiteRows rows2;
    rc = mt::build_site_rows(g, dx.data(), dy.data(), dz.data(), nzAir, origin, 1, off.data(),
                             nullptr, nullptr, rows2, err);
    check(rc == MT_OK || !err.empty(), "site off the grid: rows or an error message");
  }

  // ---- oracle on the same grid: AAii equal to the library's CSR
  sp2o_problem* p = sp2o_create(nx, ny, nzAir, nzEarth, dx.data(), dy.data(), dz.data(),
                                lc.data(), air);
  check(p != nullptr, "oracle create");
  int sz[9];
  sp2o_sizes(p, sz);
  check(sz[1] == g.nInt, "oracle nInt");
  int nr, nc;
  const int nnz = sp2o_csr_nnz(p, 0, &nr, &nc);
  std::vector<int> orow(nr + 1), ocol(nnz);
  std::vector<double> oval(nnz);
  sp2o_csr_get(p, 0, orow.data(), ocol.data(), oval.data());
  {
    // the same operator: every row within 1e-14 of its largest entry (the
    // library assembles the stencil directly, in its own summation order)
    bool same = nr == g.nInt;
    for (int r = 0; same && r < nr; ++r) {
      std::vector<std::pair<int, double>> d;
      double mx = 0.0;
      for (int q = orow[r]; q < orow[r + 1]; ++q) {
        d.push_back({ocol[q], oval[q]});
        mx = std::max(mx, std::abs(oval[q]));
      }
      for (int q = row[r]; q < row[r + 1]; ++q) d.push_back({col[q], -val[q]});
      std::sort(d.begin(), d.end());
      for (size_t q = 0; q < d.size();) {
        double sum = 0.0;
        size_t e = q;
        for (; e < d.size() && d[e].first == d[q].first; ++e) sum += d[e].second;
        if (std::abs(sum) > 1e-14 * mx) same = false;
        q = e;
      }
    }
    check(same, "AAii: oracle == library");
  }
  const int n = sz[1];
  const double om = 2 * M_PI * 0.3;
  sp2o_update_freq(p, om);
  std::vector<zdouble> b(n), x(n), y(n), z(n);
  for (int i = 0; i < n; ++i) b[i] = zmake(U(rng) - 0.5, U(rng) - 0.5);
  sp2o_mult_aii(p, 0, b.data(), y.data());
  sp2o_mult_aii(p, 1, b.data(), y.data());
  sp2o_pc_lsolve(p, b.data(), y.data());
  sp2o_pc_usolve(p, y.data(), z.data());
  sp2o_pc_lsolve_h(p, b.data(), y.data());
  sp2o_pc_usolve_h(p, y.data(), z.data());
  int it = 0;
  std::vector<double> rerr(400);
  std::fill(x.begin(), x.end(), zmake(0.0, 0.0));
  int rc = sp2o_bicg(p, b.data(), x.data(), 200, 1e-10, 1, 0, &it, rerr.data());
  check(rc == 0, "oracle cuBiCG converges");
  std::fill(x.begin(), x.end(), zmake(0.0, 0.0));
  rc = sp2o_qmr(p, b.data(), x.data(), 400, 1e-10, &it, rerr.data());
  check(rc == 0, "oracle QMR converges");
  sp2o_divcor_setup(p);
  std::vector<zdouble> phi(sz[3]), eall(sz[0]), sall(sz[0]);
  sp2o_divc(p, eall.data(), phi.data());
  sp2o_grad(p, phi.data(), eall.data());
  for (auto& v : sall) v = zmake(U(rng) - 0.5, 0.0);
  sp2o_rhs_fwd(p, sall.data(), b.data());
  sp2o_rhs_trn(p, om, sall.data(), b.data());
  sp2o_post_trn(p, b.data(), eall.data());
  std::vector<double> dm2(lc.size(), 0.0);
  std::vector<zdouble> e2(size_t(2) * sz[0]);
  for (auto& v : e2) v = zmake(U(rng) - 0.5, U(rng) - 0.5);
  sp2o_pmultT(p, om, 2, e2.data(), e2.data(), dm2.data());
  std::vector<zdouble> e0(sz[0]);
  int nswap = 0;
  rc = sp2o_bc_x0(nx, ny, nzAir, nzEarth, dx.data(), dy.data(), dz.data(), lc.data(), om, 1,
                  e0.data(), &nswap);
  check(rc == 0, "oracle BC_x0_WS");
  sp2o_destroy(p);

  if (g_fail) {
    std::fprintf(stderr, "%d check(s) failed\n", g_fail);
    return 1;
  }
  std::printf("host sanitizer check: ok (n=%d)\n", n);
  return 0;
}
