// libmtsp2: context, operator upload, BiCGStab drivers and the C ABI
// (include/mtsp2.h).  Kernels live in mt_kernels.hpp, mt_spmv.hpp,
// mt_dilu.hpp and mt_divcor.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>
#include <memory>
#include <string>
#include <vector>

#include "mt_bc2d.hpp"
#include "mt_brick.hpp"
#include "mt_cond.hpp"
#include "mt_datafunc.hpp"
#include "mt_dilu.hpp"
#include "mt_divcor.hpp"
#include "mt_host_ops.hpp"
#include "mt_internal.h"
#include "mt_kernels.hpp"
#include "mt_qmr.hpp"
#include "mt_sens.hpp"
#include "mt_sites.hpp"
#include "mt_spmv.hpp"

using namespace mt;

// ============================================================== context
// Scratch of the allocation-free LmultT (mt_lmultT_noalloc): one array per
// GPU, alive until the process ends, carved into the per-(frequency, site)
// functional coefficients (14 complex each), the data and presence flags (up
// to 8 components each) and the frequencies.  Host code never uses this name
// as a device address: mt_ctx_create resolves the device's copy with
// hipGetSymbolAddress and every use goes through that pointer.
constexpr int kNoAllocMaxSiteFreq = 16384;   // nfreq * nsite at most
constexpr int kNoAllocMaxFreq = 1024;
constexpr size_t kNoAllocCoefOff = 0;
constexpr size_t kNoAllocDataOff = kNoAllocCoefOff + size_t(kNoAllocMaxSiteFreq) * 14 * sizeof(Cd);
constexpr size_t kNoAllocExistOff = kNoAllocDataOff + size_t(kNoAllocMaxSiteFreq) * 8 * sizeof(double);
constexpr size_t kNoAllocOmegaOff = kNoAllocExistOff + size_t(kNoAllocMaxSiteFreq) * 8;
constexpr size_t kNoAllocBytes = kNoAllocOmegaOff + size_t(kNoAllocMaxFreq) * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

struct DeviceOperator {
  bool ready = false;
  double* coef64 = nullptr;  // [15][nE]
  float* coef32 = nullptr;
  double* vms64 = nullptr;   // [nE]
  float* vms32 = nullptr;
  double* coefT64 = nullptr;  // transposed stencil (lazy)
  float* coefT32 = nullptr;
  // row-major copies [nE][16] (15 couplings + vms1) for the k-marching SpMV
  double* row64 = nullptr;
  float* row32 = nullptr;
  double* rowT64 = nullptr;
  float* rowT32 = nullptr;
  int* imap = nullptr;       // interior -> all-edge
  int off[3][15];
};

struct TilePlan {
  int tx = 8, ty = 8, tz = 8;
  int nlev = 0;
  int nlat = 0;
  int ntall = 0;              // tiles over all lattices (dense tile ids 0..ntall-1)
  LatticeInfo lat[3];
  std::vector<int> lvl_off;   // host: start of each tile level in tiles[]
  int4* tiles = nullptr;      // device
  TileShape* shape = nullptr; // device
};

// Divergence-correction state (node lattice), built on first use.
struct DivcorData {
  bool ready = false;
  int64_t nN = 0;
  double* coefN = nullptr;  // [7][nN] VDsG stencil
  double* wE = nullptr;     // [nE] DualFaceA*sigma_E on interior edges
  double* invL = nullptr;   // [nE] 1/EdgeL
  Z* dinv = nullptr;        // [nN] 1/pivot (real, stored complex)
  TilePlan tiles;
};

// receivers of the data functionals (mt_receivers_set)
struct SiteData {
  bool ready = false;
  int nsite = 0, nu = 0;
  int64_t nE = 0;
  int *ptr = nullptr, *idx = nullptr, *uedge = nullptr, *gptr = nullptr, *grow = nullptr;
  double *val = nullptr, *rot = nullptr, *gval = nullptr;
};

static void free_sites(SiteData& s) {
  hipFree(s.ptr); hipFree(s.idx); hipFree(s.uedge); hipFree(s.gptr); hipFree(s.grow);
  hipFree(s.val); hipFree(s.rot); hipFree(s.gval);
  s = SiteData();
}

// Bricks of the mixed solver's inner solves (red-black DILU): the tile
// geometry, the brick -> all-edge map and the brick-ordered FP32 couplings of
// A and A^T (row-major for the SpMV, slots 1..6 for the sweeps).
struct BrickData {
  bool ready = false;
  BrickGeom g{};
  int* bmap = nullptr;                   // [nB] all-edge index, -1 past a partial tile
  float* rows[2] = {nullptr, nullptr};   // [nB][16] of A, A^T
  float* slots[2] = {nullptr, nullptr};  // [6][nB] of A, A^T
};

static void free_brick(BrickData& b) {
  hipFree(b.bmap);
  for (int q = 0; q < 2; ++q) {
    hipFree(b.rows[q]);
    hipFree(b.slots[q]);
  }
  b = BrickData();
}

struct mt_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  HostOperator hop;
  DeviceOperator dop;
  TilePlan tiles;
  TilePlan tiles_rb;  // red-black tile ordering of the same lattices
  BrickData bk;       // bricks of the mixed solver's inner solves
  bool brick = true;  // inner solves on bricks (mixed, red-black DILU, 4-system SpMV)
  int bk_sg = 0;      // brick SpMV: systems per block (0: all active systems)
  bool bk_pairs = true;  // bricks: p-hat, s-hat interleaved by system pair (pair SpMV)
  DivcorData dc;
  SiteData sites;
  std::map<std::string, std::pair<void*, size_t>> ws;  // named device workspaces
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int G_override = 0;           // row chunks per system (mt_set_tuning; 0: rows_total)
  int rows_total = 0;           // row chunks over the active systems (mt_autotune; 0: 2048)
  int num_cu = 256;             // compute units of the device
  int spmv_variant = 3;         // 0 row-parallel SpMV; 1 k-marching for FP32
                                // solves, row-parallel for FP64; 2 k-marching;
                                // 3 multi-system rows for FP32, row-parallel FP64
  int km_ks = 16;               // k-marching SpMV: layers per k segment
  int ms_S = 4;                 // multi-system row SpMV (variant 3): systems per block
  bool rhs_ready = false;       // CCib uploaded for mt_rhs_fwd_dev (workspaces rhs_*)
  // mt_update_cond: the device operator is newer than the host copies in hop
  // (refreshed by ensure_host before host-side use); cond_ready: the
  // sigma-free assembly data are on the device (workspaces cond_*)
  bool host_stale = false;
  bool cond_ready = false;
  // sensitivity kernels: sens_fields -- sigma_E, sigma_N and the node weights
  // of the current model are in the cond_* workspaces; sens_geom -- the
  // sigma-free data (ghost volumes, edge -> row map, CCib transposed) are in
  // the sens_* workspaces
  bool sens_fields = false;
  bool sens_geom = false;
  int rb_variant = 6;           // red-black DILU: 1 persistent single-system
                                // workgroups; 4 four systems per workgroup;
                                // 5 the same in the lean LDS layout; 6 (on
                                // bricks) persistent four-system workgroups
                                // loading the next item while they sweep
  // kernel timing (mt_timing_enable): HIP events on `stream` around every
  // SpMV launch and every DILU sweep of the BiCGStab engine
  bool timing = false;
  struct KTime { long long launches = 0; double ms = 0.0, bytes = 0.0; } kt[2];
  std::vector<hipEvent_t> tev;  // event pool
  int64_t nnz_stencil = -1;     // stored AAii couplings (cached)
  unsigned char* scratch_dev = nullptr;  // this device's copy of `scratch`
};

static int mt_fail(mt_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

// Named workspace, grown on demand, zero-filled on (re)allocation.
static int ws_get(mt_ctx* ctx, const std::string& name, size_t bytes, void** out) {
  auto it = ctx->ws.find(name);
  if (it != ctx->ws.end() && it->second.second >= bytes) {
    *out = it->second.first;
    return MT_OK;
  }
  if (it != ctx->ws.end()) {
    hipFree(it->second.first);
    ctx->ws.erase(it);
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    return mt_fail(ctx, MT_E_OOM, "hipMalloc failed for workspace " + name + " (" +
                                      std::to_string(bytes) + " bytes)");
  }
  MT_HIP_CHECK(ctx, hipMemsetAsync(p, 0, bytes ? bytes : 16, ctx->stream));
  // the stream is non-blocking: callers upload into a fresh workspace with
  // synchronous copies on the null stream, which must not race the clear
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->ws[name] = {p, bytes};
  *out = p;
  return MT_OK;
}
template <class P>
static int ws(mt_ctx* ctx, const std::string& name, size_t count, P** out) {
  void* p = nullptr;
  int rc = ws_get(ctx, name, count * sizeof(P), &p);
  *out = static_cast<P*>(p);
  return rc;
}

static void free_dop(DeviceOperator& d) {
  hipFree(d.coef64); hipFree(d.coef32); hipFree(d.vms64); hipFree(d.vms32);
  hipFree(d.coefT64); hipFree(d.coefT32); hipFree(d.imap);
  hipFree(d.row64); hipFree(d.row32); hipFree(d.rowT64); hipFree(d.rowT32);
  d = DeviceOperator();
}

// neighbour offsets of each slot relative to its base address
static void build_offsets(const Grid& g, int off[3][15]) {
  int A[3], B[3];
  for (int c = 0; c < 3; ++c) {
    int d;
    g.dims(c, A[c], B[c], d);
  }
  for (int c = 0; c < 3; ++c) {
    Nb nb[kSlots];
    stencil_neighbours(c, nb);
    for (int s = 0; s < kSlots; ++s) {
      const int cb = nb[s].comp;
      off[c][s] = nb[s].di + A[cb] * (nb[s].dj + B[cb] * nb[s].dk);
    }
  }
}

template <class V>
static StencilView<V> view(const mt_ctx* ctx, bool transpose) {
  StencilView<V> A;
  const Grid& g = ctx->hop.g;
  if constexpr (sizeof(V) == 8) {
    A.coef = reinterpret_cast<const V*>(transpose ? ctx->dop.coefT64 : ctx->dop.coef64);
    A.rowc = reinterpret_cast<const V*>(transpose ? ctx->dop.rowT64 : ctx->dop.row64);
    A.vms1 = reinterpret_cast<const V*>(ctx->dop.vms64);
  } else {
    A.coef = reinterpret_cast<const V*>(transpose ? ctx->dop.coefT32 : ctx->dop.coef32);
    A.rowc = reinterpret_cast<const V*>(transpose ? ctx->dop.rowT32 : ctx->dop.row32);
    A.vms1 = reinterpret_cast<const V*>(ctx->dop.vms32);
  }
  A.nE = g.nE;
  A.nXi = g.nXi; A.nYi = g.nYi; A.nInt = g.nInt;
  A.nx = g.nx; A.ny = g.ny; A.nz = g.nz;
  A.nXe = g.nXe; A.nYe = g.nYe;
  std::memcpy(A.off, ctx->dop.off, sizeof(A.off));
  return A;
}

// Tile plan for the DILU sweeps: the tiles of all lattices grouped by tile
// level ti+tj+tk.
// rb = true: two groups instead, the red (ti+tj+tk even) and the black tiles.
static int build_tile_plan(mt_ctx* ctx, TilePlan& tp, int nlat, const LatticeInfo* lat,
                           bool rb = false) {
  hipFree(tp.tiles);
  hipFree(tp.shape);
  tp.tiles = nullptr;
  tp.shape = nullptr;
  tp.nlat = nlat;
  std::vector<std::vector<int4>> byLev;
  int tbase = 0;
  for (int c = 0; c < nlat; ++c) {
    tp.lat[c] = lat[c];
    const int NTx = (lat[c].A + tp.tx - 1) / tp.tx, NTy = (lat[c].B + tp.ty - 1) / tp.ty,
              NTz = (lat[c].C + tp.tz - 1) / tp.tz;
    tp.lat[c].ntx = NTx;
    tp.lat[c].nty = NTy;
    tp.lat[c].tbase = tbase;
    tbase += NTx * NTy * NTz;
    for (int tk = 0; tk < NTz; ++tk)
      for (int tj = 0; tj < NTy; ++tj)
        for (int ti = 0; ti < NTx; ++ti) {
          const int L = rb ? ((ti + tj + tk) & 1) : ti + tj + tk;
          if ((int)byLev.size() <= L) byLev.resize(L + 1);
          byLev[L].push_back(make_int4(c, ti, tj, tk));
        }
  }
  tp.ntall = tbase;
  tp.nlev = (int)byLev.size();
  tp.lvl_off.assign(tp.nlev + 1, 0);
  std::vector<int4> flat;
  for (int L = 0; L < tp.nlev; ++L) {
    tp.lvl_off[L] = (int)flat.size();
    flat.insert(flat.end(), byLev[L].begin(), byLev[L].end());
  }
  tp.lvl_off[tp.nlev] = (int)flat.size();
  TileShape sh{};
  sh.tx = tp.tx; sh.ty = tp.ty; sh.tz = tp.tz;
  sh.npts = tp.tx * tp.ty * tp.tz;
  sh.nlev = tp.tx + tp.ty + tp.tz - 2;
  int w = 0;
  for (int h = 0; h < sh.nlev; ++h) {
    sh.lvl[h] = w;
    for (int lk = 0; lk < tp.tz; ++lk)
      for (int lj = 0; lj < tp.ty; ++lj)
        for (int li = 0; li < tp.tx; ++li)
          if (li + lj + lk == h) sh.pts[w++] = unsigned(li) | (unsigned(lj) << 8) | (unsigned(lk) << 16);
  }
  sh.lvl[sh.nlev] = w;
  if (tp.tx == 8 && tp.ty == 8 && tp.tz == 8) {
    // lane order of the lean red-black sweeps: fill each 16-lane group of a
    // level with points whose box slots differ modulo 16 (8-byte bank pairs)
    // while such points remain, then with the rest in level order
    for (int h = 0; h < sh.nlev; ++h) {
      std::vector<unsigned int> rem(sh.pts + sh.lvl[h], sh.pts + sh.lvl[h + 1]);
      int slot = sh.lvl[h];
      auto put = [&](unsigned int v) {
        const int li = v & 0xff, lj = (v >> 8) & 0xff, lk = (v >> 16) & 0xff;
        sh.rpts[slot] = unsigned(li) | (unsigned(lj) << 3) | (unsigned(lk) << 6) | (unsigned(slot) << 9);
        sh.rslot[li + 8 * (lj + 8 * lk)] = (unsigned short)slot;
        ++slot;
      };
      while (!rem.empty()) {
        bool used[16] = {};
        int n = 0;
        for (size_t r = 0; r < rem.size() && n < 16;) {
          const unsigned int v = rem[r];
          const int b = rb_box(v & 0xff, (v >> 8) & 0xff, (v >> 16) & 0xff) & 15;
          if (used[b]) { ++r; continue; }
          used[b] = true;
          put(v);
          rem.erase(rem.begin() + r);
          ++n;
        }
        for (; n < 16 && !rem.empty(); ++n) {
          put(rem.front());
          rem.erase(rem.begin());
        }
      }
    }
  }
  // the red-black sweeps give each lane one point per hyperplane
  for (int h = 0; h < sh.nlev; ++h)
    if (sh.lvl[h + 1] - sh.lvl[h] > 64)
      return mt_fail(ctx, MT_E_STATE, "tile hyperplane wider than a wave");
  if (tp.tx == 8 && tp.ty == 8 && tp.tz == 8)
    for (int h = 0; h < sh.nlev; ++h)
      for (int l = 0; l < 64; ++l)
        sh.lpts[h][l] = sh.lvl[h] + l < sh.lvl[h + 1] ? sh.rpts[sh.lvl[h] + l] : 0xffffffffu;

  MT_HIP_CHECK(ctx, hipMalloc(&tp.tiles, sizeof(int4) * std::max<size_t>(1, flat.size())));
  MT_HIP_CHECK(ctx, hipMalloc(&tp.shape, sizeof(TileShape)));
  MT_HIP_CHECK(ctx, hipMemcpy(tp.tiles, flat.data(), sizeof(int4) * flat.size(), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(tp.shape, &sh, sizeof(TileShape), hipMemcpyHostToDevice));
  return MT_OK;
}

// The three interior edge lattices (x, y, z blocks of PC_setup).
static void edge_lattices(const Grid& g, LatticeInfo lat[3]) {
  lat[0] = {g.nx, g.ny - 1, g.nz - 1, 0, 1, 1, g.nx, g.ny + 1, 0};
  lat[1] = {g.nx - 1, g.ny, g.nz - 1, 1, 0, 1, g.nx + 1, g.ny, g.nXe};
  lat[2] = {g.nx - 1, g.ny - 1, g.nz, 1, 1, 0, g.nx + 1, g.ny + 1, g.nXe + g.nYe};
}

static void free_workspaces(mt_ctx* ctx) {
  hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->ws) hipFree(kv.second.first);
  ctx->ws.clear();
}

static int upload_operator(mt_ctx* ctx) {
  const Grid& g = ctx->hop.g;
  free_dop(ctx->dop);
  // Padded vectors rely on zero boundary entries and a zero tail; buffers laid
  // out for another grid would carry stale values there, so drop them all.
  free_workspaces(ctx);
  ctx->rhs_ready = false;
  ctx->cond_ready = false;
  ctx->sens_fields = false;
  ctx->sens_geom = false;
  ctx->host_stale = false;
  ctx->nnz_stencil = -1;  // counted again for the new operator on first use
  hipFree(ctx->dc.coefN); hipFree(ctx->dc.wE); hipFree(ctx->dc.invL); hipFree(ctx->dc.dinv);
  hipFree(ctx->dc.tiles.tiles); hipFree(ctx->dc.tiles.shape);
  ctx->dc = DivcorData();
  free_brick(ctx->bk);
  DeviceOperator& d = ctx->dop;
  const size_t nc = size_t(kSlots) * g.nE;
  std::vector<float> tmp(nc);
  MT_HIP_CHECK(ctx, hipMalloc(&d.coef64, nc * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&d.coef32, nc * sizeof(float)));
  MT_HIP_CHECK(ctx, hipMalloc(&d.vms64, g.nE * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&d.vms32, g.nE * sizeof(float)));
  MT_HIP_CHECK(ctx, hipMalloc(&d.imap, g.nInt * sizeof(int)));
  MT_HIP_CHECK(ctx, hipMemcpy(d.coef64, ctx->hop.coef.data(), nc * sizeof(double), hipMemcpyHostToDevice));
  for (size_t q = 0; q < nc; ++q) tmp[q] = float(ctx->hop.coef[q]);
  MT_HIP_CHECK(ctx, hipMemcpy(d.coef32, tmp.data(), nc * sizeof(float), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(d.vms64, ctx->hop.vms1.data(), g.nE * sizeof(double), hipMemcpyHostToDevice));
  for (int64_t q = 0; q < g.nE; ++q) tmp[q] = float(ctx->hop.vms1[q]);
  MT_HIP_CHECK(ctx, hipMemcpy(d.vms32, tmp.data(), g.nE * sizeof(float), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(d.imap, ctx->hop.imap.data(), g.nInt * sizeof(int), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMalloc(&d.row64, size_t(16) * g.nE * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&d.row32, size_t(16) * g.nE * sizeof(float)));
  k_rows_from_slots<double><<<2048, 256, 0, ctx->stream>>>(d.coef64, d.vms64, g.nE, d.row64);
  k_rows_from_slots<float><<<2048, 256, 0, ctx->stream>>>(d.coef32, d.vms32, g.nE, d.row32);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  build_offsets(g, d.off);
  // the k-marching SpMV carries its own copy of the neighbour table
  for (int c = 0; c < 3; ++c) {
    Nb nb[kSlots];
    stencil_neighbours(c, nb);
    for (int sl = 0; sl < kSlots; ++sl)
      if (km_nb(c, sl, 0) != nb[sl].comp || km_nb(c, sl, 1) != nb[sl].di ||
          km_nb(c, sl, 2) != nb[sl].dj || km_nb(c, sl, 3) != nb[sl].dk)
        return mt_fail(ctx, MT_E_STATE, "k-marching SpMV neighbour table mismatch");
  }
  d.ready = true;
  LatticeInfo lat[3];
  edge_lattices(g, lat);
  int rc = build_tile_plan(ctx, ctx->tiles, 3, lat);
  if (rc) return rc;
  return build_tile_plan(ctx, ctx->tiles_rb, 3, lat, true);
}

// After mt_update_cond the host copies of the sigma-dependent operator data
// (coefficients, sigma_E/N, the RHS-form GDii, the earth classes) are rebuilt
// once, on first host-side use, by the host assembly of the current model
// (which gives the same coefficients as the device, bit for bit).
static int ensure_host(mt_ctx* ctx) {
  if (!ctx->host_stale) return MT_OK;
  HostOperator& H = ctx->hop;
  const std::vector<double> dx = H.dx, dy = H.dy, dz = H.dz, lc = H.logcond;
  Status st = assemble(H, H.g.nx, H.g.ny, H.nzAir, H.nzEarth, dx.data(), dy.data(), dz.data(),
                       lc.data(), H.airLog);
  if (st.code) return mt_fail(ctx, st.code, st.msg);
  ctx->host_stale = false;
  return MT_OK;
}

static int ensure_transpose(mt_ctx* ctx) {
  if (ctx->dop.coefT64) return MT_OK;
  { int rc = ensure_host(ctx); if (rc) return rc; }
  const Grid& g = ctx->hop.g;
  std::vector<double> cT;
  transpose_stencil(ctx->hop, cT);
  const size_t nc = size_t(kSlots) * g.nE;
  std::vector<float> t32(nc);
  for (size_t q = 0; q < nc; ++q) t32[q] = float(cT[q]);
  MT_HIP_CHECK(ctx, hipMalloc(&ctx->dop.coefT64, nc * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&ctx->dop.coefT32, nc * sizeof(float)));
  MT_HIP_CHECK(ctx, hipMemcpy(ctx->dop.coefT64, cT.data(), nc * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(ctx->dop.coefT32, t32.data(), nc * sizeof(float), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMalloc(&ctx->dop.rowT64, size_t(16) * g.nE * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&ctx->dop.rowT32, size_t(16) * g.nE * sizeof(float)));
  k_rows_from_slots<double><<<2048, 256, 0, ctx->stream>>>(ctx->dop.coefT64, ctx->dop.vms64, g.nE,
                                                           ctx->dop.rowT64);
  k_rows_from_slots<float><<<2048, 256, 0, ctx->stream>>>(ctx->dop.coefT32, ctx->dop.vms32, g.nE,
                                                          ctx->dop.rowT32);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

// The bricks of the inner solves (geometry and map once per grid, the
// couplings of A or A^T once per operator).
static int ensure_brick(mt_ctx* ctx, bool transpose) {
  BrickData& B = ctx->bk;
  if (!B.ready) {
    const TilePlan& tp = ctx->tiles_rb;
    if (tp.tx != 8 || tp.ty != 8 || tp.tz != 8 || tp.nlat != 3)
      return mt_fail(ctx, MT_E_STATE, "bricks need the 8x8x8 red-black tile plan");
    if (int64_t(tp.ntall) * kBkPts * 16 >= (int64_t(1) << 31))
      return mt_fail(ctx, MT_E_ARG, "grid too large for 32-bit brick indices");
    BrickGeom& G = B.g;
    for (int c = 0; c < 3; ++c) {
      const LatticeInfo& L = tp.lat[c];
      G.A[c] = L.A; G.B[c] = L.B; G.C[c] = L.C;
      G.oi[c] = L.oi; G.oj[c] = L.oj; G.ok[c] = L.ok;
      G.ntx[c] = L.ntx; G.nty[c] = L.nty; G.ntz[c] = (L.C + 7) / 8;
      G.tbase[c] = L.tbase;
    }
    G.tbase[3] = tp.ntall;
    G.nB = int64_t(tp.ntall) * kBkPts;
    std::vector<int> bm(size_t(G.nB), -1);
    for (int c = 0; c < 3; ++c) {
      const LatticeInfo& L = tp.lat[c];
      for (int tk = 0; tk < G.ntz[c]; ++tk)
        for (int tj = 0; tj < G.nty[c]; ++tj)
          for (int ti = 0; ti < G.ntx[c]; ++ti) {
            const int64_t t = L.tbase + ti + int64_t(L.ntx) * (tj + int64_t(L.nty) * tk);
            for (int p = 0; p < kBkPts; ++p) {
              const int i = ti * 8 + (p & 7), j = tj * 8 + ((p >> 3) & 7), k = tk * 8 + (p >> 6);
              if (i >= L.A || j >= L.B || k >= L.C) continue;
              bm[size_t(t * kBkPts + p)] =
                  int(L.base + (i + L.oi) + int64_t(L.a) * ((j + L.oj) + int64_t(L.b) * (k + L.ok)));
            }
          }
    }
    MT_HIP_CHECK(ctx, hipMalloc(&B.bmap, bm.size() * sizeof(int)));
    MT_HIP_CHECK(ctx, hipMemcpy(B.bmap, bm.data(), bm.size() * sizeof(int), hipMemcpyHostToDevice));
    B.ready = true;
  }
  const int q = transpose ? 1 : 0;
  if (!B.rows[q]) {
    if (transpose) { int rc = ensure_transpose(ctx); if (rc) return rc; }
    const StencilView<float> A = view<float>(ctx, transpose);
    const int64_t nB = B.g.nB;
    MT_HIP_CHECK(ctx, hipMalloc(&B.rows[q], size_t(nB) * 16 * sizeof(float)));
    MT_HIP_CHECK(ctx, hipMalloc(&B.slots[q], size_t(nB) * 6 * sizeof(float)));
    k_bk_rows<float><<<4096, BS, 0, ctx->stream>>>(B.bmap, nB, A.rowc, B.rows[q]);
    k_bk_slots<float><<<2048, BS, 0, ctx->stream>>>(B.bmap, nB, A.nE, A.coef, B.slots[q]);
    MT_HIP_CHECK(ctx, hipGetLastError());
    MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  }
  return MT_OK;
}

// ====================================================== preconditioner data
// Inverse pivots (DILU) or inverse diagonal (Jacobi) per unique frequency.
// The DILU pivots of the transposed operator are those of A^T's blocks.
template <class T>
static int build_dinv(mt_ctx* ctx, int pc, bool transpose, const std::vector<double>& freqs,
                      C2<T>** out) {
  const Grid& g = ctx->hop.g;
  const int nf = (int)freqs.size();
  std::string key = std::string("dinv") + (sizeof(T) == 8 ? "64" : "32");
  int rc = ws(ctx, key, size_t(nf) * g.nE, out);
  if (rc) return rc;
  if (pc == PCK_NONE) return MT_OK;
  if (pc == PCK_JACOBI) {
    double* dfreq;
    if ((rc = ws(ctx, "freqs", nf, &dfreq))) return rc;
    MT_HIP_CHECK(ctx, hipMemcpyAsync(dfreq, freqs.data(), nf * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    k_jacobi_dinv<T, double><<<1024, 256, 0, ctx->stream>>>(view<double>(ctx, transpose), dfreq, nf, *out);
    MT_HIP_CHECK(ctx, hipGetLastError());
    return MT_OK;
  }
  // DILU pivots on the device, in FP64, for all frequencies in one
  // level-scheduled sweep (PC_setup per frequency, as UpdateFreq does)
  const bool rb = pc == PCK_DILU_RB;
  const TilePlan& tp = rb ? ctx->tiles_rb : ctx->tiles;
  double* dfreq;
  if ((rc = ws(ctx, "freqs", nf, &dfreq))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(dfreq, freqs.data(), nf * sizeof(double), hipMemcpyHostToDevice,
                                   ctx->stream));
  Z* p64;
  if (sizeof(T) == 8) {
    p64 = reinterpret_cast<Z*>(*out);
  } else if ((rc = ws(ctx, "piv64", size_t(nf) * g.nE, &p64))) {
    return rc;
  }
  PivotArgs pa;
  pa.coef = transpose ? ctx->dop.coefT64 : ctx->dop.coef64;
  pa.cstride = g.nE;
  pa.vms1 = ctx->dop.vms64;
  for (int c = 0; c < 3; ++c) pa.lat[c] = tp.lat[c];
  pa.omegas = dfreq;
  pa.nf = nf;
  pa.dinv = p64;
  for (int L = 0; L < tp.nlev; ++L) {
    pa.tiles = tp.tiles + tp.lvl_off[L];
    pa.ntiles = tp.lvl_off[L + 1] - tp.lvl_off[L];
    if (pa.ntiles == 0) continue;
    if (!rb)
      k_dilu_pivots<8, 8, 8><<<pa.ntiles * nf, 64, 0, ctx->stream>>>(pa, tp.shape);
    else if (L == 0)
      k_dilu_pivots_rb<8, 8, 8, false><<<pa.ntiles * nf, 64, 0, ctx->stream>>>(pa, tp.shape);
    else
      k_dilu_pivots_rb<8, 8, 8, true><<<pa.ntiles * nf, 64, 0, ctx->stream>>>(pa, tp.shape);
  }
  if (sizeof(T) != 8) {
    const int64_t n = int64_t(nf) * g.nE;
    k_cast_pivots<T><<<int(std::min<int64_t>(4096, (n + 255) / 256)), 256, 0, ctx->stream>>>(p64, *out, n);
  }
  MT_HIP_CHECK(ctx, hipGetLastError());
  return MT_OK;
}

// DILU apply for the active systems with status `req`: out = U^-1 L^-1 in,
// over the lattices of `tp` with coefficients coef[slot*cstride + addr]: the
// forward (lower) sweep over the tile levels, then the backward (upper) sweep
// in place.
template <class T, class V>
static DiluArgs<T, V> dilu_args(const TilePlan& tp, const V* coef, int64_t cstride, int64_t vstride,
                                const int* act, int nact, const SysScal* sc, const C2<T>* in,
                                C2<T>* out, const C2<T>* dinv, int req) {
  DiluArgs<T, V> a;
  a.coef = coef;
  a.cstride = cstride;
  a.vstride = vstride;
  for (int c = 0; c < 3; ++c) a.lat[c] = tp.lat[c];
  a.act = act; a.nact = nact; a.sc = sc; a.in = in; a.out = out; a.dinv = dinv;
  a.in2 = nullptr;
  a.status_req = req;
  a.tiles = nullptr;
  a.ntiles = 0;
  return a;
}

template <class T, class V>
static void dilu_fwd(mt_ctx* ctx, const TilePlan& tp, DiluArgs<T, V> a) {
  for (int L = 0; L < tp.nlev; ++L) {
    a.tiles = tp.tiles + tp.lvl_off[L];
    a.ntiles = tp.lvl_off[L + 1] - tp.lvl_off[L];
    k_dilu_ro<T, V, true><<<a.ntiles * a.nact, 64, 0, ctx->stream>>>(a, tp.shape);
  }
}

template <class T, class V>
static void dilu_bwd(mt_ctx* ctx, const TilePlan& tp, DiluArgs<T, V> a) {
  a.in = a.out;
  for (int L = tp.nlev - 1; L >= 0; --L) {
    a.tiles = tp.tiles + tp.lvl_off[L];
    a.ntiles = tp.lvl_off[L + 1] - tp.lvl_off[L];
    k_dilu_ro<T, V, false><<<a.ntiles * a.nact, 64, 0, ctx->stream>>>(a, tp.shape);
  }
}

// Red-black tile ordering: red forward, black forward+backward, red
// backward, each one persistent launch.
static int rb_grid(mt_ctx* ctx, int nwork) {
  // 8 blocks per CU (the LDS of one block is 18-28 KB), a multiple of 8 XCDs
  int g = std::min(nwork, ctx->num_cu * 8);
  g = std::max(8, (g + 7) / 8 * 8);
  return g;
}

template <class T, class V>
static void dilu_rb_apply(mt_ctx* ctx, const TilePlan& tp, DiluArgs<T, V> a) {
  if (!a.xf) {  // every caller provides the packed-face buffer (rb_faces)
    ctx->err = "red-black DILU launched without its face buffer";
    return;
  }
  const int nl = tp.nlev;  // 2: red, black (a lattice of one tile has no black level)
  for (int step = 0; step < 3; ++step) {
    const int L = (step == 1) ? 1 : 0;
    if (L >= nl) continue;
    a.tiles = tp.tiles + tp.lvl_off[L];
    a.ntiles = tp.lvl_off[L + 1] - tp.lvl_off[L];
    if (a.ntiles == 0) continue;
    if (ctx->rb_variant >= 4) {
      constexpr int NW = 4;
      const int64_t nwork = int64_t(a.ntiles) * ((a.nact + NW - 1) / NW);
      const int g = int(8 * ((nwork + 7) / 8));
      if (ctx->rb_variant >= 5) {
        if (step == 0) k_dilu_rbl<T, V, RB_RED_FWD, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else if (step == 1) k_dilu_rbl<T, V, RB_BLACK_FB, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else k_dilu_rbl<T, V, RB_RED_BWD, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
      } else {
        if (step == 0) k_dilu_rbm<T, V, RB_RED_FWD, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else if (step == 1) k_dilu_rbm<T, V, RB_BLACK_FB, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else k_dilu_rbm<T, V, RB_RED_BWD, NW><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
      }
      continue;
    }
    const int g = rb_grid(ctx, a.ntiles * a.nact);
    if (step == 0) k_dilu_rb2<T, V, RB_RED_FWD><<<g, 64, 0, ctx->stream>>>(a, tp.shape);
    else if (step == 1) k_dilu_rb2<T, V, RB_BLACK_FB><<<g, 64, 0, ctx->stream>>>(a, tp.shape);
    else k_dilu_rb2<T, V, RB_RED_BWD><<<g, 64, 0, ctx->stream>>>(a, tp.shape);
  }
}

// Packed i-face buffer of the red-black sweeps for systems 0..nsys-1
// (DiluArgs::xf); sets a.xf / a.xfsys / a.xfhalf.
// The red-black sweeps on bricks (k_dilu_bk): the three launches of
// dilu_rb_apply, four systems of one tile per workgroup.
template <class T, class V>
static void dilu_bk_apply(mt_ctx* ctx, const TilePlan& tp, DiluArgs<T, V> a) {
  if (!a.xf) {
    ctx->err = "red-black DILU launched without its face buffer";
    return;
  }
  constexpr int NW = 4;
  for (int step = 0; step < 3; ++step) {
    const int L = (step == 1) ? 1 : 0;
    if (L >= tp.nlev) continue;
    a.tiles = tp.tiles + tp.lvl_off[L];
    a.ntiles = tp.lvl_off[L + 1] - tp.lvl_off[L];
    if (a.ntiles == 0) continue;
    const int64_t nwork = int64_t(a.ntiles) * ((a.nact + NW - 1) / NW);
    const int g = int(8 * ((nwork + 7) / 8));
    // the result pair-interleaved (zout: FP32 inner solves with paired x)
    auto launch = [&](auto zp) {
      constexpr bool ZP = decltype(zp)::value;
      if (ctx->rb_variant == 6) {
        // persistent: three workgroups per CU (their LDS), a multiple of 8 XCDs
        const int gp = std::min(g, std::max(8, ctx->num_cu * 3 / 8 * 8));
        if (step == 0 && a.pu_r)   // (the p-update folded into the red forward sweep)
          k_dilu_bkp<T, V, RB_RED_FWD, NW, ZP, true><<<gp, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else if (step == 0) k_dilu_bkp<T, V, RB_RED_FWD, NW, ZP><<<gp, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else if (step == 1) k_dilu_bkp<T, V, RB_BLACK_FB, NW, ZP><<<gp, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        else k_dilu_bkp<T, V, RB_RED_BWD, NW, ZP><<<gp, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
        return;
      }
      if (step == 0) k_dilu_bk<T, V, RB_RED_FWD, NW, ZP><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
      else if (step == 1) k_dilu_bk<T, V, RB_BLACK_FB, NW, ZP><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
      else k_dilu_bk<T, V, RB_RED_BWD, NW, ZP><<<g, 64 * NW, 0, ctx->stream>>>(a, tp.shape);
    };
    if constexpr (sizeof(T) == 4) {
      if (a.zout) {
        launch(std::true_type());
        continue;
      }
    }
    launch(std::false_type());
  }
}

template <class T, class V>
static int rb_faces(mt_ctx* ctx, int nsys, DiluArgs<T, V>& a) {
  const TilePlan& tp = ctx->tiles_rb;
  a.xfsys = int64_t(tp.ntall) * kRbFaces;
  a.xfhalf = int64_t(nsys) * a.xfsys;
  return ws(ctx, sizeof(T) == 8 ? "rb_faces64" : "rb_faces32", size_t(2 * a.xfhalf), &a.xf);
}

template <class T, class V>
static void dilu_sweeps(mt_ctx* ctx, const TilePlan& tp, const V* coef, int64_t cstride,
                        int64_t vstride, const int* act, int nact, const SysScal* sc,
                        const C2<T>* in, C2<T>* out, const C2<T>* dinv, int req) {
  const DiluArgs<T, V> a =
      dilu_args<T, V>(tp, coef, cstride, vstride, act, nact, sc, in, out, dinv, req);
  dilu_fwd<T, V>(ctx, tp, a);
  dilu_bwd<T, V>(ctx, tp, a);
}

template <class T, class V>
static void dilu_apply(mt_ctx* ctx, bool transpose, const int* act, int nact, const SysScal* sc,
                       const C2<T>* in, C2<T>* out, const C2<T>* dinv, int req) {
  const Grid& g = ctx->hop.g;
  dilu_sweeps<T, V>(ctx, ctx->tiles, view<V>(ctx, transpose).coef, g.nE, g.nE, act, nact, sc,
                    in, out, dinv, req);
}

static int64_t stencil_nnz(mt_ctx* ctx) {
  if (ctx->nnz_stencil < 0) {
    if (ensure_host(ctx)) return 0;
    const Grid& g = ctx->hop.g;
    int64_t z = 0;
    for (int e : ctx->hop.imap)
      for (int s = 0; s < kSlots; ++s) z += ctx->hop.coef[size_t(s) * g.nE + e] != 0.0;
    ctx->nnz_stencil = z;
  }
  return ctx->nnz_stencil;
}

// ======================================================= BiCGStab engine
// Row chunks per system of the row-parallel kernels (vector updates, row
// SpMV): a fixed count (mt_set_tuning), else a total over the active systems
// (mt_autotune's choice, default 2048).
static int pick_G(const mt_ctx* ctx, int nact) {
  if (ctx->G_override > 0) return ctx->G_override;
  const int total = ctx->rows_total > 0 ? ctx->rows_total : 2048;
  int G = total / std::max(1, nact);
  G = std::max(8, (G / 8) * 8);
  return G;
}

// Chunks of the brick SpMV: at most kBtChunks, a multiple of 8.
static int brick_chunks(const BrickGeom& g) { return std::min(kBtChunks, (g.tbase[3] + 7) / 8 * 8); }

// Per-system partial sums any reduction kernel may store: the row kernels'
// chunks, the k-marching SpMV's work items, the brick SpMV's chunks.
static int partials_G(const mt_ctx* ctx, const KmGeom& km) {
  return std::max(std::max(pick_G(ctx, 1), km.Gpad), kBtChunks);
}

// Work items of the k-marching SpMV for the grid of ctx.
static KmGeom km_geom(const mt_ctx* ctx) {
  const Grid& g = ctx->hop.g;
  KmGeom k;
  k.nti = (g.nx + 1 + KM_TI - 1) / KM_TI;
  k.ntj = (g.ny + 1 + KM_TJ - 1) / KM_TJ;
  k.ks = std::max(1, ctx->km_ks);
  k.nseg = (g.nz + 1 + k.ks - 1) / k.ks;
  k.Greal = k.nti * k.ntj * k.nseg;
  k.Gpad = (k.Greal + 7) / 8 * 8;
  return k;
}

template <class T, class V>
struct Engine {
  mt_ctx* ctx;
  IterArgs<T, V> a;
  int nsys;
  bool transpose;
  bool use_graph;
  bool recheck = true;  // FP64 solves re-check the true residual; inner solves do not
  hipGraphExec_t gexec = nullptr;
  int g_nact = -1, g_chunk = -1;
  bool use_km = true;  // k-marching SpMV (set from the context at solve start)
  KmGeom km{};
  int ms = 0;          // >0: the multi-system row SpMV with this many systems per block
  C2<T>* xf = nullptr;  // red-black DILU packed faces (rb_faces)
  int64_t xfsys = 0, xfhalf = 0;
  // inner solves on bricks: the geometry, the brick-ordered couplings (rows
  // for the SpMV, slots 1..6 for the sweeps); a.nE is then the brick length
  bool brick = false;
  BrickGeom bg{};
  const V* bk_rows = nullptr;
  const V* bk_slots = nullptr;
  // paired x (a.ppair > 0, FP32 on bricks): the active systems' pairs
  // {2m, 2m+1, active index of each or -1}, rebuilt with the active list
  int4* d_pairs = nullptr;
  std::vector<int4> h_pairs;
  int npair = 0, g_npair = -1;
  void set_pairs(const std::vector<int>& act) {
    h_pairs.clear();
    for (int si = 0; si < int(act.size()); ++si) {
      const int s = act[si], m = s >> 1;
      if (h_pairs.empty() || h_pairs.back().x != 2 * m) h_pairs.push_back({2 * m, 2 * m + 1, -1, -1});
      if (s & 1) h_pairs.back().w = si;
      else h_pairs.back().z = si;
    }
    npair = int(h_pairs.size());
    // the source outlives the copy: callers synchronise before the next set
    if (npair)
      hipMemcpyAsync(d_pairs, h_pairs.data(), npair * sizeof(int4), hipMemcpyHostToDevice, ctx->stream);
  }
  // timing: event pairs (family, first, second) recorded by one chunk; with a
  // graph they are recorded by the graph's event nodes on every replay
  struct TPair { int fam, e0, e1; };
  std::vector<TPair> tpairs;
  int tnext = 0;
  std::vector<int> last_iter;

  int trec() {
    hipEventRecord(ctx->tev[tnext], ctx->stream);
    return tnext++;
  }
  void apply_pc(const DiluArgs<T, V>& d) {
    if (a.pc == PCK_DILU_RB) {
      dilu_rb_apply<T, V>(ctx, ctx->tiles_rb, d);
    } else {
      dilu_fwd<T, V>(ctx, ctx->tiles, d);
      dilu_bwd<T, V>(ctx, ctx->tiles, d);
    }
  }
  // out = M^-1 (in - alpha*in2) (in2 may be null); with zout (bricks, paired
  // x) the result goes to zout pair-interleaved and out holds the red tiles'
  // intermediate w
  // pu: the red forward sweep also forms p (k_pupdate) for the red tiles
  void precond(const C2<T>* in, C2<T>* out, const C2<T>* in2 = nullptr, C2<T>* zout = nullptr,
               bool pu = false) {
    if (a.pc != PCK_DILU && a.pc != PCK_DILU_RB) return;
    const Grid& g = ctx->hop.g;
    const TilePlan& tp = a.pc == PCK_DILU_RB ? ctx->tiles_rb : ctx->tiles;
    DiluArgs<T, V> d =
        brick ? dilu_args<T, V>(tp, bk_slots, a.nE, a.nE, a.act, a.nact, a.sc, in, out, a.dinv, ST_RUN)
              : dilu_args<T, V>(tp, view<V>(ctx, transpose).coef, g.nE, g.nE, a.act, a.nact, a.sc,
                                in, out, a.dinv, ST_RUN);
    d.in2 = in2;
    d.sout = (brick && in2) ? a.s : nullptr;
    d.zout = brick ? zout : nullptr;
    d.zpair = a.ppair;
    if (pu) {
      d.pu_r = a.r;
      d.pu_v = a.v;
      d.pu_p = a.p;
      d.pu_rt = a.rt;
    }
    d.xf = xf;
    d.xfsys = xfsys;
    d.xfhalf = xfhalf;
    const int e0 = ctx->timing ? trec() : -1;
    if (brick) dilu_bk_apply<T, V>(ctx, tp, d);
    else apply_pc(d);
    if (ctx->timing) tpairs.push_back({MT_KT_PRECOND, e0, trec()});
  }
  // the SpMV and its dot partials; returns the partial count per system
  template <int WHICH>
  int spmv(int nblk) {
    const int e0 = ctx->timing ? trec() : -1;
    int G;
    if (brick) {
      // one block per chunk of tiles for all systems: at most 2048 chunks
      // (the partials buffer), a multiple of 8 (XCDs)
      IterArgs<T, V> b = a;
      b.G = brick_chunks(bg);
      const int sg = ctx->bk_sg > 0 ? ctx->bk_sg : a.nact;
      if constexpr (std::is_same<T, float>::value && std::is_same<V, float>::value) {
        if (a.ppair) {
          // pairs per block (0: every active pair in one group)
          const int spg = ctx->bk_sg > 0 ? std::max(1, ctx->bk_sg / 2) : std::max(1, npair);
          const int ngrp = (npair + spg - 1) / spg;
          k_spmv_btp<WHICH><<<b.G * ngrp, kBtThreads, 0, ctx->stream>>>(b, bg, bk_rows, d_pairs,
                                                                       npair, spg);
          if (ctx->timing) tpairs.push_back({MT_KT_SPMV, e0, trec()});
          return b.G;
        }
      }
      const int ngrp = (a.nact + sg - 1) / sg;
      k_spmv_bt<T, V, WHICH><<<b.G * ngrp, kBtThreads, 0, ctx->stream>>>(b, bg, bk_rows, sg);
      G = b.G;
    } else if (ms > 0) {
      const int nb = ((a.nact + ms - 1) / ms) * a.G;
      if (ms == 4) k_spmv_ms<T, V, WHICH, 4><<<nb, BS, 0, ctx->stream>>>(a);
      else k_spmv_ms<T, V, WHICH, 8><<<nb, BS, 0, ctx->stream>>>(a);
      G = a.G;
    } else if (use_km) {
      k_spmv_km<T, V, WHICH><<<a.nact * km.Gpad, KM_BS, 0, ctx->stream>>>(a, km);
      G = km.Gpad;
    } else {
      k_spmv_dot<T, V, WHICH><<<nblk, BS, 0, ctx->stream>>>(a);
      G = a.G;
    }
    if (ctx->timing) tpairs.push_back({MT_KT_SPMV, e0, trec()});
    return G;
  }

  // the p-update folded into the first application's red forward sweep:
  // inner solves on bricks with the persistent red-black sweeps
  bool pfold() const {
    return brick && a.fused && a.pc == PCK_DILU_RB && ctx->rb_variant == 6 && ctx->tiles_rb.nlev == 2;
  }
  void launch_iteration() {
    const int nblk = a.nact * a.G;
    hipStream_t s = ctx->stream;
    // (paired x: p-hat to phi, the sweeps' w in ph)
    if (pfold()) {
      // p of the black tiles here, of the red ones in the red forward sweep
      const TilePlan& tp = ctx->tiles_rb;
      k_pupdate_tiles<T, V><<<nblk, BS, 0, s>>>(a, tp.tiles + tp.lvl_off[1], tp.lvl_off[2] - tp.lvl_off[1],
                                                tp.lat[0], tp.lat[1], tp.lat[2]);
      precond(a.p, a.ph, nullptr, a.ppair ? a.phi : nullptr, true);
    } else {
      k_pupdate<T, V><<<nblk, BS, 0, s>>>(a);
      precond(a.p, a.ph, nullptr, a.ppair ? a.phi : nullptr);
    }
    const int g1 = spmv<1>(nblk);
    k_finalize<FIN_ALPHA><<<a.nact, BS, 0, s>>>(a.sc, a.act, a.nact, g1, a.partials, a.rerr, a.rerr_len, a.tol);
    if (a.fused) {
      precond(a.r, a.ppair ? a.ph : a.sh, a.v, a.ppair ? a.shi : nullptr);  // sh = M^-1 (r - alpha v)
    } else {
      k_supdate<T, V><<<nblk, BS, 0, s>>>(a);
      precond(a.s, a.sh);
    }
    const int g2 = spmv<2>(nblk);
    k_finalize<FIN_OMEGA><<<a.nact, BS, 0, s>>>(a.sc, a.act, a.nact, g2, a.partials, a.rerr, a.rerr_len, a.tol);
    if (a.fused) k_rupdate_fused<T, V><<<nblk, BS, 0, s>>>(a);
    else k_rupdate<T, V><<<nblk, BS, 0, s>>>(a);
    k_finalize<FIN_RUPDATE><<<a.nact, BS, 0, s>>>(a.sc, a.act, a.nact, a.G, a.partials, a.rerr, a.rerr_len, a.tol);
    if (recheck) {
      k_recheck<T, V><<<nblk, BS, 0, s>>>(a);
      k_finalize<FIN_RECHECK><<<a.nact, BS, 0, s>>>(a.sc, a.act, a.nact, a.G, a.partials, a.rerr, a.rerr_len, a.tol);
    }
  }

  // event pool for one chunk: 2 SpMV + 2 DILU applications per iteration
  int ensure_events(int chunk) {
    const size_t need = size_t(chunk) * 8 + 4;  // 4 pairs per iteration
    while (ctx->tev.size() < need) {
      hipEvent_t e;
      MT_HIP_CHECK(ctx, hipEventCreate(&e));
      ctx->tev.push_back(e);
    }
    return MT_OK;
  }

  int run_chunk(int chunk) {
    if (ctx->timing) { int rc = ensure_events(chunk); if (rc) return rc; }
    // HIP cannot time events recorded by graph nodes: timing launches eagerly
    if (!use_graph || ctx->timing) {
      tpairs.clear();
      tnext = 0;
      for (int c = 0; c < chunk; ++c) launch_iteration();
      MT_HIP_CHECK(ctx, hipGetLastError());
      return MT_OK;
    }
    if (gexec && (g_nact != a.nact || g_chunk != chunk || g_npair != npair)) {
      hipGraphExecDestroy(gexec);
      gexec = nullptr;
    }
    if (!gexec) {
      hipGraph_t graph;
      tpairs.clear();
      tnext = 0;
      MT_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      for (int c = 0; c < chunk; ++c) launch_iteration();
      MT_HIP_CHECK(ctx, hipStreamEndCapture(ctx->stream, &graph));
      MT_HIP_CHECK(ctx, hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      hipGraphDestroy(graph);
      g_nact = a.nact;
      g_chunk = chunk;
      g_npair = npair;
    }
    MT_HIP_CHECK(ctx, hipGraphLaunch(gexec, ctx->stream));
    return MT_OK;
  }

  // Iteration baseline of every system (host copy of the scalars).
  void begin(const std::vector<SysScal>& h) {
    last_iter.resize(h.size());
    for (size_t q = 0; q < h.size(); ++q) last_iter[q] = h[q].iter;
  }
  // After a chunk and the host read of the scalars: add the event times of the
  // chunk and the algorithmic bytes of the iterations it ran.  Per iteration
  // a system takes part in 2 SpMV launches and 2 preconditioner applications.
  int account(const std::vector<SysScal>& h) {
    if (!ctx->timing) return MT_OK;
    long long units = 0;
    for (size_t q = 0; q < h.size() && q < last_iter.size(); ++q) {
      units += std::max(0, h[q].iter - last_iter[q]);
      last_iter[q] = h[q].iter;
    }
    const Grid& g = ctx->hop.g;
    const double sv = sizeof(V), sc2 = 2.0 * sizeof(T), n = double(g.nInt);
    const TilePlan& tp = a.pc == PCK_DILU_RB ? ctx->tiles_rb : ctx->tiles;
    const int pc_launches = a.pc == PCK_DILU_RB ? 3 : 2 * tp.nlev;
    int cnt[2] = {0, 0};
    for (const TPair& tq : tpairs) {
      float ms = 0.f;
      MT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->tev[tq.e0], ctx->tev[tq.e1]));
      mt_ctx::KTime& k = ctx->kt[tq.fam];
      k.ms += ms;
      k.launches += tq.fam == MT_KT_SPMV ? 1 : pc_launches;
      ++cnt[tq.fam];
    }
    // SpMV, per launch: stored couplings + frequency-free diagonal; per system:
    // x read, y written, the dot partner read.  Preconditioner application,
    // per application: the six lattice couplings; per system: the right-hand
    // side(s) read, the result written, the pivots read.
    if (cnt[MT_KT_SPMV])
      ctx->kt[MT_KT_SPMV].bytes += cnt[MT_KT_SPMV] * (double(stencil_nnz(ctx)) * sv + n * sv) +
                                   2.0 * units * 3.0 * n * sc2;
    // (the second application of an iteration reads r and v when fused, and
    // on bricks writes s = r - alpha v as well; with the p-update folded, the
    // first one's red forward sweep reads r and v and writes p for the red
    // tiles, half of the points)
    if (cnt[MT_KT_PRECOND])
      ctx->kt[MT_KT_PRECOND].bytes += cnt[MT_KT_PRECOND] * 6.0 * n * sv +
                                      units * (3.0 + (pfold() ? 1.5 : 0.0) +
                                               (a.fused ? (a.s_ok ? 5.0 : 4.0) : 3.0)) * n * sc2;
    return MT_OK;
  }
  ~Engine() {
    if (gexec) hipGraphExecDestroy(gexec);
  }
};

// Host view of the per-system status after a sync.
static int read_scal(mt_ctx* ctx, const SysScal* d, int n, std::vector<SysScal>& h) {
  h.resize(n);
  MT_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), d, n * sizeof(SysScal), hipMemcpyDeviceToHost, ctx->stream));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

// Unique frequencies -> per-system frequency index.
static void unique_freqs(const double* omegas, int nsys, std::vector<double>& freqs,
                         std::vector<int>& fidx) {
  freqs.clear();
  fidx.assign(nsys, 0);
  for (int s = 0; s < nsys; ++s) {
    int f = -1;
    for (size_t q = 0; q < freqs.size(); ++q)
      if (freqs[q] == omegas[s]) { f = int(q); break; }
    if (f < 0) { f = int(freqs.size()); freqs.push_back(omegas[s]); }
    fidx[s] = f;
  }
}

static int restart_of(const mt_solve_opts& o) { return o.restart > 0 ? o.restart : 120; }

// Compacts the active list to systems whose status is one of `keep`; returns
// the new count and rewrites the device list.
static int compact(mt_ctx* ctx, const std::vector<SysScal>& h, int* d_act, std::vector<int>& act,
                   bool (*keep)(int)) {
  std::vector<int> na;
  for (int s : act)
    if (keep(h[s].status)) na.push_back(s);
  const bool changed = na.size() != act.size();
  act.swap(na);
  // the source is the caller's list, which outlives the copy (every caller
  // synchronises the stream before it compacts again)
  if (changed && !act.empty())
    hipMemcpyAsync(d_act, act.data(), act.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream);
  return int(act.size());
}
static bool keep_running(int st) { return st == ST_RUN || st == ST_RECHECK; }

// FP64 BiCGStab on nsys padded systems (b, x device, system-major), the
// semantics of cuBiCG.  Fills final SysScal records and the rerr history.
static int solve_fp64(mt_ctx* ctx, int nsys, const double* omegas, Z* b, Z* x,
                      const mt_solve_opts& o, std::vector<SysScal>& fin,
                      std::vector<double>* rerr_out) {
  const Grid& g = ctx->hop.g;
  const int64_t nE = g.nE;
  const size_t nv = size_t(g.vlen(nsys));
  const bool tr = o.op_transpose != 0;
  if (tr) { int rc = ensure_transpose(ctx); if (rc) return rc; }
  std::vector<double> freqs;
  std::vector<int> fidx;
  unique_freqs(omegas, nsys, freqs, fidx);
  Engine<double, double> E;
  E.ctx = ctx;
  E.nsys = nsys;
  E.transpose = tr;
  E.use_graph = o.use_graph != 0;
  IterArgs<double, double>& a = E.a;
  a.A = view<double>(ctx, tr);
  a.pc = o.precond;
  a.nE = nE;
  int rc = 0;
  Z* const* dummy = nullptr;
  (void)dummy;
  if ((rc = ws(ctx, "f64_r", nv, &a.r)) || (rc = ws(ctx, "f64_rt", nv, &a.rt)) ||
      (rc = ws(ctx, "f64_p", nv, &a.p)) || (rc = ws(ctx, "f64_ph", nv, &a.ph)) ||
      (rc = ws(ctx, "f64_v", nv, &a.v)) || (rc = ws(ctx, "f64_s", nv, &a.s)) ||
      (rc = ws(ctx, "f64_sh", nv, &a.sh)) || (rc = ws(ctx, "f64_t", nv, &a.t)))
    return rc;
  Z* xbest = nullptr;
  if (o.best_iterate && (rc = ws(ctx, "f64_xbest", nv, &xbest))) return rc;
  a.xbest = xbest;
  a.x = x;
  a.b = b;
  C2<double>* dinv = nullptr;
  if ((rc = build_dinv<double>(ctx, o.precond, tr, freqs, &dinv))) return rc;
  a.dinv = dinv;
  if (o.precond == PCK_DILU_RB) {
    DiluArgs<double, double> fa;
    if ((rc = rb_faces(ctx, nsys, fa))) return rc;
    E.xf = fa.xf; E.xfsys = fa.xfsys; E.xfhalf = fa.xfhalf;
  }
  E.use_km = ctx->spmv_variant == 2;
  E.km = km_geom(ctx);
  const int Gmax = partials_G(ctx, E.km);
  SysScal* d_sc;
  int* d_act;
  if ((rc = ws(ctx, "f64_sc", nsys, &d_sc)) || (rc = ws(ctx, "act", nsys, &d_act)) ||
      (rc = ws(ctx, "partials", size_t(nsys) * Gmax * 2 + 64, &a.partials)))
    return rc;
  const int rlen = o.maxit + 1;
  double* d_rerr = nullptr;
  if ((rc = ws(ctx, "f64_rerr", size_t(nsys) * rlen, &d_rerr))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(d_rerr, 0, sizeof(double) * nsys * rlen, ctx->stream));
  a.rerr = d_rerr;
  a.rerr_len = rlen;
  std::vector<SysScal> h(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&h[s], 0, sizeof(SysScal));
    h[s].freq_omega = omegas[s];
    h[s].fidx = fidx[s];
    h[s].maxit = o.maxit;
    h[s].restart = restart_of(o);
    h[s].inner = 0;
  }
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, h.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  std::vector<int> act(nsys);
  for (int s = 0; s < nsys; ++s) act[s] = s;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, act.data(), nsys * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  a.sc = d_sc;
  a.act = d_act;
  a.nact = nsys;
  a.G = pick_G(ctx, nsys);
  a.tol = o.tol;
  k_init<double, double><<<a.nact * a.G, BS, 0, ctx->stream>>>(a);
  k_finalize<FIN_INIT><<<a.nact, BS, 0, ctx->stream>>>(a.sc, a.act, a.nact, a.G, a.partials, a.rerr,
                                                  a.rerr_len, a.tol);
  MT_HIP_CHECK(ctx, hipGetLastError());
  if ((rc = read_scal(ctx, d_sc, nsys, h))) return rc;
  E.begin(h);
  const int chunk = std::max(1, o.chunk);
  while (true) {
    const int n = compact(ctx, h, d_act, act, keep_running);
    if (n == 0) break;
    if (n != a.nact) {
      a.nact = n;
      a.G = pick_G(ctx, n);
    }
    if ((rc = E.run_chunk(chunk))) return rc;
    if ((rc = read_scal(ctx, d_sc, nsys, h))) return rc;
    if ((rc = E.account(h))) return rc;
  }
  // best iterate for systems that did not converge (solver.f90:1911-1919)
  for (int s = 0; s < nsys; ++s) {
    if (h[s].status == ST_CONV || !o.best_iterate) continue;
    if (!h[s].best_due)
      MT_HIP_CHECK(ctx, hipMemcpyAsync(x + size_t(s) * nE, xbest + size_t(s) * nE,
                                       nE * sizeof(Z), hipMemcpyDeviceToDevice, ctx->stream));
  }
  if (rerr_out) {
    rerr_out->resize(size_t(nsys) * rlen);
    MT_HIP_CHECK(ctx, hipMemcpyAsync(rerr_out->data(), d_rerr, sizeof(double) * nsys * rlen,
                                     hipMemcpyDeviceToHost, ctx->stream));
  }
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  fin = h;
  return MT_OK;
}

// Mixed precision: FP32 BiCGStab inner solves inside FP64 defect correction.
static int solve_mixed(mt_ctx* ctx, int nsys, const double* omegas, Z* b, Z* x,
                       const mt_solve_opts& o, std::vector<SysScal>& fin_out,
                       std::vector<SysScal>* fin_in, std::vector<double>* rerr_out) {
  const Grid& g = ctx->hop.g;
  const int64_t nE = g.nE;
  const size_t nv = size_t(g.vlen(nsys));
  const bool tr = o.op_transpose != 0;
  if (tr) { int rc = ensure_transpose(ctx); if (rc) return rc; }
  std::vector<double> freqs;
  std::vector<int> fidx;
  unique_freqs(omegas, nsys, freqs, fidx);
  using F = C2<float>;
  Engine<float, float> E;
  E.ctx = ctx;
  E.nsys = nsys;
  E.transpose = tr;
  E.use_graph = o.use_graph != 0;
  E.recheck = false;
  IterArgs<float, float>& a = E.a;
  a.A = view<float>(ctx, tr);
  a.pc = o.precond;
  int rc = 0;
  // inner solves on bricks: the red-black DILU with the 4-system SpMV
  const bool bricks = ctx->brick && o.precond == PCK_DILU_RB && ctx->spmv_variant == 3 &&
                      ctx->ms_S == 4;
  if (bricks && (rc = ensure_brick(ctx, tr))) return rc;
  const int64_t nI = bricks ? ctx->bk.g.nB : nE;   // inner vector stride
  const size_t nvi = bricks ? size_t(nI) * nsys : nv;
  a.nE = nI;
  E.brick = bricks;
  if (bricks) {
    E.bg = ctx->bk.g;
    E.bk_rows = ctx->bk.rows[tr ? 1 : 0];
    E.bk_slots = ctx->bk.slots[tr ? 1 : 0];
  }
  const char* nm[9] = {"f32_x", "f32_r", "f32_rt", "f32_p", "f32_ph", "f32_v", "f32_s", "f32_sh", "f32_t"};
  const char* nmb[9] = {"bk_x", "bk_r", "bk_rt", "bk_p", "bk_ph", "bk_v", "bk_s", "bk_sh", "bk_t"};
  const char* const* n9 = bricks ? nmb : nm;
  F* d32;
  if ((rc = ws(ctx, n9[0], nvi, &d32)) || (rc = ws(ctx, n9[1], nvi, &a.r)) ||
      (rc = ws(ctx, n9[2], nvi, &a.rt)) || (rc = ws(ctx, n9[3], nvi, &a.p)) ||
      (rc = ws(ctx, n9[4], nvi, &a.ph)) || (rc = ws(ctx, n9[5], nvi, &a.v)) ||
      (rc = ws(ctx, n9[6], nvi, &a.s)) || (rc = ws(ctx, n9[7], nvi, &a.sh)) ||
      (rc = ws(ctx, n9[8], nvi, &a.t)))
    return rc;
  // the outer correction's FP32 residual and correction in the all-edge
  // layout (permuted to and from the bricks around each inner solve)
  F *r_nat = a.r, *d_nat = d32;
  if (bricks && ((rc = ws(ctx, "f32_r", nv, &r_nat)) || (rc = ws(ctx, "f32_x", nv, &d_nat))))
    return rc;
  a.x = d32;
  a.xbest = nullptr;
  a.b = nullptr;
  a.fused = (o.precond == PCK_DILU || o.precond == PCK_DILU_RB) ? 1 : 0;
  a.s_ok = (bricks && a.fused) ? 1 : 0;   // the brick sweeps store s
  if (bricks && ctx->bk_pairs) {
    // p-hat and s-hat interleaved by system pair, read by the pair SpMV
    const size_t npv = size_t(nI) * 2 * ((nsys + 1) / 2);
    if ((rc = ws(ctx, "bk_phi", npv, &a.phi)) || (rc = ws(ctx, "bk_shi", npv, &a.shi)) ||
        (rc = ws(ctx, "bk_pairs", size_t((nsys + 1) / 2), &E.d_pairs)))
      return rc;
    a.ppair = 2 * nI;
  }
  C2<float>* dinv = nullptr;
  if ((rc = build_dinv<float>(ctx, o.precond, tr, freqs, &dinv))) return rc;
  if (bricks) {
    // the pivots of every frequency in brick order
    const int64_t nB = ctx->bk.g.nB;
    C2<float>* db;
    if ((rc = ws(ctx, "bk_dinv", size_t(nB) * freqs.size(), &db))) return rc;
    k_bk_gather<float><<<4096, BS, 0, ctx->stream>>>(ctx->bk.bmap, nB, int(freqs.size()), nE, nB,
                                                    dinv, db);
    MT_HIP_CHECK(ctx, hipGetLastError());
    dinv = db;
  }
  a.dinv = dinv;
  if (o.precond == PCK_DILU_RB) {
    DiluArgs<float, float> fa;
    if ((rc = rb_faces(ctx, nsys, fa))) return rc;
    E.xf = fa.xf; E.xfsys = fa.xfsys; E.xfhalf = fa.xfhalf;
  }
  E.use_km = ctx->spmv_variant == 1 || ctx->spmv_variant == 2;
  E.ms = ctx->spmv_variant == 3 ? ctx->ms_S : 0;
  E.km = km_geom(ctx);
  const int Gmax = partials_G(ctx, E.km);
  SysScal *d_so, *d_si;
  int* d_act;
  if ((rc = ws(ctx, "mx_so", nsys, &d_so)) || (rc = ws(ctx, "mx_si", nsys, &d_si)) ||
      (rc = ws(ctx, "act", nsys, &d_act)) ||
      (rc = ws(ctx, "partials", size_t(nsys) * Gmax * 2 + 64, &a.partials)))
    return rc;
  a.rerr = nullptr;
  a.tol = 0.0;
  a.rerr_len = 0;
  a.sc = d_si;
  a.act = d_act;
  OuterArgs oa;
  oa.A = view<double>(ctx, tr);
  oa.so = d_so;
  oa.si = d_si;
  oa.act = d_act;
  oa.nE = nE;
  oa.b = b;
  oa.x = x;
  oa.r32 = r_nat;
  oa.d32 = d_nat;
  oa.partials = a.partials;
  oa.tol = o.tol;
  oa.inner_tol = o.inner_tol > 0 ? o.inner_tol : 3e-4;
  oa.max_outer = o.max_outer > 0 ? o.max_outer : 30;
  oa.maxit_total = o.maxit;
  const int rlen = oa.max_outer + 2;
  double* d_rerr;
  if ((rc = ws(ctx, "mx_rerr", size_t(nsys) * rlen, &d_rerr))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(d_rerr, 0, sizeof(double) * nsys * rlen, ctx->stream));
  oa.rerr = d_rerr;
  oa.rerr_len = rlen;
  std::vector<SysScal> ho(nsys), hi(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&ho[s], 0, sizeof(SysScal));
    ho[s].freq_omega = omegas[s];
    ho[s].fidx = fidx[s];
    ho[s].maxit = o.maxit;
    ho[s].restart = restart_of(o);
    std::memset(&hi[s], 0, sizeof(SysScal));
    hi[s].status = ST_CONV;
    hi[s].restart = restart_of(o);
  }
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_so, ho.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_si, hi.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  std::vector<int> act(nsys);
  for (int s = 0; s < nsys; ++s) act[s] = s;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, act.data(), nsys * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  oa.nact = nsys;
  oa.G = pick_G(ctx, nsys);
  k_outer_bnorm<<<oa.nact * oa.G, BS, 0, ctx->stream>>>(oa);
  k_outer_fin<0><<<oa.nact, BS, 0, ctx->stream>>>(oa);
  MT_HIP_CHECK(ctx, hipGetLastError());
  if ((rc = read_scal(ctx, d_so, nsys, ho))) return rc;
  const int chunk = std::max(1, o.chunk);
  while (true) {
    // outer: FP64 residual of the systems still refining
    const int n = compact(ctx, ho, d_act, act, [](int st) { return st == ST_RUN; });
    if (n == 0) break;
    oa.nact = n;
    oa.G = pick_G(ctx, n);
    k_outer_resid<<<oa.nact * oa.G, BS, 0, ctx->stream>>>(oa);
    k_outer_fin<1><<<oa.nact, BS, 0, ctx->stream>>>(oa);
    MT_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = read_scal(ctx, d_so, nsys, ho))) return rc;
    std::vector<int> outer_act;
    for (int s : act)
      if (ho[s].status == ST_RUN) outer_act.push_back(s);
    if (outer_act.empty()) break;
    // inner: FP32 BiCGStab on the refining systems until each stops
    if ((rc = read_scal(ctx, d_si, nsys, hi))) return rc;
    E.begin(hi);
    std::vector<int> iact = outer_act;
    MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, iact.data(), iact.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    a.nact = int(iact.size());
    a.G = pick_G(ctx, a.nact);
    if (a.ppair) E.set_pairs(iact);
    if (bricks) {
      k_bk_to_brick<<<4096, BS, 0, ctx->stream>>>(d_act, a.nact, ctx->bk.bmap, nI, nE, r_nat, a.r, a.x);
      MT_HIP_CHECK(ctx, hipGetLastError());
    }
    while (true) {
      if ((rc = E.run_chunk(chunk))) return rc;
      if ((rc = read_scal(ctx, d_si, nsys, hi))) return rc;
      if ((rc = E.account(hi))) return rc;
      const int ni = compact(ctx, hi, d_act, iact, keep_running);
      if (ni == 0) break;
      if (ni != a.nact) {
        a.nact = ni;
        a.G = pick_G(ctx, ni);
        if (a.ppair) E.set_pairs(iact);
      }
    }
    // x += d / scale on the systems that refined
    MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, outer_act.data(), outer_act.size() * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    act = outer_act;
    oa.nact = int(outer_act.size());
    oa.G = pick_G(ctx, oa.nact);
    if (bricks) {
      k_bk_from_brick<<<4096, BS, 0, ctx->stream>>>(d_act, oa.nact, ctx->bk.bmap, nI, nE, a.x, d_nat);
      MT_HIP_CHECK(ctx, hipGetLastError());
    }
    k_outer_update<<<oa.nact * oa.G, BS, 0, ctx->stream>>>(oa);
    MT_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = read_scal(ctx, d_so, nsys, ho))) return rc;
  }
  if (rerr_out) {
    rerr_out->resize(size_t(nsys) * rlen);
    MT_HIP_CHECK(ctx, hipMemcpy(rerr_out->data(), d_rerr, sizeof(double) * nsys * rlen, hipMemcpyDeviceToHost));
  }
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if ((rc = read_scal(ctx, d_so, nsys, ho))) return rc;
  fin_out = ho;
  if (fin_in) *fin_in = hi;
  return MT_OK;
}

// =============================================================== C ABI
static int need_op(mt_ctx* ctx) {
  if (!ctx) return MT_E_ARG;
  if (!ctx->dop.ready) return mt_fail(ctx, MT_E_STATE, "operator not set (mt_operator_set_csr or mt_assemble)");
  return MT_OK;
}

static int status_code(int st) {
  switch (st) {
    case ST_CONV: return MT_OK;
    case ST_BREAK: return MT_E_BREAKDOWN;
    default: return MT_E_MAXIT;
  }
}

// Scatter host interior vectors (nsys x nInt complex) into padded device arrays.
static int upload_padded(mt_ctx* ctx, int nsys, const mt_zdouble* h, Z* d, const char* tmpname) {
  const Grid& g = ctx->hop.g;
  Z* tmp;
  int rc = ws(ctx, tmpname, size_t(nsys) * g.nInt, &tmp);
  if (rc) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(tmp, h, sizeof(Z) * nsys * g.nInt, hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(d, 0, sizeof(Z) * nsys * g.nE, ctx->stream));
  k_scatter<double, double><<<1024, 256, 0, ctx->stream>>>(ctx->dop.imap, g.nInt, g.nE, nsys, tmp, d);
  MT_HIP_CHECK(ctx, hipGetLastError());
  return MT_OK;
}
static int download_padded(mt_ctx* ctx, int nsys, const Z* d, mt_zdouble* h, const char* tmpname) {
  const Grid& g = ctx->hop.g;
  Z* tmp;
  int rc = ws(ctx, tmpname, size_t(nsys) * g.nInt, &tmp);
  if (rc) return rc;
  k_gather<double, double><<<1024, 256, 0, ctx->stream>>>(ctx->dop.imap, g.nInt, g.nE, nsys, d, tmp);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipMemcpyAsync(h, tmp, sizeof(Z) * nsys * g.nInt, hipMemcpyDeviceToHost, ctx->stream));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

extern "C" {

void mt_default_opts(mt_solve_opts* o) {
  o->maxit = 1200;
  o->tol = 1e-7;
  o->precond = MT_PC_DILU;
  o->precision = MT_PREC_FP64;
  o->restart = 120;
  o->best_iterate = 1;
  o->inner_tol = 3e-4;
  o->max_outer = 30;
  o->op_transpose = 0;
  o->chunk = 10;
  o->use_graph = 1;
}

int mt_device_count(int* count) {
  if (!count) return MT_E_ARG;
  if (hipGetDeviceCount(count) != hipSuccess) {
    (void)hipGetLastError();
    *count = 0;
    return MT_E_DEVICE;
  }
  return MT_OK;
}

int mt_ctx_create(int device, mt_ctx** out) {
  if (!out) return MT_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
    (void)hipGetLastError();
    return MT_E_DEVICE;
  }
  mt_ctx* ctx = new mt_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return MT_E_DEVICE;
  }
  for (auto& e : ctx->ev) hipEventCreate(&e);
  // the device address of this GPU's scratch array (the allocation-free
  // entry points use it; the array's host-side name is no device address)
  if (hipGetSymbolAddress(reinterpret_cast<void**>(&ctx->scratch_dev), HIP_SYMBOL(scratch)) !=
          hipSuccess ||
      !ctx->scratch_dev) {
    (void)hipGetLastError();
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return MT_E_DEVICE;
  }
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
      ncu > 0)
    ctx->num_cu = ncu;
  (void)hipGetLastError();
  *out = ctx;
  return MT_OK;
}

int mt_ctx_destroy(mt_ctx* ctx) {
  if (!ctx) return MT_OK;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  free_workspaces(ctx);
  free_dop(ctx->dop);
  hipFree(ctx->tiles.tiles);
  hipFree(ctx->tiles.shape);
  hipFree(ctx->tiles_rb.tiles);
  hipFree(ctx->tiles_rb.shape);
  hipFree(ctx->dc.coefN); hipFree(ctx->dc.wE); hipFree(ctx->dc.invL); hipFree(ctx->dc.dinv);
  hipFree(ctx->dc.tiles.tiles); hipFree(ctx->dc.tiles.shape);
  free_sites(ctx->sites);
  free_brick(ctx->bk);
  for (auto& e : ctx->ev) if (e) hipEventDestroy(e);
  for (auto& e : ctx->tev) hipEventDestroy(e);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return MT_OK;
}

const char* mt_last_error(const mt_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mt_set_grid(mt_ctx* ctx, int nx, int ny, int nz) {
  if (!ctx) return MT_E_ARG;
  if (nx < 2 || ny < 2 || nz < 2) return mt_fail(ctx, MT_E_ARG, "grid needs at least 2 cells per axis");
  ctx->hop = HostOperator();
  ctx->hop.g.set(nx, ny, nz);
  free_dop(ctx->dop);
  ctx->nnz_stencil = -1;
  return MT_OK;
}

int mt_operator_set_csr(mt_ctx* ctx, int n, int nnz, const int* rowptr, const int* col, int base,
                        const double* val, const double* vms1) {
  if (!ctx || !rowptr || !col || !val || !vms1 || (base != 0 && base != 1))
    return mt_fail(ctx, MT_E_ARG, "mt_operator_set_csr: bad arguments");
  hipSetDevice(ctx->device);
  Status st = csr_to_stencil(ctx->hop, n, nnz, rowptr, col, base, val, vms1);
  if (st.code) return mt_fail(ctx, st.code, st.msg);
  return upload_operator(ctx);
}

int mt_assemble(mt_ctx* ctx, int nx, int ny, int nzAir, int nzEarth, const double* dx,
                const double* dy, const double* dz, const double* logcond, double air_log) {
  if (!ctx || !dx || !dy || !dz || !logcond) return mt_fail(ctx, MT_E_ARG, "mt_assemble: null argument");
  hipSetDevice(ctx->device);
  ctx->hop = HostOperator();
  Status st = assemble(ctx->hop, nx, ny, nzAir, nzEarth, dx, dy, dz, logcond, air_log);
  if (st.code) return mt_fail(ctx, st.code, st.msg);
  return upload_operator(ctx);
}

int mt_host_assemble(int nx, int ny, int nzAir, int nzEarth, const double* dx, const double* dy,
                     const double* dz, const double* logcond, double air_log, int* n_int, int* nnz,
                     int* rowptr, int* col, double* val, double* vms1) {
  if (!dx || !dy || !dz || !logcond || !n_int || !nnz) return MT_E_ARG;
  HostOperator H;
  Status st = assemble(H, nx, ny, nzAir, nzEarth, dx, dy, dz, logcond, air_log);
  if (st.code) return st.code;
  std::vector<int> r, c;
  std::vector<double> v;
  stencil_to_csr(H, r, c, v);
  *n_int = int(H.g.nInt);
  *nnz = int(c.size());
  if (!rowptr) return MT_OK;
  std::memcpy(rowptr, r.data(), r.size() * sizeof(int));
  if (col) std::memcpy(col, c.data(), c.size() * sizeof(int));
  if (val) std::memcpy(val, v.data(), v.size() * sizeof(double));
  if (vms1)
    for (int64_t q = 0; q < H.g.nInt; ++q) vms1[q] = H.vms1[H.imap[q]];
  return MT_OK;
}

int mt_host_dilu_pivots(int nx, int ny, int nz, int n, int nnz, const int* rowptr, const int* col,
                        int base, const double* val, const double* vms1, double omega,
                        mt_zdouble* dinv) {
  if (!rowptr || !col || !val || !vms1 || !dinv) return MT_E_ARG;
  HostOperator H;
  H.g.set(nx, ny, nz);
  Status st = csr_to_stencil(H, n, nnz, rowptr, col, base, val, vms1);
  if (st.code) return st.code;
  std::vector<zc> dv;
  dilu_inv_pivots(H, omega, dv);
  for (int64_t q = 0; q < H.g.nInt; ++q) dinv[q] = dv[H.imap[q]];
  return MT_OK;
}

int mt_operator_size(const mt_ctx* ctx, int* n_int, int* n_edge, int* nnz) {
  if (!ctx || !ctx->dop.ready) return MT_E_STATE;
  const Grid& g = ctx->hop.g;
  if (n_int) *n_int = int(g.nInt);
  if (n_edge) *n_edge = int(g.nE);
  if (nnz) {
    if (ensure_host(const_cast<mt_ctx*>(ctx))) return MT_E_STATE;
    int64_t z = 0;
    for (int e : ctx->hop.imap)
      for (int s = 0; s < kSlots; ++s) z += ctx->hop.coef[size_t(s) * g.nE + e] != 0.0;
    *nnz = int(z);
  }
  return MT_OK;
}

int mt_operator_get_csr(mt_ctx* ctx, int* rowptr, int* col, double* val, double* vms1) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if ((rc = ensure_host(ctx))) return rc;
  std::vector<int> r, c;
  std::vector<double> v;
  stencil_to_csr(ctx->hop, r, c, v);
  if (rowptr) std::memcpy(rowptr, r.data(), r.size() * sizeof(int));
  if (col) std::memcpy(col, c.data(), c.size() * sizeof(int));
  if (val) std::memcpy(val, v.data(), v.size() * sizeof(double));
  if (vms1)
    for (int64_t q = 0; q < ctx->hop.g.nInt; ++q) vms1[q] = ctx->hop.vms1[ctx->hop.imap[q]];
  return MT_OK;
}

int mt_bicg(mt_ctx* ctx, double omega, const mt_zdouble* b, mt_zdouble* x, int maxit, double tol,
            int precond, int* niter, double* rerr, int* failed) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!b || !x || maxit < 1 || precond < MT_PC_DILU || precond > MT_PC_DILU_RB)
    return mt_fail(ctx, MT_E_ARG, "mt_bicg: bad arguments");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  mt_solve_opts o;
  mt_default_opts(&o);
  o.maxit = maxit;
  o.tol = tol;
  o.precond = precond;
  Z *db, *dx;
  if ((rc = ws(ctx, "io_b", g.vlen(1), &db)) || (rc = ws(ctx, "io_x", g.vlen(1), &dx))) return rc;
  if ((rc = upload_padded(ctx, 1, b, db, "io_tmp")) || (rc = upload_padded(ctx, 1, x, dx, "io_tmp"))) return rc;
  std::vector<SysScal> fin;
  std::vector<double> hist;
  if ((rc = solve_fp64(ctx, 1, &omega, db, dx, o, fin, rerr ? &hist : nullptr))) return rc;
  const SysScal& s = fin[0];
  if (s.iter < 0) return mt_fail(ctx, MT_E_NAN_RHS, "right-hand side contains NaN");
  if ((rc = download_padded(ctx, 1, dx, x, "io_tmp"))) return rc;
  // KSPiter%niter: iterations on convergence, maxit otherwise (solver.f90:1911-1919)
  if (niter) *niter = (s.status == ST_CONV) ? s.iter : maxit;
  if (failed) *failed = (s.status == ST_BREAK);
  if (rerr) for (int i = 0; i < maxit; ++i) rerr[i] = hist[i];
  // like cuBiCG, non-convergence is reported through niter/failed, not an error
  return MT_OK;
}

static int solve_batch_common(mt_ctx* ctx, int nsys, const double* omegas, Z* db, Z* dx,
                              const mt_solve_opts* opts, mt_sys_stats* stats) {
  mt_solve_opts o;
  if (opts) o = *opts; else mt_default_opts(&o);
  if (o.maxit < 1 || nsys < 1 || nsys > 64) return mt_fail(ctx, MT_E_ARG, "solve: 1 <= nsys <= 64, maxit >= 1");
  if (o.precond < MT_PC_DILU || o.precond > MT_PC_DILU_RB)
    return mt_fail(ctx, MT_E_ARG, "solve: unknown preconditioner");
  if (o.precision != MT_PREC_FP64 && o.precision != MT_PREC_MIXED)
    return mt_fail(ctx, MT_E_ARG, "solve: unknown precision");
  std::vector<SysScal> fin;
  int rc;
  if (o.precision == MT_PREC_MIXED) {
    std::vector<SysScal> fin_in;
    if ((rc = solve_mixed(ctx, nsys, omegas, db, dx, o, fin, &fin_in, nullptr))) return rc;
    if (stats)
      for (int s = 0; s < nsys; ++s) {
        stats[s].niter = fin[s].tot_inner;
        stats[s].nouter = fin[s].iter;
        stats[s].status = fin[s].iter < 0 ? MT_E_NAN_RHS : status_code(fin[s].status);
        stats[s].relres = fin[s].bnorm > 0 ? fin[s].rnorm / fin[s].bnorm : 0.0;
      }
  } else {
    if ((rc = solve_fp64(ctx, nsys, omegas, db, dx, o, fin, nullptr))) return rc;
    if (stats)
      for (int s = 0; s < nsys; ++s) {
        stats[s].niter = fin[s].iter;
        stats[s].nouter = 0;
        stats[s].status = fin[s].iter < 0 ? MT_E_NAN_RHS : status_code(fin[s].status);
        stats[s].relres = fin[s].bnorm > 0 ? fin[s].rnorm / fin[s].bnorm : 0.0;
      }
  }
  return MT_OK;
}

int mt_solve_batch(mt_ctx* ctx, int nsys, const double* omegas, const mt_zdouble* b, mt_zdouble* x,
                   const mt_solve_opts* opts, mt_sys_stats* stats) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!omegas || !b || !x || nsys < 1) return mt_fail(ctx, MT_E_ARG, "mt_solve_batch: bad arguments");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  Z *db, *dx;
  if ((rc = ws(ctx, "io_b", g.vlen(nsys), &db)) || (rc = ws(ctx, "io_x", g.vlen(nsys), &dx))) return rc;
  if ((rc = upload_padded(ctx, nsys, b, db, "io_tmp")) || (rc = upload_padded(ctx, nsys, x, dx, "io_tmp"))) return rc;
  if ((rc = solve_batch_common(ctx, nsys, omegas, db, dx, opts, stats))) return rc;
  return download_padded(ctx, nsys, dx, x, "io_tmp");
}

int mt_solve_batch_dev(mt_ctx* ctx, int nsys, const double* omegas, const void* b_dev, void* x_dev,
                       const mt_solve_opts* opts, mt_sys_stats* stats) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!omegas || !b_dev || !x_dev || nsys < 1) return mt_fail(ctx, MT_E_ARG, "mt_solve_batch_dev: bad arguments");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  Z *db, *dx;
  if ((rc = ws(ctx, "io_b", g.vlen(nsys), &db)) || (rc = ws(ctx, "io_x", g.vlen(nsys), &dx))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(db, 0, sizeof(Z) * nsys * g.nE, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(dx, 0, sizeof(Z) * nsys * g.nE, ctx->stream));
  k_scatter<double, double><<<1024, 256, 0, ctx->stream>>>(ctx->dop.imap, g.nInt, g.nE, nsys, (const Z*)b_dev, db);
  k_scatter<double, double><<<1024, 256, 0, ctx->stream>>>(ctx->dop.imap, g.nInt, g.nE, nsys, (const Z*)x_dev, dx);
  MT_HIP_CHECK(ctx, hipGetLastError());
  if ((rc = solve_batch_common(ctx, nsys, omegas, db, dx, opts, stats))) return rc;
  k_gather<double, double><<<1024, 256, 0, ctx->stream>>>(ctx->dop.imap, g.nInt, g.nE, nsys, dx, (Z*)x_dev);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_dev_alloc(mt_ctx* ctx, size_t bytes, void** ptr) {
  if (!ctx || !ptr) return MT_E_ARG;
  hipSetDevice(ctx->device);
  if (hipMalloc(ptr, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    return mt_fail(ctx, MT_E_OOM, "mt_dev_alloc: hipMalloc failed");
  }
  return MT_OK;
}
int mt_dev_free(mt_ctx* ctx, void* ptr) {
  if (!ctx) return MT_E_ARG;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  MT_HIP_CHECK(ctx, hipFree(ptr));
  return MT_OK;
}
int mt_dev_copy(mt_ctx* ctx, void* dst, const void* src, size_t bytes, int kind) {
  if (!ctx || kind < 0 || kind > 2) return MT_E_ARG;
  hipSetDevice(ctx->device);
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  MT_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, bytes, k, ctx->stream));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_apply(mt_ctx* ctx, double omega, int op, const mt_zdouble* x, mt_zdouble* y) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!x || !y) return mt_fail(ctx, MT_E_ARG, "mt_apply: null vector");
  hipSetDevice(ctx->device);
  if (op && (rc = ensure_transpose(ctx))) return rc;
  const Grid& g = ctx->hop.g;
  Z *dx, *dy;
  if ((rc = ws(ctx, "ap_x", g.vlen(1), &dx)) || (rc = ws(ctx, "ap_y", g.vlen(1), &dy))) return rc;
  if ((rc = upload_padded(ctx, 1, x, dx, "io_tmp"))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(dy, 0, sizeof(Z) * g.nE, ctx->stream));
  // op 1: A^T (same diagonal); op 2: A^H = AAii^T - i*ISIGN*w*vms1, the
  // reference's Mult_Aii(adjt=.true.) (modelOperator3D.f90:424-431)
  k_apply<double, double><<<1024, BS, 0, ctx->stream>>>(view<double>(ctx, op != 0), dx, dy,
                                                          op == 2 ? -omega : omega);
  MT_HIP_CHECK(ctx, hipGetLastError());
  return download_padded(ctx, 1, dy, y, "io_tmp");
}

int mt_precond_apply(mt_ctx* ctx, double omega, int precond, const mt_zdouble* x, mt_zdouble* y) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!x || !y) return mt_fail(ctx, MT_E_ARG, "mt_precond_apply: null vector");
  if (precond < MT_PC_DILU || precond > MT_PC_DILU_RB)
    return mt_fail(ctx, MT_E_ARG, "mt_precond_apply: unknown preconditioner");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  std::vector<double> freqs(1, omega);
  C2<double>* dinv;
  if ((rc = build_dinv<double>(ctx, precond, false, freqs, &dinv))) return rc;
  Z *dx, *dy;
  if ((rc = ws(ctx, "ap_x", g.vlen(1), &dx)) || (rc = ws(ctx, "ap_y", g.vlen(1), &dy))) return rc;
  if ((rc = upload_padded(ctx, 1, x, dx, "io_tmp"))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(dy, 0, sizeof(Z) * g.nE, ctx->stream));
  SysScal h;
  std::memset(&h, 0, sizeof(h));
  h.status = ST_RUN;
  SysScal* d_sc;
  int* d_act;
  if ((rc = ws(ctx, "pc_sc", 1, &d_sc)) || (rc = ws(ctx, "pc_act", 1, &d_act))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(d_act, 0, sizeof(int), ctx->stream));
  if (precond == MT_PC_DILU) {
    dilu_apply<double, double>(ctx, false, d_act, 1, d_sc, dx, dy, dinv, ST_RUN);
  } else if (precond == MT_PC_DILU_RB) {
    DiluArgs<double, double> d = dilu_args<double, double>(
        ctx->tiles_rb, ctx->dop.coef64, g.nE, g.nE, d_act, 1, d_sc, dx, dy, dinv, ST_RUN);
    if ((rc = rb_faces(ctx, 1, d))) return rc;
    dilu_rb_apply<double, double>(ctx, ctx->tiles_rb, d);
  } else if (precond == MT_PC_JACOBI) {
    k_diag_mul<<<2048, BS, 0, ctx->stream>>>(dinv, dx, dy, g.nE);
  } else {
    MT_HIP_CHECK(ctx, hipMemcpyAsync(dy, dx, sizeof(Z) * g.nE, hipMemcpyDeviceToDevice, ctx->stream));
  }
  MT_HIP_CHECK(ctx, hipGetLastError());
  return download_padded(ctx, 1, dy, y, "io_tmp");
}


// The inner solves' brick kernels on one vector: y = A x (what 0, k_spmv_bt)
// or y = M^-1 x (what 1, k_dilu_bk with the red-black DILU of A(omega)), of A
// or of A^T, in FP32 on bricks; x and y in EDGEi order, FP64.
int mt_inner_apply(mt_ctx* ctx, double omega, int what, int transpose, const mt_zdouble* x,
                   mt_zdouble* y) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!x || !y || (what != 0 && what != 1) || (transpose != 0 && transpose != 1))
    return mt_fail(ctx, MT_E_ARG, "mt_inner_apply: bad arguments");
  hipSetDevice(ctx->device);
  const bool tr = transpose != 0;
  if ((rc = ensure_brick(ctx, tr))) return rc;
  const Grid& g = ctx->hop.g;
  const int64_t nB = ctx->bk.g.nB, nE = g.nE;
  Z *dx, *dy;
  C2<float> *xn, *yn, *xb, *yb, *zb;
  if ((rc = ws(ctx, "ap_x", g.vlen(1), &dx)) || (rc = ws(ctx, "ap_y", g.vlen(1), &dy)) ||
      (rc = ws(ctx, "ia_xn", nE, &xn)) || (rc = ws(ctx, "ia_yn", nE, &yn)) ||
      (rc = ws(ctx, "ia_xb", nB, &xb)) || (rc = ws(ctx, "ia_yb", nB, &yb)) ||
      (rc = ws(ctx, "ia_zb", nB, &zb)))
    return rc;
  if ((rc = upload_padded(ctx, 1, x, dx, "io_tmp"))) return rc;
  SysScal h;
  std::memset(&h, 0, sizeof(h));
  h.status = ST_RUN;
  h.freq_omega = omega;
  SysScal* d_sc;
  int* d_act;
  if ((rc = ws(ctx, "pc_sc", 1, &d_sc)) || (rc = ws(ctx, "pc_act", 1, &d_act))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(d_act, 0, sizeof(int), ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(yb, 0, sizeof(C2<float>) * nB, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(yn, 0, sizeof(C2<float>) * nE, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemsetAsync(zb, 0, sizeof(C2<float>) * nB, ctx->stream));
  k_bk_narrow<<<2048, BS, 0, ctx->stream>>>(dx, xn, nE);
  k_bk_gather<float><<<2048, BS, 0, ctx->stream>>>(ctx->bk.bmap, nB, 1, nE, nB, xn, xb);
  if (what == 0) {
    IterArgs<float, float> a{};
    a.A = view<float>(ctx, tr);
    a.sc = d_sc;
    a.act = d_act;
    a.nact = 1;
    a.nE = nB;
    a.ph = xb;
    a.v = yb;
    a.rt = zb;
    Z* part;
    a.G = brick_chunks(ctx->bk.g);
    if ((rc = ws(ctx, "ia_part", size_t(a.G) * 2, &part))) return rc;
    a.partials = part;
    k_spmv_bt<float, float, 1><<<a.G, kBtThreads, 0, ctx->stream>>>(a, ctx->bk.g, ctx->bk.rows[tr ? 1 : 0], 1);
  } else {
    std::vector<double> freqs(1, omega);
    C2<float>* dinv;
    if ((rc = build_dinv<float>(ctx, PCK_DILU_RB, tr, freqs, &dinv))) return rc;
    C2<float>* db;
    if ((rc = ws(ctx, "ia_dinv", nB, &db))) return rc;
    k_bk_gather<float><<<2048, BS, 0, ctx->stream>>>(ctx->bk.bmap, nB, 1, nE, nB, dinv, db);
    DiluArgs<float, float> d = dilu_args<float, float>(ctx->tiles_rb, ctx->bk.slots[tr ? 1 : 0], nB, nB,
                                                       d_act, 1, d_sc, xb, yb, db, ST_RUN);
    if ((rc = rb_faces(ctx, 1, d))) return rc;
    dilu_bk_apply<float, float>(ctx, ctx->tiles_rb, d);
  }
  k_bk_from_brick<<<2048, BS, 0, ctx->stream>>>(d_act, 1, ctx->bk.bmap, nB, nE, yb, yn);
  MT_HIP_CHECK(ctx, hipMemsetAsync(dy, 0, sizeof(Z) * g.nE, ctx->stream));
  k_bk_widen<<<2048, BS, 0, ctx->stream>>>(yn, dy, nE);
  MT_HIP_CHECK(ctx, hipGetLastError());
  if (!ctx->err.empty() && ctx->err.find("face buffer") != std::string::npos)
    return mt_fail(ctx, MT_E_STATE, ctx->err);
  return download_padded(ctx, 1, dy, y, "io_tmp");
}

int mt_rhs_fwd(mt_ctx* ctx, const mt_zdouble* e_all, mt_zdouble* b) {
  if (!ctx || !e_all || !b) return mt_fail(ctx, MT_E_ARG, "mt_rhs_fwd: null argument");
  const HostOperator& H = ctx->hop;
  if (!H.have_model) return mt_fail(ctx, MT_E_STATE, "mt_rhs_fwd needs mt_assemble");
  for (int64_t r = 0; r < H.g.nInt; ++r) {
    mt_zdouble s = 0.0;
    for (int q = H.ccib_row[r]; q < H.ccib_row[r + 1]; ++q) s += H.ccib_val[q] * e_all[H.ccib_col[q]];
    b[r] = -s;
  }
  return MT_OK;
}

// Mult_Aib for a batch on the device: one thread per (interior row, system),
// b = -sum over the row's CCib entries in their stored order; x0 = e(EDGEi).
__global__ void k_rhs_fwd(int nsys, int64_t nInt, int64_t nE, const int* __restrict__ row,
                          const int* __restrict__ col, const double* __restrict__ val,
                          const int* __restrict__ imap, const Z* __restrict__ e, Z* __restrict__ b,
                          Z* __restrict__ x0) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nInt * nsys) return;
  const int s = int(t / nInt);
  const int64_t r = t - int64_t(s) * nInt;
  const Z* __restrict__ es = e + int64_t(s) * nE;
  Z acc = {0.0, 0.0};
  for (int q = row[r]; q < row[r + 1]; ++q) {
    const Z v = es[col[q]];
    acc.x += val[q] * v.x;
    acc.y += val[q] * v.y;
  }
  b[t] = {-acc.x, -acc.y};
  if (x0) x0[t] = es[imap[r]];
}

int mt_rhs_fwd_dev(mt_ctx* ctx, int nsys, const void* e_dev, void* b_dev, void* x0_dev) {
  if (!ctx || !e_dev || !b_dev || nsys < 1) return mt_fail(ctx, MT_E_ARG, "mt_rhs_fwd_dev: bad arguments");
  const HostOperator& H = ctx->hop;
  if (!H.have_model || !ctx->dop.ready) return mt_fail(ctx, MT_E_STATE, "mt_rhs_fwd_dev needs mt_assemble");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  int *drow, *dcol;
  double* dval;
  int rc;
  if ((rc = ws(ctx, "rhs_row", H.ccib_row.size(), &drow)) ||
      (rc = ws(ctx, "rhs_col", std::max<size_t>(1, H.ccib_col.size()), &dcol)) ||
      (rc = ws(ctx, "rhs_val", std::max<size_t>(1, H.ccib_val.size()), &dval)))
    return rc;
  if (!ctx->rhs_ready) {
    MT_HIP_CHECK(ctx, hipMemcpy(drow, H.ccib_row.data(), H.ccib_row.size() * sizeof(int), hipMemcpyHostToDevice));
    if (!H.ccib_col.empty()) {
      MT_HIP_CHECK(ctx, hipMemcpy(dcol, H.ccib_col.data(), H.ccib_col.size() * sizeof(int), hipMemcpyHostToDevice));
      MT_HIP_CHECK(ctx, hipMemcpy(dval, H.ccib_val.data(), H.ccib_val.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    ctx->rhs_ready = true;
  }
  const int64_t nt = H.g.nInt * nsys;
  k_rhs_fwd<<<unsigned((nt + 255) / 256), 256, 0, ctx->stream>>>(
      nsys, H.g.nInt, H.g.nE, drow, dcol, dval, ctx->dop.imap, (const Z*)e_dev, (Z*)b_dev, (Z*)x0_dev);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_bc_x0_dev(mt_ctx* ctx, int nsys, const double* omegas, const int* modes, void* e_dev) {
  if (!ctx || !omegas || !modes || !e_dev || nsys < 1)
    return mt_fail(ctx, MT_E_ARG, "mt_bc_x0_dev: bad arguments");
  const HostOperator& H = ctx->hop;
  if (!H.have_model) return mt_fail(ctx, MT_E_STATE, "mt_bc_x0_dev needs mt_assemble");
  const Grid& g = H.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nza = H.nzAir;
  if (nx < 2 || ny < 2 || nz < 3 || H.nzEarth < 1)
    return mt_fail(ctx, MT_E_ARG, "mt_bc_x0_dev: the grid needs nx, ny >= 2 and nz >= 3");
  for (int s = 0; s < nsys; ++s)
    if ((modes[s] != 1 && modes[s] != 2) || !(omegas[s] > 0.0))
      return mt_fail(ctx, MT_E_ARG, "mt_bc_x0_dev: mode must be 1 or 2 and omega > 0");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const size_t lds = bc_lds_bytes(nz);
  int maxlds = 0;
  MT_HIP_CHECK(ctx, hipDeviceGetAttribute(&maxlds, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device));
  if (lds + 64 > size_t(maxlds))
    return mt_fail(ctx, MT_E_ARG, "mt_bc_x0_dev: nz too large for the slice window in LDS");
  // host side: earth conductivities as ModelParamToCell forms them, block distances, items
  const int64_t nce = int64_t(nx) * ny * H.nzEarth;
  std::vector<double> sig(nce);
  for (int64_t q = 0; q < nce; ++q) sig[q] = std::exp(H.logcond[q]);
  auto cdist = [](const std::vector<double>& d) {
    const int n = int(d.size());
    std::vector<double> c(n + 1);
    for (int i = 1; i < n; ++i) c[i] = d[i] + d[i - 1];
    c[0] = d[0];
    c[n] = d[n - 1];
    return c;
  };
  const std::vector<double> cx = cdist(H.dx), cy = cdist(H.dy), cz = cdist(H.dz);
  std::vector<int2> items;
  for (int s = 0; s < nsys; ++s) {
    const int nsl = modes[s] == 2 ? nx : ny;
    for (int k = 1; k <= nsl; ++k) items.push_back(make_int2(s, k));
  }
  const int nitem = int(items.size());
  const int nhmax = std::max(nx, ny), b = nz - 1;
  const int64_t nmax = int64_t(nhmax - 1) * b;
  const int per_cu = 2 * (lds + 64) <= size_t(160 * 1024) ? 2 : 1;   // workgroups per CU by LDS (160 KB)
  const int G = std::max(1, std::min(nitem, ctx->num_cu * per_cu));
  BcArgs a;
  a.nx = nx; a.ny = ny; a.nz = nz; a.nza = nza;
  a.nXe = g.nXe; a.nE = g.nE;
  a.nitem = nitem;
  a.ncolMax = nhmax + 1;
  a.mmbMax = 2 * nhmax + 2 * nz + 1;
  a.uStride = nmax * nz;
  a.yStride = nmax;
  double *ddx, *ddy, *ddz, *dcx, *dcy, *dcz, *dsig, *dom;
  int *dmode, *dflag;
  int2* ditem;
  double2 *dexb, *dside, *du, *dy;
  int rc;
  if ((rc = ws(ctx, "bc_dx", nx, &ddx)) || (rc = ws(ctx, "bc_dy", ny, &ddy)) ||
      (rc = ws(ctx, "bc_dz", nz, &ddz)) || (rc = ws(ctx, "bc_cx", nx + 1, &dcx)) ||
      (rc = ws(ctx, "bc_cy", ny + 1, &dcy)) || (rc = ws(ctx, "bc_cz", nz + 1, &dcz)) ||
      (rc = ws(ctx, "bc_sig", nce, &dsig)) || (rc = ws(ctx, "bc_om", nsys, &dom)) ||
      (rc = ws(ctx, "bc_mode", nsys, &dmode)) || (rc = ws(ctx, "bc_item", nitem, &ditem)) ||
      (rc = ws(ctx, "bc_flag", nitem, &dflag)) ||
      (rc = ws(ctx, "bc_exb", size_t(nitem) * a.mmbMax, &dexb)) ||
      (rc = ws(ctx, "bc_side", size_t(nitem) * 4 * nz, &dside)) ||
      (rc = ws(ctx, "bc_u", size_t(G) * a.uStride, &du)) ||
      (rc = ws(ctx, "bc_y", size_t(G) * a.yStride, &dy)))
    return rc;
  MT_HIP_CHECK(ctx, hipMemcpy(ddx, H.dx.data(), nx * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(ddy, H.dy.data(), ny * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(ddz, H.dz.data(), nz * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dcx, cx.data(), (nx + 1) * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dcy, cy.data(), (ny + 1) * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dcz, cz.data(), (nz + 1) * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dsig, sig.data(), nce * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dom, omegas, nsys * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dmode, modes, nsys * sizeof(int), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(ditem, items.data(), nitem * sizeof(int2), hipMemcpyHostToDevice));
  a.dx = ddx; a.dy = ddy; a.dz = ddz; a.cx = dcx; a.cy = dcy; a.cz = dcz;
  a.sig = dsig; a.omega = dom; a.mode = dmode; a.item = ditem;
  a.exb = dexb; a.side = dside; a.ucol = du; a.yx = dy; a.flag = dflag;
  a.e = static_cast<double2*>(e_dev);
  MT_HIP_CHECK(ctx, hipMemsetAsync(e_dev, 0, size_t(nsys) * g.nE * sizeof(double2), ctx->stream));
  const int64_t n1 = int64_t(nitem) * a.ncolMax;
  k_bc_te1d<<<unsigned((n1 + 255) / 256), 256, 0, ctx->stream>>>(a);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipFuncSetAttribute(reinterpret_cast<const void*>(&k_bc_te2d),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  k_bc_te2d<<<unsigned(G), 256, lds, ctx->stream>>>(a);
  MT_HIP_CHECK(ctx, hipGetLastError());
  std::vector<int> flags(nitem);
  MT_HIP_CHECK(ctx, hipMemcpyAsync(flags.data(), dflag, nitem * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < nitem; ++k)
    if (flags[k])
      return mt_fail(ctx, MT_E_PIVOT, "mt_bc_x0_dev: a 2-D slice needs row interchanges (system " +
                                          std::to_string(items[k].x) + ", slice " +
                                          std::to_string(items[k].y) + "); fields written without them");
  return MT_OK;
}

// Host-buffer forms of the sensitivity entry points: the arrays are staged
// through device workspaces and the device kernels do the work.
static int stage_up(mt_ctx* ctx, const char* name, const void* h, size_t bytes, void** d) {
  int rc = ws_get(ctx, name, bytes, d);
  if (rc) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(*d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_rhs_trn(mt_ctx* ctx, double omega, const mt_zdouble* s_all, mt_zdouble* b) {
  if (!ctx || !s_all || !b) return mt_fail(ctx, MT_E_ARG, "mt_rhs_trn: null argument");
  if (!ctx->hop.have_model) return mt_fail(ctx, MT_E_STATE, "mt_rhs_trn needs mt_assemble");
  const Grid& g = ctx->hop.g;
  void *ds, *db;
  int rc;
  if ((rc = stage_up(ctx, "io_s", s_all, g.nE * sizeof(Z), &ds)) ||
      (rc = ws_get(ctx, "io_b", g.nInt * sizeof(Z), &db)))
    return rc;
  if ((rc = mt_rhs_trn_dev(ctx, 1, &omega, ds, db, 1))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpy(b, db, g.nInt * sizeof(Z), hipMemcpyDeviceToHost));
  return MT_OK;
}

int mt_post_trn(mt_ctx* ctx, const mt_zdouble* ei, mt_zdouble* e_all) {
  if (!ctx || !ei || !e_all) return mt_fail(ctx, MT_E_ARG, "mt_post_trn: null argument");
  if (!ctx->hop.have_model) return mt_fail(ctx, MT_E_STATE, "mt_post_trn needs mt_assemble");
  const Grid& g = ctx->hop.g;
  void *di, *de;
  int rc;
  if ((rc = stage_up(ctx, "io_b", ei, g.nInt * sizeof(Z), &di)) ||
      (rc = ws_get(ctx, "io_s", g.nE * sizeof(Z), &de)))
    return rc;
  if ((rc = mt_post_trn_dev(ctx, 1, di, de))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpy(e_all, de, g.nE * sizeof(Z), hipMemcpyDeviceToHost));
  return MT_OK;
}

int mt_pmultT(mt_ctx* ctx, double omega, int npol, const mt_zdouble* e0, const mt_zdouble* lam,
              double* dm, int accumulate) {
  if (!ctx || !e0 || !lam || !dm || npol < 1) return mt_fail(ctx, MT_E_ARG, "mt_pmultT: bad arguments");
  if (!ctx->hop.have_model) return mt_fail(ctx, MT_E_STATE, "mt_pmultT needs mt_assemble");
  const HostOperator& H = ctx->hop;
  const size_t nv = size_t(npol) * H.g.nE * sizeof(Z);
  const size_t nc = size_t(H.g.nx) * H.g.ny * H.nzEarth * sizeof(double);
  void *de, *dl, *dd;
  int rc;
  if ((rc = stage_up(ctx, "io_e", e0, nv, &de)) || (rc = stage_up(ctx, "io_l", lam, nv, &dl)) ||
      (rc = stage_up(ctx, "io_dm", dm, nc, &dd)))
    return rc;
  if ((rc = mt_pmultT_dev(ctx, 1, &omega, npol, de, dl, dd, accumulate))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpy(dm, dd, nc, hipMemcpyDeviceToHost));
  return MT_OK;
}

static int ensure_divcor(mt_ctx* ctx) {
  DivcorData& dc = ctx->dc;
  if (dc.ready) return MT_OK;
  { int rc = ensure_host(ctx); if (rc) return rc; }
  const HostOperator& H = ctx->hop;
  const Grid& g = H.g;
  std::vector<double> coefN, wE, dinv;
  divcor_node_stencil(H, coefN, wE);
  divcor_inv_pivots(H, coefN, dinv);
  dc.nN = int64_t(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  std::vector<double> invL(g.nE);
  for (int64_t e = 0; e < g.nE; ++e) invL[e] = 1.0 / H.edgeL[e];
  std::vector<Z> dz(dc.nN);
  for (int64_t n = 0; n < dc.nN; ++n) dz[n] = {dinv[n], 0.0};
  MT_HIP_CHECK(ctx, hipMalloc(&dc.coefN, coefN.size() * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&dc.wE, g.nE * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&dc.invL, g.nE * sizeof(double)));
  MT_HIP_CHECK(ctx, hipMalloc(&dc.dinv, dc.nN * sizeof(Z)));
  MT_HIP_CHECK(ctx, hipMemcpy(dc.coefN, coefN.data(), coefN.size() * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dc.wE, wE.data(), g.nE * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dc.invL, invL.data(), g.nE * sizeof(double), hipMemcpyHostToDevice));
  MT_HIP_CHECK(ctx, hipMemcpy(dc.dinv, dz.data(), dc.nN * sizeof(Z), hipMemcpyHostToDevice));
  LatticeInfo lat = {g.nx - 1, g.ny - 1, g.nz - 1, 1, 1, 1, g.nx + 1, g.ny + 1, 0};
  int rc = build_tile_plan(ctx, dc.tiles, 1, &lat);
  if (rc) return rc;
  dc.ready = true;
  return MT_OK;
}

// SdivCorr (EMsolve3D.f90:507-613): phi = VDsG^-1 (VDs ei - phi0) by PCG with
// the node DILU, then ei -= Gai phi.  VDs does not depend on omega (the
// reference's VomegaMuSig/(mu0 omega)/V_E is sigma_E).
int mt_gradient(mt_ctx* ctx, double omega, int npol, const mt_zdouble* e_all,
                const mt_zdouble* lam, double* dm, int mode) {
  if (!ctx || !e_all || !lam || !dm || npol < 1 || mode < 0 || mode > 1)
    return mt_fail(ctx, MT_E_ARG, "mt_gradient: bad arguments");
  if (!ctx->hop.have_model) return mt_fail(ctx, MT_E_STATE, "mt_gradient needs mt_assemble");
  const HostOperator& H = ctx->hop;
  const size_t nc = size_t(H.g.nx) * H.g.ny * H.nzEarth * sizeof(double);
  void *de, *dl, *dd;
  int rc;
  if ((rc = stage_up(ctx, "io_e", e_all, size_t(npol) * H.g.nE * sizeof(Z), &de)) ||
      (rc = stage_up(ctx, "io_l", lam, size_t(npol) * H.g.nInt * sizeof(Z), &dl)) ||
      (rc = stage_up(ctx, "io_dm", dm, nc, &dd)))
    return rc;
  if ((rc = mt_gradient_dev(ctx, 1, &omega, npol, de, dl, dd, mode, 1))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpy(dm, dd, nc, hipMemcpyDeviceToHost));
  return MT_OK;
}

int mt_host_gradient(int nx, int ny, int nzAir, int nzEarth, const double* dx, const double* dy,
                     const double* dz, const double* logcond, double air_log, double omega,
                     int npol, const mt_zdouble* e_all, const mt_zdouble* lam, double* dm,
                     int mode) {
  if (!dx || !dy || !dz || !logcond || !e_all || !lam || !dm || npol < 1 || mode < 0 || mode > 1)
    return MT_E_ARG;
  HostOperator H;
  Status st = assemble(H, nx, ny, nzAir, nzEarth, dx, dy, dz, logcond, air_log);
  if (st.code) return st.code;
  grad_assemble(H, omega, npol, reinterpret_cast<const zc*>(e_all),
                reinterpret_cast<const zc*>(lam), mode, dm);
  return MT_OK;
}

int mt_divcor(mt_ctx* ctx, double omega, mt_zdouble* ei, const mt_zdouble* phi0, int maxit,
              double tol, int* niter, double* relres) {
  (void)omega;
  if (!ctx || !ei || maxit < 1) return mt_fail(ctx, MT_E_ARG, "mt_divcor: bad arguments");
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!ctx->hop.have_model) return mt_fail(ctx, MT_E_STATE, "mt_divcor needs mt_assemble");
  hipSetDevice(ctx->device);
  if ((rc = ensure_divcor(ctx))) return rc;
  const Grid& g = ctx->hop.g;
  DivcorData& dc = ctx->dc;
  NodeArgs a;
  std::memset(&a, 0, sizeof(a));
  a.coefN = dc.coefN; a.wE = dc.wE; a.invL = dc.invL;
  a.nN = dc.nN; a.nE = g.nE; a.nXe = g.nXe; a.nYe = g.nYe; a.nXi = g.nXi; a.nYi = g.nYi;
  a.nInt = g.nInt; a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
  a.G = 512;
  a.tol = tol;
  const size_t nNp = size_t(dc.nN + g.tail());
  if ((rc = ws(ctx, "dc_e", g.vlen(1), &a.e)) || (rc = ws(ctx, "dc_b", nNp, &a.b)) ||
      (rc = ws(ctx, "dc_x", nNp, &a.x)) || (rc = ws(ctx, "dc_r", nNp, &a.r)) ||
      (rc = ws(ctx, "dc_s", nNp, &a.s)) || (rc = ws(ctx, "dc_p", nNp, &a.p)) ||
      (rc = ws(ctx, "dc_q", nNp, &a.q)) || (rc = ws(ctx, "dc_sc", 1, &a.sc)) ||
      (rc = ws(ctx, "partials", size_t(pick_G(ctx, 1)) * 2 + 1024, &a.partials)) ||
      (rc = ws(ctx, "dc_rerr", size_t(maxit) + 1, &a.rerr)))
    return rc;
  int* d_act;
  if ((rc = ws(ctx, "dc_act", 1, &d_act))) return rc;
  MT_HIP_CHECK(ctx, hipMemsetAsync(d_act, 0, sizeof(int), ctx->stream));
  if ((rc = upload_padded(ctx, 1, ei, a.e, "io_tmp"))) return rc;
  Z* phi0d = nullptr;
  if (phi0) {
    // phi0 on interior nodes (NODEi order) -> padded node vector
    if ((rc = ws(ctx, "dc_phi0", nNp, &phi0d))) return rc;
    std::vector<Z> h(dc.nN, Z{0.0, 0.0});
    int64_t q = 0;
    for (int k = 1; k < g.nz; ++k)
      for (int j = 1; j < g.ny; ++j)
        for (int i = 1; i < g.nx; ++i, ++q)
          h[i + int64_t(g.nx + 1) * (j + int64_t(g.ny + 1) * k)] = {phi0[q].real(), phi0[q].imag()};
    MT_HIP_CHECK(ctx, hipMemcpy(phi0d, h.data(), dc.nN * sizeof(Z), hipMemcpyHostToDevice));
  }
  a.phi0 = phi0d;
  SysScal h;
  std::memset(&h, 0, sizeof(h));
  h.maxit = maxit;
  h.status = ST_RUN;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(a.sc, &h, sizeof(h), hipMemcpyHostToDevice, ctx->stream));
  const int nb = a.G;
  k_divc<<<nb, BS, 0, ctx->stream>>>(a, a.b);
  k_pcg_init<<<nb, BS, 0, ctx->stream>>>(a);
  k_pcg_fin<PF_INIT><<<1, BS, 0, ctx->stream>>>(a);
  MT_HIP_CHECK(ctx, hipGetLastError());
  std::vector<SysScal> hs(1);
  if ((rc = read_scal(ctx, a.sc, 1, hs))) return rc;
  if (hs[0].bnorm == 0.0) {  // divergence already zero: no correction (SdivCorr)
    if (niter) *niter = 0;
    if (relres) *relres = 0.0;
    return MT_OK;
  }
  while (hs[0].status == ST_RUN) {
    for (int c = 0; c < 10; ++c) {
      dilu_sweeps<double, double>(ctx, dc.tiles, dc.coefN, dc.nN, dc.nN, d_act, 1, a.sc, a.r, a.s,
                                  dc.dinv, ST_RUN);
      k_pcg_rs<<<nb, BS, 0, ctx->stream>>>(a);
      k_pcg_fin<PF_DELTA><<<1, BS, 0, ctx->stream>>>(a);
      k_pcg_pq<<<nb, BS, 0, ctx->stream>>>(a);
      k_pcg_fin<PF_ALPHA><<<1, BS, 0, ctx->stream>>>(a);
      k_pcg_xr<<<nb, BS, 0, ctx->stream>>>(a);
      k_pcg_fin<PF_R><<<1, BS, 0, ctx->stream>>>(a);
    }
    MT_HIP_CHECK(ctx, hipGetLastError());
    if ((rc = read_scal(ctx, a.sc, 1, hs))) return rc;
  }
  k_grad_sub<<<1024, BS, 0, ctx->stream>>>(a, view<double>(ctx, false));
  MT_HIP_CHECK(ctx, hipGetLastError());
  if (niter) *niter = hs[0].iter;
  if (relres) *relres = hs[0].rnorm / hs[0].bnorm;
  return download_padded(ctx, 1, a.e, ei, "io_tmp");
}

int mt_bench_spmv(mt_ctx* ctx, int nsys, int precision, int reps, double* ms_per_launch,
                  double* bytes_per_launch) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (nsys < 1 || nsys > 64 || reps < 1) return mt_fail(ctx, MT_E_ARG, "mt_bench_spmv: bad arguments");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  const size_t nv = size_t(g.vlen(nsys));
  std::vector<SysScal> h(nsys);
  std::vector<int> act(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&h[s], 0, sizeof(SysScal));
    h[s].status = ST_RUN;
    h[s].freq_omega = 2.0 * kPi * (1.0 + s);
    h[s].rho = {1.0, 0.0};
    act[s] = s;
  }
  SysScal* d_sc;
  int* d_act;
  Z* partials;
  const KmGeom km = km_geom(ctx);
  const bool use_km = ctx->spmv_variant == 2 || (ctx->spmv_variant == 1 && precision == MT_PREC_MIXED);
  const int msS = (ctx->spmv_variant == 3 && precision == MT_PREC_MIXED) ? ctx->ms_S : 0;
  // what the mixed solver's inner solves run: the brick SpMV
  const bool bricks = ctx->brick && msS == 4;
  if (bricks && (rc = ensure_brick(ctx, false))) return rc;
  if ((rc = ws(ctx, "bn_sc", nsys, &d_sc)) || (rc = ws(ctx, "bn_act", nsys, &d_act)) ||
      (rc = ws(ctx, "partials", size_t(nsys) * partials_G(ctx, km) * 2 + 64,
               &partials)))
    return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, h.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, act.data(), nsys * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  const int G = pick_G(ctx, nsys);
  float ms = 0.f;
  double bytes = 0.0;
  // algorithmic bytes: stored coefficients (nnz values) + diagonal + per system
  // x read, y written, rt read for the fused dot
  int nnz = 0;
  mt_operator_size(ctx, nullptr, nullptr, &nnz);
  auto run = [&](auto tagT) -> int {
    using T = decltype(tagT);
    IterArgs<T, T> a;
    std::memset(&a, 0, sizeof(a));
    a.A = view<T>(ctx, false);
    a.sc = d_sc; a.act = d_act; a.nact = nsys; a.G = G; a.pc = PCK_NONE; a.nE = g.nE;
    a.partials = partials;
    C2<T>* vin; C2<T>* vout; C2<T>* vrt;
    int r2;
    // (bricks: an even number of systems' room, the pair kernel's x)
    const size_t nvb = bricks ? size_t(ctx->bk.g.nB) * (nsys + (nsys & 1)) : nv;
    if ((r2 = ws(ctx, sizeof(T) == 8 ? "bn_in64" : (bricks ? "bn_inb" : "bn_in32"), nvb, &vin)) ||
        (r2 = ws(ctx, sizeof(T) == 8 ? "bn_out64" : (bricks ? "bn_outb" : "bn_out32"), nvb, &vout)) ||
        (r2 = ws(ctx, sizeof(T) == 8 ? "bn_rt64" : (bricks ? "bn_rtb" : "bn_rt32"), nvb, &vrt)))
      return r2;
    a.ph = vin; a.v = vout; a.rt = vrt;
    int4* d_pairs = nullptr;
    const int npair = (nsys + 1) / 2;
    const bool paired = bricks && ctx->bk_pairs && sizeof(T) == 4;
    if (bricks) {
      a.nE = ctx->bk.g.nB;
      a.G = brick_chunks(ctx->bk.g);
    }
    if (paired) {
      // x of every pair interleaved in vin, as the inner solves keep p-hat
      std::vector<int4> hp(npair);
      for (int m = 0; m < npair; ++m) hp[m] = {2 * m, 2 * m + 1, 2 * m, 2 * m + 1 < nsys ? 2 * m + 1 : -1};
      if ((r2 = ws(ctx, "bn_pairs", size_t(npair), &d_pairs))) return r2;
      MT_HIP_CHECK(ctx, hipMemcpy(d_pairs, hp.data(), npair * sizeof(int4), hipMemcpyHostToDevice));
      a.phi = vin;
      a.ppair = 2 * ctx->bk.g.nB;
    }
    auto launch = [&]() {
      if (paired) {
        if constexpr (sizeof(T) == 4) {
          const int sg = ctx->bk_sg > 0 ? ctx->bk_sg : nsys;
          const int spg = std::max(1, sg / 2);
          k_spmv_btp<1><<<a.G * ((npair + spg - 1) / spg), kBtThreads, 0, ctx->stream>>>(
              a, ctx->bk.g, reinterpret_cast<const float*>(ctx->bk.rows[0]), d_pairs, npair, spg);
        }
      } else if (bricks) {
        const int sg = ctx->bk_sg > 0 ? ctx->bk_sg : nsys;
        k_spmv_bt<T, T, 1><<<a.G * ((nsys + sg - 1) / sg), kBtThreads, 0, ctx->stream>>>(
            a, ctx->bk.g, reinterpret_cast<const T*>(ctx->bk.rows[0]), sg);
      }
      else if (msS == 4) k_spmv_ms<T, T, 1, 4><<<((nsys + 3) / 4) * G, BS, 0, ctx->stream>>>(a);
      else if (msS == 8) k_spmv_ms<T, T, 1, 8><<<((nsys + 7) / 8) * G, BS, 0, ctx->stream>>>(a);
      else if (use_km) k_spmv_km<T, T, 1><<<nsys * km.Gpad, KM_BS, 0, ctx->stream>>>(a, km);
      else k_spmv_dot<T, T, 1><<<nsys * G, BS, 0, ctx->stream>>>(a);
    };
    launch();  // warm-up
    MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int q = 0; q < reps; ++q) launch();
    MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    MT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[1]));
    MT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    bytes = double(nnz) * sizeof(T) + double(g.nInt) * sizeof(T) +
            double(nsys) * g.nInt * 3.0 * 2.0 * sizeof(T);
    return MT_OK;
  };
  rc = (precision == MT_PREC_MIXED) ? run(float()) : run(double());
  if (rc) return rc;
  if (ms_per_launch) *ms_per_launch = double(ms) / reps;
  if (bytes_per_launch) *bytes_per_launch = bytes;
  return MT_OK;
}

int mt_bench_precond(mt_ctx* ctx, int nsys, int precision, int precond, int reps,
                     double* ms_per_apply, double* bytes_per_apply) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (nsys < 1 || nsys > 64 || reps < 1 || (precond != MT_PC_DILU && precond != MT_PC_DILU_RB))
    return mt_fail(ctx, MT_E_ARG, "mt_bench_precond: bad arguments");
  hipSetDevice(ctx->device);
  const Grid& g = ctx->hop.g;
  const size_t nv = size_t(g.vlen(nsys));
  // two systems (polarizations) per frequency, as in a forward batch
  const int nf = (nsys + 1) / 2;
  std::vector<double> freqs(nf);
  for (int f = 0; f < nf; ++f) freqs[f] = 2.0 * kPi * std::pow(10.0, -3.0 + 5.0 * f / std::max(1, nf - 1));
  std::vector<SysScal> h(nsys);
  std::vector<int> act(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&h[s], 0, sizeof(SysScal));
    h[s].status = ST_RUN;
    h[s].fidx = s / 2;
    h[s].freq_omega = freqs[s / 2];
    h[s].alpha = {1.0, 0.0};
    act[s] = s;
  }
  SysScal* d_sc;
  int* d_act;
  if ((rc = ws(ctx, "bp_sc", nsys, &d_sc)) || (rc = ws(ctx, "bp_act", nsys, &d_act))) return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, h.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, act.data(), nsys * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  float ms = 0.f;
  auto run = [&](auto tagT) -> int {
    using T = decltype(tagT);
    C2<T>* dinv;
    int r2 = build_dinv<T>(ctx, precond, false, freqs, &dinv);
    if (r2) return r2;
    C2<T>* vin;
    C2<T>* vout;
    if ((r2 = ws(ctx, sizeof(T) == 8 ? "bp_in64" : "bp_in32", nv, &vin)) ||
        (r2 = ws(ctx, sizeof(T) == 8 ? "bp_out64" : "bp_out32", nv, &vout)))
      return r2;
    const TilePlan& tp = precond == MT_PC_DILU_RB ? ctx->tiles_rb : ctx->tiles;
    DiluArgs<T, T> d = dilu_args<T, T>(tp, view<T>(ctx, false).coef, g.nE, g.nE, d_act,
                                       nsys, d_sc, vin, vout, dinv, ST_RUN);
    // what the mixed solver's inner solves run: the sweeps on bricks
    const bool bricks = sizeof(T) == 4 && precond == MT_PC_DILU_RB && ctx->brick &&
                        ctx->spmv_variant == 3 && ctx->ms_S == 4;
    if (bricks) {
      if ((r2 = ensure_brick(ctx, false))) return r2;
      const int64_t nB = ctx->bk.g.nB;
      C2<T> *bin, *bout, *bdv;
      if ((r2 = ws(ctx, "bp_inb", size_t(nB) * nsys, &bin)) ||
          (r2 = ws(ctx, "bp_outb", size_t(nB) * nsys, &bout)) ||
          (r2 = ws(ctx, "bp_dinvb", size_t(nB) * nf, &bdv)))
        return r2;
      k_bk_gather<T><<<4096, BS, 0, ctx->stream>>>(ctx->bk.bmap, nB, nf, g.nE, nB, dinv, bdv);
      d = dilu_args<T, T>(tp, reinterpret_cast<const T*>(ctx->bk.slots[0]), nB, nB, d_act, nsys,
                          d_sc, bin, bout, bdv, ST_RUN);
    }
    if (precond == MT_PC_DILU_RB && (r2 = rb_faces(ctx, nsys, d))) return r2;
    auto apply = [&]() {
      if (bricks) {
        dilu_bk_apply<T, T>(ctx, tp, d);
      } else if (precond == MT_PC_DILU_RB) {
        dilu_rb_apply<T, T>(ctx, tp, d);
      } else {
        dilu_fwd<T, T>(ctx, tp, d);
        dilu_bwd<T, T>(ctx, tp, d);
      }
    };
    apply();  // warm-up
    MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    for (int q = 0; q < reps; ++q) apply();
    MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    MT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[1]));
    MT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
    if (bytes_per_apply)
      *bytes_per_apply = 6.0 * double(g.nInt) * sizeof(T) + double(nsys) * g.nInt * 3.0 * 2.0 * sizeof(T);
    return MT_OK;
  };
  rc = (precision == MT_PREC_MIXED) ? run(float()) : run(double());
  if (rc) return rc;
  if (ms_per_apply) *ms_per_apply = double(ms) / reps;
  return MT_OK;
}

int mt_set_dilu_variant(mt_ctx* ctx, int variant) {
  if (!ctx || (variant != 1 && variant != 4 && variant != 5 && variant != 6))
    return mt_fail(ctx, MT_E_ARG, "mt_set_dilu_variant: variant is 1, 4, 5 or 6");
  ctx->rb_variant = variant;
  return MT_OK;
}

int mt_set_inner_layout(mt_ctx* ctx, int bricks, int systems_per_block) {
  if (!ctx || bricks < 0 || bricks > 2 || systems_per_block < 0 || systems_per_block > 64)
    return mt_fail(ctx, MT_E_ARG, "mt_set_inner_layout: bricks is 0, 1 or 2, systems_per_block 0..64");
  ctx->brick = bricks != 0;
  ctx->bk_pairs = bricks == 2;
  ctx->bk_sg = systems_per_block;
  return MT_OK;
}

int mt_set_spmv_variant(mt_ctx* ctx, int variant, int layers_per_segment) {
  if (!ctx || variant < 0 || variant > 3 || layers_per_segment < 0 ||
      (variant == 3 && layers_per_segment != 0 && layers_per_segment != 4 && layers_per_segment != 8))
    return mt_fail(ctx, MT_E_ARG, "mt_set_spmv_variant: variant 0..3, layers >= 0 (4 or 8 systems for 3)");
  ctx->spmv_variant = variant;
  if (variant == 3) {
    if (layers_per_segment > 0) ctx->ms_S = layers_per_segment;
  } else if (layers_per_segment > 0) {
    ctx->km_ks = layers_per_segment;
  }
  return MT_OK;
}

int mt_set_tuning(mt_ctx* ctx, int groups_per_system) {
  if (!ctx) return MT_E_ARG;
  ctx->G_override = groups_per_system > 0 ? ((groups_per_system + 7) / 8) * 8 : 0;
  return MT_OK;
}

// One inner iteration's vector updates (k_pupdate + k_rupdate_fused) on
// nsys systems of n entries (the inner layout), ms per pair of launches.
static int time_updates(mt_ctx* ctx, int nsys, int64_t n, int reps, double* ms_out) {
  std::vector<SysScal> h(nsys);
  std::vector<int> act(nsys);
  for (int s = 0; s < nsys; ++s) {
    std::memset(&h[s], 0, sizeof(SysScal));
    h[s].status = ST_RUN;
    h[s].alpha = {0.5, 0.0};
    h[s].omg = {0.25, 0.0};
    h[s].beta = {0.125, 0.0};
    h[s].rho = {1.0, 0.0};
    act[s] = s;
  }
  SysScal* d_sc;
  int* d_act;
  C2<float>* v;
  Z* part;
  IterArgs<float, float> a{};
  a.G = pick_G(ctx, nsys);
  const size_t nv = size_t(n) * nsys;
  int rc;
  if ((rc = ws(ctx, "tu_sc", nsys, &d_sc)) || (rc = ws(ctx, "tu_act", nsys, &d_act)) ||
      (rc = ws(ctx, "tu_vec", nv * 8, &v)) || (rc = ws(ctx, "tu_part", size_t(nsys) * a.G * 2 + 64, &part)))
    return rc;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_sc, h.data(), nsys * sizeof(SysScal), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemcpyAsync(d_act, act.data(), nsys * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  a.sc = d_sc; a.act = d_act; a.nact = nsys; a.nE = n; a.pc = PCK_DILU_RB; a.fused = 1;
  a.partials = part;
  a.x = v; a.r = v + nv; a.rt = v + 2 * nv; a.p = v + 3 * nv; a.ph = v + 4 * nv;
  a.v = v + 5 * nv; a.sh = v + 6 * nv; a.t = v + 7 * nv;
  const int nblk = nsys * a.G;
  auto run = [&]() {
    k_pupdate<float, float><<<nblk, BS, 0, ctx->stream>>>(a);
    k_rupdate_fused<float, float><<<nblk, BS, 0, ctx->stream>>>(a);
  };
  run();  // warm-up
  MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
  for (int q = 0; q < reps; ++q) run();
  MT_HIP_CHECK(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
  MT_HIP_CHECK(ctx, hipEventSynchronize(ctx->ev[1]));
  float ms = 0.f;
  MT_HIP_CHECK(ctx, hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  *ms_out = double(ms) / reps;
  return MT_OK;
}

int mt_autotune(mt_ctx* ctx, int nsys, int* variant) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (nsys < 1 || nsys > 64) return mt_fail(ctx, MT_E_ARG, "mt_autotune: nsys is 1..64");
  hipSetDevice(ctx->device);
  const bool bricks = ctx->brick && ctx->spmv_variant == 3 && ctx->ms_S == 4;
  // 1. the brick SpMV's systems per workgroup
  if (bricks) {
    double best = 1e30;
    int bestS = ctx->bk_sg;
    for (int sg : {0, 8, 4}) {
      ctx->bk_sg = sg;
      double ms = 0.0;
      if ((rc = mt_bench_spmv(ctx, nsys, MT_PREC_MIXED, 5, &ms, nullptr))) return rc;
      if (ms < best) { best = ms; bestS = sg; }
    }
    ctx->bk_sg = bestS;
  }
  // 2. the row chunks of the vector updates, as a total over the systems
  int64_t n = ctx->hop.g.nE;
  if (bricks) {
    if ((rc = ensure_brick(ctx, false))) return rc;
    n = ctx->bk.g.nB;
  }
  const int keepG = ctx->G_override;
  ctx->G_override = 0;
  double best = 1e30;
  int bestT = 2048;
  for (int T : {1024, 2048, 4096, 8192, 16384, 32768}) {
    if (T / nsys < 8) continue;
    ctx->rows_total = T;
    double ms = 0.0;
    if ((rc = time_updates(ctx, nsys, n, 5, &ms))) {
      ctx->G_override = keepG;
      return rc;
    }
    if (ms < best) { best = ms; bestT = T; }
  }
  ctx->rows_total = bestT;
  if (variant) *variant = bestT;
  return MT_OK;
}

int mt_timing_enable(mt_ctx* ctx, int on) {
  if (!ctx) return MT_E_ARG;
  ctx->timing = on != 0;
  for (auto& k : ctx->kt) k = mt_ctx::KTime();
  return MT_OK;
}

int mt_kernel_timing(const mt_ctx* ctx, int family, long long* launches, double* ms,
                     double* bytes) {
  if (!ctx || family < 0 || family > 1) return MT_E_ARG;
  const mt_ctx::KTime& k = ctx->kt[family];
  if (launches) *launches = k.launches;
  if (ms) *ms = k.ms;
  if (bytes) *bytes = k.bytes;
  return MT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Data functionals: receivers, responses, adjoint sources
// ---------------------------------------------------------------------------
template <class T>
static int dev_upload(mt_ctx* ctx, const std::vector<T>& v, T** out) {
  MT_HIP_CHECK(ctx, hipMalloc(out, std::max<size_t>(v.size(), 1) * sizeof(T)));
  if (!v.empty())
    MT_HIP_CHECK(ctx, hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return MT_OK;
}

// real components per data type, dataTypes.f90:87-222
static int dtype_ncomp(int dt) {
  static const int nc[12] = {0, 8, 4, 4, 8, 4, 4, 2, 2, 2, 2, 2};
  return (dt >= 1 && dt <= 11) ? nc[dt] : 0;
}

int mt_receivers_set(mt_ctx* ctx, int nx, int ny, int nzAir, int nzEarth, const double* dx,
                     const double* dy, const double* dz, const double* origin, int nsite,
                     const double* x, const double* r, const double* azimuth) {
  if (!ctx || !dx || !dy || !dz || !x || nsite < 1 || nx < 2 || ny < 2 || nzAir < 0 || nzEarth < 1)
    return mt_fail(ctx, MT_E_ARG, "mt_receivers_set: bad arguments");
  Grid g;
  g.set(nx, ny, nzAir + nzEarth);
  const double org0[3] = {0.0, 0.0, 0.0};
  SiteRows rows;
  std::string err;
  // without reference locations every site is its own reference
  int rc = build_site_rows(g, dx, dy, dz, nzAir, origin ? origin : org0, nsite, x, r ? r : x,
                           azimuth, rows, err);
  if (rc) return mt_fail(ctx, rc, "mt_receivers_set: " + err);
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  free_sites(ctx->sites);
  SiteData& S = ctx->sites;
  if ((rc = dev_upload(ctx, rows.ptr, &S.ptr)) || (rc = dev_upload(ctx, rows.idx, &S.idx)) ||
      (rc = dev_upload(ctx, rows.val, &S.val)) || (rc = dev_upload(ctx, rows.rot, &S.rot)) ||
      (rc = dev_upload(ctx, rows.uedge, &S.uedge)) || (rc = dev_upload(ctx, rows.gptr, &S.gptr)) ||
      (rc = dev_upload(ctx, rows.grow, &S.grow)) || (rc = dev_upload(ctx, rows.gval, &S.gval))) {
    free_sites(ctx->sites);
    return rc;
  }
  S.nsite = nsite;
  S.nu = int(rows.uedge.size());
  S.nE = g.nE;
  S.ready = true;
  return MT_OK;
}

// comb = 0 without hipMemsetAsync (a fill kernel, safe inside captured work)
__global__ void k_zero_cd(int64_t n, Cd* __restrict__ out) {
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x)
    out[q] = {0.0, 0.0};
}

static int site_launch(mt_ctx* ctx, int dtype, int nfreq, const double* omegas, const void* e_dev,
                       double* resp_host, const double* d, const unsigned char* exist,
                       void* comb_dev, bool noalloc = false) {
  const SiteData& S = ctx->sites;
  const int ncomp = dtype_ncomp(dtype);
  if (!ncomp) return mt_fail(ctx, MT_E_ARG, "data type must be 1..11 (dataTypes.f90 numbering)");
  if (!S.ready) return mt_fail(ctx, MT_E_STATE, "no receivers: call mt_receivers_set first");
  if (nfreq < 1 || !omegas || !e_dev) return mt_fail(ctx, MT_E_ARG, "bad frequency batch");
  for (int f = 0; f < nfreq; ++f)
    if (!(omegas[f] > 0.0)) return mt_fail(ctx, MT_E_ARG, "omegas must be positive");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const bool adj = comb_dev != nullptr;
  const size_t nsd = size_t(nfreq) * S.nsite;
  void *w_om, *w_out, *w_d = nullptr, *w_ex = nullptr;
  int rc;
  if (noalloc) {
    // every buffer in this device's scratch array, nothing allocated
    if (!adj) return mt_fail(ctx, MT_E_ARG, "the allocation-free path is LmultT's");
    if (nfreq > kNoAllocMaxFreq || nsd > size_t(kNoAllocMaxSiteFreq))
      return mt_fail(ctx, MT_E_ARG, "mt_lmultT_noalloc: nfreq * nsite = " + std::to_string(nsd) +
                                        " (nfreq " + std::to_string(nfreq) + ") exceeds its scratch: at most " +
                                        std::to_string(kNoAllocMaxSiteFreq) + " (nfreq at most " +
                                        std::to_string(kNoAllocMaxFreq) + ")");
    unsigned char* base = ctx->scratch_dev;
    if (!base) return mt_fail(ctx, MT_E_STATE, "scratch array address not resolved");
    w_out = base + kNoAllocCoefOff;
    w_d = base + kNoAllocDataOff;
    w_om = base + kNoAllocOmegaOff;
    if (exist) w_ex = base + kNoAllocExistOff;
  } else {
    if ((rc = ws_get(ctx, "df_omega", nfreq * sizeof(double), &w_om))) return rc;
    if ((rc = ws_get(ctx, "df_out", adj ? nsd * 14 * sizeof(Cd) : nsd * ncomp * sizeof(double), &w_out)))
      return rc;
    if (adj) {
      if ((rc = ws_get(ctx, "df_d", nsd * ncomp * sizeof(double), &w_d))) return rc;
      if (exist && (rc = ws_get(ctx, "df_exist", nsd * ncomp, &w_ex))) return rc;
    }
  }
  MT_HIP_CHECK(ctx, hipMemcpyAsync(w_om, omegas, nfreq * sizeof(double), hipMemcpyHostToDevice,
                                   ctx->stream));
  if (adj) {
    MT_HIP_CHECK(ctx, hipMemcpyAsync(w_d, d, nsd * ncomp * sizeof(double), hipMemcpyHostToDevice,
                                     ctx->stream));
    if (exist)
      MT_HIP_CHECK(ctx, hipMemcpyAsync(w_ex, exist, nsd * ncomp, hipMemcpyHostToDevice, ctx->stream));
  }
  SiteArgs a;
  a.nsite = S.nsite; a.nfreq = nfreq; a.dtype = dtype; a.ncomp = ncomp; a.want_coef = adj;
  a.omegas = (const double*)w_om;
  a.e = (const Cd*)e_dev;
  a.nE = S.nE;
  a.ptr = S.ptr; a.idx = S.idx; a.val = S.val; a.rot = S.rot;
  a.resp = adj ? nullptr : (double*)w_out;
  a.d = (const double*)w_d;
  a.exist = (const unsigned char*)w_ex;
  a.coef = adj ? (Cd*)w_out : nullptr;
  const int nt = nfreq * S.nsite;
  hipLaunchKernelGGL(k_site_eval, dim3((nt + 63) / 64), dim3(64), 0, ctx->stream, a);
  MT_HIP_CHECK(ctx, hipGetLastError());
  if (adj) {
    const int64_t ncomb = int64_t(nfreq) * 2 * S.nE;
    if (noalloc)
      k_zero_cd<<<unsigned(std::min<int64_t>(4096, (ncomb + 255) / 256)), 256, 0, ctx->stream>>>(
          ncomb, (Cd*)comb_dev);
    else
      MT_HIP_CHECK(ctx, hipMemsetAsync(comb_dev, 0, size_t(ncomb) * sizeof(Cd), ctx->stream));
    const int64_t ng = int64_t(nfreq) * S.nu;
    if (ng > 0) {
      hipLaunchKernelGGL(k_lmultT_gather, dim3(unsigned((ng + 255) / 256)), dim3(256), 0, ctx->stream,
                         nfreq, S.nu, S.nsite, S.nE, (const double*)w_om, S.uedge, S.gptr, S.grow,
                         S.gval, (const Cd*)w_out, (Cd*)comb_dev);
      MT_HIP_CHECK(ctx, hipGetLastError());
    }
  } else {
    MT_HIP_CHECK(ctx, hipMemcpyAsync(resp_host, w_out, nsd * ncomp * sizeof(double),
                                     hipMemcpyDeviceToHost, ctx->stream));
  }
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_data_resp(mt_ctx* ctx, int dtype, int nfreq, const double* omegas, const void* e_dev,
                 double* resp) {
  if (!ctx || !resp) return mt_fail(ctx, MT_E_ARG, "mt_data_resp: null argument");
  return site_launch(ctx, dtype, nfreq, omegas, e_dev, resp, nullptr, nullptr, nullptr);
}

int mt_lmultT(mt_ctx* ctx, int dtype, int nfreq, const double* omegas, const void* e_dev,
              const double* d, const unsigned char* exist, void* comb_dev) {
  if (!ctx || !d || !comb_dev) return mt_fail(ctx, MT_E_ARG, "mt_lmultT: null argument");
  return site_launch(ctx, dtype, nfreq, omegas, e_dev, nullptr, d, exist, comb_dev);
}

int mt_lmultT_noalloc(mt_ctx* ctx, int dtype, int nfreq, const double* omegas, const void* e_dev,
                      const double* d, const unsigned char* exist, void* comb_dev) {
  if (!ctx || !d || !comb_dev) return mt_fail(ctx, MT_E_ARG, "mt_lmultT_noalloc: null argument");
  return site_launch(ctx, dtype, nfreq, omegas, e_dev, nullptr, d, exist, comb_dev, true);
}

// ---------------------------------------------------------------------------
// UpdateCond on the device
// ---------------------------------------------------------------------------
// The UpdateCond kernels' arguments over the cond_* workspaces: the grid's
// geometry (uploaded once per grid) and, when logcond is not null, the cell
// conductivities of that model (exp on the host, as the host assembly forms
// them; air layers at exp(air_logcond)).
static int cond_args(mt_ctx* ctx, const double* logcond, double air_logcond, CondArgs& a) {
  HostOperator& H = ctx->hop;
  const Grid& g = H.g;
  const int nx = g.nx, ny = g.ny, nz = g.nz, nza = H.nzAir;
  const int64_t nN = int64_t(nx + 1) * (ny + 1) * (nz + 1), ncell = int64_t(nx) * ny * nz;
  int rc;
  a.nx = nx; a.ny = ny; a.nz = nz; a.nzAir = nza;
  a.nXe = g.nXe; a.nYe = g.nYe; a.nE = g.nE; a.nInt = g.nInt;
  double *ddx, *ddy, *ddz, *dlx, *dly, *dlz, *dcc, *dcs, *dsE, *dsN, *dm2;
  int* dslot;
  unsigned char* dearth;
  if ((rc = ws(ctx, "cond_dx", nx, &ddx)) || (rc = ws(ctx, "cond_dy", ny, &ddy)) ||
      (rc = ws(ctx, "cond_dz", nz, &ddz)) || (rc = ws(ctx, "cond_delx", nx + 1, &dlx)) ||
      (rc = ws(ctx, "cond_dely", ny + 1, &dly)) || (rc = ws(ctx, "cond_delz", nz + 1, &dlz)) ||
      (rc = ws(ctx, "cond_cc", size_t(kSlots) * g.nE, &dcc)) || (rc = ws(ctx, "cond_slot", 36, &dslot)) ||
      (rc = ws(ctx, "cond_csig", ncell, &dcs)) || (rc = ws(ctx, "cond_sige", g.nE, &dsE)) ||
      (rc = ws(ctx, "cond_earth", g.nE, &dearth)) || (rc = ws(ctx, "cond_sign", nN, &dsN)) ||
      (rc = ws(ctx, "cond_m2", nN, &dm2)))
    return rc;
  if (!ctx->cond_ready) {
    auto half = [](const std::vector<double>& d) {
      const int n = int(d.size());
      std::vector<double> h(n + 1);
      h[0] = d[0];
      for (int i = 1; i < n; ++i) h[i] = d[i - 1] + d[i];
      h[n] = d[n - 1];
      for (auto& v : h) v /= 2.0;
      return h;
    };
    const std::vector<double> hx = half(H.dx), hy = half(H.dy), hz = half(H.dz);
    int slot[3][2][6];
    graddiv_slots(slot);
    MT_HIP_CHECK(ctx, hipMemcpy(ddx, H.dx.data(), nx * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(ddy, H.dy.data(), ny * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(ddz, H.dz.data(), nz * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(dlx, hx.data(), (nx + 1) * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(dly, hy.data(), (ny + 1) * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(dlz, hz.data(), (nz + 1) * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(dcc, H.coefCC.data(), size_t(kSlots) * g.nE * sizeof(double), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(dslot, &slot[0][0][0], 36 * sizeof(int), hipMemcpyHostToDevice));
    ctx->cond_ready = true;
  }
  if (logcond) {
    std::vector<double> cs(ncell);
    const double airC = std::exp(air_logcond);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i)
          cs[i + int64_t(nx) * (j + int64_t(ny) * k)] =
              k < nza ? airC : std::exp(logcond[i + int64_t(nx) * (j + int64_t(ny) * (k - nza))]);
    MT_HIP_CHECK(ctx, hipMemcpy(dcs, cs.data(), ncell * sizeof(double), hipMemcpyHostToDevice));
  }
  a.dx = ddx; a.dy = ddy; a.dz = ddz; a.delX = dlx; a.delY = dly; a.delZ = dlz;
  a.csig = dcs; a.imap = ctx->dop.imap; a.coefCC = dcc; a.slot = dslot;
  a.sigE = dsE; a.earthE = dearth; a.sigN = dsN; a.m2 = dm2;
  a.coef = nullptr; a.vms1 = nullptr;
  return MT_OK;
}

int mt_update_cond(mt_ctx* ctx, const double* logcond, double air_logcond) {
  if (!ctx || !logcond) return mt_fail(ctx, MT_E_ARG, "mt_update_cond: null argument");
  int rc = need_op(ctx);
  if (rc) return rc;
  HostOperator& H = ctx->hop;
  if (!H.have_model || H.coefCC.empty())
    return mt_fail(ctx, MT_E_STATE, "mt_update_cond needs mt_assemble on this grid");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const Grid& g = H.g;
  const int nx = g.nx, ny = g.ny;
  CondArgs a;
  if ((rc = cond_args(ctx, logcond, air_logcond, a))) return rc;
  DeviceOperator& d = ctx->dop;
  a.coef = d.coef64; a.vms1 = d.vms64;
  k_cond_edges<<<2048, 256, 0, ctx->stream>>>(a);
  k_cond_nodes<<<1024, 256, 0, ctx->stream>>>(a);
  k_cond_rows<<<2048, 256, 0, ctx->stream>>>(a);
  const int64_t nc = int64_t(kSlots) * g.nE;
  k_cast_array<float><<<4096, 256, 0, ctx->stream>>>(d.coef64, d.coef32, nc);
  k_cast_array<float><<<1024, 256, 0, ctx->stream>>>(d.vms64, d.vms32, g.nE);
  k_rows_from_slots<double><<<2048, 256, 0, ctx->stream>>>(d.coef64, d.vms64, g.nE, d.row64);
  k_rows_from_slots<float><<<2048, 256, 0, ctx->stream>>>(d.coef32, d.vms32, g.nE, d.row32);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  // what was derived from the previous model
  hipFree(d.coefT64); hipFree(d.coefT32); hipFree(d.rowT64); hipFree(d.rowT32);
  d.coefT64 = nullptr; d.coefT32 = nullptr; d.rowT64 = nullptr; d.rowT32 = nullptr;
  free_brick(ctx->bk);
  hipFree(ctx->dc.coefN); hipFree(ctx->dc.wE); hipFree(ctx->dc.invL); hipFree(ctx->dc.dinv);
  hipFree(ctx->dc.tiles.tiles); hipFree(ctx->dc.tiles.shape);
  ctx->dc = DivcorData();
  ctx->nnz_stencil = -1;
  H.logcond.assign(logcond, logcond + int64_t(nx) * ny * H.nzEarth);
  H.airLog = air_logcond;
  ctx->host_stale = true;
  ctx->sens_fields = true;   // sigma_E, sigma_N, m2 of this model are in cond_*
  return MT_OK;
}

// ---------------------------------------------------------------------------
// Sensitivity path on the device: FWDsolve3D's forward and TRN
// post-processing, the TRN right-hand side, PmultT and the gradient assembly
// ---------------------------------------------------------------------------

// The model fields (sigma_E, sigma_N, node weights; from the last
// mt_update_cond, else computed here for the assembled model) and the
// sigma-free data of the sensitivity kernels: ghost volumes, the all-edge ->
// interior-row map and CCib transposed (rows in EDGEb order).
struct SensData {
  CondArgs a;
  double *gE, *gN, *tE, *tN, *bt_val;
  unsigned char* cntE;
  int *rowOf, *bmap, *bt_row, *bt_col;
  int64_t nN, nBdry;
};

static int ensure_sens(mt_ctx* ctx, SensData& S) {
  if (!ctx) return MT_E_ARG;
  int rc = need_op(ctx);
  if (rc) return rc;
  HostOperator& H = ctx->hop;
  if (!H.have_model || H.coefCC.empty())
    return mt_fail(ctx, MT_E_STATE, "the sensitivity entry points need mt_assemble");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const Grid& g = H.g;
  const int64_t nN = int64_t(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  const int64_t nB = int64_t(H.bmap.size());
  S.nN = nN;
  S.nBdry = nB;
  if ((rc = cond_args(ctx, ctx->sens_fields ? nullptr : H.logcond.data(), H.airLog, S.a))) return rc;
  if (!ctx->sens_fields) {
    double* vtmp;
    if ((rc = ws(ctx, "sens_vms", g.nE, &vtmp))) return rc;
    CondArgs b = S.a;
    b.vms1 = vtmp;
    k_cond_edges<<<2048, 256, 0, ctx->stream>>>(b);
    k_cond_nodes<<<1024, 256, 0, ctx->stream>>>(b);
    MT_HIP_CHECK(ctx, hipGetLastError());
    ctx->sens_fields = true;
  }
  if ((rc = ws(ctx, "sens_ge", g.nE, &S.gE)) || (rc = ws(ctx, "sens_gn", nN, &S.gN)) ||
      (rc = ws(ctx, "sens_cnt", g.nE, &S.cntE)) || (rc = ws(ctx, "sens_rowof", g.nE, &S.rowOf)) ||
      (rc = ws(ctx, "sens_bmap", std::max<int64_t>(1, nB), &S.bmap)) ||
      (rc = ws(ctx, "sens_btrow", nB + 1, &S.bt_row)) || (rc = ws(ctx, "sens_te", g.nE, &S.tE)) ||
      (rc = ws(ctx, "sens_tn", nN, &S.tN)))
    return rc;
  if (!ctx->sens_geom) {
    k_sens_ghost<<<2048, 256, 0, ctx->stream>>>(S.a, S.gE, S.cntE, S.gN);
    MT_HIP_CHECK(ctx, hipGetLastError());
    std::vector<int> rowOf(g.nE, 0);
    for (int64_t r = 0; r < g.nInt; ++r) rowOf[H.imap[r]] = int(r);
    // CCib^T: for each boundary edge, its (interior row, value) pairs in
    // increasing row order
    std::vector<int> bpos(g.nE, -1);
    for (int64_t b = 0; b < nB; ++b) bpos[H.bmap[b]] = int(b);
    std::vector<int> cnt(nB + 1, 0);
    for (size_t q = 0; q < H.ccib_col.size(); ++q) cnt[bpos[H.ccib_col[q]] + 1]++;
    for (int64_t b = 0; b < nB; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> bcol(H.ccib_col.size()), fill(cnt.begin(), cnt.end() - 1);
    std::vector<double> bval(H.ccib_col.size());
    for (int64_t r = 0; r < g.nInt; ++r)
      for (int q = H.ccib_row[r]; q < H.ccib_row[r + 1]; ++q) {
        const int b = bpos[H.ccib_col[q]];
        bcol[fill[b]] = int(r);
        bval[fill[b]] = H.ccib_val[q];
        ++fill[b];
      }
    const size_t nt = std::max<size_t>(1, bcol.size());
    if ((rc = ws(ctx, "sens_btcol", nt, &S.bt_col)) || (rc = ws(ctx, "sens_btval", nt, &S.bt_val)))
      return rc;
    MT_HIP_CHECK(ctx, hipMemcpy(S.rowOf, rowOf.data(), g.nE * sizeof(int), hipMemcpyHostToDevice));
    if (nB) MT_HIP_CHECK(ctx, hipMemcpy(S.bmap, H.bmap.data(), nB * sizeof(int), hipMemcpyHostToDevice));
    MT_HIP_CHECK(ctx, hipMemcpy(S.bt_row, cnt.data(), (nB + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (!bcol.empty()) {
      MT_HIP_CHECK(ctx, hipMemcpy(S.bt_col, bcol.data(), bcol.size() * sizeof(int), hipMemcpyHostToDevice));
      MT_HIP_CHECK(ctx, hipMemcpy(S.bt_val, bval.data(), bval.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    ctx->sens_geom = true;
  } else {
    const size_t nt = std::max<size_t>(1, H.ccib_col.size());
    if ((rc = ws(ctx, "sens_btcol", nt, &S.bt_col)) || (rc = ws(ctx, "sens_btval", nt, &S.bt_val)))
      return rc;
  }
  return MT_OK;
}

static unsigned grid_for(int64_t n) { return unsigned(std::min<int64_t>(8192, (n + 255) / 256)); }

int mt_post_fwd_dev(mt_ctx* ctx, int nsys, const void* x_dev, void* e_dev) {
  if (!ctx || !x_dev || !e_dev || nsys < 1) return mt_fail(ctx, MT_E_ARG, "mt_post_fwd_dev: bad arguments");
  int rc = need_op(ctx);
  if (rc) return rc;
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  const Grid& g = ctx->hop.g;
  k_post_fwd<<<grid_for(g.nInt * nsys), 256, 0, ctx->stream>>>(nsys, g.nInt, g.nE, ctx->dop.imap,
                                                              (const Z*)x_dev, (Z*)e_dev);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_rhs_trn_dev(mt_ctx* ctx, int nsys, const double* omegas, const void* s_dev, void* b_dev,
                   int gd_correct) {
  if (!ctx || !s_dev || !b_dev || nsys < 1 || (gd_correct && !omegas))
    return mt_fail(ctx, MT_E_ARG, "mt_rhs_trn_dev: bad arguments");
  for (int s = 0; gd_correct && s < nsys; ++s)
    if (!(omegas[s] > 0.0)) return mt_fail(ctx, MT_E_ARG, "mt_rhs_trn_dev: omega must be > 0");
  SensData S;
  int rc = ensure_sens(ctx, S);
  if (rc) return rc;
  Z* div = nullptr;
  if (gd_correct) {
    // iOmegaMuInv = 1 / (i ISIGN omega mu0) per system, divided as the host does
    std::vector<Z> inv(nsys);
    for (int s = 0; s < nsys; ++s) {
      const mt_zdouble v = 1.0 / mt_zdouble(0.0, kIsign * omegas[s] * kMu0);
      inv[s] = {v.real(), v.imag()};
    }
    Z* dinv;
    if ((rc = ws(ctx, "sens_inv", nsys, &dinv)) || (rc = ws(ctx, "sens_div", size_t(nsys) * S.nN, &div)))
      return rc;
    MT_HIP_CHECK(ctx, hipMemcpy(dinv, inv.data(), nsys * sizeof(Z), hipMemcpyHostToDevice));
    k_trn_div<<<grid_for(S.nN * nsys), 256, 0, ctx->stream>>>(S.a, nsys, dinv, (const Z*)s_dev, div);
  }
  k_trn_rhs<<<grid_for(S.a.nInt * nsys), 256, 0, ctx->stream>>>(S.a, nsys, gd_correct ? 1 : 0,
                                                                 (const Z*)s_dev, div, (Z*)b_dev);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_post_trn_dev(mt_ctx* ctx, int nsys, const void* ei_dev, void* e_dev) {
  if (!ctx || !ei_dev || !e_dev || nsys < 1) return mt_fail(ctx, MT_E_ARG, "mt_post_trn_dev: bad arguments");
  SensData S;
  int rc = ensure_sens(ctx, S);
  if (rc) return rc;
  k_post_trn<<<grid_for((S.a.nInt + S.nBdry) * nsys), 256, 0, ctx->stream>>>(
      S.a, nsys, S.bmap, S.nBdry, S.bt_row, S.bt_col, S.bt_val, (const Z*)ei_dev, (Z*)e_dev);
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_pmultT_dev(mt_ctx* ctx, int nfreq, const double* omegas, int npol, const void* e0_dev,
                  const void* lam_dev, void* dm_dev, int accumulate) {
  if (!ctx || !omegas || !e0_dev || !lam_dev || !dm_dev || nfreq < 1 || npol < 1)
    return mt_fail(ctx, MT_E_ARG, "mt_pmultT_dev: bad arguments");
  SensData S;
  int rc = ensure_sens(ctx, S);
  if (rc) return rc;
  const CondArgs& a = S.a;
  const int64_t ncell = int64_t(a.nx) * a.ny * (a.nz - a.nzAir);
  const Z* e0 = (const Z*)e0_dev;
  const Z* lam = (const Z*)lam_dev;
  for (int f = 0; f < nfreq; ++f) {
    const int64_t off = int64_t(f) * npol * a.nE;
    k_pmult_cells<<<grid_for(ncell), 256, 0, ctx->stream>>>(a, npol, omegas[f], e0 + off, lam + off,
                                                           S.gE, S.cntE, (double*)dm_dev,
                                                           (accumulate || f > 0) ? 1 : 0);
  }
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

int mt_gradient_dev(mt_ctx* ctx, int nfreq, const double* omegas, int npol, const void* e_dev,
                    const void* lam_dev, void* dm_dev, int mode, int accumulate) {
  if (!ctx || !omegas || !e_dev || !lam_dev || !dm_dev || nfreq < 1 || npol < 1 || mode < 0 ||
      mode > 1)
    return mt_fail(ctx, MT_E_ARG, "mt_gradient_dev: bad arguments");
  SensData S;
  int rc = ensure_sens(ctx, S);
  if (rc) return rc;
  const CondArgs& a = S.a;
  const int64_t ncell = int64_t(a.nx) * a.ny * (a.nz - a.nzAir);
  Z *D, *Sn;
  if ((rc = ws(ctx, "sens_dnode", S.nN, &D)) || (rc = ws(ctx, "sens_snode", S.nN, &Sn))) return rc;
  if (!accumulate) MT_HIP_CHECK(ctx, hipMemsetAsync(dm_dev, 0, ncell * sizeof(double), ctx->stream));
  const Z* e = (const Z*)e_dev;
  const Z* lam = (const Z*)lam_dev;
  for (int f = 0; f < nfreq; ++f) {
    MT_HIP_CHECK(ctx, hipMemsetAsync(S.tE, 0, a.nE * sizeof(double), ctx->stream));
    MT_HIP_CHECK(ctx, hipMemsetAsync(S.tN, 0, S.nN * sizeof(double), ctx->stream));
    for (int p = 0; p < npol; ++p) {
      const Z* es = e + (int64_t(f) * npol + p) * a.nE;
      const Z* ls = lam + (int64_t(f) * npol + p) * a.nInt;
      if (mode) k_grad_nodes<<<grid_for(S.nN), 256, 0, ctx->stream>>>(a, S.rowOf, ls, es, D, Sn);
      k_grad_accum<<<grid_for(a.nInt + S.nN), 256, 0, ctx->stream>>>(a, omegas[f], mode, ls, es, D, Sn,
                                                                     S.tE, S.tN);
    }
    k_grad_cells<<<grid_for(ncell), 256, 0, ctx->stream>>>(a, mode, S.tE, S.tN, S.gE, S.gN,
                                                          (double*)dm_dev);
  }
  MT_HIP_CHECK(ctx, hipGetLastError());
  MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return MT_OK;
}

// ---------------------------------------------------------------------------
// QMR (the reference's default solver, solver.f90:140-370) on the device
// ---------------------------------------------------------------------------
// One FP64 system on padded all-edge vectors.  M1 = L and M2 = U of the block
// DILU in the reference ordering (the level-scheduled tile sweeps); the
// adjoint solves use the same sweeps on the transposed stencil with the
// conjugated pivots: with q = conj(p) (the pivots of A^H's blocks) and c the
// couplings of A^T,
//   U^-H x = Q^-1 Lsweep_{c,q}(x),    L^-H w = Usweep_{c,q}(Q w),
// where Lsweep / Usweep are the forward / backward DILU sweeps.  Scalars are
// formed on the host from per-block partial sums in a fixed order.
namespace {
struct QmrEngine {
  mt_ctx* ctx;
  int64_t n;          // all-edge length the vector kernels cover
  int G = 512;        // blocks of the inner-product kernel
  bool pc;            // block DILU (else M1 = M2 = I)
  double omega;
  Z *dinv = nullptr, *dinvH = nullptr, *part = nullptr;
  SysScal* sc = nullptr;
  int* act = nullptr;
  std::vector<Z> hpart;
  unsigned grid() const { return unsigned(std::min<int64_t>(4096, (n + 255) / 256)); }

  int dots(int nd, const Z* a0, const Z* b0, const Z* a1, const Z* b1, const Z* a2, const Z* b2,
           mt_zdouble* out) {
    k_q_dots<<<G, QMR_BS, 0, ctx->stream>>>(n, nd, a0, b0, a1, b1, a2, b2, part);
    MT_HIP_CHECK(ctx, hipGetLastError());
    hpart.resize(size_t(3) * G);
    MT_HIP_CHECK(ctx, hipMemcpyAsync(hpart.data(), part, sizeof(Z) * nd * G, hipMemcpyDeviceToHost,
                                     ctx->stream));
    MT_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < nd; ++k) {
      mt_zdouble s = 0.0;
      for (int g = 0; g < G; ++g) s += mt_zdouble(hpart[size_t(k) * G + g].x, hpart[size_t(k) * G + g].y);
      out[k] = s;
    }
    return MT_OK;
  }
  DiluArgs<double, double> dargs(bool transposed, const Z* in, Z* out, const Z* dv) {
    const Grid& g = ctx->hop.g;
    return dilu_args<double, double>(ctx->tiles, view<double>(ctx, transposed).coef, g.nE, g.nE, act, 1,
                                     sc, in, out, dv, ST_RUN);
  }
  void copy(const Z* in, Z* out) { k_q_diag<<<grid(), 256, 0, ctx->stream>>>(n, 0, in, nullptr, out); }
  // y = L^-1 x
  void m1(const Z* x, Z* y) {
    if (!pc) { copy(x, y); return; }
    dilu_fwd<double, double>(ctx, ctx->tiles, dargs(false, x, y, dinv));
  }
  // y = U^-1 x
  void m2(const Z* x, Z* y) {
    copy(x, y);
    if (pc) dilu_bwd<double, double>(ctx, ctx->tiles, dargs(false, y, y, dinv));
  }
  // y = L^-H x
  void m1h(const Z* x, Z* y) {
    if (!pc) { copy(x, y); return; }
    k_q_diag<<<grid(), 256, 0, ctx->stream>>>(n, 1, x, dinvH, y);
    dilu_bwd<double, double>(ctx, ctx->tiles, dargs(true, y, y, dinvH));
  }
  // y = U^-H x (t: scratch)
  void m2h(const Z* x, Z* y, Z* t) {
    if (!pc) { copy(x, y); return; }
    dilu_fwd<double, double>(ctx, ctx->tiles, dargs(true, x, t, dinvH));
    k_q_diag<<<grid(), 256, 0, ctx->stream>>>(n, 2, t, dinvH, y);
  }
  void apply(bool adjoint, const Z* x, Z* y) {
    // A x, or A^H x = AAii^T x - i ISIGN omega vms1 x
    k_apply<double, double><<<1024, BS, 0, ctx->stream>>>(view<double>(ctx, adjoint), x, y,
                                                          adjoint ? -omega : omega);
  }
};
}  // namespace

int mt_qmr(mt_ctx* ctx, double omega, const mt_zdouble* b, mt_zdouble* x, int maxit, double tol,
           int precond, int* niter, double* rerr, int* failed) {
  int rc = need_op(ctx);
  if (rc) return rc;
  if (!b || !x || maxit < 1 || (precond != MT_PC_DILU && precond != MT_PC_NONE))
    return mt_fail(ctx, MT_E_ARG, "mt_qmr: bad arguments (precond is MT_PC_DILU or MT_PC_NONE)");
  MT_HIP_CHECK(ctx, hipSetDevice(ctx->device));
  if ((rc = ensure_transpose(ctx))) return rc;
  const Grid& g = ctx->hop.g;
  const size_t nv = size_t(g.vlen(1));
  QmrEngine E;
  E.ctx = ctx;
  E.n = g.nE;
  E.pc = precond == MT_PC_DILU;
  E.omega = omega;
  static const char* names[] = {"q_x", "q_b", "q_r", "q_vt", "q_y", "q_z", "q_wt", "q_v", "q_w",
                                "q_yt", "q_zt", "q_p", "q_q", "q_pt", "q_d", "q_s", "q_t"};
  Z* v[17];
  for (int k = 0; k < 17; ++k)
    if ((rc = ws(ctx, names[k], nv, &v[k]))) return rc;
  Z *X = v[0], *B = v[1], *R = v[2], *VT = v[3], *Y = v[4], *Zv = v[5], *WT = v[6], *V = v[7], *W = v[8];
  Z *YT = v[9], *ZT = v[10], *P = v[11], *Q = v[12], *PT = v[13], *D = v[14], *S = v[15], *T = v[16];
  if ((rc = ws(ctx, "q_part", size_t(3) * E.G, &E.part)) || (rc = ws(ctx, "q_sc", 1, &E.sc)) ||
      (rc = ws(ctx, "q_act", 1, &E.act)))
    return rc;
  SysScal hs;
  std::memset(&hs, 0, sizeof(hs));
  hs.status = ST_RUN;
  hs.fidx = 0;
  hs.freq_omega = omega;
  const int zero = 0;
  MT_HIP_CHECK(ctx, hipMemcpyAsync(E.sc, &hs, sizeof(hs), hipMemcpyHostToDevice, ctx->stream));
  MT_HIP_CHECK(ctx, hipMemcpyAsync(E.act, &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  if (E.pc) {
    C2<double>* dv = nullptr;
    if ((rc = build_dinv<double>(ctx, PCK_DILU, false, std::vector<double>{omega}, &dv))) return rc;
    E.dinv = reinterpret_cast<Z*>(dv);
    if ((rc = ws(ctx, "q_dinvh", nv, &E.dinvH))) return rc;
    k_q_conj<<<E.grid(), 256, 0, ctx->stream>>>(g.nE, E.dinv, E.dinvH);
  }
  if ((rc = upload_padded(ctx, 1, b, B, "io_tmp")) || (rc = upload_padded(ctx, 1, x, X, "io_tmp")))
    return rc;
  const unsigned gr = E.grid();
  for (int i = 0; i < maxit; ++i) rerr[i] = 0.0;
  // R = b - A x
  E.apply(false, X, R);
  k_q_axpy<<<gr, 256, 0, ctx->stream>>>(g.nE, Z{1.0, 0.0}, B, R, R);
  mt_zdouble dd[3];
  if ((rc = E.dots(2, B, B, R, R, nullptr, nullptr, dd))) return rc;
  const double bnorm = std::sqrt(dd[0]).real();
  double rnorm = std::sqrt(dd[1]).real();
  double rnormin = rnorm;
  if (std::isnan(std::abs(bnorm))) {
    *niter = 0;
    *failed = 1;
    return mt_fail(ctx, MT_E_NAN_RHS, "mt_qmr: the right-hand side contains NaN");
  }
  int iter = 1;
  rerr[0] = double(float(rnorm / bnorm));
  E.copy(R, VT);
  E.m1(VT, Y);
  E.copy(R, WT);
  E.m2h(WT, Zv, T);
  if ((rc = E.dots(2, Y, Y, Zv, Zv, nullptr, nullptr, dd))) return rc;
  mt_zdouble RHO = std::sqrt(dd[0]), PSI = std::sqrt(dd[1]);
  mt_zdouble GAMM = 1.0, ETA = -1.0, EPSIL = 0.0, THET = 0.0, THET1 = 0.0;
  int fl = 0;
  auto zc = [](mt_zdouble a) { return Z{a.real(), a.imag()}; };
  while (rerr[iter - 1] > tol && iter < maxit) {
    if (RHO == 0.0 || PSI == 0.0) { fl = 1; break; }
    const mt_zdouble rhoInv = (1.0 / RHO) * mt_zdouble(1.0, 0.0);
    const mt_zdouble psiInv = (1.0 / PSI) * mt_zdouble(1.0, 0.0);
    k_q_scale<<<gr, 256, 0, ctx->stream>>>(g.nE, zc(rhoInv), zc(psiInv), VT, WT, V, W, Y, Zv);
    if ((rc = E.dots(1, Zv, Y, nullptr, nullptr, nullptr, nullptr, dd))) return rc;
    const mt_zdouble DELTA = dd[0];
    if (DELTA == 0.0) { fl = 1; break; }
    E.m2(Y, YT);
    E.m1h(Zv, ZT);
    if (iter == 1) {
      E.copy(YT, P);
      E.copy(ZT, Q);
    } else {
      const mt_zdouble PDE = PSI * DELTA / EPSIL;
      const mt_zdouble RDE = RHO * std::conj(DELTA / EPSIL);
      k_q_axpy<<<gr, 256, 0, ctx->stream>>>(g.nE, zc(PDE), YT, P, P);
      k_q_axpy<<<gr, 256, 0, ctx->stream>>>(g.nE, zc(RDE), ZT, Q, Q);
    }
    E.apply(false, P, PT);
    if ((rc = E.dots(1, Q, PT, nullptr, nullptr, nullptr, nullptr, dd))) return rc;
    EPSIL = dd[0];
    if (EPSIL == 0.0) { fl = 1; break; }
    const mt_zdouble BETA = EPSIL / DELTA;
    if (BETA == 0.0) { fl = 1; break; }
    k_q_axpy<<<gr, 256, 0, ctx->stream>>>(g.nE, zc(BETA), PT, V, VT);
    const mt_zdouble RHO1 = RHO;
    E.m1(VT, Y);
    E.apply(true, Q, WT);
    k_q_axpy<<<gr, 256, 0, ctx->stream>>>(g.nE, zc(std::conj(BETA)), WT, W, WT);
    E.m2h(WT, Zv, T);
    if ((rc = E.dots(2, Y, Y, Zv, Zv, nullptr, nullptr, dd))) return rc;
    RHO = std::sqrt(dd[0]);
    PSI = std::sqrt(dd[1]);
    if (iter > 1) THET1 = THET;
    THET = RHO / (GAMM * std::abs(BETA));
    const mt_zdouble GAMM1 = GAMM;
    GAMM = 1.0 / std::sqrt(1.0 + THET * THET);
    if (GAMM == 0.0) break;
    ETA = -(ETA * RHO1 * GAMM * GAMM / (BETA * GAMM1 * GAMM1));
    const mt_zdouble TM2 = THET1 * THET1 * GAMM * GAMM;
    k_q_update<<<gr, 256, 0, ctx->stream>>>(g.nE, iter == 1 ? 1 : 0, zc(ETA), zc(TM2), P, PT, D, S, X, R);
    if ((rc = E.dots(1, R, R, nullptr, nullptr, nullptr, nullptr, dd))) return rc;
    rnorm = std::sqrt(dd[0]).real();
    if (rnorm < rnormin) rnormin = rnorm;
    iter = iter + 1;
    rerr[iter - 1] = double(float(rnorm / bnorm));
  }
  *niter = iter;
  rerr[iter - 1] = double(float(rnormin / bnorm));
  *failed = fl;
  if ((rc = download_padded(ctx, 1, X, x, "io_tmp"))) return rc;
  if (fl) return mt_fail(ctx, MT_E_BREAKDOWN, "mt_qmr: breakdown (rho, psi, delta, epsilon or beta = 0)");
  if (iter >= maxit) return mt_fail(ctx, MT_E_MAXIT, "mt_qmr: not converged within maxit");
  return MT_OK;
}
