// extern "C" layer of the MI355X friction solve path (include/sobogus_hip.h).
// Problem set-up (tiling, transposition, W) is driven from here, by default on the
// device (sb_symbolic.hip, sb_setup.hip), else on the host (sb_host.cpp); the
// coloring runs on the host; every solve, product and residual runs as HIP
// kernels on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <functional>
#include <queue>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/sobogus_hip.h"
#include "sb_host.hpp"
#include "sb_problem.hpp"
#include "sb_spmv.hpp"
#include "sb_symbolic.hpp"

using sb::DevProblem;
using sb::TileMat;

namespace {

constexpr int kThreads = 256;     // default block size of the persistent kernels
constexpr int kMaxThreads = 512;  // __launch_bounds__ of the persistent kernels
constexpr size_t kSyncWords = 2 * sb::kSyncGroups * sb::kSyncStride + sb::kSyncStride;

#define SB_HIP(expr)                                                                      \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      last_error_string() = std::string(hipGetErrorString(e_)) + " (" #expr ", line " +   \
                            std::to_string(__LINE__) + ")";                               \
      return e_ == hipErrorOutOfMemory ? SB_ERR_OOM : SB_ERR_HIP;                         \
    }                                                                                     \
  } while (0)

std::string& last_error_string() {
  static thread_local std::string s;
  return s;
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    n = count;
    if (count == 0) return hipSuccess;
    return hipMalloc((void**)&p, sizeof(T) * count);
  }
  hipError_t upload(const T* src, size_t count, hipStream_t s) {
    hipError_t e = alloc(count);
    if (e != hipSuccess || count == 0) return e;
    return hipMemcpyAsync(p, src, sizeof(T) * count, hipMemcpyHostToDevice, s);
  }
  hipError_t upload(const std::vector<T>& v, hipStream_t s) { return upload(v.data(), v.size(), s); }
  void swap(DBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

std::vector<float> to_float(const std::vector<double>& v) {
  std::vector<float> f(v.size());
  for (size_t k = 0; k < v.size(); ++k) f[k] = (float)v[k];
  return f;
}

struct DevTiles {
  DBuf<int> tstart, brow, bcol, tlong, long_row, long_t0, long_t1, rowptr;
  DBuf<double> vals;
  DBuf<float> vals32;  // FP32 copy for the mixed-precision solvers
  sbhost::Tiles host;  // sizes, and the index arrays while hostArrays (values dropped after upload)
  bool pairs = false;  // every row exactly two blocks, in order (the contact rows of H)
  int nlong = 0;       // long rows
  bool hostArrays = true;  // false after a build on the device, until download()
  hipError_t upload(sbhost::Tiles&& t, hipStream_t s) {
    host = std::move(t);
    nlong = (int)host.long_row.size();
    hostArrays = true;
    pairs = host.nblocks == 2 * host.nrows && host.long_row.empty() && host.nblocks % 2 == 0;
    for (int i = 0; pairs && i <= host.nrows; ++i) pairs = host.rowptr[i] == 2 * i;
    hipError_t e;
    if ((e = tstart.upload(host.tstart, s)) != hipSuccess) return e;
    if ((e = brow.upload(host.brow, s)) != hipSuccess) return e;
    if ((e = bcol.upload(host.bcol, s)) != hipSuccess) return e;
    if ((e = tlong.upload(host.tlong, s)) != hipSuccess) return e;
    if ((e = long_row.upload(host.long_row, s)) != hipSuccess) return e;
    if ((e = long_t0.upload(host.long_t0, s)) != hipSuccess) return e;
    if ((e = long_t1.upload(host.long_t1, s)) != hipSuccess) return e;
    if ((e = rowptr.upload(host.rowptr, s)) != hipSuccess) return e;
    if ((e = vals.upload(host.vals, s)) != hipSuccess) return e;
    {
      const std::vector<float> f = to_float(host.vals);
      if ((e = vals32.upload(f, s)) != hipSuccess) return e;
      if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    }
    e = hipStreamSynchronize(s);
    std::vector<double>().swap(host.vals);
    return e;
  }
  template <typename T>
  TileMat<T> viewT(const T* v) const {
    TileMat<T> m;
    m.nrows = host.nrows;
    m.ncols = host.ncols;
    m.nblocks = host.nblocks;
    m.ntiles = host.ntiles;
    m.scan_steps = host.scan_steps;
    m.tstart = tstart.p;
    m.brow = brow.p;
    m.bcol = bcol.p;
    m.tlong = tlong.p;
    m.vals = v;
    m.pairs = pairs ? 1 : 0;
    m.nlong = nlong;
    m.long_row = long_row.p;
    m.long_t0 = long_t0.p;
    m.long_t1 = long_t1.p;
    return m;
  }
  TileMat<double> view() const { return viewT<double>(vals.p); }
  TileMat<float> view32() const { return viewT<float>(vals32.p); }
  size_t carryNeed(int br) const { return (size_t)host.ntiles * br; }
  // host copies of the index arrays of a layout built on the device
  hipError_t download(hipStream_t s) {
    if (hostArrays) return hipSuccess;
    auto get = [&](std::vector<int>& v, const DBuf<int>& d, size_t n) -> hipError_t {
      v.resize(n);
      return n ? hipMemcpyAsync(v.data(), d.p, sizeof(int) * n, hipMemcpyDeviceToHost, s) : hipSuccess;
    };
    hipError_t e;
    if ((e = get(host.tstart, tstart, (size_t)host.ntiles + 1)) != hipSuccess) return e;
    if ((e = get(host.brow, brow, (size_t)host.nblocks)) != hipSuccess) return e;
    if ((e = get(host.bcol, bcol, (size_t)host.nblocks)) != hipSuccess) return e;
    if ((e = get(host.tlong, tlong, (size_t)host.ntiles)) != hipSuccess) return e;
    if ((e = get(host.long_row, long_row, (size_t)nlong)) != hipSuccess) return e;
    if ((e = get(host.long_t0, long_t0, (size_t)nlong)) != hipSuccess) return e;
    if ((e = get(host.long_t1, long_t1, (size_t)nlong)) != hipSuccess) return e;
    if ((e = get(host.rowptr, rowptr, (size_t)host.nrows + 1)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    hostArrays = true;
    return hipSuccess;
  }
};

int cub_tmp(DBuf<char>& tmp, size_t bytes) {
  if (tmp.n < bytes) SB_HIP(tmp.alloc(bytes));
  return SB_OK;
}
int read_int(const int* d, int* h, hipStream_t s) {
  SB_HIP(hipMemcpyAsync(h, d, sizeof(int), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}
// out[0] = 0, out[1..n] = inclusive prefix sums of in (n + 1 entries)
int scan_exclusive(const int* in, DBuf<int>& out, int n, DBuf<char>& tmp, hipStream_t s) {
  SB_HIP(out.alloc((size_t)n + 1));
  SB_HIP(hipMemsetAsync(out.p, 0, sizeof(int), s));
  if (n == 0) return SB_OK;
  size_t bytes = 0;
  SB_HIP(sb::scan_inclusive_int(in, out.p + 1, n, nullptr, bytes, s));
  const int rc = cub_tmp(tmp, bytes);
  if (rc) return rc;
  SB_HIP(sb::scan_inclusive_int(in, out.p + 1, n, tmp.p, bytes, s));
  return SB_OK;
}

// The tile layout of a BSR pattern (rowptr rp, columns ci: device arrays) built on
// the device, element for element the layout sbhost::buildTiles builds on the host;
// src receives each tile block's source block (-1: an empty row's zero block). The
// values are allocated, not filled.
int build_tiles_device(DevTiles& T, int nrows, int ncols, int br, int bc, const int* rp, const int* ci,
                       DBuf<int>& src, DBuf<char>& tmp, hipStream_t s) {
  int rc;
  DBuf<int> tlen, next, ntl, nlg, tbase, lbase, entry, mx;
  DBuf<unsigned char> isStart, exits;
  SB_HIP(tlen.alloc(nrows));
  SB_HIP(sb::launch_row_len(nrows, rp, tlen.p, s));
  if ((rc = scan_exclusive(tlen.p, T.rowptr, nrows, tmp, s))) return rc;
  SB_HIP(mx.alloc(1));
  {
    size_t bytes = 0;
    SB_HIP(sb::reduce_max_int(tlen.p, mx.p, nrows, nullptr, bytes, s));
    if ((rc = cub_tmp(tmp, bytes))) return rc;
    SB_HIP(sb::reduce_max_int(tlen.p, mx.p, nrows, tmp.p, bytes, s));
  }
  int nblocks = 0, maxLen = 0;
  SB_HIP(hipMemcpyAsync(&maxLen, mx.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if ((rc = read_int(T.rowptr.p + nrows, &nblocks, s))) return rc;
  maxLen = std::min(maxLen, 64);
  SB_HIP(T.brow.alloc(nblocks));
  SB_HIP(T.bcol.alloc(nblocks));
  SB_HIP(src.alloc(nblocks));
  SB_HIP(sb::launch_tile_expand(nrows, nblocks, T.rowptr.p, rp, ci, T.brow.p, T.bcol.p, src.p, s));
  SB_HIP(next.alloc(nrows));
  SB_HIP(sb::launch_tile_next(nrows, T.rowptr.p, tlen.p, next.p, s));
  const int C = sb::chain_chunk_rows(nrows), nchunks = (nrows + C - 1) / C;
  SB_HIP(exits.alloc((size_t)nchunks * 64));
  SB_HIP(entry.alloc(nchunks));
  SB_HIP(isStart.alloc(nrows));
  SB_HIP(sb::launch_tile_chain(nrows, next.p, exits.p, entry.p, isStart.p, s));
  SB_HIP(ntl.alloc(nrows));
  SB_HIP(nlg.alloc(nrows));
  SB_HIP(sb::launch_tile_counts(nrows, tlen.p, isStart.p, ntl.p, nlg.p, s));
  if ((rc = scan_exclusive(ntl.p, tbase, nrows, tmp, s))) return rc;
  if ((rc = scan_exclusive(nlg.p, lbase, nrows, tmp, s))) return rc;
  int ntiles = 0, nlong = 0;
  SB_HIP(hipMemcpyAsync(&ntiles, tbase.p + nrows, sizeof(int), hipMemcpyDeviceToHost, s));
  if ((rc = read_int(lbase.p + nrows, &nlong, s))) return rc;
  SB_HIP(T.tstart.alloc((size_t)ntiles + 1));
  SB_HIP(T.tlong.alloc(ntiles));
  SB_HIP(T.long_row.alloc(nlong));
  SB_HIP(T.long_t0.alloc(nlong));
  SB_HIP(T.long_t1.alloc(nlong));
  SB_HIP(sb::launch_tile_fill(nrows, nblocks, ntiles, T.rowptr.p, tlen.p, next.p, isStart.p, tbase.p, lbase.p,
                              T.tstart.p, T.tlong.p, T.long_row.p, T.long_t0.p, T.long_t1.p, s));
  SB_HIP(T.vals.alloc((size_t)nblocks * br * bc));
  SB_HIP(T.vals32.alloc((size_t)nblocks * br * bc));
  SB_HIP(hipStreamSynchronize(s));  // the local buffers are freed on return
  T.host = sbhost::Tiles();
  T.host.nrows = nrows;
  T.host.ncols = ncols;
  T.host.br = br;
  T.host.bc = bc;
  T.host.nblocks = nblocks;
  T.host.ntiles = ntiles;
  T.host.scan_steps = sb::ceil_log2_host(std::max(1, maxLen));
  T.nlong = nlong;
  T.pairs = nblocks == 2 * nrows && maxLen == 2;
  T.hostArrays = false;
  return SB_OK;
}

}  // namespace

struct sb_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int ncu = 0;
  std::string name;
  int coop = 0;          // device supports cooperative launches
  bool coopLaunch = true; // launch persistent kernels cooperatively (SOBOGUS_LAUNCH=regular: plain launch)
  // persistent launches since creation: cooperative (residency guaranteed) and plain
  // (refused cooperative launch, SOBOGUS_LAUNCH=regular, or a batch graph); lastPlain:
  // the last solve launched at least one kernel plainly
  int64_t coopLaunches = 0, plainLaunches = 0;
  bool lastPlain = false;
  DBuf<unsigned> sync;  // group counters, top counter, release words
  DBuf<int> error;
  DBuf<double> partials, out;
  int maxGrid = 0;  // capacity of partials (blocks)
  double phase_us[4] = {0, 0, 0, 0};  // last PG solve: operator work/wait, update work/wait
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evm = nullptr;  // evm: between the stages of a mixed solve
};

struct sb_bsr {
  int nbr = 0, nbc = 0, br = 0, bc = 0, flags = 0;
  DevTiles fwd, tr;  // product and transposed product (symmetric: both equal the expanded matrix)
  DBuf<double> carry, xbuf, ybuf;
  std::vector<int> lrp, lci;  // symmetric lower pattern (split-row multiply)
  DBuf<int> d_lrp, d_lci;
  DBuf<double> d_lvals;
};

struct sb_problem {
  sb_ctx* ctx = nullptr;
  int nc = 0, nb = 0;
  std::vector<int> hrp, hci;
  std::vector<double> hv, minv, b, mu, diag, scal;
  bool hasW = false;
  std::vector<int> wrp, wci;
  std::vector<double> wv;
  DevTiles H, HT, W;
  DBuf<double> d_minv, d_b, d_mu, d_diag, d_scal, d_ones;
  DBuf<float> d_minv32;
  // workspace (FP64 solvers and kernels)
  DBuf<double> X[4], MX[4], G[3], xbest, wbody, carry, rio, hist, y, terms, local, reg, ndx;
  DBuf<double> cad[5];  // Cadoux fixed point: original b, iterate, shift, operator image, velocities
  // device set-up (sb_problem_update_values): the H^T tiles' source blocks, the W tiles'
  // product terms, the BSR index of H, and the current H and Minv values (row-major);
  // hostStale: the host copies (hv, minv, b, mu, diag, scal, wv) lag the device
  std::vector<int> htsrc;
  bool devSetup = false, hostStale = false, devRawValid = false, wFromH = true;
  DBuf<int> d_hsrc, d_htsrc, d_wtptr, d_wterm, d_hrp, d_hci;
  // pattern built on the device (sb_problem_create_device): the host copies of the
  // pattern (hrp, hci, wrp, wci, the tiles' index arrays, htsrc) are downloaded by
  // the host paths that need them (hostPattern false until then); the lower
  // triangle of W and its tiles' source blocks stay on the device
  bool hostPattern = true;
  bool hostSetup = false;  // created with SB_PROBLEM_HOST_SETUP: pattern changes stay on the host
  bool dualOnly = false;   // created from W itself (sb_problem_create_from_w): no H, no Minv
  int nnz = 0;        // blocks of H
  int64_t nwl = 0;    // blocks of W's lower triangle (0 without W)
  DBuf<int> d_wrp, d_wci, d_wsrc;
  // assembly order keys: okey[c] is contact c's index when W was assembled (a contact
  // permutation keeps each W block's orientation); wtermOkey: the product terms of
  // the W tiles follow them (device set-up) rather than the indices (host set-up)
  DBuf<int> d_okey;
  bool wtermOkey = false;
  DBuf<double> d_hraw, d_mraw;
  // workspace of the FP32 (mixed-precision) projected-gradient solver
  DBuf<float> X32[4], MX32[4], G32[3], xbest32, wbody32, carry32;
  DBuf<unsigned char> skip;
  DBuf<int> colors;
  // matrix-free Gauss-Seidel: Minv_k H_ik^T per H block and per-contact block ranges,
  // rebuilt when the uploaded data changes (version)
  DBuf<double> dmt;
  DBuf<int> hrow;
  int version = 0, pgsVersion = -1;
  // matrix-free projected gradient: factored tiles F = H L and F^T (Minv = L L^T), FP64
  // and FP32, current while fVersion == version; factored false: some Minv_k is not
  // positive definite and the solvers apply Minv between the passes
  DBuf<double> fH, fHT;
  DBuf<float> fH32, fHT32;
  DBuf<int> ferr;
  DBuf<double> fL;  // the Cholesky factors L_k (element-major), kept so a refresh allocates nothing
  int fVersion = -1;
  bool factored = false;
  bool useFactored() const { return fVersion == version && factored; }
  DBuf<double> out;  // [2][16] solver results (FP32 stage, FP64 stage)
  // derivatives, Coulomb law: De Saxce directions recorded by an unroll (branch_mode 1) for
  // its replay (branch_mode 2): [adjDirK][nc][2]
  DBuf<double> adjDir;
  int adjDirK = 0;
  DevProblem<double> view(bool withScaling) const {
    DevProblem<double> v;
    v.nc = nc;
    v.nb = nb;
    v.H = H.view();
    v.HT = HT.view();
    v.W = W.view();
    v.minv = d_minv.p;
    v.b = d_b.p;
    v.mu = d_mu.p;
    v.diag = d_diag.p;
    v.scal = withScaling ? d_scal.p : d_ones.p;
    v.hasW = hasW;
    return v;
  }
  DevProblem<float> view32(bool withScaling) const {
    DevProblem<float> v;
    v.nc = nc;
    v.nb = nb;
    v.H = H.view32();
    v.HT = HT.view32();
    v.W = W.view32();
    v.minv = d_minv32.p;
    v.b = d_b.p;
    v.mu = d_mu.p;
    v.diag = d_diag.p;
    v.scal = withScaling ? d_scal.p : d_ones.p;
    v.hasW = hasW;
    return v;
  }
  // Solver arguments of one problem for storage type T; x_io, hist and out are
  // set by the caller.
  void fill(sb::PGProblem<double>& pb, bool withScaling) const {
    pb.P = view(withScaling);
    if (useFactored()) {
      pb.P.H = H.viewT<double>(fH.p);
      pb.P.HT = HT.viewT<double>(fHT.p);
      pb.P.minv_folded = true;
    }
    for (int k = 0; k < 4; ++k) {
      pb.X[k] = X[k].p;
      pb.MX[k] = MX[k].p;
    }
    for (int k = 0; k < 3; ++k) pb.G[k] = G[k].p;
    pb.xbest = xbest.p;
    pb.wbody = wbody.p;
    pb.carry = carry.p;
  }
  void fill(sb::PGProblem<float>& pb, bool withScaling) const {
    pb.P = view32(withScaling);
    if (useFactored()) {
      pb.P.H = H.viewT<float>(fH32.p);
      pb.P.HT = HT.viewT<float>(fHT32.p);
      pb.P.minv_folded = true;
    }
    for (int k = 0; k < 4; ++k) {
      pb.X[k] = X32[k].p;
      pb.MX[k] = MX32[k].p;
    }
    for (int k = 0; k < 3; ++k) pb.G[k] = G32[k].p;
    pb.xbest = xbest32.p;
    pb.wbody = wbody32.p;
    pb.carry = carry32.p;
  }
};

namespace {

int check_ctx(sb_ctx* ctx) {
  if (!ctx) return SB_ERR_INVALID;
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) {
    last_error_string() = hipGetErrorString(e);
    return SB_ERR_HIP;
  }
  return SB_OK;
}

int ensure_grid_scratch(sb_ctx* ctx, int blocks) {
  if (blocks <= ctx->maxGrid) return SB_OK;
  SB_HIP(ctx->partials.alloc((size_t)2 * blocks * sb::kMaxPartials));
  ctx->maxGrid = blocks;
  return SB_OK;
}

// Element-major 6x6 inverse-mass blocks.
std::vector<double> minv_element_major(int nb, const std::vector<double>& minv) {
  std::vector<double> em((size_t)36 * nb);
  for (int k = 0; k < nb; ++k)
    for (int e = 0; e < 36; ++e) em[(size_t)e * nb + k] = minv[(size_t)k * 36 + e];
  return em;
}

// Values that sb_problem_update may replace: W (when assembled), Minv, b, mu,
// the diagonal blocks and the row scalings.
int upload_values(sb_problem* P) {
  hipStream_t s = P->ctx->stream;
  const int nc = P->nc, nb = P->nb;
  ++P->version;
  P->devRawValid = false;  // the host copies are the current values
  if (P->hasW) {
    std::vector<int> frp, fci;
    std::vector<double> fv;
    sbhost::expandSymmetric(nc, 3, P->wrp.data(), P->wci.data(), P->wv.data(), frp, fci, fv);
    SB_HIP(P->W.upload(sbhost::buildTiles(nc, nc, 3, 3, frp.data(), fci.data(), fv.data()), s));
  }
  {
    const std::vector<double> em = minv_element_major(nb, P->minv);
    SB_HIP(P->d_minv.upload(em, s));
    const std::vector<float> em32 = to_float(em);
    SB_HIP(P->d_minv32.upload(em32, s));
    SB_HIP(hipStreamSynchronize(s));
  }
  SB_HIP(P->d_b.upload(P->b, s));
  SB_HIP(P->d_mu.upload(P->mu, s));
  // diagonal blocks: from the assembled W when present, else from H Minv H^T
  if (P->hasW) {
    P->diag.assign((size_t)9 * nc, 0.0);
    for (int i = 0; i < nc; ++i)
      for (int p = P->wrp[i]; p < P->wrp[i + 1]; ++p)
        if (P->wci[p] == i) std::memcpy(&P->diag[(size_t)9 * i], &P->wv[(size_t)9 * p], 9 * sizeof(double));
  } else {
    sbhost::diagBlocks(nc, P->hrp.data(), P->hci.data(), P->hv.data(), P->minv.data(), P->diag);
  }
  P->scal.resize(nc);
  for (int i = 0; i < nc; ++i) {
    const double* d = &P->diag[(size_t)9 * i];
    P->scal[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3.0);  // estimate_block_scaling
  }
  SB_HIP(P->d_diag.upload(P->diag, s));
  SB_HIP(P->d_scal.upload(P->scal, s));
  std::vector<double> ones(nc, 1.0);
  SB_HIP(P->d_ones.upload(ones, s));
  return SB_OK;
}

// Solver workspace of a problem (FP64 and FP32 iterates, products, carries).
int alloc_workspace(sb_problem* P) {
  const int nc = P->nc, nb = P->nb;
  const size_t n3 = 3 * (size_t)nc;
  for (int k = 0; k < 4; ++k) {
    SB_HIP(P->X[k].alloc(n3));
    SB_HIP(P->MX[k].alloc(n3));
  }
  for (int k = 0; k < 3; ++k) SB_HIP(P->G[k].alloc(n3));
  SB_HIP(P->xbest.alloc(n3));
  SB_HIP(P->rio.alloc(n3));
  SB_HIP(P->y.alloc(n3));
  SB_HIP(P->terms.alloc(nc));
  SB_HIP(P->wbody.alloc(6 * (size_t)nb));
  size_t carry = std::max({P->H.carryNeed(3), P->HT.carryNeed(6), P->W.carryNeed(3), (size_t)8});
  SB_HIP(P->carry.alloc(carry));
  for (int k = 0; k < 4; ++k) {
    SB_HIP(P->X32[k].alloc(n3));
    SB_HIP(P->MX32[k].alloc(n3));
  }
  for (int k = 0; k < 3; ++k) SB_HIP(P->G32[k].alloc(n3));
  SB_HIP(P->xbest32.alloc(n3));
  SB_HIP(P->wbody32.alloc(6 * (size_t)nb));
  SB_HIP(P->carry32.alloc(carry));
  SB_HIP(P->out.alloc(32));
  return SB_OK;
}

int upload_problem(sb_problem* P) {
  hipStream_t s = P->ctx->stream;
  const int nc = P->nc, nb = P->nb;
  ++P->version;
  P->devSetup = false;  // index maps follow the (possibly new) pattern
  P->hostPattern = true;
  P->nnz = P->hrp[nc];
  P->nwl = P->hasW ? (int64_t)P->wci.size() : 0;
  // H contact-major, H^T body-major
  SB_HIP(P->H.upload(sbhost::buildTiles(nc, nb, 3, 6, P->hrp.data(), P->hci.data(), P->hv.data()), s));
  {
    std::vector<int> trp, tci, tsrc;
    std::vector<double> tv;
    sbhost::transposeBSR(nc, nb, 3, 6, P->hrp.data(), P->hci.data(), P->hv.data(), trp, tci, tv, &tsrc);
    sbhost::Tiles ht = sbhost::buildTiles(nb, nc, 6, 3, trp.data(), tci.data(), tv.data());
    P->htsrc.assign(ht.src.size(), -1);
    for (size_t q = 0; q < ht.src.size(); ++q)
      if (ht.src[q] >= 0) P->htsrc[q] = tsrc[ht.src[q]];
    SB_HIP(P->HT.upload(std::move(ht), s));
  }
  int rc = upload_values(P);
  if (rc) return rc;
  if ((rc = alloc_workspace(P))) return rc;
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}

int copy_in(sb_ctx* ctx, double* dst, const double* src, size_t n, int on_device) {
  SB_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                        ctx->stream));
  return SB_OK;
}
int copy_out(sb_ctx* ctx, double* dst, const double* src, size_t n, int on_device) {
  SB_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                        ctx->stream));
  return SB_OK;
}

// Persistent grid: at most the co-resident blocks (cooperative launch), by
// default one block per CU (the barrier cost grows with the number of blocks),
// and no more blocks than there are tiles of work.
int grid_for(sb_ctx* ctx, const void* func, int threads, int requested, int work_waves) {
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, func, threads, 0) != hipSuccess || perCU < 1) perCU = 1;
  const int maxCoop = perCU * ctx->ncu;
  const int wpb = threads / 64;
  int want = requested > 0 ? requested : std::min(ctx->ncu, (work_waves + wpb - 1) / wpb);
  want = std::max(1, std::min(want, maxCoop));
  return want;
}

int block_threads_for(int requested, int dflt) {
  if (requested == 0) return dflt;
  return (requested == 256 || requested == 512) ? requested : -1;
}

int read_status(sb_ctx* ctx, double* out, int nout) {
  SB_HIP(hipMemcpyAsync(out, ctx->out.p, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

// Persistent kernels need every block resident at once (their grid barrier is
// a software barrier; no hardware wave-sync resource is used). A cooperative
// launch guarantees residency. The runtime refuses it with hipErrorIllegalState
// while another process holds the device's cooperative queue; the same grid,
// sized to the occupancy limit, is then launched plainly - on an otherwise idle
// device it is resident, and if it is not, the bounded barrier wait ends the
// solve with SB_ERR_BARRIER_TIMEOUT instead of hanging.
hipError_t launch_persistent(sb_ctx* ctx, const void* fn, int grid, int threads, void** args, hipStream_t s) {
  if (ctx->coopLaunch) {
    const hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(threads), args, 0, s);
    if (e != hipErrorIllegalState) {
      if (e == hipSuccess) ++ctx->coopLaunches;
      return e;
    }
    (void)hipGetLastError();
  }
  const hipError_t e = hipLaunchKernel(fn, dim3(grid), dim3(threads), args, 0, s);
  if (e == hipSuccess) {
    ++ctx->plainLaunches;
    ctx->lastPlain = true;
  }
  return e;
}

// Device state of a persistent launch: barrier words (kSyncWords), the sticky error
// word, reduction partials for up to maxBlocks blocks, and the solver results.
struct GridState {
  unsigned* sync = nullptr;
  int* error = nullptr;
  double* partials = nullptr;
  int maxBlocks = 0;
  double* out = nullptr;  // [32]: two stages' results (single-problem PG solves)
};

// Fixed-size device array for sb_pg_solve_noalloc: the grid state of its persistent
// launches (barrier words, error word, reduction partials of up to kNoAllocBlocks
// workgroups, the two stages' results). Every GPU holds its own copy until the process
// ends; host code reaches it only through the per-device address from
// hipGetSymbolAddress (the array's name in host code is not a device address).
constexpr int kNoAllocBlocks = 2048;
constexpr size_t kNoAllocSyncBytes = ((kSyncWords * sizeof(unsigned) + 255) / 256) * 256;
constexpr size_t kNoAllocErrorOff = kNoAllocSyncBytes;
constexpr size_t kNoAllocOutOff = kNoAllocErrorOff + 256;          // 32 doubles: 256 bytes
constexpr size_t kNoAllocPartialsOff = kNoAllocOutOff + 32 * sizeof(double);
constexpr size_t kNoAllocBytes = kNoAllocPartialsOff + (size_t)2 * kNoAllocBlocks * sb::kMaxPartials * sizeof(double);
static_assert(kNoAllocPartialsOff % 256 == 0, "scratch regions stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratchMutex;
std::vector<void*> g_scratchPtr;  // per device: address of scratch (null: not resolved)

// Resolves this device's scratch address once; the device must be current.
int resolve_scratch(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratchMutex);
  if ((int)g_scratchPtr.size() <= device) g_scratchPtr.resize(device + 1, nullptr);
  if (!g_scratchPtr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      last_error_string() = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return SB_ERR_HIP;
    }
    g_scratchPtr[device] = ptr;
  }
  *out = g_scratchPtr[device];
  return SB_OK;
}
sb::GridSync make_sync(const GridState& st, int blocks) {
  sb::GridSync g;
  g.group = st.sync;
  g.release = st.sync + sb::kSyncGroups * sb::kSyncStride;
  g.top = st.sync + 2 * sb::kSyncGroups * sb::kSyncStride;
  g.error = st.error;
  g.partials = st.partials;
  g.nblocks = (unsigned)blocks;
  g.timeout_ticks = 100ull * 1000 * 1000 * 20;  // 20 s at 100 MHz
  return g;
}
sb::GridSync make_sync(sb_ctx* ctx, int blocks) {
  GridState st;
  st.sync = ctx->sync.p;
  st.error = ctx->error.p;
  st.partials = ctx->partials.p;
  st.maxBlocks = ctx->maxGrid;
  return make_sync(st, blocks);
}

int check_pg_options(const sb_pg_options* o, bool hasW, bool batch = false) {
  if (o->variant < SB_PG_STANDARD || o->variant > SB_PG_SPG) return SB_ERR_INVALID;
  // the batch kernel runs the APGD / Descent / SPG body only
  if (batch && (o->variant == SB_PG_STANDARD || o->variant == SB_PG_CONJUGATED)) return SB_ERR_UNSUPPORTED;
  if (o->op != SB_OP_ASSEMBLED && o->op != SB_OP_MATRIX_FREE) return SB_ERR_INVALID;
  if (o->op == SB_OP_ASSEMBLED && !hasW) return SB_ERR_INVALID;
  if (o->max_iters < 1 || o->ls_iters < 1 || !(o->tol >= 0)) return SB_ERR_INVALID;
  if (o->precision != SB_PRECISION_FP64 && o->precision != SB_PRECISION_MIXED) return SB_ERR_INVALID;
  if (o->precision == SB_PRECISION_MIXED && !(o->mixed_tol_factor >= 0 && o->mixed_tol_factor <= 1))
    return SB_ERR_INVALID;
  if (o->precision == SB_PRECISION_MIXED && o->mixed_stall_iters < 0) return SB_ERR_INVALID;
  return SB_OK;
}

// A problem created from W alone has no matrix-free operator.
int check_pg_problem(const sb_pg_options* o, const sb_problem* P) {
  if (o->op == SB_OP_MATRIX_FREE && P->dualOnly) return SB_ERR_UNSUPPORTED;
  return check_pg_options(o, P->hasW);
}

sb::PGParams make_params(const sb_pg_options* o, bool mf) {
  sb::PGParams p;
  p.variant = o->variant;
  p.coulomb = o->law == SB_LAW_COULOMB;
  p.inf_norm = o->inf_norm;
  const bool useScal = (o->scaling == SB_SCALING_DIAGONAL) || (o->scaling == SB_SCALING_REFERENCE && !mf);
  p.use_scaling = useScal ? 1 : 0;
  p.tol = o->tol;
  p.max_iters = o->max_iters;
  p.ls_iters = o->ls_iters;
  p.ls_opt = o->ls_optimistic;
  p.ls_pess = o->ls_pessimistic;
  p.ls_armijo = o->ls_armijo;
  p.stall_iters = 0;
  return p;
}

// Single-problem matrix-free APGD / Descent / SPG solves form each line-search
// candidate in a phase of its own (pg_kernel_cand) from this many contacts on
constexpr int kCandPhaseContacts = 200000;

template <typename T>
const void* pg_grid_fn(bool mf, bool candPhase, int variant) {
  if (variant == SB_PG_STANDARD || variant == SB_PG_CONJUGATED)
    return mf ? (const void*)&sb::pg_kernel_sc<true, T> : (const void*)&sb::pg_kernel_sc<false, T>;
  if (mf && candPhase) return (const void*)&sb::pg_kernel_cand<T>;
  return mf ? (const void*)&sb::pg_kernel<true, T> : (const void*)&sb::pg_kernel<false, T>;
}

// Launches one persistent single-problem solver kernel (grid barrier state reset first)
// with the context's grid state (grown on demand), or with a fixed one (fixed != null:
// a grid larger than its capacity is refused, nothing is allocated).
template <typename Args>
int launch_pg_grid(sb_ctx* ctx, const GridState* fixed, const void* fn, Args* a, int threads, int requested,
                   int workWaves) {
  const int grid = grid_for(ctx, fn, threads, requested, workWaves);
  GridState st;
  if (fixed) {
    if (grid > fixed->maxBlocks) {
      last_error_string() = "persistent grid of " + std::to_string(grid) + " workgroups exceeds the fixed grid state (" +
                            std::to_string(fixed->maxBlocks) + ")";
      return SB_ERR_LIMIT;
    }
    st = *fixed;
  } else {
    int rc = ensure_grid_scratch(ctx, grid);
    if (rc) return rc;
    st.sync = ctx->sync.p;
    st.error = ctx->error.p;
    st.partials = ctx->partials.p;
    st.maxBlocks = ctx->maxGrid;
  }
  a->gs = make_sync(st, grid);
  SB_HIP(sb::launch_reset_sync(st.sync, (int)kSyncWords, st.error, ctx->stream));
  void* args[] = {a};
  SB_HIP(launch_persistent(ctx, fn, grid, threads, args, ctx->stream));
  return SB_OK;
}

// smallest eigenvalue of a symmetric 3x3 block (Jacobi rotations)
double min_eig_sym3(const double* A) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (A[3 * i + j] + A[3 * j + i]);
  for (int sweep = 0; sweep < 50; ++sweep) {
    if (a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2] < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0) continue;
        const double th = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = std::copysign(1.0, th) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  return std::min(a[0][0], std::min(a[1][1], a[2][2]));
}

__global__ void k_split_row(const int* lrp, const int* lci, const double* lv, int n, int row, const double* x,
                            double* res) {
  // one thread: res += sum_{j != row} W_{row,j} x_j over the lower pattern (and its transpose)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double acc[3] = {res[0], res[1], res[2]};
  for (int i = 0; i < n; ++i)
    for (int p = lrp[i]; p < lrp[i + 1]; ++p) {
      const int j = lci[p];
      const double* B = lv + (size_t)9 * p;
      if (i == row && j < row) {
        for (int r = 0; r < 3; ++r) acc[r] += B[3 * r] * x[3 * j] + B[3 * r + 1] * x[3 * j + 1] + B[3 * r + 2] * x[3 * j + 2];
      } else if (j == row && i > row) {
        for (int r = 0; r < 3; ++r) acc[r] += B[r] * x[3 * i] + B[3 + r] * x[3 * i + 1] + B[6 + r] * x[3 * i + 2];
      }
    }
  res[0] = acc[0];
  res[1] = acc[1];
  res[2] = acc[2];
}

}  // namespace

namespace {
// Host copies of a pattern built on the device: H's row pointer and columns, the
// tiles' index arrays and source blocks, and W's lower-triangle pattern.
int sync_pattern(sb_problem* P) {
  if (P->hostPattern) return SB_OK;
  hipStream_t s = P->ctx->stream;
  const int nc = P->nc;
  P->hrp.resize((size_t)nc + 1);
  P->hci.resize((size_t)P->nnz);
  SB_HIP(hipMemcpyAsync(P->hrp.data(), P->d_hrp.p, sizeof(int) * P->hrp.size(), hipMemcpyDeviceToHost, s));
  if (P->nnz)
    SB_HIP(hipMemcpyAsync(P->hci.data(), P->d_hci.p, sizeof(int) * P->hci.size(), hipMemcpyDeviceToHost, s));
  SB_HIP(P->H.download(s));
  SB_HIP(P->HT.download(s));
  P->H.host.src.resize((size_t)P->H.host.nblocks);
  P->htsrc.resize((size_t)P->HT.host.nblocks);
  SB_HIP(hipMemcpyAsync(P->H.host.src.data(), P->d_hsrc.p, sizeof(int) * P->H.host.src.size(),
                        hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(P->htsrc.data(), P->d_htsrc.p, sizeof(int) * P->htsrc.size(), hipMemcpyDeviceToHost, s));
  if (P->hasW) {
    SB_HIP(P->W.download(s));
    P->W.host.src.resize((size_t)P->W.host.nblocks);
    SB_HIP(hipMemcpyAsync(P->W.host.src.data(), P->d_wsrc.p, sizeof(int) * P->W.host.src.size(),
                          hipMemcpyDeviceToHost, s));
    P->wrp.resize((size_t)nc + 1);
    P->wci.resize((size_t)P->nwl);
    SB_HIP(hipMemcpyAsync(P->wrp.data(), P->d_wrp.p, sizeof(int) * P->wrp.size(), hipMemcpyDeviceToHost, s));
    if (P->nwl)
      SB_HIP(hipMemcpyAsync(P->wci.data(), P->d_wci.p, sizeof(int) * P->wci.size(), hipMemcpyDeviceToHost, s));
  }
  SB_HIP(hipStreamSynchronize(s));
  P->hostPattern = true;
  return SB_OK;
}

// Host copies of the values after device-side updates (paths that rebuild or read
// them on the host: permutation, W export, host updates, GS local matrices, the
// matrix-free GS's per-block products).
int sync_host(sb_problem* P) {
  int rc = sync_pattern(P);
  if (rc) return rc;
  if (!P->hostStale) return SB_OK;
  hipStream_t s = P->ctx->stream;
  const int nc = P->nc, nb = P->nb;
  P->hv.resize((size_t)P->nnz * 18);
  P->minv.resize((size_t)nb * 36);
  P->b.resize((size_t)nc * 3);
  P->mu.resize((size_t)nc);
  SB_HIP(hipMemcpyAsync(P->hv.data(), P->d_hraw.p, sizeof(double) * P->hv.size(), hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(P->minv.data(), P->d_mraw.p, sizeof(double) * 36 * (size_t)nb, hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(P->b.data(), P->d_b.p, sizeof(double) * 3 * (size_t)nc, hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(P->mu.data(), P->d_mu.p, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost, s));
  P->diag.resize((size_t)9 * nc);
  P->scal.resize(nc);
  SB_HIP(hipMemcpyAsync(P->diag.data(), P->d_diag.p, sizeof(double) * 9 * (size_t)nc, hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(P->scal.data(), P->d_scal.p, sizeof(double) * (size_t)nc, hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  if (P->hasW && P->wFromH) {
    // the lower triangle of W as the solvers hold it: the first blocks of each row of
    // the two-triangle tiles (a row's lower part precedes its upper part)
    const sbhost::Tiles& T = P->W.host;
    const size_t nq = (size_t)T.nblocks;
    std::vector<double> tv(9 * nq);
    SB_HIP(hipMemcpyAsync(tv.data(), P->W.vals.p, sizeof(double) * tv.size(), hipMemcpyDeviceToHost, s));
    SB_HIP(hipStreamSynchronize(s));
    P->wv.resize((size_t)9 * P->wci.size());
    for (int i = 0; i < nc; ++i)
      for (int t = P->wrp[i]; t < P->wrp[i + 1]; ++t) {
        const size_t q = (size_t)T.rowptr[i] + (size_t)(t - P->wrp[i]);
        for (int e = 0; e < 9; ++e) P->wv[(size_t)9 * t + e] = tv[(size_t)e * nq + q];
      }
  }
  P->hostStale = false;
  return SB_OK;
}

// Index maps of the device set-up, built from the host structures of the current
// pattern: the source block of every H and H^T tile block, the BSR index of H, and
// the product terms (p_i, p_j) of every expanded W tile block.
int setup_maps(sb_problem* P) {
  if (P->devSetup) return SB_OK;
  hipStream_t s = P->ctx->stream;
  SB_HIP(P->d_hsrc.upload(P->H.host.src, s));
  SB_HIP(P->d_htsrc.upload(P->htsrc, s));
  SB_HIP(P->d_hrp.upload(P->hrp, s));
  SB_HIP(P->d_hci.upload(P->hci, s));
  if (P->hasW) {
    const sbhost::Tiles& T = P->W.host;
    std::vector<int> tptr(1, 0), term;
    for (int q = 0; q < T.nblocks; ++q) {
      if (T.src[q] >= 0) {
        // block (i, j): terms of the lower block (max, min) in the order the host assembly sums them
        const int i = std::max(T.brow[q], T.bcol[q]), j = std::min(T.brow[q], T.bcol[q]);
        for (int a = P->hrp[i]; a < P->hrp[i + 1]; ++a)
          for (int b = P->hrp[j]; b < P->hrp[j + 1]; ++b)
            if (P->hci[a] == P->hci[b]) {
              term.push_back(a);
              term.push_back(b);
            }
      }
      tptr.push_back((int)term.size() / 2);
    }
    if (term.empty()) term = {0, 0};
    SB_HIP(P->d_wtptr.upload(tptr, s));
    SB_HIP(P->d_wterm.upload(term, s));
  }
  SB_HIP(hipStreamSynchronize(s));
  P->devSetup = true;
  return SB_OK;
}

// The pattern-dependent set-up on the device, from the raw arrays in d_hrp, d_hci,
// d_hraw, d_mraw, d_b, d_mu (nc, nb, nnz and hasW set): tiles of H, the stable
// transpose and tiles of H^T, with W the pattern of its lower triangle, both
// triangles' tiles and the product terms of its blocks; then the numeric phase
// (tile values, Minv layout, diagonal blocks, scalings, W) and the workspace.
// Every array equals what the host set-up (upload_problem) builds.
int build_device(sb_problem* P) {
  hipStream_t s = P->ctx->stream;
  const int nc = P->nc, nb = P->nb, nnz = P->nnz;
  int rc;
  DBuf<char> tmp;
  ++P->version;
  if ((rc = build_tiles_device(P->H, nc, nb, 3, 6, P->d_hrp.p, P->d_hci.p, P->d_hsrc, tmp, s))) return rc;
  // H^T: H's blocks sorted by body, stably (within a body, in contact order)
  DBuf<unsigned> tkeys;
  DBuf<int> iota, tsrc, trp, tci, htsrcT;
  SB_HIP(tkeys.alloc(nnz));
  SB_HIP(iota.alloc(nnz));
  SB_HIP(tsrc.alloc(nnz));
  SB_HIP(trp.alloc((size_t)nb + 1));
  SB_HIP(tci.alloc(nnz));
  SB_HIP(sb::launch_iota(nnz, iota.p, s));
  if (nnz > 0) {
    const int bits = std::max(1, sb::ceil_log2_host(nb));
    size_t bytes = 0;
    const unsigned* keys = reinterpret_cast<const unsigned*>(P->d_hci.p);  // columns are non-negative
    SB_HIP(sb::sort_pairs_u32(keys, tkeys.p, iota.p, tsrc.p, nnz, bits, nullptr, bytes, s));
    if ((rc = cub_tmp(tmp, bytes))) return rc;
    SB_HIP(sb::sort_pairs_u32(keys, tkeys.p, iota.p, tsrc.p, nnz, bits, tmp.p, bytes, s));
  }
  SB_HIP(sb::launch_lower_bounds_u32(nb + 1, tkeys.p, nnz, trp.p, s));
  SB_HIP(sb::launch_rows_of(nnz, tsrc.p, nc, P->d_hrp.p, tci.p, s));
  if ((rc = build_tiles_device(P->HT, nb, nc, 6, 3, trp.p, tci.p, htsrcT, tmp, s))) return rc;
  SB_HIP(P->d_htsrc.alloc((size_t)P->HT.host.nblocks));
  SB_HIP(sb::launch_compose_src(P->HT.host.nblocks, htsrcT.p, tsrc.p, P->d_htsrc.p, s));
  P->nwl = 0;
  if (P->hasW) {
    // lower triangle: for contact i, the contacts j <= i sharing a body (sorted, unique)
    DBuf<int> cnt, cptr, flags, pos, uptr, flen, frp, fci, tcnt;
    DBuf<uint64_t> keys, skeys, ukeys, lk, slk;
    DBuf<long long> total;
    SB_HIP(cnt.alloc(nc));
    SB_HIP(sb::launch_w_cand_count(nc, P->d_hrp.p, P->d_hci.p, trp.p, tci.p, cnt.p, s));
    SB_HIP(total.alloc(1));
    {
      size_t bytes = 0;
      SB_HIP(sb::reduce_sum_int64(cnt.p, total.p, nc, nullptr, bytes, s));
      if ((rc = cub_tmp(tmp, bytes))) return rc;
      SB_HIP(sb::reduce_sum_int64(cnt.p, total.p, nc, tmp.p, bytes, s));
    }
    long long ncand = 0;
    SB_HIP(hipMemcpyAsync(&ncand, total.p, sizeof(long long), hipMemcpyDeviceToHost, s));
    SB_HIP(hipStreamSynchronize(s));
    if (ncand > (long long)std::numeric_limits<int>::max() / 2) {
      last_error_string() = "W = H Minv H^T: more than 2^30 contact pairs share a body";
      return SB_ERR_LIMIT;
    }
    const int nk = (int)ncand;
    if ((rc = scan_exclusive(cnt.p, cptr, nc, tmp, s))) return rc;
    SB_HIP(keys.alloc(nk));
    SB_HIP(skeys.alloc(nk));
    SB_HIP(sb::launch_w_cand_fill(nc, P->d_hrp.p, P->d_hci.p, trp.p, tci.p, cptr.p, keys.p, s));
    const uint64_t nn = (uint64_t)nc * (uint64_t)nc;
    if (nk > 0) {
      const int bits = std::max(1, sb::ceil_log2_host((int64_t)nn));
      size_t bytes = 0;
      SB_HIP(sb::sort_keys_u64(keys.p, skeys.p, nk, bits, nullptr, bytes, s));
      if ((rc = cub_tmp(tmp, bytes))) return rc;
      SB_HIP(sb::sort_keys_u64(keys.p, skeys.p, nk, bits, tmp.p, bytes, s));
    }
    SB_HIP(flags.alloc(nk));
    SB_HIP(sb::launch_unique_flags(nk, skeys.p, flags.p, s));
    if ((rc = scan_exclusive(flags.p, pos, nk, tmp, s))) return rc;
    int nw = 0;
    if ((rc = read_int(pos.p + nk, &nw, s))) return rc;
    SB_HIP(ukeys.alloc(nw));
    SB_HIP(sb::launch_unique_scatter(nk, skeys.p, pos.p + 1, ukeys.p, s));
    SB_HIP(P->d_wrp.alloc((size_t)nc + 1));
    SB_HIP(P->d_wci.alloc(nw));
    SB_HIP(sb::launch_key_rows(nc, ukeys.p, nw, P->d_wrp.p, P->d_wci.p, s));
    P->nwl = nw;
    // both triangles: lower row i, then the entries (j, i) below the diagonal as (i, j)
    SB_HIP(lk.alloc(nw));
    SB_HIP(slk.alloc(nw));
    SB_HIP(sb::launch_upper_keys(nc, P->d_wrp.p, P->d_wci.p, lk.p, s));
    if (nw > 0) {
      const int bits = std::max(1, sb::ceil_log2_host((int64_t)nn + 1));
      size_t bytes = 0;
      SB_HIP(sb::sort_keys_u64(lk.p, slk.p, nw, bits, nullptr, bytes, s));
      if ((rc = cub_tmp(tmp, bytes))) return rc;
      SB_HIP(sb::sort_keys_u64(lk.p, slk.p, nw, bits, tmp.p, bytes, s));
    }
    SB_HIP(uptr.alloc((size_t)nc + 1));
    SB_HIP(flen.alloc(nc));
    SB_HIP(sb::launch_upper_rows(nc, slk.p, nw, P->d_wrp.p, uptr.p, flen.p, s));
    if ((rc = scan_exclusive(flen.p, frp, nc, tmp, s))) return rc;
    int nf = 0;
    if ((rc = read_int(frp.p + nc, &nf, s))) return rc;
    SB_HIP(fci.alloc(nf));
    SB_HIP(sb::launch_full_fill(nc, P->d_wrp.p, P->d_wci.p, uptr.p, slk.p, frp.p, fci.p, s));
    if ((rc = build_tiles_device(P->W, nc, nc, 3, 3, frp.p, fci.p, P->d_wsrc, tmp, s))) return rc;
    // product terms of every W tile block
    const int nq = P->W.host.nblocks;
    SB_HIP(tcnt.alloc(nq));
    SB_HIP(sb::launch_w_term_count(nq, P->W.brow.p, P->W.bcol.p, P->d_wsrc.p, P->d_okey.p, P->d_hrp.p, P->d_hci.p,
                                   tcnt.p, s));
    if ((rc = scan_exclusive(tcnt.p, P->d_wtptr, nq, tmp, s))) return rc;
    int nt = 0;
    if ((rc = read_int(P->d_wtptr.p + nq, &nt, s))) return rc;
    SB_HIP(P->d_wterm.alloc(std::max<size_t>(2, 2 * (size_t)nt)));
    SB_HIP(sb::launch_w_term_fill(nq, P->W.brow.p, P->W.bcol.p, P->d_wsrc.p, P->d_okey.p, P->d_hrp.p, P->d_hci.p,
                                  P->d_wtptr.p, P->d_wterm.p, s));
    P->wtermOkey = true;
    SB_HIP(hipStreamSynchronize(s));  // the local buffers are freed on return
  }
  P->devSetup = true;
  P->devRawValid = true;
  P->hostStale = true;
  P->hostPattern = false;
  // numeric phase
  SB_HIP(P->d_minv.alloc((size_t)nb * 36));
  SB_HIP(P->d_minv32.alloc((size_t)nb * 36));
  SB_HIP(P->d_diag.alloc((size_t)nc * 9));
  SB_HIP(P->d_scal.alloc(nc));
  SB_HIP(P->d_ones.alloc(nc));
  SB_HIP(sb::launch_fill_f64((size_t)nc, 1.0, P->d_ones.p, s));
  SB_HIP(sb::launch_tile_values(3, 6, false, P->H.host.nblocks, P->d_hsrc.p, P->d_hraw.p, P->H.vals.p,
                                P->H.vals32.p, s));
  SB_HIP(sb::launch_tile_values(6, 3, true, P->HT.host.nblocks, P->d_htsrc.p, P->d_hraw.p, P->HT.vals.p,
                                P->HT.vals32.p, s));
  SB_HIP(sb::launch_minv_layout(nb, P->d_mraw.p, P->d_minv.p, P->d_minv32.p, s));
  SB_HIP(sb::launch_diag_scal(nc, P->d_hrp.p, P->d_hci.p, P->d_hraw.p, P->d_mraw.p, P->d_diag.p, P->d_scal.p, s));
  if (P->hasW)
    SB_HIP(sb::launch_w_numeric(P->W.host.nblocks, P->W.brow.p, P->W.bcol.p, P->d_wtptr.p, P->d_wterm.p,
                                P->d_okey.p, P->d_hci.p, P->d_hraw.p, P->d_mraw.p, P->W.vals.p, P->W.vals32.p, s));
  if ((rc = alloc_workspace(P))) return rc;
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}
}  // namespace

namespace {
int create_problem(sb_ctx* ctx, std::unique_ptr<sb_problem>& P, int nc, int nb, const int32_t* rp, const int32_t* ci,
                   const double* hv, const double* minv, const double* b, const double* mu, int flags);
int create_problem_device(sb_ctx* ctx, std::unique_ptr<sb_problem>& P, int nc, int nb, const int32_t* rp,
                          const int32_t* ci, const double* hv, const double* minv, const double* b, const double* mu,
                          int flags);
}  // namespace

extern "C" {

const char* sb_status_string(int status) {
  switch (status) {
    case SB_OK: return "ok";
    case SB_NOT_CONVERGED: return "not converged (tolerance not reached within max_iters)";
    case SB_ERR_INVALID: return "invalid argument";
    case SB_ERR_OOM: return "out of memory";
    case SB_ERR_HIP: return last_error_string().empty() ? "HIP runtime error" : last_error_string().c_str();
    case SB_ERR_NO_DEVICE: return "no usable GPU device";
    case SB_ERR_UNSUPPORTED: return "unsupported option on this path";
    case SB_ERR_BARRIER_TIMEOUT: return "grid barrier timed out in a persistent kernel";
    case SB_ERR_NAN: return "non-finite residual";
    case SB_ERR_LIMIT: return last_error_string().empty() ? "a fixed limit of the entry point is exceeded"
                                                          : last_error_string().c_str();
    case SB_ERR_IO: return "file could not be opened, read or written, or is not a local-problem file";
    default: return "unknown status";
  }
}

void sb_pg_options_default(sb_pg_options* o) {
  if (!o) return;
  o->variant = SB_PG_APGD;
  o->law = SB_LAW_SOC;
  o->op = SB_OP_ASSEMBLED;
  o->scaling = SB_SCALING_REFERENCE;
  o->inf_norm = 0;
  o->tol = 1e-6;
  o->max_iters = 300;
  o->ls_iters = 8;
  o->ls_optimistic = 1.25;
  o->ls_pessimistic = 0.5;
  o->ls_armijo = 0.5;
  o->grid_blocks = 0;
  o->block_threads = 0;
  o->precision = SB_PRECISION_FP64;
  o->mixed_tol_factor = 0.95;
  o->mixed_stall_iters = 100;
}

void sb_gs_options_default(sb_gs_options* o) {
  if (!o) return;
  o->law = SB_LAW_COULOMB;
  o->strategy = SB_REV_HYBRID;
  o->inf_norm = 0;
  o->try_zero = 1;
  o->tol = 1e-6;
  o->max_iters = 250;
  o->eval_every = 25;
  o->skip_tol = 1e-6;
  o->skip_iters = 10;
  o->auto_regularization = 0.0;
  o->grid_blocks = 0;
  o->block_threads = 0;
  o->precision = SB_PRECISION_FP64;
  o->mixed_tol_factor = 0.95;
}

int sb_ctx_create(sb_ctx** out, int device) {
  if (!out) return SB_ERR_INVALID;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SB_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return SB_ERR_INVALID;
  std::unique_ptr<sb_ctx> c(new (std::nothrow) sb_ctx());
  if (!c) return SB_ERR_OOM;
  c->device = device;
  SB_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  SB_HIP(hipGetDeviceProperties(&prop, device));
  c->ncu = prop.multiProcessorCount;
  c->name = prop.gcnArchName;
  SB_HIP(hipDeviceGetAttribute(&c->coop, hipDeviceAttributeCooperativeLaunch, device));
  {
    void* scratchPtr = nullptr;  // sb_pg_solve_noalloc's grid state on this device
    const int rc = resolve_scratch(device, &scratchPtr);
    if (rc != SB_OK) return rc;
  }
  // A plain launch of the same (co-resident sized) grid serves profilers that do
  // not handle cooperative launches; the barriers' bounded waits report a grid
  // that is not fully resident as SB_ERR_BARRIER_TIMEOUT instead of hanging.
  if (const char* m = std::getenv("SOBOGUS_LAUNCH")) c->coopLaunch = std::strcmp(m, "regular") != 0;
  SB_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  SB_HIP(c->sync.alloc(kSyncWords));
  SB_HIP(c->error.alloc(1));
  SB_HIP(c->out.alloc(16));
  SB_HIP(hipEventCreate(&c->ev0));
  SB_HIP(hipEventCreate(&c->ev1));
  SB_HIP(hipEventCreate(&c->evm));
  int rc = ensure_grid_scratch(c.get(), std::max(1, c->ncu * 8));
  if (rc != SB_OK) return rc;
  *out = c.release();
  return SB_OK;
}

int sb_ctx_destroy(sb_ctx* ctx) {
  if (!ctx) return SB_ERR_INVALID;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->evm) (void)hipEventDestroy(ctx->evm);
  hipStream_t s = ctx->stream;
  delete ctx;
  if (s) (void)hipStreamDestroy(s);
  return SB_OK;
}

int sb_ctx_launch_counts(sb_ctx* ctx, int64_t* cooperative, int64_t* plain) {
  if (!ctx) return SB_ERR_INVALID;
  if (cooperative) *cooperative = ctx->coopLaunches;
  if (plain) *plain = ctx->plainLaunches;
  return SB_OK;
}

int sb_ctx_info(sb_ctx* ctx, char* name, int name_len, int* n_cu) {
  if (!ctx) return SB_ERR_INVALID;
  if (name && name_len > 0) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (n_cu) *n_cu = ctx->ncu;
  return SB_OK;
}

// ------------------------------------------------------------------ BSR
int sb_bsr_create(sb_ctx* ctx, sb_bsr** A, int nbr, int nbc, int br, int bc, int flags, const int32_t* rowptr,
                  const int32_t* colidx, const double* vals) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!A || nbr < 1 || nbc < 1 || br < 1 || bc < 1 || br > 8 || bc > 8 || !rowptr) return SB_ERR_INVALID;
  const bool sym = (flags & SB_BSR_SYMMETRIC) != 0;
  if (sym && (br != bc || nbr != nbc)) return SB_ERR_INVALID;
  const int nnz = rowptr[nbr];
  if (nnz < 0 || (nnz > 0 && (!colidx || !vals))) return SB_ERR_INVALID;
  for (int i = 0; i < nbr; ++i)
    if (rowptr[i + 1] < rowptr[i]) return SB_ERR_INVALID;
  for (int p = 0; p < nnz; ++p)
    if (colidx[p] < 0 || colidx[p] >= nbc) return SB_ERR_INVALID;
  std::unique_ptr<sb_bsr> M(new (std::nothrow) sb_bsr());
  if (!M) return SB_ERR_OOM;
  M->nbr = nbr;
  M->nbc = nbc;
  M->br = br;
  M->bc = bc;
  M->flags = flags;
  hipStream_t s = ctx->stream;
  if (sym) {
    std::vector<int> frp, fci;
    std::vector<double> fv;
    sbhost::expandSymmetric(nbr, br, rowptr, colidx, vals, frp, fci, fv);
    SB_HIP(M->fwd.upload(sbhost::buildTiles(nbr, nbc, br, bc, frp.data(), fci.data(), fv.data()), s));
    if (br == 3) {
      M->lrp.assign(rowptr, rowptr + nbr + 1);
      M->lci.assign(colidx, colidx + nnz);
      SB_HIP(M->d_lrp.upload(M->lrp, s));
      SB_HIP(M->d_lci.upload(M->lci, s));
      SB_HIP(M->d_lvals.upload(vals, (size_t)nnz * 9, s));
    }
  } else {
    SB_HIP(M->fwd.upload(sbhost::buildTiles(nbr, nbc, br, bc, rowptr, colidx, vals), s));
    std::vector<int> trp, tci;
    std::vector<double> tv;
    sbhost::transposeBSR(nbr, nbc, br, bc, rowptr, colidx, vals, trp, tci, tv);
    SB_HIP(M->tr.upload(sbhost::buildTiles(nbc, nbr, bc, br, trp.data(), tci.data(), tv.data()), s));
  }
  const size_t carry = std::max({(size_t)M->fwd.host.ntiles, (size_t)M->tr.host.ntiles, (size_t)1}) * 8;
  SB_HIP(M->carry.alloc(carry));
  SB_HIP(hipStreamSynchronize(s));
  *A = M.release();
  return SB_OK;
}

int sb_bsr_destroy(sb_bsr* A) {
  if (!A) return SB_ERR_INVALID;
  delete A;
  return SB_OK;
}

int sb_bsr_multiply(sb_ctx* ctx, const sb_bsr* Ac, int transpose, double alpha, const double* x, double beta,
                    double* res, int on_device) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Ac || !x || !res) return SB_ERR_INVALID;
  sb_bsr* A = const_cast<sb_bsr*>(Ac);
  const bool sym = (A->flags & SB_BSR_SYMMETRIC) != 0;
  const bool tr = transpose && !sym;
  const DevTiles& T = tr ? A->tr : A->fwd;
  const int outRows = tr ? A->nbc : A->nbr, outBs = tr ? A->bc : A->br;
  const int inRows = tr ? A->nbr : A->nbc, inBs = tr ? A->br : A->bc;
  const size_t nin = (size_t)inRows * inBs, nout = (size_t)outRows * outBs;
  const double* dx = x;
  double* dy = res;
  hipStream_t s = ctx->stream;
  if (!on_device) {
    if (A->xbuf.n < nin) SB_HIP(A->xbuf.alloc(nin));
    if (A->ybuf.n < nout) SB_HIP(A->ybuf.alloc(nout));
    SB_HIP(hipMemcpyAsync(A->xbuf.p, x, nin * sizeof(double), hipMemcpyHostToDevice, s));
    if (beta != 0.0) SB_HIP(hipMemcpyAsync(A->ybuf.p, res, nout * sizeof(double), hipMemcpyHostToDevice, s));
    dx = A->xbuf.p;
    dy = A->ybuf.p;
  }
  SB_HIP(sb::launch_bsr_generic(T.view(), outBs, inBs, alpha, dx, beta, dy, A->carry.p, s));
  if (!on_device) SB_HIP(hipMemcpyAsync(res, dy, nout * sizeof(double), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}

int sb_bsr_split_row_multiply(sb_ctx* ctx, const sb_bsr* A, int row, const double* x, double* res3) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!A || !x || !res3 || !(A->flags & SB_BSR_SYMMETRIC) || A->br != 3 || row < 0 || row >= A->nbr)
    return SB_ERR_INVALID;
  hipStream_t s = ctx->stream;
  DBuf<double> dx, dr;
  SB_HIP(dx.upload(x, (size_t)3 * A->nbr, s));
  SB_HIP(dr.upload(res3, 3, s));
  (void)hipGetLastError();
  k_split_row<<<1, 64, 0, s>>>(A->d_lrp.p, A->d_lci.p, A->d_lvals.p, A->nbr, row, dx.p, dr.p);
  SB_HIP(hipGetLastError());
  SB_HIP(hipMemcpyAsync(res3, dr.p, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}

// ------------------------------------------------------------------ problem
int sb_problem_create(sb_ctx* ctx, sb_problem** out, int nc, int nb, const int32_t* rp, const int32_t* ci,
                      const double* hv, const double* minv, const double* b, const double* mu, int flags) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!out) return SB_ERR_INVALID;
  std::unique_ptr<sb_problem> P;
  rc = create_problem(ctx, P, nc, nb, rp, ci, hv, minv, b, mu, flags);
  if (rc) return rc;
  *out = P.release();
  return SB_OK;
}

int sb_problem_create_device(sb_ctx* ctx, sb_problem** out, int nc, int nb, const int32_t* rp, const int32_t* ci,
                             const double* hv, const double* minv, const double* b, const double* mu, int flags) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!out) return SB_ERR_INVALID;
  std::unique_ptr<sb_problem> P;
  rc = create_problem_device(ctx, P, nc, nb, rp, ci, hv, minv, b, mu, flags);
  if (rc) return rc;
  *out = P.release();
  return SB_OK;
}

int sb_problem_layout(sb_ctx* ctx, const sb_problem* Pc, int which, int64_t* sizes, int32_t* tstart, int32_t* brow,
                      int32_t* bcol, int32_t* tlong, int32_t* long_rows, int32_t* rowptr, int32_t* src,
                      double* vals) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (!P || !sizes || which < 0 || which > 2 || (which == 2 && !P->hasW)) return SB_ERR_INVALID;
  const DevTiles& T = which == 0 ? P->H : which == 1 ? P->HT : P->W;
  const sbhost::Tiles& h = T.host;
  sizes[0] = h.nrows;
  sizes[1] = h.nblocks;
  sizes[2] = h.ntiles;
  sizes[3] = T.nlong;
  sizes[4] = h.scan_steps;
  sizes[5] = T.pairs ? 1 : 0;
  hipStream_t s = ctx->stream;
  auto get = [&](int32_t* dst, const int* d, size_t n) -> int {
    if (dst && n) SB_HIP(hipMemcpyAsync(dst, d, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    return SB_OK;
  };
  const size_t nq = (size_t)h.nblocks, nt = (size_t)h.ntiles, nl = (size_t)T.nlong;
  if ((rc = get(tstart, T.tstart.p, nt + 1)) || (rc = get(brow, T.brow.p, nq)) || (rc = get(bcol, T.bcol.p, nq)) ||
      (rc = get(tlong, T.tlong.p, nt)) || (rc = get(rowptr, T.rowptr.p, (size_t)h.nrows + 1)))
    return rc;
  if (long_rows && nl) {
    if ((rc = get(long_rows, T.long_row.p, nl)) || (rc = get(long_rows + nl, T.long_t0.p, nl)) ||
        (rc = get(long_rows + 2 * nl, T.long_t1.p, nl)))
      return rc;
  }
  if (vals && nq) {
    const int bs = which == 2 ? 9 : 18;
    SB_HIP(hipMemcpyAsync(vals, T.vals.p, sizeof(double) * nq * bs, hipMemcpyDeviceToHost, s));
  }
  SB_HIP(hipStreamSynchronize(s));
  if (src) {
    if ((rc = sync_pattern(P))) return rc;
    const std::vector<int>& v = which == 0 ? P->H.host.src : which == 1 ? P->htsrc : P->W.host.src;
    if (v.size() != nq) return SB_ERR_INVALID;
    std::memcpy(src, v.data(), sizeof(int) * nq);
  }
  return SB_OK;
}

}  // extern "C"

namespace {

int create_problem(sb_ctx* ctx, std::unique_ptr<sb_problem>& P, int nc, int nb, const int32_t* rp, const int32_t* ci,
                   const double* hv, const double* minv, const double* b, const double* mu, int flags) {
  int rc;
  if (nc <= 0 || nb <= 0 || !rp || !ci || !hv || !minv || !b || !mu) return SB_ERR_INVALID;
  if (rp[0] != 0) return SB_ERR_INVALID;
  for (int i = 0; i < nc; ++i)
    if (rp[i + 1] < rp[i]) return SB_ERR_INVALID;
  const int nnz = rp[nc];
  for (int p = 0; p < nnz; ++p)
    if (ci[p] < 0 || ci[p] >= nb) return SB_ERR_INVALID;
  P.reset(new (std::nothrow) sb_problem());
  if (!P) return SB_ERR_OOM;
  P->ctx = ctx;
  P->nc = nc;
  P->nb = nb;
  P->hasW = (flags & SB_PROBLEM_ASSEMBLE_W) != 0;
  P->hostSetup = (flags & SB_PROBLEM_HOST_SETUP) != 0;
  if (!P->hostSetup) {
    // raw arrays to the device, set-up there; the host copies are made when needed
    hipStream_t s = ctx->stream;
    P->nnz = nnz;
    SB_HIP(P->d_hrp.upload(rp, (size_t)nc + 1, s));
    SB_HIP(P->d_hci.upload(ci, (size_t)nnz, s));
    SB_HIP(P->d_hraw.upload(hv, (size_t)nnz * 18, s));
    SB_HIP(P->d_mraw.upload(minv, (size_t)nb * 36, s));
    SB_HIP(P->d_b.upload(b, (size_t)nc * 3, s));
    SB_HIP(P->d_mu.upload(mu, (size_t)nc, s));
    SB_HIP(P->d_okey.alloc(nc));
    SB_HIP(sb::launch_iota(nc, P->d_okey.p, s));
    return build_device(P.get());
  }
  P->hrp.assign(rp, rp + nc + 1);
  P->hci.assign(ci, ci + nnz);
  P->hv.assign(hv, hv + (size_t)nnz * 18);
  P->minv.assign(minv, minv + (size_t)nb * 36);
  P->b.assign(b, b + 3 * (size_t)nc);
  P->mu.assign(mu, mu + nc);
  if (P->hasW)
    sbhost::assembleW(nc, nb, P->hrp.data(), P->hci.data(), P->hv.data(), P->minv.data(), P->wrp, P->wci, P->wv);
  rc = upload_problem(P.get());
  if (rc) return rc;
  return SB_OK;
}

int create_problem_device(sb_ctx* ctx, std::unique_ptr<sb_problem>& P, int nc, int nb, const int32_t* rp,
                          const int32_t* ci, const double* hv, const double* minv, const double* b, const double* mu,
                          int flags) {
  int rc;
  if (nc <= 0 || nb <= 0 || !rp || !ci || !hv || !minv || !b || !mu) return SB_ERR_INVALID;
  hipStream_t s = ctx->stream;
  DBuf<int> err;
  SB_HIP(err.alloc(1));
  SB_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
  SB_HIP(sb::launch_check_rows(nc, rp, err.p, s));
  int bad = 0, nnz = 0;
  SB_HIP(hipMemcpyAsync(&nnz, rp + nc, sizeof(int), hipMemcpyDeviceToHost, s));
  if ((rc = read_int(err.p, &bad, s))) return rc;
  if (bad || nnz < 0) return SB_ERR_INVALID;
  SB_HIP(sb::launch_check_cols(nnz, nb, ci, err.p, s));
  if ((rc = read_int(err.p, &bad, s))) return rc;
  if (bad) return SB_ERR_INVALID;
  P.reset(new (std::nothrow) sb_problem());
  if (!P) return SB_ERR_OOM;
  P->ctx = ctx;
  P->nc = nc;
  P->nb = nb;
  P->nnz = nnz;
  P->hasW = (flags & SB_PROBLEM_ASSEMBLE_W) != 0;
  auto dup = [&](auto& dst, const auto* src, size_t n) -> int {
    SB_HIP(dst.alloc(n));
    if (n) SB_HIP(hipMemcpyAsync(dst.p, src, sizeof(*src) * n, hipMemcpyDeviceToDevice, s));
    return SB_OK;
  };
  if ((rc = dup(P->d_hrp, rp, (size_t)nc + 1))) return rc;
  if ((rc = dup(P->d_hci, ci, (size_t)nnz))) return rc;
  if ((rc = dup(P->d_hraw, hv, (size_t)nnz * 18))) return rc;
  if ((rc = dup(P->d_mraw, minv, (size_t)nb * 36))) return rc;
  if ((rc = dup(P->d_b, b, (size_t)nc * 3))) return rc;
  if ((rc = dup(P->d_mu, mu, (size_t)nc))) return rc;
  SB_HIP(P->d_okey.alloc(nc));
  SB_HIP(sb::launch_iota(nc, P->d_okey.p, s));
  return build_device(P.get());
}

}  // namespace

extern "C" {

int sb_problem_set_w(sb_ctx* ctx, sb_problem* P, const int32_t* wrp, const int32_t* wci, const double* wv) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!P || !wrp || !wci || !wv || wrp[0] != 0) return SB_ERR_INVALID;
  const int nc = P->nc;
  for (int i = 0; i < nc; ++i) {
    if (wrp[i + 1] < wrp[i]) return SB_ERR_INVALID;
    for (int p = wrp[i]; p < wrp[i + 1]; ++p)
      if (wci[p] < 0 || wci[p] > i) return SB_ERR_INVALID;  // lower triangle only
  }
  if ((rc = sync_host(P))) return rc;
  P->wrp.assign(wrp, wrp + nc + 1);
  P->wci.assign(wci, wci + wrp[nc]);
  P->wv.assign(wv, wv + (size_t)wrp[nc] * 9);
  P->hasW = true;
  P->wFromH = false;
  return upload_problem(P);
}

int sb_problem_create_from_w(sb_ctx* ctx, sb_problem** out, int nc, const int32_t* rp, const int32_t* ci,
                             const double* v, const double* b, const double* mu, double prune_tol) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!out || nc <= 0 || !rp || !ci || !v || !b || !mu || !(prune_tol >= 0.0)) return SB_ERR_INVALID;
  const int n = 3 * nc;
  if (rp[0] != 0) return SB_ERR_INVALID;
  for (int i = 0; i < n; ++i)
    if (rp[i + 1] < rp[i]) return SB_ERR_INVALID;
  for (int p = 0; p < rp[n]; ++p)
    if (ci[p] < 0 || ci[p] >= n) return SB_ERR_INVALID;
  // 3x3 blocks of the lower triangle (bogus::convert into a symmetric block matrix,
  // SparseConversions.hpp:26-95): block row I reads scalar rows 3I..3I+2, each up to
  // its first entry right of the diagonal block; the diagonal block is symmetrised;
  // then blocks whose entries are all <= prune_tol in magnitude are dropped
  // (SparseBlockMatrixBase::prune, SparseBlockMatrixBase.impl.hpp:286-305).
  std::vector<int> wrp(nc + 1, 0), wci;
  std::vector<double> wv;
  std::map<int, std::array<double, 9>> row;
  for (int I = 0; I < nc; ++I) {
    row.clear();
    for (int i = 0; i < 3; ++i)
      for (int p = rp[3 * I + i]; p < rp[3 * I + i + 1]; ++p) {
        const int J = ci[p] / 3;
        if (J > I) break;
        row.try_emplace(J, std::array<double, 9>{}).first->second[3 * i + (ci[p] - 3 * J)] = v[p];
      }
    auto d = row.find(I);
    if (d != row.end()) {
      const std::array<double, 9> D = d->second;
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d->second[3 * r + c] = 0.5 * (D[3 * r + c] + D[3 * c + r]);
    }
    for (const auto& [J, B] : row) {
      bool zero = true;
      for (double x : B) zero = zero && std::fabs(x) <= prune_tol;
      if (zero) continue;
      wci.push_back(J);
      wv.insert(wv.end(), B.begin(), B.end());
    }
    wrp[I + 1] = (int)wci.size();
  }
  std::unique_ptr<sb_problem> P(new (std::nothrow) sb_problem());
  if (!P) return SB_ERR_OOM;
  // one placeholder body without contacts: H is empty and only the assembled W is used
  P->ctx = ctx;
  P->nc = nc;
  P->nb = 1;
  P->hostSetup = true;
  P->dualOnly = true;
  P->hrp.assign((size_t)nc + 1, 0);
  P->minv.assign(36, 0.0);
  for (int k = 0; k < 6; ++k) P->minv[7 * k] = 1.0;
  P->b.assign(b, b + 3 * (size_t)nc);
  P->mu.assign(mu, mu + nc);
  P->hasW = true;
  P->wFromH = false;
  P->wrp.swap(wrp);
  P->wci.swap(wci);
  P->wv.swap(wv);
  if ((rc = upload_problem(P.get()))) return rc;
  *out = P.release();
  return SB_OK;
}

int sb_merit1(sb_ctx* ctx, const sb_problem* Pc, int op, const double* r, int on_device, double* merit) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !r || !merit || (op != SB_OP_ASSEMBLED && op != SB_OP_MATRIX_FREE)) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (op == SB_OP_ASSEMBLED && !P->hasW) return SB_ERR_INVALID;
  if (op == SB_OP_MATRIX_FREE && P->dualOnly) return SB_ERR_UNSUPPORTED;
  const size_t n3 = 3 * (size_t)P->nc;
  hipStream_t s = ctx->stream;
  if ((rc = copy_in(ctx, P->rio.p, r, n3, on_device))) return rc;
  const DevProblem<double> V = P->view(true);
  if (op == SB_OP_ASSEMBLED)
    SB_HIP(sb::launch_w(V, P->rio.p, P->y.p, P->carry.p, s));
  else
    SB_HIP(sb::launch_matfree(V, P->rio.p, P->y.p, P->wbody.p, P->carry.p, s));
  SB_HIP(sb::launch_merit1_terms(P->nc, P->d_mu.p, P->y.p, P->d_b.p, P->rio.p, P->terms.p, s));
  std::vector<double> t(P->nc), bv(n3);
  SB_HIP(hipMemcpyAsync(t.data(), P->terms.p, sizeof(double) * P->nc, hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(bv.data(), P->d_b.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s));  // b is current on the device
  SB_HIP(hipStreamSynchronize(s));
  double err = 0.0, nb2 = 0.0;
  for (double x : t) err += x;
  for (double x : bv) nb2 += x * x;
  *merit = std::sqrt(err) / (1.0 + std::sqrt(nb2));
  return std::isfinite(*merit) ? SB_OK : SB_ERR_NAN;
}

int sb_problem_destroy(sb_problem* P) {
  if (!P) return SB_ERR_INVALID;
  (void)hipSetDevice(P->ctx->device);
  delete P;
  return SB_OK;
}

int sb_problem_sizes(const sb_problem* P, int* nc, int* nb, int64_t* hblocks, int64_t* wblocks) {
  if (!P) return SB_ERR_INVALID;
  if (nc) *nc = P->nc;
  if (nb) *nb = P->nb;
  if (hblocks) *hblocks = (int64_t)P->nnz;
  if (wblocks) *wblocks = P->hasW ? P->nwl : 0;
  return SB_OK;
}

int sb_problem_get_w(sb_ctx* ctx, const sb_problem* Pc, int32_t* wrp, int32_t* wci, double* wv) {
  (void)ctx;
  if (!Pc || !Pc->hasW) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  const int rc = sync_host(P);
  if (rc) return rc;
  if (wrp) std::memcpy(wrp, P->wrp.data(), sizeof(int32_t) * P->wrp.size());
  if (wci) std::memcpy(wci, P->wci.data(), sizeof(int32_t) * P->wci.size());
  if (wv) std::memcpy(wv, P->wv.data(), sizeof(double) * P->wv.size());
  return SB_OK;
}

int sb_problem_apply_permutation(sb_ctx* ctx, sb_problem* P, const int64_t* perm) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!P || !perm) return SB_ERR_INVALID;
  const int n = P->nc;
  std::vector<int> inv(n, -1);
  for (int i = 0; i < n; ++i) {
    if (perm[i] < 0 || perm[i] >= n || inv[perm[i]] >= 0) return SB_ERR_INVALID;
    inv[perm[i]] = i;
  }
  // On the device (then a new device set-up) unless the problem was set up on the host
  // or carries a W that was not assembled from H.
  if (!P->hostSetup && (!P->hasW || P->wFromH)) {
    hipStream_t s = ctx->stream;
    if (!P->devRawValid) {
      SB_HIP(P->d_hraw.upload(P->hv, s));
      SB_HIP(P->d_mraw.upload(P->minv, s));
      P->devRawValid = true;
    }
    if (P->d_hrp.n != (size_t)n + 1) {
      SB_HIP(P->d_hrp.upload(P->hrp, s));
      SB_HIP(P->d_hci.upload(P->hci, s));
    }
    std::vector<int> p32(perm, perm + n);
    DBuf<int> dperm, len, nrp, nci;
    DBuf<double> nhv, nb, nmu;
    DBuf<char> tmp;
    SB_HIP(dperm.upload(p32, s));
    SB_HIP(len.alloc(n));
    SB_HIP(sb::launch_perm_lens(n, dperm.p, P->d_hrp.p, len.p, s));
    if ((rc = scan_exclusive(len.p, nrp, n, tmp, s))) return rc;
    SB_HIP(nci.alloc((size_t)P->nnz));
    SB_HIP(nhv.alloc((size_t)P->nnz * 18));
    SB_HIP(nb.alloc((size_t)n * 3));
    SB_HIP(nmu.alloc(n));
    SB_HIP(sb::launch_perm_rows(n, dperm.p, P->d_hrp.p, P->d_hci.p, P->d_hraw.p, P->d_b.p, P->d_mu.p, nrp.p, nci.p,
                                nhv.p, nb.p, nmu.p, s));
    SB_HIP(hipStreamSynchronize(s));
    DBuf<int> nokey;
    SB_HIP(nokey.alloc(n));
    SB_HIP(sb::launch_gather_int(n, dperm.p, P->d_okey.p, nokey.p, s));
    SB_HIP(hipStreamSynchronize(s));
    P->d_hrp.swap(nrp);
    P->d_hci.swap(nci);
    P->d_hraw.swap(nhv);
    P->d_b.swap(nb);
    P->d_mu.swap(nmu);
    P->d_okey.swap(nokey);
    return build_device(P);
  }
  if ((rc = sync_host(P))) return rc;
  // H rows, b, mu: new i <- old perm[i]
  std::vector<int> rp(n + 1, 0), ci;
  std::vector<double> hv, b(3 * (size_t)n), mu(n);
  for (int i = 0; i < n; ++i) rp[i + 1] = rp[i] + (P->hrp[perm[i] + 1] - P->hrp[perm[i]]);
  ci.resize(rp[n]);
  hv.resize((size_t)rp[n] * 18);
  for (int i = 0; i < n; ++i) {
    const int src = (int)perm[i];
    int dst = rp[i];
    for (int p = P->hrp[src]; p < P->hrp[src + 1]; ++p, ++dst) {
      ci[dst] = P->hci[p];
      std::memcpy(&hv[(size_t)dst * 18], &P->hv[(size_t)p * 18], 18 * sizeof(double));
    }
    for (int k = 0; k < 3; ++k) b[3 * i + k] = P->b[3 * src + k];
    mu[i] = P->mu[src];
  }
  P->hrp.swap(rp);
  P->hci.swap(ci);
  P->hv.swap(hv);
  P->b.swap(b);
  P->mu.swap(mu);
  if (P->hasW) {
    // block (i, j) of the new W is old (perm[i], perm[j]); blocks landing above the
    // diagonal are stored transposed (SparseBlockMatrixBase::applyPermutation)
    std::vector<std::vector<std::pair<int, int>>> rows(n);
    for (int oi = 0; oi < n; ++oi)
      for (int p = P->wrp[oi]; p < P->wrp[oi + 1]; ++p) {
        int ni = inv[oi], nj = inv[P->wci[p]];
        int tag = p + 1;
        if (nj > ni) {
          std::swap(ni, nj);
          tag = -(p + 1);
        }
        rows[ni].push_back({nj, tag});
      }
    std::vector<int> wrp(n + 1, 0), wci;
    std::vector<double> wv;
    for (int i = 0; i < n; ++i) {
      std::sort(rows[i].begin(), rows[i].end());
      wrp[i + 1] = wrp[i] + (int)rows[i].size();
    }
    wci.resize(wrp[n]);
    wv.resize((size_t)wrp[n] * 9);
    for (int i = 0; i < n; ++i) {
      int dst = wrp[i];
      for (auto& e : rows[i]) {
        wci[dst] = e.first;
        const double* src = &P->wv[(size_t)(std::abs(e.second) - 1) * 9];
        double* d = &wv[(size_t)dst * 9];
        if (e.second > 0)
          std::memcpy(d, src, 9 * sizeof(double));
        else
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) d[3 * r + c] = src[3 * c + r];
        ++dst;
      }
    }
    P->wrp.swap(wrp);
    P->wci.swap(wci);
    P->wv.swap(wv);
  }
  return upload_problem(P);
}

int sb_problem_coloring(sb_ctx* ctx, const sb_problem* Pc, int mode, int64_t* perm, int64_t* colors, int* ncolors) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  const bool onHost = (mode & SB_COLORING_HOST) != 0, devOnly = (mode & SB_COLORING_DEVICE) != 0;
  mode &= ~(SB_COLORING_HOST | SB_COLORING_DEVICE);
  if (onHost && devOnly) return SB_ERR_INVALID;
  if (!Pc || !perm || !colors || !ncolors || (mode != 0 && mode != 1)) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  // First fit on the device when the coupling is the contact-body incidence of H (no
  // W, or W assembled from H); the host rule otherwise, where the device rule's limits
  // (colors, bodies per contact) are exceeded, and for the reference rule (whose
  // thousands of colors at 50k contacts make it a parity tool, not a solver input).
  if (devOnly && mode != SB_COLORING_GREEDY) return SB_ERR_UNSUPPORTED;
  if (!onHost && mode == SB_COLORING_GREEDY && (!P->hasW || P->wFromH)) {
    hipStream_t s = ctx->stream;
    const int nc = P->nc, nb = P->nb;
    if (P->d_hrp.n != (size_t)nc + 1) {  // a host set-up uploads the pattern on demand
      SB_HIP(P->d_hrp.upload(P->hrp, s));
      SB_HIP(P->d_hci.upload(P->hci, s));
    }
    if (P->d_htsrc.n != (size_t)P->HT.host.nblocks) SB_HIP(P->d_htsrc.upload(P->htsrc, s));
    DBuf<int> color, info, perm32, keys, iota, offs, deg;
    SB_HIP(color.alloc(nc));
    SB_HIP(info.alloc(3));  // [0] colors, [1] limit exceeded, [2] contacts left after a round
    SB_HIP(deg.alloc(1));
    SB_HIP(hipMemsetAsync(info.p, 0, 3 * sizeof(int), s));
    SB_HIP(hipMemsetAsync(color.p, 0xff, sizeof(int) * (size_t)nc, s));  // -1: not colored
    SB_HIP(sb::launch_max_degree(nb, P->HT.rowptr.p, deg.p, s));
    int hdeg = 0;
    if ((rc = read_int(deg.p, &hdeg, s))) return rc;
    int hinfo[3] = {0, 1, 0};  // a body with more contacts than colors: the host rule
    if (hdeg <= sb::coloring_max_colors()) {
      hinfo[1] = 0;
      // rounds in groups of four; the group's last round counts the contacts left.
      // Every round colors at least the first uncolored contact; a chain of coupled
      // contacts longer than kMaxColorRounds goes to the host rule.
      constexpr int kMaxColorRounds = 4096;
      for (int rounds = 0;; rounds += 4) {
        if (rounds >= kMaxColorRounds) {
          hinfo[1] = 1;
          break;
        }
        for (int g = 0; g < 4; ++g) {
          if (g == 3) SB_HIP(hipMemsetAsync(info.p + 2, 0, sizeof(int), s));
          SB_HIP(sb::launch_color_round(nc, P->d_hrp.p, P->d_hci.p, P->HT.rowptr.p, P->HT.bcol.p, P->d_htsrc.p,
                                        color.p, g == 3 ? info.p + 2 : deg.p, info.p + 1, s));
        }
        SB_HIP(hipMemcpyAsync(hinfo + 1, info.p + 1, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
        SB_HIP(hipStreamSynchronize(s));
        if (hinfo[1] || hinfo[2] == 0) break;
      }
    }
    if (hinfo[1] == 0) {
      DBuf<char> tmp;
      size_t bytes = 0;
      SB_HIP(sb::reduce_max_int(color.p, info.p, nc, nullptr, bytes, s));
      if ((rc = cub_tmp(tmp, bytes))) return rc;
      SB_HIP(sb::reduce_max_int(color.p, info.p, nc, tmp.p, bytes, s));
      if ((rc = read_int(info.p, &hinfo[0], s))) return rc;
      const int ncol = hinfo[0] + 1;
      // perm: the contacts by color, ascending within a color (a stable sort of the colors)
      SB_HIP(keys.alloc(nc));
      SB_HIP(iota.alloc(nc));
      SB_HIP(perm32.alloc(nc));
      SB_HIP(offs.alloc((size_t)ncol + 1));
      SB_HIP(sb::launch_iota(nc, iota.p, s));
      const int bits = std::max(1, sb::ceil_log2_host(ncol));
      bytes = 0;
      const unsigned* ck = reinterpret_cast<const unsigned*>(color.p);
      unsigned* sk = reinterpret_cast<unsigned*>(keys.p);
      SB_HIP(sb::sort_pairs_u32(ck, sk, iota.p, perm32.p, nc, bits, nullptr, bytes, s));
      if ((rc = cub_tmp(tmp, bytes))) return rc;
      SB_HIP(sb::sort_pairs_u32(ck, sk, iota.p, perm32.p, nc, bits, tmp.p, bytes, s));
      SB_HIP(sb::launch_lower_bounds_u32(ncol + 1, sk, nc, offs.p, s));
      std::vector<int> hp(nc), ho((size_t)ncol + 1);
      SB_HIP(hipMemcpyAsync(hp.data(), perm32.p, sizeof(int) * hp.size(), hipMemcpyDeviceToHost, s));
      SB_HIP(hipMemcpyAsync(ho.data(), offs.p, sizeof(int) * ho.size(), hipMemcpyDeviceToHost, s));
      SB_HIP(hipStreamSynchronize(s));
      for (int i = 0; i < nc; ++i) perm[i] = hp[i];
      for (int c = 0; c <= ncol; ++c) colors[c] = ho[c];
      *ncolors = ncol;
      return SB_OK;
    }
  }
  if (devOnly) {
    last_error_string() = "coloring: the device rule's limits are exceeded, or W was supplied";
    return SB_ERR_LIMIT;
  }
  if ((rc = sync_pattern(P))) return rc;
  std::vector<int64_t> pm, cl;
  // the pattern of W is the contact-body incidence of H: the coloring does not need W
  if (P->hasW)
    sbhost::coloring(P->nc, P->wrp.data(), P->wci.data(), mode, pm, cl);
  else
    sbhost::coloringFromH(P->nc, P->nb, P->hrp.data(), P->hci.data(), mode, pm, cl);
  std::memcpy(perm, pm.data(), sizeof(int64_t) * pm.size());
  std::memcpy(colors, cl.data(), sizeof(int64_t) * cl.size());
  *ncolors = (int)cl.size() - 1;
  return SB_OK;
}

int sb_problem_apply_w(sb_ctx* ctx, const sb_problem* Pc, int op, const double* x, double* y, int on_device) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !x || !y) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (op == SB_OP_ASSEMBLED && !P->hasW) return SB_ERR_INVALID;
  if (op != SB_OP_ASSEMBLED && P->dualOnly) return SB_ERR_UNSUPPORTED;
  const size_t n3 = 3 * (size_t)P->nc;
  hipStream_t s = ctx->stream;
  const double* dx = x;
  double* dy = y;
  if (!on_device) {
    if ((rc = copy_in(ctx, P->rio.p, x, n3, 0))) return rc;
    dx = P->rio.p;
    dy = P->y.p;
  }
  const DevProblem<double> V = P->view(true);
  if (op == SB_OP_ASSEMBLED)
    SB_HIP(sb::launch_w(V, dx, dy, P->carry.p, s));
  else
    SB_HIP(sb::launch_matfree(V, dx, dy, P->wbody.p, P->carry.p, s));
  if (!on_device && (rc = copy_out(ctx, y, dy, n3, 0))) return rc;
  SB_HIP(hipStreamSynchronize(s));
  return SB_OK;
}

int sb_fb_residual(sb_ctx* ctx, const sb_problem* Pc, int law, int op, int scaling, int inf_norm, const double* r,
                   double* residual, int on_device) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !r || !residual) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (op == SB_OP_ASSEMBLED && !P->hasW) return SB_ERR_INVALID;
  if (op != SB_OP_ASSEMBLED && P->dualOnly) return SB_ERR_UNSUPPORTED;
  const size_t n3 = 3 * (size_t)P->nc;
  hipStream_t s = ctx->stream;
  if ((rc = copy_in(ctx, P->rio.p, r, n3, on_device))) return rc;
  const DevProblem<double> V = P->view(true);
  if (op == SB_OP_ASSEMBLED)
    SB_HIP(sb::launch_w(V, P->rio.p, P->y.p, P->carry.p, s));
  else
    SB_HIP(sb::launch_matfree(V, P->rio.p, P->y.p, P->wbody.p, P->carry.p, s));
  const bool useScal = (scaling == SB_SCALING_DIAGONAL) || (scaling == SB_SCALING_REFERENCE && op == SB_OP_ASSEMBLED);
  SB_HIP(sb::launch_fb_terms(P->nc, law == SB_LAW_COULOMB, P->d_mu.p, useScal ? P->d_scal.p : nullptr, P->y.p, P->d_b.p,
                             P->rio.p, P->terms.p, s));
  std::vector<double> t(P->nc);
  SB_HIP(hipMemcpyAsync(t.data(), P->terms.p, sizeof(double) * P->nc, hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  double acc = 0.0;
  if (inf_norm) {
    for (double v : t) acc = std::max(acc, v);
  } else {
    for (double v : t) acc += v;
    acc /= (1.0 + P->nc);
  }
  *residual = acc;
  return std::isfinite(acc) ? SB_OK : SB_ERR_NAN;
}

// Factored operator tiles for the matrix-free projected-gradient solvers: with the
// Cholesky factors Minv_k = L_k L_k^T, W = F F^T for F_ik = H_ik L_k, so an operator
// application is the two tile passes F (F^T x), with no Minv blocks to read and no
// body step between them, and stays exactly symmetric in FP32. Computed on the device
// from the current tile values and Minv after the problem's values change; a Minv_k
// that is not symmetric positive definite keeps the unfactored product.
// Storage of the factored tiles (allocated once per pattern).
static bool factor_allocated(const sb_problem* P) {
  return P->fH.n == 18 * (size_t)P->H.host.nblocks && P->fHT.n == 18 * (size_t)P->HT.host.nblocks &&
         P->fH32.n == P->fH.n && P->fHT32.n == P->fHT.n && P->ferr.n == 1 && P->fL.n == 36 * (size_t)P->nb;
}
static int factor_alloc(sb_problem* P) {
  const size_t nh = 18 * (size_t)P->H.host.nblocks, nht = 18 * (size_t)P->HT.host.nblocks;
  if (P->fH.n != nh) SB_HIP(P->fH.alloc(nh));
  if (P->fH32.n != nh) SB_HIP(P->fH32.alloc(nh));
  if (P->fHT.n != nht) SB_HIP(P->fHT.alloc(nht));
  if (P->fHT32.n != nht) SB_HIP(P->fHT32.alloc(nht));
  if (P->ferr.n != 1) SB_HIP(P->ferr.alloc(1));
  if (P->fL.n != 36 * (size_t)P->nb) SB_HIP(P->fL.alloc(36 * (size_t)P->nb));
  return SB_OK;
}
// Recomputes the factored tiles into their storage (no allocation).
static int factor_refresh(sb_problem* P) {
  hipStream_t s = P->ctx->stream;
  int err = 0;
  SB_HIP(hipMemcpyAsync(P->ferr.p, &err, sizeof(int), hipMemcpyHostToDevice, s));
  SB_HIP(sb::launch_minv_chol(P->nb, P->d_minv.p, P->fL.p, P->ferr.p, s));
  SB_HIP(sb::launch_factor_tiles(false, P->H.host.nblocks, P->nb, P->H.bcol.p, P->H.vals.p, P->fL.p, P->fH.p,
                                 P->fH32.p, s));
  SB_HIP(sb::launch_factor_tiles(true, P->HT.host.nblocks, P->nb, P->HT.brow.p, P->HT.vals.p, P->fL.p, P->fHT.p,
                                 P->fHT32.p, s));
  SB_HIP(hipMemcpyAsync(&err, P->ferr.p, sizeof(int), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  P->factored = (err == 0);
  P->fVersion = P->version;
  return SB_OK;
}
static int factor_prepare(sb_problem* P) {
  if (P->fVersion == P->version) return SB_OK;
  int rc = factor_alloc(P);
  return rc ? rc : factor_refresh(P);
}

}  // extern "C"

namespace {
bool cand_phase_for(const sb_problem* P, const sb_pg_options* o) {
  const bool lineSearch = o->variant == SB_PG_APGD || o->variant == SB_PG_DESCENT || o->variant == SB_PG_SPG;
  return o->op == SB_OP_MATRIX_FREE && lineSearch && P->nc >= kCandPhaseContacts;
}

// The projected-gradient solve of a prepared problem on device vector x (warm start in,
// result out): one persistent launch per stage (FP32 then FP64 in mixed precision), grid
// state from the context or the fixed array (fixed != null), results into stats.
int pg_run(sb_ctx* ctx, sb_problem* P, const sb_pg_options* o, double* x, int histCap, const GridState* fixed,
           double* residual, sb_stats* st) {
  hipStream_t s = ctx->stream;
  const bool mf = (o->op == SB_OP_MATRIX_FREE);
  const bool mixed = (o->precision == SB_PRECISION_MIXED);
  const sb::PGParams prm = make_params(o, mf);
  const int workWaves = mf ? std::max(P->H.host.ntiles, P->HT.host.ntiles) : P->W.host.ntiles;
  const int threads = block_threads_for(o->block_threads, kMaxThreads);
  if (threads < 0) return SB_ERR_INVALID;
  double* dout = fixed ? fixed->out : P->out.p;
  int rc;
  ctx->lastPlain = false;
  SB_HIP(hipEventRecord(ctx->ev0, s));
  if (mixed) {
    // FP32 stage to tol * factor (no history: the callback stream is the FP64 one)
    sb::PGGridArgs<float> a;
    std::memset(&a, 0, sizeof(a));
    a.prm = prm;
    a.prm.tol = o->tol * (o->mixed_tol_factor > 0 ? o->mixed_tol_factor : 0.95);
    a.prm.stall_iters = o->mixed_stall_iters;
    P->fill(a.pb, true);
    a.pb.x_io = x;
    a.pb.out = dout;
    const void* fn = pg_grid_fn<float>(mf, cand_phase_for(P, o), o->variant);
    if ((rc = launch_pg_grid(ctx, fixed, fn, &a, threads, o->grid_blocks, workWaves))) return rc;
    SB_HIP(hipEventRecord(ctx->evm, s));
  }
  {
    // FP64 solve; in mixed precision it starts from the FP32 result, so it only
    // verifies the tolerance unless refinement is needed
    sb::PGGridArgs<double> a;
    std::memset(&a, 0, sizeof(a));
    a.prm = prm;
    P->fill(a.pb, true);
    a.pb.x_io = x;
    a.pb.hist = histCap > 0 ? P->hist.p : nullptr;
    a.pb.hist_cap = histCap;
    a.pb.out = dout + 16;
    const void* fn = pg_grid_fn<double>(mf, cand_phase_for(P, o), o->variant);
    if ((rc = launch_pg_grid(ctx, fixed, fn, &a, threads, o->grid_blocks, workWaves))) return rc;
  }
  SB_HIP(hipEventRecord(ctx->ev1, s));
  double out[32];
  SB_HIP(hipMemcpyAsync(out, dout, sizeof(out), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  const double* o32 = out;
  const double* o64 = out + 16;
  for (int k = 0; k < 4; ++k) ctx->phase_us[k] = o64[7 + k] + (mixed ? o32[7 + k] : 0.0);
  if (st) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st->iterations = (int)o64[1] + (mixed ? (int)o32[1] : 0);
    st->spmvs = (int)o64[2] + (mixed ? (int)o32[2] : 0);
    st->ls_trials = (int)o64[3] + (mixed ? (int)o32[3] : 0);
    st->nesterov_steps = (int)o64[4] + (mixed ? (int)o32[4] : 0);
    st->converged = (int)o64[5];
    st->local_failures = 0;
    st->residual = o64[0];
    st->device_ms = ms;
    st->history_length = (int)o64[1] + ((int)o64[5] ? 1 : 0);
    st->refine_iterations = mixed ? (int)o64[1] : 0;
    st->refine_spmvs = mixed ? (int)o64[2] : 0;
    st->mixed_residual = mixed ? o32[0] : 0.0;
    st->plain_launch = ctx->lastPlain ? 1 : 0;
    st->mixed_device_ms = 0.0;
    if (mixed) {
      float m32 = 0.f;
      (void)hipEventElapsedTime(&m32, ctx->ev0, ctx->evm);
      st->mixed_device_ms = m32;
    }
    if (histCap > 0) {
      const int hl = std::min(st->history_length, histCap);
      if (hl > 0) SB_HIP(hipMemcpy(st->history, P->hist.p, sizeof(double) * hl, hipMemcpyDeviceToHost));
    }
  }
  if (residual) *residual = o64[0];
  if (o64[6] != 0.0 || (mixed && o32[6] != 0.0)) return SB_ERR_BARRIER_TIMEOUT;
  if (!std::isfinite(o64[0])) return SB_ERR_NAN;
  return o64[5] != 0.0 ? SB_OK : SB_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int sb_pg_solve(sb_ctx* ctx, const sb_problem* Pc, const sb_pg_options* o, double* r, int on_device,
                double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !r) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if ((rc = check_pg_problem(o, P))) return rc;
  if (!ctx->coop && ctx->coopLaunch) return SB_ERR_UNSUPPORTED;
  const size_t n3 = 3 * (size_t)P->nc;
  if ((rc = copy_in(ctx, P->rio.p, r, n3, on_device))) return rc;
  const int histCap = (st && st->history && st->history_capacity > 0) ? st->history_capacity : 0;
  if (histCap > 0 && P->hist.n < (size_t)histCap) SB_HIP(P->hist.alloc(histCap));
  if (o->op == SB_OP_MATRIX_FREE && (rc = factor_prepare(P))) return rc;
  const int prc = pg_run(ctx, P, o, P->rio.p, histCap, nullptr, residual, st);
  if (prc < 0 && prc != SB_ERR_BARRIER_TIMEOUT && prc != SB_ERR_NAN) return prc;
  if ((rc = copy_out(ctx, r, P->rio.p, n3, on_device))) return rc;
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return prc;
}

int sb_pg_solve_noalloc(sb_ctx* ctx, const sb_problem* Pc, const sb_pg_options* o, double* r_dev,
                        double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !r_dev) return SB_ERR_INVALID;
  if (st && st->history && st->history_capacity > 0) return SB_ERR_INVALID;  // no history buffer here
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if ((rc = check_pg_problem(o, P))) return rc;
  if (!ctx->coop && ctx->coopLaunch) return SB_ERR_UNSUPPORTED;
  if (o->op == SB_OP_MATRIX_FREE && P->fVersion != P->version) {
    // the factored tiles are refreshed in place; their storage comes from an earlier
    // matrix-free sb_pg_solve on this problem
    if (!factor_allocated(P)) {
      last_error_string() = "the matrix-free operator's factored tiles are not allocated: solve once with sb_pg_solve "
                            "and SB_OP_MATRIX_FREE first";
      return SB_ERR_LIMIT;
    }
    if ((rc = factor_refresh(P))) return rc;
  }
  void* base = nullptr;
  if ((rc = resolve_scratch(ctx->device, &base))) return rc;
  unsigned char* b = static_cast<unsigned char*>(base);
  GridState fixed;
  fixed.sync = reinterpret_cast<unsigned*>(b);
  fixed.error = reinterpret_cast<int*>(b + kNoAllocErrorOff);
  fixed.out = reinterpret_cast<double*>(b + kNoAllocOutOff);
  fixed.partials = reinterpret_cast<double*>(b + kNoAllocPartialsOff);
  fixed.maxBlocks = kNoAllocBlocks;
  return pg_run(ctx, P, o, r_dev, 0, &fixed, residual, st);
}

// Set-up shared by the two Gauss-Seidel solvers: options, color offsets, local
// matrices with auto-regularization (GaussSeidelBase::processLocalMatrices,
// GaussSeidelBase.impl.hpp:20-40), skip counters, r, history, kernel arguments.
static int gs_prepare(sb_ctx* ctx, sb_problem* P, const sb_gs_options* o, const int64_t* colors, int ncolors,
                      double* r, int on_device, sb_stats* st, sb::GSArgs& a, int& histCap, int& maxColor) {
  if (o->max_iters < 0 || o->eval_every < 1 || o->strategy < 0 || o->strategy > 3) return SB_ERR_INVALID;
  if (o->precision != SB_PRECISION_FP64 && o->precision != SB_PRECISION_MIXED) return SB_ERR_INVALID;
  if (o->precision == SB_PRECISION_MIXED && !(o->mixed_tol_factor > 0.0 && o->mixed_tol_factor <= 1.0))
    return SB_ERR_INVALID;
  if (!ctx->coop && ctx->coopLaunch) return SB_ERR_UNSUPPORTED;
  const int n = P->nc;
  std::vector<int> col;
  if (colors && ncolors > 0) {
    if (colors[0] != 0 || colors[ncolors] != n) return SB_ERR_INVALID;
    col.resize(ncolors + 1);
    for (int c = 0; c <= ncolors; ++c) {
      if (c > 0 && colors[c] < colors[c - 1]) return SB_ERR_INVALID;
      col[c] = (int)colors[c];
    }
  } else {
    col = {0, n};
  }
  hipStream_t s = ctx->stream;
  int rc;
  std::vector<double> local((size_t)9 * n), reg(n, 0.0);  // the diagonal blocks (current on the device)
  SB_HIP(hipMemcpyAsync(local.data(), P->d_diag.p, sizeof(double) * local.size(), hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  if (o->auto_regularization > 0.0)
    for (int i = 0; i < n; ++i) {
      reg[i] = std::max(0.0, o->auto_regularization - min_eig_sym3(&local[(size_t)9 * i]));
      local[(size_t)9 * i] += reg[i];
      local[(size_t)9 * i + 4] += reg[i];
      local[(size_t)9 * i + 8] += reg[i];
    }
  SB_HIP(P->local.upload(local, s));
  SB_HIP(P->reg.upload(reg, s));
  SB_HIP(P->colors.upload(col, s));
  SB_HIP(P->skip.alloc(n));
  SB_HIP(hipMemsetAsync(P->skip.p, 0, n, s));
  if ((rc = copy_in(ctx, P->rio.p, r, 3 * (size_t)n, on_device))) return rc;
  histCap = (st && st->history && st->history_capacity > 0) ? st->history_capacity : 0;
  if (histCap > 0 && P->hist.n < (size_t)histCap) SB_HIP(P->hist.alloc(histCap));
  std::memset(&a, 0, sizeof(a));
  a.nc = n;
  a.coulomb = o->law == SB_LAW_COULOMB;
  a.strategy = o->strategy;
  a.inf_norm = o->inf_norm;
  a.try_zero = o->try_zero;
  a.tol = o->tol;
  a.max_iters = o->max_iters;
  a.eval_every = o->eval_every;
  a.skip_tol = o->skip_tol;
  a.skip_iters = std::min(o->skip_iters, (int)std::sqrt((double)n));
  a.local_tol = std::pow(2.220446049250313e-16, 0.75);  // SOCLaw::m_localTol
  a.P = P->view(true);
  a.local = P->local.p;
  a.reg = P->reg.p;
  a.colors = P->colors.p;
  a.ncolors = (int)col.size() - 1;
  a.x = P->rio.p;
  a.y = P->y.p;
  a.xbest = P->xbest.p;
  a.skip = P->skip.p;
  a.carry = P->carry.p;
  a.hist = P->hist.p;
  a.hist_cap = histCap;
  a.out = ctx->out.p;
  a.mixed = o->precision == SB_PRECISION_MIXED ? 1 : 0;
  a.stop_factor = o->mixed_tol_factor;
  a.W32 = P->W.view32();
  a.x32 = P->X32[0].p;
  a.carry32 = P->carry32.p;
  maxColor = 0;
  for (size_t c = 0; c + 1 < col.size(); ++c) maxColor = std::max(maxColor, col[c + 1] - col[c]);
  return SB_OK;
}

// Launch of a prepared Gauss-Seidel kernel, result and statistics.
static int gs_run(sb_ctx* ctx, sb_problem* P, const void* fn, int grid, int threads, void** args, double* r,
                  int on_device, double* residual, sb_stats* st, int histCap) {
  hipStream_t s = ctx->stream;
  SB_HIP(sb::launch_reset_sync(ctx->sync.p, (int)kSyncWords, ctx->error.p, s));
  ctx->lastPlain = false;
  SB_HIP(hipEventRecord(ctx->ev0, s));
  SB_HIP(launch_persistent(ctx, fn, grid, threads, args, s));
  SB_HIP(hipEventRecord(ctx->ev1, s));
  double out[9];
  int rc;
  if ((rc = read_status(ctx, out, 9))) return rc;
  const bool mixed = static_cast<sb::GSArgs*>(args[0])->mixed != 0;
  if ((rc = copy_out(ctx, r, P->rio.p, 3 * (size_t)P->nc, on_device))) return rc;
  if (st) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st->iterations = (int)out[1];
    st->spmvs = (int)out[4];
    st->ls_trials = 0;
    st->nesterov_steps = 0;
    st->converged = (int)out[2];
    st->local_failures = (int)out[3];
    st->residual = out[0];
    st->device_ms = ms;
    st->history_length = (int)out[4];
    st->refine_iterations = mixed ? (int)(out[1] - out[6]) : 0;
    st->refine_spmvs = mixed ? (int)out[8] : 0;
    st->mixed_residual = mixed ? out[7] : 0.0;
    st->plain_launch = ctx->lastPlain ? 1 : 0;
    st->mixed_device_ms = 0.0;
    if (histCap > 0) {
      const int hl = std::min(st->history_length, histCap);
      if (hl > 0) SB_HIP(hipMemcpyAsync(st->history, P->hist.p, sizeof(double) * hl, hipMemcpyDeviceToHost, s));
    }
  }
  SB_HIP(hipStreamSynchronize(s));
  if (residual) *residual = out[0];
  if (out[5] != 0.0) return SB_ERR_BARRIER_TIMEOUT;
  if (!std::isfinite(out[0])) return SB_ERR_NAN;
  return out[2] != 0.0 ? SB_OK : SB_NOT_CONVERGED;
}

int sb_gs_solve(sb_ctx* ctx, const sb_problem* Pc, const sb_gs_options* o, const int64_t* colors, int ncolors,
                double* r, int on_device, double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !r) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (!P->hasW) return SB_ERR_INVALID;
  if ((rc = sync_pattern(P))) return rc;
  // Contacts of one color are solved in parallel from each other's previous values, so
  // a color must hold no coupled pair (W_ij != 0). Given colors are checked; without
  // colors the natural order is cut into maximal runs of uncoupled contacts, which
  // sweeps the contacts exactly in the reference's sequential order.
  std::vector<int64_t> runs;
  {
    const int n = P->nc;
    std::vector<int> colorOf(n, -1);
    if (colors && ncolors > 0) {
      if (colors[0] != 0 || colors[ncolors] != n) return SB_ERR_INVALID;
      for (int c = 0; c < ncolors; ++c) {
        if (colors[c + 1] < colors[c]) return SB_ERR_INVALID;
        for (int64_t i = colors[c]; i < colors[c + 1]; ++i) colorOf[i] = c;
      }
      for (int i = 0; i < n; ++i)
        for (int p = P->wrp[i]; p < P->wrp[i + 1]; ++p)
          if (P->wci[p] != i && colorOf[P->wci[p]] == colorOf[i]) return SB_ERR_INVALID;
    } else {
      runs.push_back(0);
      int start = 0;
      for (int i = 0; i < n; ++i) {
        bool clash = false;
        for (int p = P->wrp[i]; p < P->wrp[i + 1]; ++p) clash |= P->wci[p] != i && P->wci[p] >= start;
        if (clash) {
          runs.push_back(i);
          start = i;
        }
      }
      runs.push_back(n);
      colors = runs.data();
      ncolors = (int)runs.size() - 1;
    }
  }
  sb::GSArgs a;
  int histCap = 0, maxColor = 0;
  if ((rc = gs_prepare(ctx, P, o, colors, ncolors, r, on_device, st, a, histCap, maxColor))) return rc;
  const void* fn = (const void*)&sb::gs_kernel;
  const int threads = block_threads_for(o->block_threads, kThreads);
  if (threads != 256) return threads < 0 ? SB_ERR_INVALID : SB_ERR_UNSUPPORTED;  // GS kernel: 256 threads
  const int grid = grid_for(ctx, fn, threads, o->grid_blocks, std::max(P->W.host.ntiles, (maxColor + 63) / 64));
  if ((rc = ensure_grid_scratch(ctx, grid))) return rc;
  a.gs = make_sync(ctx, grid);
  const int* wrowptr = P->W.rowptr.p;
  void* args[] = {&a, &wrowptr};
  return gs_run(ctx, P, fn, grid, threads, args, r, on_device, residual, st, histCap);
}

// Per-H-block Minv_k H_ik^T (DMtStorage<.., true>::compute, ProductGaussSeidel.impl.hpp:64-68)
// in the H tiles' block order, and each contact's first block and block count.
static int pgs_upload(sb_problem* P) {
  if (P->pgsVersion == P->version) return SB_OK;
  const int rc = sync_host(P);
  if (rc) return rc;
  const sbhost::Tiles& T = P->H.host;
  const int n = P->nc;
  const size_t nq = (size_t)T.nblocks;
  std::vector<double> dmt(18 * nq, 0.0);
  std::vector<int> hrow(2 * (size_t)n);
  for (int i = 0; i < n; ++i) {
    const int q0 = T.rowptr[i], len = P->hrp[i + 1] - P->hrp[i];
    hrow[2 * i] = q0;
    hrow[2 * i + 1] = len;
    for (int t = 0; t < len; ++t) {
      const int p = P->hrp[i] + t;
      const double* h = &P->hv[(size_t)p * 18];
      const double* m = &P->minv[(size_t)P->hci[p] * 36];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0.0;
          for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * h[c * 6 + l];
          dmt[(size_t)(r * 3 + c) * nq + q0 + t] = acc;
        }
    }
  }
  hipStream_t s = P->ctx->stream;
  SB_HIP(P->dmt.upload(dmt, s));
  SB_HIP(P->hrow.upload(hrow, s));
  SB_HIP(hipStreamSynchronize(s));
  P->pgsVersion = P->version;
  return SB_OK;
}

int sb_pgs_solve(sb_ctx* ctx, const sb_problem* Pc, const sb_gs_options* o, const int64_t* colors, int ncolors,
                 double* r, int on_device, double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !r) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (P->dualOnly) return SB_ERR_UNSUPPORTED;  // the product form needs H
  if ((rc = sync_pattern(P))) return rc;
  // Contacts of one color update the body vector in parallel, so they must share no
  // body. Given colors are checked; without colors the natural order is cut into
  // maximal runs of contacts that share no body, which sweeps the contacts exactly
  // in the reference's sequential order.
  std::vector<int64_t> runs;
  {
    std::vector<int> stamp(P->nb, -1);
    if (colors && ncolors > 0) {
      if (colors[0] != 0 || colors[ncolors] != P->nc) return SB_ERR_INVALID;
      for (int c = 0; c < ncolors; ++c) {
        if (colors[c + 1] < colors[c]) return SB_ERR_INVALID;
        for (int64_t i = colors[c]; i < colors[c + 1]; ++i)
          for (int p = P->hrp[i]; p < P->hrp[i + 1]; ++p) {
            if (stamp[P->hci[p]] == c) return SB_ERR_INVALID;
            stamp[P->hci[p]] = c;
          }
      }
    } else {
      runs.push_back(0);
      int c = 0;
      for (int i = 0; i < P->nc; ++i) {
        bool clash = false;
        for (int p = P->hrp[i]; p < P->hrp[i + 1]; ++p) clash |= stamp[P->hci[p]] == c;
        if (clash) {
          runs.push_back(i);
          ++c;
        }
        for (int p = P->hrp[i]; p < P->hrp[i + 1]; ++p) stamp[P->hci[p]] = c;
      }
      runs.push_back(P->nc);
      colors = runs.data();
      ncolors = (int)runs.size() - 1;
    }
  }
  sb::GSArgs a;
  int histCap = 0, maxColor = 0;
  if (o->precision != SB_PRECISION_FP64) return SB_ERR_UNSUPPORTED;  // mixed precision: sb_gs_solve only
  if ((rc = gs_prepare(ctx, P, o, colors, ncolors, r, on_device, st, a, histCap, maxColor))) return rc;
  if ((rc = pgs_upload(P))) return rc;
  sb::PGSArgs e;
  e.hrow = P->hrow.p;
  e.dmt = P->dmt.p;
  e.ybody = P->wbody.p;
  const void* fn = (const void*)&sb::pgs_kernel;
  const int threads = block_threads_for(o->block_threads, kThreads);
  if (threads != 256) return threads < 0 ? SB_ERR_INVALID : SB_ERR_UNSUPPORTED;  // GS kernels: 256 threads
  const int work = std::max({P->H.host.ntiles, P->HT.host.ntiles, (maxColor + 63) / 64});
  const int grid = grid_for(ctx, fn, threads, o->grid_blocks, work);
  if ((rc = ensure_grid_scratch(ctx, grid))) return rc;
  a.gs = make_sync(ctx, grid);
  void* args[] = {&a, &e};
  return gs_run(ctx, P, fn, grid, threads, args, r, on_device, residual, st, histCap);
}

// ------------------------------------------------------------------ Cadoux fixed point
// friction_problem::solveCadoux / solveCadouxVel (src/Interfaces/Cadoux.hpp:44-155,
// FrictionProblem.cpp:68-100): SOC solves whose affine term carries the De Saxce
// correction of the previous iterate. Everything stays on the device; the problem's
// b is swapped for each inner problem's and restored afterwards.
static int cadoux_run(sb_ctx* ctx, sb_problem* P, const sb_pg_options* po, const sb_gs_options* go,
                      const int64_t* colors, int ncolors, int iterations, double tighten, int velocity, double* x,
                      int on_device, double* residual, sb_stats* st) {
  const bool gs = go != nullptr;
  const int n = P->nc;
  const size_t n3 = 3 * (size_t)n;
  hipStream_t s = ctx->stream;
  const int op = gs ? SB_OP_ASSEMBLED : po->op;
  if (op == SB_OP_ASSEMBLED && !P->hasW) return SB_ERR_INVALID;
  if (iterations < 0 || !(tighten > 0.0)) return SB_ERR_INVALID;
  // the minimizer's own scaling, as its eval() uses it
  const bool useScal = gs || po->scaling == SB_SCALING_DIAGONAL ||
                       (po->scaling == SB_SCALING_REFERENCE && op == SB_OP_ASSEMBLED);
  const int infNorm = gs ? go->inf_norm : po->inf_norm;
  const double tol = gs ? go->tol : po->tol;
  // Host copies first: the inner solves run while d_b holds an inner problem's affine
  // term, and a host refresh then (the GS set-up reads the host copies) would take
  // that term for the problem's b.
  int rc;
  if ((rc = sync_host(P))) return rc;
  for (auto& b : P->cad) SB_HIP(b.alloc(n3));
  double *b0 = P->cad[0].p, *cx = P->cad[1].p, *cs = P->cad[2].p, *cw = P->cad[3].p, *cu = P->cad[4].p;
  SB_HIP(hipMemcpyAsync(b0, P->d_b.p, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
  if ((rc = copy_in(ctx, cx, x, n3, on_device))) return rc;
  const DevProblem<double> V = P->view(true);
  int spmvs = 0;
  auto applyW = [&](const double* in, double* out) -> hipError_t {
    ++spmvs;
    return op == SB_OP_ASSEMBLED ? sb::launch_w(V, in, out, P->carry.p, s)
                                 : sb::launch_matfree(V, in, out, P->wbody.p, P->carry.p, s);
  };
  // FB residual of (xa, W(.) + b0), the minimizer's norm and scaling
  std::vector<double> terms(n);
  auto evalRes = [&](bool coulomb, const double* xa, const double* wy, double& res) -> int {
    SB_HIP(sb::launch_fb_terms(n, coulomb ? 1 : 0, P->d_mu.p, useScal ? P->d_scal.p : nullptr, wy, b0, xa,
                               P->terms.p, s));
    SB_HIP(hipMemcpyAsync(terms.data(), P->terms.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    SB_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    if (infNorm) {
      for (double v : terms) acc = std::max(acc, v);
    } else {
      for (double v : terms) acc += v;
      acc /= (1.0 + n);
    }
    res = acc;
    return std::isfinite(acc) ? SB_OK : SB_ERR_NAN;
  };
  bool anyPlain = false;
  // the inner SOC solve of xi at tolerance t, with d_b already set; b restored after
  auto inner = [&](double* xi, double t) -> int {
    sb_stats is;
    std::memset(&is, 0, sizeof(is));
    double r = 0.0;
    int irc;
    if (gs) {
      sb_gs_options g = *go;
      g.law = SB_LAW_SOC;
      g.tol = t;
      irc = sb_gs_solve(ctx, P, &g, colors, ncolors, xi, 1, &r, &is);
    } else {
      sb_pg_options g = *po;
      g.law = SB_LAW_SOC;
      g.tol = t;
      irc = sb_pg_solve(ctx, P, &g, xi, 1, &r, &is);
    }
    spmvs += is.spmvs + is.refine_spmvs;
    anyPlain |= is.plain_launch != 0;
    const hipError_t e = hipMemcpyAsync(P->d_b.p, b0, n3 * sizeof(double), hipMemcpyDeviceToDevice, s);
    if (irc < 0) return irc;
    return e == hipSuccess ? SB_OK : SB_ERR_HIP;
  };
  std::vector<double> hist;
  double res = 0.0;
  int it = 0;
  SB_HIP(hipEventRecord(ctx->ev0, s));
  if (!velocity) {
    // s = W r + b; res = eval(Coulomb, s, r)
    SB_HIP(applyW(cx, cw));
    if ((rc = evalRes(true, cx, cw, res))) return rc;
    hist.push_back(res);
    while (it < iterations) {
      const double t = tighten * std::max(tol, std::min(res, 1.0));
      SB_HIP(sb::launch_cadoux_rhs(n, P->d_mu.p, cw, b0, P->d_b.p, s));  // b + s(W r + b)
      if ((rc = inner(cx, t))) return rc;
      SB_HIP(applyW(cx, cw));
      if ((rc = evalRes(true, cx, cw, res))) return rc;
      hist.push_back(res);
      ++it;
      if (res < tol) break;
    }
  } else {
    // x holds u; ustar = u + s(u), s_N = |u_T| max(0, 1/mu); res = eval(SOC, W u + b, ustar)
    SB_HIP(hipMemcpyAsync(cu, cx, n3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    SB_HIP(sb::launch_vel_shift(n, P->d_mu.p, cu, cs, s));
    SB_HIP(sb::launch_lincomb(n3, 1.0, cu, 1.0, cs, cx, s));
    SB_HIP(applyW(cu, cw));
    if ((rc = evalRes(false, cx, cw, res))) return rc;
    hist.push_back(res);
    while (it < iterations) {
      const double t = tighten * std::max(tol, std::min(res, 1.0));
      SB_HIP(applyW(cs, cw));
      SB_HIP(sb::launch_lincomb(n3, 1.0, b0, -1.0, cw, P->d_b.p, s));  // b - W s
      if ((rc = inner(cx, t))) return rc;
      SB_HIP(sb::launch_lincomb(n3, 1.0, cx, -1.0, cs, cu, s));  // u = ustar - s
      SB_HIP(sb::launch_vel_shift(n, P->d_mu.p, cx, cs, s));
      SB_HIP(sb::launch_lincomb(n3, 1.0, cu, 1.0, cs, cx, s));  // ustar = u + s
      SB_HIP(applyW(cu, cw));
      if ((rc = evalRes(false, cx, cw, res))) return rc;
      hist.push_back(res);
      ++it;
      if (it > 1 && res < tol) break;
    }
  }
  SB_HIP(hipEventRecord(ctx->ev1, s));
  if ((rc = copy_out(ctx, x, velocity ? cu : cx, n3, on_device))) return rc;
  SB_HIP(hipStreamSynchronize(s));
  const bool conv = res < tol;
  if (st) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st->iterations = it;
    st->spmvs = spmvs;
    st->ls_trials = 0;
    st->nesterov_steps = 0;
    st->converged = conv ? 1 : 0;
    st->local_failures = 0;
    st->residual = res;
    st->device_ms = ms;
    st->history_length = (int)hist.size();
    st->refine_iterations = 0;
    st->refine_spmvs = 0;
    st->mixed_residual = 0.0;
    st->plain_launch = anyPlain ? 1 : 0;
    st->mixed_device_ms = 0.0;
    if (st->history && st->history_capacity > 0)
      for (int k = 0; k < st->history_capacity && k < (int)hist.size(); ++k) st->history[k] = hist[k];
  }
  if (residual) *residual = res;
  return conv ? SB_OK : SB_NOT_CONVERGED;
}

int sb_pg_solve_cadoux(sb_ctx* ctx, const sb_problem* Pc, const sb_pg_options* o, int iterations,
                       double tol_tighten, int velocity, double* x, int on_device, double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !x) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if ((rc = check_pg_problem(o, P))) return rc;
  return cadoux_run(ctx, P, o, nullptr, nullptr, 0, iterations, tol_tighten, velocity, x, on_device, residual, st);
}

int sb_gs_solve_cadoux(sb_ctx* ctx, const sb_problem* Pc, const sb_gs_options* o, const int64_t* colors,
                       int ncolors, int iterations, double tol_tighten, int velocity, double* x, int on_device,
                       double* residual, sb_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !x) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  return cadoux_run(ctx, P, nullptr, o, colors, ncolors, iterations, tol_tighten, velocity, x, on_device, residual,
                    st);
}

// ------------------------------------------------------------------ memory helpers
int sb_problem_update(sb_ctx* ctx, sb_problem* P, const double* b, const double* mu, const double* minv) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!P) return SB_ERR_INVALID;
  if (minv && P->dualOnly) return SB_ERR_INVALID;  // no bodies
  if ((rc = sync_host(P))) return rc;
  if (b) P->b.assign(b, b + 3 * (size_t)P->nc);
  if (mu) P->mu.assign(mu, mu + P->nc);
  if (minv) {
    P->minv.assign(minv, minv + 36 * (size_t)P->nb);
    if (P->hasW)
      sbhost::assembleW(P->nc, P->nb, P->hrp.data(), P->hci.data(), P->hv.data(), P->minv.data(), P->wrp, P->wci,
                        P->wv);
  }
  rc = upload_values(P);
  if (rc) return rc;
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

void sb_adjoint_options_default(sb_adjoint_options* o) {
  if (!o) return;
  o->mode = SB_ADJOINT_UNROLL;
  o->law = SB_LAW_COULOMB;
  o->iterations = 200;
  o->checkpoint_every = 0;
  o->rho = 0.0;
  o->tol = 1e-10;
  o->max_iters = 3000;
  o->grid_blocks = 0;
  o->branch_mode = 0;
  o->krylov_dim = 40;
}

int sb_adjoint(sb_ctx* ctx, const sb_problem* Pc, const sb_adjoint_options* o, const double* r0, const double* g,
               double* r_out, double* grad_b, double* grad_mu, double* grad_minv, unsigned char* branches,
               sb_adjoint_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!Pc || !o || !r0) return SB_ERR_INVALID;
  if (Pc->dualOnly) return SB_ERR_UNSUPPORTED;  // derivatives in Minv and the operator need H
  if (o->mode != SB_ADJOINT_UNROLL && o->mode != SB_ADJOINT_IMPLICIT) return SB_ERR_INVALID;
  if (o->mode == SB_ADJOINT_IMPLICIT && !g) return SB_ERR_INVALID;
  if (o->mode == SB_ADJOINT_UNROLL && o->iterations < 1) return SB_ERR_INVALID;
  if (o->mode == SB_ADJOINT_IMPLICIT && (o->max_iters < 1 || !(o->tol > 0))) return SB_ERR_INVALID;
  if (o->mode == SB_ADJOINT_IMPLICIT && (o->krylov_dim < 1 || o->krylov_dim > sb::kMaxKrylov)) return SB_ERR_INVALID;
  if (!(o->rho >= 0) || o->checkpoint_every < 0) return SB_ERR_INVALID;
  if (o->branch_mode < 0 || o->branch_mode > 2) return SB_ERR_INVALID;
  if (o->branch_mode != 0 && (!branches || o->mode != SB_ADJOINT_UNROLL)) return SB_ERR_INVALID;
  if (!ctx->coop && ctx->coopLaunch) return SB_ERR_UNSUPPORTED;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  hipStream_t s = ctx->stream;
  const int nc = P->nc, nb = P->nb;
  const size_t n3 = 3 * (size_t)nc;
  const DevProblem<double> V = P->view(true);
  // step of the map: 1 / lambda_max(W), lambda_max by power iteration on the device
  double rho = o->rho;
  if (rho == 0.0) {
    std::vector<double> x(n3), y(n3);
    for (size_t k = 0; k < n3; ++k) x[k] = 1.0 + 0.5 * std::sin(0.7 * (double)k);
    double lam = 0.0;
    for (int it = 0; it < 40; ++it) {
      double nx = 0.0;
      for (double v : x) nx += v * v;
      nx = std::sqrt(nx);
      for (double& v : x) v /= nx;
      SB_HIP(hipMemcpyAsync(P->X[0].p, x.data(), sizeof(double) * n3, hipMemcpyHostToDevice, s));
      SB_HIP(sb::launch_matfree(V, P->X[0].p, P->y.p, P->wbody.p, P->carry.p, s));
      SB_HIP(hipMemcpyAsync(y.data(), P->y.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s));
      SB_HIP(hipStreamSynchronize(s));
      double ny = 0.0;
      for (double v : y) ny += v * v;
      lam = std::sqrt(ny);
      x.swap(y);
    }
    if (!(lam > 0)) return SB_ERR_NAN;
    rho = 1.0 / lam;
  }
  const int K = o->iterations;
  int C = o->checkpoint_every;
  if (C == 0) C = std::max(1, (int)std::ceil(std::sqrt((double)K)));
  const int nseg = (K + C - 1) / C;
  DBuf<double> r0d, gd, rK, ckpt, seg, lam, kry, mz, muu, hr, jac, gb, gmu, gminv, out;
  DBuf<unsigned char> br;
  const size_t nbr = (size_t)K * nc;
  if (o->branch_mode == 2) SB_HIP(br.upload(branches, nbr, s));
  if (o->branch_mode == 1) SB_HIP(br.alloc(nbr));
  SB_HIP(r0d.upload(r0, n3, s));
  if (g) SB_HIP(gd.upload(g, n3, s));
  SB_HIP(rK.alloc(n3));
  SB_HIP(lam.alloc(2 * n3));
  SB_HIP(mz.alloc(n3));
  SB_HIP(muu.alloc(n3));
  SB_HIP(hr.alloc(6 * (size_t)nb));
  SB_HIP(gb.alloc(n3));
  SB_HIP(gmu.alloc(nc));
  SB_HIP(gminv.alloc(36 * (size_t)nb));
  SB_HIP(out.alloc(8));
  if (o->mode == SB_ADJOINT_UNROLL) {
    SB_HIP(ckpt.alloc((size_t)(nseg + 1) * n3));
    SB_HIP(seg.alloc((size_t)std::max(C, 2) * n3));
  } else {
    SB_HIP(jac.alloc(16 * (size_t)nc));
    SB_HIP(kry.alloc((size_t)(o->krylov_dim + 1) * n3));
  }
  // Coulomb law: the De Saxce directions of every step, recorded with the branches and
  // replayed with them (kept with the problem between the two calls)
  double* sdir = nullptr;
  if (o->law == SB_LAW_COULOMB && o->mode == SB_ADJOINT_UNROLL && o->branch_mode != 0) {
    const size_t nd = 2 * (size_t)K * nc;
    if (o->branch_mode == 1) {
      if (P->adjDir.n != nd) SB_HIP(P->adjDir.alloc(nd));
      P->adjDirK = K;
    } else if (P->adjDir.n != nd || P->adjDirK != K) {
      last_error_string() = "replay of a Coulomb unroll needs the directions recorded by an unroll of the same length";
      return SB_ERR_INVALID;
    }
    sdir = P->adjDir.p;
  }
  sb::AdjArgs a;
  std::memset(&a, 0, sizeof(a));
  a.P = V;
  a.mode = o->mode;
  a.K = K;
  a.C = C;
  a.coulomb = o->law == SB_LAW_COULOMB;
  a.rho = rho;
  a.tol = o->tol;
  a.max_iters = o->max_iters;
  a.kdim = o->krylov_dim;
  a.kry = kry.p;
  a.sdir = sdir;
  a.r0 = r0d.p;
  a.g = g ? gd.p : nullptr;
  a.rK = rK.p;
  a.ckpt = ckpt.p;
  a.seg = seg.p;
  a.lam = lam.p;
  a.mz = mz.p;
  a.mu_u = muu.p;
  a.hr = hr.p;
  a.jac = jac.p;
  a.wbody = P->wbody.p;
  a.carry = P->carry.p;
  a.grad_b = gb.p;
  a.grad_mu = gmu.p;
  a.grad_minv = gminv.p;
  a.branch = br.p;
  a.branch_mode = o->branch_mode;
  a.out = out.p;
  const void* fn = (const void*)&sb::adjoint_kernel;
  const int threads = kMaxThreads;
  const int grid = grid_for(ctx, fn, threads, o->grid_blocks, std::max(P->H.host.ntiles, P->HT.host.ntiles));
  if ((rc = ensure_grid_scratch(ctx, grid))) return rc;
  a.gs = make_sync(ctx, grid);
  SB_HIP(sb::launch_reset_sync(ctx->sync.p, (int)kSyncWords, ctx->error.p, s));
  void* args[] = {&a};
  SB_HIP(hipEventRecord(ctx->ev0, s));
  SB_HIP(launch_persistent(ctx, fn, grid, threads, args, s));
  SB_HIP(hipEventRecord(ctx->ev1, s));
  double res[8];
  SB_HIP(hipMemcpyAsync(res, out.p, sizeof(res), hipMemcpyDeviceToHost, s));
  if (r_out) SB_HIP(hipMemcpyAsync(r_out, o->mode == SB_ADJOINT_UNROLL ? rK.p : r0d.p, sizeof(double) * n3,
                                   hipMemcpyDeviceToHost, s));
  if (g && grad_b) SB_HIP(hipMemcpyAsync(grad_b, gb.p, sizeof(double) * n3, hipMemcpyDeviceToHost, s));
  if (g && grad_mu) SB_HIP(hipMemcpyAsync(grad_mu, gmu.p, sizeof(double) * nc, hipMemcpyDeviceToHost, s));
  if (g && grad_minv) SB_HIP(hipMemcpyAsync(grad_minv, gminv.p, sizeof(double) * 36 * nb, hipMemcpyDeviceToHost, s));
  if (o->branch_mode == 1) SB_HIP(hipMemcpyAsync(branches, br.p, nbr, hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  if (st) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st->rho = rho;
    st->loss = g ? res[0] : 0.0;
    st->iterations = (int)res[1];
    st->change = res[2];
    st->converged = (int)res[3];
    st->device_ms = ms;
  }
  if (res[4] != 0.0) return SB_ERR_BARRIER_TIMEOUT;
  if (o->mode == SB_ADJOINT_IMPLICIT && res[3] == 0.0) return SB_NOT_CONVERGED;
  return SB_OK;
}

int sb_device_alloc(sb_ctx* ctx, void** ptr, size_t bytes) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!ptr) return SB_ERR_INVALID;
  SB_HIP(hipMalloc(ptr, bytes));
  return SB_OK;
}
int sb_device_free(sb_ctx* ctx, void* ptr) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  SB_HIP(hipFree(ptr));
  return SB_OK;
}
int sb_memcpy_h2d(sb_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  SB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}
int sb_memcpy_d2h(sb_ctx* ctx, void* dst, const void* src, size_t bytes) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  SB_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}
// Device-side refresh of a problem's values for an unchanged pattern: new H block
// values (row-major 3x6, in the problem's current contact order and BSR block
// order), inverse masses (row-major 6x6 per body), b, mu; NULL leaves a part as it
// is. Tiles of H and H^T, Minv (FP64 and FP32 copies), the diagonal blocks and row
// scalings and, with an assembled W, W's values are recomputed by kernels from
// index maps built once per pattern. The host copies are refreshed lazily, by
// the host paths that need them.
int sb_problem_update_values(sb_ctx* ctx, sb_problem* P, const double* hvals, const double* minv, const double* b,
                             const double* mu, int on_device) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!P) return SB_ERR_INVALID;
  if ((hvals || minv) && P->hasW && !P->wFromH) return SB_ERR_INVALID;  // W was supplied, not H Minv H^T
  hipStream_t s = ctx->stream;
  const int nc = P->nc, nb = P->nb;
  const size_t nh = (size_t)P->nnz * 18;
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if ((rc = setup_maps(P))) return rc;
  if (!P->devRawValid) {  // current values from the host copies
    SB_HIP(P->d_hraw.upload(P->hv, s));
    SB_HIP(P->d_mraw.upload(P->minv, s));
    P->devRawValid = true;
  }
  if (hvals) SB_HIP(hipMemcpyAsync(P->d_hraw.p, hvals, sizeof(double) * nh, kind, s));
  if (minv) SB_HIP(hipMemcpyAsync(P->d_mraw.p, minv, sizeof(double) * 36 * (size_t)nb, kind, s));
  if (b) SB_HIP(hipMemcpyAsync(P->d_b.p, b, sizeof(double) * 3 * (size_t)nc, kind, s));
  if (mu) SB_HIP(hipMemcpyAsync(P->d_mu.p, mu, sizeof(double) * (size_t)nc, kind, s));
  if (hvals) {
    SB_HIP(sb::launch_tile_values(3, 6, false, P->H.host.nblocks, P->d_hsrc.p, P->d_hraw.p, P->H.vals.p,
                                  P->H.vals32.p, s));
    SB_HIP(sb::launch_tile_values(6, 3, true, P->HT.host.nblocks, P->d_htsrc.p, P->d_hraw.p, P->HT.vals.p,
                                  P->HT.vals32.p, s));
  }
  if (minv) SB_HIP(sb::launch_minv_layout(nb, P->d_mraw.p, P->d_minv.p, P->d_minv32.p, s));
  if (hvals || minv) {
    SB_HIP(sb::launch_diag_scal(nc, P->d_hrp.p, P->d_hci.p, P->d_hraw.p, P->d_mraw.p, P->d_diag.p, P->d_scal.p, s));
    if (P->hasW)
      SB_HIP(sb::launch_w_numeric(P->W.host.nblocks, P->W.brow.p, P->W.bcol.p, P->d_wtptr.p, P->d_wterm.p,
                                  P->wtermOkey ? P->d_okey.p : nullptr, P->d_hci.p, P->d_hraw.p, P->d_mraw.p,
                                  P->W.vals.p, P->W.vals32.p, s));
  }
  SB_HIP(hipStreamSynchronize(s));
  ++P->version;
  P->hostStale = true;
  return SB_OK;
}

// Minv_k = M_k^-1 for n_bodies 6x6 row-major blocks (PrimalFrictionProblem::computeMInv,
// PrimalFrictionProblem.cpp:28-38), on the device; SB_ERR_INVALID if a block is singular.
int sb_mass_inverse(sb_ctx* ctx, int n_bodies, const double* M, double* minv, int on_device) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n_bodies < 0 || (n_bodies > 0 && (!M || !minv))) return SB_ERR_INVALID;
  if (n_bodies == 0) return SB_OK;
  hipStream_t s = ctx->stream;
  const size_t n = 36 * (size_t)n_bodies;
  DBuf<double> dM, dI;
  DBuf<int> err;
  const double* pM = M;
  double* pI = minv;
  if (!on_device) {
    SB_HIP(dM.alloc(n));
    SB_HIP(dI.alloc(n));
    SB_HIP(hipMemcpyAsync(dM.p, M, sizeof(double) * n, hipMemcpyHostToDevice, s));
    pM = dM.p;
    pI = dI.p;
  }
  SB_HIP(err.alloc(1));
  SB_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
  SB_HIP(sb::launch_mass_inverse(n_bodies, pM, pI, err.p, s));
  int e = 0;
  SB_HIP(hipMemcpyAsync(&e, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (!on_device) SB_HIP(hipMemcpyAsync(minv, dI.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  SB_HIP(hipStreamSynchronize(s));
  return e ? SB_ERR_INVALID : SB_OK;
}

int sb_memset_device(sb_ctx* ctx, void* dst, int value, size_t bytes) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!dst && bytes) return SB_ERR_INVALID;
  SB_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

int sb_synchronize(sb_ctx* ctx) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  SB_HIP(hipStreamSynchronize(ctx->stream));
  return SB_OK;
}

int sb_bench_copy(sb_ctx* ctx, size_t bytes, int reps, double* ms) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!ms || reps < 1 || bytes < 32) return SB_ERR_INVALID;
  bytes &= ~(size_t)31;
  DBuf<char> a, b;
  SB_HIP(a.alloc(bytes));
  SB_HIP(b.alloc(bytes));
  hipStream_t s = ctx->stream;
  SB_HIP(hipMemsetAsync(a.p, 0, bytes, s));
  SB_HIP(sb::launch_copy(a.p, b.p, bytes, s));  // warm-up
  SB_HIP(hipEventRecord(ctx->ev0, s));
  for (int k = 0; k < reps; ++k) SB_HIP(sb::launch_copy(a.p, b.p, bytes, s));
  SB_HIP(hipEventRecord(ctx->ev1, s));
  SB_HIP(hipEventSynchronize(ctx->ev1));
  float t = 0.f;
  SB_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = (double)t / reps;
  return SB_OK;
}

int sb_last_phase_times(sb_ctx* ctx, double* us4) {
  if (!ctx || !us4) return SB_ERR_INVALID;
  for (int k = 0; k < 4; ++k) us4[k] = ctx->phase_us[k];
  return SB_OK;
}

int sb_bench_spmv(sb_ctx* ctx, const sb_problem* Pc, int kind, int reps, double* ms) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  const bool flush = (kind & SB_BENCH_FLUSH_CACHES) != 0;
  kind &= ~SB_BENCH_FLUSH_CACHES;
  if (!Pc || !ms || reps < 1 || kind < 0 || kind > 3) return SB_ERR_INVALID;
  sb_problem* P = const_cast<sb_problem*>(Pc);
  if (P->dualOnly && kind != 0) return SB_ERR_UNSUPPORTED;
  if (kind == 0 && !P->hasW) return SB_ERR_INVALID;
  hipStream_t s = ctx->stream;
  const DevProblem<double> V = P->view(true);
  // deterministic non-trivial inputs
  {
    std::vector<double> x(3 * (size_t)P->nc), w(6 * (size_t)P->nb);
    for (size_t k = 0; k < x.size(); ++k) x[k] = std::sin(0.37 * (double)k);
    for (size_t k = 0; k < w.size(); ++k) w[k] = std::cos(0.11 * (double)k);
    SB_HIP(hipMemcpyAsync(P->X[0].p, x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, s));
    SB_HIP(hipMemcpyAsync(P->wbody.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, s));
  }
  auto launch = [&]() -> hipError_t {
    switch (kind) {
      case 0: return sb::launch_w(V, P->X[0].p, P->y.p, P->carry.p, s);
      case 1: return sb::launch_h(V, P->wbody.p, P->y.p, P->carry.p, s);
      case 2: return sb::launch_ht_minv(V, P->X[0].p, P->wbody.p, P->carry.p, s);
      default: return sb::launch_matfree(V, P->X[0].p, P->y.p, P->wbody.p, P->carry.p, s);
    }
  };
  SB_HIP(launch());  // warm-up
  if (flush) {
    // every launch timed alone after a copy of 2 x 384 MB, more than the 256 MB
    // Infinity Cache holds: the launch finds none of its data in the caches
    const size_t fbytes = (size_t)384 << 20;
    DBuf<char> fa, fb;
    SB_HIP(fa.alloc(fbytes));
    SB_HIP(fb.alloc(fbytes));
    SB_HIP(hipMemsetAsync(fa.p, 0, fbytes, s));
    double total = 0.0;
    for (int k = 0; k < reps; ++k) {
      SB_HIP(sb::launch_copy(fa.p, fb.p, fbytes, s));
      SB_HIP(hipEventRecord(ctx->ev0, s));
      SB_HIP(launch());
      SB_HIP(hipEventRecord(ctx->ev1, s));
      SB_HIP(hipEventSynchronize(ctx->ev1));
      float t = 0.f;
      SB_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
      total += t;
    }
    *ms = total / reps;
    return SB_OK;
  }
  SB_HIP(hipEventRecord(ctx->ev0, s));
  for (int k = 0; k < reps; ++k) SB_HIP(launch());
  SB_HIP(hipEventRecord(ctx->ev1, s));
  SB_HIP(hipEventSynchronize(ctx->ev1));
  float t = 0.f;
  SB_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  *ms = (double)t / reps;
  return SB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches

struct sb_batch {
  sb_ctx* ctx = nullptr;
  int n = 0;
  int64_t totalNc = 0;
  bool hasW = true;
  std::vector<std::unique_ptr<sb_problem>> probs;
  std::vector<int64_t> off;   // offset of problem k in the concatenated r (doubles)
  std::vector<int64_t> cost;  // work estimate per iteration: stored operator blocks
  std::vector<int> iterHint;  // iterations each problem took in the last full solve (0: unknown)
  DBuf<double> xio;           // [3 sum n_c]
  DBuf<double> outs;          // [2][n][16]: FP32 stage, FP64 stage
  DBuf<sb::PGProblem<float>> d32;
  DBuf<sb::PGProblem<double>> d64;
  DBuf<int> d_order;          // [n] problems grouped by class
  DBuf<int> bmap;             // [8 * CUs] class, team and rank of each block of the grid
  DBuf<unsigned> syncw;       // [kQueueWords + teams * kTeamSyncWords]: class queues, team barriers
  DBuf<double> teamPartials;  // [teams][2][kMaxTeamBlocks][kMaxPartials]
  DBuf<int> teamSlot;         // [teams]
  DBuf<int> error;
  // locality order: gperm[internal contact] = caller's contact (global indices); stage
  // holds r in the caller's order when it comes from host memory
  bool reordered = false;
  bool spreadTeams = false;   // SB_BATCH_SPREAD_TEAMS: teams on consecutive blocks
  DBuf<int> gperm;
  DBuf<double> stage;
  int maxTeams = 0;
  size_t maxSyncWords = 0;
  int descScaling = -1;       // scaling mode the descriptor arrays were built for
  bool tuned = false;
  sb_batch_config tunedCfg{0, 0, 0};
  // captured launch sequence and the plan it was captured for
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<unsigned char> graphKey;
  ~sb_batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

namespace {

constexpr unsigned kQueueWords = 32;  // class queue counters (one 128-byte line)
constexpr int kDefaultContactsPerBlock = 4000;

template <typename T, int NT>
const void* batch_fn(bool mf) {
  return mf ? (const void*)&sb::pg_batch_kernel<true, T, NT> : (const void*)&sb::pg_batch_kernel<false, T, NT>;
}
template <typename T>
const void* batch_fn(bool mf, int threads) {
  switch (threads) {
    case 256: return batch_fn<T, 256>(mf);
    case 512: return batch_fn<T, 512>(mf);
    case 1024: return batch_fn<T, 1024>(mf);
    default: return nullptr;
  }
}

int batch_occupancy(const void* fn, int threads) {
  int perCU = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, threads, 0) != hipSuccess || perCU < 1) perCU = 1;
  return perCU;
}

// Per-problem solver descriptors (device arrays); they depend only on the
// residual scaling mode.
int batch_descriptors(sb_batch* B, bool withScaling) {
  const int mode = withScaling ? 1 : 0;
  if (B->descScaling == mode) return SB_OK;
  std::vector<sb::PGProblem<float>> h32(B->n);
  std::vector<sb::PGProblem<double>> h64(B->n);
  for (int k = 0; k < B->n; ++k) {
    std::memset(&h32[k], 0, sizeof(h32[k]));
    std::memset(&h64[k], 0, sizeof(h64[k]));
    B->probs[k]->fill(h32[k], withScaling);
    B->probs[k]->fill(h64[k], withScaling);
    h32[k].x_io = h64[k].x_io = B->xio.p + B->off[k];
    h32[k].out = B->outs.p + (size_t)16 * k;
    h64[k].out = B->outs.p + (size_t)16 * (B->n + k);
  }
  // the arrays were allocated at creation: their addresses are captured in the graph
  hipStream_t s = B->ctx->stream;
  SB_HIP(hipMemcpyAsync(B->d32.p, h32.data(), sizeof(h32[0]) * B->n, hipMemcpyHostToDevice, s));
  SB_HIP(hipMemcpyAsync(B->d64.p, h64.data(), sizeof(h64[0]) * B->n, hipMemcpyHostToDevice, s));
  SB_HIP(hipStreamSynchronize(s));
  B->descScaling = mode;
  return SB_OK;
}

struct BatchPlan {
  bool mf, mixed;
  int threads, blocksPerCU, contactsPerBlock, grid;
  int nclass, nteams;
  int poolPerXcd;  // helper teams of class 0's size per XCD (0: none)
  sb::BatchClass cls[sb::kMaxBatchClasses];
  sb::PGParams prm32, prm64;
};

// Helper teams (blocks whose class has drained join teams of class 0's size on their
// XCD and take class 0's problems) unless built with SB_BATCH_HELPERS=0.
#ifndef SB_BATCH_HELPERS
#define SB_BATCH_HELPERS 1
#endif
constexpr int kXcdCount = 8;
// Team ids (sync words, partials, slots) of a plan: its own teams, then the helpers'.
int plan_team_ids(const BatchPlan& pl) { return pl.nteams + kXcdCount * pl.poolPerXcd; }
// Words of the batch's sync buffer: class queues, every team's barrier words, the
// helpers' join counters (one line per XCD) and states (one line per helper team).
size_t plan_sync_words(const BatchPlan& pl) {
  return kQueueWords + (size_t)plan_team_ids(pl) * sb::kTeamSyncWords +
         (size_t)kXcdCount * sb::kSyncStride * (1 + pl.poolPerXcd);
}

int pow2ceil(int64_t v) {
  int k = 1;
  while (k < v && k < (1 << 30)) k <<= 1;
  return k;
}

// Splits the problems into classes by team size and the grid into the classes'
// block ranges, in proportion to each class's work; fills order (problems by
// class, most expensive first).
void batch_classes(const sb_batch* B, BatchPlan& pl, std::vector<int>& order) {
  const int G = pl.grid;
  int kcap = 1;  // largest team: a power of two, at most half the grid
  while (kcap * 2 <= std::min(sb::kMaxTeamBlocks, G / 2)) kcap *= 2;
  // iteration weight of each problem relative to the median, from the last full
  // solve of this batch (1 when unknown)
  std::vector<double> wIt(B->n, 1.0);
  {
    std::vector<int> h;
    for (int v : B->iterHint)
      if (v > 0) h.push_back(v);
    if ((int)h.size() == B->n) {
      std::nth_element(h.begin(), h.begin() + h.size() / 2, h.end());
      const double med = std::max(1, h[h.size() / 2]);
      for (int p = 0; p < B->n; ++p) wIt[p] = std::max(0.25, B->iterHint[p] / med);
    }
  }
  std::vector<int> kOf(B->n);
  for (int p = 0; p < B->n; ++p) {
    int k = 1;
    if (pl.contactsPerBlock > 0) {
      const int64_t need = (int64_t)std::ceil(B->probs[p]->nc * wIt[p] / pl.contactsPerBlock);
      k = std::min(pow2ceil(need), kcap);
    }
    kOf[p] = k;
  }
  // A team's time per iteration is a latency floor (barriers and dependent memory
  // round trips, about as long as streaming kLatencyCost stored blocks per
  // workgroup) plus its share of the problem's blocks: a problem's duration on its
  // team is w (floor + cost / k).
  constexpr double kLatencyCost = 4000.0;
  auto durationOf = [&](int p) { return wIt[p] * (kLatencyCost + (double)B->cost[p] / kOf[p]); };
  // classes in decreasing team size
  std::vector<int> ks;
  for (int k = sb::kMaxTeamBlocks; k >= 1; k >>= 1)
    for (int p = 0; p < B->n; ++p)
      if (kOf[p] == k) {
        ks.push_back(k);
        break;
      }
  const int C = (int)ks.size();
  std::vector<int> count(C, 0);
  std::vector<std::vector<double>> dur(C);
  for (int p = 0; p < B->n; ++p) {
    const int c = (int)(std::find(ks.begin(), ks.end(), kOf[p]) - ks.begin());
    dur[c].push_back(durationOf(p));
    ++count[c];
  }
  for (auto& d : dur) std::sort(d.begin(), d.end(), std::greater<double>());
  // time until class c's queue drains on n teams: longest first, each problem to the
  // team that frees up first (what the queue does)
  auto makespan = [&](int c, int n) {
    std::priority_queue<double, std::vector<double>, std::greater<double>> free;
    for (int i = 0; i < n; ++i) free.push(0.0);
    double m = 0.0;
    for (double d : dur[c]) {
      const double t = free.top() + d;
      free.pop();
      free.push(t);
      m = std::max(m, t);
    }
    return m;
  };
  // teams per class: one each (their sizes are distinct powers of two, at most half
  // the grid, so they fit), then one more at a time for the class that would drain
  // last while it has problems to spare and blocks remain; blocks left over then go
  // to the other classes, longest drain first
  std::vector<int> teams(C, 1);
  std::vector<double> span(C);
  int used = 0;
  for (int c = 0; c < C; ++c) {
    used += ks[c];
    span[c] = makespan(c, 1);
  }
  for (;;) {
    const int c = (int)(std::max_element(span.begin(), span.end()) - span.begin());
    if (teams[c] >= count[c] || used + ks[c] > G) break;
    ++teams[c];
    used += ks[c];
    span[c] = makespan(c, teams[c]);
  }
  for (bool grew = true; grew;) {
    grew = false;
    int cbest = -1;
    for (int c = 0; c < C; ++c)
      if (teams[c] < count[c] && used + ks[c] <= G && (cbest < 0 || span[c] > span[cbest])) cbest = c;
    if (cbest >= 0) {
      ++teams[cbest];
      used += ks[cbest];
      span[cbest] = makespan(cbest, teams[cbest]);
      grew = true;
    }
  }
  order.clear();
  int b0 = 0, t0 = 0;
  pl.nclass = C;
  for (int c = 0; c < C; ++c) {
    sb::BatchClass& cl = pl.cls[c];
    cl.b0 = b0;
    cl.k = ks[c];
    cl.nteams = teams[c];
    cl.q0 = (int)order.size();
    std::vector<int> mine;
    for (int p = 0; p < B->n; ++p)
      if (kOf[p] == ks[c]) mine.push_back(p);
    std::stable_sort(mine.begin(), mine.end(), [&](int a, int b) { return durationOf(a) > durationOf(b); });
    order.insert(order.end(), mine.begin(), mine.end());
    cl.q1 = (int)order.size();
    cl.team0 = t0;
    b0 += teams[c] * ks[c];
    t0 += teams[c];
  }
  pl.nteams = t0;  // the grid stays whole: blocks no team uses leave at once
}

// Places the teams on the grid (sbhost::batchBlockMap): the k blocks of a team on
// one XCD; the batch kernel checks this on the device and falls back to
// L2-writing-back barriers for a team that is not.
constexpr int kXcds = kXcdCount;
std::vector<int> batch_block_map(const BatchPlan& pl, bool spread) {
  int k[sb::kMaxBatchClasses], nt[sb::kMaxBatchClasses];
  for (int c = 0; c < pl.nclass; ++c) {
    k[c] = pl.cls[c].k;
    nt[c] = pl.cls[c].nteams;
  }
  return sbhost::batchBlockMap(pl.grid, kXcds, pl.nclass, k, nt, spread);
}

int batch_plan(sb_batch* B, const sb_pg_options* o, const sb_batch_config* cfg, BatchPlan& pl,
               std::vector<int>& order) {
  sb_ctx* ctx = B->ctx;
  pl.mf = (o->op == SB_OP_MATRIX_FREE);
  pl.mixed = (o->precision == SB_PRECISION_MIXED);
  sb_batch_config c{0, 0, 0};
  if (cfg) c = *cfg;
  if (c.block_threads == 0 && B->tuned) c.block_threads = B->tunedCfg.block_threads;
  if (c.blocks_per_cu == 0 && B->tuned) c.blocks_per_cu = B->tunedCfg.blocks_per_cu;
  if (c.contacts_per_block == 0 && B->tuned) c.contacts_per_block = B->tunedCfg.contacts_per_block;
  if (c.block_threads == 0) c.block_threads = 512;
  if (c.contacts_per_block == 0) c.contacts_per_block = kDefaultContactsPerBlock;
  const void* f64 = batch_fn<double>(pl.mf, c.block_threads);
  const void* f32 = batch_fn<float>(pl.mf, c.block_threads);
  if (!f64 || !f32) return SB_ERR_INVALID;
  // teams of several blocks need every block resident: stay within the occupancy of
  // every kernel the plan launches
  int occ = batch_occupancy(f64, c.block_threads);
  if (pl.mixed) occ = std::min(occ, batch_occupancy(f32, c.block_threads));
  if (c.blocks_per_cu == 0) c.blocks_per_cu = occ;
  if (c.blocks_per_cu < 1 || c.blocks_per_cu > 8) return SB_ERR_INVALID;
  c.blocks_per_cu = std::min(c.blocks_per_cu, occ);
  pl.threads = c.block_threads;
  pl.blocksPerCU = c.blocks_per_cu;
  pl.contactsPerBlock = c.contacts_per_block;
  pl.grid = c.blocks_per_cu * ctx->ncu;
  batch_classes(B, pl, order);
  pl.poolPerXcd = (SB_BATCH_HELPERS && pl.cls[0].k >= 2) ? (pl.grid / kXcdCount) / pl.cls[0].k : 0;
  pl.prm64 = make_params(o, pl.mf);
  pl.prm32 = pl.prm64;
  pl.prm32.tol = o->tol * (o->mixed_tol_factor > 0 ? o->mixed_tol_factor : 0.95);
  pl.prm32.stall_iters = o->mixed_stall_iters;
  return SB_OK;
}

int batch_buffers(sb_batch* B, const BatchPlan& pl) {
  const int nteams = plan_team_ids(pl);
  const size_t words = plan_sync_words(pl);
  if (nteams <= B->maxTeams && words <= B->maxSyncWords) return SB_OK;
  SB_HIP(B->syncw.alloc(words));
  SB_HIP(B->teamPartials.alloc((size_t)nteams * 2 * sb::kMaxTeamBlocks * sb::kMaxPartials));
  SB_HIP(B->teamSlot.alloc(nteams));
  B->maxTeams = nteams;
  B->maxSyncWords = words;
  // the graph holds the old addresses
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  B->graphKey.clear();
  return SB_OK;
}

template <typename T>
void batch_args(sb::PGBatchArgs<T>& a, const sb_batch* B, const BatchPlan& pl, const sb::PGParams& prm) {
  std::memset(&a, 0, sizeof(a));
  a.prm = prm;
  a.nprob = B->n;
  a.order = B->d_order.p;
  a.bmap = B->bmap.p;
  a.nclass = pl.nclass;
  for (int c = 0; c < pl.nclass; ++c) a.cls[c] = pl.cls[c];
  a.queue = B->syncw.p;
  a.team_sync = B->syncw.p + kQueueWords;
  a.team_partials = B->teamPartials.p;
  a.team_slot = B->teamSlot.p;
  a.error = B->error.p;
  a.timeout_ticks = 100ull * 1000 * 1000 * 20;  // 20 s at 100 MHz
  a.nteams = pl.nteams;
  a.pool_per_xcd = pl.poolPerXcd;
  a.pool_count = B->syncw.p + kQueueWords + (size_t)plan_team_ids(pl) * sb::kTeamSyncWords;
  a.pool_state = a.pool_count + (size_t)kXcdCount * sb::kSyncStride;
}

// Captures [reset, FP32 stage, reset, FP64 stage] (or [reset, FP64 stage]) into the
// batch's graph; reused while the plan is unchanged.
int batch_graph(sb_batch* B, const BatchPlan& pl) {
  std::vector<unsigned char> key(sizeof(BatchPlan));
  std::memcpy(key.data(), &pl, sizeof(BatchPlan));
  if (B->exec && key == B->graphKey) return SB_OK;
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  if (B->graph) (void)hipGraphDestroy(B->graph);
  B->exec = nullptr;
  B->graph = nullptr;
  hipStream_t s = B->ctx->stream;
  sb::PGBatchArgs<float> a32;
  sb::PGBatchArgs<double> a64;
  batch_args(a32, B, pl, pl.prm32);
  batch_args(a64, B, pl, pl.prm64);
  const sb::PGProblem<float>* p32 = B->d32.p;
  const sb::PGProblem<double>* p64 = B->d64.p;
  const int nwords = (int)plan_sync_words(pl);
  {
    const std::vector<int> bmap = batch_block_map(pl, B->spreadTeams);
    if (bmap.empty() || (int)bmap.size() > 8 * B->ctx->ncu) return SB_ERR_INVALID;
    SB_HIP(hipMemcpyAsync(B->bmap.p, bmap.data(), sizeof(int) * bmap.size(), hipMemcpyHostToDevice, s));
    SB_HIP(hipStreamSynchronize(s));
  }
  SB_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (pl.mixed) {
    e = sb::launch_reset_sync(B->syncw.p, nwords, B->error.p, s);
    void* args[] = {&a32, &p32};
    if (e == hipSuccess)
      e = hipLaunchKernel(batch_fn<float>(pl.mf, pl.threads), dim3(pl.grid), dim3(pl.threads), args, 0, s);
  }
  if (e == hipSuccess) e = sb::launch_reset_sync(B->syncw.p, nwords, B->error.p, s);
  if (e == hipSuccess) {
    void* args[] = {&a64, &p64};
    e = hipLaunchKernel(batch_fn<double>(pl.mf, pl.threads), dim3(pl.grid), dim3(pl.threads), args, 0, s);
  }
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(s, &g);
  SB_HIP(e);
  SB_HIP(e2);
  B->graph = g;
  SB_HIP(hipGraphInstantiate(&B->exec, B->graph, nullptr, nullptr, 0));
  B->graphKey = key;
  return SB_OK;
}

int batch_run(sb_ctx* ctx, sb_batch* B, const sb_pg_options* o, const sb_batch_config* cfg, double* r, int on_device,
              sb_batch_stats* st, bool updateHints) {
  int rc = check_pg_options(o, B->hasW, true);
  if (rc) return rc;
  BatchPlan pl;
  std::memset(&pl, 0, sizeof(pl));
  std::vector<int> order;
  if ((rc = batch_plan(B, o, cfg, pl, order))) return rc;
  const bool useScal = (o->scaling == SB_SCALING_DIAGONAL) || (o->scaling == SB_SCALING_REFERENCE && !pl.mf);
  if ((rc = batch_descriptors(B, useScal))) return rc;
  if ((rc = batch_buffers(B, pl))) return rc;
  if ((rc = batch_graph(B, pl))) return rc;
  hipStream_t s = ctx->stream;
  SB_HIP(hipMemcpyAsync(B->d_order.p, order.data(), sizeof(int) * B->n, hipMemcpyHostToDevice, s));
  const size_t n3 = 3 * (size_t)B->totalNc;
  if (r && B->reordered) {
    const double* src = r;
    if (!on_device) {
      if ((rc = copy_in(ctx, B->stage.p, r, n3, 0))) return rc;
      src = B->stage.p;
    }
    SB_HIP(sb::launch_permute3(B->totalNc, B->gperm.p, src, B->xio.p, false, s));
  } else if (r) {
    if ((rc = copy_in(ctx, B->xio.p, r, n3, on_device))) return rc;
  } else {
    SB_HIP(hipMemsetAsync(B->xio.p, 0, n3 * sizeof(double), s));
  }
  SB_HIP(hipEventRecord(ctx->ev0, s));
  SB_HIP(hipGraphLaunch(B->exec, s));
  ctx->plainLaunches += pl.mixed ? 2 : 1;  // occupancy-sized grids in the graph, launched plainly
  SB_HIP(hipEventRecord(ctx->ev1, s));
  std::vector<double> outs((size_t)32 * B->n);
  int err = 0;
  SB_HIP(hipMemcpyAsync(outs.data(), B->outs.p, sizeof(double) * outs.size(), hipMemcpyDeviceToHost, s));
  SB_HIP(hipMemcpyAsync(&err, B->error.p, sizeof(int), hipMemcpyDeviceToHost, s));
  if (r && B->reordered) {
    SB_HIP(sb::launch_permute3(B->totalNc, B->gperm.p, B->xio.p, on_device ? r : B->stage.p, true, s));
    if (!on_device && (rc = copy_out(ctx, r, B->stage.p, n3, 0))) return rc;
  } else if (r && (rc = copy_out(ctx, r, B->xio.p, n3, on_device))) {
    return rc;
  }
  SB_HIP(hipStreamSynchronize(s));
  int nconv = 0, totalIt = 0, maxIt = 0, refined = 0, failed = 0;
  int64_t spmvs = 0;
  for (int k = 0; k < B->n; ++k) {
    const double* o32 = &outs[(size_t)16 * k];
    const double* o64 = &outs[(size_t)16 * (B->n + k)];
    const int it = (int)o64[1] + (pl.mixed ? (int)o32[1] : 0);
    totalIt += it;
    maxIt = std::max(maxIt, it);
    spmvs += (int64_t)o64[2] + (pl.mixed ? (int64_t)o32[2] : 0);
    const bool conv = o64[5] != 0.0;
    nconv += conv ? 1 : 0;
    if (pl.mixed && o64[1] > 0) ++refined;
    if (o64[6] != 0.0 || (pl.mixed && o32[6] != 0.0) || !std::isfinite(o64[0])) ++failed;
    if (updateHints) B->iterHint[k] = std::max(1, it);
    if (st) {
      if (st->stage_spmvs) {
        st->stage_spmvs[2 * k] = pl.mixed ? (int)o32[2] : 0;
        st->stage_spmvs[2 * k + 1] = (int)o64[2];
      }
      if (st->iterations) st->iterations[k] = it;
      if (st->phase_us)
        for (int e = 0; e < 4; ++e) {
          st->phase_us[8 * k + e] = pl.mixed ? o32[7 + e] : 0.0;
          st->phase_us[8 * k + 4 + e] = o64[7 + e];
        }
      if (st->residuals) st->residuals[k] = o64[0];
      if (st->converged) st->converged[k] = conv ? 1 : 0;
    }
  }
  if (st) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    st->n_converged = nconv;
    st->total_iterations = totalIt;
    st->max_iterations = maxIt;
    st->refined_problems = refined;
    st->total_spmvs = spmvs;
    st->device_ms = ms;
    st->config.block_threads = pl.threads;
    st->config.blocks_per_cu = pl.blocksPerCU;
    st->config.contacts_per_block = pl.contactsPerBlock;
  }
  if (err) return SB_ERR_BARRIER_TIMEOUT;
  if (failed) return SB_ERR_NAN;
  return nconv == B->n ? SB_OK : SB_NOT_CONVERGED;
}

}  // namespace

extern "C" {

int sb_batch_create(sb_ctx* ctx, sb_batch** out, int n, const int* ncs, const int* nbs, const int32_t* const* rps,
                    const int32_t* const* cis, const double* const* hvs, const double* const* minvs,
                    const double* const* bs, const double* const* mus, int flags) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!out || n < 1 || !ncs || !nbs || !rps || !cis || !hvs || !minvs || !bs || !mus) return SB_ERR_INVALID;
  std::unique_ptr<sb_batch> B(new (std::nothrow) sb_batch());
  if (!B) return SB_ERR_OOM;
  B->ctx = ctx;
  B->n = n;
  B->hasW = (flags & SB_PROBLEM_ASSEMBLE_W) != 0;
  B->probs.resize(n);
  B->off.resize(n);
  B->cost.resize(n);
  B->iterHint.assign(n, 0);
  B->reordered = (flags & SB_BATCH_CALLER_ORDER) == 0;
  B->spreadTeams = (flags & SB_BATCH_SPREAD_TEAMS) != 0;
  const int pflags = flags & SB_PROBLEM_ASSEMBLE_W;
  std::vector<int> gperm;
  for (int k = 0; k < n; ++k) {
    if (!B->reordered) {
      if ((rc = create_problem(ctx, B->probs[k], ncs[k], nbs[k], rps[k], cis[k], hvs[k], minvs[k], bs[k], mus[k],
                               pflags)))
        return rc;
    } else {
      const int nc = ncs[k], nb = nbs[k];
      if (nc <= 0 || nb <= 0 || !rps[k] || !cis[k] || !hvs[k] || !minvs[k] || !bs[k] || !mus[k] || rps[k][0] != 0)
        return SB_ERR_INVALID;
      for (int i = 0; i < nc; ++i)
        if (rps[k][i + 1] < rps[k][i]) return SB_ERR_INVALID;
      for (int p = 0; p < rps[k][nc]; ++p)
        if (cis[k][p] < 0 || cis[k][p] >= nb) return SB_ERR_INVALID;
      std::vector<int> cperm, bmap;
      sbhost::localityOrder(nc, nb, rps[k], cis[k], cperm, bmap);
      std::vector<int32_t> rp(nc + 1, 0), ci;
      std::vector<double> hv, mi((size_t)36 * nb), b((size_t)3 * nc), mu(nc);
      ci.reserve(rps[k][nc]);
      hv.reserve((size_t)18 * rps[k][nc]);
      std::vector<std::pair<int, int>> row;
      for (int i = 0; i < nc; ++i) {
        const int o = cperm[i];
        row.clear();
        for (int p = rps[k][o]; p < rps[k][o + 1]; ++p) row.push_back({bmap[cis[k][p]], p});
        std::sort(row.begin(), row.end());
        for (auto& e : row) {
          ci.push_back(e.first);
          hv.insert(hv.end(), hvs[k] + (size_t)18 * e.second, hvs[k] + (size_t)18 * e.second + 18);
        }
        rp[i + 1] = (int32_t)ci.size();
        for (int c = 0; c < 3; ++c) b[3 * (size_t)i + c] = bs[k][3 * (size_t)o + c];
        mu[i] = mus[k][o];
        gperm.push_back((int)(B->totalNc + o));
      }
      for (int j = 0; j < nb; ++j)
        std::memcpy(&mi[(size_t)36 * bmap[j]], minvs[k] + (size_t)36 * j, 36 * sizeof(double));
      if ((rc = create_problem(ctx, B->probs[k], nc, nb, rp.data(), ci.data(), hv.data(), mi.data(), b.data(),
                               mu.data(), pflags)))
        return rc;
    }
    B->off[k] = 3 * B->totalNc;
    B->totalNc += ncs[k];
    const sb_problem* P = B->probs[k].get();
    B->cost[k] = (int64_t)P->H.host.nblocks * 2 + P->HT.host.nblocks + (B->hasW ? P->W.host.nblocks : 0);
    if ((rc = factor_prepare(B->probs[k].get()))) return rc;
  }
  hipStream_t s = ctx->stream;
  SB_HIP(B->d_order.alloc(n));
  SB_HIP(B->bmap.alloc((size_t)8 * ctx->ncu));
  SB_HIP(B->xio.alloc(3 * (size_t)B->totalNc));
  SB_HIP(B->outs.alloc((size_t)32 * n));
  SB_HIP(B->error.alloc(1));
  SB_HIP(B->d32.alloc(n));
  SB_HIP(B->d64.alloc(n));
  if (B->reordered) {
    SB_HIP(B->gperm.upload(gperm, s));
    SB_HIP(B->stage.alloc(3 * (size_t)B->totalNc));
  }
  SB_HIP(hipStreamSynchronize(s));
  *out = B.release();
  return SB_OK;
}

int sb_batch_destroy(sb_batch* B) {
  if (!B) return SB_ERR_INVALID;
  (void)hipSetDevice(B->ctx->device);
  delete B;
  return SB_OK;
}

int sb_batch_sizes(const sb_batch* B, int* n, int64_t* total) {
  if (!B) return SB_ERR_INVALID;
  if (n) *n = B->n;
  if (total) *total = B->totalNc;
  return SB_OK;
}

int sb_batch_set_iteration_hints(sb_batch* B, const int* hints) {
  if (!B) return SB_ERR_INVALID;
  if (!hints) {
    B->iterHint.assign(B->n, 0);
    return SB_OK;
  }
  for (int k = 0; k < B->n; ++k)
    if (hints[k] < 0) return SB_ERR_INVALID;
  B->iterHint.assign(hints, hints + B->n);
  return SB_OK;
}

int sb_batch_get_iteration_hints(const sb_batch* B, int* hints) {
  if (!B || !hints) return SB_ERR_INVALID;
  std::copy(B->iterHint.begin(), B->iterHint.end(), hints);
  return SB_OK;
}

int sb_batch_solve(sb_ctx* ctx, sb_batch* B, const sb_pg_options* o, const sb_batch_config* cfg, double* r,
                   int on_device, sb_batch_stats* st) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!B || !o || !r || B->ctx != ctx) return SB_ERR_INVALID;
  // the plan (and so every problem's arithmetic) changes only through sb_batch_autotune
  // and sb_batch_set_iteration_hints
  return batch_run(ctx, B, o, cfg, r, on_device, st, false);
}

int sb_batch_autotune(sb_ctx* ctx, sb_batch* B, const sb_pg_options* o, int probe_iters, sb_batch_config* best,
                      double* best_ms) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!B || !o || B->ctx != ctx || probe_iters < 1) return SB_ERR_INVALID;
  const bool mf = (o->op == SB_OP_MATRIX_FREE);
  const bool mixed = (o->precision == SB_PRECISION_MIXED);
  auto timeRun = [&](const sb_pg_options* opts, const sb_batch_config& c, bool hints, double& ms) -> int {
    ms = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      sb_batch_stats st;
      std::memset(&st, 0, sizeof(st));
      const int r = batch_run(ctx, B, opts, &c, nullptr, 0, &st, hints && rep == 0);
      if (r < 0) return r;
      ms = std::min(ms, st.device_ms);
    }
    return SB_OK;
  };
  // 1. screen every geometry on probe_iters iterations per problem (the per-iteration
  //    cost of each launch shape)
  sb_pg_options po = *o;
  po.max_iters = std::min(probe_iters, o->max_iters);
  std::vector<std::pair<double, sb_batch_config>> screened;
  const int threadsList[3] = {256, 512, 1024};
  const int cpbList[7] = {-1, 8000, 4000, 2000, 1000, 750, 500};
  for (int th : threadsList) {
    int occ = batch_occupancy(batch_fn<double>(mf, th), th);
    if (mixed) occ = std::min(occ, batch_occupancy(batch_fn<float>(mf, th), th));
    for (int bpc = 1; bpc <= std::min(occ, 8); ++bpc) {
      if (th * bpc > 2048) continue;  // at most 32 waves per CU
      for (int cpb : cpbList) {
        const sb_batch_config c{th, bpc, cpb};
        double ms;
        if ((rc = timeRun(&po, c, false, ms))) return rc;
        screened.push_back({ms, c});
      }
    }
  }
  std::sort(screened.begin(), screened.end(),
            [](const std::pair<double, sb_batch_config>& a, const std::pair<double, sb_batch_config>& b) {
              return a.first < b.first;
            });
  // 2. one full solve with the best screened shape records every problem's
  //    iteration count (long-running problems then get larger teams); the best
  //    shape of each team sizing is then timed on full solves
  double ms0;
  if ((rc = timeRun(o, screened[0].second, true, ms0))) return rc;
  double bestMs = 1e300;
  sb_batch_config bestCfg = screened[0].second;
  for (int cpb : cpbList) {  // the two best screened shapes of each team sizing
    int taken = 0;
    for (const auto& sc : screened) {
      if (sc.second.contacts_per_block != cpb) continue;
      double ms;
      if ((rc = timeRun(o, sc.second, false, ms))) return rc;
      if (ms < bestMs) {
        bestMs = ms;
        bestCfg = sc.second;
      }
      if (++taken == 2) break;
    }
  }
  B->tuned = true;
  B->tunedCfg = bestCfg;
  if (best) *best = bestCfg;
  if (best_ms) *best_ms = bestMs;
  return SB_OK;
}

}  // extern "C"
