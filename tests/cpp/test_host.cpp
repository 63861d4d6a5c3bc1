// Host-side layout code (so-bogus_amd/csrc/sb_host.cpp) checked on the CPU:
// tiling invariants and an emulation of the kernels' segmented wave scan
// (sb_tiles.hpp) at each tile's recorded depth, BSR transpose, symmetric
// expansion, W = H Minv H^T assembly and its diagonal against dense products,
// and the two colorings. Integer-valued data, so every sum is exact and the
// comparisons are bit-exact. Prints "ok <n checks>" and exits 0, or names the
// first failed check and exits 1.
#include <algorithm>
#include <map>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "sb_host.hpp"

using sbhost::Tiles;

static int g_checks = 0;
#define CHECK(cond, ...)                                              \
  do {                                                                \
    ++g_checks;                                                       \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "FAILED %s:%d: %s: ", __FILE__, __LINE__, #cond); \
      std::fprintf(stderr, __VA_ARGS__);                              \
      std::fprintf(stderr, "\n");                                     \
      std::exit(1);                                                   \
    }                                                                 \
  } while (0)

struct Bsr {
  int nrows, ncols, br, bc;
  std::vector<int> rowptr, colidx;
  std::vector<double> vals;  // row-major per block
};

// Random BSR with the given row lengths (distinct sorted columns), integer values.
static Bsr randomBsr(std::mt19937& g, int ncols, int br, int bc, const std::vector<int>& lens) {
  Bsr A{(int)lens.size(), ncols, br, bc, {0}, {}, {}};
  std::uniform_int_distribution<int> val(-8, 8);
  for (int len : lens) {
    std::vector<int> cols(ncols);
    for (int j = 0; j < ncols; ++j) cols[j] = j;
    std::shuffle(cols.begin(), cols.end(), g);
    cols.resize(len);
    std::sort(cols.begin(), cols.end());
    for (int j : cols) {
      A.colidx.push_back(j);
      for (int e = 0; e < br * bc; ++e) A.vals.push_back(val(g));
    }
    A.rowptr.push_back((int)A.colidx.size());
  }
  return A;
}

static std::vector<double> denseMul(const Bsr& A, const std::vector<double>& x) {
  std::vector<double> y((size_t)A.nrows * A.br, 0.0);
  for (int i = 0; i < A.nrows; ++i)
    for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p)
      for (int r = 0; r < A.br; ++r)
        for (int c = 0; c < A.bc; ++c)
          y[(size_t)i * A.br + r] += A.vals[(size_t)p * A.br * A.bc + r * A.bc + c] * x[(size_t)A.colidx[p] * A.bc + c];
  return y;
}

// y = A x through the tiles, emulating the wave: lane l of tile t holds block
// tstart[t] + l; a segmented shfl_up scan over `steps` doublings; the row's last
// lane is its leader; long tiles leave a carry per tile that is summed per row.
static std::vector<double> tiledMul(const Tiles& T, const std::vector<double>& x) {
  const int W = 64, BR = T.br, BC = T.bc;
  std::vector<double> y((size_t)T.nrows * BR, 0.0);
  std::vector<int> written(T.nrows, 0);
  std::vector<double> carry((size_t)T.ntiles * BR, 0.0);
  for (int t = 0; t < T.ntiles; ++t) {
    const int q0 = T.tstart[t], q1 = T.tstart[t + 1];
    CHECK(q1 > q0 && q1 - q0 <= W, "tile %d spans %d blocks", t, q1 - q0);
    std::vector<int> row(W, -1);
    std::vector<std::vector<double>> v(W, std::vector<double>(BR, 0.0));
    for (int l = 0; l < W; ++l) {
      const int q = q0 + l;
      if (q >= q1) continue;
      row[l] = T.brow[q];
      for (int r = 0; r < BR; ++r)
        for (int c = 0; c < BC; ++c)
          v[l][r] += T.vals[(size_t)(r * BC + c) * T.nblocks + q] * x[(size_t)T.bcol[q] * BC + c];
    }
    const int tl = T.tlong[t];
    const int steps = tl >= 0 ? 6 : -tl - 1;
    for (int s = 0; s < steps; ++s) {
      const int d = 1 << s;
      std::vector<std::vector<double>> vu = v;  // values before this step (all lanes shuffle at once)
      for (int l = W - 1; l >= 0; --l)
        if (l >= d && row[l - d] == row[l])
          for (int r = 0; r < BR; ++r) v[l][r] += vu[l - d][r];
    }
    for (int l = 0; l < W; ++l) {
      if (row[l] < 0) continue;
      const int rn = l == W - 1 ? row[l] : row[l + 1];
      if (!(l == W - 1 || rn != row[l])) continue;
      if (tl >= 0) {
        for (int r = 0; r < BR; ++r) carry[(size_t)t * BR + r] = v[l][r];
      } else {
        CHECK(!written[row[l]], "row %d has two leaders", row[l]);
        written[row[l]] = 1;
        for (int r = 0; r < BR; ++r) y[(size_t)row[l] * BR + r] = v[l][r];
      }
    }
  }
  for (size_t k = 0; k < T.long_row.size(); ++k) {
    const int i = T.long_row[k];
    CHECK(!written[i], "long row %d also led by a whole-row tile", i);
    written[i] = 1;
    for (int t = T.long_t0[k]; t < T.long_t1[k]; ++t)
      for (int r = 0; r < BR; ++r) y[(size_t)i * BR + r] += carry[(size_t)t * BR + r];
  }
  for (int i = 0; i < T.nrows; ++i) CHECK(written[i], "row %d has no leader", i);
  return y;
}

static void checkTiles(const Bsr& A, const Tiles& T, bool transposed) {
  const int W = 64;
  const int br = transposed ? A.bc : A.br, bc = transposed ? A.br : A.bc;
  CHECK(T.tstart.front() == 0 && T.tstart.back() == T.nblocks, "tile bounds %d..%d of %d", T.tstart.front(),
        T.tstart.back(), T.nblocks);
  CHECK((int)T.tlong.size() == T.ntiles && (int)T.tstart.size() == T.ntiles + 1, "tile arrays");
  // every row present once, empty rows as one zero block
  for (int i = 0; i < A.nrows; ++i) {
    const int len = A.rowptr[i + 1] - A.rowptr[i];
    CHECK(T.rowptr[i + 1] - T.rowptr[i] == std::max(len, 1), "row %d length", i);
    for (int q = T.rowptr[i]; q < T.rowptr[i + 1]; ++q) {
      CHECK(T.brow[q] == i, "block %d row", q);
      const int p = A.rowptr[i] + (q - T.rowptr[i]);
      for (int r = 0; r < br; ++r)
        for (int c = 0; c < bc; ++c) {
          const double want = len == 0 ? 0.0
                              : transposed ? A.vals[(size_t)p * br * bc + c * br + r]
                                           : A.vals[(size_t)p * br * bc + r * bc + c];
          CHECK(T.vals[(size_t)(r * bc + c) * T.nblocks + q] == want, "value (%d,%d) of block %d", r, c, q);
        }
      if (len > 0) CHECK(T.bcol[q] == A.colidx[p], "block %d column", q);
    }
  }
  // whole-row tiles hold whole rows and record a scan depth that covers their longest row
  for (int t = 0; t < T.ntiles; ++t) {
    const int q0 = T.tstart[t], q1 = T.tstart[t + 1];
    if (T.tlong[t] >= 0) {
      const int k = T.tlong[t];
      CHECK(k < (int)T.long_row.size() && t >= T.long_t0[k] && t < T.long_t1[k], "long tile %d", t);
      CHECK(T.brow[q0] == T.long_row[k] && T.brow[q1 - 1] == T.long_row[k], "long tile %d rows", t);
      continue;
    }
    CHECK(q0 == T.rowptr[T.brow[q0]] && q1 == T.rowptr[T.brow[q1 - 1] + 1], "tile %d cuts a row", t);
    int longest = 0;
    for (int q = q0; q < q1; ++q) longest = std::max(longest, T.rowptr[T.brow[q] + 1] - T.rowptr[T.brow[q]]);
    const int steps = -T.tlong[t] - 1;
    CHECK(steps >= 0 && steps <= 6 && (1 << steps) >= longest && (steps == 0 || (1 << (steps - 1)) < longest),
          "tile %d depth %d for a row of %d", t, steps, longest);
  }
  for (size_t k = 0; k < T.long_row.size(); ++k) {
    const int i = T.long_row[k];
    CHECK(T.rowptr[i + 1] - T.rowptr[i] > W, "long row %d is short", i);
    CHECK(T.tstart[T.long_t0[k]] == T.rowptr[i] && T.tstart[T.long_t1[k]] == T.rowptr[i + 1], "long row %d tiles", i);
  }
  // products through the emulated wave scan equal the dense product
  std::mt19937 g(7);
  std::uniform_int_distribution<int> val(-4, 4);
  std::vector<double> x((size_t)T.ncols * bc);
  for (auto& e : x) e = val(g);
  std::vector<double> y = tiledMul(T, x);
  Bsr E{A.nrows, A.ncols, br, bc, A.rowptr, A.colidx, A.vals};  // the blocks the tiles hold
  if (transposed)
    for (size_t p = 0; p < A.colidx.size(); ++p)
      for (int r = 0; r < br; ++r)
        for (int c = 0; c < bc; ++c) E.vals[p * br * bc + r * bc + c] = A.vals[p * br * bc + c * br + r];
  const std::vector<double> want = denseMul(E, x);
  for (size_t e = 0; e < y.size(); ++e) CHECK(y[e] == want[e], "product entry %zu: %g vs %g", e, y[e], want[e]);
}

static void testTiling() {
  std::mt19937 g(1234);
  // rows of every length class: empty, 1, 2 (contact rows of H), up to 64, just over 64, several tiles long
  std::vector<std::vector<int>> patterns = {
      {2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2},
      {0, 1, 0, 0, 3, 64, 0, 65, 1, 2, 200, 0, 7, 33, 31, 128, 129, 5},
      {64, 64, 63, 1, 0},
      {0},
      {300},
  };
  std::uniform_int_distribution<int> len(0, 12), rare(0, 40);
  std::vector<int> mixed;
  for (int i = 0; i < 3000; ++i) mixed.push_back(rare(g) == 0 ? 60 + len(g) * 20 : len(g));
  patterns.push_back(mixed);
  for (const auto& lens : patterns) {
    const int ncols = 400;
    for (auto [br, bc] : {std::pair<int, int>{3, 6}, {6, 3}, {3, 3}}) {
      Bsr A = randomBsr(g, ncols, br, bc, lens);
      Tiles T = sbhost::buildTiles(A.nrows, A.ncols, br, bc, A.rowptr.data(), A.colidx.data(), A.vals.data());
      checkTiles(A, T, false);
    }
  }
  // transposeBlocks: values stored as the transposes of the source blocks
  Bsr A = randomBsr(g, 50, 6, 3, {1, 2, 0, 70, 3});
  Tiles T = sbhost::buildTiles(A.nrows, A.ncols, 3, 6, A.rowptr.data(), A.colidx.data(), A.vals.data(), true);
  checkTiles(A, T, true);
}

static void testTranspose() {
  std::mt19937 g(99);
  Bsr A = randomBsr(g, 37, 3, 6, {2, 0, 5, 1, 37, 2, 2, 9});
  std::vector<int> trp, tci, rrp, rci;
  std::vector<double> tv, rv;
  sbhost::transposeBSR(A.nrows, A.ncols, 3, 6, A.rowptr.data(), A.colidx.data(), A.vals.data(), trp, tci, tv);
  // (A^T)^T == A
  sbhost::transposeBSR(A.ncols, A.nrows, 6, 3, trp.data(), tci.data(), tv.data(), rrp, rci, rv);
  CHECK(rrp == A.rowptr && rci == A.colidx && rv == A.vals, "double transpose");
  // y^T (A x) == (A^T y)^T x
  Bsr At{A.ncols, A.nrows, 6, 3, trp, tci, tv};
  std::vector<double> x(37 * 6), y(8 * 3);
  std::uniform_int_distribution<int> val(-5, 5);
  for (auto& e : x) e = val(g);
  for (auto& e : y) e = val(g);
  const std::vector<double> Ax = denseMul(A, x), Aty = denseMul(At, y);
  double l = 0, r = 0;
  for (size_t e = 0; e < y.size(); ++e) l += y[e] * Ax[e];
  for (size_t e = 0; e < x.size(); ++e) r += Aty[e] * x[e];
  CHECK(l == r, "adjoint identity %g vs %g", l, r);
  for (int j = 0; j < A.ncols; ++j)
    for (int p = trp[j]; p + 1 < trp[j + 1]; ++p) CHECK(tci[p] < tci[p + 1], "transposed row %d unsorted", j);
}

// Dense n x n (block size bs) of a BSR.
static std::vector<double> dense(int n, int bs, const std::vector<int>& rp, const std::vector<int>& ci,
                                 const std::vector<double>& v) {
  std::vector<double> D((size_t)n * bs * n * bs, 0.0);
  for (int i = 0; i < n; ++i)
    for (int p = rp[i]; p < rp[i + 1]; ++p)
      for (int r = 0; r < bs; ++r)
        for (int c = 0; c < bs; ++c)
          D[(size_t)(i * bs + r) * n * bs + ci[p] * bs + c] += v[(size_t)p * bs * bs + r * bs + c];
  return D;
}

static void testAssembly() {
  std::mt19937 g(2024);
  const int nb = 30, nc = 70;
  // contacts join two distinct bodies (H rows of two 3x6 blocks), a few hub bodies
  Bsr H{nc, nb, 3, 6, {0}, {}, {}};
  std::uniform_int_distribution<int> body(0, nb - 1), hub(0, 2), coin(0, 2), val(-3, 3);
  for (int i = 0; i < nc; ++i) {
    int a = coin(g) == 0 ? hub(g) : body(g), b = body(g);
    while (b == a) b = body(g);
    if (a > b) std::swap(a, b);
    for (int j : {a, b}) {
      H.colidx.push_back(j);
      for (int e = 0; e < 18; ++e) H.vals.push_back(val(g));
    }
    H.rowptr.push_back((int)H.colidx.size());
  }
  std::vector<double> minv((size_t)nb * 36);
  for (int k = 0; k < nb; ++k)  // symmetric integer blocks
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c <= r; ++c) minv[k * 36 + r * 6 + c] = minv[k * 36 + c * 6 + r] = val(g);
  std::vector<int> wrp, wci, frp, fci;
  std::vector<double> wv, fv, diag;
  sbhost::assembleW(nc, nb, H.rowptr.data(), H.colidx.data(), H.vals.data(), minv.data(), wrp, wci, wv);
  for (int i = 0; i < nc; ++i)
    for (int p = wrp[i]; p < wrp[i + 1]; ++p) {
      CHECK(wci[p] <= i, "W block (%d,%d) above the diagonal", i, wci[p]);
      if (p > wrp[i]) CHECK(wci[p - 1] < wci[p], "W row %d unsorted", i);
    }
  sbhost::expandSymmetric(nc, 3, wrp.data(), wci.data(), wv.data(), frp, fci, fv);
  const std::vector<double> Wd = dense(nc, 3, frp, fci, fv);
  // dense H Minv H^T
  const int n3 = 3 * nc, n6 = 6 * nb;
  std::vector<double> Hd((size_t)n3 * n6, 0.0), Md((size_t)n6 * n6, 0.0);
  for (int i = 0; i < nc; ++i)
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 6; ++c) Hd[(size_t)(3 * i + r) * n6 + 6 * H.colidx[p] + c] = H.vals[p * 18 + r * 6 + c];
  for (int k = 0; k < nb; ++k)
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 6; ++c) Md[(size_t)(6 * k + r) * n6 + 6 * k + c] = minv[k * 36 + r * 6 + c];
  int nnzBlocks = 0;
  for (int i = 0; i < nc; ++i)
    for (int j = 0; j < nc; ++j) {
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0;
          for (int a = 0; a < n6; ++a) {
            if (Hd[(size_t)(3 * i + r) * n6 + a] == 0) continue;
            for (int b = 0; b < n6; ++b) acc += Hd[(size_t)(3 * i + r) * n6 + a] * Md[(size_t)a * n6 + b] * Hd[(size_t)(3 * j + c) * n6 + b];
          }
          CHECK(Wd[(size_t)(3 * i + r) * n3 + 3 * j + c] == acc, "W(%d,%d)[%d,%d] = %g vs %g", i, j, r, c,
                Wd[(size_t)(3 * i + r) * n3 + 3 * j + c], acc);
        }
      // a block is stored iff the contacts share a body
      bool share = false;
      for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p)
        for (int q = H.rowptr[j]; q < H.rowptr[j + 1]; ++q) share |= H.colidx[p] == H.colidx[q];
      if (j <= i && share) ++nnzBlocks;
    }
  CHECK(wrp[nc] == nnzBlocks, "W holds %d lower blocks, %d contact pairs share a body", wrp[nc], nnzBlocks);
  sbhost::diagBlocks(nc, H.rowptr.data(), H.colidx.data(), H.vals.data(), minv.data(), diag);
  for (int i = 0; i < nc; ++i)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c)
        CHECK(diag[i * 9 + r * 3 + c] == Wd[(size_t)(3 * i + r) * n3 + 3 * i + c], "diag block %d", i);

  // colorings: a permutation, color ranges, no two coupled contacts in one color
  for (int mode : {0, 1}) {
    std::vector<int64_t> perm, colors;
    sbhost::coloring(nc, wrp.data(), wci.data(), mode, perm, colors);
    CHECK((int)perm.size() == nc && colors.front() == 0 && colors.back() == nc, "coloring %d sizes", mode);
    std::vector<int> colorOf(nc, -1);
    for (size_t c = 0; c + 1 < colors.size(); ++c) {
      CHECK(colors[c] < colors[c + 1], "coloring %d: empty color %zu", mode, c);
      for (int64_t q = colors[c]; q < colors[c + 1]; ++q) {
        CHECK(colorOf[perm[q]] == -1, "coloring %d: contact %lld twice", mode, (long long)perm[q]);
        colorOf[perm[q]] = (int)c;
        if (q > colors[c]) CHECK(perm[q - 1] < perm[q], "coloring %d: color %zu not ascending", mode, c);
      }
    }
    for (int i = 0; i < nc; ++i)
      for (int p = wrp[i]; p < wrp[i + 1]; ++p)
        if (wci[p] != i) CHECK(colorOf[i] != colorOf[wci[p]], "coloring %d: %d and %d coupled", mode, i, wci[p]);
    // the coloring from H's contact-body incidence is the same coloring
    std::vector<int64_t> hperm, hcolors;
    sbhost::coloringFromH(nc, nb, H.rowptr.data(), H.colidx.data(), mode, hperm, hcolors);
    CHECK(hperm == perm && hcolors == colors, "coloring %d from H differs from the one from W", mode);
  }
}

// Colorings from H equal the ones from the assembled pattern on larger contact
// graphs: hub bodies, contacts with one body and with none.
static void testColoringFromH() {
  std::mt19937 g(77);
  for (int trial = 0; trial < 4; ++trial) {
    const int nb = 300 + 200 * trial, nc = 2000 + 500 * trial;
    std::uniform_int_distribution<int> body(0, nb - 1), hub(0, 4), kind(0, 19);
    std::vector<int> rp{0}, ci;
    std::vector<double> hv;
    for (int i = 0; i < nc; ++i) {
      const int k = kind(g);
      std::vector<int> bs;
      if (k == 0) {
      } else if (k == 1) {
        bs = {body(g)};
      } else {
        int a = k < 6 ? hub(g) : body(g), b = body(g);
        while (b == a) b = body(g);
        bs = {std::min(a, b), std::max(a, b)};
      }
      for (int j : bs) {
        ci.push_back(j);
        hv.insert(hv.end(), 18, 1.0);
      }
      rp.push_back((int)ci.size());
    }
    std::vector<double> minv((size_t)nb * 36, 0.0);
    for (int k = 0; k < nb; ++k)
      for (int r = 0; r < 6; ++r) minv[(size_t)k * 36 + r * 7] = 1.0;
    std::vector<int> wrp, wci;
    std::vector<double> wv;
    sbhost::assembleW(nc, nb, rp.data(), ci.data(), hv.data(), minv.data(), wrp, wci, wv);
    for (int mode : {0, 1}) {
      std::vector<int64_t> perm, colors, hperm, hcolors;
      sbhost::coloring(nc, wrp.data(), wci.data(), mode, perm, colors);
      sbhost::coloringFromH(nc, nb, rp.data(), ci.data(), mode, hperm, hcolors);
      CHECK(hperm == perm && hcolors == colors, "trial %d mode %d: %zu vs %zu colors", trial, mode, hcolors.size(),
            colors.size());
    }
  }
}

// The locality order is a pair of permutations under which H M^-1 H^T is the same
// operator: (P W P^T)(P x) = P (W x), checked exactly on integer data.
static void testLocalityOrder() {
  std::mt19937 g(5);
  const int nb = 400, nc = 1500;
  std::uniform_int_distribution<int> body(0, nb - 1), kind(0, 29), val(-3, 3);
  std::vector<int> rp{0}, ci;
  std::vector<double> hv;
  for (int i = 0; i < nc; ++i) {
    const int k = kind(g);
    std::vector<int> bs;
    if (k == 1) bs = {body(g)};
    if (k > 1) {
      int a = body(g), b = body(g);
      while (b == a) b = body(g);
      bs = {std::min(a, b), std::max(a, b)};
    }
    for (int j : bs) {
      ci.push_back(j);
      for (int e = 0; e < 18; ++e) hv.push_back(val(g));
    }
    rp.push_back((int)ci.size());
  }
  std::vector<double> minv((size_t)nb * 36);
  for (auto& v : minv) v = val(g);
  std::vector<int> cperm, bmap;
  sbhost::localityOrder(nc, nb, rp.data(), ci.data(), cperm, bmap);
  std::vector<int> seen(nc, 0), bseen(nb, 0);
  for (int i : cperm) seen[i]++;
  for (int j : bmap) bseen[j]++;
  for (int i = 0; i < nc; ++i) CHECK(seen[i] == 1, "contact %d appears %d times", i, seen[i]);
  for (int j = 0; j < nb; ++j) CHECK(bseen[j] == 1, "body %d appears %d times", j, bseen[j]);
  // permuted H (columns renumbered and sorted within each row) and Minv
  std::vector<int> rp2{0}, ci2;
  std::vector<double> hv2;
  for (int i = 0; i < nc; ++i) {
    const int o = cperm[i];
    std::vector<std::pair<int, int>> row;
    for (int p = rp[o]; p < rp[o + 1]; ++p) row.push_back({bmap[ci[p]], p});
    std::sort(row.begin(), row.end());
    for (auto& e : row) {
      ci2.push_back(e.first);
      hv2.insert(hv2.end(), hv.begin() + (size_t)e.second * 18, hv.begin() + (size_t)e.second * 18 + 18);
    }
    rp2.push_back((int)ci2.size());
  }
  std::vector<double> minv2((size_t)nb * 36);
  for (int j = 0; j < nb; ++j)
    for (int e = 0; e < 36; ++e) minv2[(size_t)bmap[j] * 36 + e] = minv[(size_t)j * 36 + e];
  // consecutive contacts: lowest bodies non-decreasing
  int prevLo = -1;
  for (int i = 0; i < nc; ++i) {
    if (rp2[i + 1] == rp2[i]) continue;
    CHECK(ci2[rp2[i]] >= prevLo, "contact %d out of order", i);
    prevLo = ci2[rp2[i]];
  }
  auto applyW = [&](const std::vector<int>& r, const std::vector<int>& c, const std::vector<double>& h,
                    const std::vector<double>& m, const std::vector<double>& x) {
    std::vector<double> z((size_t)nb * 6, 0.0), w((size_t)nb * 6, 0.0), y((size_t)nc * 3, 0.0);
    for (int i = 0; i < nc; ++i)
      for (int p = r[i]; p < r[i + 1]; ++p)
        for (int a = 0; a < 6; ++a)
          for (int rr = 0; rr < 3; ++rr) z[(size_t)c[p] * 6 + a] += h[(size_t)p * 18 + rr * 6 + a] * x[3 * i + rr];
    for (int j = 0; j < nb; ++j)
      for (int a = 0; a < 6; ++a)
        for (int b2 = 0; b2 < 6; ++b2) w[(size_t)j * 6 + a] += m[(size_t)j * 36 + a * 6 + b2] * z[(size_t)j * 6 + b2];
    for (int i = 0; i < nc; ++i)
      for (int p = r[i]; p < r[i + 1]; ++p)
        for (int rr = 0; rr < 3; ++rr)
          for (int a = 0; a < 6; ++a) y[3 * i + rr] += h[(size_t)p * 18 + rr * 6 + a] * w[(size_t)c[p] * 6 + a];
    return y;
  };
  std::vector<double> x((size_t)nc * 3), x2((size_t)nc * 3);
  for (auto& v : x) v = val(g);
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < 3; ++k) x2[3 * i + k] = x[3 * cperm[i] + k];
  const std::vector<double> y = applyW(rp, ci, hv, minv, x), y2 = applyW(rp2, ci2, hv2, minv2, x2);
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < 3; ++k) CHECK(y2[3 * i + k] == y[3 * cperm[i] + k], "W x differs at contact %d", i);
}

// Batch block map: every block holds at most one (class, team, rank); every team has
// all its ranks, on blocks of one XCD (b % 8) when placed per XCD; the XCDs' blocks
// in use differ by at most the largest team; a spread map places teams on
// consecutive blocks.
static void testBatchBlockMap() {
  struct Case {
    int grid, nclass;
    std::vector<int> k, nteams;
  };
  const std::vector<Case> cases = {
      {256, 5, {16, 8, 4, 2, 1}, {3, 5, 7, 9, 100}},
      {512, 4, {16, 4, 2, 1}, {16, 16, 32, 128}},
      {256, 1, {1}, {256}},
      {768, 3, {16, 8, 1}, {20, 20, 200}},
      {1024, 2, {8, 1}, {64, 400}},
  };
  for (const Case& cs : cases) {
    for (bool spread : {false, true}) {
      const std::vector<int> bmap =
          sbhost::batchBlockMap(cs.grid, 8, cs.nclass, cs.k.data(), cs.nteams.data(), spread);
      CHECK((int)bmap.size() == cs.grid, "grid %d: map size %d", cs.grid, (int)bmap.size());
      std::map<std::pair<int, int>, std::vector<int>> teams;  // (class, team) -> blocks by rank
      for (int b = 0; b < (int)bmap.size(); ++b) {
        if (bmap[b] < 0) continue;
        const int c = bmap[b] >> 20, t = (bmap[b] >> 4) & 0xffff, r = bmap[b] & 15;
        CHECK(c < cs.nclass && t < cs.nteams[c] && r < cs.k[c], "block %d: bad code", b);
        auto& v = teams[{c, t}];
        if ((int)v.size() < cs.k[c]) v.resize(cs.k[c], -1);
        CHECK(v[r] < 0, "class %d team %d rank %d placed twice", c, t, r);
        v[r] = b;
      }
      std::vector<int> perXcd(8, 0);
      for (int c = 0; c < cs.nclass; ++c)
        for (int t = 0; t < cs.nteams[c]; ++t) {
          auto it = teams.find({c, t});
          CHECK(it != teams.end(), "class %d team %d missing", c, t);
          if (it == teams.end()) continue;
          const std::vector<int>& v = it->second;
          for (int r = 0; r < cs.k[c]; ++r) CHECK(v[r] >= 0, "class %d team %d rank %d missing", c, t, r);
          if (spread) {
            for (int r = 1; r < cs.k[c]; ++r) CHECK(v[r] == v[0] + r, "spread team not consecutive");
          } else {
            for (int r = 1; r < cs.k[c]; ++r) CHECK(v[r] % 8 == v[0] % 8, "team on two XCDs");
            perXcd[v[0] % 8] += cs.k[c];
          }
        }
      if (!spread) {
        const int mx = *std::max_element(perXcd.begin(), perXcd.end());
        const int mn = *std::min_element(perXcd.begin(), perXcd.end());
        CHECK(mx - mn <= cs.k[0], "XCD blocks in use %d..%d", mn, mx);
      }
    }
  }
  // more blocks requested than the grid has: empty map
  const int k1[1] = {16}, n1[1] = {17};
  CHECK(sbhost::batchBlockMap(256, 8, 1, k1, n1, false).empty(), "overfull map not rejected");
}

int main() {
  testBatchBlockMap();
  testLocalityOrder();
  testColoringFromH();
  testTiling();
  testTranspose();
  testAssembly();
  std::printf("ok %d checks\n", g_checks);
  return 0;
}
