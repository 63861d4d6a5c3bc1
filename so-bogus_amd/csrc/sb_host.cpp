// Host-side setup of the device layout (see sb_host.hpp).
#include "sb_host.hpp"

#include <algorithm>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace sbhost {

static int ceilLog2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}

Tiles buildTiles(int nrows, int ncols, int br, int bc, const int* rowptr, const int* colidx, const double* vals,
                 bool transposeBlocks) {
  const int W = 64;
  Tiles T;
  T.nrows = nrows;
  T.ncols = ncols;
  T.br = br;
  T.bc = bc;
  const int bsz = br * bc;
  // block list in row order, one zero block for each empty row
  std::vector<int> srcBlock;  // index into vals, -1 for zero blocks
  srcBlock.reserve(rowptr[nrows] + 16);
  T.rowptr.assign(nrows + 1, 0);
  for (int i = 0; i < nrows; ++i) {
    const int len = rowptr[i + 1] - rowptr[i];
    if (len == 0) {
      srcBlock.push_back(-1);
      T.brow.push_back(i);
      T.bcol.push_back(0);
    } else {
      for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
        srcBlock.push_back(p);
        T.brow.push_back(i);
        T.bcol.push_back(colidx[p]);
      }
    }
    T.rowptr[i + 1] = (int)srcBlock.size();
  }
  T.nblocks = (int)srcBlock.size();
  T.src = srcBlock;
  // tiles: whole rows packed greedily into <= 64 blocks; longer rows get their own tiles
  // A whole-row tile records -(1 + scan steps) for its own longest row; a long-row
  // tile records the long row's index.
  T.tstart.push_back(0);
  int maxLen = 1;
  int cur = 0;     // blocks in the open tile
  int curMax = 1;  // longest row in the open tile
  auto closeTile = [&](int end) {
    T.tstart.push_back(end);
    T.tlong.push_back(-1 - ceilLog2(curMax));
    cur = 0;
    curMax = 1;
  };
  for (int i = 0; i < nrows; ++i) {
    const int len = T.rowptr[i + 1] - T.rowptr[i];
    if (len > W) {
      if (cur > 0) closeTile(T.rowptr[i]);
      const int k = (int)T.long_row.size();
      T.long_row.push_back(i);
      T.long_t0.push_back((int)T.tstart.size() - 1);
      for (int s = T.rowptr[i]; s < T.rowptr[i + 1]; s += W) {
        T.tstart.push_back(std::min(s + W, T.rowptr[i + 1]));
        T.tlong.push_back(k);
      }
      T.long_t1.push_back((int)T.tstart.size() - 1);
      maxLen = W;
      continue;
    }
    if (cur + len > W) closeTile(T.rowptr[i]);
    cur += len;
    curMax = std::max(curMax, len);
    maxLen = std::max(maxLen, len);
  }
  if (cur > 0) closeTile(T.nblocks);
  T.ntiles = (int)T.tlong.size();
  T.scan_steps = ceilLog2(maxLen);
  // element-major values
  T.vals.assign((size_t)bsz * T.nblocks, 0.0);
  for (int q = 0; q < T.nblocks; ++q) {
    const int p = srcBlock[q];
    if (p < 0) continue;
    const double* src = vals + (size_t)p * bsz;
    for (int r = 0; r < br; ++r)
      for (int c = 0; c < bc; ++c) {
        const double v = transposeBlocks ? src[c * br + r] : src[r * bc + c];
        T.vals[(size_t)(r * bc + c) * T.nblocks + q] = v;
      }
  }
  return T;
}

void transposeBSR(int nrows, int ncols, int br, int bc, const int* rowptr, const int* colidx, const double* vals,
                  std::vector<int>& trowptr, std::vector<int>& tcolidx, std::vector<double>& tvals,
                  std::vector<int>* tsrc) {
  const int nnz = rowptr[nrows];
  trowptr.assign(ncols + 1, 0);
  for (int p = 0; p < nnz; ++p) trowptr[colidx[p] + 1]++;
  for (int j = 0; j < ncols; ++j) trowptr[j + 1] += trowptr[j];
  tcolidx.assign(nnz, 0);
  tvals.assign((size_t)nnz * br * bc, 0.0);
  if (tsrc) tsrc->assign(nnz, -1);
  std::vector<int> fill(trowptr.begin(), trowptr.end() - 1);
  for (int i = 0; i < nrows; ++i)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int j = colidx[p];
      const int q = fill[j]++;
      tcolidx[q] = i;
      if (tsrc) (*tsrc)[q] = p;
      const double* src = vals + (size_t)p * br * bc;
      double* dst = &tvals[(size_t)q * br * bc];
      for (int r = 0; r < br; ++r)
        for (int c = 0; c < bc; ++c) dst[c * br + r] = src[r * bc + c];
    }
}

void expandSymmetric(int n, int bs, const int* rowptr, const int* colidx, const double* vals, std::vector<int>& frowptr,
                     std::vector<int>& fcolidx, std::vector<double>& fvals) {
  const int b2 = bs * bs;
  std::vector<int> cnt(n, 0);
  for (int i = 0; i < n; ++i)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int j = colidx[p];
      cnt[i]++;
      if (j != i) cnt[j]++;
    }
  frowptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) frowptr[i + 1] = frowptr[i] + cnt[i];
  std::vector<std::pair<int, int>> ent(frowptr[n]);  // (col, source block; negative = transposed)
  std::vector<int> fill(frowptr.begin(), frowptr.end() - 1);
  for (int i = 0; i < n; ++i)
    for (int p = rowptr[i]; p < rowptr[i + 1]; ++p) {
      const int j = colidx[p];
      ent[fill[i]++] = {j, p + 1};
      if (j != i) ent[fill[j]++] = {i, -(p + 1)};
    }
  fcolidx.resize(frowptr[n]);
  fvals.resize((size_t)frowptr[n] * b2);
  for (int i = 0; i < n; ++i) {
    std::sort(ent.begin() + frowptr[i], ent.begin() + frowptr[i + 1]);
    for (int q = frowptr[i]; q < frowptr[i + 1]; ++q) {
      fcolidx[q] = ent[q].first;
      const int s = ent[q].second;
      const double* src = vals + (size_t)(std::abs(s) - 1) * b2;
      double* dst = &fvals[(size_t)q * b2];
      if (s > 0)
        std::memcpy(dst, src, sizeof(double) * b2);
      else
        for (int r = 0; r < bs; ++r)
          for (int c = 0; c < bs; ++c) dst[r * bs + c] = src[c * bs + r];
    }
  }
}

void assembleW(int nc, int nb, const int* hrowptr, const int* hcolidx, const double* hvals, const double* minv,
               std::vector<int>& wrowptr, std::vector<int>& wcolidx, std::vector<double>& wvals) {
  std::vector<int> trp, tci;
  std::vector<double> tv;
  transposeBSR(nc, nb, 3, 6, hrowptr, hcolidx, hvals, trp, tci, tv);
  // body-major index: for body k, the contacts touching it (tci) and the H block index
  std::vector<int> tpos(trp[nb]);
  {
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < nc; ++i)
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) tpos[fill[hcolidx[p]]++] = p;
  }
  const int nnzH = hrowptr[nc];
  // T_p = Minv_k H_p^T (6x3)
  std::vector<double> T((size_t)nnzH * 18);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < nnzH; ++p) {
    const double* h = hvals + (size_t)p * 18;
    const double* m = minv + (size_t)hcolidx[p] * 36;
    double* t = &T[(size_t)p * 18];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0;
        for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * h[c * 6 + l];
        t[r * 3 + c] = acc;
      }
  }
  std::vector<std::vector<int>> cols(nc);
#pragma omp parallel for schedule(dynamic, 512)
  for (int i = 0; i < nc; ++i) {
    auto& c = cols[i];
    for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) {
      const int k = hcolidx[p];
      for (int q = trp[k]; q < trp[k + 1]; ++q)
        if (tci[q] <= i) c.push_back(tci[q]);
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
  }
  wrowptr.assign(nc + 1, 0);
  for (int i = 0; i < nc; ++i) wrowptr[i + 1] = wrowptr[i] + (int)cols[i].size();
  wcolidx.resize(wrowptr[nc]);
  wvals.assign((size_t)wrowptr[nc] * 9, 0.0);
#pragma omp parallel for schedule(dynamic, 512)
  for (int i = 0; i < nc; ++i) {
    const int base = wrowptr[i];
    const auto& c = cols[i];
    for (size_t t = 0; t < c.size(); ++t) wcolidx[base + t] = c[t];
    for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) {
      const int k = hcolidx[p];
      const double* h = hvals + (size_t)p * 18;
      for (int q = trp[k]; q < trp[k + 1]; ++q) {
        const int j = tci[q];
        if (j > i) continue;
        const int slot = base + int(std::lower_bound(c.begin(), c.end(), j) - c.begin());
        const double* t = &T[(size_t)tpos[q] * 18];
        double* w = &wvals[(size_t)slot * 9];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) {
            double acc = 0;
            for (int l = 0; l < 6; ++l) acc += h[r * 6 + l] * t[l * 3 + cc];
            w[r * 3 + cc] += acc;
          }
      }
    }
  }
}

void diagBlocks(int nc, const int* hrowptr, const int* hcolidx, const double* hvals, const double* minv,
                std::vector<double>& diag) {
  // W_ii summed as the assembly sums it: pairs (a, b) of contact i's blocks at the
  // same body, a outer, b inner; T = Minv_k H_b^T first, then H_a T
  diag.assign((size_t)nc * 9, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nc; ++i) {
    double* d = &diag[(size_t)i * 9];
    for (int a = hrowptr[i]; a < hrowptr[i + 1]; ++a)
      for (int b = hrowptr[i]; b < hrowptr[i + 1]; ++b) {
        if (hcolidx[a] != hcolidx[b]) continue;
        const double* ha = hvals + (size_t)a * 18;
        const double* hb = hvals + (size_t)b * 18;
        const double* m = minv + (size_t)hcolidx[a] * 36;
        double t[6][3];
        for (int r = 0; r < 6; ++r)
          for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * hb[c * 6 + l];
            t[r][c] = acc;
          }
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int l = 0; l < 6; ++l) acc += ha[r * 6 + l] * t[l][c];
            d[r * 3 + c] += acc;
          }
      }
  }
}

void coloring(int n, const int* wrowptr, const int* wcolidx, int mode, std::vector<int64_t>& perm,
              std::vector<int64_t>& colors) {
  std::vector<std::vector<int>> members;
  if (mode == 0) {
    // Coloring::compute (Coloring.impl.hpp:20-62): contact i joins the first color c such
    // that lower_bound(color_c, j) == end for every lower neighbour j < i. Colors hold
    // ascending indices, so that holds iff last(color_c) < min_j.
    std::vector<int> last;
    for (int i = 0; i < n; ++i) {
      int jmin = std::numeric_limits<int>::max();
      for (int p = wrowptr[i]; p < wrowptr[i + 1]; ++p)
        if (wcolidx[p] < i) jmin = std::min(jmin, wcolidx[p]);
      size_t c = 0;
      for (; c < members.size(); ++c)
        if (jmin == std::numeric_limits<int>::max() || last[c] < jmin) break;
      if (c == members.size()) {
        members.emplace_back();
        last.push_back(-1);
      }
      members[c].push_back(i);
      last[c] = i;
    }
  } else {
    // first fit: smallest color not used by a lower neighbour
    std::vector<int> color(n, -1), mark;
    for (int i = 0; i < n; ++i) {
      for (int p = wrowptr[i]; p < wrowptr[i + 1]; ++p) {
        const int j = wcolidx[p];
        if (j < i) {
          const int cj = color[j];
          if (cj >= (int)mark.size()) mark.resize(cj + 1, -1);
          mark[cj] = i;
        }
      }
      int c = 0;
      while (c < (int)mark.size() && mark[c] == i) ++c;
      color[i] = c;
      if (c >= (int)members.size()) members.resize(c + 1);
      members[c].push_back(i);
    }
  }
  perm.clear();
  colors.assign(members.size() + 1, 0);
  for (size_t c = 0; c < members.size(); ++c) {
    colors[c + 1] = colors[c] + (int64_t)members[c].size();
    for (int i : members[c]) perm.push_back(i);
  }
}

void coloringFromH(int nc, int nb, const int* hrowptr, const int* hcolidx, int mode, std::vector<int64_t>& perm,
                   std::vector<int64_t>& colors) {
  std::vector<std::vector<int>> members;
  if (mode == 0) {
    // Coloring::compute as written: the smallest lower neighbour of contact i is the
    // first contact of one of its bodies, when that contact precedes i
    std::vector<int> first(nb, std::numeric_limits<int>::max());
    for (int i = 0; i < nc; ++i)
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) first[hcolidx[p]] = std::min(first[hcolidx[p]], i);
    std::vector<int> last;
    for (int i = 0; i < nc; ++i) {
      int jmin = std::numeric_limits<int>::max();
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p)
        if (first[hcolidx[p]] < i) jmin = std::min(jmin, first[hcolidx[p]]);
      size_t c = 0;
      for (; c < members.size(); ++c)
        if (jmin == std::numeric_limits<int>::max() || last[c] < jmin) break;
      if (c == members.size()) {
        members.emplace_back();
        last.push_back(-1);
      }
      members[c].push_back(i);
      last[c] = i;
    }
  } else {
    // first fit over the colors already used at the contact's bodies
    std::vector<std::vector<int>> used(nb);
    std::vector<int> mark;
    for (int i = 0; i < nc; ++i) {
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p)
        for (int c : used[hcolidx[p]]) {
          if (c >= (int)mark.size()) mark.resize(c + 1, -1);
          mark[c] = i;
        }
      int c = 0;
      while (c < (int)mark.size() && mark[c] == i) ++c;
      if (c >= (int)members.size()) members.resize(c + 1);
      members[c].push_back(i);
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) {
        auto& u = used[hcolidx[p]];
        if (std::find(u.begin(), u.end(), c) == u.end()) u.push_back(c);
      }
    }
  }
  perm.clear();
  colors.assign(members.size() + 1, 0);
  for (size_t c = 0; c < members.size(); ++c) {
    colors[c + 1] = colors[c] + (int64_t)members[c].size();
    for (int i : members[c]) perm.push_back(i);
  }
}

void localityOrder(int nc, int nb, const int* hrowptr, const int* hcolidx, std::vector<int>& cperm,
                   std::vector<int>& bmap) {
  const int none = std::numeric_limits<int>::max();
  auto keys = [&](const std::vector<int>& map, std::vector<std::pair<int, int>>& k) {
    k.resize(nc);
    for (int i = 0; i < nc; ++i) {
      int lo = none, hi = none;
      for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p) {
        const int j = map.empty() ? hcolidx[p] : map[hcolidx[p]];
        lo = (lo == none) ? j : std::min(lo, j);
        hi = (hi == none) ? j : std::max(hi, j);
      }
      k[i] = {lo, hi};
    }
  };
  auto order = [&](const std::vector<std::pair<int, int>>& k, std::vector<int>& perm) {
    perm.resize(nc);
    for (int i = 0; i < nc; ++i) perm[i] = i;
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return k[a] < k[b]; });
  };
  std::vector<std::pair<int, int>> k;
  std::vector<int> perm0;
  keys({}, k);
  order(k, perm0);
  bmap.assign(nb, -1);
  int next = 0;
  for (int i : perm0)
    for (int p = hrowptr[i]; p < hrowptr[i + 1]; ++p)
      if (bmap[hcolidx[p]] < 0) bmap[hcolidx[p]] = next++;
  for (int j = 0; j < nb; ++j)
    if (bmap[j] < 0) bmap[j] = next++;
  keys(bmap, k);
  order(k, cperm);
}


std::vector<int> batchBlockMap(int grid, int nxcd, int nclass, const int* k, const int* nteams, bool spread) {
  int kmax = 1;
  for (int c = 0; c < nclass; ++c) kmax = std::max(kmax, k[c]);
  const int nx = (!spread && nxcd > 1 && grid % nxcd == 0 && (grid / nxcd) % kmax == 0) ? nxcd : 1;
  const int per = grid / nx;
  std::vector<int> bmap(grid, -1), used(nx, 0);
  for (int c = 0; c < nclass; ++c)
    for (int t = 0; t < nteams[c]; ++t) {
      int x = 0;
      for (int y = 1; y < nx; ++y)
        if (used[y] < used[x]) x = y;
      if (used[x] + k[c] > per) return {};
      for (int r = 0; r < k[c]; ++r) bmap[(size_t)(used[x] + r) * nx + x] = (c << 20) | (t << 4) | r;
      used[x] += k[c];
    }
  return bmap;
}
}  // namespace sbhost
