// Tiled block-sparse row gather ("lane per block") for wave64.
//
// A block-sparse matrix with BR x BC blocks is stored as:
//   * blocks in row order (the reference's compressed major index,
//     src/Core/Block/CompressedSparseBlockIndex.hpp:19-144), with the row and
//     column of every block (brow, bcol);
//   * values element-major: vals[e * nblocks + q] is element e (row-major in the
//     block) of block q, so lane l of a wave reading block q0 + l reads
//     consecutive addresses for every e (coalesced HBM traffic);
//   * tiles of at most 64 consecutive blocks made of whole rows. A row longer
//     than 64 blocks (a hub) gets tiles of its own ("long" tiles), whose partial
//     sums go to a carry buffer and are combined by a fix-up pass.
// One wave processes one tile: lane l multiplies block q0 + l with its column
// vector, then a segmented inclusive scan over lanes (rows are contiguous)
// leaves each row's sum in the row's last lane (the "leader").
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sb_device.hpp"

namespace sb {

template <typename T>
struct TileMat {
  int nrows = 0, ncols = 0, nblocks = 0, ntiles = 0;
  int scan_steps = 0;          // ceil(log2(longest row inside a tile))
  const int* tstart = nullptr; // [ntiles + 1] block offsets
  const int* brow = nullptr;   // [nblocks]
  const int* bcol = nullptr;   // [nblocks]
  const int* tlong = nullptr;  // [ntiles] < 0: whole rows, -(1 + scan steps of the tile);
                               // >= 0: the tile holds a piece of that long row
  const T* vals = nullptr;     // [BR*BC][nblocks]
  int pairs = 0;               // 1: every row holds exactly two blocks, row i's at 2i and 2i + 1
  int nlong = 0;               // long rows
  const int* long_row = nullptr;  // [nlong]
  const int* long_t0 = nullptr;   // [nlong] first tile
  const int* long_t1 = nullptr;   // [nlong] one past last tile
};

struct NoPre {  // row prefetch that loads nothing
  struct type {};
  __device__ type operator()(int) const { return type{}; }
};

// Per-lane indices of one tile: the block range and, for lane l, the row and
// column of block q0 + l (row -1 past the tile's end).
struct TileIdx {
  int q0, q1, row, col, tl;
};
template <typename T>
__device__ __forceinline__ TileIdx tile_indices(const TileMat<T>& A, int t) {
  TileIdx ix;
  ix.q0 = uni(gld(A.tstart + t));
  ix.q1 = uni(gld(A.tstart + t + 1));
  ix.tl = uni(gld(A.tlong + t));
  const int q = ix.q0 + lane_id();
  if (q < ix.q1) {
    ix.row = gld(A.brow + q);
    ix.col = gld(A.bcol + q);
  } else {
    ix.row = -1;
    ix.col = 0;
  }
  return ix;
}

template <int BR, typename Acc, typename Leader, typename Row>
__device__ __forceinline__ void tile_finish(int t, const TileIdx& ix, Acc v[BR], Leader& lead, const Row& rp,
                                            Acc* carry);

// Processes tile t of A whose indices are ix. load(col, xc[BC]) returns the
// column vector (in the accumulation type Acc). pre(row) is evaluated by every
// active lane (so the row's own data is fetched while the gathers are in flight);
// lead(row, v[BR], pre) consumes complete row sums in the row's last lane.
// Partial sums of long tiles go to carry[t * BR + r].
// cv != nullptr: the block values come from the wave's LDS copy cv[e * 64 + lane].
template <int BR, int BC, typename T, typename Acc, bool NT = false, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ void tile_gather(const TileMat<T>& A, int t, const TileIdx& ix, Loader& load, Pre& pre,
                                            Leader& lead, Acc* carry, const T* cv = nullptr) {
  const int l = lane_id();
  const int q = ix.q0 + l;
  const bool act = ix.row >= 0;
  const int row = ix.row;
  Acc v[BR];
#pragma unroll
  for (int r = 0; r < BR; ++r) v[r] = Acc(0);
  decltype(pre(0)) rp{};
  if (act) {
    // Issue order matters: the row data and the block values are requested before
    // the column gather, whose loader may branch on the gathered values (a
    // projection) - loads placed after such branches wait for them.
    rp = pre(row);
    const T* vq = A.vals + q;
    const size_t nb = (size_t)A.nblocks;
    T blk[BR * BC];
    if (cv != nullptr) {
#pragma unroll
      for (int e = 0; e < BR * BC; ++e) blk[e] = cv[e * SB_WAVE + l];
    } else {
#pragma unroll
      for (int e = 0; e < BR * BC; ++e)
        blk[e] = NT ? __builtin_nontemporal_load((const gmem<T>*)(vq + (size_t)e * nb)) : gld(vq + (size_t)e * nb);
    }
    Acc xc[BC];
    load(ix.col, xc);
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      Acc acc = Acc(0);
#pragma unroll
      for (int c = 0; c < BC; ++c) acc += Acc(blk[r * BC + c]) * xc[c];
      v[r] = acc;
    }
  }
  tile_finish<BR>(t, ix, v, lead, rp, carry);
}

// The end of a tile: the segmented scan over the lanes' block products v, then the
// row leaders (or, for a piece of a long row, its carry).
template <int BR, typename Acc, typename Leader, typename Row>
__device__ __forceinline__ void tile_finish(int t, const TileIdx& ix, Acc v[BR], Leader& lead, const Row& rp,
                                            Acc* carry) {
  const int l = lane_id();
  const int row = ix.row;
  const bool act = row >= 0;
  const int steps = ix.tl >= 0 ? 6 : -ix.tl - 1;  // a long tile is one row across the wave
  for (int s = 0; s < steps; ++s) {
    const int d = 1 << s;
    const int ru = __shfl_up(row, d, SB_WAVE);
    Acc vu[BR];
#pragma unroll
    for (int r = 0; r < BR; ++r) vu[r] = __shfl_up(v[r], d, SB_WAVE);
    if (l >= d && ru == row) {
#pragma unroll
      for (int r = 0; r < BR; ++r) v[r] += vu[r];
    }
  }
  const int rn = __shfl_down(row, 1, SB_WAVE);
  const bool leader = act && (l == SB_WAVE - 1 || rn != row);
  if (leader) {
    if (ix.tl >= 0) {
#pragma unroll
      for (int r = 0; r < BR; ++r) cst(&carry[(size_t)t * BR + r], v[r]);
    } else {
      lead(row, v, rp);
    }
  }
}

// A loader may split its work into fetch(j), which issues the loads of column j's
// vector and returns them, and finish(raw, xc), which uses them (declared by a
// member kSplit = true); tile_pair issues both tiles' fetches before either finish.
template <typename L, typename = void>
struct SplitLoader {
  static constexpr bool value = false;
};
template <typename L>
struct SplitLoader<L, std::void_t<decltype(L::kSplit)>> {
  static constexpr bool value = L::kSplit;
};

// Tiles ta and tb of A together (a pass without row data, e.g. the H^T pass of the
// matrix-free operator): both tiles' block values and column vectors are requested
// before either tile's products, scan and leaders, so a wave has two tiles' loads in
// flight. An index set with row -1 in every lane (past the last tile) does nothing.
// cva != nullptr: tile ta's block values come from the wave's LDS copy cva[e * 64 + lane].
template <int BR, int BC, typename T, typename Acc, typename Loader, typename Leader>
__device__ __forceinline__ void tile_pair(const TileMat<T>& A, int ta, const TileIdx& ia, int tb, const TileIdx& ib,
                                          Loader& load, Leader& lead, Acc* carry, const T* cva = nullptr) {
  const int l = lane_id();
  const bool aa = ia.row >= 0, ab = ib.row >= 0;
  const size_t nb = (size_t)A.nblocks;
  T ba[BR * BC], bb[BR * BC];
  if (aa) {
    if (cva != nullptr) {
#pragma unroll
      for (int e = 0; e < BR * BC; ++e) ba[e] = cva[e * SB_WAVE + l];
    } else {
      const T* vq = A.vals + ia.q0 + l;
#pragma unroll
      for (int e = 0; e < BR * BC; ++e) ba[e] = gld(vq + (size_t)e * nb);
    }
  }
  if (ab) {
    const T* vq = A.vals + ib.q0 + l;
#pragma unroll
    for (int e = 0; e < BR * BC; ++e) bb[e] = gld(vq + (size_t)e * nb);
  }
  Acc xa[BC], xb[BC];
  if constexpr (SplitLoader<Loader>::value) {
    decltype(load.fetch(0)) ra{}, rb{};
    if (aa) ra = load.fetch(ia.col);
    if (ab) rb = load.fetch(ib.col);
    if (aa) load.finish(ra, xa);
    if (ab) load.finish(rb, xb);
  } else {
    if (aa) load(ia.col, xa);
    if (ab) load(ib.col, xb);
  }
  Acc va[BR], vb[BR];
#pragma unroll
  for (int r = 0; r < BR; ++r) va[r] = vb[r] = Acc(0);
  if (aa) {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      Acc acc = Acc(0);
#pragma unroll
      for (int c = 0; c < BC; ++c) acc += Acc(ba[r * BC + c]) * xa[c];
      va[r] = acc;
    }
  }
  if (ab) {
#pragma unroll
    for (int r = 0; r < BR; ++r) {
      Acc acc = Acc(0);
#pragma unroll
      for (int c = 0; c < BC; ++c) acc += Acc(bb[r * BC + c]) * xb[c];
      vb[r] = acc;
    }
  }
  const NoPre::type none{};
  tile_finish<BR>(ta, ia, va, lead, none, carry);
  tile_finish<BR>(tb, ib, vb, lead, none, carry);
}

// LDS copy of the block values and column indices of the first unit of rows that a
// wave takes in pairs_loop (64 rows: [18][64] 2-vectors, then [64] index pairs), kept
// across the operator applications of one problem: the units a wave takes depend only
// on the matrix, the team's wave count and the wave's index. lds == nullptr: no copy.
// After them, the indices of the first two tiles the wave takes in tile_loop_pairs
// (its H^T pass): per tile q0, q1, tl, then each lane's row and column.
template <typename T>
struct PairsCache {
  typedef T T2 __attribute__((ext_vector_type(2)));
  static constexpr int kHWords = 18 * 64 + 64 * 8 / (int)sizeof(T2);  // H-pass unit, T2 units
  static constexpr int kTInts = 8 + 4 * 64;                            // H^T first pair's indices
  static constexpr int kWords = kHWords + kTInts * 4 / (int)sizeof(T2);
  T2* lds;
  bool valid;   // lds holds the current H's unit
  bool tvalid;  // lds holds the current H^T's first pair of tile indices
  // optional: the block values of the first H^T tile ([18][64]; grid teams only)
  T* htv = nullptr;
  bool htvalid = false;
};

// Tiles first, first + stride, ... of A two at a time (tile_pair), with the indices
// of the next two loaded before the current two are processed. Cached: the first
// pair's indices come from (or, on the first call for a matrix, go to) the wave's
// PairsCache pc (pc->lds == nullptr: no copy).
template <int BR, int BC, typename T, typename Acc, bool Cached = false, bool CacheVals = false, typename Loader,
          typename Leader>
__device__ __forceinline__ void tile_loop_pairs(const TileMat<T>& A, int first, int stride, Loader& load,
                                                Leader& lead, Acc* carry, PairsCache<T>* pc = nullptr) {
  if (first >= A.ntiles) return;
  auto idx = [&](int t) {
    if (t < A.ntiles) return tile_indices(A, t);
    TileIdx none;
    none.q0 = none.q1 = 0;
    none.row = -1;
    none.col = 0;
    none.tl = -1;
    return none;
  };
  TileIdx ca, cb;
  if constexpr (Cached) {
    int* ti = pc->lds != nullptr ? reinterpret_cast<int*>(pc->lds + PairsCache<T>::kHWords) : nullptr;
    const int l = lane_id();
    if (ti != nullptr && pc->tvalid) {
      ca.q0 = uni(ti[0]);
      ca.q1 = uni(ti[1]);
      ca.tl = uni(ti[2]);
      cb.q0 = uni(ti[3]);
      cb.q1 = uni(ti[4]);
      cb.tl = uni(ti[5]);
      ca.row = ti[8 + l];
      ca.col = ti[8 + SB_WAVE + l];
      cb.row = ti[8 + 2 * SB_WAVE + l];
      cb.col = ti[8 + 3 * SB_WAVE + l];
    } else {
      ca = idx(first);
      cb = idx(first + stride);
      if (ti != nullptr) {
        if (l == 0) {
          ti[0] = ca.q0;
          ti[1] = ca.q1;
          ti[2] = ca.tl;
          ti[3] = cb.q0;
          ti[4] = cb.q1;
          ti[5] = cb.tl;
        }
        ti[8 + l] = ca.row;
        ti[8 + SB_WAVE + l] = ca.col;
        ti[8 + 2 * SB_WAVE + l] = cb.row;
        ti[8 + 3 * SB_WAVE + l] = cb.col;
        pc->tvalid = true;
      }
    }
  } else {
    ca = idx(first);
    cb = idx(first + stride);
  }
  int t0 = first;
  if constexpr (Cached && CacheVals) {
    // CacheVals: the first tile's block values from LDS (copied there by the first call)
    T* cv = pc->htv;
    if (cv != nullptr) {
      if (!pc->htvalid) {
        if (ca.row >= 0) {
          const T* vq = A.vals + ca.q0 + lane_id();
#pragma unroll
          for (int e = 0; e < BR * BC; ++e) cv[e * SB_WAVE + lane_id()] = gld(vq + (size_t)e * A.nblocks);
        }
        pc->htvalid = true;
      }
      const int tn = first + 2 * stride;
      TileIdx na = ca, nb = cb;
      if (tn < A.ntiles) {
        na = idx(tn);
        nb = idx(tn + stride);
      }
      tile_pair<BR, BC, T, Acc>(A, first, ca, first + stride, cb, load, lead, carry, cv);
      ca = na;
      cb = nb;
      t0 = tn;
    }
  }
  for (int t = t0; t < A.ntiles; t += 2 * stride) {
    const int tn = t + 2 * stride;
    TileIdx na = ca, nb = cb;
    if (tn < A.ntiles) {
      na = idx(tn);
      nb = idx(tn + stride);
    }
    tile_pair<BR, BC, T, Acc>(A, t, ca, t + stride, cb, load, lead, carry);
    ca = na;
    cb = nb;
  }
}

// Tiles first, first + stride, ... of A, software-pipelined: the indices of the
// next tile are loaded before the current tile's gathers are issued, so each tile
// costs one dependent memory round trip instead of three.
#ifndef SB_TILE_PIPELINE
#define SB_TILE_PIPELINE 1
#endif
// Block values of the solver's tiles read with non-temporal loads (streamed past
// L2, which then keeps the gathered vectors).
#ifndef SB_SOLVER_NT
#define SB_SOLVER_NT 0
#endif
template <int BR, int BC, typename T, typename Acc, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ void tile_loop(const TileMat<T>& A, int first, int stride, Loader& load, Pre& pre,
                                          Leader& lead, Acc* carry) {
#if !SB_TILE_PIPELINE
  for (int t = first; t < A.ntiles; t += stride) tile_gather<BR, BC, T, Acc, SB_SOLVER_NT>(A, t, tile_indices(A, t), load, pre, lead, carry);
  return;
#endif
  if (first >= A.ntiles) return;
  TileIdx cur = tile_indices(A, first);
  for (int t = first; t < A.ntiles; t += stride) {
    const int tn = t + stride;
    TileIdx nxt = cur;
    if (tn < A.ntiles) nxt = tile_indices(A, tn);
    tile_gather<BR, BC, T, Acc, SB_SOLVER_NT>(A, t, cur, load, pre, lead, carry);
    cur = nxt;
  }
}

// LDS copy of the first tile a wave takes in one tile_loop pass (its block values,
// [BR*BC][64], then each lane's row and column and the tile's q0, q1, tl), kept
// across the operator applications of one problem like PairsCache. Two passes per
// wave (H and H^T, or W): regions 0 and 1.
template <typename T>
struct TileCache {
  static constexpr int kMaxVals = 18;  // BR * BC of the largest blocks (3 x 6)
  static constexpr int kPassBytes = kMaxVals * 64 * (int)sizeof(T) + (2 * 64 + 8) * 4;
  static constexpr int kBytes = 2 * kPassBytes;  // per wave
  char* lds;  // this wave's slice, or nullptr: no copy
  bool valid[2];
};

// tile_loop with the first tile's indices and block values read from region `pass`
// of the wave's TileCache; the first call for a matrix copies them there first.
template <int BR, int BC, typename T, typename Acc, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ void tile_loop_cached(const TileMat<T>& A, int first, int stride, Loader& load, Pre& pre,
                                                 Leader& lead, Acc* carry, TileCache<T>& tc, int pass) {
  static_assert(BR * BC <= TileCache<T>::kMaxVals, "tile cache too small");
  if (first >= A.ntiles) return;
  if (tc.lds == nullptr) {
    tile_loop<BR, BC, T, Acc>(A, first, stride, load, pre, lead, carry);
    return;
  }
  char* base = tc.lds + pass * TileCache<T>::kPassBytes;
  T* cv = reinterpret_cast<T*>(base);
  int* ci = reinterpret_cast<int*>(base + TileCache<T>::kMaxVals * 64 * sizeof(T));
  const int l = lane_id();
  if (!tc.valid[pass]) {
    const TileIdx ix = tile_indices(A, first);
    if (l == 0) {
      ci[128] = ix.q0;
      ci[129] = ix.q1;
      ci[130] = ix.tl;
    }
    ci[l] = ix.row;
    ci[64 + l] = ix.col;
    if (ix.row >= 0) {
      const T* vq = A.vals + ix.q0 + l;
#pragma unroll
      for (int e = 0; e < BR * BC; ++e) cv[e * SB_WAVE + l] = gld(vq + (size_t)e * A.nblocks);
    }
    tc.valid[pass] = true;
  }
  TileIdx cur;
  cur.q0 = uni(ci[128]);
  cur.q1 = uni(ci[129]);
  cur.tl = uni(ci[130]);
  cur.row = ci[l];
  cur.col = ci[64 + l];
  {
    const int tn = first + stride;
    TileIdx nxt = cur;
    if (tn < A.ntiles) nxt = tile_indices(A, tn);
    tile_gather<BR, BC, T, Acc, SB_SOLVER_NT>(A, first, cur, load, pre, lead, carry, cv);
    cur = nxt;
  }
  for (int t = first + stride; t < A.ntiles; t += stride) {
    const int tn = t + stride;
    TileIdx nxt = cur;
    if (tn < A.ntiles) nxt = tile_indices(A, tn);
    tile_gather<BR, BC, T, Acc, SB_SOLVER_NT>(A, t, cur, load, pre, lead, carry);
    cur = nxt;
  }
}

// One lane per row for a matrix whose rows all hold two blocks, blocks 2i and 2i + 1
// of row i (TileMat::pairs): rows 64 u .. 64 u + 63 for units u = first, first +
// stride, ... Element e of a row's two blocks is one 2-vector load; every lane
// does its row's leader work (no scan, no idle lanes). The sums equal tile_gather's
// (each block's products in order, then the second block's sum plus the first's).
template <typename T, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ void pairs_loop(const TileMat<T>& A, int first, int stride, Loader& load, Pre& pre,
                                           Leader& lead) {
  typedef T T2 __attribute__((ext_vector_type(2)));
  const int nunits = (A.nrows + SB_WAVE - 1) / SB_WAVE;
  const size_t nb2 = (size_t)A.nblocks / 2;
  for (int u = first; u < nunits; u += stride) {
    const int i = u * SB_WAVE + lane_id();
    if (i >= A.nrows) continue;
    auto rp = pre(i);
    const T2* v2 = reinterpret_cast<const T2*>(A.vals) + i;
    T2 e[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) e[k] = gld(v2 + (size_t)k * nb2);
    const int ca = gld(A.bcol + 2 * i), cb = gld(A.bcol + 2 * i + 1);
    T xa[6], xb[6];
    load(ca, xa);
    load(cb, xb);
    T v[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T aa = T(0), ab = T(0);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        aa += e[r * 6 + c].x * xa[c];
        ab += e[r * 6 + c].y * xb[c];
      }
      v[r] = ab + aa;
    }
    lead(i, v, rp);
  }
}

// pairs_loop with the first unit's values read from (or, on the first call for a
// matrix, copied into) the wave's LDS slice; the values, and so the sums, are those
// of pairs_loop.
template <typename T, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ void pairs_loop_cached(const TileMat<T>& A, int first, int stride, Loader& load, Pre& pre,
                                                  Leader& lead, PairsCache<T>& pc) {
  typedef T T2 __attribute__((ext_vector_type(2)));
  const int nunits = (A.nrows + SB_WAVE - 1) / SB_WAVE;
  const size_t nb2 = (size_t)A.nblocks / 2;
  for (int u = first; u < nunits; u += stride) {
    const int i = u * SB_WAVE + lane_id();
    if (i >= A.nrows) continue;
    auto rp = pre(i);
    const T2* v2 = reinterpret_cast<const T2*>(A.vals) + i;
    T2 e[18];
    int ca, cb;
    if (u == first && pc.lds != nullptr) {
      T2* c = pc.lds + lane_id();
      int2* ci = reinterpret_cast<int2*>(pc.lds + 18 * SB_WAVE) + lane_id();
      if (pc.valid) {
        const int2 cc = *ci;
        ca = cc.x;
        cb = cc.y;
#pragma unroll
        for (int k = 0; k < 18; ++k) e[k] = c[k * SB_WAVE];
      } else {
        ca = gld(A.bcol + 2 * i);
        cb = gld(A.bcol + 2 * i + 1);
#pragma unroll
        for (int k = 0; k < 18; ++k) e[k] = gld(v2 + (size_t)k * nb2);
        *ci = make_int2(ca, cb);
#pragma unroll
        for (int k = 0; k < 18; ++k) c[k * SB_WAVE] = e[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 18; ++k) e[k] = gld(v2 + (size_t)k * nb2);
      ca = gld(A.bcol + 2 * i);
      cb = gld(A.bcol + 2 * i + 1);
    }
    T xa[6], xb[6];
    load(ca, xa);
    load(cb, xb);
    T v[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      T aa = T(0), ab = T(0);
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        aa += e[r * 6 + c].x * xa[c];
        ab += e[r * 6 + c].y * xb[c];
      }
      v[r] = ab + aa;
    }
    lead(i, v, rp);
  }
  pc.valid = true;
}

// Sum of the carries of long row k (one wave), fixed order: lane-strided partial
// sums (four tiles in flight per lane) then the xor tree. Result valid in every
// lane. Coherent: the carries were written by other blocks of the same kernel and
// no cache invalidation separates the reads from the writes; otherwise plain loads.
template <int BR, typename T, typename Acc, bool Coherent = true>
__device__ __forceinline__ void long_row_sum(const TileMat<T>& A, int k, const Acc* carry, Acc v[BR]) {
  const int t0 = gld(A.long_t0 + k), t1 = gld(A.long_t1 + k);
#pragma unroll
  for (int r = 0; r < BR; ++r) v[r] = Acc(0);
  auto ld = [&](size_t i) { return Coherent ? cld(carry + i) : gld(carry + i); };
  int t = t0 + lane_id();
  for (; t + 3 * SB_WAVE < t1; t += 4 * SB_WAVE) {
    Acc c[4][BR];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < BR; ++r) c[j][r] = ld((size_t)(t + j * SB_WAVE) * BR + r);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < BR; ++r) v[r] += c[j][r];
  }
  for (; t < t1; t += SB_WAVE) {
#pragma unroll
    for (int r = 0; r < BR; ++r) v[r] += ld((size_t)t * BR + r);
  }
#pragma unroll
  for (int r = 0; r < BR; ++r) v[r] = wave_sum(v[r]);
}

}  // namespace sb
