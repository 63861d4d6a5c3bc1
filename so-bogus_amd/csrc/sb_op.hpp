// One application of the Delassus operator inside a persistent solver, by a
// team: the assembled W (3x3 blocks), or matrix-free H (Minv (H^T x)) with a
// caller-provided body step between the H^T and the H passes.
#pragma once
#include <hip/hip_runtime.h>

#include "sb_problem.hpp"
#include "sb_tiles.hpp"

namespace sb {

template <typename T>
constexpr bool kPairsPass = sizeof(T) == 4;

// H^T pass with two tiles per wave in flight (tile_loop_pairs)
#ifndef SB_HT_TILE_PAIRS32
#define SB_HT_TILE_PAIRS32 1
#endif
#ifndef SB_HT_TILE_PAIRS64
#define SB_HT_TILE_PAIRS64 0
#endif
template <typename T>
constexpr bool kHtTilePairs = sizeof(T) == 4 ? SB_HT_TILE_PAIRS32 != 0 : SB_HT_TILE_PAIRS64 != 0;

// Minv_k v for the element-major [36][nb] storage.
template <typename T>
__device__ __forceinline__ void minv_apply(const T* minv, int nb, int k, const T v[6], T w[6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < 6; ++c) acc += gld(minv + (size_t)(r * 6 + c) * nb + k) * v[c];
    w[r] = acc;
  }
}

// M x with x given by the loader, finished by the leader on complete contact rows
// (with the row data fetched by pre). Matrix-free: body(k, v) receives (H^T x)_k
// and must store the body vector that the H pass reads from wbody[6k..6k+6).
// Contains 0 (assembled, no long rows), 1 (matrix-free) or 2 (matrix-free with
// hub bodies) internal team barriers.
template <bool MF, typename T, typename Team, typename Loader, typename Pre, typename Leader, typename Body>
__device__ __forceinline__ bool apply_operator(const DevProblem<T>& P, const T* wbody, T* carry, Team& tm, Loader& ld,
                                               Pre& pre, Leader& lead, Body& body) {
  if (!MF) {
    if constexpr (Team::kTileCache && sizeof(T) == 8)
      tile_loop_cached<3, 3, T, T>(P.W, tm.gw, tm.nwaves, ld, pre, lead, carry, tm.tcache, 0);
    else
      tile_loop<3, 3, T, T>(P.W, tm.gw, tm.nwaves, ld, pre, lead, carry);
    if (P.W.nlong > 0) {
      if (!tm.sync()) return false;
      for (int k = tm.gw; k < P.W.nlong; k += tm.nwaves) {
        T v[3];
        long_row_sum<3, T, T, Team::kCoherentReads>(P.W, k, carry, v);
        if (lane_id() == 0) {
          const int row = gld(P.W.long_row + k);
          lead(row, v, pre(row));
        }
      }
    }
    return true;
  }
  NoPre nopre;
  auto leadA = [&](int k, T v[6], NoPre::type) { body(k, v); };
  if constexpr (kHtTilePairs<T> && Team::kPairsCache && sizeof(T) == 4)
    tile_loop_pairs<6, 3, T, T, true, Team::kTileCache>(P.HT, tm.gw, tm.nwaves, ld, leadA, carry, &tm.pcache);
  else if constexpr (kHtTilePairs<T>)
    tile_loop_pairs<6, 3, T, T>(P.HT, tm.gw, tm.nwaves, ld, leadA, carry);
  else if constexpr (Team::kTileCacheHT && sizeof(T) == 8)
    tile_loop_cached<6, 3, T, T>(P.HT, tm.gw, tm.nwaves, ld, nopre, leadA, carry, tm.tcache, 1);
  else
    tile_loop<6, 3, T, T>(P.HT, tm.gw, tm.nwaves, ld, nopre, leadA, carry);
  if (!tm.sync()) return false;
  if (P.HT.nlong > 0) {
    for (int k = tm.gw; k < P.HT.nlong; k += tm.nwaves) {
      T v[6];
      long_row_sum<6, T, T, Team::kCoherentReads>(P.HT, k, carry, v);
      if (lane_id() == 0) body(gld(P.HT.long_row + k), v);
    }
    if (!tm.sync()) return false;
  }
  auto ldB = [&](int k, T xc[6]) {
#pragma unroll
    for (int r = 0; r < 6; ++r) xc[r] = Team::ld(&wbody[6 * k + r]);
  };
  // FP32: one lane per contact when every row of H holds two blocks (FP64 keeps the
  // tiled pass: its 36 values per lane would cost the kernels their occupancy)
  if (kPairsPass<T> && P.H.pairs) {
    if constexpr (Team::kPairsCache && sizeof(T) == 4)
      pairs_loop_cached<T>(P.H, tm.gw, tm.nwaves, ldB, pre, lead, tm.pcache);
    else
      pairs_loop<T>(P.H, tm.gw, tm.nwaves, ldB, pre, lead);
  } else if constexpr (Team::kTileCache && sizeof(T) == 8) {
    tile_loop_cached<3, 6, T, T>(P.H, tm.gw, tm.nwaves, ldB, pre, lead, carry, tm.tcache, 0);
  } else {
    tile_loop<3, 6, T, T>(P.H, tm.gw, tm.nwaves, ldB, pre, lead, carry);
  }
  return true;
}

}  // namespace sb
