// Teams: the set of threads that solves one problem together.
//
//  * GridTeam  - co-resident blocks (a whole persistent grid, or k blocks of the
//                batched kernel on one XCD); phases are separated by the
//                fence-free grid barrier and reductions go through the
//                deterministic grid reduction (Barrier) over the team's blocks.
//  * BlockTeam - one workgroup; phases are separated by __syncthreads and
//                reductions go through LDS. Used by the batched solver, where
//                every workgroup takes whole problems from a queue.
//
// Both expose the same interface to the solver bodies: the wave index and wave
// count used to stride over tiles, the thread index and thread count used to
// stride over contacts, sync(), reduce<NV>() leaving the results in red[], and
// st/ld/cld for vectors that other members of the team read in later phases.
#pragma once
#include <hip/hip_runtime.h>

#include "sb_device.hpp"
#include "sb_tiles.hpp"

namespace sb {

// Batch teams keep each wave's first unit of FP32 H-pass block values in LDS for the
// duration of a problem (pairs_loop_cached); the grid kernels do not.
#ifndef SB_PAIRS_CACHE
#define SB_PAIRS_CACHE 1
#endif
constexpr int kPairsCacheWaves = 4;  // 256-thread workgroups: 42 KB of LDS
constexpr int kGridCacheWaves = 8;   // 512-thread workgroups (FP64: 156 KB of LDS)

// Acquire = true: the barrier invalidates this XCD's caches and vectors written by
// other blocks are read with plain loads. Acquire = false: no invalidation; such
// vectors are read with coherent loads. Plain = true: vectors are written with
// plain stores (they stay in this XCD's L2); the barrier then writes the L2 back
// before arriving unless every block of the team runs on one XCD (bar.one_xcd).
// Scalar = true: reduction results and the solver's uniform state are read into scalar
// registers (uni): fewer vector registers, one more wave per SIMD for the batch kernel,
// at the cost of scalar spills (the single-problem grid kernel, one block per CU, keeps
// them in vector registers).
// Cache = true: the single-problem grid kernels' teams, whose waves keep operator data
// in LDS for the whole solve (FP32: PairsCache; FP64: TileCache).
template <bool Acquire, bool Plain = false, bool Scalar = false, bool Cache = false>
struct GridTeamT {
  // vectors written by other blocks must be read with coherent loads
  static constexpr bool kCoherentReads = !Acquire || SB_COHERENT_GATHER;
  static constexpr bool kScalarState = Scalar;
  static constexpr bool kPairsCache = (Scalar || Cache) && SB_PAIRS_CACHE != 0;
  // FP64 tile copies: the H (or W) pass for grid and batch teams, the H^T pass for grid
  // teams only (a batch workgroup's LDS holds one region per wave)
  static constexpr bool kTileCache = (Scalar || Cache) && SB_PAIRS_CACHE != 0;
  static constexpr bool kTileCacheHT = Cache && SB_PAIRS_CACHE != 0;
  PairsCache<float> pcache{nullptr, false, false};
  TileCache<double> tcache{nullptr, {false, false}};
  Barrier bar;
  int gw, nwaves, tid, nthreads;
  bool lead;
  double red[kMaxPartials];
  // The blocks rank = 0..g.nblocks-1 of a team (the whole grid, or a group of
  // blocks of the batched kernel) that share the barrier state g.
  __device__ GridTeamT(const GridSync& g, unsigned rank) {
    bar.gs = g;
    bar.rank = rank;
    bar.acquire = Acquire;
    bar.plain_stores = Plain;
    const int wpb = blockDim.x / SB_WAVE;
    gw = Scalar ? uni((int)rank * wpb + wave_in_block()) : (int)rank * wpb + wave_in_block();
    nwaves = (int)g.nblocks * wpb;
    tid = (int)rank * blockDim.x + threadIdx.x;
    nthreads = (int)g.nblocks * blockDim.x;
    lead = rank == 0 && threadIdx.x == 0;
  }
  __device__ bool leader() const { return lead; }
  __device__ __forceinline__ bool sync() { return bar.sync(); }
  // Plain-store team over a whole grid: checks through one barrier that block b runs
  // on XCD b mod 8 (round-robin dispatch; the word after the top counter collects
  // mismatches) and, if every block does, lets each barrier write every XCD's L2 back
  // once, from the XCD's last arriving block, instead of once per block.
  __device__ __forceinline__ bool group_by_xcd() {
    __shared__ unsigned s_mis;
    unsigned* mis = bar.gs.top + 1;
    if (threadIdx.x == 0 && xcd_id() != bar.rank % kXcdGroups)
      __hip_atomic_fetch_or(mis, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!sync()) return false;
    if (threadIdx.x == 0) s_mis = __hip_atomic_load(mis, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    bar.xcd_wb = Plain && s_mis == 0u;
    return true;
  }
  template <int NV>
  __device__ __forceinline__ bool reduce(const double (&v)[NV], unsigned max_mask) {
    if (!bar.reduce<NV>(v, max_mask)) return false;
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k] = Scalar ? uni(bar.red[k]) : bar.red[k];
    return true;
  }
  template <typename T>
  __device__ static void st(T* p, T v) {
    if (Plain)
      gst(p, v);
    else
      cst(p, v);
  }
  template <typename T>
  __device__ static T ld(const T* p) {
    if (Acquire) return mld(p);
    return cld(p);
  }
  template <typename T>
  __device__ static T cload(const T* p) {
    return cld(p);
  }
};
using GridTeam = GridTeamT<true>;
// Plain stores; L2 write-back at each barrier, once per XCD after group_by_xcd().
using GridTeamWB = GridTeamT<true, true>;

// dst[k] = src[k] (converted) for k = tid, tid + nthreads, ... < n, by a team; src
// was written before the team's last barrier (or by the same thread). Groups of
// eight loads are issued before their stores, so they are in flight together (a
// load/store pair per iteration would wait one memory round trip each).
template <typename Team, typename TD, typename TS>
__device__ __forceinline__ void team_copy(const Team& tm, TD* dst, const TS* src, int n) {
  constexpr int U = 8;
  const int st = tm.nthreads;
  int k = tm.tid;
  for (; k + (U - 1) * st < n; k += U * st) {
    TS v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = Team::ld(src + k + u * st);
#pragma unroll
    for (int u = 0; u < U; ++u) Team::st(dst + k + u * st, (TD)v[u]);
  }
  for (; k < n; k += st) Team::st(dst + k, (TD)Team::ld(src + k));
}

struct BlockTeam {
  static constexpr bool kCoherentReads = false;
  static constexpr bool kScalarState = true;  // used by the batch kernel only
  static constexpr bool kPairsCache = SB_PAIRS_CACHE != 0;
  static constexpr bool kTileCache = SB_PAIRS_CACHE != 0;
  static constexpr bool kTileCacheHT = false;
  PairsCache<float> pcache{nullptr, false, false};
  TileCache<double> tcache{nullptr, {false, false}};
  int gw, nwaves, tid, nthreads;
  double red[kMaxPartials];
  __device__ BlockTeam() {
    gw = uni(wave_in_block());
    nwaves = blockDim.x / SB_WAVE;
    tid = threadIdx.x;
    nthreads = blockDim.x;
  }
  __device__ bool leader() const { return threadIdx.x == 0; }
  // Workgroup barrier; __syncthreads orders global-memory accesses of the
  // workgroup (all its waves share one CU and its L1).
  __device__ __forceinline__ bool sync() {
    __syncthreads();
    return true;
  }
  // Fixed-order reduction: xor tree inside each wave, then the per-wave values
  // in wave order. Every thread ends with the same bits in red[].
  template <int NV>
  __device__ __forceinline__ bool reduce(const double (&v)[NV], unsigned max_mask) {
    __shared__ double s_p[NV][16];  // up to 1024 threads
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double w = ((max_mask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
      if (lane_id() == 0) s_p[k][wave_in_block()] = w;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const bool isMax = (max_mask >> k) & 1u;
      double acc = 0.0;
      for (int q = 0; q < nwaves; ++q) acc = isMax ? (s_p[k][q] > acc ? s_p[k][q] : acc) : acc + s_p[k][q];
      red[k] = uni(acc);
    }
    __syncthreads();
    return true;
  }
  template <typename T>
  __device__ static void st(T* p, T v) {
    gst(p, v);
  }
  template <typename T>
  __device__ static T ld(const T* p) {
    return gld(p);
  }
  template <typename T>
  __device__ static T cload(const T* p) {
    return gld(p);
  }
};

}  // namespace sb
