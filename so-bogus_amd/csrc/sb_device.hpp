// Device utilities for the CDNA4 (gfx950) solve path: wave64 shuffles and
// reductions, a bounded grid barrier with a deterministic in-barrier reduction,
// and the persistent-kernel bookkeeping shared by the solver kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SB_WAVE 64

namespace sb {

// ------------------------------------------------------------------ shuffles
__device__ __forceinline__ double shfl_up_d(double v, int d) { return __shfl_up(v, d, SB_WAVE); }
__device__ __forceinline__ int shfl_up_i(int v, int d) { return __shfl_up(v, d, SB_WAVE); }
__device__ __forceinline__ int shfl_down_i(int v, int d) { return __shfl_down(v, d, SB_WAVE); }
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, SB_WAVE); }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (SB_WAVE - 1); }
__device__ __forceinline__ int wave_in_block() { return threadIdx.x / SB_WAVE; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int m = SB_WAVE / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, SB_WAVE);
  return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int m = SB_WAVE / 2; m >= 1; m >>= 1) {
    const T o = __shfl_xor(v, m, SB_WAVE);
    v = v > o ? v : o;
  }
  return v;
}

// ------------------------------------------------------------------ barrier
// Grid-wide barrier for persistent kernels. All blocks of the grid must be
// co-resident (cooperative launch). The last block to arrive reduces the
// per-block partial sums in a fixed order (so every run and every block sees
// the same bits), publishes them, and opens the barrier. Waiting is bounded by
// a real-time deadline: on timeout the error word is set and every block
// leaves the solver loop, so a broken schedule ends instead of hanging.
struct GridSync {
  unsigned* count;     // arrivals in the current generation
  unsigned* gen;       // generation counter
  int* error;          // non-zero after a timeout (sticky)
  double* partials;    // [nblocks][kMaxPartials]
  double* reduced;     // [kMaxPartials] published results
  unsigned nblocks;
  uint64_t timeout_ticks;  // s_memrealtime ticks (100 MHz)
};
constexpr int kMaxPartials = 8;

__device__ __forceinline__ uint64_t realtime() { return __builtin_amdgcn_s_memrealtime(); }

// kind bit k set => partial k is reduced with max instead of sum.
__device__ inline bool grid_reduce_sync(const GridSync& gs, int nvals, unsigned max_mask) {
  __shared__ unsigned s_last;
  __shared__ int s_err;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = __hip_atomic_load(gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    const unsigned arrived = __hip_atomic_fetch_add(gs.count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (arrived == gs.nblocks - 1) ? 1u : 0u;
    s_err = 0;
    if (!s_last) {
      const uint64_t t0 = realtime();
      while (__hip_atomic_load(gs.gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        __builtin_amdgcn_s_sleep(2);
        if (__hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (realtime() - t0 > gs.timeout_ticks) {
          __hip_atomic_store(gs.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
      }
    }
    __threadfence();
  }
  __syncthreads();
  if (s_last) {
    // fixed-order reduction of the partials by the whole last block
    __shared__ double s_red[kMaxPartials][SB_WAVE];
    const int nthreads = blockDim.x;
    for (int k = 0; k < nvals; ++k) {
      const bool isMax = (max_mask >> k) & 1u;
      double acc = 0.0;
      for (unsigned b = threadIdx.x; b < gs.nblocks; b += nthreads) {
        const double v = __hip_atomic_load(&gs.partials[(size_t)b * kMaxPartials + k], __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
        acc = isMax ? (v > acc ? v : acc) : acc + v;
      }
      // combine across the block: first within waves (fixed xor tree), then waves in order
      acc = isMax ? wave_max(acc) : wave_sum(acc);
      if (lane_id() == 0) s_red[k][wave_in_block()] = acc;
    }
    __syncthreads();
    if (threadIdx.x < (unsigned)nvals) {
      const int k = threadIdx.x;
      const bool isMax = (max_mask >> k) & 1u;
      double acc = 0.0;
      const int nw = (nthreads + SB_WAVE - 1) / SB_WAVE;
      for (int w = 0; w < nw; ++w) acc = isMax ? (s_red[k][w] > acc ? s_red[k][w] : acc) : acc + s_red[k][w];
      __hip_atomic_store(&gs.reduced[k], acc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      __hip_atomic_store(gs.count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_add(gs.gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) s_err = __hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return s_err == 0;
}

// Block-level partial reduction: every thread contributes v[0..nvals); thread 0
// of the block writes the block's partials (sum, or max for bits of max_mask).
template <int NV>
__device__ inline void block_partials(const GridSync& gs, const double (&v)[NV], unsigned max_mask) {
  __shared__ double s_p[NV][16];
  double w[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k] = ((max_mask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s_p[k][wave_in_block()] = w[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + SB_WAVE - 1) / SB_WAVE;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const bool isMax = (max_mask >> k) & 1u;
      double acc = 0.0;
      for (int q = 0; q < nw; ++q) acc = isMax ? (s_p[k][q] > acc ? s_p[k][q] : acc) : acc + s_p[k][q];
      gs.partials[(size_t)blockIdx.x * kMaxPartials + k] = acc;
    }
  }
}

__device__ __forceinline__ double read_reduced(const GridSync& gs, int k) {
  return __hip_atomic_load(&gs.reduced[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace sb
