// Device utilities for the CDNA4 (gfx950) solve path: wave64 shuffles and
// reductions, XCD-coherent loads/stores, and a bounded fence-free grid barrier
// with a deterministic grid reduction for persistent kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SB_WAVE 64

// 1: values written by other blocks are read with agent-scope loads (sc1) and the
//    barrier does not invalidate caches, so read-only matrix data stays in L2
//    across phases. 0: the barrier invalidates L1/L2 (acquire) and such values
//    are read with plain cached loads.
#ifndef SB_COHERENT_GATHER
#define SB_COHERENT_GATHER 0
#endif
// 1: reductions whose blocks share one XCD exchange their partial sums through that
//    XCD's L2 (plain stores, and plain loads after the barrier's L1 invalidation).
#ifndef SB_XCD_LOCAL_PARTIALS
#define SB_XCD_LOCAL_PARTIALS 1
#endif

namespace sb {

// ------------------------------------------------------------------ shuffles
__device__ __forceinline__ double shfl_up_d(double v, int d) { return __shfl_up(v, d, SB_WAVE); }
__device__ __forceinline__ int shfl_up_i(int v, int d) { return __shfl_up(v, d, SB_WAVE); }
__device__ __forceinline__ int shfl_down_i(int v, int d) { return __shfl_down(v, d, SB_WAVE); }
__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, SB_WAVE); }

__device__ __forceinline__ int lane_id() { return threadIdx.x & (SB_WAVE - 1); }

// A value that every lane of the wave holds (the result of a team reduction, a
// decision taken from one): read from the first active lane into scalar registers,
// so that the compiler keeps it out of the vector register file.
__device__ __forceinline__ double uni(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float uni(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
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

// ------------------------------------------------------------------ coherent access
// MI355X has one L2 per XCD and the L2s are not coherent with each other for
// ordinary device memory. Values that one block writes and another block reads
// later in the same persistent kernel are stored at agent scope (relaxed atomic
// store: global_store with the sc1 policy bit, written through to memory), so no
// dirty copy stays behind in a private L2 and no L2 write-back is needed at the
// barrier. After each barrier every block invalidates its L1/L2 once (acquire
// fence), and loads of such values in the next phase are then ordinary cached
// loads. cld() (agent-scope load) is kept for reads that may not follow a
// barrier of the reading block.
//
// Every access goes through the global address space explicitly: the solvers'
// arrays live in device memory, but pointers read from memory (the per-problem
// descriptors of the batched kernel) carry no address space, and the compiler
// would otherwise emit flat instructions, which are slower and are waited on
// together with LDS and scalar accesses.
template <typename T>
using gmem = __attribute__((address_space(1))) T;
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  return *(const gmem<T>*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(gmem<T>*)p = v;
}
__device__ __forceinline__ void cst(double* p, double v) {
  __hip_atomic_store((gmem<double>*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double cld(const double* p) {
  return __hip_atomic_load((gmem<double>*)const_cast<double*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cst(float* p, float v) {
  __hip_atomic_store((gmem<float>*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float cld(const float* p) {
  return __hip_atomic_load((gmem<float>*)const_cast<float*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cst(int* p, int v) {
  __hip_atomic_store((gmem<int>*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Load of a value another block wrote in an earlier phase of the same kernel.
template <typename T>
__device__ __forceinline__ T mld(const T* p) {
#if SB_COHERENT_GATHER
  return cld(p);
#else
  return gld(p);
#endif
}

// ------------------------------------------------------------------ barrier
// Grid-wide barrier for persistent kernels whose blocks are all co-resident
// (cooperative launch). Counters are never reset during a launch (barrier
// number e completes when a counter reaches a multiple of e + 1), so there is
// no reset race. Every wave first waits for its own memory operations
// (s_waitcnt 0), so the coherent stores issued before the barrier have
// completed when the block arrives. Waiting is bounded by a real-time
// deadline: on timeout the error word is set and every block leaves the solver
// loop, so a broken schedule ends instead of hanging.
struct GridSync {
  unsigned* group;     // [kSyncGroups][kSyncStride] monotonic arrival counters per group
  unsigned* top;       // monotonic count of completed groups
  unsigned* release;   // [kSyncGroups][kSyncStride] epoch released to each group
  int* error;          // non-zero after a timeout (sticky)
  double* partials;    // [2][nblocks][kMaxPartials], alternating between reductions
  unsigned nblocks;
  uint64_t timeout_ticks;  // s_memrealtime ticks (100 MHz)
};
constexpr unsigned kSyncGroups = 16;  // blocks are grouped by blockIdx % kSyncGroups
constexpr unsigned kXcdGroups = 8;    // XCDs of the device: with round-robin dispatch, block b runs on XCD b % 8
constexpr unsigned kSyncStride = 32;  // one 128-byte line per counter
constexpr int kMaxPartials = 12;

__device__ __forceinline__ uint64_t realtime() { return __builtin_amdgcn_s_memrealtime(); }

// Index of the XCD (accelerator complex die) this wave runs on (HW_REG_XCC_ID).
__device__ __forceinline__ unsigned xcd_id() { return __builtin_amdgcn_s_getreg((15 << 11) | 20) & 15u; }

struct Barrier {
  GridSync gs;
  unsigned rank = 0;         // this block's index among the gs.nblocks blocks (must be set)
  bool acquire = true;       // invalidate L1/L2 after the barrier (readers use plain loads); false:
                             // readers use coherent loads and the caches are left alone
  bool one_xcd = false;      // every block of the barrier runs on one XCD, whose L2 they
                             // share: plain stores need no L2 write-back, and the
                             // reduction's partial sums go through that L2 too
  bool plain_stores = false; // vectors are written with plain stores: unless one_xcd,
                             // the L2 is written back before arriving (release)
  bool xcd_wb = false;       // with plain_stores: block b runs on XCD b mod 8 (checked by the
                             // caller); blocks are grouped by XCD and the last block of each
                             // XCD to arrive writes that XCD's L2 back, once for all its blocks
  unsigned epoch = 0;
  unsigned nred = 0;         // reductions done (selects the partials buffer)
  double red[kMaxPartials];  // results of the last reduction (identical in every block)

  // Two-level arrival: a block increments its group's counter; the group's last
  // block increments the top counter; the last group writes the new epoch into
  // every group's release word, which that group's blocks poll. Contention on any
  // one address is limited to a group (or the kSyncGroups group leaders).
  __device__ __forceinline__ bool sync() {
    __shared__ int s_ok;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
      if (plain_stores && !one_xcd && !xcd_wb) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      // the grouping must stay the same for every barrier of a launch (the counters are
      // monotonic): plain-store teams are grouped by rank % 8 from their first barrier on,
      // before xcd_wb or one_xcd is known
      const unsigned ng = plain_stores ? kXcdGroups : kSyncGroups;
      const unsigned G = gs.nblocks < ng ? gs.nblocks : ng;
      const unsigned g = rank % G;
      const unsigned gsize = gs.nblocks / G + ((g < gs.nblocks % G) ? 1u : 0u);
      const unsigned old = __hip_atomic_fetch_add(&gs.group[g * kSyncStride], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
      if (old + 1 == (epoch + 1) * gsize) {
        // every block of this XCD has stored its data (each waited for its stores
        // before arriving): one write-back of the XCD's L2 publishes them all
        if (xcd_wb) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned t = __hip_atomic_fetch_add(gs.top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t + 1 == (epoch + 1) * G)
          for (unsigned k = 0; k < G; ++k)
            __hip_atomic_store(&gs.release[k * kSyncStride], epoch + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      const uint64_t t0 = realtime();
      int ok = 1;
      for (;;) {
        // the sticky error word is checked first: no block passes a barrier after a timeout
        if (__hip_atomic_load(gs.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          ok = 0;
          break;
        }
        if (__hip_atomic_load(&gs.release[g * kSyncStride], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= epoch + 1)
          break;
        if (realtime() - t0 > gs.timeout_ticks) {
          __hip_atomic_store(gs.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      s_ok = ok;
#if !SB_COHERENT_GATHER
      // Acquire (agent scope): drop this CU's L1 (and stale L2 copies) of lines
      // written by other blocks before the barrier, so that plain cached loads in
      // the next phase read fresh values. Blocks on one XCD read through the L2
      // their plain stores went to. (A workgroup-scope invalidation, buffer_inv sc0,
      // leaves L1 alone when a workgroup's waves share one CU: a reduction through
      // plain stores then read stale partial sums.)
      if (acquire) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#endif
    }
    __syncthreads();
    ++epoch;
    return s_ok != 0;
  }

  // Reduces v[0..NV) over the whole grid (sum, or max for the bits of max_mask)
  // and leaves the results in red[] of every block. Block partials are combined
  // in the same fixed order in every block, so all blocks see the same bits.
  template <int NV>
  __device__ __forceinline__ bool reduce(const double (&v)[NV], unsigned max_mask) {
    __shared__ double s_p[NV][16];
    __shared__ double s_r[NV];
    // Two partials buffers: a block can start writing the buffer of reduction n+2
    // only after every block has passed the barrier of reduction n+1, i.e. after
    // every block has finished reading the buffer of reduction n.
    double* part = gs.partials + (size_t)(nred & 1u) * gs.nblocks * kMaxPartials;
    ++nred;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double w = ((max_mask >> k) & 1u) ? wave_max(v[k]) : wave_sum(v[k]);
      if (lane_id() == 0) s_p[k][wave_in_block()] = w;
    }
    __syncthreads();
    if (threadIdx.x < NV) {
      const int k = threadIdx.x;
      const bool isMax = (max_mask >> k) & 1u;
      const int nw = (blockDim.x + SB_WAVE - 1) / SB_WAVE;
      double acc = 0.0;
      for (int q = 0; q < nw; ++q) acc = isMax ? (s_p[k][q] > acc ? s_p[k][q] : acc) : acc + s_p[k][q];
#if SB_XCD_LOCAL_PARTIALS
      if (one_xcd)
        gst(&part[(size_t)rank * kMaxPartials + k], acc);
      else
#endif
        cst(&part[(size_t)rank * kMaxPartials + k], acc);
    }
    if (!sync()) return false;
    // every wave reduces one partial index over all blocks: lane l takes blocks l, l+64,
    // l+128, l+192 of each group of 256 (loads issued together), then the xor tree
    const int nw = blockDim.x / SB_WAVE;
    for (int k = wave_in_block(); k < NV; k += nw) {
      const bool isMax = (max_mask >> k) & 1u;
      double acc = 0.0;
      for (unsigned b0 = 0; b0 < gs.nblocks; b0 += 4 * SB_WAVE) {
        double x[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const unsigned b = b0 + j * SB_WAVE + lane_id();
          x[j] = (b < gs.nblocks) ? ((SB_XCD_LOCAL_PARTIALS && one_xcd) ? gld(&part[(size_t)b * kMaxPartials + k])
                                             : cld(&part[(size_t)b * kMaxPartials + k]))
                                  : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) acc = isMax ? (x[j] > acc ? x[j] : acc) : acc + x[j];
      }
      acc = isMax ? wave_max(acc) : wave_sum(acc);
      if (lane_id() == 0) s_r[k] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k] = s_r[k];
    __syncthreads();
    return true;
  }
};

}  // namespace sb
