// Symbolic set-up on the device: the tile layout of a BSR pattern (the layout the
// host builds in sbhost::buildTiles), the stable transposition of H, and the pattern
// of W = H Minv H^T (lower triangle, then both triangles) with the product terms of
// W's numeric phase. The reference computes these on the host: the transpose in
// SparseBlockMatrixBase::transpose / cacheTranspose
// (src/Core/Block/SparseBlockMatrixBase.impl.hpp:169-202), the product pattern in
// SparseMatrixMatrixProduct (src/Core/Block/SparseMatrixMatrixProduct.impl.hpp:124-212).
// Every result equals the host-built one element for element.
//
// The one sequential step of the tiling, the greedy packing of whole rows into
// tiles of at most 64 blocks, is a chain i -> next(i) over the rows (next(i): the
// first row that does not fit in a tile starting at row i, at most 64 rows on).
// It is followed in chunks of rows: for each chunk, the 64 possible entry offsets
// are walked in parallel to the chunk's exit offset, one workgroup composes the
// chunks' exits in order, and each chunk then marks its tile starts.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "sb_symbolic.hpp"

namespace sb {

namespace {
constexpr int kTileBlocks = 64;   // blocks per tile: one wave
constexpr int kMaxChunks = 512;   // chunks of the chain (exits in LDS: 32 KB)

inline int grid_n(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 16384)); }

__device__ __forceinline__ int ceil_log2(int v) {
  int s = 0;
  while ((1 << s) < v) ++s;
  return s;
}
// first t in [lo, hi) with a[t] > key (a ascending), hi if none
__device__ __forceinline__ int upper_bound_i(const int* a, int lo, int hi, int key) {
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (a[mid] > key) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}
// largest i in [0, n) with a[i] <= key, given a[0] <= key (a non-decreasing)
__device__ __forceinline__ int row_of(const int* a, int n, int key) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = lo + ((hi - lo) >> 1);
    if (a[mid] <= key) lo = mid;
    else hi = mid;
  }
  return lo;
}
template <typename K>
__device__ __forceinline__ int lower_bound_k(const K* a, int n, K key) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
}  // namespace

int ceil_log2_host(int64_t v) {
  int s = 0;
  while (((int64_t)1 << s) < v) ++s;
  return s;
}

// ------------------------------------------------------------------ hipCUB steps
hipError_t scan_inclusive_int(const int* in, int* out, int n, void* tmp, size_t& bytes, hipStream_t s) {
  return hipcub::DeviceScan::InclusiveSum(tmp, bytes, in, out, n, s);
}
hipError_t reduce_max_int(const int* in, int* out, int n, void* tmp, size_t& bytes, hipStream_t s) {
  return hipcub::DeviceReduce::Max(tmp, bytes, in, out, n, s);
}
hipError_t reduce_sum_int64(const int* in, long long* out, int n, void* tmp, size_t& bytes, hipStream_t s) {
  return hipcub::DeviceReduce::Reduce(tmp, bytes, in, out, n, hipcub::Sum(), 0LL, s);
}
hipError_t sort_pairs_u32(const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int end_bit,
                          void* tmp, size_t& bytes, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, s);
}
hipError_t sort_keys_u64(const uint64_t* kin, uint64_t* kout, int n, int end_bit, void* tmp, size_t& bytes,
                         hipStream_t s) {
  return hipcub::DeviceRadixSort::SortKeys(tmp, bytes, kin, kout, n, 0, end_bit, s);
}

// ------------------------------------------------------------------ checks
__global__ void k_check_rows(int nrows, const int* __restrict__ rp, int* err) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x)
    if ((i == 0 && rp[0] != 0) || rp[i + 1] < rp[i]) *err = 1;
}
__global__ void k_check_cols(int nnz, int ncols, const int* __restrict__ ci, int* err) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += gridDim.x * blockDim.x)
    if (ci[p] < 0 || ci[p] >= ncols) *err = 1;
}
hipError_t launch_check_rows(int nrows, const int* rp, int* err, hipStream_t s) {
  (void)hipGetLastError();
  if (nrows > 0) k_check_rows<<<grid_n(nrows), 256, 0, s>>>(nrows, rp, err);
  return hipGetLastError();
}
hipError_t launch_check_cols(int nnz, int ncols, const int* ci, int* err, hipStream_t s) {
  (void)hipGetLastError();
  if (nnz > 0) k_check_cols<<<grid_n(nnz), 256, 0, s>>>(nnz, ncols, ci, err);
  return hipGetLastError();
}

// ------------------------------------------------------------------ tiles
__global__ void k_row_len(int nrows, const int* __restrict__ rp, int* __restrict__ tlen) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x)
    tlen[i] = max(1, rp[i + 1] - rp[i]);
}
hipError_t launch_row_len(int nrows, const int* rp, int* tlen, hipStream_t s) {
  (void)hipGetLastError();
  if (nrows > 0) k_row_len<<<grid_n(nrows), 256, 0, s>>>(nrows, rp, tlen);
  return hipGetLastError();
}

__global__ void k_tile_expand(int nrows, int nblocks, const int* __restrict__ trp, const int* __restrict__ rp,
                              const int* __restrict__ ci, int* __restrict__ brow, int* __restrict__ bcol,
                              int* __restrict__ src) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nblocks; q += gridDim.x * blockDim.x) {
    const int i = row_of(trp, nrows, q);  // trp strictly increasing, trp[0] = 0
    const int t = q - trp[i];
    const int p0 = rp[i], len = rp[i + 1] - p0;
    brow[q] = i;
    bcol[q] = len > 0 ? ci[p0 + t] : 0;
    src[q] = len > 0 ? p0 + t : -1;
  }
}
hipError_t launch_tile_expand(int nrows, int nblocks, const int* trp, const int* rp, const int* ci, int* brow,
                              int* bcol, int* src, hipStream_t s) {
  (void)hipGetLastError();
  if (nblocks > 0) k_tile_expand<<<grid_n(nblocks), 256, 0, s>>>(nrows, nblocks, trp, rp, ci, brow, bcol, src);
  return hipGetLastError();
}

// A row longer than a tile has tiles of its own, so the next tile starts at the
// next row. Otherwise the tile takes rows i, i + 1, ... while their blocks fit in
// 64: next(i) is the first j with trp[j + 1] - trp[i] > 64, at most i + 64 (every
// row holds at least one block), or nrows.
__global__ void k_tile_next(int nrows, const int* __restrict__ trp, const int* __restrict__ tlen,
                            int* __restrict__ next) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x) {
    if (tlen[i] > kTileBlocks) {
      next[i] = i + 1;
      continue;
    }
    const int limit = trp[i] + kTileBlocks;
    int lo = i + 1, hi = min(i + kTileBlocks, nrows);
    while (lo < hi) {
      const int mid = lo + ((hi - lo) >> 1);
      if (trp[mid + 1] > limit) hi = mid;
      else lo = mid + 1;
    }
    next[i] = lo;
  }
}
hipError_t launch_tile_next(int nrows, const int* trp, const int* tlen, int* next, hipStream_t s) {
  (void)hipGetLastError();
  if (nrows > 0) k_tile_next<<<grid_n(nrows), 256, 0, s>>>(nrows, trp, tlen, next);
  return hipGetLastError();
}

int chain_chunk_rows(int nrows) {
  const int per = (int)((nrows + (int64_t)kMaxChunks - 1) / kMaxChunks);
  return std::max(1024, per);
}

// One workgroup of 64 lanes per chunk: lane e follows the chain from row c0 + e to
// the first chain row at or past the chunk's end, c1 + exit (exit < 64, since a
// step advances at most 64 rows).
__global__ void __launch_bounds__(64) k_chain_exits(int nrows, int C, const int* __restrict__ next,
                                                    unsigned char* __restrict__ exits) {
  const int c = blockIdx.x;
  const int c0 = c * C, c1 = min(nrows, c0 + C);
  int r = c0 + (int)threadIdx.x;
  while (r < c1) r = next[r];
  exits[(size_t)c * 64 + threadIdx.x] = (unsigned char)(r - c1);
}
// Entry offsets of the chunks in order: the chain starts at row 0.
__global__ void __launch_bounds__(256) k_chain_entries(int nchunks, const unsigned char* __restrict__ exits,
                                                       int* __restrict__ entry) {
  __shared__ unsigned char s_ex[kMaxChunks * 64];
  for (int k = threadIdx.x; k < nchunks * 64; k += blockDim.x) s_ex[k] = exits[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    int e = 0;
    for (int c = 0; c < nchunks; ++c) {
      entry[c] = e;
      e = s_ex[c * 64 + e];
    }
  }
}
__global__ void k_chain_mark(int nrows, int C, int nchunks, const int* __restrict__ next,
                             const int* __restrict__ entry, unsigned char* __restrict__ is_start) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunks) return;
  const int c0 = c * C, c1 = min(nrows, c0 + C);
  for (int r = c0 + entry[c]; r < c1; r = next[r]) is_start[r] = 1;
}
hipError_t launch_tile_chain(int nrows, const int* next, unsigned char* exits, int* entry, unsigned char* is_start,
                             hipStream_t s) {
  (void)hipGetLastError();
  if (nrows <= 0) return hipSuccess;
  const int C = chain_chunk_rows(nrows);
  const int nchunks = (nrows + C - 1) / C;
  if (nchunks > kMaxChunks) return hipErrorInvalidValue;
  hipError_t e = hipMemsetAsync(is_start, 0, (size_t)nrows, s);
  if (e != hipSuccess) return e;
  k_chain_exits<<<nchunks, 64, 0, s>>>(nrows, C, next, exits);
  k_chain_entries<<<1, 256, 0, s>>>(nchunks, exits, entry);
  k_chain_mark<<<(nchunks + 63) / 64, 64, 0, s>>>(nrows, C, nchunks, next, entry, is_start);
  return hipGetLastError();
}

__global__ void k_tile_counts(int nrows, const int* __restrict__ tlen, const unsigned char* __restrict__ is_start,
                              int* __restrict__ ntl, int* __restrict__ nlg) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x) {
    const int L = tlen[i];
    const bool lng = L > kTileBlocks;
    ntl[i] = is_start[i] ? (lng ? (L + kTileBlocks - 1) / kTileBlocks : 1) : 0;
    nlg[i] = lng ? 1 : 0;
  }
}
hipError_t launch_tile_counts(int nrows, const int* tlen, const unsigned char* is_start, int* ntl, int* nlg,
                              hipStream_t s) {
  (void)hipGetLastError();
  if (nrows > 0) k_tile_counts<<<grid_n(nrows), 256, 0, s>>>(nrows, tlen, is_start, ntl, nlg);
  return hipGetLastError();
}

// Tile t spans blocks [tstart[t], tstart[t + 1]). A tile of whole rows records
// -(1 + scan steps of its longest row); each tile of a long row records the row's
// index among the long rows, whose tiles are [long_t0, long_t1).
__global__ void k_tile_fill(int nrows, int nblocks, int ntiles, const int* __restrict__ trp,
                            const int* __restrict__ tlen, const int* __restrict__ next,
                            const unsigned char* __restrict__ is_start, const int* __restrict__ tbase,
                            const int* __restrict__ lbase, int* __restrict__ tstart, int* __restrict__ tlong,
                            int* __restrict__ long_row, int* __restrict__ long_t0, int* __restrict__ long_t1) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += gridDim.x * blockDim.x) {
    if (i == 0) tstart[ntiles] = nblocks;
    if (!is_start[i]) continue;
    const int t = tbase[i], L = tlen[i];
    if (L > kTileBlocks) {
      const int k = lbase[i], m = (L + kTileBlocks - 1) / kTileBlocks;
      long_row[k] = i;
      long_t0[k] = t;
      long_t1[k] = t + m;
      for (int j = 0; j < m; ++j) {
        tstart[t + j] = trp[i] + kTileBlocks * j;
        tlong[t + j] = k;
      }
    } else {
      int mx = 1;
      for (int r = i; r < next[i]; ++r) mx = max(mx, tlen[r]);
      tstart[t] = trp[i];
      tlong[t] = -1 - ceil_log2(mx);
    }
  }
}
hipError_t launch_tile_fill(int nrows, int nblocks, int ntiles, const int* trp, const int* tlen, const int* next,
                            const unsigned char* is_start, const int* tbase, const int* lbase, int* tstart,
                            int* tlong, int* long_row, int* long_t0, int* long_t1, hipStream_t s) {
  (void)hipGetLastError();
  if (nrows > 0)
    k_tile_fill<<<grid_n(nrows), 256, 0, s>>>(nrows, nblocks, ntiles, trp, tlen, next, is_start, tbase, lbase, tstart,
                                              tlong, long_row, long_t0, long_t1);
  return hipGetLastError();
}

// ------------------------------------------------------------------ transposition
__global__ void k_iota(int n, int* __restrict__ v) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) v[k] = k;
}
hipError_t launch_iota(int n, int* v, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_iota<<<grid_n(n), 256, 0, s>>>(n, v);
  return hipGetLastError();
}

__global__ void k_lower_bounds_u32(int nout, const unsigned* __restrict__ keys, int n, int* __restrict__ out) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < nout; j += gridDim.x * blockDim.x)
    out[j] = lower_bound_k<unsigned>(keys, n, (unsigned)j);
}
hipError_t launch_lower_bounds_u32(int nout, const unsigned* keys, int n, int* out, hipStream_t s) {
  (void)hipGetLastError();
  if (nout > 0) k_lower_bounds_u32<<<grid_n(nout), 256, 0, s>>>(nout, keys, n, out);
  return hipGetLastError();
}

__global__ void k_rows_of(int n, const int* __restrict__ src, int nrows, const int* __restrict__ rp,
                          int* __restrict__ rows) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    rows[q] = row_of(rp, nrows, src[q]);
}
hipError_t launch_rows_of(int n, const int* src, int nrows, const int* rp, int* rows, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_rows_of<<<grid_n(n), 256, 0, s>>>(n, src, nrows, rp, rows);
  return hipGetLastError();
}

__global__ void k_compose_src(int n, const int* __restrict__ ssrc, const int* __restrict__ tsrc,
                              int* __restrict__ out) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x)
    out[q] = ssrc[q] >= 0 ? tsrc[ssrc[q]] : -1;
}
hipError_t launch_compose_src(int n, const int* ssrc, const int* tsrc, int* out, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_compose_src<<<grid_n(n), 256, 0, s>>>(n, ssrc, tsrc, out);
  return hipGetLastError();
}

__global__ void k_fill_f64(size_t n, double v, double* __restrict__ out) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    out[k] = v;
}
hipError_t launch_fill_f64(size_t n, double v, double* out, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_fill_f64<<<grid_n((int64_t)n), 256, 0, s>>>(n, v, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ W pattern
// The contacts at body k are tci[trp[k] .. trp[k + 1]), ascending (stable transpose of
// a row-ordered H).
__global__ void k_w_cand_count(int nc, const int* __restrict__ hrp, const int* __restrict__ hci,
                               const int* __restrict__ trp, const int* __restrict__ tci, int* __restrict__ cnt) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    int c = 0;
    for (int p = hrp[i]; p < hrp[i + 1]; ++p) {
      const int k = hci[p];
      c += upper_bound_i(tci, trp[k], trp[k + 1], i) - trp[k];
    }
    cnt[i] = c;
  }
}
hipError_t launch_w_cand_count(int nc, const int* hrp, const int* hci, const int* trp, const int* tci, int* cnt,
                               hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_w_cand_count<<<grid_n(nc), 256, 0, s>>>(nc, hrp, hci, trp, tci, cnt);
  return hipGetLastError();
}

// cptr: row i's keys start at cptr[i] (exclusive prefix sums, nc + 1 entries)
__global__ void k_w_cand_fill(int nc, const int* __restrict__ hrp, const int* __restrict__ hci,
                              const int* __restrict__ trp, const int* __restrict__ tci, const int* __restrict__ cptr,
                              uint64_t* __restrict__ keys) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    int o = cptr[i];
    for (int p = hrp[i]; p < hrp[i + 1]; ++p) {
      const int k = hci[p];
      const int q1 = upper_bound_i(tci, trp[k], trp[k + 1], i);
      for (int q = trp[k]; q < q1; ++q) keys[o++] = (uint64_t)i * (uint64_t)nc + (uint64_t)tci[q];
    }
  }
}
hipError_t launch_w_cand_fill(int nc, const int* hrp, const int* hci, const int* trp, const int* tci,
                              const int* cptr, uint64_t* keys, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_w_cand_fill<<<grid_n(nc), 256, 0, s>>>(nc, hrp, hci, trp, tci, cptr, keys);
  return hipGetLastError();
}

__global__ void k_unique_flags(int n, const uint64_t* __restrict__ keys, int* __restrict__ flags) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    flags[t] = (t == 0 || keys[t] != keys[t - 1]) ? 1 : 0;
}
hipError_t launch_unique_flags(int n, const uint64_t* keys, int* flags, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_unique_flags<<<grid_n(n), 256, 0, s>>>(n, keys, flags);
  return hipGetLastError();
}
// pos: inclusive prefix sums of the flags
__global__ void k_unique_scatter(int n, const uint64_t* __restrict__ keys, const int* __restrict__ pos,
                                 uint64_t* __restrict__ out) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
    if (t == 0 || keys[t] != keys[t - 1]) out[pos[t] - 1] = keys[t];
}
hipError_t launch_unique_scatter(int n, const uint64_t* keys, const int* pos, uint64_t* out, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_unique_scatter<<<grid_n(n), 256, 0, s>>>(n, keys, pos, out);
  return hipGetLastError();
}

__global__ void k_key_rows(int nc, const uint64_t* __restrict__ keys, int n, int* __restrict__ rp,
                           int* __restrict__ ci) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (int i = tid; i <= nc; i += stride) rp[i] = lower_bound_k<uint64_t>(keys, n, (uint64_t)i * (uint64_t)nc);
  for (int t = tid; t < n; t += stride) ci[t] = (int)(keys[t] % (uint64_t)nc);
}
hipError_t launch_key_rows(int nc, const uint64_t* keys, int n, int* rp, int* ci, hipStream_t s) {
  (void)hipGetLastError();
  k_key_rows<<<grid_n(std::max(nc + 1, n)), 256, 0, s>>>(nc, keys, n, rp, ci);
  return hipGetLastError();
}

__global__ void k_upper_keys(int nc, const int* __restrict__ wrp, const int* __restrict__ wci,
                             uint64_t* __restrict__ keys) {
  const uint64_t n = (uint64_t)nc;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    for (int t = wrp[i]; t < wrp[i + 1]; ++t) {
      const int j = wci[t];
      keys[t] = j != i ? (uint64_t)j * n + (uint64_t)i : n * n;
    }
}
hipError_t launch_upper_keys(int nc, const int* wrp, const int* wci, uint64_t* keys, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_upper_keys<<<grid_n(nc), 256, 0, s>>>(nc, wrp, wci, keys);
  return hipGetLastError();
}

__global__ void k_upper_rows(int nc, const uint64_t* __restrict__ ukeys, int n, const int* __restrict__ wrp,
                             int* __restrict__ uptr, int* __restrict__ flen) {
  const uint64_t m = (uint64_t)nc;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= nc; i += gridDim.x * blockDim.x) {
    const int u0 = lower_bound_k<uint64_t>(ukeys, n, (uint64_t)i * m);
    uptr[i] = u0;
    if (i < nc) {
      const int u1 = lower_bound_k<uint64_t>(ukeys, n, (uint64_t)(i + 1) * m);
      flen[i] = (wrp[i + 1] - wrp[i]) + (u1 - u0);
    }
  }
}
hipError_t launch_upper_rows(int nc, const uint64_t* ukeys, int n, const int* wrp, int* uptr, int* flen,
                             hipStream_t s) {
  (void)hipGetLastError();
  k_upper_rows<<<grid_n(nc + 1), 256, 0, s>>>(nc, ukeys, n, wrp, uptr, flen);
  return hipGetLastError();
}

// frp: full row pointer (exclusive prefix sums of the lengths, nc + 1 entries)
__global__ void k_full_fill(int nc, const int* __restrict__ wrp, const int* __restrict__ wci,
                            const int* __restrict__ uptr, const uint64_t* __restrict__ ukeys,
                            const int* __restrict__ frp, int* __restrict__ fci) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    int o = frp[i];
    for (int t = wrp[i]; t < wrp[i + 1]; ++t) fci[o++] = wci[t];
    for (int u = uptr[i]; u < uptr[i + 1]; ++u) fci[o++] = (int)(ukeys[u] % (uint64_t)nc);
  }
}
hipError_t launch_full_fill(int nc, const int* wrp, const int* wci, const int* uptr, const uint64_t* ukeys,
                            const int* frp, int* fci, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_full_fill<<<grid_n(nc), 256, 0, s>>>(nc, wrp, wci, uptr, ukeys, frp, fci);
  return hipGetLastError();
}

// Terms of block (i, j): a over the H blocks of contact hi, b over those of contact lo,
// at the same body, in that nesting order (the order the host assembly sums them),
// where hi is the one of i, j with the larger assembly order key okey (the contact's
// index when W was assembled: a later permutation keeps each block's orientation,
// as a symmetric permutation of the assembled W does).
__device__ __forceinline__ void w_block_rows(const int* __restrict__ okey, int i, int j, int& hi, int& lo) {
  const bool iHi = okey[i] >= okey[j];
  hi = iHi ? i : j;
  lo = iHi ? j : i;
}
__global__ void k_w_term_count(int nq, const int* __restrict__ brow, const int* __restrict__ bcol,
                               const int* __restrict__ src, const int* __restrict__ okey, const int* __restrict__ hrp,
                               const int* __restrict__ hci, int* __restrict__ cnt) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    int c = 0;
    if (src[q] >= 0) {
      int i, j;
      w_block_rows(okey, brow[q], bcol[q], i, j);
      for (int a = hrp[i]; a < hrp[i + 1]; ++a)
        for (int b = hrp[j]; b < hrp[j + 1]; ++b) c += hci[a] == hci[b] ? 1 : 0;
    }
    cnt[q] = c;
  }
}
hipError_t launch_w_term_count(int nq, const int* brow, const int* bcol, const int* src, const int* okey,
                               const int* hrp, const int* hci, int* cnt, hipStream_t s) {
  (void)hipGetLastError();
  if (nq > 0) k_w_term_count<<<grid_n(nq), 256, 0, s>>>(nq, brow, bcol, src, okey, hrp, hci, cnt);
  return hipGetLastError();
}
// tptr: exclusive prefix sums (nq + 1 entries)
__global__ void k_w_term_fill(int nq, const int* __restrict__ brow, const int* __restrict__ bcol,
                              const int* __restrict__ src, const int* __restrict__ okey, const int* __restrict__ hrp,
                              const int* __restrict__ hci, const int* __restrict__ tptr, int* __restrict__ term) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    if (src[q] < 0) continue;
    int o = tptr[q];
    int i, j;
    w_block_rows(okey, brow[q], bcol[q], i, j);
    for (int a = hrp[i]; a < hrp[i + 1]; ++a)
      for (int b = hrp[j]; b < hrp[j + 1]; ++b)
        if (hci[a] == hci[b]) {
          term[2 * o] = a;
          term[2 * o + 1] = b;
          ++o;
        }
  }
}
hipError_t launch_w_term_fill(int nq, const int* brow, const int* bcol, const int* src, const int* okey,
                              const int* hrp, const int* hci, const int* tptr, int* term, hipStream_t s) {
  (void)hipGetLastError();
  if (nq > 0) k_w_term_fill<<<grid_n(nq), 256, 0, s>>>(nq, brow, bcol, src, okey, hrp, hci, tptr, term);
  return hipGetLastError();
}

// ------------------------------------------------------------------ permutation
// New contact i is old contact perm[i] (DualFrictionProblem::applyPermutation): row
// lengths, then each row's columns and 18 values, b and mu.
__global__ void k_perm_lens(int nc, const int* __restrict__ perm, const int* __restrict__ rp, int* __restrict__ len) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x)
    len[i] = rp[perm[i] + 1] - rp[perm[i]];
}
__global__ void k_perm_rows(int nc, const int* __restrict__ perm, const int* __restrict__ rp,
                            const int* __restrict__ ci, const double* __restrict__ hv, const double* __restrict__ b,
                            const double* __restrict__ mu, const int* __restrict__ nrp, int* __restrict__ nci,
                            double* __restrict__ nhv, double* __restrict__ nb, double* __restrict__ nmu) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    const int src = perm[i];
    int dst = nrp[i];
    for (int p = rp[src]; p < rp[src + 1]; ++p, ++dst) {
      nci[dst] = ci[p];
      for (int e = 0; e < 18; ++e) nhv[(size_t)dst * 18 + e] = hv[(size_t)p * 18 + e];
    }
    for (int k = 0; k < 3; ++k) nb[3 * (size_t)i + k] = b[3 * (size_t)src + k];
    nmu[i] = mu[src];
  }
}
__global__ void k_gather_int(int n, const int* __restrict__ idx, const int* __restrict__ src, int* __restrict__ dst) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[idx[i]];
}
hipError_t launch_gather_int(int n, const int* idx, const int* src, int* dst, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) k_gather_int<<<grid_n(n), 256, 0, s>>>(n, idx, src, dst);
  return hipGetLastError();
}
hipError_t launch_perm_lens(int nc, const int* perm, const int* rp, int* len, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_perm_lens<<<grid_n(nc), 256, 0, s>>>(nc, perm, rp, len);
  return hipGetLastError();
}
hipError_t launch_perm_rows(int nc, const int* perm, const int* rp, const int* ci, const double* hv, const double* b,
                            const double* mu, const int* nrp, int* nci, double* nhv, double* nb, double* nmu,
                            hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_perm_rows<<<grid_n(nc), 256, 0, s>>>(nc, perm, rp, ci, hv, b, mu, nrp, nci, nhv, nb, nmu);
  return hipGetLastError();
}

// ------------------------------------------------------------------ coloring
// First-fit coloring of the contacts (coupled: sharing a body): contact i takes the
// smallest color not used by a contact j < i coupled to it. The rule is sequential
// in the contacts, but contact i's color depends only on its earlier neighbours'
// colors, so rounds of a parallel kernel give exactly the sequential result: each
// round colors every contact whose earlier neighbours are all colored (a color once
// written is final; a contact that reads an unfinished neighbour waits for the next
// round). The rounds follow the longest chain of increasing coupled contacts.
constexpr int kColorWords = 4;  // colors < 256

// deg: the largest number of contacts at one body (atomicMax), from the body rows of H^T
__global__ void k_max_degree(int nb, const int* __restrict__ trp, int* __restrict__ deg) {
  int d = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x)
    d = max(d, trp[k + 1] - trp[k]);
  atomicMax(deg, d);
}

// One round. Body k's contacts are tc[q], q in [trp[k], trp[k + 1]), ascending;
// tsrc[q] < 0 marks an empty body row's zero block. left: contacts still uncolored;
// info: set if a contact needs more than 64 * kColorWords colors.
__global__ void __launch_bounds__(256) k_color_round(int nc, const int* __restrict__ hrp, const int* __restrict__ hci,
                                                     const int* __restrict__ trp, const int* __restrict__ tc,
                                                     const int* __restrict__ tsrc, int* color, int* left, int* info) {
  int nleft = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    if (__hip_atomic_load(&color[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= 0) continue;
    unsigned long long used[kColorWords] = {0ull, 0ull, 0ull, 0ull};
    bool ready = true;
    for (int p = hrp[i]; p < hrp[i + 1] && ready; ++p) {
      const int k = hci[p];
      for (int q = trp[k]; q < trp[k + 1]; ++q) {
        const int j = tc[q];
        if (j >= i || tsrc[q] < 0) break;
        const int c = __hip_atomic_load(&color[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (c < 0) {
          ready = false;
          break;
        }
#pragma unroll
        for (int w = 0; w < kColorWords; ++w)
          if ((c >> 6) == w) used[w] |= 1ull << (c & 63);
      }
    }
    if (!ready) {
      ++nleft;
      continue;
    }
    int c = 64 * kColorWords;
#pragma unroll
    for (int w = kColorWords - 1; w >= 0; --w)
      if (~used[w]) c = 64 * w + (int)__builtin_ctzll(~used[w]);
    if (c >= 64 * kColorWords) {
      *info = 1;
      continue;
    }
    __hip_atomic_store(&color[i], c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (nleft) atomicAdd(left, nleft);
}

hipError_t launch_max_degree(int nb, const int* trp, int* deg, hipStream_t s) {
  (void)hipGetLastError();
  hipError_t e = hipMemsetAsync(deg, 0, sizeof(int), s);
  if (e != hipSuccess) return e;
  k_max_degree<<<grid_n(nb), 256, 0, s>>>(nb, trp, deg);
  return hipGetLastError();
}
hipError_t launch_color_round(int nc, const int* hrp, const int* hci, const int* trp, const int* tc, const int* tsrc,
                              int* color, int* left, int* info, hipStream_t s) {
  (void)hipGetLastError();
  k_color_round<<<grid_n(nc), 256, 0, s>>>(nc, hrp, hci, trp, tc, tsrc, color, left, info);
  return hipGetLastError();
}
int coloring_max_colors() { return 64 * kColorWords; }

}  // namespace sb
