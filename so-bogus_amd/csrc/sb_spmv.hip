// Stand-alone block SpMV kernels on the tiled layout (one wave per tile):
//   * y = alpha op(A) x + beta y for a generic BSR (SparseBlockMatrixBase::multiply,
//     src/Core/Block/SparseMatrixVectorProduct.impl.hpp:382-432), block sizes <= 8;
//   * the friction operators: W x (assembled), H w, Minv H^T r, and the matrix-free
//     W x = H (Minv (H^T x)) (ProductVector::multiply, :355-376).
// Symmetric matrices are stored with both triangles expanded and transposed
// products use a cached transposed copy, so every product is a pure gather: no
// atomics, deterministic results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "sb_law.hpp"
#include "sb_problem.hpp"
#include "sb_spmv.hpp"

namespace sb {

// ------------------------------------------------------------------ generic BSR
template <int MAXB>
__global__ void __launch_bounds__(256) k_bsr_generic(TileMat<double> A, int br, int bc, double alpha,
                                                     const double* __restrict__ x, double beta, double* y,
                                                     double* carry) {
  const int wpb = blockDim.x / SB_WAVE;
  const int t = blockIdx.x * wpb + wave_in_block();
  if (t >= A.ntiles) return;
  const int l = lane_id();
  const int q0 = A.tstart[t], q1 = A.tstart[t + 1];
  const int q = q0 + l;
  const bool act = q < q1;
  int row = -1;
  double v[MAXB];
  for (int r = 0; r < MAXB; ++r) v[r] = 0.0;
  if (act) {
    row = A.brow[q];
    const int col = A.bcol[q];
    for (int r = 0; r < br; ++r) {
      double acc = 0.0;
      for (int c = 0; c < bc; ++c) acc += A.vals[(size_t)(r * bc + c) * A.nblocks + q] * x[(size_t)col * bc + c];
      v[r] = acc;
    }
  }
  for (int s = 0; s < A.scan_steps; ++s) {
    const int d = 1 << s;
    const int ru = __shfl_up(row, d, SB_WAVE);
    for (int r = 0; r < MAXB; ++r) {
      const double vu = __shfl_up(v[r], d, SB_WAVE);
      if (l >= d && ru == row) v[r] += vu;
    }
  }
  const int rn = __shfl_down(row, 1, SB_WAVE);
  if (act && (l == SB_WAVE - 1 || rn != row)) {
    if (A.tlong[t] >= 0) {
      for (int r = 0; r < br; ++r) carry[(size_t)t * MAXB + r] = v[r];
    } else {
      for (int r = 0; r < br; ++r) {
        const size_t o = (size_t)row * br + r;
        y[o] = alpha * v[r] + (beta == 0.0 ? 0.0 : beta * y[o]);
      }
    }
  }
}

__global__ void k_bsr_generic_long(TileMat<double> A, int br, double alpha, double beta, double* y,
                                   const double* carry) {
  const int wpb = blockDim.x / SB_WAVE;
  const int k = blockIdx.x * wpb + wave_in_block();
  if (k >= A.nlong) return;
  const int t0 = A.long_t0[k], t1 = A.long_t1[k], row = A.long_row[k];
  for (int r = 0; r < br; ++r) {
    double acc = 0.0;
    for (int t = t0 + lane_id(); t < t1; t += SB_WAVE) acc += carry[(size_t)t * 8 + r];
    acc = wave_sum(acc);
    if (lane_id() == 0) {
      const size_t o = (size_t)row * br + r;
      y[o] = alpha * acc + (beta == 0.0 ? 0.0 : beta * y[o]);
    }
  }
}

hipError_t launch_bsr_generic(const TileMat<double>& A, int br, int bc, double alpha, const double* x, double beta,
                              double* y, double* carry, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  const int threads = 256, wpb = threads / SB_WAVE;
  if (A.ntiles > 0)
    k_bsr_generic<8><<<(A.ntiles + wpb - 1) / wpb, threads, 0, s>>>(A, br, bc, alpha, x, beta, y, carry);
  if (A.nlong > 0) k_bsr_generic_long<<<(A.nlong + wpb - 1) / wpb, threads, 0, s>>>(A, br, alpha, beta, y, carry);
  return hipGetLastError();
}

// ------------------------------------------------------------------ friction operators
namespace {
struct VecLoader3 {
  const double* x;
  __device__ void operator()(int j, double xc[3]) const {
    xc[0] = x[3 * j];
    xc[1] = x[3 * j + 1];
    xc[2] = x[3 * j + 2];
  }
};
struct VecLoader6 {
  const double* x;
  __device__ void operator()(int j, double xc[6]) const {
#pragma unroll
    for (int r = 0; r < 6; ++r) xc[r] = x[6 * j + r];
  }
};
struct Store3 {
  double* y;
  __device__ void operator()(int i, double v[3], NoPre::type = {}) const {
    y[3 * i] = v[0];
    y[3 * i + 1] = v[1];
    y[3 * i + 2] = v[2];
  }
};
struct StoreMinv {  // w_k = Minv_k v
  double* w;
  const double* minv;
  int nb;
  __device__ void operator()(int k, double v[6], NoPre::type = {}) const {
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) acc += minv[(size_t)(r * 6 + c) * nb + k] * v[c];
      w[6 * k + r] = acc;
    }
  }
};
}  // namespace

template <int BR, int BC, typename LD, typename LE>
__global__ void __launch_bounds__(256) k_tiles(TileMat<double> A, LD ld, LE le, double* carry) {
  const int wpb = blockDim.x / SB_WAVE;
  const int t = blockIdx.x * wpb + wave_in_block();
  if (t >= A.ntiles) return;
  NoPre pre;
  // the block values are streamed once: non-temporal loads keep the gathered vectors cached
  tile_gather<BR, BC, double, double, true>(A, t, tile_indices(A, t), ld, pre, le, carry);
}

// One block per long row: the row's carries (written by the previous kernel) are
// summed by every thread with four tiles in flight, then across the block in a
// fixed order.
template <int BR, typename LE>
__global__ void __launch_bounds__(256) k_long(TileMat<double> A, LE le, const double* __restrict__ carry) {
  __shared__ double s_v[BR][4];
  const int k = blockIdx.x;
  const int t0 = A.long_t0[k], t1 = A.long_t1[k];
  double v[BR];
#pragma unroll
  for (int r = 0; r < BR; ++r) v[r] = 0.0;
  int t = t0 + threadIdx.x;
  for (; t + 3 * 256 < t1; t += 4 * 256) {
    double c[4][BR];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < BR; ++r) c[j][r] = carry[(size_t)(t + j * 256) * BR + r];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < BR; ++r) v[r] += c[j][r];
  }
  for (; t < t1; t += 256) {
#pragma unroll
    for (int r = 0; r < BR; ++r) v[r] += carry[(size_t)t * BR + r];
  }
#pragma unroll
  for (int r = 0; r < BR; ++r) {
    v[r] = wave_sum(v[r]);
    if (lane_id() == 0) s_v[r][wave_in_block()] = v[r];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int r = 0; r < BR; ++r) v[r] = (s_v[r][0] + s_v[r][1]) + (s_v[r][2] + s_v[r][3]);
    le(A.long_row[k], v);
  }
}

template <int BR, int BC, typename LD, typename LE>
static void run_tiles(const TileMat<double>& A, LD ld, LE le, double* carry, hipStream_t s) {
  const int threads = 256, wpb = threads / SB_WAVE;
  if (A.ntiles > 0) k_tiles<BR, BC><<<(A.ntiles + wpb - 1) / wpb, threads, 0, s>>>(A, ld, le, carry);
  if (A.nlong > 0) k_long<BR><<<A.nlong, 256, 0, s>>>(A, le, carry);
}

hipError_t launch_w(const DevProblem<double>& P, const double* x, double* y, double* carry, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  run_tiles<3, 3>(P.W, VecLoader3{x}, Store3{y}, carry, s);
  return hipGetLastError();
}
// y = H w for an H whose rows all hold two blocks (a contact between two bodies),
// blocks 2i and 2i + 1 for row i: one lane per row, no scan. Element e of the two
// blocks is adjacent in the element-major layout, so a lane reads both with one
// 16-byte non-temporal load (a wave: 1 KB contiguous per element); a body vector is
// three 16-byte loads. The row sum is formed as in the tiled kernel (each block's
// products in order, then the second block's sum plus the first's): the same bits.
// A wave's 64 results (1536 contiguous bytes of y) go through LDS and are written
// with 16-byte non-temporal stores. Aligned: w and y 16-byte aligned, else 8-byte
// accesses.
typedef double dbl2 __attribute__((ext_vector_type(2)));
template <bool Aligned>
__global__ void __launch_bounds__(256) k_h_pairs(TileMat<double> A, const double* __restrict__ w,
                                                 double* __restrict__ y) {
  __shared__ double sy[4][3 * 64];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nc = A.nrows;
  if (i < nc) {
    const size_t nb2 = (size_t)A.nblocks / 2;
    const dbl2* v2 = reinterpret_cast<const dbl2*>(A.vals) + i;
    const int ca = A.bcol[2 * i], cb = A.bcol[2 * i + 1];
    dbl2 e[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) e[k] = __builtin_nontemporal_load(v2 + (size_t)k * nb2);
    double xa[6], xb[6];
    if (Aligned) {
      const dbl2* wa = reinterpret_cast<const dbl2*>(w + 6 * (size_t)ca);
      const dbl2* wb = reinterpret_cast<const dbl2*>(w + 6 * (size_t)cb);
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const dbl2 a2 = wa[k], b2 = wb[k];
        xa[2 * k] = a2.x;
        xa[2 * k + 1] = a2.y;
        xb[2 * k] = b2.x;
        xb[2 * k + 1] = b2.y;
      }
    } else {
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        xa[c] = w[6 * (size_t)ca + c];
        xb[c] = w[6 * (size_t)cb + c];
      }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      double aa = 0.0, ab = 0.0;
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        aa += e[r * 6 + c].x * xa[c];
        ab += e[r * 6 + c].y * xb[c];
      }
      sy[wv][3 * l + r] = ab + aa;
    }
  }
  __syncthreads();  // every thread of the block reaches it
  const int i0 = i - l;
  if (i0 < nc) {
    const int nd = 3 * min(64, nc - i0);  // the wave's results
    double* yo = y + 3 * (size_t)i0;
    if (Aligned) {
      dbl2* yo2 = reinterpret_cast<dbl2*>(yo);
      for (int k = l; k < nd / 2; k += 64) {
        const dbl2 v = {sy[wv][2 * k], sy[wv][2 * k + 1]};
        __builtin_nontemporal_store(v, yo2 + k);
      }
      if ((nd & 1) && l == 0) yo[nd - 1] = sy[wv][nd - 1];
    } else {
      for (int k = l; k < nd; k += 64) yo[k] = sy[wv][k];
    }
  }
}

hipError_t launch_h(const DevProblem<double>& P, const double* w, double* y, double* carry, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  if (P.H.pairs) {
    const bool aligned = ((uintptr_t)w % 16 == 0) && ((uintptr_t)y % 16 == 0);
    const unsigned grid = (unsigned)((P.H.nrows + 255) / 256);
    if (P.H.nrows > 0) {
      if (aligned) k_h_pairs<true><<<grid, 256, 0, s>>>(P.H, w, y);
      else k_h_pairs<false><<<grid, 256, 0, s>>>(P.H, w, y);
    }
  } else {
    run_tiles<3, 6>(P.H, VecLoader6{w}, Store3{y}, carry, s);
  }
  return hipGetLastError();
}
hipError_t launch_ht_minv(const DevProblem<double>& P, const double* r, double* w, double* carry, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  run_tiles<6, 3>(P.HT, VecLoader3{r}, StoreMinv{w, P.minv, P.nb}, carry, s);
  return hipGetLastError();
}
hipError_t launch_matfree(const DevProblem<double>& P, const double* x, double* y, double* wbody, double* carry,
                          hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  const hipError_t e = launch_ht_minv(P, x, wbody, carry, s);
  if (e != hipSuccess) return e;
  return launch_h(P, wbody, y, carry, s);
}

// ------------------------------------------------------------------ residual
// u = (W r or matrix-free) + b is formed first; this kernel reduces the FB terms.
__global__ void __launch_bounds__(256) k_fb_terms(int nc, int coulomb, const double* __restrict__ mu,
                                                  const double* __restrict__ scal, const double* __restrict__ u,
                                                  const double* __restrict__ b, const double* __restrict__ r,
                                                  double* terms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const double s = scal ? scal[i] : 1.0;
  const double x[3] = {s * r[3 * i], s * r[3 * i + 1], s * r[3 * i + 2]};
  const double y[3] = {u[3 * i] + b[3 * i], u[3 * i + 1] + b[3 * i + 1], u[3 * i + 2] + b[3 * i + 2]};
  terms[i] = law_eval<double>(coulomb != 0, mu[i], x, y);
}

hipError_t launch_fb_terms(int nc, int coulomb, const double* mu, const double* scal, const double* u,
                           const double* b, const double* r, double* terms, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  if (nc > 0) k_fb_terms<<<(nc + 255) / 256, 256, 0, s>>>(nc, coulomb, mu, scal, u, b, r, terms);
  return hipGetLastError();
}

// FCLib's merit1 terms (apps/FCLibLoader.cpp:40-61): with u_i = (W r)_i + b_i,
// x = r_i - u_i, x_N -= mu_i |u_iT|, x projected on the cone of aperture mu_i
// (SOCLaw::projectOnConstraint); the term is |x - r_i|^2.
__global__ void __launch_bounds__(256) k_merit1_terms(int nc, const double* __restrict__ mu,
                                                      const double* __restrict__ wr, const double* __restrict__ b,
                                                      const double* __restrict__ r, double* __restrict__ terms) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const double u0 = wr[3 * i] + b[3 * i], u1 = wr[3 * i + 1] + b[3 * i + 1], u2 = wr[3 * i + 2] + b[3 * i + 2];
  const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
  double x0 = r0 - u0, x1 = r1 - u1, x2 = r2 - u2;
  x0 -= mu[i] * sqrt(u1 * u1 + u2 * u2);
  project<double>(mu[i], x0, x1, x2);
  const double d0 = x0 - r0, d1 = x1 - r1, d2 = x2 - r2;
  terms[i] = d0 * d0 + d1 * d1 + d2 * d2;
}
hipError_t launch_merit1_terms(int nc, const double* mu, const double* wr, const double* b, const double* r,
                               double* terms, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_merit1_terms<<<(nc + 255) / 256, 256, 0, s>>>(nc, mu, wr, b, r, terms);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Cadoux fixed point
// bout = b0 + s(Wx + b0): the affine term of the next SOC problem, s the De Saxce
// correction (mu |u_T|, 0, 0) of u = W x + b0 (Cadoux.hpp:72-76).
__global__ void __launch_bounds__(256) k_cadoux_rhs(int nc, const double* __restrict__ mu,
                                                    const double* __restrict__ wx, const double* __restrict__ b0,
                                                    double* __restrict__ bout) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  const double u1 = wx[3 * i + 1] + b0[3 * i + 1], u2 = wx[3 * i + 2] + b0[3 * i + 2];
  bout[3 * i] = b0[3 * i] + cov_normal<double>(true, mu[i], u1, u2);
  bout[3 * i + 1] = b0[3 * i + 1];
  bout[3 * i + 2] = b0[3 * i + 2];
}
// s = (|v_T| max(0, 1/mu), 0, 0) (Cadoux.hpp:119-123, 139-143)
__global__ void __launch_bounds__(256) k_vel_shift(int nc, const double* __restrict__ mu, const double* __restrict__ v,
                                                   double* __restrict__ sv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nc) return;
  sv[3 * i] = sqrt(v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * fmax(0.0, 1.0 / mu[i]);
  sv[3 * i + 1] = 0.0;
  sv[3 * i + 2] = 0.0;
}
// out = a x + c y
__global__ void __launch_bounds__(256) k_lincomb(size_t n, double a, const double* __restrict__ x, double c,
                                                 const double* __restrict__ y, double* __restrict__ out) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    out[k] = a * x[k] + c * y[k];
}

// Contact vectors between two orders: gather dst[i] = src[perm[i]] or scatter
// dst[perm[i]] = src[i], 3 doubles per contact.
__global__ void __launch_bounds__(256) k_permute3(int64_t n, const int* __restrict__ perm,
                                                  const double* __restrict__ src, double* __restrict__ dst,
                                                  int scatter) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = perm[i];
    const int64_t from = scatter ? i : j, to = scatter ? j : i;
    dst[3 * to] = src[3 * from];
    dst[3 * to + 1] = src[3 * from + 1];
    dst[3 * to + 2] = src[3 * from + 2];
  }
}
hipError_t launch_permute3(int64_t n, const int* perm, const double* src, double* dst, bool scatter, hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) {
    const int grid = (int)std::min<int64_t>((n + 255) / 256, 8192);
    k_permute3<<<grid, 256, 0, s>>>(n, perm, src, dst, scatter ? 1 : 0);
  }
  return hipGetLastError();
}

hipError_t launch_cadoux_rhs(int nc, const double* mu, const double* wx, const double* b0, double* bout,
                             hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_cadoux_rhs<<<(nc + 255) / 256, 256, 0, s>>>(nc, mu, wx, b0, bout);
  return hipGetLastError();
}
hipError_t launch_vel_shift(int nc, const double* mu, const double* v, double* sv, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_vel_shift<<<(nc + 255) / 256, 256, 0, s>>>(nc, mu, v, sv);
  return hipGetLastError();
}
hipError_t launch_lincomb(size_t n, double a, const double* x, double c, const double* y, double* out,
                          hipStream_t s) {
  (void)hipGetLastError();
  if (n > 0) {
    const int grid = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_lincomb<<<grid, 256, 0, s>>>(n, a, x, c, y, out);
  }
  return hipGetLastError();
}

// Plain streaming copy (bandwidth and counter calibration).
typedef float f32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy(const f32x4* __restrict__ src, f32x4* __restrict__ dst, size_t n) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + k), dst + k);
}

// One 16-byte element per thread, non-temporal both ways, over a grid that covers
// the buffer (grid-stride only beyond 2^22 workgroups): measured on MI355X with
// 1 GiB each way, 6.5 TB/s, against 6.2 TB/s for plain 16-byte accesses, 5.8 TB/s
// for 32-byte ones, 5.0-5.4 TB/s for 1k-16k workgroups that loop and 4.85 TB/s for
// hipMemcpyAsync. bytes: a multiple of 16.
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  if (bytes % 16 != 0) return hipErrorInvalidValue;
  const size_t n = bytes / 16;
  const size_t grid = std::min<size_t>((n + 255) / 256, (size_t)1 << 22);
  k_copy<<<dim3((unsigned)std::max<size_t>(grid, 1)), 256, 0, s>>>((const f32x4*)src, (f32x4*)dst, n);
  return hipGetLastError();
}

// Zeroes the barrier words before a persistent launch (a kernel, not a memset, so
// that the sequence can be captured in a graph).
__global__ void k_reset_sync(unsigned* words, int n, int* error) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) words[k] = 0u;
  if (threadIdx.x == 0) *error = 0;
}

hipError_t launch_reset_sync(unsigned* words, int n, int* error, hipStream_t s) {
  (void)hipGetLastError();  // clear a stale error: report this launch only
  k_reset_sync<<<1, 256, 0, s>>>(words, n, error);
  return hipGetLastError();
}

}  // namespace sb
