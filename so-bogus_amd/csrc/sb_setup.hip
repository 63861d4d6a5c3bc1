// Problem set-up on the device: refresh of every value-dependent array of a
// problem whose sparsity pattern is unchanged (new H values, inverse masses,
// affine term, friction coefficients), and the per-body inverse of the mass
// matrices. The reference does this work on the host at every setMatrix /
// computeFrom (PrimalFrictionProblem::computeMInv, PrimalFrictionProblem.cpp:28-38;
// DualFrictionProblem::computeFrom and the W product, FrictionProblem.cpp:30-40,
// SparseMatrixMatrixProduct.impl.hpp:295-349 numeric phase); the symbolic parts
// (tiling, transposition, the pattern of W) are computed once on the host when the
// problem is created and reused here through index maps.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sb_problem.hpp"
#include "sb_spmv.hpp"

namespace sb {

namespace {
inline int grid_for_n(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 16384); }
}  // namespace

// Tile values of a BR x BC block matrix from row-major blocks raw[p][BR*BC]
// (transposed source blocks raw[p][BC*BR] when Transposed), element-major
// vals[e * nq + q]; src[q] < 0 marks an empty row's zero block.
template <int BR, int BC, bool Transposed>
__global__ void __launch_bounds__(256) k_tile_values(int nq, const int* __restrict__ src,
                                                     const double* __restrict__ raw, double* __restrict__ vals,
                                                     float* __restrict__ vals32) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    const int p = src[q];
#pragma unroll
    for (int r = 0; r < BR; ++r)
#pragma unroll
      for (int c = 0; c < BC; ++c) {
        const double v = p < 0 ? 0.0 : (Transposed ? raw[(size_t)p * BR * BC + c * BR + r]
                                                   : raw[(size_t)p * BR * BC + r * BC + c]);
        vals[(size_t)(r * BC + c) * nq + q] = v;
        vals32[(size_t)(r * BC + c) * nq + q] = (float)v;
      }
  }
}

// Minv_k row-major [nb][36] -> element-major [36][nb] (FP64 and FP32).
__global__ void __launch_bounds__(256) k_minv_layout(int nb, const double* __restrict__ raw, double* __restrict__ em,
                                                     float* __restrict__ em32) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x)
#pragma unroll
    for (int e = 0; e < 36; ++e) {
      const double v = raw[(size_t)36 * k + e];
      em[(size_t)e * nb + k] = v;
      em32[(size_t)e * nb + k] = (float)v;
    }
}

// 3x3 block sum_terms H_pi Minv_k H_pj^T, k the body both blocks share.
__device__ __forceinline__ void w_block(const double* __restrict__ hraw, const double* __restrict__ mraw, int pi,
                                        int pj, int k, double acc[9]) {
  // No contraction into fused multiply-adds: the same rounding as the host assembly.
#pragma clang fp contract(off)
  const double* hi = hraw + (size_t)18 * pi;
  const double* hj = hraw + (size_t)18 * pj;
  const double* m = mraw + (size_t)36 * k;
  // t = Minv_k H_pj^T (6x3)
  double t[6][3];
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double a = 0.0;
#pragma unroll
      for (int l = 0; l < 6; ++l) a += m[r * 6 + l] * hj[c * 6 + l];
      t[r][c] = a;
    }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double a = 0.0;
#pragma unroll
      for (int l = 0; l < 6; ++l) a += hi[r * 6 + l] * t[l][c];
      acc[r * 3 + c] += a;
    }
}

// Diagonal blocks W_ii = sum over pairs (a, b) of contact i's H blocks at the same
// body k of H_a Minv_k H_b^T (row-major [nc][9]; the diagonal blocks of the
// assembled W, summed in its order) and the row scalings s_i = max(1, tr(W_ii) / 3)
// (ConstrainedSolverBase.impl.hpp:78-118).
__global__ void __launch_bounds__(256) k_diag_scal(int nc, const int* __restrict__ hrp, const int* __restrict__ hci,
                                                   const double* __restrict__ hraw, const double* __restrict__ mraw,
                                                   double* __restrict__ diag, double* __restrict__ scal) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += gridDim.x * blockDim.x) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = hrp[i]; a < hrp[i + 1]; ++a)
      for (int b = hrp[i]; b < hrp[i + 1]; ++b)
        if (hci[a] == hci[b]) w_block(hraw, mraw, a, b, hci[a], acc);
#pragma unroll
    for (int e = 0; e < 9; ++e) diag[(size_t)9 * i + e] = acc[e];
    scal[i] = fmax(1.0, (acc[0] + acc[4] + acc[8]) / 3.0);
  }
}

// Numeric phase of W = H Minv H^T on the expanded (both triangles) tiled pattern.
// Block q is the sum over its terms wterm[wtptr[q] .. wtptr[q+1]) = (p_a, p_b) of
// H_pa Minv_k H_pb^T (k the shared body), p_a of contact hi and p_b of contact lo;
// it is stored transposed where the block's row is lo (the transpose of block
// (hi, lo), whose terms it holds, as the lower-triangle storage of the symmetric W
// implies). hi is the contact with the larger index, or with okey the larger
// assembly order key (okey[c]: contact c's index when W was assembled).
__global__ void __launch_bounds__(256) k_w_numeric(int nq, const int* __restrict__ brow, const int* __restrict__ bcol,
                                                   const int* __restrict__ wtptr, const int2* __restrict__ wterm,
                                                   const int* __restrict__ okey, const int* __restrict__ hci,
                                                   const double* __restrict__ hraw,
                                                   const double* __restrict__ mraw, double* __restrict__ vals,
                                                   float* __restrict__ vals32) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int t = wtptr[q]; t < wtptr[q + 1]; ++t) {
      const int2 pq = wterm[t];
      w_block(hraw, mraw, pq.x, pq.y, hci[pq.x], acc);
    }
    const bool upper = okey ? okey[bcol[q]] > okey[brow[q]] : bcol[q] > brow[q];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double v = upper ? acc[c * 3 + r] : acc[r * 3 + c];
        vals[(size_t)(r * 3 + c) * nq + q] = v;
        vals32[(size_t)(r * 3 + c) * nq + q] = (float)v;
      }
  }
}

// Minv_k = M_k^-1 for 6x6 blocks (row-major), Gauss-Jordan elimination with partial
// pivoting on [M | I], one lane per body; a zero pivot flags the body (err = 1).
__global__ void __launch_bounds__(128) k_mass_inverse(int nb, const double* __restrict__ M, double* __restrict__ Minv,
                                                      int* __restrict__ err) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x) {
    double a[6][12];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 12; ++j) a[i][j] = j < 6 ? M[(size_t)36 * k + 6 * i + j] : (j - 6 == i ? 1.0 : 0.0);
    bool singular = false;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      // pivot row: the largest |a[i][c]|, i >= c (first maximum); swapped by selection
      int p = c;
      double big = fabs(a[c][c]);
#pragma unroll
      for (int i = c + 1; i < 6; ++i)
        if (fabs(a[i][c]) > big) {
          big = fabs(a[i][c]);
          p = i;
        }
      singular |= big == 0.0;
#pragma unroll
      for (int i = c + 1; i < 6; ++i)
        if (i == p)
#pragma unroll
          for (int j = 0; j < 12; ++j) {
            const double t = a[c][j];
            a[c][j] = a[i][j];
            a[i][j] = t;
          }
      const double inv = 1.0 / a[c][c];
#pragma unroll
      for (int j = 0; j < 12; ++j) a[c][j] *= inv;
#pragma unroll
      for (int i = 0; i < 6; ++i)
        if (i != c) {
          const double m = a[i][c];
#pragma unroll
          for (int j = 0; j < 12; ++j) a[i][j] -= m * a[c][j];
        }
    }
    if (singular) {
      atomicOr(err, 1);
      continue;
    }
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) Minv[(size_t)36 * k + 6 * i + j] = a[i][j + 6];
  }
}

// Factored operator tiles. With Minv_k = L_k L_k^T (Cholesky, L_k lower triangular),
// W = H Minv H^T = F F^T for F_ik = H_ik L_k: the matrix-free product is then the
// two tile passes F (F^T x) with no body step, exactly symmetric in any precision.
// k_minv_chol: L_k from the element-major Minv [36][nb] into element-major [36][nb]
// (upper part zero); Minv_k not symmetric (beyond 1e-12 of its largest diagonal
// entry: the factors use its lower triangle only) or a pivot that is not positive
// (Minv_k not positive definite) sets err.
__global__ void __launch_bounds__(256) k_minv_chol(int nb, const double* __restrict__ minv, double* __restrict__ L,
                                                   int* __restrict__ err) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x) {
    double a[6][6], l[6][6];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) {
        a[r][c] = minv[(size_t)(r * 6 + c) * nb + k];
        l[r][c] = 0.0;
      }
    double dmax = 0.0, asym = 0.0;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      dmax = fmax(dmax, fabs(a[r][r]));
#pragma unroll
      for (int c = 0; c < r; ++c) asym = fmax(asym, fabs(a[r][c] - a[c][r]));
    }
    bool ok = asym <= 1e-12 * dmax;
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      double d = a[c][c];
#pragma unroll
      for (int m = 0; m < c; ++m) d -= l[c][m] * l[c][m];
      ok = ok && d > 0.0;
      const double ljj = sqrt(d > 0.0 ? d : 1.0);
      l[c][c] = ljj;
#pragma unroll
      for (int r = c + 1; r < 6; ++r) {
        double t = a[r][c];
#pragma unroll
        for (int m = 0; m < c; ++m) t -= l[r][m] * l[c][m];
        l[r][c] = t / ljj;
      }
    }
    if (!ok) atomicOr(err, 1);
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) L[(size_t)(r * 6 + c) * nb + k] = l[r][c];
  }
}

// F tiles: Transposed = false: 3x6 blocks F_q = H_q L_k, k = bcol[q] (contact-major
// H tiles); Transposed = true: 6x3 blocks F_q = L_k^T HT_q, k = brow[q] (body-major
// H^T tiles). Element-major in and out, FP64 and FP32.
template <bool Transposed>
__global__ void __launch_bounds__(256) k_factor_tiles(int nq, int nb, const int* __restrict__ body,
                                                      const double* __restrict__ v, const double* __restrict__ L,
                                                      double* __restrict__ f, float* __restrict__ f32) {
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < nq; q += gridDim.x * blockDim.x) {
    const int k = body[q];
    double h[18], l[36];
#pragma unroll
    for (int e = 0; e < 18; ++e) h[e] = v[(size_t)e * nq + q];
#pragma unroll
    for (int e = 0; e < 36; ++e) l[e] = L[(size_t)e * nb + k];
#pragma unroll
    for (int r = 0; r < (Transposed ? 6 : 3); ++r)
#pragma unroll
      for (int c = 0; c < (Transposed ? 3 : 6); ++c) {
        double acc = 0.0;
        if (Transposed) {
#pragma unroll
          for (int m = r; m < 6; ++m) acc += l[m * 6 + r] * h[m * 3 + c];  // (L^T)_rm HT_mc
        } else {
#pragma unroll
          for (int m = c; m < 6; ++m) acc += h[r * 6 + m] * l[m * 6 + c];  // H_rm L_mc
        }
        const int e = Transposed ? r * 3 + c : r * 6 + c;
        f[(size_t)e * nq + q] = acc;
        f32[(size_t)e * nq + q] = (float)acc;
      }
  }
}

hipError_t launch_minv_chol(int nb, const double* minv, double* L, int* err, hipStream_t s) {
  (void)hipGetLastError();
  if (nb > 0) k_minv_chol<<<grid_for_n(nb), 256, 0, s>>>(nb, minv, L, err);
  return hipGetLastError();
}

hipError_t launch_factor_tiles(bool transposed, int nq, int nb, const int* body, const double* v, const double* L,
                               double* f, float* f32, hipStream_t s) {
  (void)hipGetLastError();
  if (nq <= 0) return hipSuccess;
  if (transposed)
    k_factor_tiles<true><<<grid_for_n(nq), 256, 0, s>>>(nq, nb, body, v, L, f, f32);
  else
    k_factor_tiles<false><<<grid_for_n(nq), 256, 0, s>>>(nq, nb, body, v, L, f, f32);
  return hipGetLastError();
}

hipError_t launch_tile_values(int br, int bc, bool transposed, int nq, const int* src, const double* raw,
                              double* vals, float* vals32, hipStream_t s) {
  (void)hipGetLastError();
  if (nq <= 0) return hipSuccess;
  const int g = grid_for_n(nq);
  if (br == 3 && bc == 6 && !transposed)
    k_tile_values<3, 6, false><<<g, 256, 0, s>>>(nq, src, raw, vals, vals32);
  else if (br == 6 && bc == 3 && transposed)
    k_tile_values<6, 3, true><<<g, 256, 0, s>>>(nq, src, raw, vals, vals32);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_minv_layout(int nb, const double* raw, double* em, float* em32, hipStream_t s) {
  (void)hipGetLastError();
  if (nb > 0) k_minv_layout<<<grid_for_n(nb), 256, 0, s>>>(nb, raw, em, em32);
  return hipGetLastError();
}

hipError_t launch_diag_scal(int nc, const int* hrp, const int* hci, const double* hraw, const double* mraw,
                            double* diag, double* scal, hipStream_t s) {
  (void)hipGetLastError();
  if (nc > 0) k_diag_scal<<<grid_for_n(nc), 256, 0, s>>>(nc, hrp, hci, hraw, mraw, diag, scal);
  return hipGetLastError();
}

hipError_t launch_w_numeric(int nq, const int* brow, const int* bcol, const int* wtptr, const int* wterm,
                            const int* okey, const int* hci, const double* hraw, const double* mraw, double* vals,
                            float* vals32, hipStream_t s) {
  (void)hipGetLastError();
  if (nq > 0)
    k_w_numeric<<<grid_for_n(nq), 256, 0, s>>>(nq, brow, bcol, wtptr, reinterpret_cast<const int2*>(wterm), okey,
                                                hci, hraw, mraw, vals, vals32);
  return hipGetLastError();
}

hipError_t launch_mass_inverse(int nb, const double* M, double* Minv, int* err, hipStream_t s) {
  (void)hipGetLastError();
  if (nb > 0) k_mass_inverse<<<(int)std::min<int64_t>((nb + 127) / 128, 16384), 128, 0, s>>>(nb, M, Minv, err);
  return hipGetLastError();
}

}  // namespace sb
