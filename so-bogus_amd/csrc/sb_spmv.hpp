// Launchers of the stand-alone block SpMV and helper kernels (sb_spmv.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "sb_problem.hpp"

namespace sb {

hipError_t launch_bsr_generic(const TileMat<double>& A, int br, int bc, double alpha, const double* x, double beta,
                              double* y, double* carry, hipStream_t s);
hipError_t launch_w(const DevProblem<double>& P, const double* x, double* y, double* carry, hipStream_t s);
hipError_t launch_h(const DevProblem<double>& P, const double* w, double* y, double* carry, hipStream_t s);
hipError_t launch_ht_minv(const DevProblem<double>& P, const double* r, double* w, double* carry, hipStream_t s);
hipError_t launch_matfree(const DevProblem<double>& P, const double* x, double* y, double* wbody, double* carry,
                          hipStream_t s);
// problem set-up on the device (sb_setup.hip)
hipError_t launch_minv_chol(int nb, const double* minv, double* L, int* err, hipStream_t s);
hipError_t launch_factor_tiles(bool transposed, int nq, int nb, const int* body, const double* v, const double* L,
                               double* f, float* f32, hipStream_t s);
hipError_t launch_tile_values(int br, int bc, bool transposed, int nq, const int* src, const double* raw,
                              double* vals, float* vals32, hipStream_t s);
hipError_t launch_minv_layout(int nb, const double* raw, double* em, float* em32, hipStream_t s);
hipError_t launch_diag_scal(int nc, const int* hrp, const int* hci, const double* hraw, const double* mraw,
                            double* diag, double* scal, hipStream_t s);
// okey: the contacts' assembly order keys (nullptr: their indices)
hipError_t launch_w_numeric(int nq, const int* brow, const int* bcol, const int* wtptr, const int* wterm,
                            const int* okey, const int* hci, const double* hraw, const double* mraw, double* vals,
                            float* vals32, hipStream_t s);
hipError_t launch_mass_inverse(int nb, const double* M, double* Minv, int* err, hipStream_t s);
hipError_t launch_permute3(int64_t n, const int* perm, const double* src, double* dst, bool scatter, hipStream_t s);
hipError_t launch_cadoux_rhs(int nc, const double* mu, const double* wx, const double* b0, double* bout,
                             hipStream_t s);
hipError_t launch_vel_shift(int nc, const double* mu, const double* v, double* sv, hipStream_t s);
hipError_t launch_lincomb(size_t n, double a, const double* x, double c, const double* y, double* out,
                          hipStream_t s);
hipError_t launch_merit1_terms(int nc, const double* mu, const double* wr, const double* b, const double* r,
                               double* terms, hipStream_t s);
hipError_t launch_fb_terms(int nc, int coulomb, const double* mu, const double* scal, const double* u,
                           const double* b, const double* r, double* terms, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
hipError_t launch_reset_sync(unsigned* words, int n, int* error, hipStream_t s);

template <bool MF, typename T>
__global__ void pg_kernel(PGGridArgs<T> a);
template <bool MF, typename T>
__global__ void pg_kernel_sc(PGGridArgs<T> a);
template <typename T>
__global__ void pg_kernel_cand(PGGridArgs<T> a);
template <bool MF, typename T, int NT>
__global__ void pg_batch_kernel(PGBatchArgs<T> a, const PGProblem<T>* __restrict__ probs);
__global__ void gs_kernel(GSArgs a, const int* wrowptr);
__global__ void pgs_kernel(GSArgs a, PGSArgs e);
__global__ void adjoint_kernel(AdjArgs a);

}  // namespace sb
