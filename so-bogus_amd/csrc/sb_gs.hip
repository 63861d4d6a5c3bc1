// Persistent contact-colored projected block Gauss-Seidel.
//
// GaussSeidel::solveWithLinearConstraints / innerLoop
// (src/Core/BlockSolvers/GaussSeidel.impl.hpp:62-124, 144-205) with
// GaussSeidelBase::evalAndKeepBest / tryZero (GaussSeidelBase.impl.hpp:44-77).
// Contacts are permuted so that every color is a contiguous range; contacts of
// one color do not share a body, so they are solved in parallel, one lane per
// contact: lb = b_i - reg_i x_i + sum_{j != i} W_ij x_j, then the FP64 local
// solver, the damped update and the sleep heuristic. A grid barrier separates
// colors. The reference step-size reference ndxRef is a racy running maximum;
// here each color reads the value left by the previous colors and the maxima
// of the color are merged at the color's barrier (deterministic).
#include <hip/hip_runtime.h>

#include "sb_law.hpp"
#include "sb_local.hpp"
#include "sb_op.hpp"
#include "sb_problem.hpp"
#include "sb_team.hpp"

namespace sb {

namespace {
// read-only inputs
__device__ __forceinline__ void ld3(const double* p, int i, double v[3]) {
  v[0] = p[3 * i];
  v[1] = p[3 * i + 1];
  v[2] = p[3 * i + 2];
}
// x is updated inside the kernel: stores are coherent (cst) and every read of a
// value written by another block follows a barrier (whose acquire dropped stale
// cached copies), so reads are plain loads.
__device__ __forceinline__ void cld3(const double* p, int i, double v[3]) {
  v[0] = mld(p + 3 * i);
  v[1] = mld(p + 3 * i + 1);
  v[2] = mld(p + 3 * i + 2);
}
// Lanes per contact in a color phase: a contact gets g lanes (a power of two, at most
// kSpread, as many as the grid holds for the color in one pass), of which the first
// min(g, kGroup) gather the blocks of the contact's W row together (lane k: blocks k,
// k + g', ...), so a row costs one dependent memory round trip for its columns and
// one for its values and x_j; their sums are combined by butterfly shuffles, and the
// first lane solves. Spreading the contacts over every wave of the grid (one per wave
// when the color is small enough) keeps the local solves of different contacts out of
// one another's divergent paths: a wave runs the union of its solving lanes' paths.
constexpr int kGroup = 16;
constexpr int kSpread = 64;
template <typename T>
__device__ __forceinline__ void group_gather(const TileMat<T>& W, size_t nnz, int q0, int q1, int i, int sub, int g,
                                             const T* x, T acc[3]) {
  acc[0] = acc[1] = acc[2] = T(0);
  for (int q = q0 + sub; q < q1; q += g) {
    const int j = W.bcol[q];
    if (j == i) continue;
    T xj[3], w[9];
#pragma unroll
    for (int c = 0; c < 3; ++c) xj[c] = mld(x + 3 * j + c);
#pragma unroll
    for (int e = 0; e < 9; ++e) w[e] = W.vals[(size_t)e * nnz + q];
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += w[3 * r + 0] * xj[0] + w[3 * r + 1] * xj[1] + w[3 * r + 2] * xj[2];
  }
  for (int m = g >> 1; m >= 1; m >>= 1)
#pragma unroll
    for (int r = 0; r < 3; ++r) acc[r] += __shfl_xor(acc[r], m, kGroup);
}
}  // namespace

// One launch runs one or two stages of sweeps. FP64: the reference's loop. Mixed
// precision: a first stage whose off-diagonal gathers (W_ij x_j) and evaluation
// product read FP32 copies of W and x (the local right-hand side and the local
// solve stay FP64, as does the FB evaluation of the FP32 product) runs to
// tol * stop_factor, or until an evaluation improves on the stage's best by less
// than 1%; it hands its best iterate to an FP64 stage (skip state and step
// reference reset), whose first FP64 evaluation follows one sweep (the final
// refinement sweep), then every eval_every sweeps, within max(eval_every,
// max_iters - FP32 sweeps) sweeps. The result is the best FP64-evaluated iterate of
// that stage (the handed-over one included).
template <bool F32>
struct StageTag {
  static constexpr bool value = F32;
};

__global__ void __launch_bounds__(256) gs_kernel(GSArgs a, const int* __restrict__ wrowptr) {
  const int wpb = blockDim.x / SB_WAVE;
  const int gw = blockIdx.x * wpb + wave_in_block();
  const int nwaves = gridDim.x * wpb;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int nthreads = gridDim.x * blockDim.x;
  const int n = a.nc, n3 = 3 * n;
  const bool coulomb = a.coulomb != 0;
  const unsigned resMax = a.inf_norm ? 1u : 0u;
  const TileMat<double>& W = a.P.W;
  const TileMat<float>& W32 = a.W32;
  const size_t nnz = (size_t)W.nblocks;
  Barrier bar;
  bar.gs = a.gs;
  bar.rank = blockIdx.x;
  const double absSkipTol = fmin(a.skip_tol, a.tol);
  int nfail = 0, nEval = 0;

  // Sweeps until an evaluation is below stopTol or maxIt sweeps; errBest / x_best of
  // the stage; returns false on a barrier failure.
  auto stage = [&](auto tag, bool tryZero, int maxIt, double stopTol, int firstEval, int& itDone, double& errBest,
                   double& errLast, bool& converged) -> bool {
    constexpr bool kF32 = decltype(tag)::value;
    // y = W x + b ; err = eval(y, x) [and err0 = eval(b, 0) when asked]
    auto eval_phase = [&](bool withZero, double& err, double& err0) -> bool {
      double part[2] = {0.0, 0.0};
      NoPre nopre;
      auto lead = [&](int i, auto* v, NoPre::type) {
        double b[3], xi[3];
        ld3(a.P.b, i, b);
        cld3(a.x, i, xi);
        const double s = a.P.scal[i], mu = a.P.mu[i];
        const double y[3] = {(double)v[0] + b[0], (double)v[1] + b[1], (double)v[2] + b[2]};
        const double xs[3] = {s * xi[0], s * xi[1], s * xi[2]};
        const double r = law_eval<double>(coulomb, mu, xs, y);
        part[0] = resMax ? fmax(part[0], r) : part[0] + r;
        if (withZero) {
          const double z[3] = {0.0, 0.0, 0.0};
          const double r0 = law_eval<double>(coulomb, mu, z, b);
          part[1] = resMax ? fmax(part[1], r0) : part[1] + r0;
        }
      };
      if constexpr (kF32) {
        auto ld = [&](int j, float xc[3]) {
          xc[0] = mld(a.x32 + 3 * j);
          xc[1] = mld(a.x32 + 3 * j + 1);
          xc[2] = mld(a.x32 + 3 * j + 2);
        };
        tile_loop<3, 3, float, float>(W32, gw, nwaves, ld, nopre, lead, a.carry32);
        if (W32.nlong > 0) {
          if (!bar.sync()) return false;
          for (int k = gw; k < W32.nlong; k += nwaves) {
            float v[3];
            long_row_sum<3, float, float>(W32, k, a.carry32, v);
            if (lane_id() == 0) lead(W32.long_row[k], v, NoPre::type{});
          }
        }
      } else {
        auto ld = [&](int j, double xc[3]) { cld3(a.x, j, xc); };
        tile_loop<3, 3, double, double>(W, gw, nwaves, ld, nopre, lead, a.carry);
        if (W.nlong > 0) {
          if (!bar.sync()) return false;
          for (int k = gw; k < W.nlong; k += nwaves) {
            double v[3];
            long_row_sum<3, double, double>(W, k, a.carry, v);
            if (lane_id() == 0) lead(W.long_row[k], v, NoPre::type{});
          }
        }
      }
      if (!bar.reduce<2>(part, resMax | (resMax << 1))) return false;
      err = bar.red[0];
      err0 = bar.red[1];
      if (!a.inf_norm) {
        err /= (1.0 + n);
        err0 /= (1.0 + n);
      }
      return true;
    };

    errBest = 1.7976931348623157e308;
    converged = false;
    itDone = 0;
    double err = 0.0, err0 = 0.0;
    bool ok = eval_phase(tryZero, err, err0);
    if (ok) {
      if (err < errBest) {
        errBest = err;
        for (int k = tid; k < n3; k += nthreads) cst(&a.xbest[k], cld(&a.x[k]));
      }
      if (tryZero && err0 < errBest) {
        errBest = err0;
        for (int k = tid; k < n3; k += nthreads) {
          cst(&a.xbest[k], 0.0);
          cst(&a.x[k], 0.0);
          if constexpr (kF32) cst(&a.x32[k], 0.0f);
        }
      }
      if (blockIdx.x == 0 && threadIdx.x == 0 && nEval < a.hist_cap) a.hist[nEval] = errBest;
      // x/xbest writes above are consumed by other blocks in the color phases
      ok = bar.sync();
    }
    errLast = err;
    ++nEval;

    double ndxRef = 0.0, prevBest = errBest;
    int it = 1;
    for (; ok && it <= maxIt; ++it) {
      for (int col = 0; col < a.ncolors && ok; ++col) {
        const int c0 = a.colors[col], c1 = a.colors[col + 1];
        double part[1] = {0.0};
        // gs consecutive lanes per contact (gs divides 64 and nthreads is a multiple of 64:
        // a group never straddles a wave and leaves the loop as a whole); the gather's
        // width g and so its sums do not depend on gs
        int gs = kSpread;
        while (gs > 1 && (c1 - c0) * gs > nthreads) gs >>= 1;
        const int g = gs < kGroup ? gs : kGroup;  // the gathering lanes of a contact
        for (int sg = tid; sg < (c1 - c0) * gs; sg += nthreads) {
          const int i = c0 + sg / gs, sub = sg % gs;
          if (sub >= g) continue;
          const unsigned char sk = a.skip[i];
          if (sk) {
            if (sub == 0) a.skip[i] = (unsigned char)(sk - 1);
            continue;
          }
          // the solving lane's own data first, then the group's gather
          double lx[3] = {0.0, 0.0, 0.0}, bi[3] = {0.0, 0.0, 0.0}, reg = 0.0, A[3][3];
          if (sub == 0) {
            cld3(a.x, i, lx);
            ld3(a.P.b, i, bi);
            reg = a.reg[i];
#pragma unroll
            for (int e = 0; e < 9; ++e) A[e / 3][e % 3] = a.local[(size_t)9 * i + e];
          }
          double lb[3], x0[3];
          if constexpr (kF32) {
            float acc[3];
            group_gather<float>(W32, nnz, wrowptr[i], wrowptr[i + 1], i, sub, g, a.x32, acc);
            lb[0] = (double)acc[0];
            lb[1] = (double)acc[1];
            lb[2] = (double)acc[2];
          } else {
            group_gather<double>(W, nnz, wrowptr[i], wrowptr[i + 1], i, sub, g, a.x, lb);
          }
          if (sub != 0) continue;
          x0[0] = lx[0];
          x0[1] = lx[1];
          x0[2] = lx[2];
#pragma unroll
          for (int r = 0; r < 3; ++r) lb[r] += bi[r] - reg * lx[r];
          const double ldx0[3] = {-lx[0], -lx[1], -lx[2]};
          const double s = a.P.scal[i];
          const double lres = local::solve(coulomb, a.strategy, A, lb, lx, a.P.mu[i], a.local_tol, s);
          const bool lok = a.local_tol > lres;
          double ldx[3] = {ldx0[0] + lx[0], ldx0[1] + lx[1], ldx0[2] + lx[2]};
          if (!lok) {
            ldx[0] *= 0.5;
            ldx[1] *= 0.5;
            ldx[2] *= 0.5;
            ++nfail;
          }
          // x_i was read above (lx before the local solve); write x_i + ldx
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const double v = x0[r] + ldx[r];
            cst(&a.x[3 * i + r], v);
            if constexpr (kF32) cst(&a.x32[3 * i + r], (float)v);
          }
          const double s2 = s * s;
          const double nx2 = s2 * (lx[0] * lx[0] + lx[1] * lx[1] + lx[2] * lx[2]);
          const double ndx2 = s2 * (ldx[0] * ldx[0] + ldx[1] * ldx[1] + ldx[2] * ldx[2]);
          part[0] = fmax(part[0], ndx2);
          if (fmin(nx2, ndx2) < absSkipTol || ndx2 < a.skip_tol * fmin(nx2, ndxRef))
            a.skip[i] = (unsigned char)a.skip_iters;
        }
        ok = bar.reduce<1>(part, 1u);
        if (ok) ndxRef = fmax(ndxRef, bar.red[0]);
      }
      if (!ok) break;
      if (it % a.eval_every == 0 || it == firstEval) {
        double e0;
        ok = eval_phase(false, err, e0);
        if (!ok) break;
        errLast = err;
        if (blockIdx.x == 0 && threadIdx.x == 0 && nEval < a.hist_cap) a.hist[nEval] = err;
        ++nEval;
        if (err < errBest) {
          errBest = err;
          for (int k = tid; k < n3; k += nthreads) cst(&a.xbest[k], cld(&a.x[k]));
        }
        if (err < stopTol) {
          converged = true;
          break;
        }
        // the FP32 stage also ends when an evaluation no longer improves on the best by
        // 1% (the rounding floor of its gathers is reached)
        if (kF32 && !(err < 0.99 * prevBest)) break;
        prevBest = fmin(prevBest, err);
        ndxRef /= a.eval_every;
        // xbest must be complete before the next sweep modifies x
        ok = bar.sync();
      }
    }
    itDone = it > maxIt ? maxIt : it;
    return ok;
  };

  double errBest = 0.0, errLast = 0.0, err1 = 0.0;
  bool converged = false, ok = true;
  int it = 0, it1 = 0, nEval1 = 0;
  if (a.mixed) {
    for (int k = tid; k < n3; k += nthreads) cst(&a.x32[k], (float)cld(&a.x[k]));
    ok = bar.sync();
    bool conv1 = false;
    double best1 = 0.0;
    if (ok)
      ok = stage(StageTag<true>{}, a.try_zero != 0, a.max_iters, a.tol * a.stop_factor, 0, it1, best1, err1, conv1);
    nEval1 = nEval;
    // hand over the stage's best iterate; every contact active again
    if (ok) {
      for (int k = tid; k < n3; k += nthreads) cst(&a.x[k], cld(&a.xbest[k]));
      for (int k = tid; k < n; k += nthreads) a.skip[k] = 0;
      ok = bar.sync();
    }
    if (ok)
      ok = stage(StageTag<false>{}, false, max(a.eval_every, a.max_iters - it1), a.tol, 1, it, errBest, errLast,
                 converged);
  } else {
    ok = stage(StageTag<false>{}, a.try_zero != 0, a.max_iters, a.tol, 0, it, errBest, errLast, converged);
  }
  // maxIters exhausted: x = x_best (GaussSeidel.impl.hpp:202). Mixed precision always
  // returns the FP64 stage's best iterate (its verification evaluation included), so
  // the reported residual is that of the returned r.
  if (ok && (!converged || a.mixed))
    for (int k = tid; k < n3; k += nthreads) cst(&a.x[k], cld(&a.xbest[k]));
  // count local failures over the grid
  {
    double part[1] = {(double)nfail};
    const bool ok2 = ok && bar.reduce<1>(part, 0u);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      a.out[0] = errBest;
      a.out[1] = (double)(it + it1);
      a.out[2] = converged ? 1.0 : 0.0;
      a.out[3] = ok2 ? bar.red[0] : -1.0;
      a.out[4] = (double)nEval;
      a.out[5] = (ok && ok2) ? 0.0 : 1.0;
      a.out[6] = (double)it1;
      a.out[7] = err1;
      a.out[8] = (double)(nEval - nEval1);
    }
  }
}

// Matrix-free (product) Gauss-Seidel: ProductGaussSeidel::solve with D = Minv
// and D H^T precomputed (src/Core/BlockSolvers/ProductGaussSeidel.impl.hpp:128-304,
// reached through PrimalFrictionProblem::solveWith(ProductGaussSeidelType&),
// PrimalFrictionProblem.cpp:40-59). W is never formed: the solver keeps the body
// vector y = Minv H^T x; contact i's local right-hand side is
// lb = b_i - A_i x_i + H_i y (A_i = W_ii + regularization), and after its update
// y_k += (Minv_k H_ik^T) dx_i for each of its bodies. Contacts of one color share
// no body, so their y updates do not overlap and the colors run as in gs_kernel
// (one lane per contact, a grid barrier per color, ndxRef merged per color). The
// evaluation every eval_every sweeps recomputes y = Minv (H^T x) from x (as the
// reference recomputes D H^T x there) with the matrix-free operator phases and
// evaluates u = H y + b.
__global__ void __launch_bounds__(256) pgs_kernel(GSArgs a, PGSArgs e) {
  GridTeam tm(a.gs, blockIdx.x);
  const int tid = tm.tid, nthreads = tm.nthreads;
  const int n = a.nc, n3 = 3 * n, nb = a.P.nb;
  const bool coulomb = a.coulomb != 0;
  const unsigned resMax = a.inf_norm ? 1u : 0u;
  const TileMat<double>& H = a.P.H;
  const size_t nnz = (size_t)H.nblocks;

  // y = Minv H^T x (kept in e.ybody); u = H y + b; err = eval(u, x) [err0 = eval(b, 0)]
  auto eval_phase = [&](bool withZero, double& err, double& err0) -> bool {
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, double xc[3]) { cld3(a.x, j, xc); };
    NoPre nopre;
    auto body = [&](int k, double v[6]) {
      double w[6];
      minv_apply<double>(a.P.minv, nb, k, v, w);
#pragma unroll
      for (int r = 0; r < 6; ++r) cst(&e.ybody[6 * k + r], w[r]);
    };
    auto lead = [&](int i, double v[3], NoPre::type) {
      double b[3], xi[3];
      ld3(a.P.b, i, b);
      cld3(a.x, i, xi);
      const double s = a.P.scal[i], mu = a.P.mu[i];
      const double y[3] = {v[0] + b[0], v[1] + b[1], v[2] + b[2]};
      const double xs[3] = {s * xi[0], s * xi[1], s * xi[2]};
      const double r = law_eval<double>(coulomb, mu, xs, y);
      part[0] = resMax ? fmax(part[0], r) : part[0] + r;
      if (withZero) {
        const double z[3] = {0.0, 0.0, 0.0};
        const double r0 = law_eval<double>(coulomb, mu, z, b);
        part[1] = resMax ? fmax(part[1], r0) : part[1] + r0;
      }
    };
    if (!apply_operator<true, double>(a.P, e.ybody, a.carry, tm, ld, nopre, lead, body)) return false;
    if (!tm.reduce<2>(part, resMax | (resMax << 1))) return false;
    err = tm.red[0];
    err0 = tm.red[1];
    if (!a.inf_norm) {
      err /= (1.0 + n);
      err0 /= (1.0 + n);
    }
    return true;
  };

  double errBest = 1.7976931348623157e308;
  double err = 0.0, err0 = 0.0;
  bool ok = eval_phase(a.try_zero != 0, err, err0);
  if (ok) {
    if (err < errBest) {
      errBest = err;
      for (int k = tid; k < n3; k += nthreads) cst(&a.xbest[k], cld(&a.x[k]));
    }
    // tryZero (GaussSeidelBase.impl.hpp:61-77): a better zero start also zeroes y
    if (a.try_zero && err0 < errBest) {
      errBest = err0;
      for (int k = tid; k < n3; k += nthreads) {
        cst(&a.xbest[k], 0.0);
        cst(&a.x[k], 0.0);
      }
      for (int k = tid; k < 6 * nb; k += nthreads) cst(&e.ybody[k], 0.0);
    }
    if (tm.leader() && a.hist_cap > 0) a.hist[0] = errBest;
    ok = tm.sync();
  }

  const double absSkipTol = fmin(a.skip_tol, a.tol);
  double ndxRef = 0.0;
  int it = 1, nfail = 0, nEval = 1;
  bool converged = false;
  for (; ok && it <= a.max_iters; ++it) {
    for (int col = 0; col < a.ncolors && ok; ++col) {
      const int c0 = a.colors[col], c1 = a.colors[col + 1];
      double part[1] = {0.0};
      // one lane per contact, the contacts spread over the waves (see kSpread)
      int gs = kSpread;
      while (gs > 1 && (c1 - c0) * gs > nthreads) gs >>= 1;
      for (int sg = tid; sg < (c1 - c0) * gs; sg += nthreads) {
        if (sg % gs != 0) continue;
        const int i = c0 + sg / gs;
        if (a.skip[i]) {
          a.skip[i] = (unsigned char)(a.skip[i] - 1);
          continue;
        }
        double lx[3], x0[3];
        cld3(a.x, i, lx);
        x0[0] = lx[0];
        x0[1] = lx[1];
        x0[2] = lx[2];
        double A[3][3];
#pragma unroll
        for (int k = 0; k < 9; ++k) A[k / 3][k % 3] = a.local[(size_t)9 * i + k];
        double lb[3];
        ld3(a.P.b, i, lb);
#pragma unroll
        for (int r = 0; r < 3; ++r) lb[r] -= A[r][0] * lx[0] + A[r][1] * lx[1] + A[r][2] * lx[2];
        // rowMultiply: lb += H_i y, keeping y_k for the update below (contact i is
        // the only one of its color at body k)
        const int q0 = e.hrow[2 * i], nq = e.hrow[2 * i + 1];
        double yk[2][6];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          if (t >= nq) break;
          const int q = q0 + t;
          const int k = H.bcol[q];
#pragma unroll
          for (int c = 0; c < 6; ++c) yk[t][c] = mld(&e.ybody[6 * k + c]);
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += H.vals[(size_t)(r * 6 + c) * nnz + q] * yk[t][c];
            lb[r] += acc;
          }
        }
        for (int t = 2; t < nq; ++t) {  // contacts on more than two bodies
          const int q = q0 + t;
          const int k = H.bcol[q];
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc += H.vals[(size_t)(r * 6 + c) * nnz + q] * mld(&e.ybody[6 * k + c]);
            lb[r] += acc;
          }
        }
        const double ldx0[3] = {-lx[0], -lx[1], -lx[2]};
        const double s = a.P.scal[i];
        const double lres = local::solve(coulomb, a.strategy, A, lb, lx, a.P.mu[i], a.local_tol, s);
        const bool lok = a.local_tol > lres;
        double ldx[3] = {ldx0[0] + lx[0], ldx0[1] + lx[1], ldx0[2] + lx[2]};
        if (!lok) {
          ldx[0] *= 0.5;
          ldx[1] *= 0.5;
          ldx[2] *= 0.5;
          ++nfail;
        }
        cst(&a.x[3 * i], x0[0] + ldx[0]);
        cst(&a.x[3 * i + 1], x0[1] + ldx[1]);
        cst(&a.x[3 * i + 2], x0[2] + ldx[2]);
        // colMultiply: y_k += (Minv_k H_ik^T) dx
        auto dmt_row = [&](int q, int r) {
          return e.dmt[(size_t)(r * 3 + 0) * nnz + q] * ldx[0] + e.dmt[(size_t)(r * 3 + 1) * nnz + q] * ldx[1] +
                 e.dmt[(size_t)(r * 3 + 2) * nnz + q] * ldx[2];
        };
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          if (t >= nq) break;
          const int q = q0 + t;
          const int k = H.bcol[q];
#pragma unroll
          for (int r = 0; r < 6; ++r) cst(&e.ybody[6 * k + r], yk[t][r] + dmt_row(q, r));
        }
        for (int t = 2; t < nq; ++t) {
          const int q = q0 + t;
          const int k = H.bcol[q];
#pragma unroll
          for (int r = 0; r < 6; ++r) cst(&e.ybody[6 * k + r], mld(&e.ybody[6 * k + r]) + dmt_row(q, r));
        }
        const double s2 = s * s;
        const double nx2 = s2 * (lx[0] * lx[0] + lx[1] * lx[1] + lx[2] * lx[2]);
        const double ndx2 = s2 * (ldx[0] * ldx[0] + ldx[1] * ldx[1] + ldx[2] * ldx[2]);
        part[0] = fmax(part[0], ndx2);
        if (fmin(nx2, ndx2) < absSkipTol || ndx2 < a.skip_tol * fmin(nx2, ndxRef))
          a.skip[i] = (unsigned char)a.skip_iters;
      }
      ok = tm.reduce<1>(part, 1u);
      if (ok) ndxRef = fmax(ndxRef, tm.red[0]);
    }
    if (!ok) break;
    if (it % a.eval_every == 0) {
      double e0;
      ok = eval_phase(false, err, e0);
      if (!ok) break;
      if (tm.leader() && nEval < a.hist_cap) a.hist[nEval] = err;
      ++nEval;
      if (err < errBest) {
        errBest = err;
        for (int k = tid; k < n3; k += nthreads) cst(&a.xbest[k], cld(&a.x[k]));
      }
      if (err < a.tol) {
        converged = true;
        break;
      }
      ndxRef /= a.eval_every;
      ok = tm.sync();
    }
  }
  // maxIters exhausted: x = x_best (ProductGaussSeidel.impl.hpp:301)
  if (ok && !converged)
    for (int k = tid; k < n3; k += nthreads) cst(&a.x[k], cld(&a.xbest[k]));
  {
    double part[1] = {(double)nfail};
    const bool ok2 = ok && tm.reduce<1>(part, 0u);
    if (tm.leader()) {
      a.out[0] = errBest;
      a.out[1] = (double)(it > a.max_iters ? a.max_iters : it);
      a.out[2] = converged ? 1.0 : 0.0;
      a.out[3] = ok2 ? tm.red[0] : -1.0;
      a.out[4] = (double)nEval;
      a.out[5] = (ok && ok2) ? 0.0 : 1.0;
    }
  }
}

}  // namespace sb
