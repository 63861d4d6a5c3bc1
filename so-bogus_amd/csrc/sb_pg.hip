// Persistent projected-gradient solver (APGD / Descent / SPG) for one dual
// friction problem, fused with the per-contact cone projection, De Saxce
// change of variable and Fischer-Burmeister residual.
//
// Algorithm: PgMethod<APGD|Descent> and PgMethod<SPG>
// (src/Core/BlockSolvers/ProjectedGradient.impl.hpp:74-155, 288-351), with
// test_residual / guess_alpha / nesterov_inertia (:21-52) and the residual of
// ConstrainedSolverBase::eval (ConstrainedSolverBase.impl.hpp:21-74).
//
// Every block of a co-resident grid runs the same control flow. Each phase is
// one pass over the data (a block-SpMV gather with the per-contact work fused
// into the row leaders, or a per-contact update) followed by a grid barrier
// that also reduces the phase's dot products deterministically; the decisions
// (line-search acceptance, restart, stopping) are then taken redundantly by
// every block from the same reduced values. The line-search candidate
// xs = P(x - alpha g) is formed on the fly inside the gather.
//
// Differences from the reference's floating-point sequence (same mathematics):
//  * after a Nesterov step the product M x is formed as
//    M xs + beta (M xs - M prev) from the two products already computed
//    (M is linear), instead of a new SpMV;
//  * dot products are summed in a fixed parallel order.
#include <hip/hip_runtime.h>

#include "sb_law.hpp"
#include "sb_problem.hpp"

namespace sb {

namespace {

struct Ctl {  // grid geometry of this thread
  int gw, nwaves, tid, nthreads;
};

// Read-only inputs (b, x_io): plain loads.
__device__ __forceinline__ void ld3(const double* p, int i, double v[3]) {
  v[0] = p[3 * i];
  v[1] = p[3 * i + 1];
  v[2] = p[3 * i + 2];
}
// Vectors written by other blocks in earlier phases.
__device__ __forceinline__ void mld3(const double* p, int i, double v[3]) {
  v[0] = mld(p + 3 * i);
  v[1] = mld(p + 3 * i + 1);
  v[2] = mld(p + 3 * i + 2);
}
// Coherent stores of vectors that other blocks read in later phases.
__device__ __forceinline__ void cst3(double* p, int i, const double v[3]) {
  cst(p + 3 * i, v[0]);
  cst(p + 3 * i + 1, v[1]);
  cst(p + 3 * i + 2, v[2]);
}

// Minv_k v for the element-major [36][nb] storage.
__device__ __forceinline__ void minv_apply(const double* minv, int nb, int k, const double v[6], double w[6]) {
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    double acc = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc += minv[(size_t)(r * 6 + c) * nb + k] * v[c];
    w[r] = acc;
  }
}

// One operator application M x with x given by the loader, finished by the
// leader on complete contact rows (with the row data fetched by pre). Contains
// 0 (assembled, no long rows), 1 (matrix-free) or 2 (matrix-free with hub
// bodies) internal barriers.
template <bool MF, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ bool apply_op(const PGArgs& a, const Ctl& c, Barrier& bar, Loader& ld, Pre& pre,
                                         Leader& lead) {
  if (!MF) {
    for (int t = c.gw; t < a.P.W.ntiles; t += c.nwaves)
      tile_gather<3, 3, double, double>(a.P.W, t, ld, pre, lead, a.carry);
    if (a.P.W.nlong > 0) {
      if (!bar.sync()) return false;
      for (int k = c.gw; k < a.P.W.nlong; k += c.nwaves) {
        double v[3];
        long_row_sum<3, double, double>(a.P.W, k, a.carry, v);
        if (lane_id() == 0) {
          const int row = a.P.W.long_row[k];
          lead(row, v, pre(row));
        }
      }
    }
    return true;
  }
  NoPre nopre;
  auto leadA = [&](int k, double v[6], NoPre::type) {
    double w[6];
    minv_apply(a.P.minv, a.nb, k, v, w);
#pragma unroll
    for (int r = 0; r < 6; ++r) cst(&a.wbody[6 * k + r], w[r]);
  };
  for (int t = c.gw; t < a.P.HT.ntiles; t += c.nwaves)
    tile_gather<6, 3, double, double>(a.P.HT, t, ld, nopre, leadA, a.carry);
  if (!bar.sync()) return false;
  if (a.P.HT.nlong > 0) {
    for (int k = c.gw; k < a.P.HT.nlong; k += c.nwaves) {
      double v[6];
      long_row_sum<6, double, double>(a.P.HT, k, a.carry, v);
      if (lane_id() == 0) leadA(a.P.HT.long_row[k], v, NoPre::type{});
    }
    if (!bar.sync()) return false;
  }
  auto ldB = [&](int k, double xc[6]) {
#pragma unroll
    for (int r = 0; r < 6; ++r) xc[r] = mld(&a.wbody[6 * k + r]);
  };
  for (int t = c.gw; t < a.P.H.ntiles; t += c.nwaves)
    tile_gather<3, 6, double, double>(a.P.H, t, ldB, pre, lead, a.carry);
  return true;
}

__device__ __forceinline__ double nesterov(double& theta, double q) {
  const double tp = theta;
  const double tp2 = tp * tp;
  const double bb = tp2 - q;
  const double delta = sqrt(bb * bb + 4.0 * tp2);
  theta = 0.5 * (delta - bb);
  return (tp - tp2) / (tp2 + theta);
}

__device__ __forceinline__ void grid_copy(const Ctl& c, double* dst, const double* src, int n) {
  for (int k = c.tid; k < n; k += c.nthreads) cst(dst + k, cld(src + k));
}

// Residual contribution, COV, gradient and energy terms at (x_i, Mx_i).
__device__ __forceinline__ void eval_contact(bool coulomb, double mu, double s, const double b[3], const double x[3],
                                             const double Mx[3], double& res, double g[3], double& J) {
  const double y[3] = {Mx[0] + b[0], Mx[1] + b[1], Mx[2] + b[2]};
  const double xsc[3] = {s * x[0], s * x[1], s * x[2]};
  res = law_eval<double>(coulomb, mu, xsc, y);
  const double s0 = cov_normal<double>(coulomb, mu, y[1], y[2]);
  g[0] = y[0] + s0;
  g[1] = y[1];
  g[2] = y[2];
  J = x[0] * (0.5 * Mx[0] + b[0] + s0) + x[1] * (0.5 * Mx[1] + b[1]) + x[2] * (0.5 * Mx[2] + b[2]);
}

__device__ __forceinline__ void proj_step(double mu, double al, const double x[3], const double g[3], double xs[3]) {
  xs[0] = x[0] - al * g[0];
  xs[1] = x[1] - al * g[1];
  xs[2] = x[2] - al * g[2];
  project<double>(mu, xs[0], xs[1], xs[2]);
}

struct InitRow {
  double x[3], b[3], mu, s;
};
struct TrialRow {  // own-contact data of a line-search trial row
  double x[3], g[3], p[3], m[3], b[3], mu, s;
};
struct SpgRow {
  double x[3], g[3], p[3], m[3], mp[3], b[3], mu, s;
};

}  // namespace

template <bool MF>
__global__ void __launch_bounds__(512) pg_kernel(PGArgs a) {
  Ctl c;
  const int wpb = blockDim.x / SB_WAVE;
  c.gw = blockIdx.x * wpb + wave_in_block();
  c.nwaves = gridDim.x * wpb;
  c.tid = blockIdx.x * blockDim.x + threadIdx.x;
  c.nthreads = gridDim.x * blockDim.x;
  const int n3 = 3 * a.nc;
  const bool coulomb = a.coulomb != 0;
  const unsigned resMax = a.inf_norm ? 1u : 0u;  // bit of the residual partial when it is a max
  const bool spg = (a.variant == 4);
  const bool useNesterov = (a.variant == 3 || spg);
  int it = 0, spmvs = 0, lsTrials = 0, nest = 0;
  bool ok = true;
  double minRes = -1.0, res = 0.0, J = 0.0, alpha = 1.0, theta = 1.0;
  int rx = 0, rprev = 0, rg = 0;
  Barrier bar;
  bar.gs = a.gs;
  // phase-time breakdown seen by block 0 (s_memrealtime ticks): [0] operator-phase work,
  // [1] operator-phase reduction + barrier wait, [2] update-phase work, [3] update-phase wait
  uint64_t tsum[4] = {0, 0, 0, 0};
  uint64_t tp0 = 0, tp1 = 0;
  auto rowScale = [&](int i) { return a.use_scaling ? a.P.scal[i] : 1.0; };

  auto initPre = [&](int i) {
    InitRow r;
    ld3(a.x_io, i, r.x);
    r.mu = a.P.mu[i];
    project<double>(r.mu, r.x[0], r.x[1], r.x[2]);
    ld3(a.P.b, i, r.b);
    r.s = rowScale(i);
    return r;
  };
  // ---- initial phases: x <- P(x0); alpha from guess_alpha; first residual.
  {
    // M (P(x0) - 1)
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, double xc[3]) {
      double v[3];
      ld3(a.x_io, j, v);
      project<double>(a.P.mu[j], v[0], v[1], v[2]);
      xc[0] = v[0] - 1.0;
      xc[1] = v[1] - 1.0;
      xc[2] = v[2] - 1.0;
    };
    auto lead = [&](int i, double v[3], const InitRow& r) {
      part[0] += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      part[1] += (r.x[0] - 1.0) * (r.x[0] - 1.0) + (r.x[1] - 1.0) * (r.x[1] - 1.0) + (r.x[2] - 1.0) * (r.x[2] - 1.0);
    };
    ok = apply_op<MF>(a, c, bar, ld, initPre, lead) && bar.reduce<2>(part, 0u);
    ++spmvs;
    if (ok) {
      const double nMx = bar.red[0], nd = bar.red[1];
      if (nMx > 1.e-16) alpha = fmin(1.0, nd / nMx);
    }
  }
  if (ok) {
    // M P(x0), written as x with its gradient and first residual
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, double xc[3]) {
      ld3(a.x_io, j, xc);
      project<double>(a.P.mu[j], xc[0], xc[1], xc[2]);
    };
    auto lead = [&](int i, double v[3], const InitRow& r) {
      cst3(a.X[0], i, r.x);
      cst3(a.MX[0], i, v);
      double rr, g[3], Ji;
      eval_contact(coulomb, r.mu, r.s, r.b, r.x, v, rr, g, Ji);
      cst3(a.G[0], i, g);
      part[0] = resMax ? fmax(part[0], rr) : part[0] + rr;
      part[1] += Ji;
    };
    ok = apply_op<MF>(a, c, bar, ld, initPre, lead) && bar.reduce<2>(part, resMax);
    ++spmvs;
    if (ok) {
      res = bar.red[0];
      if (!a.inf_norm) res /= (1.0 + a.nc);
      J = bar.red[1];
    }
  }

  bool converged = false;
  while (ok) {
    // ---- test_residual at the current x (iteration `it`)
    if (blockIdx.x == 0 && threadIdx.x == 0 && it < a.hist_cap) a.hist[it] = res;
    if (it == 0 || res < minRes) {
      minRes = res;
      grid_copy(c, a.xbest, a.X[rx], n3);
    }
    if (res < a.tol) {
      converged = true;
      break;
    }
    int rxs = 0;
    while (rxs == rx || rxs == rprev) ++rxs;
    const double* Xx = a.X[rx];
    const double* Gx = a.G[rg];
    if (!spg) {
      // ---- APGD / Descent: backtracking line search (ProjectedGradient.impl.hpp:117-133)
      alpha *= a.ls_opt;
      double Js = 0, decr = 0, nd = 0, decr2 = 0, resN = 0, JN = 0;
      auto pre = [&](int i) {
        TrialRow r;
        mld3(Xx, i, r.x);
        mld3(Gx, i, r.g);
        mld3(a.X[rprev], i, r.p);
        mld3(a.MX[rx], i, r.m);
        ld3(a.P.b, i, r.b);
        r.mu = a.P.mu[i];
        r.s = rowScale(i);
        return r;
      };
      for (int ls = 0; ls < a.ls_iters && ok; ++ls) {
        double part[6] = {0, 0, 0, 0, 0, 0};
        const double al = alpha;
        auto ld = [&](int j, double xc[3]) {
          double xv[3], gv[3];
          mld3(Xx, j, xv);
          mld3(Gx, j, gv);
          proj_step(a.P.mu[j], al, xv, gv, xc);
        };
        auto lead = [&](int i, double v[3], const TrialRow& r) {
          double xs[3];
          proj_step(r.mu, al, r.x, r.g, xs);
          cst3(a.X[rxs], i, xs);
          cst3(a.MX[rxs], i, v);
          const double s0 = cov_normal<double>(coulomb, r.mu, r.m[1] + r.b[1], r.m[2] + r.b[2]);
          part[0] += xs[0] * (0.5 * v[0] + r.b[0] + s0) + xs[1] * (0.5 * v[1] + r.b[1]) +
                     xs[2] * (0.5 * v[2] + r.b[2]);
          const double d[3] = {xs[0] - r.x[0], xs[1] - r.x[1], xs[2] - r.x[2]};
          part[1] += r.g[0] * d[0] + r.g[1] * d[1] + r.g[2] * d[2];
          part[2] += d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          part[3] += (xs[0] - r.p[0]) * r.g[0] + (xs[1] - r.p[1]) * r.g[1] + (xs[2] - r.p[2]) * r.g[2];
          double rr, gn[3], Jn;
          eval_contact(coulomb, r.mu, r.s, r.b, xs, v, rr, gn, Jn);
          cst3(a.G[1 - rg], i, gn);
          part[4] = resMax ? fmax(part[4], rr) : part[4] + rr;
          part[5] += Jn;
        };
        tp0 = realtime();
        ok = apply_op<MF>(a, c, bar, ld, pre, lead);
        tp1 = realtime();
        ok = ok && bar.reduce<6>(part, resMax << 4);
        tsum[0] += tp1 - tp0;
        tsum[1] += realtime() - tp1;
        ++spmvs;
        ++lsTrials;
        if (!ok) break;
        Js = bar.red[0];
        decr = bar.red[1];
        nd = bar.red[2];
        decr2 = bar.red[3];
        resN = bar.red[4];
        JN = bar.red[5];
        if (Js < J + decr + 0.5 / alpha * nd) break;
        alpha *= a.ls_pess;
      }
      if (!ok) break;
      // ---- Nesterov extrapolation or restart (:135-149)
      if (useNesterov && decr2 < 0) {
        const double beta = nesterov(theta, 0.0);
        int rout = 0;
        while (rout == rxs || rout == rprev) ++rout;
        tp0 = realtime();
        double part[2] = {0.0, 0.0};
        for (int i = c.tid; i < a.nc; i += c.nthreads) {
          double xs[3], pv[3], Ms[3], Mp[3], xn[3], Mn[3], b[3];
          mld3(a.X[rxs], i, xs);
          mld3(a.X[rprev], i, pv);
          mld3(a.MX[rxs], i, Ms);
          mld3(a.MX[rprev], i, Mp);
          ld3(a.P.b, i, b);
          for (int k = 0; k < 3; ++k) {
            xn[k] = xs[k] + beta * (xs[k] - pv[k]);
            Mn[k] = Ms[k] + beta * (Ms[k] - Mp[k]);
          }
          cst3(a.X[rout], i, xn);
          cst3(a.MX[rout], i, Mn);
          double rr, g[3], Ji;
          eval_contact(coulomb, a.P.mu[i], rowScale(i), b, xn, Mn, rr, g, Ji);
          cst3(a.G[rg], i, g);
          part[0] = resMax ? fmax(part[0], rr) : part[0] + rr;
          part[1] += Ji;
        }
        tp1 = realtime();
        ok = bar.reduce<2>(part, resMax);
        tsum[2] += tp1 - tp0;
        tsum[3] += realtime() - tp1;
        if (!ok) break;
        res = bar.red[0];
        J = bar.red[1];
        rx = rout;
        rprev = rxs;
        ++nest;
      } else {
        rx = rxs;
        rprev = rxs;
        rg = 1 - rg;
        res = resN;
        J = JN;
        theta = 1.0;
      }
      if (!a.inf_norm) res /= (1.0 + a.nc);
    } else {
      // ---- SPG (ProjectedGradient.impl.hpp:316-346): one product per iteration.
      double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const double al = alpha;
      auto pre = [&](int i) {
        SpgRow r;
        mld3(Xx, i, r.x);
        mld3(Gx, i, r.g);
        mld3(a.X[rprev], i, r.p);
        mld3(a.MX[rx], i, r.m);
        mld3(a.MX[rprev], i, r.mp);
        ld3(a.P.b, i, r.b);
        r.mu = a.P.mu[i];
        r.s = rowScale(i);
        return r;
      };
      auto ld = [&](int j, double xc[3]) {
        double xv[3], gv[3];
        mld3(Xx, j, xv);
        mld3(Gx, j, gv);
        proj_step(a.P.mu[j], al, xv, gv, xc);
      };
      auto lead = [&](int i, double v[3], const SpgRow& r) {
        double xs[3];
        proj_step(r.mu, al, r.x, r.g, xs);
        cst3(a.X[rxs], i, xs);
        cst3(a.MX[rxs], i, v);
        double ad = 0, ae = 0, cd = 0, ce = 0, dd = 0, de = 0, ee = 0, d2 = 0;
        for (int k = 0; k < 3; ++k) {
          const double av = xs[k] - r.x[k], cv = xs[k] - r.p[k];
          const double dv = v[k] - r.m[k], ev = v[k] - r.mp[k];
          d2 += cv * r.g[k];
          ad += av * dv;
          ae += av * ev;
          cd += cv * dv;
          ce += cv * ev;
          dd += dv * dv;
          de += dv * ev;
          ee += ev * ev;
        }
        part[0] += d2;
        part[1] += ad;
        part[2] += ae + cd;
        part[3] += ce;
        part[4] += dd;
        part[5] += de;
        part[6] += ee;
        double rr, gn[3], Jn;
        eval_contact(coulomb, r.mu, r.s, r.b, xs, v, rr, gn, Jn);
        cst3(a.G[1 - rg], i, gn);
        part[7] = resMax ? fmax(part[7], rr) : part[7] + rr;
      };
      tp0 = realtime();
      ok = apply_op<MF>(a, c, bar, ld, pre, lead);
      tp1 = realtime();
      ok = ok && bar.reduce<8>(part, resMax << 7);
      tsum[0] += tp1 - tp0;
      tsum[1] += realtime() - tp1;
      ++spmvs;
      if (!ok) break;
      const double decr2 = bar.red[0];
      double wz, zz;
      if (decr2 < 0.0) {
        const double beta = nesterov(theta, 0.0);
        wz = bar.red[1] + beta * bar.red[2] + beta * beta * bar.red[3];
        zz = bar.red[4] + 2.0 * beta * bar.red[5] + beta * beta * bar.red[6];
        int rout = 0;
        while (rout == rxs || rout == rprev) ++rout;
        tp0 = realtime();
        double p2[1] = {0.0};
        for (int i = c.tid; i < a.nc; i += c.nthreads) {
          double xs[3], pv[3], Ms[3], Mp[3], xn[3], Mn[3], b[3];
          mld3(a.X[rxs], i, xs);
          mld3(a.X[rprev], i, pv);
          mld3(a.MX[rxs], i, Ms);
          mld3(a.MX[rprev], i, Mp);
          ld3(a.P.b, i, b);
          for (int k = 0; k < 3; ++k) {
            xn[k] = xs[k] + beta * (xs[k] - pv[k]);
            Mn[k] = Ms[k] + beta * (Ms[k] - Mp[k]);
          }
          cst3(a.X[rout], i, xn);
          cst3(a.MX[rout], i, Mn);
          double rr, g[3], Ji;
          eval_contact(coulomb, a.P.mu[i], rowScale(i), b, xn, Mn, rr, g, Ji);
          cst3(a.G[rg], i, g);
          p2[0] = resMax ? fmax(p2[0], rr) : p2[0] + rr;
        }
        tp1 = realtime();
        ok = bar.reduce<1>(p2, resMax);
        tsum[2] += tp1 - tp0;
        tsum[3] += realtime() - tp1;
        if (!ok) break;
        res = bar.red[0];
        rx = rout;
        rprev = rxs;
        ++nest;
      } else {
        wz = bar.red[1];
        zz = bar.red[4];
        rx = rxs;
        rprev = rxs;
        rg = 1 - rg;
        res = bar.red[7];
        theta = 1.0;
      }
      if (!a.inf_norm) res /= (1.0 + a.nc);
      alpha = wz / zz;
      alpha = fmin(1.e6, fmax(1.e-6, alpha));
    }
    ++it;
    if (it >= a.max_iters) break;
  }
  // x = x_best (ProjectedGradient.impl.hpp:152)
  if (ok) grid_copy(c, a.x_io, a.xbest, n3);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    a.out[0] = minRes;
    a.out[1] = (double)it;
    a.out[2] = (double)spmvs;
    a.out[3] = (double)lsTrials;
    a.out[4] = (double)nest;
    a.out[5] = converged ? 1.0 : 0.0;
    a.out[6] = ok ? 0.0 : 1.0;
    for (int k = 0; k < 4; ++k) a.out[7 + k] = 0.01 * (double)tsum[k];  // microseconds (100 MHz clock)
  }
}

template __global__ void pg_kernel<false>(PGArgs);
template __global__ void pg_kernel<true>(PGArgs);

}  // namespace sb
