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

struct Ctl {  // per-thread copy of the grid-uniform solver state
  int gw, nwaves, tid, nthreads;
};

__device__ __forceinline__ void load3(const double* p, int i, double v[3]) {
  v[0] = p[3 * i];
  v[1] = p[3 * i + 1];
  v[2] = p[3 * i + 2];
}
__device__ __forceinline__ void store3(double* p, int i, const double v[3]) {
  p[3 * i] = v[0];
  p[3 * i + 1] = v[1];
  p[3 * i + 2] = v[2];
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
// leader on complete contact rows. Contains 0 (assembled, no long rows),
// 1 (matrix-free) or 2 (matrix-free with hub bodies) internal barriers.
template <bool MF, typename Loader, typename Leader>
__device__ __forceinline__ bool apply_op(const PGArgs& a, const Ctl& c, Loader& ld, Leader& lead) {
  if (!MF) {
    for (int t = c.gw; t < a.P.W.ntiles; t += c.nwaves) tile_gather<3, 3, double, double>(a.P.W, t, ld, lead, a.carry);
    if (a.P.W.nlong > 0) {
      if (!grid_reduce_sync(a.gs, 0, 0)) return false;
      for (int k = c.gw; k < a.P.W.nlong; k += c.nwaves) {
        double v[3];
        long_row_sum<3, double, double>(a.P.W, k, a.carry, v);
        if (lane_id() == 0) lead(a.P.W.long_row[k], v);
      }
    }
    return true;
  }
  auto leadA = [&](int k, double v[6]) {
    double w[6];
    minv_apply(a.P.minv, a.nb, k, v, w);
#pragma unroll
    for (int r = 0; r < 6; ++r) a.wbody[6 * k + r] = w[r];
  };
  for (int t = c.gw; t < a.P.HT.ntiles; t += c.nwaves) tile_gather<6, 3, double, double>(a.P.HT, t, ld, leadA, a.carry);
  if (!grid_reduce_sync(a.gs, 0, 0)) return false;
  if (a.P.HT.nlong > 0) {
    for (int k = c.gw; k < a.P.HT.nlong; k += c.nwaves) {
      double v[6];
      long_row_sum<6, double, double>(a.P.HT, k, a.carry, v);
      if (lane_id() == 0) leadA(a.P.HT.long_row[k], v);
    }
    if (!grid_reduce_sync(a.gs, 0, 0)) return false;
  }
  auto ldB = [&](int k, double xc[6]) {
#pragma unroll
    for (int r = 0; r < 6; ++r) xc[r] = a.wbody[6 * k + r];
  };
  for (int t = c.gw; t < a.P.H.ntiles; t += c.nwaves) tile_gather<3, 6, double, double>(a.P.H, t, ldB, lead, a.carry);
  return true;
}

template <int NV>
__device__ __forceinline__ bool finish_phase(const PGArgs& a, double (&part)[NV], unsigned maxmask) {
  block_partials<NV>(a.gs, part, maxmask);
  return grid_reduce_sync(a.gs, NV, maxmask);
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
  for (int k = c.tid; k < n; k += c.nthreads) dst[k] = src[k];
}

__device__ __forceinline__ double contact_scaling(const PGArgs& a, int i) { return a.use_scaling ? a.P.scal[i] : 1.0; }

// Residual contribution, COV, gradient and energy terms at (x_i, Mx_i).
__device__ __forceinline__ void eval_contact(const PGArgs& a, int i, const double x[3], const double Mx[3],
                                             double& res, double g[3], double& J) {
  const double mu = a.P.mu[i];
  double b[3];
  load3(a.P.b, i, b);
  const double y[3] = {Mx[0] + b[0], Mx[1] + b[1], Mx[2] + b[2]};
  const double s = contact_scaling(a, i);
  const double xsc[3] = {s * x[0], s * x[1], s * x[2]};
  res = law_eval<double>(a.coulomb != 0, mu, xsc, y);
  const double s0 = cov_normal<double>(a.coulomb != 0, mu, y[1], y[2]);
  g[0] = y[0] + s0;
  g[1] = y[1];
  g[2] = y[2];
  J = x[0] * (0.5 * Mx[0] + b[0] + s0) + x[1] * (0.5 * Mx[1] + b[1]) + x[2] * (0.5 * Mx[2] + b[2]);
}

}  // namespace

template <bool MF>
__global__ void __launch_bounds__(256) pg_kernel(PGArgs a) {
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

  // ---- initial phases: x <- P(x0); alpha from guess_alpha; first residual.
  {
    // M (P(x0) - 1)
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, double xc[3]) {
      double v[3];
      load3(a.x_io, j, v);
      project<double>(a.P.mu[j], v[0], v[1], v[2]);
      xc[0] = v[0] - 1.0;
      xc[1] = v[1] - 1.0;
      xc[2] = v[2] - 1.0;
    };
    auto lead = [&](int i, double v[3]) {
      double xi[3];
      load3(a.x_io, i, xi);
      project<double>(a.P.mu[i], xi[0], xi[1], xi[2]);
      part[0] += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      part[1] += (xi[0] - 1.0) * (xi[0] - 1.0) + (xi[1] - 1.0) * (xi[1] - 1.0) + (xi[2] - 1.0) * (xi[2] - 1.0);
    };
    ok = apply_op<MF>(a, c, ld, lead) && finish_phase<2>(a, part, 0u);
    ++spmvs;
    if (ok) {
      const double nMx = read_reduced(a.gs, 0), nd = read_reduced(a.gs, 1);
      if (nMx > 1.e-16) alpha = fmin(1.0, nd / nMx);
    }
  }
  if (ok) {
    // M P(x0), written as x with its gradient and first residual
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, double xc[3]) {
      load3(a.x_io, j, xc);
      project<double>(a.P.mu[j], xc[0], xc[1], xc[2]);
    };
    auto lead = [&](int i, double v[3]) {
      double xi[3];
      load3(a.x_io, i, xi);
      project<double>(a.P.mu[i], xi[0], xi[1], xi[2]);
      store3(a.X[0], i, xi);
      store3(a.MX[0], i, v);
      double r, g[3], Ji;
      eval_contact(a, i, xi, v, r, g, Ji);
      store3(a.G[0], i, g);
      part[0] = resMax ? fmax(part[0], r) : part[0] + r;
      part[1] += Ji;
    };
    ok = apply_op<MF>(a, c, ld, lead) && finish_phase<2>(a, part, resMax);
    ++spmvs;
    if (ok) {
      res = read_reduced(a.gs, 0);
      if (!a.inf_norm) res /= (1.0 + a.nc);
      J = read_reduced(a.gs, 1);
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
    if (!spg) {
      // ---- APGD / Descent: backtracking line search (ProjectedGradient.impl.hpp:117-133)
      alpha *= a.ls_opt;
      int rxs = 0;
      while (rxs == rx || rxs == rprev) ++rxs;
      double Js = 0, decr = 0, nd = 0, decr2 = 0, resN = 0, JN = 0;
      bool accepted = false;
      for (int ls = 0; ls < a.ls_iters && ok; ++ls) {
        double part[6] = {0, 0, 0, 0, 0, 0};
        const double al = alpha;
        const double* Xx = a.X[rx];
        const double* Gx = a.G[rg];
        auto ld = [&](int j, double xc[3]) {
          double xv[3], gv[3];
          load3(Xx, j, xv);
          load3(Gx, j, gv);
          xc[0] = xv[0] - al * gv[0];
          xc[1] = xv[1] - al * gv[1];
          xc[2] = xv[2] - al * gv[2];
          project<double>(a.P.mu[j], xc[0], xc[1], xc[2]);
        };
        auto lead = [&](int i, double v[3]) {
          double xv[3], gv[3], xs[3], pv[3], Mx[3], b[3];
          load3(Xx, i, xv);
          load3(Gx, i, gv);
          ld(i, xs);
          store3(a.X[rxs], i, xs);
          store3(a.MX[rxs], i, v);
          load3(a.X[rprev], i, pv);
          load3(a.MX[rx], i, Mx);
          load3(a.P.b, i, b);
          const double mu = a.P.mu[i];
          const double s0 = cov_normal<double>(coulomb, mu, Mx[1] + b[1], Mx[2] + b[2]);
          part[0] += xs[0] * (0.5 * v[0] + b[0] + s0) + xs[1] * (0.5 * v[1] + b[1]) + xs[2] * (0.5 * v[2] + b[2]);
          const double d[3] = {xs[0] - xv[0], xs[1] - xv[1], xs[2] - xv[2]};
          part[1] += gv[0] * d[0] + gv[1] * d[1] + gv[2] * d[2];
          part[2] += d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          part[3] += (xs[0] - pv[0]) * gv[0] + (xs[1] - pv[1]) * gv[1] + (xs[2] - pv[2]) * gv[2];
          double r, gn[3], Jn;
          eval_contact(a, i, xs, v, r, gn, Jn);
          store3(a.G[1 - rg], i, gn);
          part[4] = resMax ? fmax(part[4], r) : part[4] + r;
          part[5] += Jn;
        };
        ok = apply_op<MF>(a, c, ld, lead) && finish_phase<6>(a, part, resMax << 4);
        ++spmvs;
        ++lsTrials;
        if (!ok) break;
        Js = read_reduced(a.gs, 0);
        decr = read_reduced(a.gs, 1);
        nd = read_reduced(a.gs, 2);
        decr2 = read_reduced(a.gs, 3);
        resN = read_reduced(a.gs, 4);
        JN = read_reduced(a.gs, 5);
        if (Js < J + decr + 0.5 / alpha * nd) {
          accepted = true;
          break;
        }
        alpha *= a.ls_pess;
      }
      (void)accepted;
      if (!ok) break;
      // ---- Nesterov extrapolation or restart (:135-149)
      if (useNesterov && decr2 < 0) {
        const double beta = nesterov(theta, 0.0);
        int rout = 0;
        while (rout == rxs || rout == rprev) ++rout;
        double part[2] = {0.0, 0.0};
        for (int i = c.tid; i < a.nc; i += c.nthreads) {
          double xs[3], pv[3], Ms[3], Mp[3], xn[3], Mn[3];
          load3(a.X[rxs], i, xs);
          load3(a.X[rprev], i, pv);
          load3(a.MX[rxs], i, Ms);
          load3(a.MX[rprev], i, Mp);
          for (int k = 0; k < 3; ++k) {
            xn[k] = xs[k] + beta * (xs[k] - pv[k]);
            Mn[k] = Ms[k] + beta * (Ms[k] - Mp[k]);
          }
          store3(a.X[rout], i, xn);
          store3(a.MX[rout], i, Mn);
          double r, g[3], Ji;
          eval_contact(a, i, xn, Mn, r, g, Ji);
          store3(a.G[rg], i, g);
          part[0] = resMax ? fmax(part[0], r) : part[0] + r;
          part[1] += Ji;
        }
        ok = finish_phase<2>(a, part, resMax);
        if (!ok) break;
        res = read_reduced(a.gs, 0);
        J = read_reduced(a.gs, 1);
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
      int rxs = 0;
      while (rxs == rx || rxs == rprev) ++rxs;
      double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const double al = alpha;
      const double* Xx = a.X[rx];
      const double* Gx = a.G[rg];
      auto ld = [&](int j, double xc[3]) {
        double xv[3], gv[3];
        load3(Xx, j, xv);
        load3(Gx, j, gv);
        xc[0] = xv[0] - al * gv[0];
        xc[1] = xv[1] - al * gv[1];
        xc[2] = xv[2] - al * gv[2];
        project<double>(a.P.mu[j], xc[0], xc[1], xc[2]);
      };
      auto lead = [&](int i, double v[3]) {
        double xv[3], gv[3], xs[3], pv[3], Mx[3], Mp[3];
        load3(Xx, i, xv);
        load3(Gx, i, gv);
        ld(i, xs);
        store3(a.X[rxs], i, xs);
        store3(a.MX[rxs], i, v);
        load3(a.X[rprev], i, pv);
        load3(a.MX[rx], i, Mx);
        load3(a.MX[rprev], i, Mp);
        double ad = 0, ae = 0, cd = 0, ce = 0, dd = 0, de = 0, ee = 0, d2 = 0;
        for (int k = 0; k < 3; ++k) {
          const double av = xs[k] - xv[k], cv = xs[k] - pv[k];
          const double dv = v[k] - Mx[k], ev = v[k] - Mp[k];
          d2 += cv * gv[k];
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
        double r, gn[3], Jn;
        eval_contact(a, i, xs, v, r, gn, Jn);
        store3(a.G[1 - rg], i, gn);
        part[7] = resMax ? fmax(part[7], r) : part[7] + r;
      };
      ok = apply_op<MF>(a, c, ld, lead) && finish_phase<8>(a, part, resMax << 7);
      ++spmvs;
      if (!ok) break;
      const double decr2 = read_reduced(a.gs, 0);
      double wz, zz;
      if (decr2 < 0.0) {
        const double beta = nesterov(theta, 0.0);
        wz = read_reduced(a.gs, 1) + beta * read_reduced(a.gs, 2) + beta * beta * read_reduced(a.gs, 3);
        zz = read_reduced(a.gs, 4) + 2.0 * beta * read_reduced(a.gs, 5) + beta * beta * read_reduced(a.gs, 6);
        int rout = 0;
        while (rout == rxs || rout == rprev) ++rout;
        double p2[1] = {0.0};
        for (int i = c.tid; i < a.nc; i += c.nthreads) {
          double xs[3], pv[3], Ms[3], Mp[3], xn[3], Mn[3];
          load3(a.X[rxs], i, xs);
          load3(a.X[rprev], i, pv);
          load3(a.MX[rxs], i, Ms);
          load3(a.MX[rprev], i, Mp);
          for (int k = 0; k < 3; ++k) {
            xn[k] = xs[k] + beta * (xs[k] - pv[k]);
            Mn[k] = Ms[k] + beta * (Ms[k] - Mp[k]);
          }
          store3(a.X[rout], i, xn);
          store3(a.MX[rout], i, Mn);
          double r, g[3], Ji;
          eval_contact(a, i, xn, Mn, r, g, Ji);
          store3(a.G[rg], i, g);
          p2[0] = resMax ? fmax(p2[0], r) : p2[0] + r;
        }
        ok = finish_phase<1>(a, p2, resMax);
        if (!ok) break;
        res = read_reduced(a.gs, 0);
        rx = rout;
        rprev = rxs;
        ++nest;
      } else {
        wz = read_reduced(a.gs, 1);
        zz = read_reduced(a.gs, 4);
        rx = rxs;
        rprev = rxs;
        rg = 1 - rg;
        res = read_reduced(a.gs, 7);
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
  }
}

template __global__ void pg_kernel<false>(PGArgs);
template __global__ void pg_kernel<true>(PGArgs);

}  // namespace sb
