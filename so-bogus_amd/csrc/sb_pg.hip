// Projected-gradient solver (APGD / Descent / SPG) for dual friction problems,
// fused with the per-contact cone projection, De Saxce change of variable and
// Fischer-Burmeister residual.
//
// Algorithm: PgMethod<APGD|Descent> and PgMethod<SPG>
// (src/Core/BlockSolvers/ProjectedGradient.impl.hpp:79-155, 291-351), with
// test_residual / guess_alpha / nesterov_inertia (:22-54) and the residual of
// ConstrainedSolverBase::eval (ConstrainedSolverBase.impl.hpp:23-74).
//
// The solver body is written once for a team (the threads solving one problem:
// a whole persistent grid, or one workgroup of the batched kernel) and a storage
// type T (double, or float for the mixed-precision bulk). Each phase is one pass
// over the data (a block-SpMV gather with the per-contact work fused into the row
// leaders, or a per-contact update) ended by a team reduction of the phase's dot
// products; the decisions (line-search acceptance, restart, stopping) are then
// taken redundantly by every member from the same reduced values. The
// line-search candidate xs = P(x - alpha g) is formed on the fly inside the
// gather.
//
// Precision: with T = float the matrices, the iterates, the products and the
// projection are FP32; b, mu, the row scalings, the Fischer-Burmeister residual,
// the energy J and every dot product and decision are FP64 (the FP32 values are
// widened exactly before use).
//
// Differences from the reference's floating-point sequence (same mathematics):
//  * after a Nesterov step the product M x is formed as
//    M xs + beta (M xs - M prev) from the two products already computed
//    (M is linear), instead of a new SpMV;
//  * dot products are summed in a fixed parallel order;
//  * the first residual is evaluated before guess_alpha (the reference evaluates
//    it at iteration 0 just after); a warm start that already meets the
//    tolerance therefore costs one operator application instead of two.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "sb_law.hpp"
#include "sb_op.hpp"
#include "sb_problem.hpp"
#include "sb_team.hpp"

namespace sb {

namespace {

template <typename T>
__device__ __forceinline__ void dld3(const double* p, int i, double v[3]) {
  v[0] = gld(p + 3 * i);
  v[1] = gld(p + 3 * i + 1);
  v[2] = gld(p + 3 * i + 2);
}
template <typename Team, typename T>
__device__ __forceinline__ void tld3(const T* p, int i, T v[3]) {
  v[0] = Team::ld(p + 3 * i);
  v[1] = Team::ld(p + 3 * i + 1);
  v[2] = Team::ld(p + 3 * i + 2);
}
template <typename Team, typename T>
__device__ __forceinline__ void tst3(T* p, int i, const T v[3]) {
  Team::st(p + 3 * i, v[0]);
  Team::st(p + 3 * i + 1, v[1]);
  Team::st(p + 3 * i + 2, v[2]);
}
template <typename T>
__device__ __forceinline__ void widen3(const T v[3], double d[3]) {
  d[0] = (double)v[0];
  d[1] = (double)v[1];
  d[2] = (double)v[2];
}

// One operator application M x (assembled W, or H (Minv (H^T x)) with the
// default body step: Minv applied - or, with factored tiles, F (F^T x) and the
// body vector F^T x stored as it is - for the H pass).
template <bool MF, typename T, typename Team, typename Loader, typename Pre, typename Leader>
__device__ __forceinline__ bool apply_op(const PGProblem<T>& pb, Team& tm, Loader& ld, Pre& pre, Leader& lead) {
  const DevProblem<T>& P = pb.P;
  auto body = [&](int k, T v[6]) {
    if (P.minv_folded) {
#pragma unroll
      for (int r = 0; r < 6; ++r) Team::st(&pb.wbody[6 * k + r], v[r]);
      return;
    }
    T w[6];
    minv_apply<T>(P.minv, P.nb, k, v, w);
#pragma unroll
    for (int r = 0; r < 6; ++r) Team::st(&pb.wbody[6 * k + r], w[r]);
  };
  return apply_operator<MF, T>(P, pb.wbody, pb.carry, tm, ld, pre, lead, body);
}

__device__ __forceinline__ double nesterov(double& theta, double q) {
  const double tp = theta;
  const double tp2 = tp * tp;
  const double bb = tp2 - q;
  const double delta = sqrt(bb * bb + 4.0 * tp2);
  theta = 0.5 * (delta - bb);
  return (tp - tp2) / (tp2 + theta);
}

// Residual contribution, COV, gradient and energy terms at (x_i, Mx_i), FP64.
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

template <typename T>
__device__ __forceinline__ void proj_step(T mu, T al, const T x[3], const T g[3], T xs[3]) {
  xs[0] = x[0] - al * g[0];
  xs[1] = x[1] - al * g[1];
  xs[2] = x[2] - al * g[2];
  project<T>(mu, xs[0], xs[1], xs[2]);
}

// Column loader of a line-search trial's H^T pass: the candidate
// P(x_j - alpha g_j) of contact j (split into the loads and the projection, so
// that a pass with two tiles in flight issues both tiles' loads first).
template <typename Team, typename T>
struct CandidateLoader {
  static constexpr bool kSplit = true;
  const T* X;
  const T* G;
  const double* mu;
  T al;
  struct Raw {
    T x[3], g[3], mu;
  };
  __device__ __forceinline__ Raw fetch(int j) const {
    Raw r;
    tld3<Team, T>(X, j, r.x);
    tld3<Team, T>(G, j, r.g);
    r.mu = (T)gld(mu + j);
    return r;
  }
  __device__ __forceinline__ void finish(const Raw& r, T xc[3]) const { proj_step<T>(r.mu, al, r.x, r.g, xc); }
  __device__ __forceinline__ void operator()(int j, T xc[3]) const { finish(fetch(j), xc); }
};

// Contact-major phase of a line-search trial (pg_body's CandPhase): the candidate
// P(x_i - alpha g_i) of every contact into Xs. Team-synchronised by the caller.
template <typename Team, typename T>
__device__ __forceinline__ void candidate_phase(const Team& tm, int nc, const CandidateLoader<Team, T>& ld, T* Xs) {
  for (int i = tm.tid; i < nc; i += tm.nthreads) {
    T xs[3];
    ld(i, xs);
    tst3<Team, T>(Xs, i, xs);
  }
}
// Column loader of an H^T pass over a stored vector.
template <typename Team, typename T>
struct StoredLoader {
  const T* X;
  __device__ __forceinline__ void operator()(int j, T xc[3]) const { tld3<Team, T>(X, j, xc); }
};

struct InitRow {
  double x[3], b[3], mu, s;
};
template <typename T>
struct TrialRow {  // own-contact data of a line-search trial row
  T x[3], g[3], p[3], m[3], mp[3];
  double b[3], mu, s;
};
template <typename T>
struct SpgRow {
  T x[3], g[3], p[3], m[3], mp[3];
  double b[3], mu, s;
};

// Projected warm start of contact i (FP64), rounded to the storage type.
template <typename T>
__device__ __forceinline__ void start_point(const PGProblem<T>& pb, int i, T xc[3]) {
  double v[3];
  dld3<T>(pb.x_io, i, v);
  project<double>(gld(pb.P.mu + i), v[0], v[1], v[2]);
  xc[0] = (T)v[0];
  xc[1] = (T)v[1];
  xc[2] = (T)v[2];
}

}  // namespace

// The solve of one problem by one team. tsum accumulates phase times
// (s_memrealtime ticks): [0] operator-phase work, [1] operator-phase reduction
// and wait, [2] update-phase work, [3] update-phase wait.
// Fuse: the Nesterov point is formed speculatively inside every line-search trial
// (no update phase; pays where a team barrier is expensive: the whole grid), or by
// a separate element-wise phase after the accepted trial (cheaper per trial and in
// registers: teams synchronised by __syncthreads or few blocks).
// CandPhase (matrix-free): each line-search trial first stores its candidate
// P(x - alpha g) in a contact-major phase, and the H^T pass gathers that one vector
// instead of x, g and mu (fewer cache lines where the vectors do not fit the L2
// caches); otherwise the candidate is formed inside the gather. Same arithmetic.
template <bool MF, typename T, typename Team, bool Fuse, bool CandPhase = false>
__device__ __forceinline__ void pg_body(const PGParams& prm, const PGProblem<T>& pb, Team& tm, uint64_t tsum[4]) {
  const DevProblem<T>& P = pb.P;
  const int nc = P.nc;
  const int n3 = 3 * nc;
  const bool coulomb = prm.coulomb != 0;
  const unsigned resMax = prm.inf_norm ? 1u : 0u;  // bit of the residual partial when it is a max
  const bool spg = (prm.variant == 4);
  const bool useNesterov = (prm.variant == 3 || spg);
  int it = 0, spmvs = 0, lsTrials = 0, nest = 0;
  bool ok = true;
  double minRes = -1.0, res = 0.0, J = 0.0, alpha = 1.0, theta = 1.0;
  int rx = 0, rprev = 0, rg = 0;
  uint64_t tp0 = 0, tp1 = 0;
  // uniform solver state in scalar registers where the team asks for it
  auto U = [](double v) { return Team::kScalarState ? uni(v) : v; };
  auto rowScale = [&](int i) { return prm.use_scaling ? gld(P.scal + i) : 1.0; };

  auto initPre = [&](int i) {
    InitRow r;
    T xr[3];
    start_point<T>(pb, i, xr);
    widen3<T>(xr, r.x);
    r.mu = gld(P.mu + i);
    dld3<T>(P.b, i, r.b);
    r.s = rowScale(i);
    return r;
  };
  // ---- x <- P(x0): M x, gradient and the first residual
  {
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, T xc[3]) { start_point<T>(pb, j, xc); };
    auto lead = [&](int i, T v[3], const InitRow& r) {
      T xr[3] = {(T)r.x[0], (T)r.x[1], (T)r.x[2]};
      tst3<Team, T>(pb.X[0], i, xr);
      tst3<Team, T>(pb.MX[0], i, v);
      double vd[3], rr, g[3], Ji;
      widen3<T>(v, vd);
      eval_contact(coulomb, r.mu, r.s, r.b, r.x, vd, rr, g, Ji);
      const T gt[3] = {(T)g[0], (T)g[1], (T)g[2]};
      tst3<Team, T>(pb.G[0], i, gt);
      part[0] = resMax ? fmax(part[0], rr) : part[0] + rr;
      part[1] += Ji;
    };
    ok = apply_op<MF, T>(pb, tm, ld, initPre, lead) && tm.template reduce<2>(part, resMax);
    ++spmvs;
    if (ok) {
      res = tm.red[0];
      res = U(prm.inf_norm ? res : res / (1.0 + nc));
      J = tm.red[1];
    }
  }
  // ---- guess_alpha: alpha = min(1, |P(x0) - 1|^2 / |M (P(x0) - 1)|^2), only when iterating
  if (ok && !(res < prm.tol) && prm.max_iters > 0) {
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, T xc[3]) {
      start_point<T>(pb, j, xc);
      xc[0] -= T(1);
      xc[1] -= T(1);
      xc[2] -= T(1);
    };
    auto lead = [&](int i, T v[3], const InitRow& r) {
      double vd[3];
      widen3<T>(v, vd);
      part[0] += vd[0] * vd[0] + vd[1] * vd[1] + vd[2] * vd[2];
      part[1] += (r.x[0] - 1.0) * (r.x[0] - 1.0) + (r.x[1] - 1.0) * (r.x[1] - 1.0) + (r.x[2] - 1.0) * (r.x[2] - 1.0);
    };
    ok = apply_op<MF, T>(pb, tm, ld, initPre, lead) && tm.template reduce<2>(part, 0u);
    ++spmvs;
    if (ok) {
      const double nMx = tm.red[0], nd = tm.red[1];
      if (nMx > 1.e-16) alpha = U(fmin(1.0, nd / nMx));
    }
  }

  bool converged = false;
  double stallRef = 1.7976931348623157e308;  // best residual at the last 1% improvement
  int lastImprove = 0;
  while (ok) {
    // ---- test_residual at the current x (iteration `it`)
    if (tm.leader() && it < pb.hist_cap) gst(pb.hist + it, res);
    if (it == 0 || res < minRes) {
      minRes = res;
      team_copy(tm, pb.xbest, pb.X[rx], n3);
    }
    if (res < prm.tol) {
      converged = true;
      break;
    }
    if (minRes < 0.99 * stallRef) {
      stallRef = minRes;
      lastImprove = it;
    }
    // FP32 stage of a mixed solve: hand over to FP64 once the best residual has not
    // decreased by 1% for stall_iters iterations near the tolerance (the FP32
    // rounding floor), or for 4 x stall_iters anywhere (an APGD plateau far above
    // the tolerance is not a precision floor and usually ends by itself)
    if (prm.stall_iters > 0) {
      const int since = it - lastImprove;
      if ((since >= prm.stall_iters && minRes < 10.0 * prm.tol) || since >= 4 * prm.stall_iters) break;
    }
    // slots: the candidate xs and the Nesterov point xn take iterate slots other than
    // the current x and the previous iterate (which other members still read), and
    // gradient slots other than the current one
    int rxs = 0;
    while (rxs == rx || rxs == rprev) ++rxs;
    int rout = 0;
    while (rout == rx || rout == rprev || rout == rxs) ++rout;
    const int rgs = (rg + 1) % 3, rgn = (rg + 2) % 3;
    const T* Xx = pb.X[rx];
    const T* Gx = pb.G[rg];
    // Nesterov step for this iteration (ProjectedGradient.impl.hpp:135-149): beta
    // depends only on theta, so every trial also forms xn = xs + beta (xs - prev),
    // M xn = M xs + beta (M xs - M prev), g(xn), its residual and energy; the
    // iteration then adopts xn (decr2 < 0) or restarts at xs, without another pass.
    double thetaN = theta;
    const double beta = U(useNesterov ? nesterov(thetaN, 0.0) : 0.0);
    thetaN = U(thetaN);
    // contact i's Nesterov point from its candidate (T arithmetic as the candidate)
    auto nesterov_point = [&](int i, const T xst[3], const T v[3], const T pv[3], const T mp[3], double mu,
                              double sc, const double b[3], double& rr, double& Jn) {
      T xnt[3], Mnt[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        xnt[k] = (T)((double)xst[k] + beta * ((double)xst[k] - (double)pv[k]));
        Mnt[k] = (T)((double)v[k] + beta * ((double)v[k] - (double)mp[k]));
      }
      tst3<Team, T>(pb.X[rout], i, xnt);
      tst3<Team, T>(pb.MX[rout], i, Mnt);
      double xn[3], Mn[3], g[3];
      widen3<T>(xnt, xn);
      widen3<T>(Mnt, Mn);
      eval_contact(coulomb, mu, sc, b, xn, Mn, rr, g, Jn);
      const T gt[3] = {(T)g[0], (T)g[1], (T)g[2]};
      tst3<Team, T>(pb.G[rgn], i, gt);
    };
    if (!spg) {
      // ---- APGD / Descent: backtracking line search (ProjectedGradient.impl.hpp:117-133)
      alpha = U(alpha * prm.ls_opt);
      double Js = 0, decr = 0, nd = 0, decr2 = 0, resN = 0, JN = 0, resX = 0, JX = 0;
      auto pre = [&](int i) {
        TrialRow<T> r;
        tld3<Team, T>(Xx, i, r.x);
        tld3<Team, T>(Gx, i, r.g);
        tld3<Team, T>(pb.X[rprev], i, r.p);
        tld3<Team, T>(pb.MX[rx], i, r.m);
        if (Fuse && useNesterov) tld3<Team, T>(pb.MX[rprev], i, r.mp);
        dld3<T>(P.b, i, r.b);
        r.mu = gld(P.mu + i);
        r.s = rowScale(i);
        return r;
      };
      for (int ls = 0; ls < prm.ls_iters && ok; ++ls) {
        double part[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const T al = (T)alpha;
        CandidateLoader<Team, T> ld{Xx, Gx, P.mu, al};
        constexpr bool cph = MF && CandPhase;
        auto lead = [&](int i, T v[3], const TrialRow<T>& r) {
          T xst[3];
          proj_step<T>((T)r.mu, al, r.x, r.g, xst);
          if (!cph) tst3<Team, T>(pb.X[rxs], i, xst);
          tst3<Team, T>(pb.MX[rxs], i, v);
          double xs[3], vd[3], x[3], g[3], p[3], m[3];
          widen3<T>(xst, xs);
          widen3<T>(v, vd);
          widen3<T>(r.x, x);
          widen3<T>(r.g, g);
          widen3<T>(r.p, p);
          widen3<T>(r.m, m);
          const double s0 = cov_normal<double>(coulomb, r.mu, m[1] + r.b[1], m[2] + r.b[2]);
          part[0] += xs[0] * (0.5 * vd[0] + r.b[0] + s0) + xs[1] * (0.5 * vd[1] + r.b[1]) +
                     xs[2] * (0.5 * vd[2] + r.b[2]);
          const double d[3] = {xs[0] - x[0], xs[1] - x[1], xs[2] - x[2]};
          part[1] += g[0] * d[0] + g[1] * d[1] + g[2] * d[2];
          part[2] += d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
          part[3] += (xs[0] - p[0]) * g[0] + (xs[1] - p[1]) * g[1] + (xs[2] - p[2]) * g[2];
          double rr, gn[3], Jn;
          eval_contact(coulomb, r.mu, r.s, r.b, xs, vd, rr, gn, Jn);
          const T gt[3] = {(T)gn[0], (T)gn[1], (T)gn[2]};
          tst3<Team, T>(pb.G[rgs], i, gt);
          part[4] = resMax ? fmax(part[4], rr) : part[4] + rr;
          part[5] += Jn;
          if (Fuse && useNesterov) {
            double rx2, Jx;
            nesterov_point(i, xst, v, r.p, r.mp, r.mu, r.s, r.b, rx2, Jx);
            part[6] = resMax ? fmax(part[6], rx2) : part[6] + rx2;
            part[7] += Jx;
          }
        };
        tp0 = realtime();
        if constexpr (cph) {
          candidate_phase<Team, T>(tm, nc, ld, pb.X[rxs]);
          StoredLoader<Team, T> lds{pb.X[rxs]};
          ok = tm.sync() && apply_op<MF, T>(pb, tm, lds, pre, lead);
        } else {
          ok = apply_op<MF, T>(pb, tm, ld, pre, lead);
        }
        tp1 = realtime();
        ok = ok && tm.template reduce<8>(part, (resMax << 4) | (resMax << 6));
        tsum[0] += tp1 - tp0;
        tsum[1] += realtime() - tp1;
        ++spmvs;
        ++lsTrials;
        if (!ok) break;
        Js = tm.red[0];
        decr = tm.red[1];
        nd = tm.red[2];
        decr2 = tm.red[3];
        resN = tm.red[4];
        JN = tm.red[5];
        resX = tm.red[6];
        JX = tm.red[7];
        if (Js < J + decr + 0.5 / alpha * nd) break;
        alpha = U(alpha * prm.ls_pess);
      }
      if (!ok) break;
      if (useNesterov && decr2 < 0) {
        if (!Fuse) {
          // element-wise Nesterov update of the accepted candidate
          tp0 = realtime();
          double p2[2] = {0.0, 0.0};
          for (int i = tm.tid; i < nc; i += tm.nthreads) {
            T xst[3], pvt[3], Mst[3], Mpt[3];
            tld3<Team, T>(pb.X[rxs], i, xst);
            tld3<Team, T>(pb.X[rprev], i, pvt);
            tld3<Team, T>(pb.MX[rxs], i, Mst);
            tld3<Team, T>(pb.MX[rprev], i, Mpt);
            double b[3], rr, Jn;
            dld3<T>(P.b, i, b);
            nesterov_point(i, xst, Mst, pvt, Mpt, gld(P.mu + i), rowScale(i), b, rr, Jn);
            p2[0] = resMax ? fmax(p2[0], rr) : p2[0] + rr;
            p2[1] += Jn;
          }
          tp1 = realtime();
          ok = tm.template reduce<2>(p2, resMax);
          tsum[2] += tp1 - tp0;
          tsum[3] += realtime() - tp1;
          if (!ok) break;
          resX = tm.red[0];
          JX = tm.red[1];
        }
        res = resX;
        J = JX;
        rx = rout;
        rprev = rxs;
        rg = rgn;
        theta = thetaN;
        ++nest;
      } else {
        rx = rxs;
        rprev = rxs;
        rg = rgs;
        res = resN;
        J = JN;
        theta = 1.0;
      }
      res = U(prm.inf_norm ? res : res / (1.0 + nc));
    } else {
      // ---- SPG (ProjectedGradient.impl.hpp:316-346): one product per iteration.
      double part[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      const T al = (T)alpha;
      auto pre = [&](int i) {
        SpgRow<T> r;
        tld3<Team, T>(Xx, i, r.x);
        tld3<Team, T>(Gx, i, r.g);
        tld3<Team, T>(pb.X[rprev], i, r.p);
        tld3<Team, T>(pb.MX[rx], i, r.m);
        tld3<Team, T>(pb.MX[rprev], i, r.mp);
        dld3<T>(P.b, i, r.b);
        r.mu = gld(P.mu + i);
        r.s = rowScale(i);
        return r;
      };
      CandidateLoader<Team, T> ld{Xx, Gx, P.mu, al};
      constexpr bool cph = MF && CandPhase;
      auto lead = [&](int i, T v[3], const SpgRow<T>& r) {
        T xst[3];
        proj_step<T>((T)r.mu, al, r.x, r.g, xst);
        if (!cph) tst3<Team, T>(pb.X[rxs], i, xst);
        tst3<Team, T>(pb.MX[rxs], i, v);
        double xs[3], vd[3], x[3], g[3], p[3], m[3], mp[3];
        widen3<T>(xst, xs);
        widen3<T>(v, vd);
        widen3<T>(r.x, x);
        widen3<T>(r.g, g);
        widen3<T>(r.p, p);
        widen3<T>(r.m, m);
        widen3<T>(r.mp, mp);
        double ad = 0, ae = 0, cd = 0, ce = 0, dd = 0, de = 0, ee = 0, d2 = 0;
        for (int k = 0; k < 3; ++k) {
          const double av = xs[k] - x[k], cv = xs[k] - p[k];
          const double dv = vd[k] - m[k], ev = vd[k] - mp[k];
          d2 += cv * g[k];
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
        eval_contact(coulomb, r.mu, r.s, r.b, xs, vd, rr, gn, Jn);
        const T gt[3] = {(T)gn[0], (T)gn[1], (T)gn[2]};
        tst3<Team, T>(pb.G[rgs], i, gt);
        part[7] = resMax ? fmax(part[7], rr) : part[7] + rr;
        if (Fuse) {
          double rx2, Jx;
          nesterov_point(i, xst, v, r.p, r.mp, r.mu, r.s, r.b, rx2, Jx);
          part[8] = resMax ? fmax(part[8], rx2) : part[8] + rx2;
        }
      };
      tp0 = realtime();
      if constexpr (cph) {
        candidate_phase<Team, T>(tm, nc, ld, pb.X[rxs]);
        StoredLoader<Team, T> lds{pb.X[rxs]};
        ok = tm.sync() && apply_op<MF, T>(pb, tm, lds, pre, lead);
      } else {
        ok = apply_op<MF, T>(pb, tm, ld, pre, lead);
      }
      tp1 = realtime();
      ok = ok && tm.template reduce<9>(part, (resMax << 7) | (resMax << 8));
      tsum[0] += tp1 - tp0;
      tsum[1] += realtime() - tp1;
      ++spmvs;
      if (!ok) break;
      const double decr2 = tm.red[0];
      double wz, zz;
      if (decr2 < 0.0) {
        wz = tm.red[1] + beta * tm.red[2] + beta * beta * tm.red[3];
        zz = tm.red[4] + 2.0 * beta * tm.red[5] + beta * beta * tm.red[6];
        res = tm.red[8];
        if (!Fuse) {
          tp0 = realtime();
          double p2[1] = {0.0};
          for (int i = tm.tid; i < nc; i += tm.nthreads) {
            T xst[3], pvt[3], Mst[3], Mpt[3];
            tld3<Team, T>(pb.X[rxs], i, xst);
            tld3<Team, T>(pb.X[rprev], i, pvt);
            tld3<Team, T>(pb.MX[rxs], i, Mst);
            tld3<Team, T>(pb.MX[rprev], i, Mpt);
            double b[3], rr, Jn;
            dld3<T>(P.b, i, b);
            nesterov_point(i, xst, Mst, pvt, Mpt, gld(P.mu + i), rowScale(i), b, rr, Jn);
            p2[0] = resMax ? fmax(p2[0], rr) : p2[0] + rr;
          }
          tp1 = realtime();
          ok = tm.template reduce<1>(p2, resMax);
          tsum[2] += tp1 - tp0;
          tsum[3] += realtime() - tp1;
          if (!ok) break;
          res = tm.red[0];
        }
        rx = rout;
        rprev = rxs;
        rg = rgn;
        theta = thetaN;
        ++nest;
      } else {
        wz = tm.red[1];
        zz = tm.red[4];
        rx = rxs;
        rprev = rxs;
        rg = rgs;
        res = tm.red[7];
        theta = 1.0;
      }
      res = U(prm.inf_norm ? res : res / (1.0 + nc));
      alpha = wz / zz;
      alpha = U(fmin(1.e6, fmax(1.e-6, alpha)));
    }
    ++it;
    if (it >= prm.max_iters) break;
  }
  // x = x_best (ProjectedGradient.impl.hpp:152); the same thread wrote xbest[k] or
  // a barrier separates the writer
  if (ok)
    team_copy(tm, pb.x_io, pb.xbest, n3);
  if (tm.leader()) {
    gst(pb.out + 0, minRes);
    gst(pb.out + 1, (double)it);
    gst(pb.out + 2, (double)spmvs);
    gst(pb.out + 3, (double)lsTrials);
    gst(pb.out + 4, (double)nest);
    gst(pb.out + 5, converged ? 1.0 : 0.0);
    gst(pb.out + 6, ok ? 0.0 : 1.0);
    for (int k = 0; k < 4; ++k) gst(pb.out + 7 + k, 0.01 * (double)tsum[k]);  // microseconds (100 MHz clock)
  }
}

// PgMethod<Standard> and PgMethod<Conjugated> (ProjectedGradient.impl.hpp:157-209,
// 211-286) with armijo_ls (:54-72), for one team. x, M x, the gradient y = M x + b + s
// and J are kept as in pg_body; the line-search candidate P(x + alpha dir) is formed
// inside the gather, the direction computed on the fly from stored vectors:
//  * Standard: dir = beta (x - P(x - y)), beta = -y.pg / (pg . M pg), which takes one
//    operator application on pg before the line search;
//  * Conjugated: dir = -y or beta dir_prev - y (Polak-Ribiere or Hestenes-Stiefel beta
//    from the projected gradients x - P(x - y) of this and the previous iteration),
//    with the test dir . y >= 0 evaluated from reduced dot products.
// Slots: x and its candidate alternate in X[0..1], MX[0..1], G[0..1]; Conjugated keeps
// the previous direction in X[2] and the previous projected gradient in X[3].
template <bool MF, typename T, typename Team>
__device__ __forceinline__ void pg_body_sc(const PGParams& prm, const PGProblem<T>& pb, Team& tm) {
  const DevProblem<T>& P = pb.P;
  const int nc = P.nc;
  const int n3 = 3 * nc;
  const bool coulomb = prm.coulomb != 0;
  const unsigned resMax = prm.inf_norm ? 1u : 0u;
  const bool conj = (prm.variant == 2);
  int it = 0, spmvs = 0, lsTrials = 0;
  bool ok = true;
  double minRes = -1.0, res = 0.0, J = 0.0, alpha = 1.0;
  int rx = 0, rg = 0;
  auto rowScale = [&](int i) { return prm.use_scaling ? gld(P.scal + i) : 1.0; };
  auto initPre = [&](int i) {
    InitRow r;
    T xr[3];
    start_point<T>(pb, i, xr);
    widen3<T>(xr, r.x);
    r.mu = gld(P.mu + i);
    dld3<T>(P.b, i, r.b);
    r.s = rowScale(i);
    return r;
  };
  // ---- x <- P(x0): M x, gradient and the first residual
  {
    double part[2] = {0.0, 0.0};
    auto ld = [&](int j, T xc[3]) { start_point<T>(pb, j, xc); };
    auto lead = [&](int i, T v[3], const InitRow& r) {
      T xr[3] = {(T)r.x[0], (T)r.x[1], (T)r.x[2]};
      tst3<Team, T>(pb.X[0], i, xr);
      tst3<Team, T>(pb.MX[0], i, v);
      if (conj) {
        const T z3[3] = {T(0), T(0), T(0)};
        tst3<Team, T>(pb.X[3], i, xr);  // prev_proj_grad = x
        tst3<Team, T>(pb.X[2], i, z3);  // no previous direction yet
      }
      double vd[3], rr, g[3], Ji;
      widen3<T>(v, vd);
      eval_contact(coulomb, r.mu, r.s, r.b, r.x, vd, rr, g, Ji);
      const T gt[3] = {(T)g[0], (T)g[1], (T)g[2]};
      tst3<Team, T>(pb.G[0], i, gt);
      part[0] = resMax ? fmax(part[0], rr) : part[0] + rr;
      part[1] += Ji;
    };
    ok = apply_op<MF, T>(pb, tm, ld, initPre, lead) && tm.template reduce<2>(part, resMax);
    ++spmvs;
    if (ok) {
      res = tm.red[0];
      if (!prm.inf_norm) res /= (1.0 + nc);
      J = tm.red[1];
    }
  }
  // x - P(x - y) of contact j (the projected gradient)
  auto projGrad = [&](T mu, const T x[3], const T y[3], T pgv[3]) {
    T z[3] = {x[0] - y[0], x[1] - y[1], x[2] - y[2]};
    project<T>(mu, z[0], z[1], z[2]);
    pgv[0] = x[0] - z[0];
    pgv[1] = x[1] - z[1];
    pgv[2] = x[2] - z[2];
  };
  bool converged = false;
  double stallRef = 1.7976931348623157e308;
  int lastImprove = 0;
  double prevN2 = 0.0;
  while (ok) {
    // ---- test_residual at the current x
    if (tm.leader() && it < pb.hist_cap) gst(pb.hist + it, res);
    if (it == 0 || res < minRes) {
      minRes = res;
      team_copy(tm, pb.xbest, pb.X[rx], n3);
    }
    if (res < prm.tol) {
      converged = true;
      break;
    }
    if (minRes < 0.99 * stallRef) {
      stallRef = minRes;
      lastImprove = it;
    }
    if (prm.stall_iters > 0) {
      const int since = it - lastImprove;
      if ((since >= prm.stall_iters && minRes < 10.0 * prm.tol) || since >= 4 * prm.stall_iters) break;
    }
    const int rxs = 1 - rx, rgs = 1 - rg;
    const T* Xx = pb.X[rx];
    const T* Gx = pb.G[rg];
    T* dirPrev = pb.X[2];
    T* pgPrev = pb.X[3];
    // direction: dir_j = cx * pg_j (Standard), or cd * dirPrev_j - y_j (Conjugated)
    double cx = 0.0, cd = 0.0;
    if (!conj) {
      // beta = -y . pg / (pg . M pg), one operator application on pg
      double part[2] = {0.0, 0.0};
      auto ld = [&](int j, T xc[3]) {
        T xv[3], gv[3];
        tld3<Team, T>(Xx, j, xv);
        tld3<Team, T>(Gx, j, gv);
        projGrad((T)gld(P.mu + j), xv, gv, xc);
      };
      auto pre = [&](int i) {
        SpgRow<T> r;
        tld3<Team, T>(Xx, i, r.x);
        tld3<Team, T>(Gx, i, r.g);
        r.mu = gld(P.mu + i);
        return r;
      };
      auto lead = [&](int i, T v[3], const SpgRow<T>& r) {
        T pgv[3];
        projGrad((T)r.mu, r.x, r.g, pgv);
        double pd[3], gd[3], vd[3];
        widen3<T>(pgv, pd);
        widen3<T>(r.g, gd);
        widen3<T>(v, vd);
        part[0] += gd[0] * pd[0] + gd[1] * pd[1] + gd[2] * pd[2];
        part[1] += pd[0] * vd[0] + pd[1] * vd[1] + pd[2] * vd[2];
      };
      ok = apply_op<MF, T>(pb, tm, ld, pre, lead) && tm.template reduce<2>(part, 0u);
      ++spmvs;
      if (!ok) break;
      cx = -tm.red[0] / tm.red[1];
      alpha = fmin(1.0, alpha * prm.ls_opt);
    } else {
      // projected gradient, its norm and the conjugation dot products (one pass)
      double part[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
      for (int i = tm.tid; i < nc; i += tm.nthreads) {
        T xv[3], gv[3], pgv[3], pp[3], dp[3];
        tld3<Team, T>(Xx, i, xv);
        tld3<Team, T>(Gx, i, gv);
        tld3<Team, T>(pgPrev, i, pp);
        tld3<Team, T>(dirPrev, i, dp);
        projGrad((T)gld(P.mu + i), xv, gv, pgv);
        tst3<Team, T>(pgPrev, i, pgv);
        double pd[3], ppd[3], dpd[3], gd[3];
        widen3<T>(pgv, pd);
        widen3<T>(pp, ppd);
        widen3<T>(dp, dpd);
        widen3<T>(gv, gd);
        for (int c = 0; c < 3; ++c) {
          part[0] += pd[c] * pd[c];
          part[1] += dpd[c] * (pd[c] - ppd[c]);
          part[2] += pd[c] * ppd[c];
          part[3] += dpd[c] * gd[c];
          part[4] += gd[c] * gd[c];
        }
      }
      ok = tm.template reduce<5>(part, 0u);
      if (!ok) break;
      const double ng2 = tm.red[0];
      if (prevN2 == 0.0) {
        cd = 0.0;
      } else {
        const double den = tm.red[1], xp = tm.red[2];
        const double beta = (den < 1.e-12) ? fmax(0.0, ng2 - xp) / prevN2 : fmax(0.0, ng2 - xp) / den;
        // dir . y = beta dirPrev . y - y . y; a non-descent direction falls back to -y
        cd = (beta * tm.red[3] - tm.red[4] >= 0.0) ? 0.0 : beta;
      }
      prevN2 = ng2;
      alpha *= prm.ls_opt;
    }
    // ---- armijo_ls: xs = P(x + alpha dir)
    double Js = J, decr = 0.0, resN = 0.0, JN = 0.0;
    auto dirOf = [&](int j, const T xv[3], const T gv[3], T mu, T d[3]) {
      if (!conj) {
        T pgv[3];
        projGrad(mu, xv, gv, pgv);
        for (int c = 0; c < 3; ++c) d[c] = (T)cx * pgv[c];
      } else {
        T dp[3] = {T(0), T(0), T(0)};
        if (cd != 0.0) tld3<Team, T>(dirPrev, j, dp);
        for (int c = 0; c < 3; ++c) d[c] = (T)cd * dp[c] - gv[c];
      }
    };
    auto pre = [&](int i) {
      TrialRow<T> r;
      tld3<Team, T>(Xx, i, r.x);
      tld3<Team, T>(Gx, i, r.g);
      tld3<Team, T>(pb.MX[rx], i, r.m);
      dld3<T>(P.b, i, r.b);
      r.mu = gld(P.mu + i);
      r.s = rowScale(i);
      return r;
    };
    for (int ls = 0; ls < prm.ls_iters && ok; ++ls) {
      double part[4] = {0, 0, 0, 0};
      const T al = (T)alpha;
      auto ld = [&](int j, T xc[3]) {
        T xv[3], gv[3], d[3];
        tld3<Team, T>(Xx, j, xv);
        tld3<Team, T>(Gx, j, gv);
        const T mu = (T)gld(P.mu + j);
        dirOf(j, xv, gv, mu, d);
        for (int c = 0; c < 3; ++c) xc[c] = xv[c] + al * d[c];
        project<T>(mu, xc[0], xc[1], xc[2]);
      };
      auto lead = [&](int i, T v[3], const TrialRow<T>& r) {
        T d[3], xst[3];
        dirOf(i, r.x, r.g, (T)r.mu, d);
        for (int c = 0; c < 3; ++c) xst[c] = r.x[c] + al * d[c];
        project<T>((T)r.mu, xst[0], xst[1], xst[2]);
        tst3<Team, T>(pb.X[rxs], i, xst);
        tst3<Team, T>(pb.MX[rxs], i, v);
        double xs[3], vd[3], x[3], g[3], m[3];
        widen3<T>(xst, xs);
        widen3<T>(v, vd);
        widen3<T>(r.x, x);
        widen3<T>(r.g, g);
        widen3<T>(r.m, m);
        const double s0 = cov_normal<double>(coulomb, r.mu, m[1] + r.b[1], m[2] + r.b[2]);
        part[0] += xs[0] * (0.5 * vd[0] + r.b[0] + s0) + xs[1] * (0.5 * vd[1] + r.b[1]) +
                   xs[2] * (0.5 * vd[2] + r.b[2]);
        part[1] += g[0] * (xs[0] - x[0]) + g[1] * (xs[1] - x[1]) + g[2] * (xs[2] - x[2]);
        double rr, gn[3], Jn;
        eval_contact(coulomb, r.mu, r.s, r.b, xs, vd, rr, gn, Jn);
        const T gt[3] = {(T)gn[0], (T)gn[1], (T)gn[2]};
        tst3<Team, T>(pb.G[rgs], i, gt);
        part[2] = resMax ? fmax(part[2], rr) : part[2] + rr;
        part[3] += Jn;
      };
      ok = apply_op<MF, T>(pb, tm, ld, pre, lead) && tm.template reduce<4>(part, resMax << 2);
      ++spmvs;
      ++lsTrials;
      if (!ok) break;
      Js = tm.red[0];
      decr = tm.red[1];
      resN = tm.red[2];
      JN = tm.red[3];
      if (Js < J + prm.ls_armijo * decr) break;
      alpha *= prm.ls_pess;
    }
    if (!ok) break;
    if (conj) {
      // prev_dir = (xs - x) / alpha (the line search's last alpha)
      for (int i = tm.tid; i < nc; i += tm.nthreads) {
        T xv[3], xsv[3], d[3];
        tld3<Team, T>(Xx, i, xv);
        tld3<Team, T>(pb.X[rxs], i, xsv);
        for (int c = 0; c < 3; ++c) d[c] = (T)(((double)xsv[c] - (double)xv[c]) / alpha);
        tst3<Team, T>(dirPrev, i, d);
      }
      ok = tm.sync();
      if (!ok) break;
    }
    rx = rxs;
    rg = rgs;
    res = resN;
    J = JN;
    if (!prm.inf_norm) res /= (1.0 + nc);
    ++it;
    if (it >= prm.max_iters) break;
  }
  if (ok) team_copy(tm, pb.x_io, pb.xbest, n3);
  if (tm.leader()) {
    gst(pb.out + 0, minRes);
    gst(pb.out + 1, (double)it);
    gst(pb.out + 2, (double)spmvs);
    gst(pb.out + 3, (double)lsTrials);
    gst(pb.out + 4, 0.0);
    gst(pb.out + 5, converged ? 1.0 : 0.0);
    gst(pb.out + 6, ok ? 0.0 : 1.0);
    for (int k = 0; k < 4; ++k) gst(pb.out + 7 + k, 0.0);
  }
}

// ------------------------------------------------------------------ kernels

// Single-problem teams: vectors written with plain stores and each
// XCD's L2 written back once per barrier (GridTeamWB), or written through to memory
// store by store (GridTeam)
#ifndef SB_GRID_XCD_WRITEBACK
#define SB_GRID_XCD_WRITEBACK 1
#endif
using GridTeamSel =
    std::conditional_t<SB_GRID_XCD_WRITEBACK != 0, GridTeamT<true, true, false, true>, GridTeamT<true, false, false, true>>;

// The grid team's per-wave LDS caches of operator data (waves beyond kGridCacheWaves,
// or a team without caches, read everything from memory)
#define SB_GRID_CACHE_ATTACH(MFV, T, tm)                                                      \
  if constexpr ((MFV) && sizeof(T) == 4 && GridTeamSel::kPairsCache) {                          \
    __shared__ PairsCache<float>::T2 s_pc[kGridCacheWaves][PairsCache<float>::kWords];          \
    __shared__ float s_ht[kGridCacheWaves][18 * SB_WAVE];                                       \
    if (wave_in_block() < kGridCacheWaves) {                                                    \
      tm.pcache.lds = s_pc[wave_in_block()];                                                    \
      tm.pcache.htv = s_ht[wave_in_block()];                                                    \
    }                                                                                           \
  } else if constexpr (sizeof(T) == 8 && GridTeamSel::kTileCache) {                             \
    __shared__ double s_tc[kGridCacheWaves][TileCache<double>::kBytes / sizeof(double)];        \
    if (wave_in_block() < kGridCacheWaves) tm.tcache.lds = reinterpret_cast<char*>(s_tc[wave_in_block()]); \
  }

// One problem on the whole co-resident grid (the solver state stays in vector
// registers: at one workgroup per CU the scalar variant's spills cost 3-4 %, and more
// workgroups per CU did not shorten an operator application).
// APGD, Descent, SPG.
template <bool MF, typename T>
__global__ void __launch_bounds__(512) pg_kernel(PGGridArgs<T> a) {
  GridTeamSel tm(a.gs, blockIdx.x);
  SB_GRID_CACHE_ATTACH(MF, T, tm)
  if (SB_GRID_XCD_WRITEBACK && !tm.group_by_xcd()) return;
  uint64_t tsum[4] = {0, 0, 0, 0};
  pg_body<MF, T, GridTeamSel, true>(a.prm, a.pb, tm, tsum);
}

// Standard, Conjugated (a kernel of their own: with both solver bodies in one kernel
// the FP64 matrix-free kernel spilled registers).
template <bool MF, typename T>
__global__ void __launch_bounds__(512) pg_kernel_sc(PGGridArgs<T> a) {
  GridTeamSel tm(a.gs, blockIdx.x);
  SB_GRID_CACHE_ATTACH(MF, T, tm)
  if (SB_GRID_XCD_WRITEBACK && !tm.group_by_xcd()) return;
  pg_body_sc<MF, T, GridTeamSel>(a.prm, a.pb, tm);
}

// Teams of several workgroups are placed on one XCD. They write their vectors (and
// their reductions' partial sums) with plain stores, kept in the XCD's L2, and read
// what other workgroups wrote with plain loads after the barrier's acquire. A team
// found spread over several XCDs also writes the L2 back before each barrier.
// Measured on the 512-problem batch: coherent loads without invalidation
// (SB_BATCH_TEAM_ACQUIRE=0) 2x slower than invalidation; coherent (written-through)
// stores (SB_BATCH_PLAIN_STORES=0) 16 % slower than plain ones.
#ifndef SB_BATCH_TEAM_ACQUIRE
#define SB_BATCH_TEAM_ACQUIRE 1
#endif
#ifndef SB_BATCH_PLAIN_STORES
#define SB_BATCH_PLAIN_STORES 1
#endif
using BatchTeamSel =
    GridTeamT<SB_BATCH_TEAM_ACQUIRE != 0, SB_BATCH_PLAIN_STORES != 0 && SB_BATCH_TEAM_ACQUIRE != 0, true>;

// Many problems. The block map names each block's class, team and rank; its
// team takes the next problem of the class's queue and solves it alone. Teams
// never wait for other teams. A team of k > 1 blocks needs its blocks resident
// together (the grid is sized to the occupancy limit); its barrier waits are
// bounded and set the sticky error flag on timeout.
template <typename T>
__device__ __forceinline__ GridSync batch_team_sync(const PGBatchArgs<T>& a, int gt, int k) {
  GridSync g;
  g.group = a.team_sync + (size_t)gt * kTeamSyncWords;
  g.release = g.group + kSyncGroups * kSyncStride;
  g.top = g.group + 2 * kSyncGroups * kSyncStride;
  g.error = a.error;
  g.partials = a.team_partials + (size_t)gt * 2 * kMaxTeamBlocks * kMaxPartials;
  g.nblocks = (unsigned)k;
  g.timeout_ticks = a.timeout_ticks;
  return g;
}

// A block whose own class has no problems left joins a helper team of class 0's size
// on its XCD (the j-th block of the XCD to join: team j / K, rank j % K). The team's
// last member starts it (forming -> running) unless a waiting member has found class
// 0's queue empty and cancelled it (forming -> cancelled); each state change is one
// compare-and-swap, so every member sees the same outcome. A running helper team
// solves class 0's problems exactly as class 0's own teams do (same size, same rank
// order: the same arithmetic), with its blocks on one XCD. Waits are bounded: a team
// still forming after the deadline is cancelled (nothing is lost: its blocks had no
// work of their own). Returns whether this block joined a running team (gt, rank).
template <typename T>
__device__ __forceinline__ bool batch_helper_join(const PGBatchArgs<T>& a, int& s_next, int& gt, unsigned& rank) {
  const BatchClass cl = a.cls[0];
  const int K = cl.k;
  const int nq = cl.q1 - cl.q0;
  if (a.pool_per_xcd <= 0 || K < 2) return false;
  __syncthreads();
  if (threadIdx.x == 0)
    s_next = (__hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0 &&
              (int)__hip_atomic_load(&a.queue[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < nq)
                 ? 1
                 : 0;
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(s_next) == 0) return false;
  const unsigned x = xcd_id() & 7u;
  __syncthreads();
  if (threadIdx.x == 0) s_next = (int)atomicAdd(&a.pool_count[x * kSyncStride], 1u);
  __syncthreads();
  const int j = __builtin_amdgcn_readfirstlane(s_next);
  const int t = j / K;
  rank = (unsigned)(j % K);
  if (t >= a.pool_per_xcd) return false;
  const int pt = (int)x * a.pool_per_xcd + t;
  unsigned* state = a.pool_state + (size_t)pt * kSyncStride;
  __syncthreads();
  if (threadIdx.x == 0) {
    int go = 0;
    if ((int)rank == K - 1) {
      unsigned expected = 0u;
      go = __hip_atomic_compare_exchange_strong(state, &expected, 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
               ? 1
               : 0;
    } else {
      const uint64_t t0 = realtime();
      for (;;) {
        const unsigned v = __hip_atomic_load(state, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v == 1u) {
          go = 1;
          break;
        }
        if (v == 2u) break;
        const bool drained =
            (int)__hip_atomic_load(&a.queue[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= nq ||
            __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
            realtime() - t0 > a.timeout_ticks;
        if (drained) {
          unsigned expected = 0u;
          if (__hip_atomic_compare_exchange_strong(state, &expected, 2u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT))
            break;  // cancelled by this member
          continue;  // the last member started the team first: go
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    s_next = go;
  }
  __syncthreads();
  if (__builtin_amdgcn_readfirstlane(s_next) == 0) return false;
  gt = a.nteams + pt;
  return true;
}

template <bool MF, typename T, int NT>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT <= 256 ? 3 : 1)))
pg_batch_kernel(PGBatchArgs<T> a, const PGProblem<T>* __restrict__ probs) {
  // the Nesterov point is formed inside the line-search trials in both stages (with the
  // uniform state in scalar registers the FP32 stage keeps 3 waves per SIMD with it:
  // 153 VGPRs; mixed batch 103.5 -> 99.6 ms against a separate update phase)
  constexpr bool kFuse = true;
  __shared__ int s_next;
  const int code = a.bmap[blockIdx.x];
  if (code < 0) return;  // a block no team uses
  // FP32: the first unit of H-pass block values and column indices, and the indices of
  // the first two H^T tiles, of each of the first kPairsCacheWaves waves stay in LDS
  // while the team solves a problem (10.5 KB per wave)
  PairsCache<float> pc{nullptr, false, false};
  TileCache<double> tc{nullptr, {false, false}};
  if constexpr (sizeof(T) == 4 && BlockTeam::kPairsCache) {
    __shared__ PairsCache<float>::T2 s_pc[kPairsCacheWaves][PairsCache<float>::kWords];
    const int w = wave_in_block();
    if (w < kPairsCacheWaves) pc.lds = s_pc[w];
  } else if constexpr (MF && sizeof(T) == 8 && BlockTeam::kTileCache) {
    // FP64: the first H-pass tile of each of the first kPairsCacheWaves waves (region 0 of
    // a TileCache only: 9.8 KB per wave)
    __shared__ double s_tc[kPairsCacheWaves][TileCache<double>::kPassBytes / sizeof(double)];
    const int w = wave_in_block();
    if (w < kPairsCacheWaves) tc.lds = reinterpret_cast<char*>(s_tc[w]);
  }
  // this block's team: the plan's (class, team, rank), later a helper team of class 0
  int c = code >> 20;
  unsigned rank = (unsigned)(code & 15);
  int gt = a.cls[c].team0 + ((code >> 4) & 0xffff);
  for (;;) {
    const BatchClass cl = a.cls[c];
    const int nq = cl.q1 - cl.q0;
    if (cl.k == 1) {
      for (;;) {
        if (threadIdx.x == 0) s_next = (int)atomicAdd(&a.queue[c], 1u);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_next);
        __syncthreads();
        if (q >= nq) break;
        const int p = __builtin_amdgcn_readfirstlane(a.order[cl.q0 + q]);
        BlockTeam tm;
        tm.pcache = pc;
        tm.tcache = tc;
        uint64_t tsum[4] = {0, 0, 0, 0};
        pg_body<MF, T, BlockTeam, kFuse>(a.prm, probs[p], tm, tsum);
        __syncthreads();
      }
    } else {
      BatchTeamSel tm(batch_team_sync(a, gt, cl.k), rank);
      // Which XCDs the team's blocks run on: one bit each, gathered through the first
      // barrier. A team on one XCD (a helper team always is) shares that XCD's L2: its
      // plain stores are visible to its other blocks without an L2 write-back.
      unsigned* xcds = tm.bar.gs.top + 1;
      if (threadIdx.x == 0)
        __hip_atomic_fetch_or(xcds, 1u << xcd_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (!tm.sync()) return;
      if (threadIdx.x == 0) s_next = (int)__hip_atomic_load(xcds, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __syncthreads();
      const unsigned xm = (unsigned)__builtin_amdgcn_readfirstlane(s_next);
      tm.bar.one_xcd = (xm & (xm - 1)) == 0;
      for (;;) {
        // the team leader draws the next problem and hands it over through the barrier
        if (rank == 0 && threadIdx.x == 0) cst(&a.team_slot[gt], (int)atomicAdd(&a.queue[c], 1u));
        if (!tm.sync()) return;
        if (threadIdx.x == 0) s_next = __hip_atomic_load(&a.team_slot[gt], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(s_next);
        // the slot is rewritten only after the next problem's first barrier, which every
        // member passes after this read
        if (q >= nq) break;
        const int p = __builtin_amdgcn_readfirstlane(a.order[cl.q0 + q]);
        tm.pcache = pc;
        tm.tcache = tc;
        uint64_t tsum[4] = {0, 0, 0, 0};
        pg_body<MF, T, BatchTeamSel, kFuse>(a.prm, probs[p], tm, tsum);
        if (__hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
      }
    }
    // this block's class has drained: help class 0 (the largest teams) if it has not
    if (c == 0 || !batch_helper_join(a, s_next, gt, rank)) return;
    c = 0;
  }
}

// The matrix-free single-problem solver (APGD, Descent, SPG) with each line-search
// candidate formed in a contact-major phase of its own.
template <typename T>
__global__ void __launch_bounds__(512) pg_kernel_cand(PGGridArgs<T> a) {
  GridTeamSel tm(a.gs, blockIdx.x);
  SB_GRID_CACHE_ATTACH(true, T, tm)
  if (SB_GRID_XCD_WRITEBACK && !tm.group_by_xcd()) return;
  uint64_t tsum[4] = {0, 0, 0, 0};
  pg_body<true, T, GridTeamSel, true, true>(a.prm, a.pb, tm, tsum);
}

template __global__ void pg_kernel_cand<double>(PGGridArgs<double>);
template __global__ void pg_kernel_cand<float>(PGGridArgs<float>);
template __global__ void pg_kernel<false, double>(PGGridArgs<double>);
template __global__ void pg_kernel<true, double>(PGGridArgs<double>);
template __global__ void pg_kernel<false, float>(PGGridArgs<float>);
template __global__ void pg_kernel<true, float>(PGGridArgs<float>);
template __global__ void pg_kernel_sc<false, double>(PGGridArgs<double>);
template __global__ void pg_kernel_sc<true, double>(PGGridArgs<double>);
template __global__ void pg_kernel_sc<false, float>(PGGridArgs<float>);
template __global__ void pg_kernel_sc<true, float>(PGGridArgs<float>);

#define SB_BATCH_INST(MF, T, NT) \
  template __global__ void pg_batch_kernel<MF, T, NT>(PGBatchArgs<T>, const PGProblem<T>* __restrict__);
SB_BATCH_INST(false, double, 256)
SB_BATCH_INST(true, double, 256)
SB_BATCH_INST(false, float, 256)
SB_BATCH_INST(true, float, 256)
SB_BATCH_INST(false, double, 512)
SB_BATCH_INST(true, double, 512)
SB_BATCH_INST(false, float, 512)
SB_BATCH_INST(true, float, 512)
SB_BATCH_INST(false, double, 1024)
SB_BATCH_INST(true, double, 1024)
SB_BATCH_INST(false, float, 1024)
SB_BATCH_INST(true, float, 1024)
#undef SB_BATCH_INST

}  // namespace sb
