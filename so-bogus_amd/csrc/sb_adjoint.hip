// Reverse-mode derivatives of the friction solve (no counterpart in the
// reference, which has no differentiation).
//
// The differentiated map is the projected fixed-point step of the dual problem
//   T(r) = P_mu(z),  z = r - rho (u + s_mu(u)),  u = W r + b,
// with P_mu the projection on the cones K_mu (SOCLaw::projectOnConstraint,
// SOCLaw.impl.hpp:47-76) and s_mu(u) = mu |u_T| e_N the De Saxce term
// (SOCLaw.hpp:70-80; 0 for the SOC law). Its fixed points are the solutions the
// projected-gradient solvers converge to. W = H Minv H^T is applied matrix-free.
// For a loss L = g . r the kernels return dL/db, dL/dmu and dL/dMinv_k (6x6).
//
//  * unroll: r_{k+1} = T(r_k) for k < K from a given r_0, L = g . r_K; the
//    reverse sweep visits k = K-1 .. 0 with lambda_K = g:
//      mu_z = J_P(z_k)^T lambda_{k+1}
//      mu_u = -rho (I + J_s(u_k))^T mu_z                    (= dL/du_k)
//      dL/db += mu_u,  dL/dmu_i += lambda_i . dP/dmu - rho |u_T| mu_z,N
//      dL/dMinv_k += (H^T mu_u)_k (H^T r_k)_k^T
//      lambda_k = mu_z + W mu_u
//    Iterates are recomputed segment by segment from checkpoints stored every C
//    steps by the forward pass (memory (K/C + C) vectors instead of K).
//    With branch_mode 1 the projection branch of every (step, contact) is recorded
//    and, for the Coulomb law, the direction d = u_T / |u_T| of the De Saxce term;
//    branch_mode 2 replays both: each contact's projection formula is the recorded
//    branch's and its De Saxce term is mu (d . u_T), which has the value and the
//    gradient of mu |u_T| at the recorded point. The replayed map is then smooth in
//    the parameters (no kink of the cone projection and none of |u_T| at u_T = 0,
//    which sticking contacts approach), so central differences of it check the
//    gradient of the recorded run exactly.
//  * implicit: at a fixed point r* = T(r*), with A = dT/dr frozen at r*, the adjoint
//    system (I - A^T) lambda = g is solved by restarted GMRES (Arnoldi with
//    classical Gram-Schmidt applied twice, Givens rotations; the Hessenberg matrix
//    is updated redundantly by every block from identical reduced values) to a
//    relative residual tol; the gradients are (dT/dtheta)^T lambda*. I - A^T is
//    ill-conditioned (its smallest eigenvalues are rho times the smallest of W on
//    the sticking contacts' subspace: thousands of Neumann iterations), and singular
//    only where the fixed point is not isolated.
#include <hip/hip_runtime.h>

#include "sb_law.hpp"
#include "sb_op.hpp"
#include "sb_problem.hpp"
#include "sb_team.hpp"

namespace sb {

namespace {

template <typename Team>
__device__ __forceinline__ void ld3t(const double* p, int i, double v[3]) {
  v[0] = Team::ld(p + 3 * i);
  v[1] = Team::ld(p + 3 * i + 1);
  v[2] = Team::ld(p + 3 * i + 2);
}
template <typename Team>
__device__ __forceinline__ void st3t(double* p, int i, const double v[3]) {
  Team::st(p + 3 * i, v[0]);
  Team::st(p + 3 * i + 1, v[1]);
  Team::st(p + 3 * i + 2, v[2]);
}

// Direction d = u_T / |u_T| of the De Saxce term (0 at u_T = 0) and |u_T|.
__device__ __forceinline__ void tangent_dir(const double u[3], double d[2], double& ut) {
  ut = sqrt(u[1] * u[1] + u[2] * u[2]);
  d[0] = ut > 0.0 ? u[1] / ut : 0.0;
  d[1] = ut > 0.0 ? u[2] / ut : 0.0;
}

// z = r - rho (u + mu te e_N), te = |u_T| (or d . u_T for a replayed direction).
__device__ __forceinline__ void step_point(bool coulomb, double mu, double rho, const double r[3], const double u[3],
                                           double te, double z[3]) {
  const double s0 = coulomb ? mu * te : 0.0;
  z[0] = r[0] - rho * (u[0] + s0);
  z[1] = r[1] - rho * u[1];
  z[2] = r[2] - rho * u[2];
}

// mu_z = J^T lam, mu_u = -rho (I + J_s)^T mu_z with J_s = e_N (mu d)^T.
__device__ __forceinline__ void adjoint_point(bool coulomb, double mu, double rho, const double J[9], const double d[2],
                                              const double lam[3], double mz[3], double mu_u[3]) {
  mz[0] = J[0] * lam[0] + J[3] * lam[1] + J[6] * lam[2];
  mz[1] = J[1] * lam[0] + J[4] * lam[1] + J[7] * lam[2];
  mz[2] = J[2] * lam[0] + J[5] * lam[1] + J[8] * lam[2];
  double a1 = 0.0, a2 = 0.0;
  if (coulomb && mu > 0.0) {
    a1 = mu * d[0];
    a2 = mu * d[1];
  }
  mu_u[0] = -rho * mz[0];
  mu_u[1] = -rho * (mz[1] + a1 * mz[0]);
  mu_u[2] = -rho * (mz[2] + a2 * mz[0]);
}

struct FwdRow {
  double r[3], b[3], mu;
};
struct RevRow {
  double r[3], b[3], lam[3], mu;
};

}  // namespace

// Forward: K steps of T from r0 (checkpoints every C steps), the loss g . r_K.
// Reverse (unless g is null): the sweep above. Implicit: Neumann iterations at r0.
template <typename Team>
__device__ void adjoint_body(const AdjArgs& a, Team& tm) {
  const DevProblem<double>& P = a.P;
  const int nc = P.nc, n3 = 3 * nc;
  const bool coulomb = a.coulomb != 0;
  const double rho = a.rho;
  bool ok = true;
  auto bodyDefault = [&](int k, double v[6]) {
    double w[6];
    minv_apply<double>(P.minv, P.nb, k, v, w);
#pragma unroll
    for (int r = 0; r < 6; ++r) Team::st(&a.wbody[6 * k + r], w[r]);
  };
  // branch of contact i at step k: from the tests (recorded when asked) or replayed
  auto branch_of = [&](int k, int i, double mu, const double z[3]) -> int {
    if (a.branch_mode == 2) return a.branch[(size_t)k * nc + i];
    const int br = project_branch<double>(mu, z);
    if (a.branch_mode == 1) a.branch[(size_t)k * nc + i] = (unsigned char)br;
    return br;
  };
  // De Saxce term of contact i at step k: its direction d and te = |u_T| (recorded when
  // asked), or the replayed direction and te = d . u_T
  const bool useDir = coulomb && a.sdir != nullptr;
  auto desaxce = [&](int k, int i, const double u[3], double d[2], double& te) {
    if (useDir && a.branch_mode == 2) {
      const double* sd = a.sdir + 2 * ((size_t)k * nc + i);
      d[0] = sd[0];
      d[1] = sd[1];
      te = d[0] * u[1] + d[1] * u[2];
      return;
    }
    tangent_dir(u, d, te);
    if (useDir && a.branch_mode == 1) {
      double* sd = a.sdir + 2 * ((size_t)k * nc + i);
      sd[0] = d[0];
      sd[1] = d[1];
    }
  };
  // r_dst = T(r_src), step k
  auto forward_step = [&](const double* src, double* dst, int k) -> bool {
    auto ld = [&](int j, double xc[3]) { ld3t<Team>(src, j, xc); };
    auto pre = [&](int i) {
      FwdRow w;
      ld3t<Team>(src, i, w.r);
      w.b[0] = P.b[3 * i];
      w.b[1] = P.b[3 * i + 1];
      w.b[2] = P.b[3 * i + 2];
      w.mu = P.mu[i];
      return w;
    };
    auto lead = [&](int i, double v[3], const FwdRow& w) {
      const double u[3] = {v[0] + w.b[0], v[1] + w.b[1], v[2] + w.b[2]};
      double z[3], d[2], te;
      desaxce(k, i, u, d, te);
      step_point(coulomb, w.mu, rho, w.r, u, te, z);
      project_on_branch<double>(branch_of(k, i, w.mu, z), w.mu, z);
      st3t<Team>(dst, i, z);
    };
    return apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, pre, lead, bodyDefault) && tm.sync();
  };
  auto copy = [&](double* dst, const double* src) {
    team_copy(tm, dst, src, n3);
  };
  // zero the gradients
  if (a.g) {
    for (int k = tm.tid; k < n3; k += tm.nthreads) Team::st(a.grad_b + k, 0.0);
    for (int k = tm.tid; k < nc; k += tm.nthreads) Team::st(a.grad_mu + k, 0.0);
    for (int k = tm.tid; k < 36 * P.nb; k += tm.nthreads) Team::st(a.grad_minv + k, 0.0);
  }
  int iters = 0;
  double change = 0.0;
  bool converged = false;
  if (a.mode == 0) {
    // ---- forward with checkpoints
    copy(a.ckpt, a.r0);
    ok = tm.sync();
    const int C = a.C;
    for (int k = 0; k < a.K && ok; ++k) {
      const double* src = (k % C == 0) ? a.ckpt + (size_t)(k / C) * n3 : a.seg + (size_t)(k & 1) * n3;
      double* dst = ((k + 1) % C == 0 || k + 1 == a.K) ? a.ckpt + (size_t)((k + 1) / C + ((k + 1) % C ? 1 : 0)) * n3
                                                       : a.seg + (size_t)((k + 1) & 1) * n3;
      ok = forward_step(src, dst, k);
    }
    // r_K sits in checkpoint slot ceil(K / C)
    const double* rK = a.ckpt + (size_t)((a.K + C - 1) / C) * n3;
    if (ok) copy(a.rK, rK);
    if (ok && a.g) {
      double part[1] = {0.0};
      for (int k = tm.tid; k < n3; k += tm.nthreads) part[0] += a.g[k] * Team::cload(rK + k);
      ok = tm.template reduce<1>(part, 0u);
      if (ok && tm.leader()) a.out[0] = tm.red[0];
      // ---- reverse sweep, segment by segment
      int cur = 0;
      copy(a.lam, a.g);
      ok = ok && tm.sync();
      const int nseg = (a.K + C - 1) / C;
      for (int sgi = nseg - 1; sgi >= 0 && ok; --sgi) {
        const int k0 = sgi * C, k1 = (k0 + C < a.K) ? k0 + C : a.K;
        // recompute r_k0 .. r_{k1-1} into seg[0 .. k1-k0)
        copy(a.seg, a.ckpt + (size_t)sgi * n3);
        ok = tm.sync();
        for (int k = k0; k + 1 < k1 && ok; ++k)
          ok = forward_step(a.seg + (size_t)(k - k0) * n3, a.seg + (size_t)(k - k0 + 1) * n3, k);
        for (int k = k1 - 1; k >= k0 && ok; --k) {
          const double* rk = a.seg + (size_t)(k - k0) * n3;
          const double* lam = a.lam + (size_t)cur * n3;
          double* lamN = a.lam + (size_t)(1 - cur) * n3;
          // op A: u_k = W r_k + b, keeping (H^T r_k); local adjoint at contact level
          {
            auto ld = [&](int j, double xc[3]) { ld3t<Team>(rk, j, xc); };
            auto body = [&](int kb, double v[6]) {
#pragma unroll
              for (int r = 0; r < 6; ++r) Team::st(&a.hr[6 * kb + r], v[r]);
              bodyDefault(kb, v);
            };
            auto pre = [&](int i) {
              RevRow w;
              ld3t<Team>(rk, i, w.r);
              ld3t<Team>(lam, i, w.lam);
              w.b[0] = P.b[3 * i];
              w.b[1] = P.b[3 * i + 1];
              w.b[2] = P.b[3 * i + 2];
              w.mu = P.mu[i];
              return w;
            };
            auto lead = [&](int i, double v[3], const RevRow& w) {
              const double u[3] = {v[0] + w.b[0], v[1] + w.b[1], v[2] + w.b[2]};
              double z[3], d[2], ut, J[9], dmu[3], mz[3], muu[3];
              if (useDir && a.branch_mode == 2) {
                const double* sd = a.sdir + 2 * ((size_t)k * nc + i);
                d[0] = sd[0];
                d[1] = sd[1];
                ut = d[0] * u[1] + d[1] * u[2];
              } else {
                tangent_dir(u, d, ut);
              }
              step_point(coulomb, w.mu, rho, w.r, u, ut, z);
              project_jacobian<double>(w.mu, z, J, dmu,
                                       a.branch_mode == 2 ? (int)a.branch[(size_t)k * nc + i] : -1);
              adjoint_point(coulomb, w.mu, rho, J, d, w.lam, mz, muu);
              st3t<Team>(a.mz, i, mz);
              st3t<Team>(a.mu_u, i, muu);
              // a row has one leader thread in every step (static tile schedule), and
              // team barriers separate the steps: every old value is loaded first (no
              // store between the loads, so they are in flight together), then stored
              double gb[3];
#pragma unroll
              for (int c = 0; c < 3; ++c) gb[c] = Team::ld(a.grad_b + 3 * i + c);
              const double gmu = Team::ld(a.grad_mu + i);
#pragma unroll
              for (int c = 0; c < 3; ++c) Team::st(a.grad_b + 3 * i + c, gb[c] + muu[c]);
              Team::st(a.grad_mu + i, gmu + w.lam[0] * dmu[0] + w.lam[1] * dmu[1] + w.lam[2] * dmu[2] -
                                          (coulomb ? rho * ut * mz[0] : 0.0));
            };
            ok = apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, pre, lead, body) && tm.sync();
          }
          // op B: lambda_k = mu_z + W mu_u; dL/dMinv += (H^T mu_u)(H^T r_k)^T
          if (ok) {
            auto ld = [&](int j, double xc[3]) { ld3t<Team>(a.mu_u, j, xc); };
            auto body = [&](int kb, double v[6]) {
              double hr[6];
#pragma unroll
              for (int c = 0; c < 6; ++c) hr[c] = Team::ld(&a.hr[6 * kb + c]);
              double* gm = a.grad_minv + (size_t)36 * kb;
              // one leader per body row in every step, steps separated by team barriers:
              // the 36 old values are loaded together, then updated
              double g[36];
#pragma unroll
              for (int e = 0; e < 36; ++e) g[e] = Team::ld(gm + e);
#pragma unroll
              for (int r = 0; r < 6; ++r)
#pragma unroll
                for (int c = 0; c < 6; ++c) Team::st(gm + 6 * r + c, g[6 * r + c] + v[r] * hr[c]);
              bodyDefault(kb, v);
            };
            auto pre = [&](int i) {
              double m[3];
              ld3t<Team>(a.mz, i, m);
              return FwdRow{{m[0], m[1], m[2]}, {0, 0, 0}, 0.0};
            };
            auto lead = [&](int i, double v[3], const FwdRow& w) {
              const double l[3] = {w.r[0] + v[0], w.r[1] + v[1], w.r[2] + v[2]};
              st3t<Team>(lamN, i, l);
            };
            ok = apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, pre, lead, body) && tm.sync();
          }
          cur = 1 - cur;
        }
      }
      iters = a.K;
      converged = ok;
    }
  } else {
    // ---- implicit: frozen Jacobians at r* = r0
    {
      auto ld = [&](int j, double xc[3]) { ld3t<Team>(a.r0, j, xc); };
      auto body = [&](int kb, double v[6]) {
#pragma unroll
        for (int r = 0; r < 6; ++r) Team::st(&a.hr[6 * kb + r], v[r]);
        bodyDefault(kb, v);
      };
      auto pre = [&](int i) {
        FwdRow w;
        ld3t<Team>(a.r0, i, w.r);
        w.b[0] = P.b[3 * i];
        w.b[1] = P.b[3 * i + 1];
        w.b[2] = P.b[3 * i + 2];
        w.mu = P.mu[i];
        return w;
      };
      auto lead = [&](int i, double v[3], const FwdRow& w) {
        const double u[3] = {v[0] + w.b[0], v[1] + w.b[1], v[2] + w.b[2]};
        double z[3], d[2], ut, J[9], dmu[3];
        tangent_dir(u, d, ut);
        step_point(coulomb, w.mu, rho, w.r, u, ut, z);
        project_jacobian<double>(w.mu, z, J, dmu);
        double* jr = a.jac + (size_t)16 * i;
#pragma unroll
        for (int e = 0; e < 9; ++e) Team::st(jr + e, J[e]);
        Team::st(jr + 9, dmu[0]);
        Team::st(jr + 10, dmu[1]);
        Team::st(jr + 11, dmu[2]);
        Team::st(jr + 12, u[0]);
        Team::st(jr + 13, u[1]);
        Team::st(jr + 14, u[2]);
        Team::st(jr + 15, ut);
      };
      ok = apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, pre, lead, body) && tm.sync();
    }
    // mu_z, mu_u of a vector (contact-local, Jacobians at r*)
    auto local = [&](const double* lam) {
      for (int i = tm.tid; i < nc; i += tm.nthreads) {
        const double* jr = a.jac + (size_t)16 * i;
        double J[9], u[3], d[2], ut, l[3], mz[3], muu[3];
#pragma unroll
        for (int e = 0; e < 9; ++e) J[e] = Team::ld(jr + e);
        u[0] = Team::ld(jr + 12);
        u[1] = Team::ld(jr + 13);
        u[2] = Team::ld(jr + 14);
        tangent_dir(u, d, ut);
        ld3t<Team>(lam, i, l);
        adjoint_point(coulomb, P.mu[i], rho, J, d, l, mz, muu);
        st3t<Team>(a.mz, i, mz);
        st3t<Team>(a.mu_u, i, muu);
      }
    };
    // dst = (I - A^T) src = src - mu_z(src) - W mu_u(src); ends with a team barrier
    auto applySys = [&](const double* src, double* dst) -> bool {
      local(src);
      if (!tm.sync()) return false;
      auto ld = [&](int j, double xc[3]) { ld3t<Team>(a.mu_u, j, xc); };
      auto pre = [&](int i) {
        RevRow w;
        ld3t<Team>(a.mz, i, w.r);
        ld3t<Team>(src, i, w.lam);
        return w;
      };
      auto lead = [&](int i, double v[3], const RevRow& w) {
        const double o[3] = {w.lam[0] - w.r[0] - v[0], w.lam[1] - w.r[1] - v[1], w.lam[2] - w.r[2] - v[2]};
        st3t<Team>(dst, i, o);
      };
      return apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, pre, lead, bodyDefault) && tm.sync();
    };
    const int m = a.kdim;
    double* V = a.kry;
    auto Vk = [&](int j) { return V + (size_t)j * n3; };
    double* lamv = a.lam;       // the solution
    double* tmp = a.lam + n3;   // (I - A^T) lam
    // Hessenberg matrix, rotations and rotated right-hand side: identical in every block
    __shared__ double sH[(kMaxKrylov + 1) * kMaxKrylov];
    __shared__ double sCs[kMaxKrylov], sSn[kMaxKrylov], sG[kMaxKrylov + 1], sY[kMaxKrylov], sHc[kMaxKrylov + 1];
    __shared__ double h1[kMaxKrylov + 1], h2[kMaxKrylov + 1];  // the two Gram-Schmidt passes' coefficients
    // h[0..cnt) = V_i . w for i < cnt (h in LDS), reduced in groups of kMaxPartials
    auto dots = [&](int cnt, const double* w, double* h) -> bool {
      for (int c0 = 0; c0 < cnt; c0 += kMaxPartials) {
        const int nq = (cnt - c0 < kMaxPartials) ? cnt - c0 : kMaxPartials;
        double part[kMaxPartials];
#pragma unroll
        for (int q = 0; q < kMaxPartials; ++q) part[q] = 0.0;
        for (int k = tm.tid; k < n3; k += tm.nthreads) {
          const double wk = Team::ld(w + k);
#pragma unroll
          for (int q = 0; q < kMaxPartials; ++q)
            if (q < nq) part[q] += Team::ld(Vk(c0 + q) + k) * wk;
        }
        if (!tm.template reduce<kMaxPartials>(part, 0u)) return false;
        if (threadIdx.x == 0)
          for (int q = 0; q < nq; ++q) h[c0 + q] = tm.red[q];
        __syncthreads();
      }
      return true;
    };
    double gn2[1] = {0.0};
    for (int k = tm.tid; k < n3; k += tm.nthreads) gn2[0] += a.g[k] * a.g[k];
    ok = ok && tm.template reduce<1>(gn2, 0u);
    const double gnorm = ok ? sqrt(tm.red[0]) : 0.0;
    for (int k = tm.tid; k < n3; k += tm.nthreads) Team::st(lamv + k, 0.0);
    ok = ok && tm.sync();
    bool first = true;
    double relres = 1.0;
    while (ok && gnorm > 0.0) {
      // true residual r = g - (I - A^T) lam into V_0 and its norm
      if (!first && !(ok = applySys(lamv, tmp))) break;
      double rn2[1] = {0.0};
      for (int k = tm.tid; k < n3; k += tm.nthreads) {
        const double rk = a.g[k] - (first ? 0.0 : Team::ld(tmp + k));
        Team::st(Vk(0) + k, rk);
        rn2[0] += rk * rk;
      }
      first = false;
      if (!(ok = tm.template reduce<1>(rn2, 0u))) break;
      const double beta = sqrt(tm.red[0]);
      relres = beta / gnorm;
      change = relres;
      if (relres < a.tol) {
        converged = true;
        break;
      }
      if (iters >= a.max_iters) break;
      for (int k = tm.tid; k < n3; k += tm.nthreads) Team::st(Vk(0) + k, Team::ld(Vk(0) + k) / beta);
      if (threadIdx.x == 0) {
        sG[0] = beta;
        for (int q = 1; q <= m; ++q) sG[q] = 0.0;
      }
      if (!(ok = tm.sync())) break;
      int j = 0;
      for (; j < m && iters < a.max_iters;) {
        double* w = Vk(j + 1);
        if (!(ok = applySys(Vk(j), w))) break;
        // classical Gram-Schmidt, twice
        if (!(ok = dots(j + 1, w, h1))) break;
        for (int k = tm.tid; k < n3; k += tm.nthreads) {
          double acc = Team::ld(w + k);
          for (int q = 0; q <= j; ++q) acc -= h1[q] * Team::ld(Vk(q) + k);
          Team::st(w + k, acc);
        }
        if (!(ok = dots(j + 1, w, h2))) break;
        double wn2[1] = {0.0};
        for (int k = tm.tid; k < n3; k += tm.nthreads) {
          double acc = Team::ld(w + k);
          for (int q = 0; q <= j; ++q) acc -= h2[q] * Team::ld(Vk(q) + k);
          Team::st(w + k, acc);
          wn2[0] += acc * acc;
        }
        if (!(ok = tm.template reduce<1>(wn2, 0u))) break;
        const double hn = sqrt(tm.red[0]);
        if (hn > 0.0)
          for (int k = tm.tid; k < n3; k += tm.nthreads) Team::st(w + k, Team::ld(w + k) / hn);
        if (threadIdx.x == 0) {
          // column j: previous rotations, then the rotation that zeroes H[j+1][j]
          for (int q = 0; q <= j; ++q) sHc[q] = h1[q] + h2[q];
          sHc[j + 1] = hn;
          for (int q = 0; q < j; ++q) {
            const double t0 = sCs[q] * sHc[q] + sSn[q] * sHc[q + 1];
            sHc[q + 1] = -sSn[q] * sHc[q] + sCs[q] * sHc[q + 1];
            sHc[q] = t0;
          }
          const double den = sqrt(sHc[j] * sHc[j] + sHc[j + 1] * sHc[j + 1]);
          const double c = den > 0.0 ? sHc[j] / den : 1.0, sn = den > 0.0 ? sHc[j + 1] / den : 0.0;
          sCs[j] = c;
          sSn[j] = sn;
          sHc[j] = den;
          for (int q = 0; q <= j; ++q) sH[q * kMaxKrylov + j] = sHc[q];
          sG[j + 1] = -sn * sG[j];
          sG[j] = c * sG[j];
        }
        if (!(ok = tm.sync())) break;
        ++iters;
        ++j;
        relres = fabs(sG[j]) / gnorm;
        if (relres < a.tol || !(hn > 0.0)) break;
      }
      if (!ok) break;
      // y = R^{-1} g (upper triangular j x j), lam += sum_q y_q V_q
      if (threadIdx.x == 0) {
        for (int q = j - 1; q >= 0; --q) {
          double acc = sG[q];
          for (int c2 = q + 1; c2 < j; ++c2) acc -= sH[q * kMaxKrylov + c2] * sY[c2];
          const double dq = sH[q * kMaxKrylov + q];
          sY[q] = dq != 0.0 ? acc / dq : 0.0;
        }
      }
      __syncthreads();
      for (int k = tm.tid; k < n3; k += tm.nthreads) {
        double acc = Team::ld(lamv + k);
        for (int q = 0; q < j; ++q) acc += sY[q] * Team::ld(Vk(q) + k);
        Team::st(lamv + k, acc);
      }
      if (!(ok = tm.sync())) break;
    }
    // gradients from lambda*
    if (ok) {
      const double* lam = lamv;
      local(lam);
      for (int i = tm.tid; i < nc; i += tm.nthreads) {
        const double* jr = a.jac + (size_t)16 * i;
        double l[3], mz[3], muu[3];
        ld3t<Team>(lam, i, l);
        ld3t<Team>(a.mz, i, mz);
        ld3t<Team>(a.mu_u, i, muu);
        Team::st(a.grad_b + 3 * i, muu[0]);
        Team::st(a.grad_b + 3 * i + 1, muu[1]);
        Team::st(a.grad_b + 3 * i + 2, muu[2]);
        Team::st(a.grad_mu + i, l[0] * Team::ld(jr + 9) + l[1] * Team::ld(jr + 10) + l[2] * Team::ld(jr + 11) -
                                    (coulomb ? rho * Team::ld(jr + 15) * mz[0] : 0.0));
      }
      ok = tm.sync();
      if (ok) {
        auto ld = [&](int j, double xc[3]) { ld3t<Team>(a.mu_u, j, xc); };
        auto body = [&](int kb, double v[6]) {
          double hr[6];
#pragma unroll
          for (int c = 0; c < 6; ++c) hr[c] = Team::ld(&a.hr[6 * kb + c]);
#pragma unroll
          for (int r = 0; r < 6; ++r)
#pragma unroll
            for (int c = 0; c < 6; ++c) Team::st(&a.grad_minv[(size_t)36 * kb + 6 * r + c], v[r] * hr[c]);
          bodyDefault(kb, v);
        };
        NoPre nopre;
        auto lead = [&](int, double*, NoPre::type) {};
        ok = apply_operator<true, double>(P, a.wbody, a.carry, tm, ld, nopre, lead, body);
      }
      // the loss at r*
      double part[1] = {0.0};
      for (int k = tm.tid; k < n3; k += tm.nthreads) part[0] += a.g[k] * a.r0[k];
      ok = ok && tm.template reduce<1>(part, 0u);
      if (ok && tm.leader()) a.out[0] = tm.red[0];
    }
  }
  if (tm.leader()) {
    a.out[1] = (double)iters;
    a.out[2] = change;
    a.out[3] = converged ? 1.0 : 0.0;
    a.out[4] = ok ? 0.0 : 1.0;
  }
}

// Plain vector stores, each XCD's L2 written back once per barrier (as the
// single-problem solvers).
__global__ void __launch_bounds__(512) adjoint_kernel(AdjArgs a) {
  GridTeamWB tm(a.gs, blockIdx.x);
  if (!tm.group_by_xcd()) return;
  adjoint_body<GridTeamWB>(a, tm);
}

}  // namespace sb
