// One-contact analytic/Newton solver in FP64, per lane (device).
//
// LocalSOCSolver<3,double,DeSaxceCOV,Strategy>::solve
//   (src/Extra/SOC/LocalSOCSolver.impl.hpp:243-318) with
//   * AnalyticLocalSOCSolver::solveOrthogonality for Coulomb (:55-102, [DBB11] App. B.1)
//     and SOC (:149-209);
//   * the real roots of the quartic from the eigenvalues of its 4x4 companion
//     matrix (src/Core/Utils/Polynomial.impl.hpp:56-109), computed here by an
//     implicit double-shift QR iteration on the (already Hessenberg) companion;
//   * NonSmoothNewton::solve (src/Core/Utils/NonSmoothNewton.impl.hpp:23-79) on
//     F(x) = FB_mu(s x, A x + b [+ COV]) with its Jacobian
//     (src/Extra/SOC/FischerBurmeister.impl.hpp:44-67, 69-140, 161-187);
//   * complete-pivoting LU solves for the 3x3 and 2x2 systems.
#pragma once
#include <hip/hip_runtime.h>

#include "sb_law.hpp"

namespace sb {
namespace local {

constexpr double kEps = 2.220446049250313e-16;

enum Strategy { PureNewton = 0, PureEnumerative = 1, Hybrid = 2, RevHybrid = 3 };

__device__ __forceinline__ bool isz(double s) { return fabs(s) < kEps; }

// ---------------------------------------------------------------- LU (complete pivoting)
// Solves A x = b; unknowns beyond the numerical rank (pivots <= N*eps*max pivot) are 0.
// Pivot rows/columns are runtime values, but every array access uses a compile-time
// index (row/column swaps select the matching entry), so a, c and the column
// permutation stay in registers.
template <int N>
__device__ inline void lu_solve(const double Ain[N][N], const double bin[N], double x[N]) {
  double a[N][N], c[N];
  int cperm[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    c[i] = bin[i];
    cperm[i] = i;
#pragma unroll
    for (int j = 0; j < N; ++j) a[i][j] = Ain[i][j];
  }
  int npiv = N;
  double maxpiv = 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (k >= npiv) break;
    double big = -1.0;
    int rb = k, cb = k;
#pragma unroll
    for (int j = k; j < N; ++j)
#pragma unroll
      for (int i = k; i < N; ++i) {
        const double v = fabs(a[i][j]);
        if (v > big) {
          big = v;
          rb = i;
          cb = j;
        }
      }
    if (big == 0.0) {
      npiv = k;
      break;
    }
    maxpiv = fmax(maxpiv, big);
#pragma unroll
    for (int r = k + 1; r < N; ++r)
      if (r == rb) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
          const double t = a[k][j];
          a[k][j] = a[r][j];
          a[r][j] = t;
        }
        const double t = c[k];
        c[k] = c[r];
        c[r] = t;
      }
#pragma unroll
    for (int q = k + 1; q < N; ++q)
      if (q == cb) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
          const double t = a[i][k];
          a[i][k] = a[i][q];
          a[i][q] = t;
        }
        const int t = cperm[k];
        cperm[k] = cperm[q];
        cperm[q] = t;
      }
#pragma unroll
    for (int i = k + 1; i < N; ++i) {
      a[i][k] /= a[k][k];
#pragma unroll
      for (int j = k + 1; j < N; ++j) a[i][j] -= a[i][k] * a[k][j];
    }
  }
  const double thr = maxpiv * (N * kEps);
  int rank = 0;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (i < npiv) rank += (fabs(a[i][i]) > thr) ? 1 : 0;
  // forward substitution (unit lower), full size
#pragma unroll
  for (int i = 1; i < N; ++i)
#pragma unroll
    for (int j = 0; j < i; ++j) c[i] -= a[i][j] * c[j];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    if (i < rank) {
#pragma unroll
      for (int j = i + 1; j < N; ++j)
        if (j < rank) c[i] -= a[i][j] * c[j];
      c[i] /= a[i][i];
    }
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double v = (i < rank) ? c[i] : 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p)
      if (cperm[i] == p) x[p] = v;
  }
}

// ---------------------------------------------------------------- quartic roots
__device__ inline void eig2(double a, double b, double c, double d, double& r1, double& i1, double& r2, double& i2) {
  const double p = 0.5 * (a - d);
  const double q = p * p + b * c;
  if (q >= 0) {
    const double z = p + copysign(sqrt(q), p);
    r1 = d + z;
    r2 = (z != 0.0) ? d - b * c / z : d + z;
    i1 = i2 = 0.0;
  } else {
    r1 = r2 = d + p;
    i1 = sqrt(-q);
    i2 = -i1;
  }
}

// One implicit double-shift (Francis) step on the active window [LO, HI] of the
// Hessenberg matrix h, shifts s (sum) and t (product). Every index is a
// compile-time constant, so h stays in registers (a runtime-indexed 4x4 array
// would live in scratch memory and put a memory round trip on every access).
template <int LO, int HI>
__device__ __forceinline__ void francis_step(double (&h)[4][4], double s, double t) {
  double x = h[LO][LO] * h[LO][LO] + h[LO][LO + 1] * h[LO + 1][LO] - s * h[LO][LO] + t;
  double y = h[LO + 1][LO] * (h[LO][LO] + h[LO + 1][LO + 1] - s);
  double z = h[LO + 1][LO] * h[LO + 2][LO + 1];
#pragma unroll
  for (int k = LO; k <= HI - 2; ++k) {
    const double alpha = -copysign(sqrt(x * x + y * y + z * z), x);
    const double v0 = x - alpha, v1 = y, v2 = z;
    const double vv = v0 * v0 + v1 * v1 + v2 * v2;
    if (vv != 0.0) {
      const double beta = 2.0 / vv;
      const int q = (k - 1 > LO) ? k - 1 : LO;
#pragma unroll
      for (int j = q; j <= HI; ++j) {
        const double d = beta * (v0 * h[k][j] + v1 * h[k + 1][j] + v2 * h[k + 2][j]);
        h[k][j] -= d * v0;
        h[k + 1][j] -= d * v1;
        h[k + 2][j] -= d * v2;
      }
      const int r = (k + 3 < HI) ? k + 3 : HI;
#pragma unroll
      for (int i = LO; i <= r; ++i) {
        const double d = beta * (h[i][k] * v0 + h[i][k + 1] * v1 + h[i][k + 2] * v2);
        h[i][k] -= d * v0;
        h[i][k + 1] -= d * v1;
        h[i][k + 2] -= d * v2;
      }
    }
    x = h[k + 1][k];
    y = h[k + 2][k];
    if (k < HI - 2) z = h[k + 3][k];
  }
  {
    const double alpha = -copysign(sqrt(x * x + y * y), x);
    const double v0 = x - alpha, v1 = y;
    const double vv = v0 * v0 + v1 * v1;
    if (vv != 0.0) {
      const double beta = 2.0 / vv;
      constexpr int k = HI - 1;
#pragma unroll
      for (int j = (k - 1 > LO ? k - 1 : LO); j <= HI; ++j) {
        const double d = beta * (v0 * h[k][j] + v1 * h[k + 1][j]);
        h[k][j] -= d * v0;
        h[k + 1][j] -= d * v1;
      }
#pragma unroll
      for (int i = LO; i <= HI; ++i) {
        const double d = beta * (h[i][k] * v0 + h[i][k + 1] * v1);
        h[i][k] -= d * v0;
        h[i][k + 1] -= d * v1;
      }
    }
  }
#pragma unroll
  for (int i = LO + 2; i <= HI; ++i)
#pragma unroll
    for (int j = LO; j < i - 1; ++j) h[i][j] = 0.0;
}

// h[i][j] for runtime i, j in 0..3 by selection over the (register) entries
__device__ __forceinline__ double hsel(const double (&h)[4][4], int i, int j) {
  double v = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) v = (a == i && b == j) ? h[a][b] : v;
  return v;
}

// Eigenvalues of an n x n (n <= 4) upper-Hessenberg matrix h (modified in place):
// implicit double-shift QR with deflation (Golub & Van Loan, Alg. 7.5.1/7.5.2),
// after scaling h by its largest |entry|; ad-hoc shifts after 10 and 20 stagnant
// iterations; deflated 2x2 blocks by eig2. Runtime indices only select among
// registers (hsel, the switch over the active window).
__device__ inline void hess_eig(int n, double (&h)[4][4], double (&wr)[4], double (&wi)[4]) {
  double scale = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i < n && j < n) scale = fmax(scale, fabs(h[i][j]));
  if (scale == 0.0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (i < n) wr[i] = wi[i] = 0.0;
    return;
  }
  const double inv = 1.0 / scale;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (i < n && j < n) h[i][j] *= inv;
  auto setw = [&](int k, double r, double im) {
#pragma unroll
    for (int a = 0; a < 4; ++a)
      if (a == k) {
        wr[a] = r;
        wi[a] = im;
      }
  };
  int hi = n - 1, iter = 0, total = 0;
  const int maxTotal = 40 * n;
  while (hi >= 0) {
    // deflation: the largest lo <= hi whose subdiagonal entry is negligible (else 0)
    int lo = 0;
#pragma unroll
    for (int l = 3; l >= 1; --l) {
      if (l <= hi && lo == 0) {
        const double s = fabs(h[l - 1][l - 1]) + fabs(h[l][l]);
        if (fabs(h[l][l - 1]) <= kEps * (s == 0.0 ? 1.0 : s)) {
          h[l][l - 1] = 0.0;
          lo = l;
        }
      }
    }
    // the search above stops at the first negligible entry from hi downwards
    if (lo == hi) {
      setw(hi, hsel(h, hi, hi) * scale, 0.0);
      --hi;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {
      double r1, i1, r2, i2;
      eig2(hsel(h, hi - 1, hi - 1), hsel(h, hi - 1, hi), hsel(h, hi, hi - 1), hsel(h, hi, hi), r1, i1, r2, i2);
      setw(hi - 1, r1 * scale, i1 * scale);
      setw(hi, r2 * scale, i2 * scale);
      hi -= 2;
      iter = 0;
      continue;
    }
    if (total >= maxTotal) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (k >= lo && k <= hi) {
          wr[k] = h[k][k] * scale;
          wi[k] = 0.0;
        }
      hi = lo - 1;
      continue;
    }
    ++iter;
    ++total;
    double s, t;
    const double hmm = hsel(h, hi - 1, hi - 1), hhh = hsel(h, hi, hi);
    if (iter == 10 || iter == 20) {
      const double e = fabs(hsel(h, hi, hi - 1)) + fabs(hsel(h, hi - 1, hi - 2));
      s = 1.5 * e;
      t = e * e;
    } else {
      s = hmm + hhh;
      t = hmm * hhh - hsel(h, hi - 1, hi) * hsel(h, hi, hi - 1);
    }
    if (lo == 0 && hi == 2)
      francis_step<0, 2>(h, s, t);
    else if (lo == 0 && hi == 3)
      francis_step<0, 3>(h, s, t);
    else
      francis_step<1, 3>(h, s, t);
  }
}

enum RootFilter { StrictlyPositive = 0, AllRootsF = 2 };

// Real roots (|Im| < eps) of the monic x^4 + c3 x^3 + c2 x^2 + c1 x + c0.
// Appends v to roots[0..n) (runtime n, register-resident array).
__device__ __forceinline__ void push_root(double (&roots)[4], int& n, double v) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if (a == n) roots[a] = v;
  ++n;
}

__device__ inline int quartic_real_roots(const double c[4], double (&roots)[4], RootFilter filter) {
  double h[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) h[i][j] = 0.0;
#pragma unroll
  for (int i = 1; i < 4; ++i) h[i][i - 1] = 1.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) h[i][3] = -c[i];
  double wr[4], wi[4];
  hess_eig(4, h, wr, wi);
  int n = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (isz(wi[i]) && !(filter == StrictlyPositive && wr[i] <= 0.0)) push_root(roots, n, wr[i]);
  return n;
}

// Real roots of c[D] x^D + ... + c[0] with a possibly vanishing leading coefficient.
__device__ inline int poly_real_roots_degenerate(double c[5], double (&roots)[4]) {
  int D = 4;
  while (D > 0 && isz(c[D])) --D;
  if (D == 0) {
    roots[0] = 0.0;
    return isz(c[0]) ? 1 : 0;
  }
  double cD = c[4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
    if (a == D) cD = c[a];
  const double inv = 1.0 / cD;
  double h[4][4], wr[4], wi[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) h[i][j] = (i < D && j < D) ? ((j == D - 1) ? -(c[i] * inv) : (i == j + 1 ? 1.0 : 0.0)) : 0.0;
  hess_eig(D, h, wr, wi);
  int n = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < D && isz(wi[i])) push_root(roots, n, wr[i]);
  return n;
}

// ---------------------------------------------------------------- FB Jacobian
// FBBaseFunction<3>::computeJacobian on the cone of aperture mu.
__device__ inline void fb_jacobian(double mu, const double x[3], const double y[3], double fb[3], double dx[3][3],
                                   double dy[3][3]) {
  if (isz(mu)) {
    const double z = sqrt(x[0] * x[0] + y[0] * y[0]);
    fb[0] = x[0] + y[0] - z;
    fb[1] = x[1];
    fb[2] = x[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dx[i][j] = (i == j) ? 1.0 : 0.0;
        dy[i][j] = 0.0;
      }
    if (!isz(z)) {
      dx[0][0] = 1.0 - x[0] / z;
      dy[0][0] = 1.0 - y[0] / z;
    }
    return;
  }
  const double xh[3] = {mu * x[0], x[1], x[2]};
  const double yh[3] = {y[0], mu * y[1], mu * y[2]};
  const double z20 = xh[0] * xh[0] + xh[1] * xh[1] + xh[2] * xh[2] + yh[0] * yh[0] + yh[1] * yh[1] + yh[2] * yh[2];
  const double z21 = xh[0] * xh[1] + yh[0] * yh[1];
  const double z22 = xh[0] * xh[2] + yh[0] * yh[2];
  const double nz2t = sqrt(z21 * z21 + z22 * z22);
  double d1, d2;
  if (isz(nz2t)) {
    d1 = 1.0;
    d2 = 0.0;
  } else {
    d1 = z21 / nz2t;
    d2 = z22 / nz2t;
  }
  const double rl1 = sqrt(fmax(0.0, z20 - 2.0 * nz2t));
  const double rl2 = sqrt(z20 + 2.0 * nz2t);
  const double z0 = 0.5 * (rl1 + rl2), zt = 0.5 * (rl2 - rl1);
  const double z[3] = {z0, zt * d1, zt * d2};
  for (int k = 0; k < 3; ++k) fb[k] = xh[k] + yh[k] - z[k];
  if (isz(rl2)) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dx[i][j] = dy[i][j] = 0.0;
  } else if (isz(rl1)) {
    const double izn = 1.0 / (xh[0] * xh[0] + yh[0] * yh[0]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dx[i][j] = (i == j) ? (1.0 - xh[0] * izn) : 0.0;
        dy[i][j] = (i == j) ? (1.0 - yh[0] * izn) : 0.0;
      }
  } else {
    const double det = rl1 * rl2;
    const double id = 1.0 / det;
    // inverse of the arrow matrix L_z, scaled: iL = [[z0, -zt^T], [-zt, (det I + zt zt^T)/z0]] / det
    const double iL[3][3] = {{z[0] * id, -z[1] * id, -z[2] * id},
                             {-z[1] * id, (det + z[1] * z[1]) / z[0] * id, (z[1] * z[2]) / z[0] * id},
                             {-z[2] * id, (z[2] * z[1]) / z[0] * id, (det + z[2] * z[2]) / z[0] * id}};
    const double Lx[3][3] = {{xh[0], xh[1], xh[2]}, {xh[1], xh[0], 0.0}, {xh[2], 0.0, xh[0]}};
    const double Ly[3][3] = {{yh[0], yh[1], yh[2]}, {yh[1], yh[0], 0.0}, {yh[2], 0.0, yh[0]}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double sx = 0.0, sy = 0.0;
        for (int k = 0; k < 3; ++k) {
          sx += iL[i][k] * Lx[k][j];
          sy += iL[i][k] * Ly[k][j];
        }
        dx[i][j] = ((i == j) ? 1.0 : 0.0) - sx;
        dy[i][j] = ((i == j) ? 1.0 : 0.0) - sy;
      }
  }
  for (int i = 0; i < 3; ++i) {  // chain rule of the mu-scalings
    dx[i][0] *= mu;
    dy[i][1] *= mu;
    dy[i][2] *= mu;
  }
}

// F(x) = FB(mu, s x, A x + b + COV) and, optionally, dF/dx.
struct LocalFn {
  bool coulomb;
  double mu, scaling;
  double A[3][3], b[3];

  __device__ double phi(const double x[3]) const {
    double y[3], xs[3];
    for (int i = 0; i < 3; ++i) {
      y[i] = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i];
      xs[i] = scaling * x[i];
    }
    return 0.5 * law_eval<double>(coulomb, mu, xs, y);
  }
  __device__ void jacobian(const double x[3], double F[3], double dF[3][3]) const {
    double xs[3], y[3];
    for (int i = 0; i < 3; ++i) {
      xs[i] = scaling * x[i];
      y[i] = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i];
    }
    double s = 0.0;
    if (coulomb) {
      s = sqrt(y[1] * y[1] + y[2] * y[2]);
      y[0] += mu * s;
    }
    double dy[3][3];
    fb_jacobian(mu, xs, y, F, dF, dy);
    if (coulomb && !isz(s)) {
      for (int i = 0; i < 3; ++i) {
        const double c = dy[i][0] * (mu / s);
        dy[i][1] += c * y[1];
        dy[i][2] += c * y[2];
      }
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double acc = dF[i][j] * scaling;
        for (int k = 0; k < 3; ++k) acc += dy[i][k] * A[k][j];
        dF[i][j] = acc;
      }
  }
};

// Damped non-smooth Newton on Phi = |F|^2 / 2; the descent test uses dF^T x as
// in the reference. Returns the best Phi; x holds the best iterate.
__device__ inline double newton(const LocalFn& f, double tol, int maxIters, double x[3]) {
  const double sigma2 = 1.e-4, damp = 0.5;
  const double phiInit = f.phi(x);
  if (phiInit < tol) return phiInit;
  const double zero[3] = {0.0, 0.0, 0.0};
  const double phiZero = f.phi(zero);
  double best[3], phiBest;
  if (phiZero < phiInit) {
    phiBest = phiZero;
    x[0] = x[1] = x[2] = 0.0;
    best[0] = best[1] = best[2] = 0.0;
  } else {
    phiBest = phiInit;
    best[0] = x[0];
    best[1] = x[1];
    best[2] = x[2];
  }
  if (phiZero < tol) return phiZero;
  for (int it = 0; it < maxIters; ++it) {
    double F[3], dF[3][3];
    f.jacobian(x, F, dF);
    const double phi = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    if (phi < tol) return phi;
    if (phi < phiBest) {
      phiBest = phi;
      best[0] = x[0];
      best[1] = x[1];
      best[2] = x[2];
    }
    double dphi[3];
    for (int j = 0; j < 3; ++j) dphi[j] = dF[0][j] * x[0] + dF[1][j] * x[1] + dF[2][j] * x[2];
    const double mF[3] = {-F[0], -F[1], -F[2]};
    double dx[3];
    lu_solve<3>(dF, mF, dx);
    const double proj = dx[0] * dphi[0] + dx[1] * dphi[1] + dx[2] * dphi[2];
    const double ndx2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double ndp2 = dphi[0] * dphi[0] + dphi[1] * dphi[1] + dphi[2] * dphi[2];
    const double step = (proj > 0 || proj * proj < sigma2 * ndx2 * ndp2) ? damp : 1.0;
    for (int i = 0; i < 3; ++i) x[i] += step * dx[i];
  }
  x[0] = best[0];
  x[1] = best[1];
  x[2] = best[2];
  return phiBest;
}

// Analytic sliding solution (Coulomb) / cone-boundary solution (SOC).
__device__ inline bool solve_orthogonality(bool coulomb, const double W[3][3], const double b[3], double r[3],
                                           double mu) {
  if (coulomb) {
    const double wN = W[0][0];
    if (wN < kEps) return false;
    const double wT0 = W[1][0], wT1 = W[2][0];
    const double bN = b[0];
    const double q0 = wT0 / wN, q1 = wT1 / wN;
    const double W00 = W[1][1] - q0 * wT0, W01 = W[1][2] - q0 * wT1;
    const double W10 = W[2][1] - q1 * wT0, W11 = W[2][2] - q1 * wT1;
    const double bb0 = b[1] / bN - q0, bb1 = b[2] / bN - q1;
    const double A = (W00 + W11) - (wT0 * bb0 + wT1 * bb1);
    const double B0 = W11 * bb0 - W10 * bb1, B1 = W00 * bb1 - W01 * bb0;
    const double C = (W00 * W11 - W01 * W10) - (wT0 * B0 + wT1 * B1);
    const double D = wN * wN / (mu * mu);
    const double c[4] = {C * C - D * (B0 * B0 + B1 * B1), 2.0 * (C * A - D * (bb0 * B0 + bb1 * B1)),
                         2.0 * C + A * A - D * (bb0 * bb0 + bb1 * bb1), 2.0 * A};
    double roots[4];
    const int n = quartic_real_roots(c, roots, StrictlyPositive);
    if (n == 0) return false;
    double alpha = roots[0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
      if (i < n) alpha = fmin(alpha, roots[i]);
    const double M[2][2] = {{W00 + alpha, W01}, {W10, W11 + alpha}};
    const double rhs[2] = {-bN * bb0, -bN * bb1};
    double rt[2];
    lu_solve<2>(M, rhs, rt);
    r[1] = rt[0];
    r[2] = rt[1];
    r[0] = sqrt(rt[0] * rt[0] + rt[1] * rt[1]) / mu;
    return true;
  }
  const double A = (b[1] * W[1][2] - b[2] * W[1][1]) * mu * mu + b[0] * W[0][2] - b[2] * W[0][0];
  const double B = (b[0] * W[1][2] + b[1] * W[0][2] - 2.0 * b[2] * W[0][1]) * mu;
  const double C = 2.0 * ((b[1] * W[2][2] - b[2] * W[1][2]) * mu * mu - b[0] * W[0][1] + b[1] * W[0][0]);
  const double D = 2.0 * (b[0] * (W[1][1] - W[2][2]) - b[1] * W[0][1] + b[2] * W[0][2]) * mu;
  const double E = -6.0 * (b[0] * W[1][2] - b[1] * W[0][2]) * mu;
  double c[5] = {A + B, C - D, E, C + D, B - A};
  double roots[4];
  const int n = poly_real_roots_degenerate(c, roots);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (i >= n) break;
    const double t = roots[i];
    const double CT = (1.0 - t * t) / (1.0 + t * t);
    const double ST = 2.0 * t / (1.0 + t * t);
    const double dir[3] = {1.0, mu * CT, mu * ST};
    double den0 = 0.0, den1 = 0.0;
    for (int k = 0; k < 3; ++k) {
      den0 += (mu * W[k][1] + CT * W[k][0]) * dir[k];
      den1 += (mu * W[k][2] + ST * W[k][0]) * dir[k];
    }
    double rN;
    if (fabs(den0) > fabs(den1)) {
      if (isz(den0)) continue;
      rN = -(CT * b[0] + b[1] * mu) / den0;
    } else {
      if (isz(den1)) continue;
      rN = -(ST * b[0] + b[2] * mu) / den1;
    }
    if (rN <= 0.0) continue;
    for (int k = 0; k < 3; ++k) r[k] = rN * dir[k];
    const double uN = W[0][0] * r[0] + W[1][0] * r[1] + W[2][0] * r[2] + b[0];
    if (uN > 0.0) return true;
  }
  return false;
}

// LocalSOCSolver::solve. Returns the local residual; x is warm start and result.
__device__ inline double solve(bool coulomb, int strat, const double A[3][3], const double b[3], double x[3],
                               double mu, double tol, double scaling) {
  if (mu < 0.0) {
    const double mb[3] = {-b[0], -b[1], -b[2]};
    lu_solve<3>(A, mb, x);
    double r2 = 0.0;
    for (int i = 0; i < 3; ++i) {
      const double v = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i];
      r2 += v * v;
    }
    return r2;
  }
  LocalFn f;
  f.coulomb = coulomb;
  f.mu = mu;
  f.scaling = scaling;
  for (int i = 0; i < 3; ++i) {
    f.b[i] = b[i];
    for (int j = 0; j < 3; ++j) f.A[i][j] = A[i][j];
  }
  const int newtonIters = 20;
  if (strat == PureNewton) return newton(f, tol, newtonIters, x);
  if (b[0] >= (coulomb ? 0.0 : mu * sqrt(b[1] * b[1] + b[2] * b[2]))) {  // take-off
    x[0] = x[1] = x[2] = 0.0;
    return 0.0;
  }
  if (isz(mu)) {  // frictionless
    if (A[0][0] < kEps) {
      x[0] = x[1] = x[2] = 0.0;
      return b[0] * b[0];
    }
    x[1] = x[2] = 0.0;
    x[0] = -b[0] / A[0][0];
    return 0.0;
  }
  double res = 0.0;
  if (strat == Hybrid) {
    res = newton(f, tol, newtonIters, x);
    if (res < tol) return res;
  }
  const double x0[3] = {x[0], x[1], x[2]};
  const double mb[3] = {-b[0], -b[1], -b[2]};
  lu_solve<3>(A, mb, x);
  if (mu * x[0] >= sqrt(x[1] * x[1] + x[2] * x[2])) return 0.0;  // sticking
  if (!solve_orthogonality(coulomb, A, b, x, mu)) {
    x[0] = x0[0];
    x[1] = x0[1];
    x[2] = x0[2];
  }
  if (strat == RevHybrid) {
    res = newton(f, tol, newtonIters, x);
  } else if (strat == Hybrid) {
    const double refined = newton(f, tol, newtonIters, x);
    if (refined <= res) return refined;
    x[0] = x0[0];
    x[1] = x0[1];
    x[2] = x0[2];
  }
  return res;
}

}  // namespace local
}  // namespace sb
