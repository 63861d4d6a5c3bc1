// cpu_ref — FP64 CPU restatement of the So-bogus Coulomb-friction solve path.
// TEST INFRASTRUCTURE ONLY (the oracle); see cpu_ref.hpp for the rules of use.
#include "cpu_ref.hpp"

#include <algorithm>
#include <array>
#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <limits>

namespace cpuref {

static const double kEps = std::numeric_limits<double>::epsilon();
// NumTraits<double>::isZero (src/Core/Utils/NumTraits.hpp:19-24): |s| < epsilon, absolute.
static inline bool isZero(double s) { return std::fabs(s) < kEps; }
static inline double sq(double a) { return a * a; }

// =============================================================== FullPivLU
// Complete pivoting: at step k the largest |entry| of the trailing corner is
// searched column by column (first maximum wins), rows and columns are swapped,
// and the corner is updated by a rank-1 elimination. The rank counts pivots
// larger than maxPivot * N * epsilon; solve() sets the unknowns past the rank to 0.
template <int N>
void FullPivLU<N>::compute(const double A[N][N]) {
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) lu[i][j] = A[i][j];
  nonzeroPivots = N;
  maxPivot = 0;
  for (int k = 0; k < N; ++k) {
    double big = -1;
    int rb = k, cb = k;
    for (int j = k; j < N; ++j)
      for (int i = k; i < N; ++i)
        if (std::fabs(lu[i][j]) > big) {
          big = std::fabs(lu[i][j]);
          rb = i;
          cb = j;
        }
    if (big == 0.0) {
      nonzeroPivots = k;
      for (int i = k; i < N; ++i) rowTransp[i] = colTransp[i] = i;
      break;
    }
    if (big > maxPivot) maxPivot = big;
    rowTransp[k] = rb;
    colTransp[k] = cb;
    if (rb != k)
      for (int j = 0; j < N; ++j) std::swap(lu[k][j], lu[rb][j]);
    if (cb != k)
      for (int i = 0; i < N; ++i) std::swap(lu[i][k], lu[i][cb]);
    if (k < N - 1) {
      for (int i = k + 1; i < N; ++i) lu[i][k] /= lu[k][k];
      for (int j = k + 1; j < N; ++j)
        for (int i = k + 1; i < N; ++i) lu[i][j] -= lu[i][k] * lu[k][j];
    }
  }
  const double thr = maxPivot * (N * kEps);
  rank = 0;
  for (int i = 0; i < nonzeroPivots; ++i) rank += (std::fabs(lu[i][i]) > thr);
}

template <int N>
void FullPivLU<N>::solve(const double b[N], double x[N]) const {
  if (rank == 0) {
    for (int i = 0; i < N; ++i) x[i] = 0;
    return;
  }
  double c[N];
  for (int i = 0; i < N; ++i) c[i] = b[i];
  for (int k = 0; k < N; ++k) std::swap(c[k], c[rowTransp[k]]);
  for (int i = 1; i < N; ++i)
    for (int j = 0; j < i; ++j) c[i] -= lu[i][j] * c[j];
  for (int i = rank - 1; i >= 0; --i) {
    for (int j = i + 1; j < rank; ++j) c[i] -= lu[i][j] * c[j];
    c[i] /= lu[i][i];
  }
  int ind[N];
  for (int i = 0; i < N; ++i) ind[i] = i;
  for (int k = 0; k < N; ++k) std::swap(ind[k], ind[colTransp[k]]);
  for (int i = 0; i < rank; ++i) x[ind[i]] = c[i];
  for (int i = rank; i < N; ++i) x[ind[i]] = 0;
}

template struct FullPivLU<2>;
template struct FullPivLU<3>;

// =============================================================== eigenvalues
// Eigenvalues of an upper-Hessenberg matrix (n <= 4) by the implicit
// double-shift (Francis) QR iteration, Golub & Van Loan, Matrix Computations,
// Alg. 7.5.1/7.5.2, with the matrix scaled by its largest |entry| first and
// ad-hoc exceptional shifts after 10 and 20 stagnant iterations. Deflated 2x2
// blocks give either two real eigenvalues (imaginary part exactly 0) or a
// complex pair. This plays the role of Eigen::EigenSolver in
// src/Core/Utils/Polynomial.impl.hpp:64.
static void eig2x2(double a, double b, double c, double d, double& r1, double& i1, double& r2, double& i2) {
  const double p = 0.5 * (a - d);
  const double q = p * p + b * c;
  if (q >= 0) {
    const double z = p + std::copysign(std::sqrt(q), p);
    r1 = d + z;
    r2 = (z != 0) ? d - b * c / z : d + z;
    i1 = i2 = 0;
  } else {
    r1 = r2 = d + p;
    i1 = std::sqrt(-q);
    i2 = -i1;
  }
}

void hessenbergEigenvalues(int n, const double* Hin, double* wr, double* wi) {
  double h[4][4];
  double scale = 0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) scale = std::max(scale, std::fabs(Hin[i * n + j]));
  if (scale == 0) {
    for (int i = 0; i < n; ++i) wr[i] = wi[i] = 0;
    return;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) h[i][j] = Hin[i * n + j] / scale;

  int hi = n - 1;  // active window is [lo, hi]
  int iter = 0, total = 0;
  const int maxTotal = 40 * n;
  while (hi >= 0) {
    // find lo: smallest index such that h[k][k-1] is negligible for k = lo
    int lo = hi;
    while (lo > 0) {
      const double s = std::fabs(h[lo - 1][lo - 1]) + std::fabs(h[lo][lo]);
      if (std::fabs(h[lo][lo - 1]) <= kEps * (s == 0 ? 1.0 : s)) {
        h[lo][lo - 1] = 0;
        break;
      }
      --lo;
    }
    if (lo == hi) {  // 1x1 block
      wr[hi] = h[hi][hi] * scale;
      wi[hi] = 0;
      --hi;
      iter = 0;
      continue;
    }
    if (lo == hi - 1) {  // 2x2 block
      double r1, i1, r2, i2;
      eig2x2(h[hi - 1][hi - 1], h[hi - 1][hi], h[hi][hi - 1], h[hi][hi], r1, i1, r2, i2);
      wr[hi - 1] = r1 * scale;
      wi[hi - 1] = i1 * scale;
      wr[hi] = r2 * scale;
      wi[hi] = i2 * scale;
      hi -= 2;
      iter = 0;
      continue;
    }
    if (total >= maxTotal) {  // no convergence: report the diagonal of the window
      for (int k = lo; k <= hi; ++k) {
        wr[k] = h[k][k] * scale;
        wi[k] = 0;
      }
      hi = lo - 1;
      continue;
    }
    ++iter;
    ++total;
    // shifts from the trailing 2x2 of the window (s = trace, t = det)
    double s, t;
    if (iter == 10 || iter == 20) {
      const double e = std::fabs(h[hi][hi - 1]) + std::fabs(h[hi - 1][hi - 2]);
      s = 1.5 * e;
      t = e * e;
    } else {
      s = h[hi - 1][hi - 1] + h[hi][hi];
      t = h[hi - 1][hi - 1] * h[hi][hi] - h[hi - 1][hi] * h[hi][hi - 1];
    }
    // first column of (H - s1 I)(H - s2 I) restricted to the window
    double x = h[lo][lo] * h[lo][lo] + h[lo][lo + 1] * h[lo + 1][lo] - s * h[lo][lo] + t;
    double y = h[lo + 1][lo] * (h[lo][lo] + h[lo + 1][lo + 1] - s);
    double z = h[lo + 1][lo] * h[lo + 2][lo + 1];
    for (int k = lo; k <= hi - 2; ++k) {
      // Householder reflector P = I - beta v v^T mapping (x,y,z) to a multiple of e1
      const double alpha = -std::copysign(std::sqrt(x * x + y * y + z * z), x);
      double v0 = x - alpha, v1 = y, v2 = z;
      const double vv = v0 * v0 + v1 * v1 + v2 * v2;
      if (vv != 0) {
        const double beta = 2.0 / vv;
        const int q = std::max(lo, k - 1);
        for (int j = q; j <= hi; ++j) {
          const double d = beta * (v0 * h[k][j] + v1 * h[k + 1][j] + v2 * h[k + 2][j]);
          h[k][j] -= d * v0;
          h[k + 1][j] -= d * v1;
          h[k + 2][j] -= d * v2;
        }
        const int r = std::min(k + 3, hi);
        for (int i = lo; i <= r; ++i) {
          const double d = beta * (h[i][k] * v0 + h[i][k + 1] * v1 + h[i][k + 2] * v2);
          h[i][k] -= d * v0;
          h[i][k + 1] -= d * v1;
          h[i][k + 2] -= d * v2;
        }
      }
      x = h[k + 1][k];
      y = h[k + 2][k];
      if (k < hi - 2) z = h[k + 3][k];
    }
    {  // final 2-vector reflector
      const double alpha = -std::copysign(std::sqrt(x * x + y * y), x);
      const double v0 = x - alpha, v1 = y;
      const double vv = v0 * v0 + v1 * v1;
      if (vv != 0) {
        const double beta = 2.0 / vv;
        const int k = hi - 1;
        for (int j = std::max(lo, k - 1); j <= hi; ++j) {
          const double d = beta * (v0 * h[k][j] + v1 * h[k + 1][j]);
          h[k][j] -= d * v0;
          h[k + 1][j] -= d * v1;
        }
        for (int i = lo; i <= hi; ++i) {
          const double d = beta * (h[i][k] * v0 + h[i][k + 1] * v1);
          h[i][k] -= d * v0;
          h[i][k + 1] -= d * v1;
        }
      }
    }
    // clean fill-in below the subdiagonal
    for (int i = lo + 2; i <= hi; ++i)
      for (int j = lo; j < i - 1; ++j) h[i][j] = 0;
  }
}

unsigned realRootsMonic(int D, const double* coeffs, double* roots, RootFilter filter) {
  // Companion matrix (Polynomial.impl.hpp:34-62): sub-diagonal identity, last column -coeffs.
  double C[16];
  for (int i = 0; i < D * D; ++i) C[i] = 0;
  for (int i = 1; i < D; ++i) C[i * D + (i - 1)] = 1;
  for (int i = 0; i < D; ++i) C[i * D + (D - 1)] = -coeffs[i];
  double wr[4], wi[4];
  hessenbergEigenvalues(D, C, wr, wi);
  unsigned count = 0;
  for (int i = 0; i < D; ++i) {
    if (isZero(wi[i])) {  // "real" means |Im| < DBL_EPSILON (Polynomial.impl.hpp:68)
      const bool discard = (filter == StrictlyPositiveRoots && wr[i] <= 0) ||
                           (filter == StrictlyNegativeRoots && wr[i] >= 0);
      if (!discard) roots[count++] = wr[i];
    }
  }
  return count;
}

unsigned realRootsDegenerate(int D, double* coeffs, double* roots, RootFilter filter) {
  // PossiblyDegenerateRootsFinder (Polynomial.impl.hpp:90-109)
  if (D == 0) {
    roots[0] = 0;
    return filter == AllRoots && isZero(coeffs[0]);
  }
  if (isZero(coeffs[D])) return realRootsDegenerate(D - 1, coeffs, roots, filter);
  const double inv = 1.0 / coeffs[D];
  for (int k = 0; k < D; ++k) coeffs[k] *= inv;
  return realRootsMonic(D, coeffs, roots, filter);
}

// =============================================================== FB function
// Canonical-cone FB with optional Jacobian (FischerBurmeister.impl.hpp:69-140,
// [Daviet et al 2011] App. A.1): fb = x + y - (x o x + y o y)^{1/2} via the
// spectral decomposition of z2 = x o x + y o y.
template <bool Jac>
static void fbCanonical(const double x[3], const double y[3], double fb[3], double dx[3][3], double dy[3][3]) {
  double z2[3];
  z2[0] = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  z2[1] = x[0] * x[1] + y[0] * y[1];
  z2[2] = x[0] * x[2] + y[0] * y[2];
  const double nz2t = std::sqrt(z2[1] * z2[1] + z2[2] * z2[2]);
  double w1[3], w2[3];
  w1[0] = w2[0] = .5;
  if (isZero(nz2t)) {
    w1[1] = -.5;
    w1[2] = 0;
    w2[1] = .5;
    w2[2] = 0;
  } else {
    w1[1] = -.5 * z2[1] / nz2t;
    w1[2] = -.5 * z2[2] / nz2t;
    w2[1] = .5 * z2[1] / nz2t;
    w2[2] = .5 * z2[2] / nz2t;
  }
  const double rl1 = std::sqrt(std::max(0.0, z2[0] - 2 * nz2t));
  const double rl2 = std::sqrt(z2[0] + 2 * nz2t);
  double z[3];
  for (int k = 0; k < 3; ++k) {
    z[k] = rl1 * w1[k] + rl2 * w2[k];
    fb[k] = x[k] + y[k] - z[k];
  }
  if (!Jac) return;
  if (isZero(rl2)) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) dx[i][j] = dy[i][j] = 0;
    return;
  }
  if (isZero(rl1)) {
    const double izn = 1. / (x[0] * x[0] + y[0] * y[0]);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dx[i][j] = (i == j) ? (1. - x[0] * izn) : 0;
        dy[i][j] = (i == j) ? (1. - y[0] * izn) : 0;
      }
    return;
  }
  const double det = rl1 * rl2;
  double iL[3][3];
  iL[0][0] = z[0];
  iL[0][1] = -z[1];
  iL[0][2] = -z[2];
  iL[1][0] = -z[1];
  iL[2][0] = -z[2];
  iL[1][1] = (det + z[1] * z[1]) / z[0];
  iL[1][2] = (z[1] * z[2]) / z[0];
  iL[2][1] = (z[2] * z[1]) / z[0];
  iL[2][2] = (det + z[2] * z[2]) / z[0];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) iL[i][j] /= det;
  const double Lx[3][3] = {{x[0], x[1], x[2]}, {x[1], x[0], 0}, {x[2], 0, x[0]}};
  const double Ly[3][3] = {{y[0], y[1], y[2]}, {y[1], y[0], 0}, {y[2], 0, y[0]}};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double sx = 0, sy = 0;
      for (int k = 0; k < 3; ++k) {
        sx += iL[i][k] * Lx[k][j];
        sy += iL[i][k] * Ly[k][j];
      }
      dx[i][j] = (i == j ? 1.0 : 0.0) - sx;
      dy[i][j] = (i == j ? 1.0 : 0.0) - sy;
    }
}

void fbBase(double mu, const double x[3], const double y[3], double fb[3]) {
  // FBBaseFunction::compute (FischerBurmeister.impl.hpp:29-42)
  if (isZero(mu)) {
    fb[0] = x[0] + y[0] - std::sqrt(x[0] * x[0] + y[0] * y[0]);
    fb[1] = x[1];
    fb[2] = x[2];
  } else {
    const double xh[3] = {mu * x[0], x[1], x[2]};
    const double yh[3] = {y[0], mu * y[1], mu * y[2]};
    fbCanonical<false>(xh, yh, fb, nullptr, nullptr);
  }
}

void fbBaseJacobian(double mu, const double x[3], const double y[3], double fb[3], double dx[3][3],
                    double dy[3][3]) {
  // FBBaseFunction::computeJacobian (FischerBurmeister.impl.hpp:44-67)
  if (isZero(mu)) {
    const double z = std::sqrt(x[0] * x[0] + y[0] * y[0]);
    fb[0] = x[0] + y[0] - z;
    fb[1] = x[1];
    fb[2] = x[2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dx[i][j] = (i == j) ? 1.0 : 0.0;
        dy[i][j] = 0;
      }
    if (!isZero(z)) {
      dx[0][0] = 1. - x[0] / z;
      dy[0][0] = 1. - y[0] / z;
    }
  } else {
    const double xh[3] = {mu * x[0], x[1], x[2]};
    const double yh[3] = {y[0], mu * y[1], mu * y[2]};
    fbCanonical<true>(xh, yh, fb, dx, dy);
    for (int i = 0; i < 3; ++i) {
      dx[i][0] *= mu;
      dy[i][1] *= mu;
      dy[i][2] *= mu;
    }
  }
}

void fbLaw(bool coulomb, double mu, const double x[3], const double y[3], double fb[3]) {
  if (coulomb) {
    const double yt[3] = {y[0] + mu * std::sqrt(y[1] * y[1] + y[2] * y[2]), y[1], y[2]};
    fbBase(mu, x, yt, fb);
  } else {
    fbBase(mu, x, y, fb);
  }
}

double lawEval(bool coulomb, double mu, const double x[3], const double y[3]) {
  if (mu < 0) return y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  double fb[3];
  fbLaw(coulomb, mu, x, y, fb);
  return fb[0] * fb[0] + fb[1] * fb[1] + fb[2] * fb[2];
}

void lawProject(double mu, double x[3]) {
  const double nxt = std::sqrt(x[1] * x[1] + x[2] * x[2]);
  const double xn = x[0];
  if (mu < 0) {
    x[0] = x[1] = x[2] = 0;
    return;
  }
  if (nxt <= mu * xn) return;  // inside the cone
  if (mu * nxt <= -xn) {       // inside the polar cone
    x[0] = x[1] = x[2] = 0;
    return;
  }
  const double den2 = (1 + mu * mu);
  const double proj = (xn + mu * nxt) / den2;
  x[0] = proj;
  x[1] *= proj * mu / nxt;
  x[2] *= proj * mu / nxt;
}

void lawCOV(bool coulomb, double mu, const double y[3], double s[3]) {
  s[0] = coulomb ? mu * std::sqrt(y[1] * y[1] + y[2] * y[2]) : 0.0;
  s[1] = s[2] = 0;
}

// =============================================================== local solver
namespace {
struct LocalFB {  // FischerBurmeister<3,double,DeSaxceCOV>(mu, A, b, scaling)
  bool coulomb;
  double mu, scaling;
  const double (*A)[3];
  const double* b;
  void compute(const double x[3], double fb[3]) const {  // FischerBurmeister.impl.hpp:154-159
    double y[3], xs[3];
    for (int i = 0; i < 3; ++i) {
      y[i] = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i];
      xs[i] = scaling * x[i];
    }
    fbLaw(coulomb, mu, xs, y, fb);
  }
  void computeJacobian(const double x[3], double fb[3], double dF[3][3]) const {  // :161-187
    double xs[3], y[3];
    for (int i = 0; i < 3; ++i) {
      xs[i] = scaling * x[i];
      y[i] = A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i];
    }
    double s = 0;
    if (coulomb) {
      s = std::sqrt(y[1] * y[1] + y[2] * y[2]);
      y[0] += mu * s;
    }
    double dy[3][3];
    fbBaseJacobian(mu, xs, y, fb, dF, dy);
    if (coulomb && !isZero(s)) {
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
}  // namespace

double nonSmoothNewton(bool coulomb, const double A[3][3], const double b[3], double mu, double scaling, double tol,
                       unsigned maxIters, double x[3]) {
  // NonSmoothNewton::solve (NonSmoothNewton.impl.hpp:23-79), restated literally,
  // including the descent test on dPhi_dx = dF_dx^T * x (as in the reference).
  const double sigma2 = 1.e-4, alpha = 0.5;
  LocalFB f{coulomb, mu, scaling, A, b};
  double F[3];
  f.compute(x, F);
  const double PhiInit = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  if (PhiInit < tol) return PhiInit;
  double xBest[3] = {0, 0, 0};
  f.compute(xBest, F);
  const double PhiZero = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
  double PhiBest;
  if (PhiZero < PhiInit) {
    PhiBest = PhiZero;
    x[0] = x[1] = x[2] = 0;
  } else {
    PhiBest = PhiInit;
    for (int i = 0; i < 3; ++i) xBest[i] = x[i];
  }
  if (PhiZero < tol) return PhiZero;
  FullPivLU<3> lu;
  for (unsigned it = 0; it < maxIters; ++it) {
    double dF[3][3];
    f.computeJacobian(x, F, dF);
    const double Phi = 0.5 * (F[0] * F[0] + F[1] * F[1] + F[2] * F[2]);
    if (Phi < tol) return Phi;
    if (Phi < PhiBest) {
      PhiBest = Phi;
      for (int i = 0; i < 3; ++i) xBest[i] = x[i];
    }
    double dPhi[3];
    for (int j = 0; j < 3; ++j) dPhi[j] = dF[0][j] * x[0] + dF[1][j] * x[1] + dF[2][j] * x[2];
    lu.compute(dF);
    const double mF[3] = {-F[0], -F[1], -F[2]};
    double dx[3];
    lu.solve(mF, dx);
    const double proj = dx[0] * dPhi[0] + dx[1] * dPhi[1] + dx[2] * dPhi[2];
    const double ndx2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
    const double ndp2 = dPhi[0] * dPhi[0] + dPhi[1] * dPhi[1] + dPhi[2] * dPhi[2];
    if (proj > 0 || proj * proj < sigma2 * ndx2 * ndp2)
      for (int i = 0; i < 3; ++i) dx[i] *= alpha;
    for (int i = 0; i < 3; ++i) x[i] += dx[i];
  }
  for (int i = 0; i < 3; ++i) x[i] = xBest[i];
  return PhiBest;
}

bool solveOrthogonality(bool coulomb, const double W[3][3], const double b[3], double r[3], double mu) {
  if (coulomb) {
    // AnalyticLocalSOCSolver<3,Scalar,true> (LocalSOCSolver.impl.hpp:55-102), [DBB11] App. B.1
    const double wN = W[0][0];
    if (wN < kEps) return false;
    const double wT[2] = {W[1][0], W[2][0]};
    const double WT[2][2] = {{W[1][1], W[1][2]}, {W[2][1], W[2][2]}};
    const double bN = b[0];
    const double bT[2] = {b[1], b[2]};
    const double wTwN[2] = {wT[0] / wN, wT[1] / wN};
    double Wbar[2][2];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) Wbar[i][j] = WT[i][j] - wTwN[i] * wT[j];
    const double bbar[2] = {bT[0] / bN - wTwN[0], bT[1] / bN - wTwN[1]};
    const double A = (Wbar[0][0] + Wbar[1][1]) - (wT[0] * bbar[0] + wT[1] * bbar[1]);
    const double B[2] = {Wbar[1][1] * bbar[0] - Wbar[1][0] * bbar[1], Wbar[0][0] * bbar[1] - Wbar[0][1] * bbar[0]};
    const double C = (Wbar[0][0] * Wbar[1][1] - Wbar[0][1] * Wbar[1][0]) - (wT[0] * B[0] + wT[1] * B[1]);
    const double D = wN * wN / (mu * mu);
    const double coeffs[4] = {C * C - D * (B[0] * B[0] + B[1] * B[1]), 2 * (C * A - D * (bbar[0] * B[0] + bbar[1] * B[1])),
                              2 * C + A * A - D * (bbar[0] * bbar[0] + bbar[1] * bbar[1]), 2 * A};
    double roots[4];
    const unsigned nRoots = realRootsMonic(4, coeffs, roots, StrictlyPositiveRoots);
    if (0 == nRoots) return false;
    double alpha = roots[0];
    for (unsigned i = 1; i != nRoots; ++i)
      if (roots[i] < alpha) alpha = roots[i];
    const double M[2][2] = {{Wbar[0][0] + alpha, Wbar[0][1]}, {Wbar[1][0], Wbar[1][1] + alpha}};
    FullPivLU<2> lu;
    lu.compute(M);
    const double rhs[2] = {-bN * bbar[0], -bN * bbar[1]};
    double rt[2];
    lu.solve(rhs, rt);
    r[1] = rt[0];
    r[2] = rt[1];
    r[0] = std::sqrt(rt[0] * rt[0] + rt[1] * rt[1]) / mu;
    return true;
  }
  // AnalyticLocalSOCSolver<3,Scalar,false> (LocalSOCSolver.impl.hpp:149-209)
  const double A = (b[1] * W[1][2] - b[2] * W[1][1]) * mu * mu + b[0] * W[0][2] - b[2] * W[0][0];
  const double B = (b[0] * W[1][2] + b[1] * W[0][2] - 2 * b[2] * W[0][1]) * mu;
  const double C = 2 * ((b[1] * W[2][2] - b[2] * W[1][2]) * mu * mu - b[0] * W[0][1] + b[1] * W[0][0]);
  const double D = 2 * (b[0] * (W[1][1] - W[2][2]) - b[1] * W[0][1] + b[2] * W[0][2]) * mu;
  const double E = -6 * (b[0] * W[1][2] - b[1] * W[0][2]) * mu;
  double coeffs[5] = {A + B, C - D, E, C + D, B - A};
  double roots[4];
  const unsigned nRoots = realRootsDegenerate(4, coeffs, roots, AllRoots);
  for (unsigned i = 0; i != nRoots; ++i) {
    const double t = roots[i];
    const double CT = (1 - t * t) / (1 + t * t);
    const double ST = 2 * t / (1 + t * t);
    const double dir[3] = {1, mu * CT, mu * ST};
    double rN;
    double den0 = 0, den1 = 0;
    for (int k = 0; k < 3; ++k) {
      den0 += (mu * W[k][1] + CT * W[k][0]) * dir[k];
      den1 += (mu * W[k][2] + ST * W[k][0]) * dir[k];
    }
    if (std::fabs(den0) > std::fabs(den1)) {
      if (isZero(den0)) continue;
      rN = -(CT * b[0] + b[1] * mu) / den0;
    } else {
      if (isZero(den1)) continue;
      rN = -(ST * b[0] + b[2] * mu) / den1;
    }
    if (rN <= 0) continue;
    for (int k = 0; k < 3; ++k) r[k] = rN * dir[k];
    const double uN = W[0][0] * r[0] + W[1][0] * r[1] + W[2][0] * r[2] + b[0];
    if (uN > 0) return true;
  }
  return false;
}

double localSolve(bool coulomb, Strategy strat, const double A[3][3], const double b[3], double x[3], double mu,
                  double tol, double scaling) {
  // LocalSOCSolver<3,double,DeSaxceCOV,Strat>::solve (LocalSOCSolver.impl.hpp:243-318)
  if (mu < 0) {
    // Reference: fullPivHouseholderQr; for the nonsingular diagonal blocks of a
    // Delassus operator this is the same solution as complete-pivoting LU.
    FullPivLU<3> lu;
    lu.compute(A);
    const double mb[3] = {-b[0], -b[1], -b[2]};
    lu.solve(mb, x);
    double r2 = 0;
    for (int i = 0; i < 3; ++i) r2 += sq(A[i][0] * x[0] + A[i][1] * x[1] + A[i][2] * x[2] + b[i]);
    return r2;
  }
  const unsigned newtonIters = 20;
  if (strat == PureNewton) return nonSmoothNewton(coulomb, A, b, mu, scaling, tol, newtonIters, x);
  const double nbT = std::sqrt(b[1] * b[1] + b[2] * b[2]);
  if (b[0] >= (coulomb ? 0 : mu * nbT)) {  // take-off
    x[0] = x[1] = x[2] = 0;
    return 0.;
  }
  if (isZero(mu)) {  // frictionless
    if (A[0][0] < kEps) {
      x[0] = x[1] = x[2] = 0;
      return b[0] * b[0];
    }
    x[1] = x[2] = 0;
    x[0] = -b[0] / A[0][0];
    return 0.;
  }
  double res = 0.;
  if (strat == Hybrid) {
    res = nonSmoothNewton(coulomb, A, b, mu, scaling, tol, newtonIters, x);
    if (res < tol) return res;
  }
  const double x0[3] = {x[0], x[1], x[2]};
  {
    FullPivLU<3> lu;
    lu.compute(A);
    const double mb[3] = {-b[0], -b[1], -b[2]};
    lu.solve(mb, x);
  }
  if (mu * x[0] >= std::sqrt(x[1] * x[1] + x[2] * x[2])) return 0.;  // sticking
  if (!solveOrthogonality(coulomb, A, b, x, mu))
    for (int i = 0; i < 3; ++i) x[i] = x0[i];
  if (strat == RevHybrid) {
    res = nonSmoothNewton(coulomb, A, b, mu, scaling, tol, newtonIters, x);
  } else if (strat == Hybrid) {
    const double refined = nonSmoothNewton(coulomb, A, b, mu, scaling, tol, newtonIters, x);
    if (refined <= res) return refined;
    for (int i = 0; i < 3; ++i) x[i] = x0[i];
  }
  return res;
}

// =============================================================== BSR products
static void buildTransposeIndex(const BSR& A, std::vector<int>& tptr, std::vector<int>& trow, std::vector<int>& tpos,
                                bool strictUpperOfSym) {
  // For a general matrix: column-major index of all blocks. For a symmetric-lower
  // matrix: index of the strictly-lower blocks by column (= strictly-upper by row).
  const int ncol = A.nbc;
  tptr.assign(ncol + 1, 0);
  for (int i = 0; i < A.nbr; ++i)
    for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
      const int j = A.colidx[p];
      if (strictUpperOfSym && j == i) continue;
      tptr[j + 1]++;
    }
  for (int j = 0; j < ncol; ++j) tptr[j + 1] += tptr[j];
  trow.assign(tptr[ncol], 0);
  tpos.assign(tptr[ncol], 0);
  std::vector<int> fill(tptr.begin(), tptr.end() - 1);
  for (int i = 0; i < A.nbr; ++i)
    for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
      const int j = A.colidx[p];
      if (strictUpperOfSym && j == i) continue;
      trow[fill[j]] = i;
      tpos[fill[j]] = p;
      fill[j]++;
    }
}

void bsrBuildUpperIndex(const BSR& A, std::vector<int>& uptr, std::vector<int>& urow, std::vector<int>& upos) {
  buildTransposeIndex(A, uptr, urow, upos, true);
}

void bsrMultiply(const BSR& A, bool transpose, const double* x, double* res, double alpha, double beta) {
  const int br = A.br, bc = A.bc;
  const int nres = (transpose && !A.symLower) ? A.nbc * bc : A.nbr * br;
  for (int i = 0; i < nres; ++i) res[i] = (beta == 0) ? 0.0 : beta * res[i];
  if (A.symLower) {
    // symmetric: stored (i,j), j<=i applied as A_ij x_j and A_ij^T x_i (j != i)
    for (int i = 0; i < A.nbr; ++i)
      for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
        const int j = A.colidx[p];
        const double* blk = A.block(p);
        for (int r = 0; r < br; ++r) {
          double acc = 0;
          for (int c = 0; c < bc; ++c) acc += blk[r * bc + c] * x[j * bc + c];
          res[i * br + r] += alpha * acc;
        }
        if (j != i)
          for (int c = 0; c < bc; ++c) {
            double acc = 0;
            for (int r = 0; r < br; ++r) acc += blk[r * bc + c] * x[i * br + r];
            res[j * bc + c] += alpha * acc;
          }
      }
    return;
  }
  if (!transpose) {
    for (int i = 0; i < A.nbr; ++i)
      for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
        const int j = A.colidx[p];
        const double* blk = A.block(p);
        for (int r = 0; r < br; ++r) {
          double acc = 0;
          for (int c = 0; c < bc; ++c) acc += blk[r * bc + c] * x[j * bc + c];
          res[i * br + r] += alpha * acc;
        }
      }
  } else {
    for (int i = 0; i < A.nbr; ++i)
      for (int p = A.rowptr[i]; p < A.rowptr[i + 1]; ++p) {
        const int j = A.colidx[p];
        const double* blk = A.block(p);
        for (int c = 0; c < bc; ++c) {
          double acc = 0;
          for (int r = 0; r < br; ++r) acc += blk[r * bc + c] * x[i * br + r];
          res[j * bc + c] += alpha * acc;
        }
      }
  }
}

void bsrSplitRowMultiply(const BSR& A, int row, const double* x, double res[3]) {
  // Symmetric 3x3 W only: res += sum_{j != row} W_{row,j} x_j, the lower part read
  // from row `row`, the upper part from the stored blocks (j,row), j > row, transposed.
  // Linear scan for the upper part: this entry point serves small known-answer tests.
  for (int p = A.rowptr[row]; p < A.rowptr[row + 1]; ++p) {
    const int j = A.colidx[p];
    if (j >= row) continue;
    const double* blk = A.block(p);
    for (int r = 0; r < 3; ++r) res[r] += blk[r * 3 + 0] * x[3 * j] + blk[r * 3 + 1] * x[3 * j + 1] + blk[r * 3 + 2] * x[3 * j + 2];
  }
  for (int j = row + 1; j < A.nbr; ++j)
    for (int p = A.rowptr[j]; p < A.rowptr[j + 1]; ++p) {
      if (A.colidx[p] != row) continue;
      const double* blk = A.block(p);
      for (int r = 0; r < 3; ++r) res[r] += blk[0 * 3 + r] * x[3 * j] + blk[1 * 3 + r] * x[3 * j + 1] + blk[2 * 3 + r] * x[3 * j + 2];
    }
}

// Indexed symmetric split-row product used by the solvers (strictly-upper part through the transpose index).
static void splitRowSym(const Problem& P, int row, const double* x, double res[3]) {
  const BSR& W = P.W;
  for (int p = W.rowptr[row]; p < W.rowptr[row + 1]; ++p) {
    const int j = W.colidx[p];
    if (j >= row) continue;
    const double* blk = W.block(p);
    for (int r = 0; r < 3; ++r)
      res[r] += blk[r * 3 + 0] * x[3 * j] + blk[r * 3 + 1] * x[3 * j + 1] + blk[r * 3 + 2] * x[3 * j + 2];
  }
  for (int q = P.wUptr[row]; q < P.wUptr[row + 1]; ++q) {
    const int j = P.wUrow[q];
    const double* blk = W.block(P.wUpos[q]);  // stored W_{j,row}; W_{row,j} = blk^T
    for (int r = 0; r < 3; ++r)
      res[r] += blk[0 * 3 + r] * x[3 * j] + blk[1 * 3 + r] * x[3 * j + 1] + blk[2 * 3 + r] * x[3 * j + 2];
  }
}

// =============================================================== friction problem
void assembleW(Problem& P) {
  // W = H (Minv H^T), keeping j <= i. Term order: for each contact pair (i, j)
  // sharing a body k, W_ij += H_ik (Minv_k H_jk^T).
  const BSR& H = P.H;
  std::vector<int> tptr, trow, tpos;
  buildTransposeIndex(H, tptr, trow, tpos, false);
  // T_p = Minv_{col(p)} H_p^T (6x3), for every H block p
  const int nnz = H.nnzb();
  std::vector<double> T((size_t)nnz * 18);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < nnz; ++p) {
    const int k = H.colidx[p];
    const double* h = H.block(p);
    const double* m = &P.minv[(size_t)k * 36];
    for (int r = 0; r < 6; ++r)
      for (int c = 0; c < 3; ++c) {
        double acc = 0;
        for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * h[c * 6 + l];
        T[(size_t)p * 18 + r * 3 + c] = acc;
      }
  }
  P.W = BSR();
  P.W.nbr = P.W.nbc = P.nc;
  P.W.br = P.W.bc = 3;
  P.W.symLower = true;
  P.W.rowptr.assign(P.nc + 1, 0);
  std::vector<std::vector<int>> cols(P.nc);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < P.nc; ++i) {
    std::vector<int>& c = cols[i];
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const int k = H.colidx[p];
      for (int q = tptr[k]; q < tptr[k + 1]; ++q)
        if (trow[q] <= i) c.push_back(trow[q]);
    }
    std::sort(c.begin(), c.end());
    c.erase(std::unique(c.begin(), c.end()), c.end());
  }
  for (int i = 0; i < P.nc; ++i) P.W.rowptr[i + 1] = P.W.rowptr[i] + (int)cols[i].size();
  P.W.colidx.resize(P.W.rowptr[P.nc]);
  P.W.vals.assign((size_t)P.W.rowptr[P.nc] * 9, 0.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < P.nc; ++i) {
    const int base = P.W.rowptr[i];
    for (size_t t = 0; t < cols[i].size(); ++t) P.W.colidx[base + t] = cols[i][t];
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const int k = H.colidx[p];
      const double* h = H.block(p);
      for (int q = tptr[k]; q < tptr[k + 1]; ++q) {
        const int j = trow[q];
        if (j > i) continue;
        const int slot = base + int(std::lower_bound(cols[i].begin(), cols[i].end(), j) - cols[i].begin());
        const double* t = &T[(size_t)tpos[q] * 18];
        double* w = &P.W.vals[(size_t)slot * 9];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double acc = 0;
            for (int l = 0; l < 6; ++l) acc += h[r * 6 + l] * t[l * 3 + c];
            w[r * 3 + c] += acc;
          }
      }
    }
  }
  P.hasW = true;
  bsrBuildUpperIndex(P.W, P.wUptr, P.wUrow, P.wUpos);
}

void computeDiag(Problem& P) {
  // W_ii = sum_k H_ik Minv_k H_ik^T, from H directly (also valid when W is not assembled).
  P.diag.assign((size_t)P.nc * 9, 0.0);
  const BSR& H = P.H;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P.nc; ++i) {
    double* d = &P.diag[(size_t)i * 9];
    if (P.hasW) {  // take the assembled block if present (identical value to the oracle's assembly)
      for (int p = P.W.rowptr[i]; p < P.W.rowptr[i + 1]; ++p)
        if (P.W.colidx[p] == i)
          for (int e = 0; e < 9; ++e) d[e] = P.W.vals[(size_t)p * 9 + e];
      continue;
    }
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const int k = H.colidx[p];
      const double* h = H.block(p);
      const double* m = &P.minv[(size_t)k * 36];
      double t[6][3];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0;
          for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * h[c * 6 + l];
          t[r][c] = acc;
        }
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0;
          for (int l = 0; l < 6; ++l) acc += h[r * 6 + l] * t[l][c];
          d[r * 3 + c] += acc;
        }
    }
  }
}

namespace {
struct MatFreeCache {  // body-major transposed index of H (cacheTranspose of H)
  std::vector<int> tptr, trow, tpos;
};
}  // namespace

static void symGather(const Problem& P, const double* x, double* y) {
  const BSR& W = P.W;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P.nc; ++i) {
    double acc[3] = {0, 0, 0};
    for (int p = W.rowptr[i]; p < W.rowptr[i + 1]; ++p) {
      const int j = W.colidx[p];
      const double* blk = W.block(p);
      for (int r = 0; r < 3; ++r)
        acc[r] += blk[r * 3 + 0] * x[3 * j] + blk[r * 3 + 1] * x[3 * j + 1] + blk[r * 3 + 2] * x[3 * j + 2];
    }
    for (int q = P.wUptr[i]; q < P.wUptr[i + 1]; ++q) {
      const int j = P.wUrow[q];
      const double* blk = W.block(P.wUpos[q]);
      for (int r = 0; r < 3; ++r)
        acc[r] += blk[0 * 3 + r] * x[3 * j] + blk[1 * 3 + r] * x[3 * j + 1] + blk[2 * 3 + r] * x[3 * j + 2];
    }
    y[3 * i] = acc[0];
    y[3 * i + 1] = acc[1];
    y[3 * i + 2] = acc[2];
  }
}

static void matFree(const Problem& P, const MatFreeCache& C, const double* x, double* y, std::vector<double>& tmp) {
  const BSR& H = P.H;
  tmp.resize((size_t)P.nb * 6);
  // w_k = Minv_k * sum_{i in k} H_ik^T x_i
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < P.nb; ++k) {
    double z[6] = {0, 0, 0, 0, 0, 0};
    for (int q = C.tptr[k]; q < C.tptr[k + 1]; ++q) {
      const int i = C.trow[q];
      const double* h = H.block(C.tpos[q]);
      for (int c = 0; c < 6; ++c) z[c] += h[0 * 6 + c] * x[3 * i] + h[1 * 6 + c] * x[3 * i + 1] + h[2 * 6 + c] * x[3 * i + 2];
    }
    const double* m = &P.minv[(size_t)k * 36];
    for (int r = 0; r < 6; ++r) {
      double acc = 0;
      for (int c = 0; c < 6; ++c) acc += m[r * 6 + c] * z[c];
      tmp[(size_t)k * 6 + r] = acc;
    }
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < P.nc; ++i) {
    double acc[3] = {0, 0, 0};
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const int k = H.colidx[p];
      const double* h = H.block(p);
      const double* w = &tmp[(size_t)k * 6];
      for (int r = 0; r < 3; ++r) {
        double s = 0;
        for (int c = 0; c < 6; ++c) s += h[r * 6 + c] * w[c];
        acc[r] += s;
      }
    }
    y[3 * i] = acc[0];
    y[3 * i + 1] = acc[1];
    y[3 * i + 2] = acc[2];
  }
}

void applyW(const Problem& P, bool useAssembled, const double* x, double* y) {
  if (useAssembled && P.hasW) {
    symGather(P, x, y);
    return;
  }
  MatFreeCache C;
  buildTransposeIndex(P.H, C.tptr, C.trow, C.tpos, false);
  std::vector<double> tmp;
  matFree(P, C, x, y, tmp);
}

double evalResidual(int n, bool coulomb, const double* mu, const double* scaling, const double* y, const double* x,
                    Norm norm) {
  // ConstrainedSolverBase::eval (ConstrainedSolverBase.impl.hpp:21-74)
  if (norm == Infinity) {
    double err = 0;
#pragma omp parallel for reduction(max : err) schedule(static)
    for (int i = 0; i < n; ++i) {
      const double s = scaling ? scaling[i] : 1.0;
      const double lx[3] = {x[3 * i] * s, x[3 * i + 1] * s, x[3 * i + 2] * s};
      err = std::max(err, lawEval(coulomb, mu[i], lx, &y[3 * i]));
    }
    return err;
  }
  double err = 0;
#pragma omp parallel for reduction(+ : err) schedule(static)
  for (int i = 0; i < n; ++i) {
    const double s = scaling ? scaling[i] : 1.0;
    const double lx[3] = {x[3 * i] * s, x[3 * i + 1] * s, x[3 * i + 2] * s};
    err += lawEval(coulomb, mu[i], lx, &y[3 * i]);
  }
  return err / (1 + n);
}

// =============================================================== projected gradient
namespace {
struct Op {  // the matrix M of ProjectedGradient (assembled W or the product expression)
  const Problem& P;
  bool assembled;
  MatFreeCache C;
  mutable std::vector<double> tmp;
  mutable unsigned count = 0;
  Op(const Problem& P_, bool a) : P(P_), assembled(a && P_.hasW) {
    if (!assembled) buildTransposeIndex(P.H, C.tptr, C.trow, C.tpos, false);
  }
  void apply(const double* x, double* y) const {
    ++count;
    if (assembled)
      symGather(P, x, y);
    else
      matFree(P, C, x, y, tmp);
  }
};

typedef std::vector<double> Vec;
static double dot(const Vec& a, const Vec& b) {
  double s = 0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
static double sqn(const Vec& a) { return dot(a, a); }

struct PG {
  const Problem& P;
  const PGOptions& o;
  Op M;
  Vec scaling;
  SolveStats* st;
  PG(const Problem& P_, const PGOptions& o_, SolveStats* s) : P(P_), o(o_), M(P_, o_.useAssembled != 0), st(s) {
    // ConstrainedSolverBase::updateScalings (ConstrainedSolverBase.impl.hpp:78-118)
    scaling.assign(P.nc, 1.0);
    const bool explicitMatrix = M.assembled;
    if (explicitMatrix || o.diagScaling) {
      for (int i = 0; i < P.nc; ++i) {
        const double* d = &P.diag[(size_t)i * 9];
        scaling[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3);
      }
    }
  }
  void project(Vec& x) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P.nc; ++i) lawProject(P.mu[i], &x[3 * i]);
  }
  void cov(const Vec& y, Vec& s) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < P.nc; ++i) lawCOV(o.coulomb != 0, P.mu[i], &y[3 * i], &s[3 * i]);
  }
  double eval(const Vec& y, const Vec& x) const {
    return evalResidual(P.nc, o.coulomb != 0, P.mu.data(), scaling.data(), y.data(), x.data(),
                        o.infNorm ? Infinity : MeanSquare);
  }
  // test_residual (ProjectedGradient.impl.hpp:41-52)
  bool testResidual(unsigned it, const Vec& x, const Vec& y, Vec& xBest, double& minRes) const {
    const double res = eval(y, x);
    if (st) st->history.push_back(res);
    if (0 == it || res < minRes) {
      xBest = x;
      minRes = res;
    }
    return res < o.tol;
  }
  void guessAlpha(const Vec& x, double& alpha) const {  // ProjectedGradient.impl.hpp:21-28
    Vec d(x.size()), Md(x.size());
    for (size_t i = 0; i < x.size(); ++i) d[i] = x[i] - 1.0;
    M.apply(d.data(), Md.data());
    const double nMx = sqn(Md);
    if (nMx > 1.e-16) alpha = std::min(1., sqn(d) / nMx);
  }
  static double nesterov(double& theta, double q) {  // :30-39
    const double tp = theta;
    const double tp2 = tp * tp;
    const double bb = tp2 - q;
    const double delta = std::sqrt(bb * bb + 4 * tp2);
    theta = .5 * (delta - bb);
    return (tp - tp2) / (tp2 + theta);
  }
  // armijo_ls (:54-72)
  void armijo(const Vec& x, const Vec& bs, const Vec& dir, const Vec& grad, double J, double& alpha, Vec& xs,
              double& Js, Vec& Mx) const {
    const size_t n = x.size();
    Js = J;
    for (unsigned ls = 0; ls < o.lsIters; ++ls, alpha *= o.lsPessimistic) {
      for (size_t i = 0; i < n; ++i) xs[i] = x[i] + alpha * dir[i];
      project(xs);
      M.apply(xs.data(), Mx.data());
      double js = 0, decr = 0;
      for (size_t i = 0; i < n; ++i) js += xs[i] * (.5 * Mx[i] + bs[i]);
      for (size_t i = 0; i < n; ++i) decr += grad[i] * (xs[i] - x[i]);
      Js = js;
      if (Js < J + o.lsArmijo * decr) break;
    }
  }

  double solveAPGD(Vec& x, bool nesterovOn, unsigned& iters) const {  // PgMethod<Descent|APGD> (:74-155)
    const size_t n = x.size();
    const Vec& b = P.b;
    Vec Mx(n), y(n), xs(n), s(n), prev, xBest;
    project(x);
    M.apply(x.data(), Mx.data());
    double thetaPrev = 1., q = 0;
    prev = x;
    double alpha = 1.;
    guessAlpha(x, alpha);
    double minRes = -1;
    unsigned it;
    for (it = 0; it < o.maxIters; ++it) {
      for (size_t i = 0; i < n; ++i) y[i] = Mx[i] + b[i];
      if (testResidual(it, x, y, xBest, minRes)) break;
      cov(y, s);
      for (size_t i = 0; i < n; ++i) y[i] += s[i];
      double J = 0;
      for (size_t i = 0; i < n; ++i) J += x[i] * (.5 * Mx[i] + b[i] + s[i]);
      double Js = J;
      alpha *= o.lsOptimistic;
      for (unsigned ls = 0; ls < o.lsIters; ++ls, alpha *= o.lsPessimistic) {
        for (size_t i = 0; i < n; ++i) xs[i] = x[i] - alpha * y[i];
        project(xs);
        M.apply(xs.data(), Mx.data());
        double js = 0, decr = 0, nd = 0;
        for (size_t i = 0; i < n; ++i) js += xs[i] * (.5 * Mx[i] + b[i] + s[i]);
        for (size_t i = 0; i < n; ++i) decr += y[i] * (xs[i] - x[i]);
        for (size_t i = 0; i < n; ++i) nd += (xs[i] - x[i]) * (xs[i] - x[i]);
        Js = js;
        if (Js < J + decr + .5 / alpha * nd) break;
      }
      double decr = 0;
      for (size_t i = 0; i < n; ++i) decr += (xs[i] - prev[i]) * y[i];
      if (nesterovOn && decr < 0) {
        const double beta = nesterov(thetaPrev, q);
        for (size_t i = 0; i < n; ++i) x[i] = xs[i] + beta * (xs[i] - prev[i]);
        M.apply(x.data(), Mx.data());
      } else {
        x = xs;
        thetaPrev = 1.;
      }
      prev = xs;
    }
    iters = it;
    x = xBest;
    return minRes;
  }

  double solveStandard(Vec& x, unsigned& iters) const {  // PgMethod<Standard> (:157-209)
    const size_t n = x.size();
    const Vec& b = P.b;
    Vec Mx(n), y(n), xs(n), s(n), pg(n), Mpg(n), bs(n), xBest;
    project(x);
    M.apply(x.data(), Mx.data());
    double alpha = 1;
    double minRes = -1;
    unsigned it;
    for (it = 0; it < o.maxIters; ++it) {
      for (size_t i = 0; i < n; ++i) y[i] = Mx[i] + b[i];
      if (testResidual(it, x, y, xBest, minRes)) break;
      cov(y, s);
      for (size_t i = 0; i < n; ++i) y[i] += s[i];
      double J = 0;
      for (size_t i = 0; i < n; ++i) J += x[i] * (.5 * Mx[i] + b[i] + s[i]);
      for (size_t i = 0; i < n; ++i) xs[i] = x[i] - y[i];
      project(xs);
      for (size_t i = 0; i < n; ++i) pg[i] = x[i] - xs[i];
      M.apply(pg.data(), Mpg.data());
      const double beta = -dot(y, pg) / dot(pg, Mpg);
      for (size_t i = 0; i < n; ++i) pg[i] *= beta;
      double Js;
      alpha = std::min(1., alpha * o.lsOptimistic);
      for (size_t i = 0; i < n; ++i) bs[i] = b[i] + s[i];
      armijo(x, bs, pg, y, J, alpha, xs, Js, Mx);
      x = xs;
    }
    iters = it;
    x = xBest;
    return minRes;
  }

  double solveConjugated(Vec& x, unsigned& iters) const {  // PgMethod<Conjugated> (:211-286)
    const size_t n = x.size();
    const Vec& b = P.b;
    Vec Mx(n), y(n), xs(n), s(n), dir(n), prevDir(n), prevPG, bs(n), xBest;
    project(x);
    M.apply(x.data(), Mx.data());
    prevPG = x;
    double prevN2 = 0, alpha = 1.;
    double minRes = -1;
    unsigned it;
    for (it = 0; it < o.maxIters; ++it) {
      for (size_t i = 0; i < n; ++i) y[i] = Mx[i] + b[i];
      if (testResidual(it, x, y, xBest, minRes)) break;
      cov(y, s);
      for (size_t i = 0; i < n; ++i) y[i] += s[i];
      double J = 0;
      for (size_t i = 0; i < n; ++i) J += x[i] * (.5 * Mx[i] + b[i] + s[i]);
      for (size_t i = 0; i < n; ++i) xs[i] = x[i] - y[i];
      project(xs);
      for (size_t i = 0; i < n; ++i) xs[i] = x[i] - xs[i];
      const double ng2 = sqn(xs);
      if (prevN2 == 0.) {
        for (size_t i = 0; i < n; ++i) dir[i] = -y[i];
      } else {
        double den = 0, xp = 0;
        for (size_t i = 0; i < n; ++i) den += prevDir[i] * (xs[i] - prevPG[i]);
        xp = dot(xs, prevPG);
        const double beta = (den < 1.e-12) ? std::max(0., (ng2 - xp)) / prevN2 : std::max(0., (ng2 - xp)) / den;
        for (size_t i = 0; i < n; ++i) dir[i] = beta * prevDir[i] - y[i];
        if (dot(dir, y) >= 0.)
          for (size_t i = 0; i < n; ++i) dir[i] = -y[i];
      }
      prevPG = xs;
      prevN2 = ng2;
      double Js;
      alpha = alpha * o.lsOptimistic;
      for (size_t i = 0; i < n; ++i) bs[i] = b[i] + s[i];
      armijo(x, bs, dir, y, J, alpha, xs, Js, Mx);
      for (size_t i = 0; i < n; ++i) prevDir[i] = (xs[i] - x[i]) / alpha;
      x = xs;
    }
    iters = it;
    x = xBest;
    return minRes;
  }

  double solveSPG(Vec& x, unsigned& iters) const {  // PgMethod<SPG> (:288-351)
    const size_t n = x.size();
    const Vec& b = P.b;
    Vec Mx(n), y(n), xs(n), s(n), z(n), w(n), prev, xBest;
    project(x);
    prev = x;
    M.apply(x.data(), Mx.data());
    double thetaPrev = 1, q = 0;
    const double aMin = 1.e-6, aMax = 1.e6;
    double alpha = 1;
    guessAlpha(x, alpha);
    double minRes = -1;
    unsigned it;
    for (it = 0; it < o.maxIters; ++it) {
      for (size_t i = 0; i < n; ++i) y[i] = Mx[i] + b[i];
      if (testResidual(it, x, y, xBest, minRes)) break;
      cov(y, s);
      for (size_t i = 0; i < n; ++i) y[i] += s[i];
      for (size_t i = 0; i < n; ++i) xs[i] = x[i] - alpha * y[i];
      project(xs);
      double d = 0;
      for (size_t i = 0; i < n; ++i) d += (xs[i] - prev[i]) * y[i];
      if (d < 0.) {
        const double beta = nesterov(thetaPrev, q);
        for (size_t i = 0; i < n; ++i) {
          w[i] = (xs[i] + beta * (xs[i] - prev[i])) - x[i];
          x[i] += w[i];
        }
      } else {
        for (size_t i = 0; i < n; ++i) w[i] = xs[i] - x[i];
        x = xs;
        thetaPrev = 1.;
      }
      for (size_t i = 0; i < n; ++i) z[i] = -Mx[i];
      M.apply(x.data(), Mx.data());
      prev = xs;
      for (size_t i = 0; i < n; ++i) z[i] += Mx[i];
      alpha = dot(w, z) / sqn(z);
      alpha = std::min(aMax, std::max(aMin, alpha));
    }
    iters = it;
    x = xBest;
    return minRes;
  }
};
}  // namespace

double pgSolve(const Problem& P, const PGOptions& o, double* xio, SolveStats* st) {
  PG pg(P, o, st);
  Vec x(xio, xio + 3 * (size_t)P.nc);
  unsigned iters = 0;
  double res = -1;
  switch (o.variant) {
    case PG_Standard: res = pg.solveStandard(x, iters); break;
    case PG_Descent: res = pg.solveAPGD(x, false, iters); break;
    case PG_Conjugated: res = pg.solveConjugated(x, iters); break;
    case PG_APGD: res = pg.solveAPGD(x, true, iters); break;
    case PG_SPG: res = pg.solveSPG(x, iters); break;
    default: return -1;  // unknown variant (ProjectedGradient.impl.hpp:373)
  }
  std::copy(x.begin(), x.end(), xio);
  if (st) {
    st->iterations = iters;
    st->spmvs = pg.M.count;
    st->residual = res;
  }
  return res;
}

// =============================================================== coloring + GS
void computeColoring(const BSR& W, std::vector<int>& perm, std::vector<int>& colors) {
  // Coloring::compute (Coloring.impl.hpp:20-62). A color conflicts with contact i
  // when std::lower_bound(color, j) != color.end() for some lower neighbour j < i,
  // i.e. when the color already holds an index >= j. Colors are sorted ascending
  // (contacts are appended in order), so that test equals last(color) >= j, and
  // over all lower neighbours it equals last(color) >= min_j. This is the
  // reference's behaviour as written (it is stricter than an exact membership
  // test, so it yields more colors, all still conflict-free).
  const int n = W.nbr;
  std::vector<std::vector<int>> tmp;
  std::vector<int> last;
  for (int i = 0; i < n; ++i) {
    int jmin = std::numeric_limits<int>::max();
    for (int p = W.rowptr[i]; p < W.rowptr[i + 1]; ++p) {
      const int j = W.colidx[p];
      if (j < i) jmin = std::min(jmin, j);
    }
    bool placed = false;
    for (size_t c = 0; c < tmp.size(); ++c) {
      const bool ok = (jmin == std::numeric_limits<int>::max()) || last[c] < jmin;
      if (ok) {
        tmp[c].push_back(i);
        last[c] = i;
        placed = true;
        break;
      }
    }
    if (!placed) {
      tmp.emplace_back(1, i);
      last.push_back(i);
    }
  }
  perm.clear();
  colors.assign(tmp.size() + 1, 0);
  for (size_t c = 0; c < tmp.size(); ++c) {
    colors[c + 1] = colors[c] + (int)tmp[c].size();
    perm.insert(perm.end(), tmp[c].begin(), tmp[c].end());
  }
}

void permuteProblem(Problem& P, const std::vector<int>& perm) {
  // new contact i is old contact perm[i] (bogus::applyPermutation semantics)
  const int n = P.nc;
  std::vector<int> inv(n);
  for (int i = 0; i < n; ++i) inv[perm[i]] = i;
  BSR H2 = P.H;
  H2.rowptr.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) H2.rowptr[i + 1] = H2.rowptr[i] + (P.H.rowptr[perm[i] + 1] - P.H.rowptr[perm[i]]);
  for (int i = 0; i < n; ++i) {
    const int src = perm[i];
    int dst = H2.rowptr[i];
    for (int p = P.H.rowptr[src]; p < P.H.rowptr[src + 1]; ++p, ++dst) {
      H2.colidx[dst] = P.H.colidx[p];
      std::memcpy(&H2.vals[(size_t)dst * 18], P.H.block(p), 18 * sizeof(double));
    }
  }
  P.H = H2;
  std::vector<double> b2(P.b.size()), mu2(P.mu.size());
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) b2[3 * i + k] = P.b[3 * perm[i] + k];
    mu2[i] = P.mu[perm[i]];
  }
  P.b.swap(b2);
  P.mu.swap(mu2);
  if (P.hasW) {
    // permute W (symmetric lower): block (i,j) of the new matrix is old (perm[i], perm[j]),
    // transposed when the new position falls in the upper triangle (SparseBlockMatrixBase.impl.hpp:328-383)
    BSR W2;
    W2.nbr = W2.nbc = n;
    W2.br = W2.bc = 3;
    W2.symLower = true;
    std::vector<std::vector<std::pair<int, std::array<double, 9>>>> rows(n);
    for (int oi = 0; oi < n; ++oi)
      for (int p = P.W.rowptr[oi]; p < P.W.rowptr[oi + 1]; ++p) {
        const int oj = P.W.colidx[p];
        int ni = inv[oi], nj = inv[oj];
        std::array<double, 9> blk;
        const double* src = P.W.block(p);
        if (nj > ni) {
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) blk[r * 3 + c] = src[c * 3 + r];
          std::swap(ni, nj);
        } else {
          for (int e = 0; e < 9; ++e) blk[e] = src[e];
        }
        rows[ni].push_back({nj, blk});
      }
    W2.rowptr.assign(n + 1, 0);
    for (int i = 0; i < n; ++i) {
      std::sort(rows[i].begin(), rows[i].end(), [](const auto& a, const auto& b) { return a.first < b.first; });
      W2.rowptr[i + 1] = W2.rowptr[i] + (int)rows[i].size();
    }
    W2.colidx.resize(W2.rowptr[n]);
    W2.vals.resize((size_t)W2.rowptr[n] * 9);
    for (int i = 0; i < n; ++i) {
      int dst = W2.rowptr[i];
      for (auto& e : rows[i]) {
        W2.colidx[dst] = e.first;
        std::memcpy(&W2.vals[(size_t)dst * 9], e.second.data(), 9 * sizeof(double));
        ++dst;
      }
    }
    P.W = W2;
    bsrBuildUpperIndex(P.W, P.wUptr, P.wUrow, P.wUpos);
  }
  if (!P.diag.empty()) {
    std::vector<double> d2(P.diag.size());
    for (int i = 0; i < n; ++i)
      for (int e = 0; e < 9; ++e) d2[(size_t)i * 9 + e] = P.diag[(size_t)perm[i] * 9 + e];
    P.diag.swap(d2);
  }
}

// Smallest eigenvalue of a symmetric 3x3 matrix (cyclic Jacobi); used for the
// auto-regularization of GaussSeidelBase::processLocalMatrices (GaussSeidelBase.impl.hpp:36-39).
static double minEigSym3(const double* A) {
  double a[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) a[i][j] = 0.5 * (A[i * 3 + j] + A[j * 3 + i]);
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
    if (off < 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (a[p][q] == 0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2 * a[p][q]);
        const double t = std::copysign(1.0, theta) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < 3; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < 3; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  return std::min(a[0][0], std::min(a[1][1], a[2][2]));
}

double gsSolve(const Problem& P, const GSOptions& o, const std::vector<int>& colors, double* xio, SolveStats* st) {
  assert(P.hasW);
  const int n = P.nc;
  const bool coulomb = o.coulomb != 0;
  const Strategy strat = (Strategy)o.strategy;
  const double localTol = std::pow(kEps, .75);  // SOCLaw::m_localTol (SOCLaw.impl.hpp:35)
  // updateLocalMatrices + processLocalMatrices (GaussSeidel.impl.hpp:37-59, GaussSeidelBase.impl.hpp:20-40)
  std::vector<double> local(P.diag), reg(n, 0.0), scaling(n, 1.0);
  for (int i = 0; i < n; ++i) {
    double* d = &local[(size_t)i * 9];
    scaling[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3);
    if (o.autoRegularization > 0.) {
      reg[i] = std::max(0., o.autoRegularization - minEigSym3(d));
      d[0] += reg[i];
      d[4] += reg[i];
      d[8] += reg[i];
    }
  }
  const Norm norm = o.infNorm ? Infinity : MeanSquare;
  Vec x(xio, xio + 3 * (size_t)n), y(3 * (size_t)n), xBest;
  const Vec& w = P.b;  // solve(): W operator of solveWithLinearConstraints is Zero, so w = b
  double errBest = std::numeric_limits<double>::max();
  unsigned spmvs = 0;
  auto evalAndKeepBest = [&](const Vec& xx) {
    symGather(P, xx.data(), y.data());
    ++spmvs;
    for (size_t k = 0; k < y.size(); ++k) y[k] += w[k];
    const double err = evalResidual(n, coulomb, P.mu.data(), scaling.data(), y.data(), xx.data(), norm);
    if (err < errBest) {
      xBest = xx;
      errBest = err;
    }
    return err;
  };
  evalAndKeepBest(x);
  if (o.tryZero) {  // GaussSeidelBase::tryZero (GaussSeidelBase.impl.hpp:61-77)
    Vec zero(x.size(), 0.0);
    const double errZero = evalResidual(n, coulomb, P.mu.data(), scaling.data(), P.b.data(), zero.data(), norm);
    if (errZero < errBest) {
      errBest = errZero;
      xBest = zero;
      x = zero;
    } else {
      x = xBest;
    }
  }
  if (st) st->history.push_back(errBest);
  std::vector<unsigned char> skip(n, 0);
  double ndxRef = 0;
  const double absSkipTol = std::min(o.skipTol, o.tol);
  const double absSkipIters = std::min(o.skipIters, (unsigned)std::sqrt((double)n));
  // With the per-color ndxRef and colors that hold no coupled pair, the contacts of a
  // color are independent of each other: they are swept in parallel, as the
  // reference's OpenMP loop over a color (GaussSeidel.impl.hpp:77-86), with results
  // identical to the sequential sweep.
  bool parallelColors = o.ndxRefPerColor != 0;
  if (parallelColors) {
    std::vector<int> colorOf(n);
    for (size_t c = 0; c + 1 < colors.size(); ++c)
      for (int i = colors[c]; i < colors[c + 1]; ++i) colorOf[i] = (int)c;
    for (int i = 0; i < n && parallelColors; ++i)
      for (int p = P.W.rowptr[i]; p < P.W.rowptr[i + 1]; ++p)
        if (P.W.colidx[p] != i && colorOf[P.W.colidx[p]] == colorOf[i]) parallelColors = false;
  }
  // one contact of the sweep; ndxMax: the per-color running maximum (ndxRefPerColor)
  auto sweepContact = [&](int i, double ndxRefColor, double& ndxMax) {
    if (skip[i]) {
      --skip[i];
      return;
    }
    double lx[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
    double lb[3];
    for (int k = 0; k < 3; ++k) lb[k] = w[3 * i + k] - reg[i] * lx[k];
    splitRowSym(P, i, x.data(), lb);
    double ldx[3] = {-lx[0], -lx[1], -lx[2]};
    const double(*A)[3] = reinterpret_cast<const double(*)[3]>(&local[(size_t)i * 9]);
    const double lres = localSolve(coulomb, strat, A, lb, lx, P.mu[i], localTol, scaling[i]);
    const bool ok = localTol > lres;
    for (int k = 0; k < 3; ++k) ldx[k] += lx[k];
    if (!ok)
      for (int k = 0; k < 3; ++k) ldx[k] *= .5;
    for (int k = 0; k < 3; ++k) x[3 * i + k] += ldx[k];
    const double s2 = scaling[i] * scaling[i];
    const double nx2 = s2 * (lx[0] * lx[0] + lx[1] * lx[1] + lx[2] * lx[2]);
    const double ndx2 = s2 * (ldx[0] * ldx[0] + ldx[1] * ldx[1] + ldx[2] * ldx[2]);
    double ref;
    if (o.ndxRefPerColor) {
      if (ndx2 > ndxMax) ndxMax = ndx2;
      ref = ndxRefColor;
    } else {
      if (ndx2 > ndxRef) ndxRef = ndx2;
      ref = ndxRef;
    }
    if (std::min(nx2, ndx2) < absSkipTol || ndx2 < o.skipTol * std::min(nx2, ref))
      skip[i] = (unsigned char)absSkipIters;
  };
  unsigned it;
  for (it = 1; it <= o.maxIters; ++it) {
    // innerLoop (GaussSeidel.impl.hpp:62-124), identity permutation within contiguous colors
    for (size_t c = 0; c + 1 < colors.size(); ++c) {
      const double ndxRefColor = ndxRef;
      double ndxRefNew = ndxRef;
      if (parallelColors) {
#pragma omp parallel for reduction(max : ndxRefNew) schedule(dynamic, 64)
        for (int i = colors[c]; i < colors[c + 1]; ++i) sweepContact(i, ndxRefColor, ndxRefNew);
      } else {
        for (int i = colors[c]; i < colors[c + 1]; ++i) sweepContact(i, ndxRefColor, ndxRefNew);
      }
      if (o.ndxRefPerColor) ndxRef = ndxRefNew;
    }
    if (0 == (it % o.evalEvery)) {
      const double err = evalAndKeepBest(x);
      if (st) st->history.push_back(err);
      if (err < o.tol) break;
      ndxRef /= o.evalEvery;
    }
  }
  if (it > o.maxIters) x = xBest;
  std::copy(x.begin(), x.end(), xio);
  if (st) {
    st->iterations = std::min(it, o.maxIters);
    st->spmvs = spmvs;
    st->residual = errBest;
  }
  return errBest;
}

// ProductGaussSeidel::solve with a block-diagonal D = Minv and the precomputed
// D H^T (src/Core/BlockSolvers/ProductGaussSeidel.impl.hpp:225-304, innerLoop
// :128-218, DMtStorage<.., true> :64-86; reached through
// PrimalFrictionProblem::solveWith(ProductGaussSeidelType&), PrimalFrictionProblem.cpp:40-59).
// The solver keeps Mx = D H^T x (one 6-vector per body): a contact's local
// right-hand side is lb = w_i - A_i x_i + H_i Mx, and after its update
// Mx_k += (D_k H_ik^T) dx_i for each of its bodies. Mx is recomputed from x at
// every evaluation. W never appears.
double pgsSolve(const Problem& P, const GSOptions& o, const std::vector<int>& colors, double* xio, SolveStats* st) {
  const int n = P.nc, nb = P.nb;
  const BSR& H = P.H;
  const bool coulomb = o.coulomb != 0;
  const Strategy strat = (Strategy)o.strategy;
  const double localTol = std::pow(kEps, .75);  // SOCLaw::m_localTol (SOCLaw.impl.hpp:35)
  // updateLocalMatrices (:147-165): A_i = sum_k H_ik D_k H_ik^T, then
  // processLocalMatrices (GaussSeidelBase.impl.hpp:20-40)
  std::vector<double> local(P.diag), reg(n, 0.0), scaling(n, 1.0);
  for (int i = 0; i < n; ++i) {
    double* d = &local[(size_t)i * 9];
    scaling[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3);
    if (o.autoRegularization > 0.) {
      reg[i] = std::max(0., o.autoRegularization - minEigSym3(d));
      d[0] += reg[i];
      d[4] += reg[i];
      d[8] += reg[i];
    }
  }
  // DMtStorage<.., true>::compute: T_p = D_k H_p^T (6x3, row-major) for block p = (i, k)
  const int nnz = H.nnzb();
  std::vector<double> T((size_t)nnz * 18);
  for (int i = 0; i < n; ++i)
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const double* h = H.block(p);
      const double* m = &P.minv[(size_t)H.colidx[p] * 36];
      for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 3; ++c) {
          double acc = 0;
          for (int l = 0; l < 6; ++l) acc += m[r * 6 + l] * h[c * 6 + l];
          T[(size_t)p * 18 + r * 3 + c] = acc;
        }
    }
  // D H^T as a body-major (row-major) matrix: the contacts of body k in ascending order
  std::vector<int> tptr, trow, tpos;
  buildTransposeIndex(H, tptr, trow, tpos, false);
  Vec Mx((size_t)nb * 6, 0.0);
  // DMtStorage::multiply (:75-80): Mx = (D H^T) x, then y += H Mx
  auto dmtMultiply = [&](const Vec& x) {
    for (int k = 0; k < nb; ++k) {
      double acc[6] = {0, 0, 0, 0, 0, 0};
      for (int q = tptr[k]; q < tptr[k + 1]; ++q) {
        const int i = trow[q];
        const double* t = &T[(size_t)tpos[q] * 18];
        for (int r = 0; r < 6; ++r) acc[r] += t[r * 3] * x[3 * i] + t[r * 3 + 1] * x[3 * i + 1] + t[r * 3 + 2] * x[3 * i + 2];
      }
      for (int r = 0; r < 6; ++r) Mx[(size_t)k * 6 + r] = acc[r];
    }
  };
  // rowMultiply (:83-86): res += H_i Mx
  auto rowMultiply = [&](int i, double res[3]) {
    for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
      const double* h = H.block(p);
      const double* m = &Mx[(size_t)H.colidx[p] * 6];
      for (int r = 0; r < 3; ++r) {
        double acc = 0;
        for (int c = 0; c < 6; ++c) acc += h[r * 6 + c] * m[c];
        res[r] += acc;
      }
    }
  };
  const Norm norm = o.infNorm ? Infinity : MeanSquare;
  Vec x(xio, xio + 3 * (size_t)n), y(3 * (size_t)n), xBest;
  const Vec& w = P.b;  // solve(): the linear-constraint operator is Zero, so w = b
  double errBest = std::numeric_limits<double>::max();
  unsigned spmvs = 0;
  auto evalAndKeepBest = [&](const Vec& xx) {
    dmtMultiply(xx);
    ++spmvs;
    for (int i = 0; i < n; ++i) {
      double yi[3] = {w[3 * i], w[3 * i + 1], w[3 * i + 2]};
      rowMultiply(i, yi);
      for (int r = 0; r < 3; ++r) y[3 * i + r] = yi[r];
    }
    const double err = evalResidual(n, coulomb, P.mu.data(), scaling.data(), y.data(), xx.data(), norm);
    if (err < errBest) {
      xBest = xx;
      errBest = err;
    }
    return err;
  };
  evalAndKeepBest(x);
  if (o.tryZero) {  // GaussSeidelBase::tryZero (GaussSeidelBase.impl.hpp:61-77); true: Mx = 0
    Vec zero(x.size(), 0.0);
    const double errZero = evalResidual(n, coulomb, P.mu.data(), scaling.data(), P.b.data(), zero.data(), norm);
    if (errZero < errBest) {
      errBest = errZero;
      xBest = zero;
      x = zero;
      std::fill(Mx.begin(), Mx.end(), 0.0);
    } else {
      x = xBest;
    }
  }
  if (st) st->history.push_back(errBest);
  std::vector<unsigned char> skip(n, 0);
  double ndxRef = 0;
  const double absSkipTol = std::min(o.skipTol, o.tol);
  const double absSkipIters = std::min(o.skipIters, (unsigned)std::sqrt((double)n));
  unsigned it;
  for (it = 1; it <= o.maxIters; ++it) {
    for (size_t c = 0; c + 1 < colors.size(); ++c) {
      const double ndxRefColor = ndxRef;
      double ndxRefNew = ndxRef;
      for (int i = colors[c]; i < colors[c + 1]; ++i) {
        if (skip[i]) {
          --skip[i];
          continue;
        }
        double lx[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        const double* A = &local[(size_t)i * 9];
        double lb[3];
        for (int r = 0; r < 3; ++r) lb[r] = w[3 * i + r] - (A[r * 3] * lx[0] + A[r * 3 + 1] * lx[1] + A[r * 3 + 2] * lx[2]);
        rowMultiply(i, lb);
        double ldx[3] = {-lx[0], -lx[1], -lx[2]};
        const double(*A3)[3] = reinterpret_cast<const double(*)[3]>(A);
        const double lres = localSolve(coulomb, strat, A3, lb, lx, P.mu[i], localTol, scaling[i]);
        const bool ok = localTol > lres;
        for (int k = 0; k < 3; ++k) ldx[k] += lx[k];
        if (!ok)
          for (int k = 0; k < 3; ++k) ldx[k] *= .5;
        for (int k = 0; k < 3; ++k) x[3 * i + k] += ldx[k];
        // colMultiply (:79-81): Mx += (D H^T)_{:, i} dx
        for (int p = H.rowptr[i]; p < H.rowptr[i + 1]; ++p) {
          const double* t = &T[(size_t)p * 18];
          double* m = &Mx[(size_t)H.colidx[p] * 6];
          for (int r = 0; r < 6; ++r) m[r] += t[r * 3] * ldx[0] + t[r * 3 + 1] * ldx[1] + t[r * 3 + 2] * ldx[2];
        }
        const double s2 = scaling[i] * scaling[i];
        const double nx2 = s2 * (lx[0] * lx[0] + lx[1] * lx[1] + lx[2] * lx[2]);
        const double ndx2 = s2 * (ldx[0] * ldx[0] + ldx[1] * ldx[1] + ldx[2] * ldx[2]);
        double ref;
        if (o.ndxRefPerColor) {
          if (ndx2 > ndxRefNew) ndxRefNew = ndx2;
          ref = ndxRefColor;
        } else {
          if (ndx2 > ndxRef) ndxRef = ndx2;
          ref = ndxRef;
        }
        if (std::min(nx2, ndx2) < absSkipTol || ndx2 < o.skipTol * std::min(nx2, ref))
          skip[i] = (unsigned char)absSkipIters;
      }
      if (o.ndxRefPerColor) ndxRef = ndxRefNew;
    }
    if (0 == (it % o.evalEvery)) {
      const double err = evalAndKeepBest(x);
      if (st) st->history.push_back(err);
      if (err < o.tol) break;
      ndxRef /= o.evalEvery;
    }
  }
  if (it > o.maxIters) x = xBest;
  std::copy(x.begin(), x.end(), xio);
  if (st) {
    st->iterations = std::min(it, o.maxIters);
    st->spmvs = spmvs;
    st->residual = errBest;
  }
  return errBest;
}

// Cadoux fixed point (src/Interfaces/Cadoux.hpp:44-91; DualFrictionProblem::solveCadoux,
// FrictionProblem.cpp:68-83, FrictionProblem.impl.hpp:76-91): a sequence of SOC problems
// whose affine term carries the De Saxce correction of the previous iterate,
// b_k = b + s(W r_k + b), each solved by the inner solver (warm-started) at a tolerance
// tightened to tolTighten * max(tol, min(res, 1)); res is the Coulomb residual of the
// iterate with the inner solver's own scaling. velocity = true is solveCadouxVel
// (Cadoux.hpp:95-155): the problem read as r = W u + b, the unknown u shifted by
// s_N = |u_T| max(0, 1/mu).
double cadouxSolve(const Problem& P, const CadouxOptions& o, const std::vector<int>& colors, double* xio,
                   SolveStats* st) {
  const int n = P.nc;
  const size_t n3 = 3 * (size_t)n;
  Problem Q = P;  // its b is replaced by the inner problems' affine terms
  // the minimizer's scaling (used by its eval): GS from the local matrices, PG by updateScalings
  std::vector<double> scaling(n, 1.0);
  const bool useDiag = o.useGS || o.pg.useAssembled || o.pg.diagScaling;
  if (useDiag)
    for (int i = 0; i < n; ++i) {
      const double* d = &P.diag[(size_t)i * 9];
      scaling[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3);
    }
  const bool assembled = o.useGS || o.pg.useAssembled;
  const Norm norm = (o.useGS ? o.gs.infNorm : o.pg.infNorm) ? Infinity : MeanSquare;
  const double tol = o.useGS ? o.gs.tol : o.pg.tol;
  auto W = [&](const Vec& x, Vec& y) { applyW(P, assembled, x.data(), y.data()); };
  unsigned spmvs = 0;
  auto inner = [&](Vec& x, double innerTol) {
    SolveStats is;
    if (o.useGS) {
      GSOptions g = o.gs;
      g.tol = innerTol;
      g.coulomb = 0;
      gsSolve(Q, g, colors, x.data(), &is);
    } else {
      PGOptions g = o.pg;
      g.tol = innerTol;
      g.coulomb = 0;
      pgSolve(Q, g, x.data(), &is);
    }
    spmvs += is.spmvs;
  };
  Vec x(xio, xio + n3), s(n3), y(n3);
  double res;
  unsigned it = 0;
  if (!o.velocity) {
    // s = W r + b; res = eval(coulomb, s, r)
    W(x, s);
    ++spmvs;
    for (size_t k = 0; k < n3; ++k) s[k] += P.b[k];
    res = evalResidual(n, true, P.mu.data(), scaling.data(), s.data(), x.data(), norm);
    if (st) st->history.push_back(res);
    for (it = 0; it < o.iterations; ++it) {
      const double innerTol = o.tolTighten * std::max(tol, std::min(res, 1.));
      for (int i = 0; i < n; ++i) {  // dualityCOV(coulomb, s, s); s += b
        double c[3];
        lawCOV(true, P.mu[i], &s[3 * i], c);
        for (int k = 0; k < 3; ++k) Q.b[3 * i + k] = c[k] + P.b[3 * i + k];
      }
      inner(x, innerTol);
      W(x, s);
      ++spmvs;
      for (size_t k = 0; k < n3; ++k) s[k] += P.b[k];
      res = evalResidual(n, true, P.mu.data(), scaling.data(), s.data(), x.data(), norm);
      if (st) st->history.push_back(res);
      if (res < tol) {
        ++it;
        break;
      }
    }
  } else {
    // x holds u; ustar = u + s, s_N = |u_T| max(0, 1/mu)
    auto shift = [&](const Vec& v, Vec& sv) {
      for (int i = 0; i < n; ++i) {
        sv[3 * i] = std::sqrt(v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]) * std::max(0., 1. / P.mu[i]);
        sv[3 * i + 1] = sv[3 * i + 2] = 0.0;
      }
    };
    Vec u = x, ustar(n3), Ws(n3);
    shift(u, s);
    for (size_t k = 0; k < n3; ++k) ustar[k] = u[k] + s[k];
    W(u, y);
    ++spmvs;
    for (size_t k = 0; k < n3; ++k) y[k] += P.b[k];
    res = evalResidual(n, false, P.mu.data(), scaling.data(), y.data(), ustar.data(), norm);
    if (st) st->history.push_back(res);
    for (it = 0; it < o.iterations; ++it) {
      const double innerTol = o.tolTighten * std::max(tol, std::min(res, 1.));
      W(s, Ws);  // Wsb = b - W s
      ++spmvs;
      for (size_t k = 0; k < n3; ++k) Q.b[k] = P.b[k] - Ws[k];
      inner(ustar, innerTol);
      for (size_t k = 0; k < n3; ++k) u[k] = ustar[k] - s[k];
      shift(ustar, s);
      for (size_t k = 0; k < n3; ++k) ustar[k] = u[k] + s[k];
      W(u, y);
      ++spmvs;
      for (size_t k = 0; k < n3; ++k) y[k] += P.b[k];
      res = evalResidual(n, false, P.mu.data(), scaling.data(), y.data(), ustar.data(), norm);
      if (st) st->history.push_back(res);
      if (it > 0 && res < tol) {
        ++it;
        break;
      }
    }
    x = u;
  }
  std::copy(x.begin(), x.end(), xio);
  if (st) {
    st->iterations = it;
    st->spmvs = spmvs;
    st->residual = res;
  }
  return res;
}

}  // namespace cpuref
