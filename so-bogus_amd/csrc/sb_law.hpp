// Per-contact second-order-cone / Coulomb law kernels (device, FP64 or FP32).
//
//  * project  : SOCLaw::projectOnConstraint      (src/Extra/SOC/SOCLaw.impl.hpp:46-72)
//  * cov      : SOCLaw::dualityCOV (De Saxce)    (src/Extra/SOC/SOCLaw.hpp:92-99)
//  * law_eval : SOCLaw::eval = |FB_mu(x, y [+ mu|y_T| e_N])|^2
//               (src/Extra/SOC/SOCLaw.hpp:55-64, src/Extra/SOC/FischerBurmeister.impl.hpp:29-42, 69-101, 142-152)
//
// Vectors are (normal, tangent1, tangent2). mu < 0 disables the contact.
#pragma once
#include <hip/hip_runtime.h>

namespace sb {

template <typename T>
struct LawConst;
template <>
struct LawConst<double> {
  static constexpr double eps = 2.220446049250313e-16;  // DBL_EPSILON: NumTraits<double>::isZero threshold
};
template <>
struct LawConst<float> {
  static constexpr float eps = 1.1920928955078125e-07f;
};

template <typename T>
__device__ __forceinline__ bool is_zero(T s) {
  return fabs(s) < LawConst<T>::eps;
}

// Projection on K_mu = { |x_T| <= mu x_N }.
template <typename T>
__device__ __forceinline__ void project(T mu, T& x0, T& x1, T& x2) {
  const T nxt = sqrt(x1 * x1 + x2 * x2);
  if (mu < T(0)) {
    x0 = x1 = x2 = T(0);
    return;
  }
  if (nxt <= mu * x0) return;  // inside the cone
  if (mu * nxt <= -x0) {       // inside the polar cone
    x0 = x1 = x2 = T(0);
    return;
  }
  const T proj = (x0 + mu * nxt) / (T(1) + mu * mu);
  const T f = proj * mu / nxt;
  x0 = proj;
  x1 *= f;
  x2 *= f;
}

// De Saxce change of variable s(y) = (mu |y_T|, 0, 0) (Coulomb) or 0 (SOC).
template <typename T>
__device__ __forceinline__ T cov_normal(bool coulomb, T mu, T y1, T y2) {
  return coulomb ? mu * sqrt(y1 * y1 + y2 * y2) : T(0);
}

// Fischer-Burmeister on the canonical cone via the spectral decomposition of
// z2 = x o x + y o y (Fukushima 2002; [Daviet et al 2011] App. A.1), returning |fb|^2.
template <typename T>
__device__ __forceinline__ T fb_canonical_sq(const T x[3], const T y[3]) {
  const T z20 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] + y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  const T z21 = x[0] * x[1] + y[0] * y[1];
  const T z22 = x[0] * x[2] + y[0] * y[2];
  const T nz2t = sqrt(z21 * z21 + z22 * z22);
  T d1, d2;  // unit tangent direction of z2
  if (is_zero(nz2t)) {
    d1 = T(1);
    d2 = T(0);
  } else {
    d1 = z21 / nz2t;
    d2 = z22 / nz2t;
  }
  const T rl1 = sqrt(fmax(T(0), z20 - T(2) * nz2t));
  const T rl2 = sqrt(z20 + T(2) * nz2t);
  // z = rl1 * (1/2, -d/2) + rl2 * (1/2, d/2)
  const T z0 = T(0.5) * (rl1 + rl2);
  const T zt = T(0.5) * (rl2 - rl1);
  const T f0 = x[0] + y[0] - z0;
  const T f1 = x[1] + y[1] - zt * d1;
  const T f2 = x[2] + y[2] - zt * d2;
  return f0 * f0 + f1 * f1 + f2 * f2;
}

// SOCLaw<3,T,coulomb>::eval(x, y): x already scaled by the row scaling.
template <typename T>
__device__ __forceinline__ T law_eval(bool coulomb, T mu, const T x[3], const T y[3]) {
  if (mu < T(0)) return y[0] * y[0] + y[1] * y[1] + y[2] * y[2];
  T yt0 = y[0];
  if (coulomb) yt0 += mu * sqrt(y[1] * y[1] + y[2] * y[2]);
  if (is_zero(mu)) {
    const T f0 = x[0] + yt0 - sqrt(x[0] * x[0] + yt0 * yt0);
    return f0 * f0 + x[1] * x[1] + x[2] * x[2];
  }
  const T xh[3] = {mu * x[0], x[1], x[2]};
  const T yh[3] = {yt0, mu * y[1], mu * y[2]};
  return fb_canonical_sq(xh, yh);
}

// Branch of the projection on K_mu taken by project(): 0 inside the cone
// (identity), 1 inside the polar cone or mu < 0 (zero), 2 onto the cone's boundary.
template <typename T>
__device__ __forceinline__ int project_branch(T mu, const T z[3]) {
  if (mu < T(0)) return 1;
  const T t = sqrt(z[1] * z[1] + z[2] * z[2]);
  if (t <= mu * z[0]) return 0;
  if (mu * t <= -z[0]) return 1;
  return 2;
}
// The projection formula of a given branch (same arithmetic as project()).
template <typename T>
__device__ __forceinline__ void project_on_branch(int branch, T mu, T z[3]) {
  if (branch == 0) return;
  const T t = sqrt(z[1] * z[1] + z[2] * z[2]);
  if (branch == 1 || !(t > T(0))) {
    z[0] = z[1] = z[2] = T(0);
    return;
  }
  const T proj = (z[0] + mu * t) / (T(1) + mu * mu);
  const T f = proj * mu / t;
  z[0] = proj;
  z[1] *= f;
  z[2] *= f;
}

// Projection on K_mu with its (generalized) Jacobian J = dP/dz (row-major 3x3)
// and dP/dmu, for the reverse-mode (adjoint) sweeps. The branch is chosen with
// the same tests as project(); on the boundary between branches the derivative
// of the branch taken is returned.
template <typename T>
__device__ __forceinline__ void project_jacobian(T mu, const T z[3], T J[9], T dmu[3], int branch = -1) {
#pragma unroll
  for (int e = 0; e < 9; ++e) J[e] = T(0);
  dmu[0] = dmu[1] = dmu[2] = T(0);
  if (branch < 0) branch = project_branch<T>(mu, z);
  const T t = sqrt(z[1] * z[1] + z[2] * z[2]);
  if (branch == 0) {  // inside the cone: identity
    J[0] = J[4] = J[8] = T(1);
    return;
  }
  if (branch == 1 || !(t > T(0))) return;  // inside the polar cone: zero
  const T den = T(1) + mu * mu;
  const T a = (z[0] + mu * t) / den;
  const T n1 = z[1] / t, n2 = z[2] / t;
  const T c = mu * a / t;
  J[0] = T(1) / den;
  J[1] = mu * n1 / den;
  J[2] = mu * n2 / den;
  J[3] = mu * n1 / den;
  J[4] = mu * mu * n1 * n1 / den + c * (T(1) - n1 * n1);
  J[5] = mu * mu * n1 * n2 / den - c * n1 * n2;
  J[6] = mu * n2 / den;
  J[7] = J[5];
  J[8] = mu * mu * n2 * n2 / den + c * (T(1) - n2 * n2);
  const T da = (t - T(2) * mu * z[0] - mu * mu * t) / (den * den);
  dmu[0] = da;
  dmu[1] = (a + mu * da) * n1;
  dmu[2] = (a + mu * da) * n2;
}

}  // namespace sb
