// The reference's own solver tests (bogus tests/GaussSeidel.cpp:19-65,
// tests/ProjectedGradient.cpp:15-83, tests/ProductGaussSeidel.cpp:16-62,
// tests/BlockMV.cpp:14-245), written against
// the C++ facade bogus_hip.hpp so that they read like the reference's, and run
// on the GPU path. Exit status 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <vector>

#include "bogus_hip.hpp"

static int g_fail = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);   \
      ++g_fail;                                                       \
    }                                                                 \
  } while (0)

static bool isApprox(const std::vector<double>& a, const std::vector<double>& b, double prec) {
  double d = 0, na = 0, nb = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    d += (a[i] - b[i]) * (a[i] - b[i]);
    na += a[i] * a[i];
    nb += b[i] * b[i];
  }
  return std::sqrt(d) <= prec * std::sqrt(std::min(na, nb));
}

// tests/SmallFrictionPb.hpp:16-85 embedded into 6-dof bodies (zero-padded H
// columns, identity-padded masses: W is unchanged).
struct SmallFrictionPb {
  bogus::PrimalFrictionProblem<3> primal;
  std::vector<double> f, w, mu, sol;
  SmallFrictionPb() {
    primal.M.assign(72, 0.0);
    for (int i = 0; i < 6; ++i) primal.M[i * 6 + i] = 1.0;
    for (int i = 0; i < 6; ++i) primal.M[36 + i * 6 + i] = (i < 2) ? 2.0 : 1.0;
    primal.H.setRows(2, 3);
    primal.H.setCols(2, 6);
    const double h00[12] = {3, 5, 7, -1, 2, 4, 5, 4, 1, 6, 9, 1};
    const double h01[6] = {-3, -6, -6, -5, -5, -8};
    const double h10[12] = {3, 7, 1, -1, 3, 6, 3, 2, 4, 4, 7, -1};
    const double h2b[12] = {6, 7, 10, 9, 5, 8, 11, 8, 4, 9, 12, 7};
    double* b00 = primal.H.insertBack(0, 0);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) b00[r * 6 + c] = h00[r * 4 + c];
    double* b01 = primal.H.insertBack(0, 1);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 2; ++c) b01[r * 6 + c] = h01[r * 2 + c];
    // H.block(2) = E^T (H10 - H2B), E = [[0,0,1],[1,0,0],[0,1,0]]
    double d[3][4];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) d[r][c] = h10[r * 4 + c] - h2b[r * 4 + c];
    const int et[3] = {1, 2, 0};  // row r of E^T D is row et[r] of D
    double* b10 = primal.H.insertBack(1, 0);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 4; ++c) b10[r * 6 + c] = d[et[r]][c];
    primal.H.finalize();
    f = {1, 2, 3, 4, 0, 0, 5, 6, 0, 0, 0, 0};
    w = {2, 1, 2, 1, 3, 3};
    mu = {0.5, 0.7};
    sol = {0.0152695, 0.0073010, 0.0022325, 0.0, 0.0, 0.0};
    primal.f = f.data();
    primal.w = w.data();
    primal.mu = mu.data();
    primal.computeMInv();
  }
};

static void testGaussSeidel() {
  SmallFrictionPb pb;
  bogus::DualFrictionProblem<3> dual;
  dual.computeFrom(pb.primal);
  bogus::GaussSeidel gs;
  const bogus::local_soc_solver::Strategy strats[4] = {
      bogus::local_soc_solver::Hybrid, bogus::local_soc_solver::PureNewton, bogus::local_soc_solver::PureEnumerative,
      bogus::local_soc_solver::RevHybrid};
  for (auto s : strats) {
    std::vector<double> x(6, 1.0);
    gs.setLocalStrategy(s);
    const double res = dual.solveWith(gs, x.data(), false);
    std::printf("GS strategy %d: residual %.3e x = %.7f %.7f %.7f %.7f %.7f %.7f\n", (int)s, res, x[0], x[1], x[2],
                x[3], x[4], x[5]);
    CHECK(res < 1.e-8);
    CHECK(isApprox(pb.sol, x, 1.e-4));
  }
  std::vector<double> x(6, 1.0);
  gs.setTol(1.e-8);
  const double res = dual.solveWith(gs, x.data(), true);  // SOC3D
  CHECK(res < 1.e-8);
  // evalWith returns the residual of a given r
  CHECK(dual.evalWith(gs, x.data(), true) < 1.e-8);
}

static void testProjectedGradient() {
  SmallFrictionPb pb;
  bogus::DualFrictionProblem<3> dual;
  dual.computeFrom(pb.primal);
  bogus::ProjectedGradient pg;
  pg.setTol(1.e-8);
  unsigned ncb = 0;
  pg.callback().connect([&](unsigned, double) { ++ncb; });
  // every variant of tests/ProjectedGradient.cpp:31-60, in the reference test's order
  const bogus::projected_gradient::Variant vs[5] = {
      bogus::projected_gradient::APGD, bogus::projected_gradient::Descent, bogus::projected_gradient::Standard,
      bogus::projected_gradient::Conjugated, bogus::projected_gradient::SPG};
  for (auto v : vs) {
    std::vector<double> x(6, 1.0);
    pg.setDefaultVariant(v);
    const double res = dual.solveWith(pg, x.data());  // staticProblem = true: SOC3D, as the reference test
    std::printf("PG variant %d: residual %.3e\n", (int)v, res);
    CHECK(res < 1.e-8);
  }
  CHECK(ncb > 0);
  // matrix-free product H Minv H^T (tests/ProjectedGradient.cpp:63-83)
  std::vector<double> x(6, 1.0);
  pg.setDefaultVariant(bogus::projected_gradient::APGD);
  pg.setMatrixFree(true);
  const double res = dual.solveWith(pg, x.data());
  CHECK(res < 1.e-8);
}

// tests/ProductGaussSeidel.cpp:44-61 (Product_GS_D_precomp) through
// PrimalFrictionProblem::solveWith(pgs, r) (PrimalFrictionProblem.cpp:40-59): no W.
static void testProductGaussSeidel() {
  SmallFrictionPb pb;
  bogus::PrimalFrictionProblem<3>::ProductGaussSeidelType pgs;
  unsigned ncb = 0;
  pgs.callback().connect([&](unsigned, double) { ++ncb; });
  for (int mode : {SB_COLORING_GREEDY, SB_COLORING_REFERENCE}) {
    std::vector<double> x(6, 1.0);
    pgs.coloring().mode = mode;
    const double res = pb.primal.solveWith(pgs, x.data());  // staticProblem = false: Coulomb
    std::printf("Product GS coloring %d: residual %.3e x = %.7f %.7f %.7f %.7f %.7f %.7f\n", mode, res, x[0], x[1],
                x[2], x[3], x[4], x[5]);
    CHECK(res < 1.e-8);
    CHECK(isApprox(pb.sol, x, 1.e-4));
  }
  CHECK(ncb > 0);
  // the same solution through the dual problem's Gauss-Seidel on the assembled W
  bogus::DualFrictionProblem<3> dual;
  dual.computeFrom(pb.primal);
  std::vector<double> x(6, 1.0);
  bogus::GaussSeidel gs;
  dual.solveWith(gs, x.data(), false);
  std::vector<double> y(6, 1.0);
  pb.primal.solveWith(pgs, y.data());
  CHECK(isApprox(x, y, 1.e-10));
}

// DualFrictionProblem::solveCadoux / solveCadouxVel (FrictionProblem.hpp:177-195) on the
// two-contact fixture: the fixed point is the Coulomb solution `sol`.
static void testCadoux() {
  SmallFrictionPb pb;
  bogus::DualFrictionProblem<3> dual;
  dual.computeFrom(pb.primal);
  bogus::GaussSeidel gs;
  gs.setTol(1.e-8);
  bogus::ProjectedGradient pg;
  pg.setTol(1.e-8);
  pg.setMaxIters(1000);
  unsigned ncb = 0;
  bogus::Signal<unsigned, double> cb;
  cb.connect([&](unsigned, double) { ++ncb; });
  std::vector<double> x(6, 0.0);
  double res = dual.solveCadoux(gs, x.data(), 50, &cb);
  std::printf("Cadoux GS: residual %.3e\n", res);
  CHECK(res < 1.e-8);
  CHECK(isApprox(pb.sol, x, 1.e-4));
  CHECK(ncb >= 2);
  std::fill(x.begin(), x.end(), 0.0);
  res = dual.solveCadoux(pg, x.data(), 50);
  std::printf("Cadoux PG: residual %.3e\n", res);
  CHECK(res < 1.e-8);
  CHECK(isApprox(pb.sol, x, 1.e-4));
  std::vector<double> u(6, 0.0);
  res = dual.solveCadouxVel(gs, u.data(), 50);
  std::printf("CadouxVel GS: residual %.3e\n", res);
  CHECK(res < 1.e-8);
}

static void testBlockMV() {
  // tests/BlockMV.cpp:14-55 SmallSBM
  bogus::SparseBlockMatrix sbm;
  sbm.setRows(5, 3);
  sbm.setCols(2, 4);
  double* b0 = sbm.insertBack(0, 1);
  for (int k = 0; k < 12; ++k) b0[k] = 1;
  double* b1 = sbm.insertBack(3, 0);
  for (int k = 0; k < 12; ++k) b1[k] = 2;
  double* b2 = sbm.insertBack(3, 1);
  for (int k = 0; k < 12; ++k) b2[k] = 3;
  sbm.finalize();
  std::vector<double> rhs(8, 1.0), res, rhs2;
  sbm.multiply<false>(rhs, res);
  const double e1[15] = {4, 4, 4, 0, 0, 0, 0, 0, 0, 20, 20, 20, 0, 0, 0};
  const double e2[8] = {120, 120, 120, 120, 192, 192, 192, 192};
  for (int k = 0; k < 15; ++k) CHECK(res[k] == e1[k]);
  sbm.multiply<true>(res, rhs2);
  for (int k = 0; k < 8; ++k) CHECK(rhs2[k] == e2[k]);
  // tests/BlockMV.cpp:192-245 symmetric
  bogus::SparseBlockMatrix ssbm(bogus::flags::SYMMETRIC);
  ssbm.setRows(3, 3);
  double* s0 = ssbm.insertBack(0, 0);
  for (int k = 0; k < 9; ++k) s0[k] = 1;
  double* s1 = ssbm.insertBack(1, 0);
  const double sec[9] = {2, 0, 0, 2, 2, 0, 2, 2, 2};
  for (int k = 0; k < 9; ++k) s1[k] = sec[k];
  double* s2 = ssbm.insertBack(2, 2);
  for (int k = 0; k < 9; ++k) s2[k] = 3;
  ssbm.finalize();
  std::vector<double> ones(9, 1.0), out;
  ssbm.multiply<false>(ones, out);
  const double se1[9] = {9, 7, 5, 2, 4, 6, 9, 9, 9};
  for (int k = 0; k < 9; ++k) CHECK(out[k] == se1[k]);
  const double se2[9] = {6, 4, 2, 2, 4, 6, 0, 0, 0};
  for (int row = 0; row < 3; ++row) {
    double seg[3] = {0, 0, 0};
    ssbm.splitRowMultiply(row, ones.data(), seg);
    for (int k = 0; k < 3; ++k) CHECK(seg[k] == se2[3 * row + k]);
  }
}

// The reference's two-contact problem given by its W (apps/FCLibLoader.cpp:64-73, the
// FCLib application's set-up): W = H MInv H^T and b = w - H MInv f computed here with plain
// loops, written as a scalar CSR with both triangles, then fromLocalProblem. The known
// Coulomb solution of tests/GaussSeidel.cpp:19-65 and SOC solves of every PG variant.
static void testLocalProblem() {
  SmallFrictionPb pb;
  const auto& H = pb.primal.H;
  const std::vector<double>& Mi = pb.primal.MInv;
  const int nc = 2;
  // G = H MInv per block, then W = G H^T (dense 6 x 6)
  double W[6][6] = {};
  for (int i = 0; i < nc; ++i)
    for (int p = H.rowPointers()[i]; p < H.rowPointers()[i + 1]; ++p)
      for (int j = 0; j < nc; ++j)
        for (int q = H.rowPointers()[j]; q < H.rowPointers()[j + 1]; ++q) {
          if (H.columnIndices()[p] != H.columnIndices()[q]) continue;
          const double* A = &H.blockValues()[(size_t)p * 18];
          const double* B = &H.blockValues()[(size_t)q * 18];
          const double* M = &Mi[(size_t)H.columnIndices()[p] * 36];
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
              double acc = 0;
              for (int k = 0; k < 6; ++k)
                for (int l = 0; l < 6; ++l) acc += A[r * 6 + k] * M[k * 6 + l] * B[c * 6 + l];
              W[3 * i + r][3 * j + c] += acc;
            }
        }
  std::vector<int> rp(1, 0), ci;
  std::vector<double> v;
  for (int r = 0; r < 6; ++r) {
    for (int c = 0; c < 6; ++c)
      if (W[r][c] != 0.0) {
        ci.push_back(c);
        v.push_back(W[r][c]);
      }
    rp.push_back((int)ci.size());
  }
  const std::vector<double> q = pb.primal.dualRhs();
  bogus::DualFrictionProblem<3> dual;
  dual.fromLocalProblem(nc, rp.data(), ci.data(), v.data(), q.data(), pb.mu.data(), 1e-14);
  bogus::GaussSeidel gs;
  std::vector<double> x(6, 1.0);
  const double res = dual.solveWith(gs, x.data(), false);
  std::printf("local problem GS: residual %.3e x = %.7f %.7f %.7f %.7f %.7f %.7f merit1 %.3e\n", res, x[0], x[1],
              x[2], x[3], x[4], x[5], dual.merit1(x.data()));
  CHECK(res < 1.e-8);
  CHECK(isApprox(pb.sol, x, 1.e-4));
  CHECK(dual.merit1(x.data()) < 1.e-6);
  bogus::ProjectedGradient pg;
  pg.setTol(1.e-8);
  for (auto vnt : {bogus::projected_gradient::APGD, bogus::projected_gradient::SPG,
                   bogus::projected_gradient::Conjugated}) {
    std::vector<double> y(6, 1.0);
    pg.setDefaultVariant(vnt);
    CHECK(dual.solveWith(pg, y.data()) < 1.e-8);
  }
}

int main() {
  try {
    testBlockMV();
    testGaussSeidel();
    testProjectedGradient();
    testProductGaussSeidel();
    testCadoux();
    testLocalProblem();
  } catch (const std::exception& e) {
    std::printf("EXCEPTION %s\n", e.what());
    return 2;
  }
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail ? 1 : 0;
}
