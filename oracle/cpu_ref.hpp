// cpu_ref — FP64 CPU restatement of the So-bogus Coulomb-friction solve path.
//
// TEST INFRASTRUCTURE ONLY. This is the correctness oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
// to check results (or as the timed CPU baseline), never as the thing measured or
// shipped. The product path (so-bogus_amd/) never links or calls it.
//
// Every routine cites the reference file:line it restates. Reference paths are
// relative to the bogus 1.4.2 source tree (src/...). The reference itself cannot
// be built in this environment (it hard-requires Eigen >= 5, which is absent),
// so this file re-derives Eigen's dense semantics where the path depends on them:
// complete-pivoting LU (FullPivLU) and the real Schur eigenvalues of a 4x4
// companion matrix (EigenSolver).
//
// Pinning (what each part is checked against) is listed in oracle/README.md.
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace cpuref {

// ---------------------------------------------------------------- dense helpers
// Complete-pivoting LU of an N x N matrix, stated after Eigen's FullPivLU
// semantics used by src/Core/Eigen/EigenLinearSolvers.hpp:36-63 (LU<> wrapper),
// src/Extra/SOC/LocalSOCSolver.impl.hpp:97,294 and
// src/Core/Utils/NonSmoothNewton.impl.hpp:55,69.
template <int N>
struct FullPivLU {
  double lu[N][N];
  int rowTransp[N];  // row transposition applied at step k
  int colTransp[N];  // col transposition applied at step k
  int rank = 0;
  int nonzeroPivots = 0;
  double maxPivot = 0;
  void compute(const double A[N][N]);
  // x = A^{-1} b in the FullPivLU sense (free unknowns set to zero if rank < N)
  void solve(const double b[N], double x[N]) const;
};

// Real roots of the monic polynomial x^D + c[D-1] x^{D-1} + ... + c[0]
// via eigenvalues of its companion matrix (src/Core/Utils/Polynomial.impl.hpp:56-75).
enum RootFilter { StrictlyPositiveRoots = 0, StrictlyNegativeRoots = 1, AllRoots = 2 };
unsigned realRootsMonic(int D, const double* coeffs, double* roots, RootFilter filter);
// Arbitrary leading coefficient; coeffs has D+1 entries (Polynomial.impl.hpp:88-109).
unsigned realRootsDegenerate(int D, double* coeffs, double* roots, RootFilter filter);
// Eigenvalues of a small real upper-Hessenberg matrix (n <= 4), row-major.
void hessenbergEigenvalues(int n, const double* H, double* wr, double* wi);

// ---------------------------------------------------------------- SOC / Coulomb law
// Fischer-Burmeister function on the cone of aperture mu
// (src/Extra/SOC/FischerBurmeister.impl.hpp:29-42, 69-101).
void fbBase(double mu, const double x[3], const double y[3], double fb[3]);
void fbBaseJacobian(double mu, const double x[3], const double y[3], double fb[3], double dx[3][3],
                    double dy[3][3]);
// FischerBurmeister<3,double,DeSaxceCOV>::compute (FischerBurmeister.impl.hpp:142-152)
void fbLaw(bool coulomb, double mu, const double x[3], const double y[3], double fb[3]);
// SOCLaw::eval (src/Extra/SOC/SOCLaw.hpp:55-64)
double lawEval(bool coulomb, double mu, const double x[3], const double y[3]);
// SOCLaw::projectOnConstraint (src/Extra/SOC/SOCLaw.impl.hpp:46-72)
void lawProject(double mu, double x[3]);
// SOCLaw::dualityCOV (SOCLaw.hpp:92-99)
void lawCOV(bool coulomb, double mu, const double y[3], double s[3]);

// Local solver strategies (src/Extra/SecondOrder.fwd.hpp:29-55)
enum Strategy { PureNewton = 0, PureEnumerative = 1, Hybrid = 2, RevHybrid = 3 };
// LocalSOCSolver<3,double,DeSaxceCOV,Strat>::solve (src/Extra/SOC/LocalSOCSolver.impl.hpp:243-318).
// Returns the local residual Phi; x is warm start and result.
double localSolve(bool coulomb, Strategy strat, const double A[3][3], const double b[3], double x[3], double mu,
                  double tol, double scaling);
// AnalyticLocalSOCSolver<3,Scalar,true|false>::solveOrthogonality (LocalSOCSolver.impl.hpp:47-103, 141-210)
bool solveOrthogonality(bool coulomb, const double W[3][3], const double b[3], double r[3], double mu);
// NonSmoothNewton<FischerBurmeister>::solve (src/Core/Utils/NonSmoothNewton.impl.hpp:23-79)
double nonSmoothNewton(bool coulomb, const double A[3][3], const double b[3], double mu, double scaling, double tol,
                       unsigned maxIters, double x[3]);

// ---------------------------------------------------------------- sparse block matrices
// Block-row compressed storage (the reference's compressed major index:
// src/Core/Block/CompressedSparseBlockIndex.hpp:19-144). Blocks are dense
// row-major br x bc. If symLower, only blocks with col <= row are stored and
// the matrix is symmetric (flags::SYMMETRIC, src/Core/Block/Constants.hpp:60-65).
struct BSR {
  int nbr = 0, nbc = 0, br = 0, bc = 0;
  bool symLower = false;
  std::vector<int> rowptr, colidx;
  std::vector<double> vals;
  int nnzb() const { return rowptr.empty() ? 0 : rowptr.back(); }
  const double* block(int p) const { return &vals[(size_t)p * br * bc]; }
};

// res = alpha * op(A) * x + beta * res, for general and symmetric-lower BSR
// (src/Core/Block/SparseMatrixVectorProduct.impl.hpp:69-175, 177-260, 382-432).
void bsrMultiply(const BSR& A, bool transpose, const double* x, double* res, double alpha, double beta);
// res += sum_{j != row} A_{row,j} x_j (symmetric splitRowMultiply, SparseMatrixVectorProduct.impl.hpp:298-320)
void bsrSplitRowMultiply(const BSR& A, int row, const double* x, double res[3]);
// Builds the transpose lookup needed by symmetric splitRowMultiply/SpMV gathers
// (the role of cacheTranspose, src/Core/Block/SparseBlockMatrixBase.impl.hpp:169-202).
void bsrBuildUpperIndex(const BSR& A, std::vector<int>& uptr, std::vector<int>& urow, std::vector<int>& upos);

// ---------------------------------------------------------------- friction problem
// The dual problem u = W r + b with W = H Minv H^T (src/Interfaces/FrictionProblem.hpp:127-214).
struct Problem {
  int nc = 0, nb = 0;
  BSR H;                    // nc x nb blocks of 3x6
  std::vector<double> minv; // nb blocks of 6x6 row-major (explicit inverse mass)
  std::vector<double> b, mu;
  bool hasW = false;
  BSR W;                    // assembled, symmetric lower, 3x3 blocks
  std::vector<int> wUptr, wUrow, wUpos; // transposed (upper) lookup of W
  std::vector<double> diag; // nc blocks of 3x3: W_ii = sum_k H_ik Minv_k H_ik^T
};

// DualFrictionProblem::computeFrom: W = H * (MInv * H^T), lower triangle
// (src/Interfaces/FrictionProblem.cpp:30-40, src/Core/Block/SparseMatrixMatrixProduct.impl.hpp:353-419).
void assembleW(Problem& P);
void computeDiag(Problem& P);
// W x (assembled symmetric if P.hasW && useAssembled, else matrix-free H (Minv (H^T x)))
void applyW(const Problem& P, bool useAssembled, const double* x, double* y);

enum Norm { MeanSquare = 0, Infinity = 1 };
// ConstrainedSolverBase::eval (src/Core/BlockSolvers/ConstrainedSolverBase.impl.hpp:21-74)
double evalResidual(int n, bool coulomb, const double* mu, const double* scaling, const double* y, const double* x,
                    Norm norm);

// ---------------------------------------------------------------- solvers
enum PGVariant { PG_Standard = 0, PG_Descent = 1, PG_Conjugated = 2, PG_APGD = 3, PG_SPG = 4 };

struct PGOptions {
  int variant = PG_APGD;
  double tol = 1e-6;
  unsigned maxIters = 300;
  unsigned lsIters = 8;
  double lsOptimistic = 1.25;
  double lsPessimistic = 0.5;
  double lsArmijo = 0.5;
  int coulomb = 0;        // 0: SOC law (staticProblem=true), 1: Coulomb (De Saxce COV)
  int useAssembled = 1;   // 1: operator is assembled W; 0: matrix-free H Minv H^T
  // Row scaling s_i of the residual. 0 (reference): max(1, tr(W_ii)/3) for the
  // assembled W, 1 for the matrix-free product (ConstrainedSolverBase.impl.hpp:83-104).
  // 1: max(1, tr(W_ii)/3) in both modes, W_ii computed from H Minv H^T.
  int diagScaling = 0;
  int infNorm = 0;
};

struct SolveStats {
  unsigned iterations = 0;
  unsigned spmvs = 0;
  double residual = 0;
  std::vector<double> history;  // residual per evaluation (the callback stream)
};

// ProjectedGradient::solve (src/Core/BlockSolvers/ProjectedGradient.impl.hpp:74-382)
double pgSolve(const Problem& P, const PGOptions& o, double* x, SolveStats* st);

struct GSOptions {
  double tol = 1e-6;
  unsigned maxIters = 250;
  unsigned evalEvery = 25;
  double skipTol = 1e-6;
  unsigned skipIters = 10;
  double autoRegularization = 0;
  int tryZero = 1;
  int coulomb = 1;
  int strategy = RevHybrid;
  int infNorm = 0;
  // 0: reference semantics (ndxRef updated in place, visible to later contacts);
  // 1: ndxRef snapshot per color (deterministic under parallel color sweeps).
  int ndxRefPerColor = 0;
};

// Coloring::compute (src/Core/BlockSolvers/Coloring.impl.hpp:20-62) on the lower
// index of symmetric W. perm lists contacts color by color; colors has #colors+1 offsets.
void computeColoring(const BSR& Wlower, std::vector<int>& perm, std::vector<int>& colors);

// GaussSeidel::solve (src/Core/BlockSolvers/GaussSeidel.impl.hpp:24-205,
// GaussSeidelBase.impl.hpp:20-76). colors: offsets into the (identity) contact order;
// pass {0, nc} for a single color.
double gsSolve(const Problem& P, const GSOptions& o, const std::vector<int>& colors, double* x, SolveStats* st);

// ProductGaussSeidel::solve with D = Minv, D H^T precomputed
// (src/Core/BlockSolvers/ProductGaussSeidel.impl.hpp:64-86, 128-304): the matrix-free
// Gauss-Seidel; W is never used. Same options and color handling as gsSolve.
double pgsSolve(const Problem& P, const GSOptions& o, const std::vector<int>& colors, double* x, SolveStats* st);

struct CadouxOptions {
  int useGS = 0;           // inner solver: GaussSeidel (gs, colors) or ProjectedGradient (pg)
  GSOptions gs;
  PGOptions pg;
  unsigned iterations = 10;
  double tolTighten = 1e-1;
  int velocity = 0;        // 0: solveCadoux (unknown r), 1: solveCadouxVel (unknown u, r = W u + b)
};
// Cadoux fixed point over SOC solves (src/Interfaces/Cadoux.hpp:44-155). The inner
// solver's law is SOC whatever its options say; st->iterations = fixed-point iterations,
// st->history = the residual before and after each.
double cadouxSolve(const Problem& P, const CadouxOptions& o, const std::vector<int>& colors, double* x,
                   SolveStats* st);

// bogus::applyPermutation (src/Core/Block/SparseBlockMatrixBase.impl.hpp:309-326): new[i] = old[perm[i]]
// and DualFrictionProblem::applyPermutation (src/Interfaces/FrictionProblem.cpp:102-114).
void permuteProblem(Problem& P, const std::vector<int>& perm);

}  // namespace cpuref
