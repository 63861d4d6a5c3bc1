// C entry points of the oracle, for ctypes (tests/ and bench.py's CPU baseline only).
// TEST INFRASTRUCTURE ONLY: see cpu_ref.hpp.
#include <omp.h>

#include <cstring>

#include "cpu_ref.hpp"

using namespace cpuref;

extern "C" {

void orc_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int orc_get_threads() { return omp_get_max_threads(); }

void* orc_problem_create(int nc, int nb, const int* rowptr, const int* colidx, const double* hvals,
                         const double* minv, const double* b, const double* mu) {
  Problem* P = new Problem();
  P->nc = nc;
  P->nb = nb;
  P->H.nbr = nc;
  P->H.nbc = nb;
  P->H.br = 3;
  P->H.bc = 6;
  P->H.rowptr.assign(rowptr, rowptr + nc + 1);
  const int nnz = rowptr[nc];
  P->H.colidx.assign(colidx, colidx + nnz);
  P->H.vals.assign(hvals, hvals + (size_t)nnz * 18);
  P->minv.assign(minv, minv + (size_t)nb * 36);
  P->b.assign(b, b + 3 * (size_t)nc);
  P->mu.assign(mu, mu + nc);
  computeDiag(*P);
  return P;
}

void orc_problem_destroy(void* p) { delete static_cast<Problem*>(p); }

int orc_problem_assemble_w(void* p) {
  Problem* P = static_cast<Problem*>(p);
  assembleW(*P);
  computeDiag(*P);
  return P->W.nnzb();
}

int orc_w_nnzb(void* p) { return static_cast<Problem*>(p)->W.nnzb(); }

void orc_w_export(void* p, int* rowptr, int* colidx, double* vals) {
  const Problem* P = static_cast<Problem*>(p);
  std::memcpy(rowptr, P->W.rowptr.data(), sizeof(int) * (P->nc + 1));
  std::memcpy(colidx, P->W.colidx.data(), sizeof(int) * P->W.nnzb());
  std::memcpy(vals, P->W.vals.data(), sizeof(double) * 9 * P->W.nnzb());
}

void orc_diag_export(void* p, double* diag) {
  const Problem* P = static_cast<Problem*>(p);
  std::memcpy(diag, P->diag.data(), sizeof(double) * 9 * P->nc);
}

void orc_problem_export(void* p, int* rowptr, int* colidx, double* hvals, double* b, double* mu) {
  const Problem* P = static_cast<Problem*>(p);
  std::memcpy(rowptr, P->H.rowptr.data(), sizeof(int) * (P->nc + 1));
  std::memcpy(colidx, P->H.colidx.data(), sizeof(int) * P->H.nnzb());
  std::memcpy(hvals, P->H.vals.data(), sizeof(double) * 18 * P->H.nnzb());
  std::memcpy(b, P->b.data(), sizeof(double) * 3 * P->nc);
  std::memcpy(mu, P->mu.data(), sizeof(double) * P->nc);
}

void orc_apply_w(void* p, int assembled, const double* x, double* y) {
  applyW(*static_cast<Problem*>(p), assembled != 0, x, y);
}

// DualFrictionProblem::evalWith (FrictionProblem.impl.hpp:58-74): u = W r + b, then eval.
double orc_eval(void* p, int coulomb, int infNorm, int diagScaling, int assembled, const double* r) {
  const Problem* P = static_cast<Problem*>(p);
  std::vector<double> u(3 * (size_t)P->nc), s(P->nc, 1.0);
  applyW(*P, assembled != 0, r, u.data());
  for (size_t k = 0; k < u.size(); ++k) u[k] += P->b[k];
  if (diagScaling)
    for (int i = 0; i < P->nc; ++i) {
      const double* d = &P->diag[(size_t)i * 9];
      s[i] = std::max(1.0, (d[0] + d[4] + d[8]) / 3);
    }
  return evalResidual(P->nc, coulomb != 0, P->mu.data(), s.data(), u.data(), r, infNorm ? Infinity : MeanSquare);
}

double orc_eval_raw(int n, int coulomb, int infNorm, const double* mu, const double* scaling, const double* y,
                    const double* x) {
  return evalResidual(n, coulomb != 0, mu, scaling, y, x, infNorm ? Infinity : MeanSquare);
}

static void copyHistory(const SolveStats& st, double* hist, int cap, int* histLen) {
  if (histLen) *histLen = (int)st.history.size();
  if (hist)
    for (int k = 0; k < cap && k < (int)st.history.size(); ++k) hist[k] = st.history[k];
}

// Line-search parameters of the following PG solves (ProjectedGradient.hpp:87-95 defaults:
// 8 trials, x1.25 optimistic, x0.5 pessimistic, Armijo 0.5).
static PGOptions g_pgLineSearch;
void orc_set_pg_line_search(int lsIters, double optimistic, double pessimistic, double armijo) {
  g_pgLineSearch.lsIters = (unsigned)lsIters;
  g_pgLineSearch.lsOptimistic = optimistic;
  g_pgLineSearch.lsPessimistic = pessimistic;
  g_pgLineSearch.lsArmijo = armijo;
}

double orc_pg_solve(void* p, int variant, double tol, int maxIters, int coulomb, int useAssembled, int diagScaling,
                    int infNorm, double* r, int* iters, int* spmvs, double* hist, int histCap, int* histLen) {
  PGOptions o;
  o.lsIters = g_pgLineSearch.lsIters;
  o.lsOptimistic = g_pgLineSearch.lsOptimistic;
  o.lsPessimistic = g_pgLineSearch.lsPessimistic;
  o.lsArmijo = g_pgLineSearch.lsArmijo;
  o.variant = variant;
  o.tol = tol;
  o.maxIters = (unsigned)maxIters;
  o.coulomb = coulomb;
  o.useAssembled = useAssembled;
  o.diagScaling = diagScaling;
  o.infNorm = infNorm;
  SolveStats st;
  const double res = pgSolve(*static_cast<Problem*>(p), o, r, &st);
  if (iters) *iters = (int)st.iterations;
  if (spmvs) *spmvs = (int)st.spmvs;
  copyHistory(st, hist, histCap, histLen);
  return res;
}

static double gsCommon(bool product, void* p, double tol, int maxIters, int evalEvery, double skipTol, int skipIters, double autoReg,
                    int tryZero, int coulomb, int strategy, int infNorm, int ndxRefPerColor, const int* colors,
                    int ncolors, double* r, int* iters, double* hist, int histCap, int* histLen) {
  const Problem* P = static_cast<Problem*>(p);
  GSOptions o;
  o.tol = tol;
  o.maxIters = (unsigned)maxIters;
  o.evalEvery = (unsigned)evalEvery;
  o.skipTol = skipTol;
  o.skipIters = (unsigned)skipIters;
  o.autoRegularization = autoReg;
  o.tryZero = tryZero;
  o.coulomb = coulomb;
  o.strategy = strategy;
  o.infNorm = infNorm;
  o.ndxRefPerColor = ndxRefPerColor;
  std::vector<int> c;
  if (colors && ncolors > 0)
    c.assign(colors, colors + ncolors + 1);
  else
    c = {0, P->nc};
  SolveStats st;
  const double res = product ? pgsSolve(*P, o, c, r, &st) : gsSolve(*P, o, c, r, &st);
  if (iters) *iters = (int)st.iterations;
  copyHistory(st, hist, histCap, histLen);
  return res;
}

double orc_gs_solve(void* p, double tol, int maxIters, int evalEvery, double skipTol, int skipIters, double autoReg,
                    int tryZero, int coulomb, int strategy, int infNorm, int ndxRefPerColor, const int* colors,
                    int ncolors, double* r, int* iters, double* hist, int histCap, int* histLen) {
  return gsCommon(false, p, tol, maxIters, evalEvery, skipTol, skipIters, autoReg, tryZero, coulomb, strategy, infNorm,
                  ndxRefPerColor, colors, ncolors, r, iters, hist, histCap, histLen);
}

// ProductGaussSeidel (matrix-free GS, no W)
double orc_pgs_solve(void* p, double tol, int maxIters, int evalEvery, double skipTol, int skipIters, double autoReg,
                     int tryZero, int coulomb, int strategy, int infNorm, int ndxRefPerColor, const int* colors,
                     int ncolors, double* r, int* iters, double* hist, int histCap, int* histLen) {
  return gsCommon(true, p, tol, maxIters, evalEvery, skipTol, skipIters, autoReg, tryZero, coulomb, strategy, infNorm,
                  ndxRefPerColor, colors, ncolors, r, iters, hist, histCap, histLen);
}

// Cadoux fixed point with the GS (use_gs = 1: colors) or APGD/SPG (use_gs = 0) inner solver.
double orc_cadoux_solve(void* p, int useGS, int variant, int useAssembled, int diagScaling, double tol, int innerMaxIters,
                        int iterations, double tolTighten, int velocity, const int* colors, int ncolors, double* r,
                        int* iters, int* spmvs, double* hist, int histCap, int* histLen) {
  const Problem* P = static_cast<Problem*>(p);
  CadouxOptions o;
  o.useGS = useGS;
  o.gs.tol = tol;
  o.gs.maxIters = (unsigned)innerMaxIters;
  o.gs.ndxRefPerColor = 1;
  o.pg.variant = variant;
  o.pg.tol = tol;
  o.pg.maxIters = (unsigned)innerMaxIters;
  o.pg.useAssembled = useAssembled;
  o.pg.diagScaling = diagScaling;
  o.iterations = (unsigned)iterations;
  o.tolTighten = tolTighten;
  o.velocity = velocity;
  std::vector<int> c;
  if (colors && ncolors > 0)
    c.assign(colors, colors + ncolors + 1);
  else
    c = {0, P->nc};
  SolveStats st;
  const double res = cadouxSolve(*P, o, c, r, &st);
  if (iters) *iters = (int)st.iterations;
  if (spmvs) *spmvs = (int)st.spmvs;
  copyHistory(st, hist, histCap, histLen);
  return res;
}

int orc_coloring(void* p, int* perm, int* colors) {
  const Problem* P = static_cast<Problem*>(p);
  std::vector<int> pm, cl;
  computeColoring(P->W, pm, cl);
  std::memcpy(perm, pm.data(), sizeof(int) * pm.size());
  std::memcpy(colors, cl.data(), sizeof(int) * cl.size());
  return (int)cl.size() - 1;
}

void orc_problem_permute(void* p, const int* perm) {
  Problem* P = static_cast<Problem*>(p);
  std::vector<int> pm(perm, perm + P->nc);
  permuteProblem(*P, pm);
}

// ---- per-contact kernels and local solver
void orc_fb(int coulomb, double mu, const double* x, const double* y, double* fb) { fbLaw(coulomb != 0, mu, x, y, fb); }
void orc_fb_base_jacobian(double mu, const double* x, const double* y, double* fb, double* dx9, double* dy9) {
  double dx[3][3], dy[3][3];
  fbBaseJacobian(mu, x, y, fb, dx, dy);
  std::memcpy(dx9, dx, sizeof(dx));
  std::memcpy(dy9, dy, sizeof(dy));
}
double orc_law_eval(int coulomb, double mu, const double* x, const double* y) { return lawEval(coulomb != 0, mu, x, y); }
void orc_project(double mu, double* x) { lawProject(mu, x); }
void orc_cov(int coulomb, double mu, const double* y, double* s) { lawCOV(coulomb != 0, mu, y, s); }

double orc_local_solve(int coulomb, int strategy, const double* A9, const double* b, double* x, double mu, double tol,
                       double scaling) {
  double A[3][3];
  std::memcpy(A, A9, sizeof(A));
  return localSolve(coulomb != 0, (Strategy)strategy, A, b, x, mu, tol, scaling);
}

int orc_solve_orthogonality(int coulomb, const double* W9, const double* b, double* r, double mu) {
  double W[3][3];
  std::memcpy(W, W9, sizeof(W));
  return solveOrthogonality(coulomb != 0, W, b, r, mu) ? 1 : 0;
}

double orc_newton(int coulomb, const double* A9, const double* b, double mu, double scaling, double tol, int maxIters,
                  double* x) {
  double A[3][3];
  std::memcpy(A, A9, sizeof(A));
  return nonSmoothNewton(coulomb != 0, A, b, mu, scaling, tol, (unsigned)maxIters, x);
}

unsigned orc_real_roots_monic(int D, const double* c, double* roots, int filter) {
  return realRootsMonic(D, c, roots, (RootFilter)filter);
}
unsigned orc_real_roots_degenerate(int D, double* c, double* roots, int filter) {
  return realRootsDegenerate(D, c, roots, (RootFilter)filter);
}
void orc_eig_hessenberg(int n, const double* H, double* wr, double* wi) { hessenbergEigenvalues(n, H, wr, wi); }

int orc_lu_solve(int n, const double* A, const double* b, double* x) {
  if (n == 3) {
    double M[3][3];
    std::memcpy(M, A, sizeof(M));
    FullPivLU<3> lu;
    lu.compute(M);
    lu.solve(b, x);
    return lu.rank;
  }
  if (n == 2) {
    double M[2][2];
    std::memcpy(M, A, sizeof(M));
    FullPivLU<2> lu;
    lu.compute(M);
    lu.solve(b, x);
    return lu.rank;
  }
  return -1;
}

// ---- generic BSR products (known-answer tests of the block SpMV semantics)
static BSR makeBSR(int nbr, int nbc, int br, int bc, int sym, const int* rowptr, const int* colidx,
                   const double* vals) {
  BSR A;
  A.nbr = nbr;
  A.nbc = nbc;
  A.br = br;
  A.bc = bc;
  A.symLower = sym != 0;
  A.rowptr.assign(rowptr, rowptr + nbr + 1);
  A.colidx.assign(colidx, colidx + rowptr[nbr]);
  A.vals.assign(vals, vals + (size_t)rowptr[nbr] * br * bc);
  return A;
}

void orc_bsr_multiply(int nbr, int nbc, int br, int bc, int sym, const int* rowptr, const int* colidx,
                      const double* vals, int transpose, const double* x, double* res, double alpha, double beta) {
  BSR A = makeBSR(nbr, nbc, br, bc, sym, rowptr, colidx, vals);
  bsrMultiply(A, transpose != 0, x, res, alpha, beta);
}

void orc_bsr_split_row(int nbr, const int* rowptr, const int* colidx, const double* vals, int row, const double* x,
                       double* res) {
  BSR A = makeBSR(nbr, nbr, 3, 3, 1, rowptr, colidx, vals);
  bsrSplitRowMultiply(A, row, x, res);
}

}  // extern "C"
