// bogus_hip.hpp — C++ host facade of the MI355X friction solve path.
//
// Mirrors the reference's C++ API for the Coulomb-friction solve path (bogus 1.4.2,
// paths relative to its src/ tree) with the same class and method names, argument
// meaning and return values, so calling code reads as with the reference:
//   SparseBlockMatrix            Core/Block/SparseBlockMatrix.hpp:18-135 (setRows, setCols,
//                                insertBack, finalize, multiply<Transpose>, splitRowMultiply)
//   ProjectedGradient            Core/BlockSolvers/ProjectedGradient.hpp:20-107
//   GaussSeidel + Coloring       Core/BlockSolvers/GaussSeidel.hpp:41-125, GaussSeidelBase.hpp:20-91,
//                                Core/BlockSolvers/Coloring.hpp:20-59
//   PrimalFrictionProblem<3>     Interfaces/FrictionProblem.hpp:55-113 (computeMInv)
//   DualFrictionProblem<3>       Interfaces/FrictionProblem.hpp:127-214 (computeFrom, solveWith,
//                                evalWith, applyPermutation, undoPermutation)
//   Signal<unsigned, double>     Core/Utils/Signal.hpp:25-54 (iteration callback)
// Differences, by design: dense blocks are plain row-major double arrays (no
// Eigen); W and the solvers live on the GPU and every compute call goes through
// the extern "C" layer of sobogus_hip.h. Hard errors (no GPU, HIP failures)
// throw bogus::hip_error; a solve that stops at max_iters returns its residual
// exactly as the reference does.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sobogus_hip.h"

namespace bogus {

struct hip_error : std::runtime_error {
  int status;
  hip_error(int s, const std::string& what)
      : std::runtime_error(what + ": " + sb_status_string(s)), status(s) {}
};

inline void sb_check(int s, const char* what) {
  if (s < 0) throw hip_error(s, what);
}

namespace flags {
enum { NONE = 0, SYMMETRIC = 1 };
}
namespace projected_gradient {
enum Variant { Standard = SB_PG_STANDARD, Descent = SB_PG_DESCENT, Conjugated = SB_PG_CONJUGATED, APGD = SB_PG_APGD,
               SPG = SB_PG_SPG };
}
namespace local_soc_solver {
enum Strategy { PureNewton = SB_PURE_NEWTON, PureEnumerative = SB_PURE_ENUMERATIVE, Hybrid = SB_HYBRID,
                RevHybrid = SB_REV_HYBRID };
}

// One HIP device/stream per thread by default (sb_ctx).
class Context {
 public:
  explicit Context(int device = 0) {
    sb_ctx* c = nullptr;
    sb_check(sb_ctx_create(&c, device), "sb_ctx_create");
    m_ctx.reset(c, [](sb_ctx* p) { sb_ctx_destroy(p); });
  }
  sb_ctx* get() const { return m_ctx.get(); }
  static Context& threadDefault() {
    thread_local Context c(0);
    return c;
  }

 private:
  std::shared_ptr<sb_ctx> m_ctx;
};

// Iteration callback (Signal<unsigned, double>): replayed after the solve from the
// device-side residual history.
template <typename... Args>
class Signal {
 public:
  using Callback = std::function<void(Args...)>;
  void connect(Callback cb) { m_cbs.push_back(std::move(cb)); }
  void disconnect() { m_cbs.clear(); }
  void trigger(Args... a) const {
    for (const auto& cb : m_cbs) cb(a...);
  }
  bool empty() const { return m_cbs.empty(); }

 private:
  std::vector<Callback> m_cbs;
};

// ------------------------------------------------------------------ SparseBlockMatrix
class SparseBlockMatrix {
 public:
  explicit SparseBlockMatrix(int fl = flags::NONE, Context& ctx = Context::threadDefault()) : m_flags(fl), m_ctx(ctx) {}

  void setRows(int nBlocks, int rowsPerBlock) {
    m_nbr = nBlocks;
    m_br = rowsPerBlock;
    if (m_flags & flags::SYMMETRIC) setCols(nBlocks, rowsPerBlock);
  }
  void setCols(int nBlocks, int colsPerBlock) {
    m_nbc = nBlocks;
    m_bc = colsPerBlock;
  }
  // Appends block (row, col); rows must be inserted in non-decreasing order and
  // columns increasing within a row. Returns the row-major block storage.
  double* insertBack(int row, int col) {
    if (row < m_lastRow || (row == m_lastRow && col <= m_lastCol)) throw std::invalid_argument("insertBack order");
    if ((m_flags & flags::SYMMETRIC) && col > row) throw std::invalid_argument("symmetric: col <= row");
    m_rows.push_back(row);
    m_colidx.push_back(col);
    m_vals.resize(m_vals.size() + (size_t)m_br * m_bc, 0.0);
    m_lastRow = row;
    m_lastCol = col;
    m_final.reset();
    return &m_vals[m_vals.size() - (size_t)m_br * m_bc];
  }
  void finalize() {
    m_rowptr.assign(m_nbr + 1, 0);
    for (int r : m_rows) m_rowptr[r + 1]++;
    for (int i = 0; i < m_nbr; ++i) m_rowptr[i + 1] += m_rowptr[i];
    sb_bsr* A = nullptr;
    sb_check(sb_bsr_create(m_ctx.get(), &A, m_nbr, m_nbc, m_br, m_bc,
                           (m_flags & flags::SYMMETRIC) ? SB_BSR_SYMMETRIC : 0, m_rowptr.data(), m_colidx.data(),
                           m_vals.data()),
             "sb_bsr_create");
    m_final.reset(A, [](sb_bsr* p) { sb_bsr_destroy(p); });
  }
  // res = alpha * op(this) * rhs + beta * res (BlockObjectBase::multiply<Transpose>)
  template <bool Transpose>
  void multiply(const double* rhs, double* res, double alpha = 1, double beta = 0) const {
    sb_check(sb_bsr_multiply(m_ctx.get(), handle(), Transpose ? 1 : 0, alpha, rhs, beta, res, 0), "multiply");
  }
  template <bool Transpose>
  void multiply(const std::vector<double>& rhs, std::vector<double>& res, double alpha = 1, double beta = 0) const {
    res.resize(Transpose ? cols() : rows(), 0.0);
    multiply<Transpose>(rhs.data(), res.data(), alpha, beta);
  }
  // res += sum_{j != row} A_{row,j} rhs_j (symmetric 3x3 matrices)
  void splitRowMultiply(int row, const double* rhs, double* res) const {
    sb_check(sb_bsr_split_row_multiply(m_ctx.get(), handle(), row, rhs, res), "splitRowMultiply");
  }
  int rowsOfBlocks() const { return m_nbr; }
  int colsOfBlocks() const { return m_nbc; }
  int rows() const { return m_nbr * m_br; }
  int cols() const { return m_nbc * m_bc; }
  size_t nBlocks() const { return m_colidx.size(); }
  const std::vector<int32_t>& rowPointers() const { return m_rowptr; }
  const std::vector<int32_t>& columnIndices() const { return m_colidx; }
  const std::vector<double>& blockValues() const { return m_vals; }
  const double* block(size_t ptr) const { return &m_vals[ptr * m_br * m_bc]; }

 private:
  const sb_bsr* handle() const {
    if (!m_final) throw std::logic_error("SparseBlockMatrix: finalize() first");
    return m_final.get();
  }
  int m_flags;
  Context& m_ctx;
  int m_nbr = 0, m_nbc = 0, m_br = 0, m_bc = 0;
  int m_lastRow = -1, m_lastCol = -1;
  std::vector<int> m_rows;
  std::vector<int32_t> m_rowptr, m_colidx;
  std::vector<double> m_vals;
  std::shared_ptr<sb_bsr> m_final;
};

// ------------------------------------------------------------------ solvers
class ProjectedGradient {
 public:
  ProjectedGradient() { sb_pg_options_default(&m_o); }
  void setTol(double t) { m_o.tol = t; }
  void setMaxIters(unsigned n) { m_o.max_iters = (int)n; }
  double tol() const { return m_o.tol; }
  unsigned maxIters() const { return (unsigned)m_o.max_iters; }
  void setLineSearchIterations(unsigned n) { m_o.ls_iters = (int)n; }
  void setLineSearchOptimisticFactor(double f) { m_o.ls_optimistic = f; }
  void setLineSearchPessimisticFactor(double f) { m_o.ls_pessimistic = f; }
  void setLineSearchArmijoCoefficient(double c) { m_o.ls_armijo = c; }
  void setDefaultVariant(projected_gradient::Variant v) { m_o.variant = v; }
  void useInfinityNorm(bool b) { m_o.inf_norm = b ? 1 : 0; }
  // GPU path: operator choice (assembled W or the matrix-free product) and residual scaling
  void setMatrixFree(bool mf) { m_o.op = mf ? SB_OP_MATRIX_FREE : SB_OP_ASSEMBLED; }
  void setScaling(int sb_scaling_mode) { m_o.scaling = sb_scaling_mode; }
  Signal<unsigned, double>& callback() { return m_cb; }
  const Signal<unsigned, double>& callback() const { return m_cb; }
  sb_pg_options options(int law) const {
    sb_pg_options o = m_o;
    o.law = law;
    return o;
  }
  const sb_stats& lastStats() const { return m_stats; }
  sb_stats& mutableStats() const { return m_stats; }

 private:
  sb_pg_options m_o;
  Signal<unsigned, double> m_cb;
  mutable sb_stats m_stats{};
};

// Coloring::update(enable, W): permutation listing contacts color by color and
// color offsets. The reference's rule is kept as SB_COLORING_REFERENCE; the GPU
// default is the exact first-fit rule (fewer colors, same guarantee).
struct Coloring {
  std::vector<std::size_t> permutation;
  std::vector<std::ptrdiff_t> colors;
  int mode = SB_COLORING_GREEDY;
  std::size_t size() const { return permutation.size(); }
  void resetPermutation() {
    for (std::size_t i = 0; i < permutation.size(); ++i) permutation[i] = i;
  }
};

class GaussSeidel {
 public:
  GaussSeidel() { sb_gs_options_default(&m_o); }
  void setTol(double t) { m_o.tol = t; }
  void setMaxIters(unsigned n) { m_o.max_iters = (int)n; }
  double tol() const { return m_o.tol; }
  void setEvalEvery(unsigned n) { m_o.eval_every = (int)n; }
  void setSkipTol(double t) { m_o.skip_tol = t; }
  void setSkipIters(unsigned n) { m_o.skip_iters = (int)n; }
  void setAutoRegularization(double r) { m_o.auto_regularization = r; }
  void doTryZeroAsWell(bool b = true) { m_o.try_zero = b ? 1 : 0; }
  void setMaxThreads(unsigned) {}  // the GPU path has no OpenMP thread count
  void useInfinityNorm(bool b) { m_o.inf_norm = b ? 1 : 0; }
  void setLocalStrategy(local_soc_solver::Strategy s) { m_o.strategy = s; }
  Coloring& coloring() { return m_coloring; }
  const Coloring& coloring() const { return m_coloring; }
  Signal<unsigned, double>& callback() { return m_cb; }
  const Signal<unsigned, double>& callback() const { return m_cb; }
  sb_gs_options options(int law) const {
    sb_gs_options o = m_o;
    o.law = law;
    return o;
  }
  const sb_stats& lastStats() const { return m_stats; }
  sb_stats& mutableStats() const { return m_stats; }

 private:
  sb_gs_options m_o;
  Coloring m_coloring;
  Signal<unsigned, double> m_cb;
  mutable sb_stats m_stats{};
};

// ProductGaussSeidel<HType, MInvType, true> (ProductGaussSeidel.hpp:20-137): the
// matrix-free Gauss-Seidel on W = H Minv H^T, W never formed. The GPU sweeps the
// contacts color by color (contacts of one color share no body); the coloring is
// computed inside solveWith, with the mode set on coloring().
class ProductGaussSeidel : public GaussSeidel {};

// ------------------------------------------------------------------ friction problems
template <unsigned Dimension>
struct PrimalFrictionProblem;
template <unsigned Dimension>
struct DualFrictionProblem;

// Primal data with 6-dof rigid bodies: M (per-body 6x6, row-major), H (3x6 blocks
// in contact coordinates), optional per-contact rotations E (3x3, row-major),
// borrowed f, w, mu (PrimalFrictionProblem, FrictionProblem.hpp:55-113).
template <>
struct PrimalFrictionProblem<3> {
  std::vector<double> M;          // n_bodies x 36
  SparseBlockMatrix H;            // n_contacts x n_bodies blocks of 3x6
  std::vector<double> E;          // n_contacts x 9, empty = identity
  const double* f = nullptr;      // 6 n_bodies
  const double* w = nullptr;      // 3 n_contacts
  const double* mu = nullptr;     // n_contacts
  std::vector<double> MInv;       // n_bodies x 36 (computeMInv)

  // Per-body inverse of M (PrimalFrictionProblem.cpp:28-38), by Gauss-Jordan
  // elimination with partial pivoting on the 6x6 blocks.
  void computeMInv() {
    const size_t nb = M.size() / 36;
    MInv.assign(nb * 36, 0.0);
    for (size_t k = 0; k < nb; ++k) {
      double a[6][12];
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 12; ++j) a[i][j] = (j < 6) ? M[k * 36 + i * 6 + j] : (j - 6 == i ? 1.0 : 0.0);
      for (int c = 0; c < 6; ++c) {
        int p = c;
        for (int i = c + 1; i < 6; ++i)
          if (std::fabs(a[i][c]) > std::fabs(a[p][c])) p = i;
        if (a[p][c] == 0.0) throw std::runtime_error("computeMInv: singular mass block");
        if (p != c)
          for (int j = 0; j < 12; ++j) std::swap(a[p][j], a[c][j]);
        const double inv = 1.0 / a[c][c];
        for (int j = 0; j < 12; ++j) a[c][j] *= inv;
        for (int i = 0; i < 6; ++i)
          if (i != c) {
            const double m = a[i][c];
            for (int j = 0; j < 12; ++j) a[i][j] -= m * a[c][j];
          }
      }
      for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) MInv[k * 36 + i * 6 + j] = a[i][j + 6];
    }
  }

  // b = E^T w - H (Minv f) (FrictionProblem.cpp:30-40, PrimalFrictionProblem.cpp:44-45)
  std::vector<double> dualRhs() const {
    const int nc = H.rowsOfBlocks(), nb = H.colsOfBlocks();
    if (MInv.size() != (size_t)nb * 36) throw std::logic_error("call computeMInv() first");
    const auto& rp = H.rowPointers();
    const auto& ci = H.columnIndices();
    const auto& hv = H.blockValues();
    std::vector<double> mf((size_t)6 * nb, 0.0);
    for (int k = 0; k < nb; ++k)
      for (int r = 0; r < 6; ++r) {
        double acc = 0;
        for (int c = 0; c < 6; ++c) acc += MInv[(size_t)k * 36 + r * 6 + c] * f[6 * k + c];
        mf[6 * k + r] = acc;
      }
    std::vector<double> b((size_t)3 * nc, 0.0);
    for (int i = 0; i < nc; ++i) {
      for (int r = 0; r < 3; ++r) {
        double ew = 0;
        if (E.empty())
          ew = w[3 * i + r];
        else
          for (int c = 0; c < 3; ++c) ew += E[(size_t)9 * i + c * 3 + r] * w[3 * i + c];
        double hm = 0;
        for (int p = rp[i]; p < rp[i + 1]; ++p)
          for (int c = 0; c < 6; ++c) hm += hv[(size_t)18 * p + r * 6 + c] * mf[6 * ci[p] + c];
        b[3 * i + r] = ew - hm;
      }
    }
    return b;
  }

  using ProductGaussSeidelType = ProductGaussSeidel;
  // PrimalFrictionProblem::solveWith(pgs, r, staticProblem = false) (PrimalFrictionProblem.cpp:40-59):
  // b from the primal data, W never assembled. The contacts are colored and permuted on a
  // device copy of the problem; r is returned in the caller's order.
  double solveWith(ProductGaussSeidelType& pgs, double* r, const bool staticProblem = false,
                   Context& ctx = Context::threadDefault()) const {
    const int nc = H.rowsOfBlocks(), nb = H.colsOfBlocks();
    const std::vector<double> b = dualRhs();
    sb_problem* P = nullptr;
    sb_check(sb_problem_create(ctx.get(), &P, nc, nb, H.rowPointers().data(), H.columnIndices().data(),
                               H.blockValues().data(), MInv.data(), b.data(), mu, 0),
             "sb_problem_create");
    std::unique_ptr<sb_problem, int (*)(sb_problem*)> hold(P, sb_problem_destroy);
    std::vector<int64_t> perm(nc), colors(nc + 1);
    int ncol = 0;
    sb_check(sb_problem_coloring(ctx.get(), P, pgs.coloring().mode, perm.data(), colors.data(), &ncol), "coloring");
    sb_check(sb_problem_apply_permutation(ctx.get(), P, perm.data()), "apply_permutation");
    pgs.coloring().permutation.assign(perm.begin(), perm.end());
    pgs.coloring().colors.assign(colors.begin(), colors.begin() + ncol + 1);
    std::vector<double> rp((size_t)3 * nc);
    for (int i = 0; i < nc; ++i)
      for (int k = 0; k < 3; ++k) rp[3 * (size_t)i + k] = r[3 * (size_t)perm[i] + k];
    const sb_gs_options o = pgs.options(staticProblem ? SB_LAW_SOC : SB_LAW_COULOMB);
    std::vector<double> hist(4096);
    sb_stats& st = pgs.mutableStats();
    st.history = hist.data();
    st.history_capacity = (int)hist.size();
    double res = 0;
    sb_check(sb_pgs_solve(ctx.get(), P, &o, colors.data(), ncol, rp.data(), 0, &res, &st), "sb_pgs_solve");
    for (int k = 0; k < st.history_length && k < (int)hist.size(); ++k)
      pgs.callback().trigger(k * o.eval_every, hist[k]);
    st.history = nullptr;
    for (int i = 0; i < nc; ++i)
      for (int k = 0; k < 3; ++k) r[3 * (size_t)perm[i] + k] = rp[3 * (size_t)i + k];
    return res;
  }
};

template <>
struct DualFrictionProblem<3> {
  using GaussSeidelType = GaussSeidel;
  using ProjectedGradientType = ProjectedGradient;
  using SignalType = Signal<unsigned, double>;

  std::vector<double> b;   // 3 n_contacts (u = W r + b)
  std::vector<double> mu;  // n_contacts

  explicit DualFrictionProblem(Context& ctx = Context::threadDefault()) : m_ctx(ctx) {}

  // W = H MInv H^T (assembled on setup), b = E^T w - H MInv f, mu (FrictionProblem.cpp:30-40)
  void computeFrom(const PrimalFrictionProblem<3>& primal) {
    const int nc = primal.H.rowsOfBlocks(), nb = primal.H.colsOfBlocks();
    if (primal.MInv.size() != (size_t)nb * 36) throw std::logic_error("computeFrom: call computeMInv() first");
    const auto& rp = primal.H.rowPointers();
    const auto& ci = primal.H.columnIndices();
    const auto& hv = primal.H.blockValues();
    b = primal.dualRhs();
    mu.assign(primal.mu, primal.mu + nc);
    sb_problem* P = nullptr;
    sb_check(sb_problem_create(m_ctx.get(), &P, nc, nb, rp.data(), ci.data(), hv.data(), primal.MInv.data(),
                               b.data(), mu.data(), SB_PROBLEM_ASSEMBLE_W),
             "sb_problem_create");
    m_P.reset(P, [](sb_problem* p) { sb_problem_destroy(p); });
    m_perm.clear();
    m_inv.clear();
  }

  // The dual problem of a local problem file, as the reference's FCLib application builds it
  // (apps/FCLibLoader.cpp:64-73: bogus::convert(W, dual.W, 3, 3); dual.W.prune(pruneTol);
  // dual.b = q; dual.mu = mu): W is the scalar CSR of the 3n x 3n matrix. Only the
  // assembled operator exists for such a problem (no H): solve with GaussSeidel or with an
  // assembled ProjectedGradient.
  void fromLocalProblem(int n, const int* rowptr, const int* colidx, const double* vals, const double* q,
                        const double* muIn, double pruneTol = 0.0) {
    b.assign(q, q + 3 * (size_t)n);
    mu.assign(muIn, muIn + n);
    sb_problem* P = nullptr;
    sb_check(sb_problem_create_from_w(m_ctx.get(), &P, n, rowptr, colidx, vals, b.data(), mu.data(), pruneTol),
             "sb_problem_create_from_w");
    m_P.reset(P, [](sb_problem* p) { sb_problem_destroy(p); });
    m_perm.clear();
    m_inv.clear();
  }

  // The FCLib application's error of r (apps/FCLibLoader.cpp:40-61): merit1 with u = W r + b.
  double merit1(const double* r) const {
    const int nc = (int)mu.size();
    std::vector<double> rp(r, r + 3 * (size_t)nc);
    if (permuted()) permuteVector(m_perm, rp);
    double m = 0;
    sb_check(sb_merit1(m_ctx.get(), m_P.get(), SB_OP_ASSEMBLED, rp.data(), 0, &m), "sb_merit1");
    return m;
  }

  // DualFrictionProblem::solveWith(gs, r, staticProblem = false)
  double solveWith(GaussSeidelType& gs, double* r, const bool staticProblem = false) const {
    const int nc = (int)mu.size();
    std::vector<double> rp(r, r + 3 * (size_t)nc);
    if (permuted()) permuteVector(m_perm, rp);
    const sb_gs_options o = gs.options(staticProblem ? SB_LAW_SOC : SB_LAW_COULOMB);
    std::vector<int64_t> colors;
    if (gs.coloring().colors.size() >= 2 && (size_t)gs.coloring().colors.back() == (size_t)nc)
      colors.assign(gs.coloring().colors.begin(), gs.coloring().colors.end());
    std::vector<double> hist(4096);
    sb_stats& st = gs.mutableStats();
    st.history = hist.data();
    st.history_capacity = (int)hist.size();
    double res = 0;
    sb_check(sb_gs_solve(m_ctx.get(), m_P.get(), &o, colors.empty() ? nullptr : colors.data(),
                         colors.empty() ? 0 : (int)colors.size() - 1, rp.data(), 0, &res, &st),
             "sb_gs_solve");
    const int every = o.eval_every;
    for (int k = 0; k < st.history_length && k < (int)hist.size(); ++k) gs.callback().trigger(k * every, hist[k]);
    st.history = nullptr;
    if (permuted()) permuteVector(m_inv, rp);
    std::copy(rp.begin(), rp.end(), r);
    return res;
  }

  // DualFrictionProblem::solveWith(pg, r, staticProblem = true)
  double solveWith(ProjectedGradientType& pg, double* r, const bool staticProblem = true) const {
    const int nc = (int)mu.size();
    std::vector<double> rp(r, r + 3 * (size_t)nc);
    if (permuted()) permuteVector(m_perm, rp);
    const sb_pg_options o = pg.options(staticProblem ? SB_LAW_SOC : SB_LAW_COULOMB);
    std::vector<double> hist((size_t)o.max_iters + 1);
    sb_stats& st = pg.mutableStats();
    st.history = hist.data();
    st.history_capacity = (int)hist.size();
    double res = 0;
    const int s = sb_pg_solve(m_ctx.get(), m_P.get(), &o, rp.data(), 0, &res, &st);
    // an unknown variant returns -1, as the reference's ProjectedGradient::solve
    if (s == SB_ERR_UNSUPPORTED || (s == SB_ERR_INVALID && (o.variant < SB_PG_STANDARD || o.variant > SB_PG_SPG)))
      return -1;
    sb_check(s, "sb_pg_solve");
    for (int k = 0; k < st.history_length && k < (int)hist.size(); ++k) pg.callback().trigger(k, hist[k]);
    st.history = nullptr;
    if (permuted()) permuteVector(m_inv, rp);
    std::copy(rp.begin(), rp.end(), r);
    return res;
  }

  // solveCadoux / solveCadouxVel (FrictionProblem.hpp:177-195, Cadoux.hpp:44-155): fixed-point
  // iterations of SOC solves with the given solver; the callback receives the residual
  // before and after each iteration.
  double solveCadoux(GaussSeidelType& gs, double* r, const unsigned fpIterations,
                     const SignalType* callback = nullptr) const {
    return cadoux(&gs, nullptr, r, fpIterations, callback, 0);
  }
  double solveCadoux(ProjectedGradientType& pg, double* r, const unsigned fpIterations,
                     const SignalType* callback = nullptr) const {
    return cadoux(nullptr, &pg, r, fpIterations, callback, 0);
  }
  double solveCadouxVel(GaussSeidelType& gs, double* u, const unsigned fpIterations,
                        const SignalType* callback = nullptr) const {
    return cadoux(&gs, nullptr, u, fpIterations, callback, 1);
  }
  double solveCadouxVel(ProjectedGradientType& pg, double* u, const unsigned fpIterations,
                        const SignalType* callback = nullptr) const {
    return cadoux(nullptr, &pg, u, fpIterations, callback, 1);
  }

  double evalWith(const GaussSeidelType&, const double* r, const bool staticProblem = false) const {
    return eval(r, staticProblem, SB_OP_ASSEMBLED, SB_SCALING_REFERENCE);
  }
  double evalWith(const ProjectedGradientType& pg, const double* r, const bool staticProblem = true) const {
    const sb_pg_options o = pg.options(0);
    return eval(r, staticProblem, o.op, o.scaling, o.inf_norm);
  }

  // Coloring::update(true, W) on the assembled W
  void computeColoring(Coloring& c) const {
    const int nc = (int)mu.size();
    std::vector<int64_t> perm(nc), colors(nc + 1);
    int ncol = 0;
    sb_check(sb_problem_coloring(m_ctx.get(), m_P.get(), c.mode, perm.data(), colors.data(), &ncol), "coloring");
    c.permutation.assign(perm.begin(), perm.end());
    c.colors.assign(colors.begin(), colors.begin() + ncol + 1);
  }

  // applyPermutation / undoPermutation (FrictionProblem.cpp:102-125): new[i] = old[perm[i]]
  void applyPermutation(const std::vector<std::size_t>& permutation) {
    if (permuted()) throw std::logic_error("applyPermutation: already permuted");
    m_perm = permutation;
    m_inv.resize(m_perm.size());
    for (std::size_t i = 0; i < m_perm.size(); ++i) m_inv[m_perm[i]] = i;
    permuteDevice(m_perm);
  }
  void undoPermutation() {
    if (!permuted()) return;
    permuteDevice(m_inv);
    m_perm.clear();
    m_inv.clear();
  }
  double cadoux(GaussSeidelType* gs, ProjectedGradientType* pg, double* x, unsigned fpIterations,
                const SignalType* callback, int velocity) const {
    const int nc = (int)mu.size();
    std::vector<double> xp(x, x + 3 * (size_t)nc);
    if (permuted()) permuteVector(m_perm, xp);
    std::vector<double> hist((size_t)fpIterations + 1);
    sb_stats st{};
    st.history = hist.data();
    st.history_capacity = (int)hist.size();
    double res = 0;
    const double tighten = 1e-1;  // Cadoux.hpp: tolTighten default
    if (gs) {
      const sb_gs_options o = gs->options(SB_LAW_SOC);
      std::vector<int64_t> colors;
      if (gs->coloring().colors.size() >= 2 && (size_t)gs->coloring().colors.back() == (size_t)nc)
        colors.assign(gs->coloring().colors.begin(), gs->coloring().colors.end());
      sb_check(sb_gs_solve_cadoux(m_ctx.get(), m_P.get(), &o, colors.empty() ? nullptr : colors.data(),
                                  colors.empty() ? 0 : (int)colors.size() - 1, (int)fpIterations, tighten, velocity,
                                  xp.data(), 0, &res, &st),
               "sb_gs_solve_cadoux");
    } else {
      const sb_pg_options o = pg->options(SB_LAW_SOC);
      sb_check(sb_pg_solve_cadoux(m_ctx.get(), m_P.get(), &o, (int)fpIterations, tighten, velocity, xp.data(), 0,
                                  &res, &st),
               "sb_pg_solve_cadoux");
    }
    if (callback)
      for (int k = 0; k < st.history_length && k < (int)hist.size(); ++k) callback->trigger((unsigned)k, hist[k]);
    if (permuted()) permuteVector(m_inv, xp);
    std::copy(xp.begin(), xp.end(), x);
    return res;
  }
  bool permuted() const { return !m_perm.empty(); }
  const std::vector<std::size_t>& permutation() const { return m_perm; }
  const std::vector<std::size_t>& invPermutation() const { return m_inv; }
  sb_problem* handle() const { return m_P.get(); }

 private:
  double eval(const double* r, bool staticProblem, int op, int scaling, int inf = 0) const {
    const int nc = (int)mu.size();
    std::vector<double> rp(r, r + 3 * (size_t)nc);
    if (permuted()) permuteVector(m_perm, rp);
    double res = 0;
    sb_check(sb_fb_residual(m_ctx.get(), m_P.get(), staticProblem ? SB_LAW_SOC : SB_LAW_COULOMB, op, scaling, inf,
                            rp.data(), &res, 0),
             "sb_fb_residual");
    return res;
  }
  void permuteDevice(const std::vector<std::size_t>& p) {
    std::vector<int64_t> q(p.begin(), p.end());
    sb_check(sb_problem_apply_permutation(m_ctx.get(), m_P.get(), q.data()), "applyPermutation");
    std::vector<double> b2(b.size()), mu2(mu.size());
    for (std::size_t i = 0; i < p.size(); ++i) {
      for (int k = 0; k < 3; ++k) b2[3 * i + k] = b[3 * p[i] + k];
      mu2[i] = mu[p[i]];
    }
    b.swap(b2);
    mu.swap(mu2);
  }
  static void permuteVector(const std::vector<std::size_t>& p, std::vector<double>& v) {
    std::vector<double> o(v.size());
    for (std::size_t i = 0; i < p.size(); ++i)
      for (int k = 0; k < 3; ++k) o[3 * i + k] = v[3 * p[i] + k];
    v.swap(o);
  }

  Context& m_ctx;
  std::shared_ptr<sb_problem> m_P;
  std::vector<std::size_t> m_perm, m_inv;
};

}  // namespace bogus
