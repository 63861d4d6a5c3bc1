/*
 * sobogus_hip.h — C ABI of the MI355X (gfx950) Coulomb-friction solve path.
 *
 * The ABI replaces the compute behind these reference interfaces
 * (bogus 1.4.2 / So-bogus; paths relative to its src/ tree):
 *   - SparseBlockMatrix storage and multiply<Transpose>(rhs, res, alpha, beta)
 *       Core/Block/SparseBlockMatrix.hpp:18-135, Core/Block/BlockObjectBase.hpp:64-71,
 *       Core/Block/SparseMatrixVectorProduct.impl.hpp:382-432
 *   - the matrix-free product H * MInv * H^T (Product / Transpose expressions)
 *       Core/Block/Expressions.hpp:18-161, SparseMatrixVectorProduct.impl.hpp:355-376, 457-462
 *   - ProjectedGradient<WType>::solve(law, b, x)
 *       Core/BlockSolvers/ProjectedGradient.hpp:36-50, ProjectedGradient.impl.hpp:355-382
 *   - GaussSeidel<WType>::solve(law, b, x, tryZeroAsWell)
 *       Core/BlockSolvers/GaussSeidel.hpp:78-79, GaussSeidel.impl.hpp:126-205
 *   - Coloring::update(true, W)
 *       Core/BlockSolvers/Coloring.hpp:28-30, Coloring.impl.hpp:20-76
 *   - ConstrainedSolverBase::eval (the Fischer-Burmeister residual)
 *       Core/BlockSolvers/ConstrainedSolverBase.impl.hpp:21-74
 *   - DualFrictionProblem<3>::{computeFrom, solveWith, evalWith, applyPermutation}
 *       Interfaces/FrictionProblem.hpp:127-214, Interfaces/FrictionProblem.cpp:30-114
 *
 * Conventions
 *   - Return value: SB_OK (0), a positive warning (SB_NOT_CONVERGED) or a
 *     negative SB_ERR_* code. sb_status_string() names it. Output residuals are
 *     always written when the call gets far enough to compute one.
 *   - Host arrays are caller-owned and copied; objects are opaque, owned by the
 *     library, and freed by the matching *_destroy.
 *   - Vectors r, b, u are 3 * n_contacts doubles, contact i at [3i, 3i+3) as
 *     (normal, tangent, tangent). Blocks are row-major.
 *   - One sb_ctx per host thread; a context owns one HIP stream on one device.
 *   - Every call is synchronous with respect to the host unless its name ends in
 *     _async (none yet).
 */
#ifndef SOBOGUS_HIP_H
#define SOBOGUS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum sb_status {
  SB_OK = 0,
  SB_NOT_CONVERGED = 1,       /* warning: tolerance not reached within max_iters */
  SB_ERR_INVALID = -1,        /* bad argument */
  SB_ERR_OOM = -2,            /* device or host allocation failed */
  SB_ERR_HIP = -3,            /* HIP runtime error */
  SB_ERR_NO_DEVICE = -4,      /* no usable GPU */
  SB_ERR_UNSUPPORTED = -5,    /* option not available on this path */
  SB_ERR_BARRIER_TIMEOUT = -6,/* a persistent kernel's grid barrier timed out */
  SB_ERR_NAN = -7,            /* non-finite residual */
  SB_ERR_LIMIT = -8,          /* a fixed limit of the entry point is exceeded (sb_status_string says which) */
  SB_ERR_IO = -9              /* a file could not be opened, read or written, or is not a local-problem file */
};

/* Solver variants (Core/BlockSolvers.fwd.hpp:29-46, projected_gradient::Variant) */
enum sb_pg_variant { SB_PG_STANDARD = 0, SB_PG_DESCENT = 1, SB_PG_CONJUGATED = 2, SB_PG_APGD = 3, SB_PG_SPG = 4 };
/* Local solver strategies (Extra/SecondOrder.fwd.hpp:29-55, local_soc_solver::Strategy) */
enum sb_local_strategy { SB_PURE_NEWTON = 0, SB_PURE_ENUMERATIVE = 1, SB_HYBRID = 2, SB_REV_HYBRID = 3 };
/* Laws: Coulomb friction (De Saxce change of variable, SOCLaw<3,double,true>) or
   plain second-order-cone complementarity (SOCLaw<3,double,false>) */
enum sb_law { SB_LAW_SOC = 0, SB_LAW_COULOMB = 1 };
/* Operator used by the projected-gradient solver. SB_OP_MATRIX_FREE: H (Minv (H^T x)),
   which the solvers apply as F (F^T x) with F = H L, Minv_k = L_k L_k^T (Cholesky per body,
   computed on the device at the first matrix-free solve after the values change); a problem
   with an Minv_k that is not symmetric positive definite keeps the three-step product. */
enum sb_operator { SB_OP_ASSEMBLED = 0, SB_OP_MATRIX_FREE = 1 };
/* Residual row scaling: reference semantics (max(1, tr(W_ii)/3) for an assembled
   W, 1 for the matrix-free product) or max(1, tr(W_ii)/3) in both modes */
enum sb_scaling { SB_SCALING_REFERENCE = 0, SB_SCALING_DIAGONAL = 1, SB_SCALING_NONE = 2 };
/* Arithmetic of the projected-gradient solvers */
enum sb_precision {
  SB_PRECISION_FP64 = 0,  /* everything FP64 (the reference's arithmetic) */
  SB_PRECISION_MIXED = 1  /* FP32 matrices, iterates, products and projection; FP64 residual,
                             dot products and decisions; then an FP64 solve warm-started from
                             the FP32 result verifies the tolerance and refines if needed */
};
/* Problem creation flags */
enum sb_problem_flags {
  SB_PROBLEM_ASSEMBLE_W = 1,
  SB_BATCH_CALLER_ORDER = 2, /* sb_batch_create: store each problem in the caller's contact and body order */
  SB_BATCH_SPREAD_TEAMS = 4, /* sb_batch_create: place a team's workgroups on consecutive blocks (spread
                                over the XCDs) instead of on one XCD; the teams' barriers then write back
                                and invalidate L2. Results are bit-identical; slower. For tests. */
  SB_PROBLEM_HOST_SETUP = 8   /* sb_problem_create: build the tiles, the transpose of H and W's pattern on the
                                 host (the default builds them on the device; both give the same arrays) */
};
/* BSR flags (bogus::flags) */
enum sb_bsr_flags { SB_BSR_SYMMETRIC = 1 /* only blocks with col <= row are given */ };
/* Coloring modes */
enum sb_coloring_mode {
  SB_COLORING_REFERENCE = 0, /* Coloring::compute as written (lower_bound test; many colors) */
  SB_COLORING_GREEDY = 1,    /* first-fit with an exact membership test (fewer colors) */
  SB_COLORING_HOST = 16,     /* or'ed into the mode: compute on the host (by default the greedy rule runs on
                                the device when W is absent or assembled from H; same result) */
  SB_COLORING_DEVICE = 32    /* or'ed into SB_COLORING_GREEDY: on the device only (SB_ERR_LIMIT where the
                                host would be needed) */
};

typedef struct sb_ctx sb_ctx;
typedef struct sb_bsr sb_bsr;
typedef struct sb_problem sb_problem;

/* ProjectedGradient parameters (ProjectedGradient.hpp:57-95, BlockSolverBase.hpp:37-43) */
typedef struct sb_pg_options {
  int variant;            /* sb_pg_variant; every variant runs on the GPU (Standard and Conjugated in
                             single-problem solves only: sb_batch_* return SB_ERR_UNSUPPORTED) */
  int law;                /* sb_law; DualFrictionProblem::solveWith(pg) defaults to SOC */
  int op;                 /* sb_operator */
  int scaling;            /* sb_scaling */
  int inf_norm;           /* ConstrainedSolverBase::useInfinityNorm */
  double tol;             /* 1e-6 */
  int max_iters;          /* 300 */
  int ls_iters;           /* 8 */
  double ls_optimistic;   /* 1.25 */
  double ls_pessimistic;  /* 0.5 */
  double ls_armijo;       /* 0.5 (Standard/Conjugated only) */
  int grid_blocks;        /* 0 = automatic; else persistent grid size (clamped) */
  int block_threads;      /* 0 = automatic (512); else 256 or 512 threads per block (the solver keeps
                             operator data in up to 156 KB of LDS per workgroup, so a grid of
                             256-thread workgroups holds one per CU, half the waves) */
  int precision;          /* sb_precision, FP64 by default */
  double mixed_tol_factor;/* mixed: the FP32 stage stops at tol * factor (0.95; its residual is evaluated in FP64
                             from the FP32 iterate and product, within 1e-4 relative of the FP64 stage's
                             verification, which refines if it is not below tol) */
  int mixed_stall_iters;  /* mixed: the FP32 stage also stops after this many iterations without a 1%
                             decrease of its best residual once that is below 10 tol (its rounding
                             floor), or after 4x as many anywhere (100) */
} sb_pg_options;

/* GaussSeidel parameters (GaussSeidelBase.hpp:39-91) */
typedef struct sb_gs_options {
  int law;                /* sb_law; DualFrictionProblem::solveWith(gs) defaults to Coulomb */
  int strategy;           /* sb_local_strategy, RevHybrid by default */
  int inf_norm;
  int try_zero;           /* doTryZeroAsWell, 1 */
  double tol;             /* 1e-6 */
  int max_iters;          /* 250 */
  int eval_every;         /* 25 */
  double skip_tol;        /* 1e-6 */
  int skip_iters;         /* 10 */
  double auto_regularization; /* 0 */
  int grid_blocks;        /* 0 = automatic */
  int block_threads;      /* 0 = automatic (256, the only size of the GS kernel) */
  int precision;          /* sb_precision, FP64 by default. Mixed (sb_gs_solve only): sweeps whose
                             off-diagonal gathers W_ij x_j read FP32 copies of W and x (local solve and
                             FB evaluation in FP64) until tol * mixed_tol_factor or a stall of their
                             residual, then FP64 sweeps from the best iterate (at least one, the final
                             refinement sweep, before the first FP64 evaluation) until tol; the returned
                             residual is an FP64 evaluation */
  double mixed_tol_factor;/* 0.95 */
} sb_gs_options;

typedef struct sb_stats {
  int iterations;         /* PG iterations / GS sweeps */
  int spmvs;              /* operator applications */
  int ls_trials;          /* line-search trials (APGD/Descent) */
  int nesterov_steps;
  int converged;
  int local_failures;     /* GS: local solves that did not reach the local tolerance */
  double residual;        /* the value returned by the reference solve() */
  double device_ms;       /* kernel time of the solve (HIP events) */
  double* history;        /* optional caller buffer: residual per evaluation (callback stream) */
  int history_capacity;
  int history_length;     /* written: number of evaluations performed */
  int refine_iterations;  /* mixed precision: iterations of the FP64 stage (0: verified at once) */
  int refine_spmvs;       /* mixed precision: FP64 operator applications (verification included) */
  double mixed_residual;  /* mixed precision: residual reached by the FP32 stage */
  int plain_launch;       /* 1: a persistent kernel of this solve was launched without the cooperative
                             launch's residency guarantee (the runtime refused it, or SOBOGUS_LAUNCH=regular);
                             such a grid is sized to the occupancy limit, and a grid that is not resident
                             ends in SB_ERR_BARRIER_TIMEOUT */
  double mixed_device_ms; /* mixed precision: kernel time of the FP32 stage (HIP events); the FP64 stage
                             took device_ms - mixed_device_ms */
} sb_stats;

const char* sb_status_string(int status);
void sb_pg_options_default(sb_pg_options* o);
void sb_gs_options_default(sb_gs_options* o);

/* ---------------------------------------------------------------- context */
int sb_ctx_create(sb_ctx** ctx, int device);
int sb_ctx_destroy(sb_ctx* ctx);
/* Device name and compute-unit count (for logs). */
int sb_ctx_info(sb_ctx* ctx, char* name, int name_len, int* n_cu);
/* Persistent-kernel launches on this context since its creation: cooperative (every
   workgroup resident, guaranteed by the runtime) and plain (a refused cooperative launch,
   SOBOGUS_LAUNCH=regular, and the batch graphs' occupancy-sized grids). No reference
   counterpart (the reference runs on the CPU). */
int sb_ctx_launch_counts(sb_ctx* ctx, int64_t* cooperative, int64_t* plain);

/* ---------------------------------------------------------------- BSR (SparseBlockMatrix) */
/* Creates a device BSR matrix of nbrows x nbcols blocks of br x bc (br, bc <= 8). With
   SB_BSR_SYMMETRIC, only the lower triangle (col <= row) is given, blocks square. */
int sb_bsr_create(sb_ctx* ctx, sb_bsr** A, int nbrows, int nbcols, int br, int bc, int flags,
                  const int32_t* rowptr, const int32_t* colidx, const double* vals);
int sb_bsr_destroy(sb_bsr* A);
/* res = alpha * op(A) * x + beta * res (SparseBlockMatrixBase::multiply<Transpose>).
   x and res are host arrays when on_device == 0, device pointers otherwise. */
int sb_bsr_multiply(sb_ctx* ctx, const sb_bsr* A, int transpose, double alpha, const double* x, double beta,
                    double* res, int on_device);
/* res += sum_{j != row} A_{row,j} x_j for a symmetric 3x3 matrix (splitRowMultiply), host arrays. */
int sb_bsr_split_row_multiply(sb_ctx* ctx, const sb_bsr* A, int row, const double* x, double* res3);

/* ---------------------------------------------------------------- friction problem */
/* A dual problem u = W r + b with W = H Minv H^T, built from the primal data:
   H: n_contacts x n_bodies BSR of 3x6 blocks (row-major), Minv: n_bodies 6x6 blocks
   (explicit inverse mass, row-major), b: 3 n_contacts, mu: n_contacts.
   With SB_PROBLEM_ASSEMBLE_W the library also assembles W (DualFrictionProblem::computeFrom).
   The arrays are copied to the device and the set-up (tiles of H and H^T, the stable transpose
   of H, W's pattern, product terms and values, diagonal blocks) runs there, unless
   SB_PROBLEM_HOST_SETUP is given; host copies of the pattern and values are made on demand by
   the calls that need them (coloring, W export, permutation, host updates). */
int sb_problem_create(sb_ctx* ctx, sb_problem** P, int n_contacts, int n_bodies, const int32_t* h_rowptr,
                      const int32_t* h_colidx, const double* h_vals, const double* minv, const double* b,
                      const double* mu, int flags);
/* The same from device arrays (a contact set produced on the GPU): checked on the device
   (row pointer non-decreasing from 0, columns in range: SB_ERR_INVALID otherwise), copied,
   and set up there. Replaces DualFrictionProblem::computeFrom (FrictionProblem.cpp:30-61) with
   its transpose (SparseBlockMatrixBase.impl.hpp:169-202) and product pattern
   (SparseMatrixMatrixProduct.impl.hpp:124-212) for a changed contact set. */
int sb_problem_create_device(sb_ctx* ctx, sb_problem** P, int n_contacts, int n_bodies, const int32_t* d_rowptr,
                             const int32_t* d_colidx, const double* d_vals, const double* d_minv, const double* d_b,
                             const double* d_mu, int flags);
/* The tile layout the solvers use for H (which 0), H^T (1) or W with both triangles (2; only
   with an assembled W): sizes[6] = rows, blocks, tiles, long rows, scan steps, 1 if every row
   holds two blocks; then, each optional (NULL: skipped), tstart (tiles + 1), brow, bcol (blocks),
   tlong (tiles), long_rows (3 x long rows: row, first tile, end tile), rowptr (rows + 1), src
   (blocks: H and H^T the H block, W the block of the two-triangle W; -1 an empty row's zero
   block) and vals ([element][block], 18 or 9 elements). No counterpart in the reference (it
   has no tiled layout): for tests and tools that compare set-ups. */
int sb_problem_layout(sb_ctx* ctx, const sb_problem* P, int which, int64_t* sizes, int32_t* tstart, int32_t* brow,
                      int32_t* bcol, int32_t* tlong, int32_t* long_rows, int32_t* rowptr, int32_t* src,
                      double* vals);
/* Supplies an assembled W (lower triangle, 3x3 row-major blocks) instead of assembling it. */
int sb_problem_set_w(sb_ctx* ctx, sb_problem* P, const int32_t* w_rowptr, const int32_t* w_colidx,
                     const double* w_vals);
/* A dual problem given by W itself, with no H or Minv: the path of a local problem read from a file
   (apps/FCLibLoader.cpp:64-73: bogus::convert(W, dual.W, 3, 3); dual.W.prune(tol); b = q; mu).
   W is the scalar CSR of the 3n x 3n matrix (columns ascending within a row, as compressed row
   storage). It is converted as bogus::convert does into a symmetric block matrix
   (Core/Eigen/SparseConversions.hpp:26-95): 3x3 blocks of the lower triangle, each scalar row read up
   to its first entry right of its diagonal block, diagonal blocks replaced by (D + D^T) / 2; then
   blocks whose entries are all <= prune_tol in magnitude are removed (SparseBlockMatrixBase::prune,
   Core/Block/SparseBlockMatrixBase.impl.hpp:286-305). Only the assembled operator exists for such a
   problem: the matrix-free operator, sb_pgs_solve, sb_adjoint and Minv updates return
   SB_ERR_UNSUPPORTED or SB_ERR_INVALID. */
int sb_problem_create_from_w(sb_ctx* ctx, sb_problem** P, int n_contacts, const int32_t* w_rowptr,
                             const int32_t* w_colidx, const double* w_vals, const double* b, const double* mu,
                             double prune_tol);
int sb_problem_destroy(sb_problem* P);
int sb_problem_sizes(const sb_problem* P, int* n_contacts, int* n_bodies, int64_t* h_blocks, int64_t* w_blocks_lower);
/* Copies out the assembled lower-triangle W (arrays sized by sb_problem_sizes). */
int sb_problem_get_w(sb_ctx* ctx, const sb_problem* P, int32_t* w_rowptr, int32_t* w_colidx, double* w_vals);
/* Contact permutation, new[i] = old[perm[i]] (DualFrictionProblem::applyPermutation). */
int sb_problem_apply_permutation(sb_ctx* ctx, sb_problem* P, const int64_t* perm);
/* Coloring of the contacts (Coloring::update(true, W), Coloring.impl.hpp:20-62); perm has
   n_contacts entries and colors n_colors + 1 offsets (allocate n_contacts + 1). Two contacts
   are coupled when they share a body (the block pattern of W), so a problem without an
   assembled W is colored from H. The greedy rule runs on the device (in rounds that give
   the sequential rule's colors), the reference rule on the host, as does the greedy rule
   where the device's limits are exceeded (a body with more than 256 contacts, more than
   256 colors, a chain of more than 4096 coupled contacts) and for a W given by
   sb_problem_set_w. */
int sb_problem_coloring(sb_ctx* ctx, const sb_problem* P, int mode, int64_t* perm, int64_t* colors, int* n_colors);

/* y = W x (op = SB_OP_ASSEMBLED) or H (Minv (H^T x)) (SB_OP_MATRIX_FREE). Host or device arrays. */
int sb_problem_apply_w(sb_ctx* ctx, const sb_problem* P, int op, const double* x, double* y, int on_device);
/* Fischer-Burmeister residual of r (DualFrictionProblem::evalWith): u = W r + b, then
   sum_i |FB(s_i r_i, u_i)|^2 / (n+1), or the max with inf_norm. */
int sb_fb_residual(sb_ctx* ctx, const sb_problem* P, int law, int op, int scaling, int inf_norm, const double* r,
                   double* residual, int on_device);
/* FCLib's merit1 of r for the dual problem (apps/FCLibLoader.cpp:40-61, the error the FCLib solver
   reports): u = W r + b with the assembled (SB_OP_ASSEMBLED) or matrix-free operator;
   merit = sqrt(sum_i |P_K(r_i - u_i - (mu_i |u_iT|, 0, 0)) - r_i|^2) / (1 + |b|), P_K the projection on
   the cone of aperture mu_i (SOCLaw::projectOnConstraint). */
int sb_merit1(sb_ctx* ctx, const sb_problem* P, int op, const double* r, int on_device, double* merit);
/* A local problem in this library's own file layout (FCLib's HDF5 container is not part of this
   build; the fields are fclib_local's for a pure 3D Coulomb problem, apps/FCLibLoader.cpp:152-218:
   W as the scalar CSR of the 3n x 3n matrix, q, mu) with an optional initial r0 (NULL: none).
   Little-endian: magic "SBLOCAL1", int32 spacedim (3), int32 n, int64 nnz, int32 flags (bit 0: r0),
   int32 0, int32 rowptr[3n+1], int32 colidx[nnz], double vals[nnz], double q[3n], double mu[n],
   [double r0[3n]]. sb_local_problem_read checks the header, the file length, the row pointer and the
   columns (SB_ERR_IO otherwise); with every array argument NULL it returns the sizes only. */
int sb_local_problem_write(const char* path, int n_contacts, const int32_t* w_rowptr, const int32_t* w_colidx,
                           const double* w_vals, const double* q, const double* mu, const double* r0);
int sb_local_problem_read(const char* path, int* n_contacts, int64_t* nnz, int* has_r0, int32_t* w_rowptr,
                          int32_t* w_colidx, double* w_vals, double* q, double* mu, double* r0);

/* ProjectedGradient::solve on the problem; r is warm start and result (host or device). */
int sb_pg_solve(sb_ctx* ctx, const sb_problem* P, const sb_pg_options* opts, double* r, int on_device,
                double* residual, sb_stats* stats);
/* sb_pg_solve on a device-resident r that allocates nothing: no hipMalloc, hipFree or other
   allocating call. No counterpart in the reference; it is for callers that cannot allocate
   device memory (tight serving or simulation loops, harnesses that forbid allocation). The
   persistent launches' grid state (barrier words, error word, reduction partials, solver
   results; about 400 KB) lives in a fixed array of the library, one per GPU, held until the
   process ends, instead of the context's on-demand buffers; the iterates use the problem's
   own workspace (allocated by sb_problem_create). Limits: a persistent grid of at most 2048
   workgroups (SB_ERR_LIMIT beyond; the automatic grid is one workgroup per CU); no residual
   history (stats->history must be NULL); a matrix-free solve needs the factored operator
   tiles allocated by one earlier sb_pg_solve with SB_OP_MATRIX_FREE on this problem
   (SB_ERR_LIMIT otherwise; after sb_problem_update_values they are recomputed in place).
   Calls that overlap on one GPU (several host threads or contexts at once) race on the
   shared array: one call at a time per GPU. Results equal sb_pg_solve's bit for bit. */
int sb_pg_solve_noalloc(sb_ctx* ctx, const sb_problem* P, const sb_pg_options* opts, double* r_dev,
                        double* residual, sb_stats* stats);
/* GaussSeidel::solve on an assembled problem whose contacts are color-contiguous:
   colors holds n_colors + 1 offsets; a color must hold no two coupled contacts (W_ij != 0),
   else SB_ERR_INVALID. With colors NULL the contacts are swept in their order (the
   reference's sequential sweep), in runs of consecutive uncoupled contacts. */
int sb_gs_solve(sb_ctx* ctx, const sb_problem* P, const sb_gs_options* opts, const int64_t* colors, int n_colors,
                double* r, int on_device, double* residual, sb_stats* stats);
/* Matrix-free Gauss-Seidel: ProductGaussSeidel<H, Minv>::solve (ProductGaussSeidel.impl.hpp:
   128-304) as reached through PrimalFrictionProblem::solveWith(ProductGaussSeidelType&, r,
   staticProblem) (PrimalFrictionProblem.cpp:40-59). Needs no assembled W; the body vector
   Minv H^T r is kept current through the sweeps. Same options and results as sb_gs_solve.
   The contacts of one color must share no body (sb_problem_coloring + permutation), else
   SB_ERR_INVALID; with colors NULL the contacts are swept in their order (the reference's
   sequential sweep), in runs of consecutive contacts that share no body. */
int sb_pgs_solve(sb_ctx* ctx, const sb_problem* P, const sb_gs_options* opts, const int64_t* colors, int n_colors,
                 double* r, int on_device, double* residual, sb_stats* stats);

/* Device-side set-up for an unchanged pattern (the per-step update of a simulation whose
   contact set is the same): new H block values (3x6 row-major, in the problem's current
   contact and block order), inverse masses (6x6 row-major per body), b and mu, any of them
   NULL to keep it; host (on_device = 0) or device arrays. The H and H^T tiles, Minv, the
   diagonal blocks, row scalings and an assembled W = H Minv H^T are recomputed on the GPU
   (the work of setMatrix / computeFrom, FrictionProblem.cpp:30-40, without the symbolic
   phase). SB_ERR_INVALID when W was supplied by sb_problem_set_w and H or Minv changes. */
int sb_problem_update_values(sb_ctx* ctx, sb_problem* P, const double* h_vals, const double* minv, const double* b,
                             const double* mu, int on_device);
/* Minv_k = M_k^-1 for n_bodies 6x6 row-major blocks on the GPU (PrimalFrictionProblem::
   computeMInv, PrimalFrictionProblem.cpp:28-38; Gauss-Jordan with partial pivoting).
   SB_ERR_INVALID if a block is singular. */
int sb_mass_inverse(sb_ctx* ctx, int n_bodies, const double* M, double* minv, int on_device);

/* Cadoux fixed point (DualFrictionProblem::solveCadoux / solveCadouxVel,
   src/Interfaces/FrictionProblem.cpp:68-100, Cadoux.hpp:44-155): up to `iterations` SOC
   solves with the inner solver, each with the affine term b + s(W r + b) (s the De Saxce
   term mu |u_T| of the previous iterate), at a tolerance tightened to
   tol_tighten * max(tol, min(res, 1)); res is the Coulomb residual of the iterate with
   the inner solver's scaling. velocity = 1: solveCadouxVel, x is u and the problem is
   read as r = W u + b. The inner options' law is ignored (SOC). stats: iterations =
   fixed-point iterations, history = residual before and after each, spmvs = all
   operator applications. Returns SB_OK when res < tol, SB_NOT_CONVERGED otherwise. */
int sb_pg_solve_cadoux(sb_ctx* ctx, const sb_problem* P, const sb_pg_options* opts, int iterations,
                       double tol_tighten, int velocity, double* x, int on_device, double* residual,
                       sb_stats* stats);
int sb_gs_solve_cadoux(sb_ctx* ctx, const sb_problem* P, const sb_gs_options* opts, const int64_t* colors,
                       int n_colors, int iterations, double tol_tighten, int velocity, double* x, int on_device,
                       double* residual, sb_stats* stats);

/* ---------------------------------------------------------------- derivatives
   Reverse-mode derivatives of L = g . r for the projected fixed-point map of the
   dual problem, T(r) = P_mu(r - rho (W r + b + s_mu(W r + b))) (s_mu: the De Saxce
   term mu |u_T| e_N of the Coulomb law, 0 for SOC), whose fixed points are the
   solutions of the projected-gradient solvers. No counterpart in the reference (it
   has no differentiation). Matrix-free (needs H and Minv). Host arrays. */
enum sb_adjoint_mode {
  SB_ADJOINT_UNROLL = 0,   /* L = g . T^K(r0): exact gradient of K steps, checkpointed */
  SB_ADJOINT_IMPLICIT = 1  /* L = g . r* at a fixed point r0 = r*: implicit function theorem, the adjoint
                              system (I - dT/dr^T) lambda = g solved by restarted GMRES */
};
typedef struct sb_adjoint_options {
  int mode;               /* sb_adjoint_mode */
  int law;                /* sb_law (Coulomb by default) */
  int iterations;         /* unroll: K (200) */
  int checkpoint_every;   /* unroll: C (0: ceil(sqrt(K))) */
  double rho;             /* step of the map; 0: 1 / lambda_max(W) by power iteration */
  double tol;             /* implicit: relative residual |g - (I - A^T) lambda| / |g| (1e-10) */
  int max_iters;          /* implicit: GMRES (Arnoldi) steps in all, over the restarts (3000) */
  int grid_blocks;        /* 0 = automatic */
  int branch_mode;        /* unroll: 0 branches from the cone tests; 1 also record them into branches
                             (and, Coulomb law, the De Saxce directions u_T / |u_T| of every step, kept
                             on the device with the problem); 2 replay them (the projection formula of
                             the recorded branch, and mu (d . u_T) with the recorded direction d for
                             mu |u_T|: the map is then one smooth piece, with the recorded run's values
                             and gradient, and finite differences see no kinks; a Coulomb replay needs
                             a record of the same K on this problem) */
  int krylov_dim;         /* implicit: GMRES restart length, 1..64 (40) */
} sb_adjoint_options;
typedef struct sb_adjoint_stats {
  double rho;             /* step used */
  double loss;            /* g . r_K (unroll) or g . r* (implicit) */
  int iterations;         /* unroll: K; implicit: Arnoldi steps */
  int converged;          /* implicit: relative residual below tol */
  double change;          /* implicit: final relative residual of the adjoint system (recomputed) */
  double device_ms;
} sb_adjoint_stats;
void sb_adjoint_options_default(sb_adjoint_options* o);
/* r0: start of the unroll or the fixed point (3 n_c). g: dL/dr (3 n_c) or NULL for
   the forward map only. r_out (optional): T^K(r0). grad_b (3 n_c), grad_mu (n_c),
   grad_minv (36 n_b, row-major 6x6 per body): outputs, each optional. branches:
   K * n_c bytes (step-major) when branch_mode is 1 or 2, else NULL. */
int sb_adjoint(sb_ctx* ctx, const sb_problem* P, const sb_adjoint_options* opts, const double* r0, const double* g,
               double* r_out, double* grad_b, double* grad_mu, double* grad_minv, unsigned char* branches,
               sb_adjoint_stats* stats);
/* Replaces b, mu and/or Minv (NULL: unchanged) of a problem in place (H kept);
   W, its diagonal and the row scalings are recomputed. */
int sb_problem_update(sb_ctx* ctx, sb_problem* P, const double* b, const double* mu, const double* minv);

/* ---------------------------------------------------------------- batches
   Many independent dual friction problems solved in one persistent launch: each
   workgroup takes whole problems from a queue, largest first, and solves them
   alone (APGD / Descent / SPG, DualFrictionProblem::solveWith(pg, r) semantics
   per problem). The launch sequence is captured once in a hipGraph and replayed.
   The reference has no batched entry point; this is the device-side equivalent
   of a loop of solveWith calls over independent problems. */
typedef struct sb_batch sb_batch;
typedef struct sb_batch_config {
  int block_threads;      /* 256, 512 or 1024 threads per workgroup; 0: tuned or default */
  int blocks_per_cu;      /* workgroups launched per CU (1..8); 0: tuned or default */
  int contacts_per_block; /* team sizing: a problem of n_c contacts is solved by
                             pow2ceil(n_c / contacts_per_block) workgroups (at most 16);
                             0: tuned or default (4000); < 0: one workgroup per problem */
} sb_batch_config;
typedef struct sb_batch_stats {
  int n_converged;        /* problems whose FP64 residual is below tol */
  int total_iterations;   /* summed over problems and stages */
  int max_iterations;     /* largest per-problem iteration count */
  int refined_problems;   /* mixed precision: problems whose FP64 stage had to iterate */
  int64_t total_spmvs;    /* operator applications summed over problems and stages */
  double device_ms;       /* graph execution time (HIP events) */
  sb_batch_config config; /* geometry of the main stage */
  int* iterations;        /* optional caller array [n_problems] */
  double* residuals;      /* optional caller array [n_problems]: FP64 residual of each result */
  int* converged;         /* optional caller array [n_problems] */
  int* stage_spmvs;       /* optional caller array [n_problems][2]: operator applications of the
                             FP32 stage (0 in FP64 precision) and of the FP64 stage */
  double* phase_us;       /* optional caller array [n_problems][2][4]: per stage (FP32, FP64), the
                             team's time in operator phases (work, then reduction and barrier
                             wait) and in update phases (work, wait), in microseconds */
} sb_batch_stats;

/* Problems k = 0..n_problems-1 as in sb_problem_create (flags: SB_PROBLEM_ASSEMBLE_W). Each
   problem is stored in a locality order (bodies renumbered by first touch, contacts sorted
   by their bodies: the operator's gathers touch fewer cache lines) unless
   SB_BATCH_CALLER_ORDER is set; r and every per-problem result stay in the caller's order. */
int sb_batch_create(sb_ctx* ctx, sb_batch** B, int n_problems, const int* n_contacts, const int* n_bodies,
                    const int32_t* const* h_rowptr, const int32_t* const* h_colidx, const double* const* h_vals,
                    const double* const* minv, const double* const* b, const double* const* mu, int flags);
int sb_batch_destroy(sb_batch* B);
int sb_batch_sizes(const sb_batch* B, int* n_problems, int64_t* total_contacts);
/* r: the problems' 3*n_contacts vectors concatenated in problem order (warm start in,
   result out); host or device (on_device) memory. cfg may be NULL. */
int sb_batch_solve(sb_ctx* ctx, sb_batch* B, const sb_pg_options* opts, const sb_batch_config* cfg, double* r,
                   int on_device, sb_batch_stats* stats);
/* Autotuner: times every launch geometry the kernels allow on this batch (each
   capped at probe_iters iterations per problem, r0 = 0), records each problem's
   iteration count in one full solve (long-running problems then get larger
   teams), times the two best geometries of each team sizing on full solves, and keeps
   the fastest for later solves of this batch. Solves are reproducible between
   calls to this function and sb_batch_set_iteration_hints: only they change the
   schedule. */
int sb_batch_autotune(sb_ctx* ctx, sb_batch* B, const sb_pg_options* opts, int probe_iters, sb_batch_config* best,
                      double* best_ms);
/* Expected iteration count of every problem (hints[n_problems], both stages together;
   0 = unknown), used by the planner as sb_batch_autotune's recorded counts are: a
   problem expected to iterate longer gets a larger team and starts earlier. Callers
   pass, for example, the counts of the previous time step's batch, or counts predicted
   from the problem sizes by a model fitted on another batch. They change the schedule
   only (every result still meets the tolerance). hints = NULL forgets them (every
   problem weighted alike). sb_batch_get_iteration_hints reads the counts in use
   (recorded by sb_batch_autotune or set here). No counterpart in the reference: it
   describes the device schedule of a loop of independent solveWith calls. */
int sb_batch_set_iteration_hints(sb_batch* B, const int* hints);
int sb_batch_get_iteration_hints(const sb_batch* B, int* hints);

/* ---------------------------------------------------------------- raw device memory helpers
   (for callers that keep vectors resident in HBM between calls) */
int sb_device_alloc(sb_ctx* ctx, void** ptr, size_t bytes);
int sb_device_free(sb_ctx* ctx, void* ptr);
int sb_memcpy_h2d(sb_ctx* ctx, void* dst, const void* src, size_t bytes);
int sb_memcpy_d2h(sb_ctx* ctx, void* dst, const void* src, size_t bytes);
int sb_synchronize(sb_ctx* ctx);
/* Sets bytes of device memory to value (stream-ordered, completes before return). */
int sb_memset_device(sb_ctx* ctx, void* dst, int value, size_t bytes);

/* ---------------------------------------------------------------- benchmarking hooks
   Times `reps` back-to-back launches of one block-SpMV kernel of the problem on the
   context's stream with HIP events; kind: 0 = assembled W x, 1 = H w, 2 = H^T r (+ Minv),
   3 = matrix-free W x (H^T+Minv, then H). Returns the mean milliseconds per launch. With
   SB_BENCH_FLUSH_CACHES or-ed into kind, every launch is timed alone after a 768 MB copy that
   evicts its data from L2 and the Infinity Cache (the product then reads it from HBM). */
#define SB_BENCH_FLUSH_CACHES 16
int sb_bench_spmv(sb_ctx* ctx, const sb_problem* P, int kind, int reps, double* ms_per_launch);
/* Times `reps` launches of a plain device-to-device copy of `bytes` (read + write
   traffic 2 * bytes per launch): achievable-bandwidth and counter calibration. */
int sb_bench_copy(sb_ctx* ctx, size_t bytes, int reps, double* ms_per_launch);
/* Phase-time breakdown of the last projected-gradient solve on this context, as seen
   by block 0 of the persistent kernel (microseconds): operator-phase work, operator-phase
   reduction + barrier wait, update-phase work, update-phase reduction + barrier wait. */
int sb_last_phase_times(sb_ctx* ctx, double* us4);

#ifdef __cplusplus
}
#endif

#endif /* SOBOGUS_HIP_H */
