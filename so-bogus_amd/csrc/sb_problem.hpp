// Device-resident problem layout and solver argument blocks (host + device).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sb_device.hpp"
#include "sb_tiles.hpp"

namespace sb {

// A dual friction problem u = W r + b resident in HBM.
//   H  : contact-major tiles, 3x6 blocks (H_ik)
//   HT : body-major tiles, 6x3 blocks (H_ik^T), for H^T r
//   with minv_folded, H and HT hold F_ik = H_ik L_k and F_ik^T instead (Minv_k = L_k L_k^T):
//        the matrix-free product is F (F^T x) with no body step
//   W  : assembled Delassus operator, both triangles expanded, 3x3 blocks (optional)
//   minv : inverse mass blocks, element-major [36][nb]
//   diag : W_ii blocks [nc][9]; scal : max(1, tr(W_ii)/3)
template <typename T>
struct DevProblem {
  int nc = 0, nb = 0;
  TileMat<T> H, HT, W;
  const T* minv = nullptr;
  const double* b = nullptr;
  const double* mu = nullptr;
  const double* diag = nullptr;
  const double* scal = nullptr;
  bool hasW = false;
  bool minv_folded = false;  // H / HT hold F = H L and F^T (Minv = L L^T): the body step is the identity
};

// Projected-gradient parameters shared by every problem of a launch.
struct PGParams {
  int variant;      // 0 Standard, 1 Descent, 2 Conjugated, 3 APGD, 4 SPG
  int coulomb;
  int inf_norm;
  int use_scaling;  // 1: residual row scaling from scal[], 0: all ones
  double tol;
  int max_iters;
  int ls_iters;
  double ls_opt, ls_pess;
  double ls_armijo;  // Standard / Conjugated: Armijo coefficient of armijo_ls
  int stall_iters;  // > 0: stop (not converged) after this many iterations without a 1% decrease of
                    // the best residual within 10 tol (4x as many above): the FP32 stage of a mixed
                    // solve hands over to FP64
};

// One problem's device data and workspace; T is the storage (and bulk
// arithmetic) type of the matrices and iterate vectors.
template <typename T>
struct PGProblem {
  DevProblem<T> P;
  T* X[4];   // iterate slots: current, previous, line-search candidate, Nesterov point
  T* MX[4];  // their operator images
  T* G[3];   // gradient slots: current, candidate, Nesterov point
  T* xbest;
  T* wbody;  // [6 nb]
  T* carry;
  double* x_io;  // in: warm start, out: result
  double* hist;
  int hist_cap;
  double* out;   // [16]: [0] residual [1] iterations [2] spmvs [3] ls trials [4] nesterov steps
                 // [5] converged [6] failed [7..10] phase times (us)
};

// Single problem, whole persistent grid.
template <typename T>
struct PGGridArgs {
  PGParams prm;
  PGProblem<T> pb;
  GridSync gs;
};

// Batch: the grid is split into classes of teams; a team is k workgroups (k = 1:
// one workgroup, synchronised by __syncthreads; k > 1: a co-resident group with
// its own barrier and reduction state), placed by the block map so that a team's
// workgroups share one XCD. Every team of a class takes problems from the class's
// queue (a range of order[], most expensive first) and solves them one at a time.
constexpr int kMaxBatchClasses = 8;
constexpr int kMaxTeamBlocks = 16;
constexpr unsigned kTeamSyncWords = 2 * kSyncGroups * kSyncStride + kSyncStride;  // per team
struct BatchClass {
  int b0;       // first block of the class (in block-map order)
  int k;        // blocks per team
  int nteams;
  int q0, q1;   // the class's problems: order[q0 .. q1)
  int team0;    // global index of the class's first team
};
template <typename T>
struct PGBatchArgs {
  PGParams prm;
  int nprob;
  const int* order;           // problem indices grouped by class, most expensive first
  const int* bmap;            // [grid] (class << 20 | team << 4 | rank) of each block, -1: idle
  int nclass;
  BatchClass cls[kMaxBatchClasses];
  unsigned* queue;            // [nclass] next position of each class's queue
  unsigned* team_sync;        // [nteams][kTeamSyncWords] barrier words of the k > 1 teams
  double* team_partials;      // [nteams][2][kMaxTeamBlocks][kMaxPartials]
  int* team_slot;             // [nteams] problem handed to the team
  int* error;                 // sticky barrier-timeout flag
  uint64_t timeout_ticks;
  // Helper teams: blocks whose own queue has drained form further teams of class 0's
  // size on their XCD and take class 0's problems. Team ids from nteams on (sync
  // words, partials, slot); pool_per_xcd = 0 turns helpers off.
  int nteams;
  int pool_per_xcd;           // helper teams per XCD at most
  unsigned* pool_count;       // [8][kSyncStride] blocks that joined a helper team, per XCD
  unsigned* pool_state;       // [8 * pool_per_xcd][kSyncStride] 0 forming, 1 running, 2 cancelled
};

// Reverse-mode derivatives of the projected fixed-point map (matrix-free, FP64).
constexpr int kMaxKrylov = 64;  // implicit mode: largest GMRES restart length
struct AdjArgs {
  DevProblem<double> P;
  int mode;             // 0: unrolled K steps from r0; 1: implicit at the fixed point r0
  int K, C;             // unroll length, checkpoint interval
  int coulomb;          // De Saxce term on (Coulomb law) or off (SOC law)
  double rho;           // step of the map
  double tol;           // implicit: relative residual |g - (I - A^T) lam| / |g| of the adjoint system
  int max_iters;        // implicit: GMRES (Arnoldi) steps in all
  int kdim;             // implicit: GMRES restart length (<= kMaxKrylov)
  const double* r0;     // [3 nc] start (unroll) or fixed point (implicit)
  const double* g;      // [3 nc] dL/dr (null: forward only)
  double* rK;           // [3 nc] out: r_K (unroll)
  double* ckpt;         // [ceil(K/C) + 1][3 nc]
  double* seg;          // [max(C, 2)][3 nc]
  double* lam;          // [2][3 nc]
  double* kry;          // implicit: [kdim + 1][3 nc] Krylov basis
  double* mz;           // [3 nc]
  double* mu_u;         // [3 nc]
  double* hr;           // [6 nb] H^T r of the current step
  double* jac;          // implicit: [nc][16] J_P (9), dP/dmu (3), u (3), |u_T|
  double* wbody;        // [6 nb]
  double* carry;
  double* grad_b;       // [3 nc] out
  double* grad_mu;      // [nc] out
  double* grad_minv;    // [nb][36] out (row-major 6x6 per body)
  unsigned char* branch;  // unroll: [K][nc] projection branch per step (null: not kept)
  int branch_mode;      // 0: branches from the projection tests, 1: record them, 2: replay them
  double* sdir;         // unroll, Coulomb law: [K][nc][2] tangential direction u_T / |u_T| of the De Saxce
                        // term per step (recorded with branch_mode 1, replayed with 2; null otherwise)
  GridSync gs;
  double* out;          // [0] loss [1] iterations [2] last relative residual [3] converged [4] failed
};

struct GSArgs {
  int nc;
  int coulomb;
  int strategy;
  int inf_norm;
  int try_zero;
  double tol;
  int max_iters;
  int eval_every;
  double skip_tol;
  int skip_iters;       // already min(skipIters, sqrt(n))
  double local_tol;
  DevProblem<double> P;
  const double* local;  // [nc][9] local matrices (regularized diagonal blocks)
  const double* reg;    // [nc]
  const int* colors;    // [ncolors + 1] offsets in the (color-contiguous) contact order
  int ncolors;
  double* x;            // in/out
  double* y;            // scratch [3 nc]
  double* xbest;
  unsigned char* skip;
  double* ndx;          // [2] ndxRef of the previous color / running max of the current color
  double* carry;
  double* hist;
  int hist_cap;
  GridSync gs;
  double* out;          // [0] residual [1] iterations [2] converged [3] local failures [4] evaluations
                        // [5] barrier failure [6] FP32-stage sweeps [7] FP32-stage residual [8] FP64 evaluations
  // mixed precision (gs_kernel): an FP32 stage (gathers of W32 and x32) to tol * stop_factor first
  int mixed;
  double stop_factor;
  TileMat<float> W32;
  float* x32;           // FP32 copy of x during the FP32 stage
  float* carry32;
};

// Extra inputs of the matrix-free (product) Gauss-Seidel.
struct PGSArgs {
  const int* hrow;      // [2 nc]: tile-order index of contact i's first H block, its number of blocks
  const double* dmt;    // [18][H.nblocks]: Minv_k H_ik^T (6x3, row-major) per H block, element-major
  double* ybody;        // [6 nb]: Minv H^T x, kept current through the sweeps
};

}  // namespace sb
