// Host-side setup of the device layout: tiling of block-sparse matrices,
// assembly of W = H Minv H^T, coloring and permutations. Plain C++ (no HIP).
#pragma once
#include <cstdint>
#include <vector>

namespace sbhost {

// Tiled lane-per-block layout (see sb_tiles.hpp). Values are element-major.
struct Tiles {
  int nrows = 0, ncols = 0, br = 0, bc = 0;
  int nblocks = 0, ntiles = 0, scan_steps = 0;
  std::vector<int> tstart, brow, bcol, tlong, long_row, long_t0, long_t1;
  std::vector<int> rowptr;   // per-row block offsets in tile order
  std::vector<int> src;      // per tile-order block: its block in the source BSR, -1 for an empty row's zero block
  std::vector<double> vals;  // [br*bc][nblocks]
};

// Builds tiles from a row-major BSR (blocks in row order, vals row-major per
// block). Empty rows receive one zero block so that every row has a leader.
Tiles buildTiles(int nrows, int ncols, int br, int bc, const int* rowptr, const int* colidx, const double* vals,
                 bool transposeBlocks = false);

// Transposes a BSR (blocks become bc x br, rows <-> cols).
// tsrc (optional): for each transposed block, the index of its source block.
void transposeBSR(int nrows, int ncols, int br, int bc, const int* rowptr, const int* colidx, const double* vals,
                  std::vector<int>& trowptr, std::vector<int>& tcolidx, std::vector<double>& tvals,
                  std::vector<int>* tsrc = nullptr);

// Expands a symmetric lower-triangle BSR of square blocks into both triangles.
void expandSymmetric(int n, int bs, const int* rowptr, const int* colidx, const double* vals, std::vector<int>& frowptr,
                     std::vector<int>& fcolidx, std::vector<double>& fvals);

// W = H Minv H^T, lower triangle, 3x3 blocks (DualFrictionProblem::computeFrom).
void assembleW(int nc, int nb, const int* hrowptr, const int* hcolidx, const double* hvals, const double* minv,
               std::vector<int>& wrowptr, std::vector<int>& wcolidx, std::vector<double>& wvals);

// Diagonal blocks W_ii: the sum over pairs of contact i's blocks at the same body,
// as the assembled W sums them.
void diagBlocks(int nc, const int* hrowptr, const int* hcolidx, const double* hvals, const double* minv,
                std::vector<double>& diag);

// Coloring of a symmetric lower-triangle block pattern.
//  mode 0: Coloring::compute as written (conflict when lower_bound(color, j) != end);
//  mode 1: first fit with an exact membership test.
void coloring(int n, const int* wrowptr, const int* wcolidx, int mode, std::vector<int64_t>& perm,
              std::vector<int64_t>& colors);

// The same two colorings computed from the contact-body incidence of H (two
// contacts are coupled iff they share a body, which is the block pattern of
// W = H Minv H^T), for problems whose W is not assembled.
void coloringFromH(int nc, int nb, const int* hrowptr, const int* hcolidx, int mode, std::vector<int64_t>& perm,
                   std::vector<int64_t>& colors);

// Locality order of a contact problem: bodies renumbered by first touch along the
// contacts sorted by (lowest body, highest body), then the contacts sorted by their
// renumbered (lowest, highest) body; contacts without bodies go last. cperm[new] = old
// contact, bmap[old] = new body. Consecutive contacts then share or neighbour their
// first body, so the gathers of both operator passes touch fewer cache lines.
void localityOrder(int nc, int nb, const int* hrowptr, const int* hcolidx, std::vector<int>& cperm,
                   std::vector<int>& bmap);


// Places the teams of a batch launch on its grid. Workgroups are dispatched to the
// XCDs round-robin (block b runs on XCD b % nxcd), so a team of k blocks
// b = nxcd * s + x, s = s0 .. s0 + k - 1, runs on one XCD. Classes c = 0..nclass-1
// (team sizes k[c], powers of two in decreasing order, nteams[c] teams each) are
// placed team by team, each onto the XCD with the fewest blocks in use: every
// XCD's share of the grid is a multiple of the largest k, so every team fits.
// Without that (or with spread) teams take consecutive blocks. Returns
// bmap[b] = c << 20 | team << 4 | rank, -1 for an idle block; empty when the
// teams need more blocks than the grid has.
std::vector<int> batchBlockMap(int grid, int nxcd, int nclass, const int* k, const int* nteams, bool spread);
}  // namespace sbhost
