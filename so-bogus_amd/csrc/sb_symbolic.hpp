// Symbolic set-up on the device (sb_symbolic.hip): the tile layout of a BSR
// pattern, its transposition, and the pattern of W = H Minv H^T with the index
// maps of W's numeric phase. Every launcher works on device arrays given by the
// caller; the sort, scan and reduction steps take a temporary buffer whose size
// they report when called with tmp == nullptr.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace sb {

// ---- library steps (hipCUB): tmp == nullptr only sets bytes
hipError_t scan_inclusive_int(const int* in, int* out, int n, void* tmp, size_t& bytes, hipStream_t s);
hipError_t reduce_max_int(const int* in, int* out, int n, void* tmp, size_t& bytes, hipStream_t s);
hipError_t sort_pairs_u32(const unsigned* kin, unsigned* kout, const int* vin, int* vout, int n, int end_bit,
                          void* tmp, size_t& bytes, hipStream_t s);
hipError_t sort_keys_u64(const uint64_t* kin, uint64_t* kout, int n, int end_bit, void* tmp, size_t& bytes,
                         hipStream_t s);

hipError_t reduce_sum_int64(const int* in, long long* out, int n, void* tmp, size_t& bytes, hipStream_t s);

// ---- pattern checks (err = 1 on a violation): rp[0] == 0 and rp non-decreasing; then,
// with the row pointer known to be valid, every column index in [0, ncols)
hipError_t launch_check_rows(int nrows, const int* rp, int* err, hipStream_t s);
hipError_t launch_check_cols(int nnz, int ncols, const int* ci, int* err, hipStream_t s);

// ---- tiles of a BSR pattern (the layout sbhost::buildTiles builds on the host)
// tlen[i] = max(1, row length): an empty row holds one zero block
hipError_t launch_row_len(int nrows, const int* rp, int* tlen, hipStream_t s);
// per tile-order block q: row, column, source block (-1: an empty row's zero block)
hipError_t launch_tile_expand(int nrows, int nblocks, const int* trp, const int* rp, const int* ci, int* brow,
                              int* bcol, int* src, hipStream_t s);
// next[i]: the row that starts the tile after a tile starting at row i
hipError_t launch_tile_next(int nrows, const int* trp, const int* tlen, int* next, hipStream_t s);
// the rows that start tiles (the chain 0, next(0), ...), marked in is_start (nrows bytes)
int chain_chunk_rows(int nrows);
hipError_t launch_tile_chain(int nrows, const int* next, unsigned char* exits, int* entry,
                             unsigned char* is_start, hipStream_t s);
// tiles and long rows per row (scanned by the caller into tbase / lbase, nrows + 1 entries)
hipError_t launch_tile_counts(int nrows, const int* tlen, const unsigned char* is_start, int* ntl, int* nlg,
                              hipStream_t s);
hipError_t launch_tile_fill(int nrows, int nblocks, int ntiles, const int* trp, const int* tlen, const int* next,
                            const unsigned char* is_start, const int* tbase, const int* lbase, int* tstart,
                            int* tlong, int* long_row, int* long_t0, int* long_t1, hipStream_t s);

// ---- transposition (stable: within a column, source blocks in their order)
hipError_t launch_iota(int n, int* v, hipStream_t s);
// first index with keys[t] >= j, for j = 0..nout-1 (keys ascending)
hipError_t launch_lower_bounds_u32(int nout, const unsigned* keys, int n, int* out, hipStream_t s);
// row of each source block: rows[q] = the row of BSR block src[q]
hipError_t launch_rows_of(int n, const int* src, int nrows, const int* rp, int* rows, hipStream_t s);
// out[q] = tsrc[ssrc[q]] (or -1 where ssrc[q] < 0)
hipError_t launch_compose_src(int n, const int* ssrc, const int* tsrc, int* out, hipStream_t s);

// ---- W = H Minv H^T: lower-triangle pattern, expansion to both triangles, product terms
// cnt[i] = candidates of row i: for every block of contact i, the contacts j <= i at its body
hipError_t launch_w_cand_count(int nc, const int* hrp, const int* hci, const int* trp, const int* tci, int* cnt,
                               hipStream_t s);
hipError_t launch_w_cand_fill(int nc, const int* hrp, const int* hci, const int* trp, const int* tci,
                              const int* cptr, uint64_t* keys, hipStream_t s);
// keys sorted: flags[t] = 1 where keys[t] differs from keys[t - 1]
hipError_t launch_unique_flags(int n, const uint64_t* keys, int* flags, hipStream_t s);
hipError_t launch_unique_scatter(int n, const uint64_t* keys, const int* pos, uint64_t* out, hipStream_t s);
// keys k = row * nc + col (sorted, unique): rp[i] = first key of row i (i = 0..nc), ci = the columns
hipError_t launch_key_rows(int nc, const uint64_t* keys, int n, int* rp, int* ci, hipStream_t s);
// one key per lower entry: col * nc + row off the diagonal, nc * nc (sorted last) on it
hipError_t launch_upper_keys(int nc, const int* wrp, const int* wci, uint64_t* keys, hipStream_t s);
// uptr[i] = first upper key of row i (i = 0..nc); flen[i] = lower row length + upper entries
hipError_t launch_upper_rows(int nc, const uint64_t* ukeys, int n, const int* wrp, int* uptr, int* flen,
                             hipStream_t s);
// full rows: lower row i (columns <= i), then its upper entries (columns > i), both ascending
hipError_t launch_full_fill(int nc, const int* wrp, const int* wci, const int* uptr, const uint64_t* ukeys,
                            const int* frp, int* fci, hipStream_t s);
// product terms of every W tile block (i, j): (a, b) over the H blocks a of hi and b of lo at
// the same body, hi / lo the one of i, j with the larger / smaller assembly order key okey
hipError_t launch_w_term_count(int nq, const int* brow, const int* bcol, const int* src, const int* okey,
                               const int* hrp, const int* hci, int* cnt, hipStream_t s);
hipError_t launch_w_term_fill(int nq, const int* brow, const int* bcol, const int* src, const int* okey,
                              const int* hrp, const int* hci, const int* tptr, int* term, hipStream_t s);

hipError_t launch_fill_f64(size_t n, double v, double* out, hipStream_t s);

// dst[i] = src[idx[i]]
hipError_t launch_gather_int(int n, const int* idx, const int* src, int* dst, hipStream_t s);
// ---- contact permutation (new i <- old perm[i]): row lengths, then rows (nrp: the new
// row pointer), b and mu
hipError_t launch_perm_lens(int nc, const int* perm, const int* rp, int* len, hipStream_t s);
hipError_t launch_perm_rows(int nc, const int* perm, const int* rp, const int* ci, const double* hv, const double* b,
                            const double* mu, const int* nrp, int* nci, double* nhv, double* nb, double* nmu,
                            hipStream_t s);

// ---- first-fit coloring of the contacts (coupled: sharing a body), in rounds:
// color[nc] starts at -1; each round colors the contacts whose earlier neighbours are
// colored and adds the others to *left; *info is set if a contact needs more than
// coloring_max_colors() colors. trp / tc / tsrc: the body rows of H^T (tiled row
// pointer, contact per block, source block, -1 for an empty row's zero block).
hipError_t launch_max_degree(int nb, const int* trp, int* deg, hipStream_t s);
hipError_t launch_color_round(int nc, const int* hrp, const int* hci, const int* trp, const int* tc, const int* tsrc,
                              int* color, int* left, int* info, hipStream_t s);
int coloring_max_colors();

int ceil_log2_host(int64_t v);

}  // namespace sb
