// Project-native file of a local (dual) friction problem: the fields that FCLib's
// fclib_local holds for a pure 3D Coulomb problem (apps/FCLibLoader.cpp:152-218 reads
// them from HDF5: W in compressed row storage, q, mu) and an optional initial r. HDF5
// is not part of this build; the layout below is this library's own, little-endian:
//
//   char     magic[8]      "SBLOCAL1"
//   int32    spacedim      3
//   int32    n_contacts    n > 0
//   int64    nnz           stored entries of W (3n x 3n)
//   int32    flags         bit 0: r0 present
//   int32    reserved      0
//   int32    rowptr[3n+1]  rowptr[0] = 0, non-decreasing, rowptr[3n] = nnz
//   int32    colidx[nnz]   in [0, 3n)
//   double   vals[nnz]
//   double   q[3n]         (b of the dual problem u = W r + b)
//   double   mu[n]
//   double   r0[3n]        when flags & 1
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/sobogus_hip.h"

namespace {

constexpr char kMagic[8] = {'S', 'B', 'L', 'O', 'C', 'A', 'L', '1'};

struct Header {
  char magic[8];
  int32_t spacedim, n;
  int64_t nnz;
  int32_t flags, reserved;
};
static_assert(sizeof(Header) == 32, "packed header");

struct File {
  std::FILE* f = nullptr;
  File(const char* path, const char* mode) : f(std::fopen(path, mode)) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

template <typename T>
bool put(std::FILE* f, const T* p, size_t n) {
  return n == 0 || std::fwrite(p, sizeof(T), n, f) == n;
}
template <typename T>
bool get(std::FILE* f, T* p, size_t n) {
  return n == 0 || std::fread(p, sizeof(T), n, f) == n;
}
// reads n values, into p when given, else past them
template <typename T>
bool get_or_skip(std::FILE* f, T* p, size_t n) {
  if (p) return get(f, p, n);
  return n == 0 || std::fseek(f, (long)(n * sizeof(T)), SEEK_CUR) == 0;
}

}  // namespace

extern "C" {

int sb_local_problem_write(const char* path, int n, const int32_t* rp, const int32_t* ci, const double* v,
                           const double* q, const double* mu, const double* r0) {
  if (!path || n <= 0 || !rp || !ci || !v || !q || !mu) return SB_ERR_INVALID;
  const int m = 3 * n;
  if (rp[0] != 0) return SB_ERR_INVALID;
  for (int i = 0; i < m; ++i)
    if (rp[i + 1] < rp[i]) return SB_ERR_INVALID;
  const int64_t nnz = rp[m];
  for (int64_t p = 0; p < nnz; ++p)
    if (ci[p] < 0 || ci[p] >= m) return SB_ERR_INVALID;
  File F(path, "wb");
  if (!F.f) return SB_ERR_IO;
  Header h;
  std::memcpy(h.magic, kMagic, 8);
  h.spacedim = 3;
  h.n = n;
  h.nnz = nnz;
  h.flags = r0 ? 1 : 0;
  h.reserved = 0;
  const bool ok = put(F.f, &h, 1) && put(F.f, rp, (size_t)m + 1) && put(F.f, ci, (size_t)nnz) &&
                  put(F.f, v, (size_t)nnz) && put(F.f, q, (size_t)m) && put(F.f, mu, (size_t)n) &&
                  (!r0 || put(F.f, r0, (size_t)m));
  if (!ok || std::fflush(F.f) != 0) return SB_ERR_IO;
  return SB_OK;
}

int sb_local_problem_read(const char* path, int* n_out, int64_t* nnz_out, int* has_r0, int32_t* rp, int32_t* ci,
                          double* v, double* q, double* mu, double* r0) {
  if (!path || !n_out || !nnz_out || !has_r0) return SB_ERR_INVALID;
  File F(path, "rb");
  if (!F.f) return SB_ERR_IO;
  Header h;
  if (!get(F.f, &h, 1) || std::memcmp(h.magic, kMagic, 8) != 0 || h.spacedim != 3 || h.n <= 0 ||
      h.n > (1 << 28) || h.nnz < 0 || (h.flags & ~1) != 0)
    return SB_ERR_IO;
  const int64_t m = 3 * (int64_t)h.n;
  // the file must hold exactly the arrays its header announces
  const int64_t want = (int64_t)sizeof(Header) + 4 * (m + 1) + 12 * h.nnz + 8 * m + 8 * (int64_t)h.n +
                       ((h.flags & 1) ? 8 * m : 0);
  if (std::fseek(F.f, 0, SEEK_END) != 0 || std::ftell(F.f) != (long)want ||
      std::fseek(F.f, (long)sizeof(Header), SEEK_SET) != 0)
    return SB_ERR_IO;
  *n_out = h.n;
  *nnz_out = h.nnz;
  *has_r0 = h.flags & 1;
  if (!rp && !ci && !v && !q && !mu && !r0) return SB_OK;  // sizes only
  std::vector<int32_t> rowptr((size_t)m + 1);
  if (!get(F.f, rowptr.data(), rowptr.size())) return SB_ERR_IO;
  if (rowptr[0] != 0 || rowptr[m] != h.nnz) return SB_ERR_IO;
  for (int64_t i = 0; i < m; ++i)
    if (rowptr[i + 1] < rowptr[i]) return SB_ERR_IO;
  if (rp) std::memcpy(rp, rowptr.data(), sizeof(int32_t) * rowptr.size());
  std::vector<int32_t> cols((size_t)h.nnz);
  if (!get(F.f, cols.data(), cols.size())) return SB_ERR_IO;
  for (int32_t c : cols)
    if (c < 0 || c >= m) return SB_ERR_IO;
  if (ci) std::memcpy(ci, cols.data(), sizeof(int32_t) * cols.size());
  if (!get_or_skip(F.f, v, (size_t)h.nnz) || !get_or_skip(F.f, q, (size_t)m) || !get_or_skip(F.f, mu, (size_t)h.n))
    return SB_ERR_IO;
  if ((h.flags & 1) && !get_or_skip(F.f, r0, (size_t)m)) return SB_ERR_IO;
  return SB_OK;
}

}  // extern "C"
