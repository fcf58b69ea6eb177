// System files: a sparse system handed over offline, e.g. assembled by femlisp and exported
// with sparse-matrix->ccs (src/matlisp/sparselu.lisp:235-333), or written by this library.
//
// Layout (little endian):
//   char    magic[8]   "FLSYS01" and a zero byte
//   int32   orientation 0: rows compressed (CSR), 1: columns compressed (CCS, femlisp's
//                       compressed-matrix with :orientation :column, compressed.lisp:37-53)
//   int32   flags       bit 0: right-hand side follows, bit 1: key permutation follows
//   int64   nrows, ncols, nnz
//   int32   starts[(orientation ? ncols : nrows) + 1], indices[nnz]
//   float64 values[nnz]
//   float64 rhs[nrows]          (flag bit 0)
//   int32   perm[nrows]         (flag bit 1; square systems) perm[k] is the canonical index of
//                               the file's row and column k, i.e. the rank of its DOF midpoint
//                               in femlisp's sort-lexicographically order (meshgen.lisp:214-217)
// Reading applies the permutation, turns columns into rows and sorts every row, so the device
// matrix is in canonical numbering with ascending columns whatever order femlisp's hash tables
// produced.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>

#include "flgpu_internal.h"

namespace flgpu {

namespace {

constexpr char kMagic[8] = {'F', 'L', 'S', 'Y', 'S', '0', '1', '\0'};

struct File {
  std::FILE *f = nullptr;
  explicit File(std::FILE *p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

void read_exact(std::FILE *f, void *p, size_t bytes, const char *what) {
  FLGPU_REQUIRE(std::fread(p, 1, bytes, f) == bytes,
                std::string("system file truncated while reading ") + what);
}

void write_exact(std::FILE *f, const void *p, size_t bytes) {
  FLGPU_REQUIRE(std::fwrite(p, 1, bytes, f) == bytes, "system file write failed");
}

}  // namespace

}  // namespace flgpu

using namespace flgpu;

extern "C" int flgpu_system_read(flgpu_ctx *ctx, const char *path, int prec, flgpu_mat **A,
                                 flgpu_vec **rhs) {
  return guard([&] {
    FLGPU_REQUIRE(ctx && path && A, "ctx, path and A must not be NULL");
    File file(std::fopen(path, "rb"));
    FLGPU_REQUIRE(file.f, std::string("cannot open system file ") + path);
    char magic[8];
    read_exact(file.f, magic, 8, "the magic");
    FLGPU_REQUIRE(std::memcmp(magic, kMagic, 8) == 0, "not a system file (bad magic)");
    int32_t orient = 0, flags = 0;
    int64_t nrows = 0, ncols = 0, nnz = 0;
    read_exact(file.f, &orient, 4, "the orientation");
    read_exact(file.f, &flags, 4, "the flags");
    read_exact(file.f, &nrows, 8, "nrows");
    read_exact(file.f, &ncols, 8, "ncols");
    read_exact(file.f, &nnz, 8, "nnz");
    FLGPU_REQUIRE(orient == 0 || orient == 1, "orientation must be 0 (rows) or 1 (columns)");
    FLGPU_REQUIRE((flags & ~3) == 0, "unknown flags");
    FLGPU_REQUIRE(nrows >= 0 && ncols >= 0 && nnz >= 0 && nrows < (1L << 31) &&
                      ncols < (1L << 31) && nnz < (1L << 31),
                  "system dimensions out of the int32 range");
    bool has_rhs = flags & 1, has_perm = flags & 2;
    FLGPU_REQUIRE(!has_perm || nrows == ncols, "a key permutation needs a square system");
    long nmajor = orient ? ncols : nrows, nminor = orient ? nrows : ncols;
    std::vector<int32_t> starts(nmajor + 1), idx(nnz);
    std::vector<double> val(nnz), b;
    std::vector<int32_t> perm;
    read_exact(file.f, starts.data(), 4 * starts.size(), "the starts");
    read_exact(file.f, idx.data(), 4 * idx.size(), "the indices");
    read_exact(file.f, val.data(), 8 * val.size(), "the values");
    if (has_rhs) {
      b.resize(nrows);
      read_exact(file.f, b.data(), 8 * b.size(), "the right-hand side");
    }
    if (has_perm) {
      perm.resize(nrows);
      read_exact(file.f, perm.data(), 4 * perm.size(), "the permutation");
      std::vector<char> seen(nrows, 0);
      for (int32_t p : perm) {
        FLGPU_REQUIRE(p >= 0 && p < nrows && !seen[p], "perm is not a permutation");
        seen[p] = 1;
      }
    }
    FLGPU_REQUIRE(starts[0] == 0 && starts[nmajor] == nnz, "starts must run from 0 to nnz");
    for (long k = 0; k < nmajor; k++)
      FLGPU_REQUIRE(starts[k + 1] >= starts[k], "starts must be non-decreasing");
    for (int32_t i : idx) FLGPU_REQUIRE(i >= 0 && i < nminor, "index out of range");
    // entries as (canonical row, canonical column, value), then CSR with ascending columns
    auto canon = [&](int32_t k) { return has_perm ? perm[k] : k; };
    std::vector<int32_t> row(nnz), col(nnz);
    for (long k = 0; k < nmajor; k++)
      for (int32_t e = starts[k]; e < starts[k + 1]; e++) {
        int32_t r = orient ? idx[e] : (int32_t)k, c = orient ? (int32_t)k : idx[e];
        row[e] = canon(r);
        col[e] = canon(c);
      }
    std::vector<int64_t> order(nnz);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b2) {
      return row[a] != row[b2] ? row[a] < row[b2] : col[a] < col[b2];
    });
    std::vector<int32_t> rowptr(nrows + 1, 0), colidx(nnz);
    std::vector<double> vals(nnz);
    for (int64_t k = 0; k < nnz; k++) {
      int64_t e = order[k];
      FLGPU_REQUIRE(k == 0 || row[e] != row[order[k - 1]] || col[e] != col[order[k - 1]],
                    "duplicate entry in the system file");
      rowptr[row[e] + 1]++;
      colidx[k] = col[e];
      vals[k] = val[e];
    }
    for (long r = 0; r < nrows; r++) rowptr[r + 1] += rowptr[r];
    flgpu_mat *M = nullptr;
    int st = flgpu_csr_upload(ctx, (int)nrows, (int)ncols, nnz, rowptr.data(), colidx.data(),
                              vals.data(), prec, &M);
    if (st != 0) {
      char msg[512];
      flgpu_last_error(msg, sizeof msg);
      throw Error{st, std::string("system upload failed: ") + msg};
    }
    std::unique_ptr<flgpu_mat, int (*)(flgpu_mat *)> guard_m(M, flgpu_mat_destroy);
    if (rhs) {
      *rhs = nullptr;
      if (has_rhs) {
        std::vector<double> bc(nrows);
        for (long k = 0; k < nrows; k++) bc[canon((int32_t)k)] = b[k];
        flgpu_vec *v = nullptr;
        FLGPU_REQUIRE(flgpu_vec_create(ctx, nrows, prec, &v) == 0, "rhs allocation failed");
        if (flgpu_vec_upload(v, bc.data()) != 0) {
          flgpu_vec_destroy(v);
          FLGPU_REQUIRE(false, "rhs upload failed");
        }
        *rhs = v;
      }
    }
    *A = guard_m.release();
    return 0;
  });
}

extern "C" int flgpu_system_write(const char *path, const flgpu_mat *A, const flgpu_vec *rhs) {
  return guard([&] {
    FLGPU_REQUIRE(path && A, "path and A must not be NULL");
    FLGPU_REQUIRE(!rhs || rhs->n == A->nrows, "rhs length differs from the matrix rows");
    std::vector<int32_t> rowptr(A->nrows + 1), colidx(A->nnz);
    std::vector<double> vals(A->nnz), b;
    FLGPU_REQUIRE(flgpu_csr_download(A, rowptr.data(), colidx.data(), vals.data()) == 0,
                  "matrix download failed");
    if (rhs) {
      b.resize(rhs->n);
      FLGPU_REQUIRE(flgpu_vec_download(rhs, b.data()) == 0, "rhs download failed");
    }
    File file(std::fopen(path, "wb"));
    FLGPU_REQUIRE(file.f, std::string("cannot create system file ") + path);
    int32_t orient = 0, flags = rhs ? 1 : 0;
    int64_t nrows = A->nrows, ncols = A->ncols, nnz = A->nnz;
    write_exact(file.f, kMagic, 8);
    write_exact(file.f, &orient, 4);
    write_exact(file.f, &flags, 4);
    write_exact(file.f, &nrows, 8);
    write_exact(file.f, &ncols, 8);
    write_exact(file.f, &nnz, 8);
    write_exact(file.f, rowptr.data(), 4 * rowptr.size());
    write_exact(file.f, colidx.data(), 4 * colidx.size());
    write_exact(file.f, vals.data(), 8 * vals.size());
    if (rhs) write_exact(file.f, b.data(), 8 * b.size());
    FLGPU_REQUIRE(std::fflush(file.f) == 0, "system file write failed");
    return 0;
  });
}
