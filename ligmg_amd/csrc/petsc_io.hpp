// PETSc binary files, the format the reference's dumps are written in
// (write_petsc, util.hpp:589-599 for matrices through MatView on a binary
// viewer; combblas_extra.hpp:26-43 for vectors): big-endian int32 class id,
// sizes, row lengths and column indices, big-endian IEEE doubles.
//
//   matrix: 1211216, rows, cols, nnz, int32[rows] row lengths,
//           int32[nnz] columns, double[nnz] values
//   vector: 1211214, n, double[n]
#pragma once
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "common.hpp"

namespace lg {
namespace pio {

constexpr int32_t MAT_CLASSID = 1211216;
constexpr int32_t VEC_CLASSID = 1211214;
constexpr int ERR_IO = 6;  // LIGMG_ERR_IO

inline uint32_t be32(uint32_t x) { return __builtin_bswap32(x); }
inline uint64_t be64(uint64_t x) { return __builtin_bswap64(x); }

struct File {
  FILE *f = nullptr;
  std::string path;
  File(const std::string &p, const char *mode) : path(p) {
    f = fopen(p.c_str(), mode);
    if (!f) throw Error(ERR_IO, "cannot open " + p);
  }
  ~File() {
    if (f) fclose(f);
  }
  File(const File &) = delete;
  File &operator=(const File &) = delete;
  void put_i32(const int32_t *a, size_t n) {
    std::vector<uint32_t> b(n);
    for (size_t i = 0; i < n; i++) b[i] = be32((uint32_t)a[i]);
    if (n && fwrite(b.data(), 4, n, f) != n) throw Error(ERR_IO, "write failed: " + path);
  }
  void put_f64(const double *a, size_t n) {
    std::vector<uint64_t> b(n);
    for (size_t i = 0; i < n; i++) {
      uint64_t u;
      memcpy(&u, a + i, 8);
      b[i] = be64(u);
    }
    if (n && fwrite(b.data(), 8, n, f) != n) throw Error(ERR_IO, "write failed: " + path);
  }
  void get_i32(int32_t *a, size_t n) {
    std::vector<uint32_t> b(n);
    if (n && fread(b.data(), 4, n, f) != n) throw Error(ERR_IO, "truncated file: " + path);
    for (size_t i = 0; i < n; i++) a[i] = (int32_t)be32(b[i]);
  }
  void get_f64(double *a, size_t n) {
    std::vector<uint64_t> b(n);
    if (n && fread(b.data(), 8, n, f) != n) throw Error(ERR_IO, "truncated file: " + path);
    for (size_t i = 0; i < n; i++) {
      const uint64_t u = be64(b[i]);
      memcpy(a + i, &u, 8);
    }
  }
};

inline bool exists(const std::string &p) {
  FILE *f = fopen(p.c_str(), "rb");
  if (!f) return false;
  fclose(f);
  return true;
}

inline void write_vec(const std::string &p, const double *x, int64_t n) {
  LG_CHECK(n < INT32_MAX, ERR_IO, "vector too long for a PETSc binary file");
  File F(p, "wb");
  const int32_t hdr[2] = {VEC_CLASSID, (int32_t)n};
  F.put_i32(hdr, 2);
  F.put_f64(x, (size_t)n);
}

// v == nullptr: unit values
inline void write_mat(const std::string &p, int64_t m, int64_t ncols, const int32_t *rp, const int32_t *ci,
                      const double *v) {
  const int64_t nnz = rp[m] - rp[0];
  LG_CHECK(m < INT32_MAX && ncols < INT32_MAX && nnz < INT32_MAX, ERR_IO,
           "matrix too large for a 32-bit PETSc binary file");
  File F(p, "wb");
  const int32_t hdr[4] = {MAT_CLASSID, (int32_t)m, (int32_t)ncols, (int32_t)nnz};
  F.put_i32(hdr, 4);
  std::vector<int32_t> len(m);
  for (int64_t i = 0; i < m; i++) len[i] = rp[i + 1] - rp[i];
  F.put_i32(len.data(), (size_t)m);
  F.put_i32(ci + rp[0], (size_t)nnz);
  if (v) {
    F.put_f64(v + rp[0], (size_t)nnz);
  } else {
    std::vector<double> ones((size_t)nnz, 1.0);
    F.put_f64(ones.data(), (size_t)nnz);
  }
}

inline std::vector<double> read_vec(const std::string &p) {
  File F(p, "rb");
  int32_t hdr[2];
  F.get_i32(hdr, 2);
  LG_CHECK(hdr[0] == VEC_CLASSID && hdr[1] >= 0, ERR_IO, "not a PETSc binary vector: " + p);
  std::vector<double> x((size_t)hdr[1]);
  F.get_f64(x.data(), x.size());
  return x;
}

struct HostCsr {
  int64_t m = 0, ncols = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
};

inline HostCsr read_mat(const std::string &p) {
  File F(p, "rb");
  int32_t hdr[4];
  F.get_i32(hdr, 4);
  LG_CHECK(hdr[0] == MAT_CLASSID && hdr[1] >= 0 && hdr[2] >= 0 && hdr[3] >= 0, ERR_IO,
           "not a PETSc binary matrix: " + p);
  HostCsr A;
  A.m = hdr[1];
  A.ncols = hdr[2];
  std::vector<int32_t> len((size_t)A.m);
  F.get_i32(len.data(), len.size());
  A.rp.assign((size_t)A.m + 1, 0);
  for (int64_t i = 0; i < A.m; i++) {
    LG_CHECK(len[i] >= 0, ERR_IO, "negative row length in " + p);
    A.rp[i + 1] = A.rp[i] + len[i];
  }
  LG_CHECK(A.rp[A.m] == hdr[3], ERR_IO, "row lengths do not add up to nnz in " + p);
  A.ci.resize((size_t)hdr[3]);
  A.v.resize((size_t)hdr[3]);
  F.get_i32(A.ci.data(), A.ci.size());
  F.get_f64(A.v.data(), A.v.size());
  for (int32_t c : A.ci) LG_CHECK(c >= 0 && c < A.ncols, ERR_IO, "column out of range in " + p);
  return A;
}

}  // namespace pio
}  // namespace lg
