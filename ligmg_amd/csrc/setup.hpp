// Device setup entry points (setup.hip).
#pragma once
#include "hierarchy.hpp"

namespace lg {

// Contract (key = row << cb | col, value) pairs into CSR (destroys the inputs).
void contract_to_csr(DevBuf<uint64_t> &keys, DevBuf<uint64_t> &vals, int64_t m, int64_t nrows,
                     int64_t ncols, DCsr &out, hipStream_t s);

// Row classes of one matrix for the row kernels: rows longer than HUGE_ROW
// entries and their HUGE_ROW-entry chunks (chunk c covers [cstart[c],
// cend[c]) of row crow[c]; a row's chunks are cptr[j] .. cptr[j+1]-1), the
// rows of WAVE + 1 .. HUGE_ROW entries (a wavefront each, canonical sums)
// and the rows longer than SEQ_LANE (a wavefront each, sequential sums).
// find_huge builds it from row pointers (the counts come back to the host);
// huge_of caches it per matrix.
constexpr int HUGE_ROW = 4096;
struct HugeRows {
  int64_t nrows = 0, nchunks = 0;
  DevBuf<int32_t> row, cptr, cstart, cend, crow;
  DevBuf<uint8_t> cacc;
  int64_t nlong = 0;  // rows of more than SEQ_LANE entries (sequential sums)
  DevBuf<int32_t> lrow;
  int64_t nmid = 0;  // rows of WAVE + 1 .. HUGE_ROW entries (canonical sums)
  DevBuf<int32_t> mrow;
};
void find_huge(int64_t n, const int32_t *rp, HugeRows &Hr, hipStream_t s);
const HugeRows &huge_of(const DCsr &A, hipStream_t s);
// the FP32 Laplacian form of A's values (v32: rounded off-diagonals, 0 on
// the diagonal; d32: -(sum of the rounded off-diagonals) in FP64)
void lap32_form(DCsr &A, hipStream_t s);

void row_sums_seq(int64_t n, const int32_t *rp, const double *v, double *out, hipStream_t s);
// Throws LIGMG_ERR_INVALID unless rp[0] = 0, rp is non-decreasing, every
// column is in [0, ncols) and, with sorted, every row's columns strictly
// increase.  Reads columns only after the row pointers check out.
void validate_csr(int64_t n, int64_t ncols, const int32_t *rp, const int32_t *ci, hipStream_t s, bool sorted);

// R = P^T (FP32 values too when P has them).
void transpose_p(const DCsr &P, int64_t ncols, DCsr &R, hipStream_t s);
// Aggregate member lists (nc x n, unit values, members ascending).
void build_members(const int32_t *agg, int64_t n, int64_t nc, DCsr &M, hipStream_t s);

// Elimination level from A; false if rejected (kept > elim_factor n).
bool setup_elimination(const DCsr &A, const ligmg_options &o, DLevel &L, DCsr &coarse, hipStream_t s);

// Strength (per stored entry, 0 where pruned) and column sums.
// tpos receives the positions of the transposed entries of A.
void compute_strength(const DCsr &A, const ligmg_options &o, int level_idx, DevBuf<double> &S,
                      DevBuf<double> &cs, DevBuf<int32_t> &tpos, hipStream_t s);
// Column sums of a strength matrix on A's pattern.
void col_sums_of(const DCsr &A, const double *S, DevBuf<double> &cs, hipStream_t s);
// Voting aggregation with given strengths; returns the number of aggregates.
int64_t run_voting(const DCsr &A, const double *S, const double *cs, const ligmg_options &o,
                   DevBuf<int32_t> &agg, DevBuf<int32_t> &ind, DevBuf<int32_t> &state, hipStream_t s);

// Aggregation level from A (strength, voting, Galerkin, member lists).
void setup_aggregation(const DCsr &A, const ligmg_options &o, int level_idx, DLevel &L, DCsr &coarse,
                       hipStream_t s);

// Renumber A's vertices by descending row length inside each block (ties
// by ascending index); perm receives internal position -> input vertex.
void degree_order(DCsr &A, const std::vector<int64_t> &blocks, DevBuf<int32_t> &perm, hipStream_t s);
// Degree permutation of A's rows inside blocks (perm: new -> old, iperm:
// old -> new), the one degree_order applies.
void degree_perm(const DCsr &A, const std::vector<int64_t> &blocks, DevBuf<int32_t> &perm, DevBuf<int32_t> &iperm,
                 hipStream_t s);
// B row i = A row rperm[i] (null: i) with its entries in their order and
// columns relabeled by ciperm (null: unchanged); FP64 values only.
void permute_rows(const DCsr &A, const int32_t *rperm, const int32_t *ciperm, int64_t ncols, DCsr &B,
                  hipStream_t s);
// B = P A P^T with P given by perm (new -> old) / iperm (old -> new) and
// every row's columns sorted; FP64 values only.  tpos (optional): the
// positions of A's transposed entries (structurally symmetric A), which
// turn the sort by (row, column) into one by column.
void permute_sorted(const DCsr &A, const int32_t *perm, const int32_t *iperm, const int32_t *tpos, DCsr &B,
                    hipStream_t s);
// out[i] = niperm[map[perm[i]]] (null: identity; negative map values kept)
void relabel_map(int64_t n, const int32_t *perm, const int32_t *map, const int32_t *niperm, int32_t *out,
                 hipStream_t s);
// dst[i] = src[perm[i]]  /  dst[perm[i]] = src[i]
void gather_vec(int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s);
void scatter_vec(int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s);

void setup_dinv(const DCsr &A, DLevel &L, bool fp32, hipStream_t s);
// FP32 P, R = P^T and Q of an elimination level from the FP32 Laplacian form
// of its matrix (L.A.v32 / L.A.d32 must exist).
void make_elim32(DLevel &L, hipStream_t s);
// perm (may be null): setup position of each row, for the start vector
double eig_estimate(const DCsr &A, const double *dinv, uint64_t seed, int level_idx, int iters,
                    const int32_t *perm, hipStream_t s);
// Dense pseudo-inverse of the coarsest level (n <= COARSE_DENSE_MAX): the
// inverse of L + sum over components C of 1_C 1_C^T / |C|, minus those terms.
constexpr int64_t COARSE_DENSE_MAX = 8192;
// coarse_pinv factors on the host below COARSE_DEVICE_MIN vertices (serial
// Cholesky, dense_pinv_host) and on the device from there (blocked
// Gauss-Jordan on the matrix cores); the result is a device n x n array.
constexpr int64_t COARSE_DEVICE_MIN = 256;
// block width of the device Gauss-Jordan (the matrix is padded to a multiple)
constexpr int COARSE_GJ_BLOCK = 128;
// gj_ms (optional): milliseconds of the device inversion kernels (0 on the
// host path).
void coarse_pinv(int64_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                 const std::vector<double> &v, DevBuf<double> &pinv, hipStream_t s, double *gj_ms = nullptr);
std::vector<double> dense_pinv_host(int64_t n, const std::vector<int32_t> &rp,
                                    const std::vector<int32_t> &ci, const std::vector<double> &v);
// The SPD matrix above (row-major n x n), its components and their sizes.
std::vector<double> coarse_dense_spd(int64_t n, const std::vector<int32_t> &rp, const std::vector<int32_t> &ci,
                                     const std::vector<double> &v, std::vector<int32_t> &comp,
                                     std::vector<int64_t> &csize);

}  // namespace lg
