// Vertex ordering of the input matrix and the vector permutations at the
// solve boundary.
//
// The reference permutes the input randomly, and only when it runs on more
// than one rank, to balance the ranks (Options::rand_perm, options.hpp:119;
// main.cpp:52-63).  On one GPU the order serves the gathers of x in every
// SpMV instead: vertices are renumbered by descending stored-entry count
// inside each connected block (ties by ascending index).  On a power-law
// graph the few thousand hubs that most rows point at then sit at the front
// of every vector, inside one XCD's L2, instead of scattered over the whole
// index range.  Coarse levels inherit the order (coarse ids follow their
// roots), so the locality carries down the hierarchy.
#include "prims.hpp"
#include "setup.hpp"

namespace lg {

namespace {

// key = block << 32 | (2^32 - 1 - row length): an ascending stable sort
// gives blocks in order, long rows first, ties by ascending index
__global__ void k_order_keys(int64_t n, const int32_t *rp, const int64_t *bptr, int nb, uint64_t *keys,
                             uint64_t *vals) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    int lo = 0, hi = nb;  // bptr[lo] <= i < bptr[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (bptr[mid] <= i) lo = mid;
      else hi = mid;
    }
    const uint32_t len = (uint32_t)(rp[i + 1] - rp[i]);
    keys[i] = ((uint64_t)lo << 32) | (uint64_t)(0xFFFFFFFFu - len);
    vals[i] = (uint64_t)i;
  }
}

__global__ void k_order_perm(int64_t n, const uint64_t *vals, const int32_t *rp, int32_t *perm,
                             int32_t *iperm, int32_t *len) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t o = (int32_t)vals[i];
    perm[i] = o;
    iperm[o] = (int32_t)i;
    len[i] = rp[o + 1] - rp[o];
  }
}

// Entry-parallel passes over a destination CSR (row pointers nrp) filled
// from a source one: body(r, nrp[r], q) handles destination position q of
// row r.  Rows of at most HUGE_ROW entries take one wavefront each; longer
// ones one wavefront per HUGE_ROW-entry piece (a second launch), so a hub
// row does not hold up the whole pass on one wavefront.
template <class Body>
__global__ void k_dst_rows(int64_t n, const int32_t *nrp, Body body) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t w0 = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / WAVE;
  for (int64_t r = w0; r < n; r += nw) {
    const int32_t qa = nrp[r], qb = nrp[r + 1];
    if (qb - qa > HUGE_ROW) continue;
    for (int32_t q = qa + lane; q < qb; q += WAVE) body(r, qa, q);
  }
}
template <class Body>
__global__ void k_dst_pieces(int64_t npieces, const int32_t *pstart, const int32_t *pend, const int32_t *prow,
                             const int32_t *nrp, Body body) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t c = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / WAVE;
  if (c >= npieces) return;
  const int32_t r = prow[c], d = nrp[r];
  for (int32_t q = pstart[c] + lane; q < pend[c]; q += WAVE) body(r, d, q);
}

// (new row << cb | new column, old position)
struct OrderBody {
  const int32_t *rp, *ci, *perm, *iperm;
  int cb;
  uint64_t *keys, *vals;
  __device__ void operator()(int64_t r, int32_t d, int32_t q) const {
    const int32_t src = rp[perm[r]] + (q - d);
    keys[q] = ((uint64_t)r << cb) | (uint64_t)iperm[ci[src]];
    vals[q] = (uint64_t)src;
  }
};

// Entries of B^T in row order (row x = A row perm[x]): key = new column in
// the low bits, x in the high 32; value = position of the transposed entry.
// A stable sort on the low bits then lists every column y (= row y of B,
// the pattern being symmetric) with its rows x ascending: B's rows, sorted.
struct OrderTBody {
  const int32_t *rp, *ci, *perm, *iperm, *tpos;
  uint64_t *keys, *vals;
  __device__ void operator()(int64_t r, int32_t d, int32_t q) const {
    const int32_t src = rp[perm[r]] + (q - d);
    keys[q] = ((uint64_t)r << 32) | (uint64_t)(uint32_t)iperm[ci[src]];
    vals[q] = (uint64_t)(uint32_t)tpos[src];
  }
};

// destination row r = source row rperm[r] (null: r), columns relabeled by
// ciperm (null: unchanged)
template <class V>
struct PermBody {
  const int32_t *rp, *ci;
  const V *v;
  const int32_t *rperm, *ciperm;
  int32_t *nci;
  V *nv;
  __device__ void operator()(int64_t r, int32_t d, int32_t q) const {
    const int32_t src = rp[rperm ? rperm[r] : (int32_t)r] + (q - d);
    const int32_t c = ci[src];
    nci[q] = ciperm ? ciperm[c] : c;
    if (v) nv[q] = v[src];
  }
};

__global__ void k_order_unpack_t(int64_t m, const uint64_t *keys, const uint64_t *vals, const double *v,
                                 int32_t *nci, double *nv) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    nci[k] = (int32_t)(keys[k] >> 32);
    nv[k] = v[vals[k]];
  }
}

__global__ void k_order_unpack(int64_t m, const uint64_t *keys, const uint64_t *vals, uint64_t mask,
                               const double *v, int32_t *nci, double *nv) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < m;
       k += (int64_t)gridDim.x * blockDim.x) {
    nci[k] = (int32_t)(keys[k] & mask);
    nv[k] = v[vals[k]];
  }
}

// B row i = A row rperm[i]; entries keep their order, columns relabeled by
// ciperm (null: unchanged).  One wavefront per row.
__global__ void k_perm_len(int64_t n, const int32_t *rp, const int32_t *rperm, int32_t *len) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t o = rperm ? rperm[i] : (int32_t)i;
    len[i] = rp[o + 1] - rp[o];
  }
}

// out[i] = niperm[map[perm[i]]] (null perm / niperm: identity)
__global__ void k_relabel(int64_t n, const int32_t *perm, const int32_t *map, const int32_t *niperm, int32_t *out) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int32_t m = map[perm ? perm[i] : i];
    out[i] = (niperm && m >= 0) ? niperm[m] : m;
  }
}

__global__ void k_gather(int64_t n, const int32_t *perm, const double *src, double *dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[perm[i]];
}

__global__ void k_scatter(int64_t n, const int32_t *perm, const double *src, double *dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dst[perm[i]] = src[i];
}

}  // namespace

void degree_perm(const DCsr &A, const std::vector<int64_t> &blocks, DevBuf<int32_t> &perm, DevBuf<int32_t> &iperm,
                 hipStream_t s) {
  const int64_t n = A.n;
  LG_CHECK(blocks.size() >= 2 && blocks.front() == 0 && blocks.back() == n, LIGMG_ERR_INVALID,
           "degree_perm: blocks must span [0, n]");
  const int nb = (int)blocks.size() - 1;
  perm.alloc(n);
  iperm.alloc(n);
  DevBuf<int32_t> len(n + 1);
  DevBuf<int64_t> bptr(blocks.size());
  to_device(bptr.p, blocks.data(), blocks.size(), s);
  DevBuf<uint64_t> keys(n), vals(n);
  hipLaunchKernelGGL(k_order_keys, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, A.rp.p, bptr.p, nb, keys.p,
                     vals.p);
  LG_LAUNCH_CHECK();
  radix_sort_pairs(keys, vals, n, 32 + bits_for((uint64_t)nb), s);
  hipLaunchKernelGGL(k_order_perm, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, vals.p, A.rp.p, perm.p,
                     iperm.p, len.p);
  LG_LAUNCH_CHECK();
  LG_HIP(hipStreamSynchronize(s));
}

template <class Body>
void for_dst_entries(int64_t n, const int32_t *nrp, const Body &body, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dst_rows<Body>, dim3(grid_for(n * WAVE, 256, 16384)), dim3(256), 0, s, n, nrp, body);
  LG_LAUNCH_CHECK();
  HugeRows Hr;
  find_huge(n, nrp, Hr, s);
  if (Hr.nchunks) {
    hipLaunchKernelGGL(k_dst_pieces<Body>, dim3((unsigned)((Hr.nchunks + 3) / 4)), dim3(256), 0, s, Hr.nchunks,
                       Hr.cstart.p, Hr.cend.p, Hr.crow.p, nrp, body);
    LG_LAUNCH_CHECK();
  }
}

void permute_rows(const DCsr &A, const int32_t *rperm, const int32_t *ciperm, int64_t ncols, DCsr &B,
                  hipStream_t s) {
  const int64_t n = A.n, m = A.nnz;
  DevBuf<int32_t> len(n + 1);
  hipLaunchKernelGGL(k_perm_len, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, A.rp.p, rperm, len.p);
  LG_LAUNCH_CHECK();
  fill_i32(len.p + n, 0, 1, s);
  B.n = n;
  B.ncols = ncols;
  B.nnz = m;
  B.rp.alloc(n + 1);
  const int64_t total = exclusive_scan_i32(len.p, B.rp.p, n + 1, s);
  LG_CHECK(total == m, LIGMG_ERR_INVALID, "permute_rows: row pointers inconsistent");
  B.ci.alloc(std::max<int64_t>(m, 1));
  if (A.v.p) B.v.alloc(std::max<int64_t>(m, 1));
  if (m > 0)
    for_dst_entries(n, B.rp.p, PermBody<double>{A.rp.p, A.ci.p, A.v.p, rperm, ciperm, B.ci.p, B.v.p}, s);
  LG_HIP(hipStreamSynchronize(s));
}

void relabel_map(int64_t n, const int32_t *perm, const int32_t *map, const int32_t *niperm, int32_t *out,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_relabel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, perm, map, niperm, out);
  LG_LAUNCH_CHECK();
}

void permute_sorted(const DCsr &A, const int32_t *perm, const int32_t *iperm, const int32_t *tpos, DCsr &B,
                    hipStream_t s) {
  const int64_t n = A.n, m = A.nnz;
  DevBuf<int32_t> len(n + 1);
  hipLaunchKernelGGL(k_perm_len, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, A.rp.p, perm, len.p);
  LG_LAUNCH_CHECK();
  B.n = n;
  B.ncols = A.ncols;
  B.rp.alloc(n + 1);
  fill_i32(len.p + n, 0, 1, s);
  const int64_t total = exclusive_scan_i32(len.p, B.rp.p, n + 1, s);
  LG_CHECK(total == m, LIGMG_ERR_INVALID, "permute_sorted: row pointers inconsistent");
  B.nnz = m;
  B.ci.alloc(std::max<int64_t>(m, 1));
  B.v.alloc(std::max<int64_t>(m, 1));
  if (m > 0 && tpos) {  // one sort by column of B^T's entries (see OrderTBody)
    const int cb = std::max(1, bits_for((uint64_t)n));
    DevBuf<uint64_t> keys(m), vals(m);
    for_dst_entries(n, B.rp.p, OrderTBody{A.rp.p, A.ci.p, perm, iperm, tpos, keys.p, vals.p}, s);
    radix_sort_pairs(keys, vals, m, cb, s);
    hipLaunchKernelGGL(k_order_unpack_t, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, m, keys.p, vals.p,
                       A.v.p, B.ci.p, B.v.p);
    LG_LAUNCH_CHECK();
  } else if (m > 0) {  // a sort by (row, column)
    const int cb = bits_for((uint64_t)n);
    DevBuf<uint64_t> keys(m), vals(m);
    for_dst_entries(n, B.rp.p, OrderBody{A.rp.p, A.ci.p, perm, iperm, cb, keys.p, vals.p}, s);
    radix_sort_pairs(keys, vals, m, 2 * cb, s);
    hipLaunchKernelGGL(k_order_unpack, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, m, keys.p, vals.p,
                       ((uint64_t)1 << cb) - 1, A.v.p, B.ci.p, B.v.p);
    LG_LAUNCH_CHECK();
  }
  LG_HIP(hipStreamSynchronize(s));
}

void degree_order(DCsr &A, const std::vector<int64_t> &blocks, DevBuf<int32_t> &perm, hipStream_t s) {
  DevBuf<int32_t> iperm;
  degree_perm(A, blocks, perm, iperm, s);
  DCsr B;
  permute_sorted(A, perm.p, iperm.p, nullptr, B, s);
  A = std::move(B);
}

void gather_vec(int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, perm, src, dst);
  LG_LAUNCH_CHECK();
}

void scatter_vec(int64_t n, const int32_t *perm, const double *src, double *dst, hipStream_t s) {
  hipLaunchKernelGGL(k_scatter, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, n, perm, src, dst);
  LG_LAUNCH_CHECK();
}

}  // namespace lg
