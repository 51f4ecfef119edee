// Deterministic device primitives: scan, stable radix sort, canonical
// segmented sums.  All run on one stream; functions that return a host value
// synchronise that stream, the others return with their work queued.
#pragma once
#include "common.hpp"

namespace lg {

// Exclusive prefix sum of in[0..n) into out[0..n) (out may alias in).
// Returns the total.
int64_t exclusive_scan_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t s);
int64_t exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, hipStream_t s);

// Stable sort of the first n (key, val) pairs by the low `bits` bits of the
// keys (keys of 2^bits or more are not allowed).  The sorted pairs are left
// in keys / vals: each may be exchanged for an internal buffer of n items.
void radix_sort_pairs(DevBuf<uint64_t> &keys, DevBuf<uint64_t> &vals, int64_t n, int bits, hipStream_t s);

// Canonical sum (the order of the CPU oracle's canon_sum) of the values of
// each run of equal keys in a sorted key array.  Writes the run keys and
// sums; returns the number of runs.  vals are doubles stored as raw bits.
int64_t segmented_canon_sum(const uint64_t *keys, const uint64_t *vals, int64_t n,
                            uint64_t *out_keys, double *out_sums, hipStream_t s);

// Sequential sum (a left fold from 0.0 in array order) of the values of
// each run of equal keys in a sorted key array: the order of a loop that
// adds the items one by one.  Same outputs and return value as
// segmented_canon_sum.
int64_t segmented_seq_sum(const uint64_t *keys, const uint64_t *vals, int64_t n,
                          uint64_t *out_keys, double *out_sums, hipStream_t s);

// out[k] = index of the k-th nonzero flag (ascending); returns the count.
int64_t compact_flags(const int32_t *flags, int64_t n, int32_t *out, hipStream_t s);

void fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);
void fill_f64(double *p, double v, int64_t n, hipStream_t s);
void fill_u8(uint8_t *p, uint8_t v, int64_t n, hipStream_t s);

}  // namespace lg
