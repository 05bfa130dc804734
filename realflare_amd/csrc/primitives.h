// Device-wide scan and stable radix sort (see primitives.hip).
#pragma once

#include "rf_common.h"

namespace rf {

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// out[i] = sum(in[0..i)) for i < n, n = *n_dev (or `capacity` when n_dev is
// NULL); *total_dev (if not NULL) = sum(in[0..n)).  Grids are sized for
// `capacity`.
size_t scan_workspace_bytes(uint64_t capacity);
int scan_exclusive_u32(const uint32_t* in, uint64_t* out, const uint64_t* n_dev, uint64_t capacity,
                       uint64_t* total_dev, void* ws, hipStream_t s);

// Stable ascending sort of keys (low `key_bits` bits significant) carrying
// vals, n = *n_dev <= capacity pairs, in place.
size_t radix_workspace_bytes(uint64_t capacity);
int radix_sort_pairs(uint32_t* keys, uint32_t* vals, const uint64_t* n_dev, uint64_t capacity,
                     int key_bits, void* ws, hipStream_t s);

}  // namespace rf
