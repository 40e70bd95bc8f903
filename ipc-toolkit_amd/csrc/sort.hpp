// Hand-written device-wide primitives for gfx950: exclusive scan, stable
// LSD radix sort of (uint64 or uint32 key, uint32 value) pairs with 8-bit digits,
// run-length encoding of sorted keys.  Declarations; kernels in sort.hip.
#pragma once

#include "common.hpp"

namespace ipcg {

// out[i] = sum(in[0..i)), returns nothing; total written to *total_dev if
// non-null (device pointer).  tmp must hold scan_tmp_bytes(n).
size_t scan_tmp_bytes(int64_t n);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total_dev,
                        void* tmp, hipStream_t s);
void exclusive_scan_u64(const uint64_t* in, uint64_t* out, int64_t n, uint64_t* total_dev,
                        void* tmp, hipStream_t s);

// Stable sort of n (key, value) pairs by the low `key_bits` bits of key.
// keys/vals are sorted in place (ping-pong through keys_alt/vals_alt).
size_t radix_tmp_bytes(int64_t n);
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt,
                      int64_t n, int key_bits, void* tmp, hipStream_t s);
// Every 8-bit pass below key_bits runs, with no host round trip and no copy
// back: returns true when the sorted pairs are in keys_alt / vals_alt.  For
// keys whose digits all vary (Morton codes, vertex ids).
bool radix_sort_pairs_all(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                          uint32_t* vals_alt, int64_t n, int key_bits, void* tmp, hipStream_t s);
bool radix_sort_pairs_all(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                          uint32_t* vals_alt, int64_t n, int key_bits, void* tmp, hipStream_t s);
// The same over the 8-bit digits that hold a bit of bit_mask (the others
// are known to be constant), again without a host round trip.
bool radix_sort_pairs_masked(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                             uint32_t* vals_alt, int64_t n, unsigned long long bit_mask,
                             void* tmp, hipStream_t s);

// Sorted keys -> unique keys, run starts (n_unique + 1 entries, last = n)
// and the number of unique keys (device scalar).  tmp: rle_tmp_bytes(n).
size_t rle_tmp_bytes(int64_t n);
void run_length_encode(const uint64_t* keys, int64_t n, uint64_t* unique_keys,
                       uint32_t* run_start, uint32_t* n_unique_dev, void* tmp, hipStream_t s);

} // namespace ipcg
