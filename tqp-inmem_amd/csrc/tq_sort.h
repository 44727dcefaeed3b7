// Internal interface of the radix sort, shared by the join and group-by drivers.
#pragma once

#include <string.h>

#include "tq_common.h"

namespace tq {

// Workspace of the digit histograms/offsets for sorting n pairs.
size_t radix_ws_bytes(int64_t n);

// Stable sort of (keys, vals) by key bits [lo_bit, hi_bit) in 8-bit digits, ping-ponging with
// the *_alt buffers.  *in_alt tells which pair of buffers holds the result.
template <typename K>
int radix_sort_pairs(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                     int lo_bit, int hi_bit, void* ws, hipStream_t s, bool* in_alt);

// Bit range [lo, hi) in which the n keys differ (lo rounded down to a digit boundary) and,
// optionally, the smallest and largest key; reads the answer back to the host, so it
// synchronises the stream.  stats_dev holds 4 words.
template <typename K>
int key_bit_range(const K* keys, int64_t n, void* stats_dev, hipStream_t s, int* lo, int* hi,
                  uint64_t* kmin = nullptr, uint64_t* kmax = nullptr);

// Digit-aligned bit range [lo, hi) in which keys with the given OR and AND differ.
void radix_bit_range(uint64_t key_or, uint64_t key_and, int* lo, int* hi);

// Copy `bytes` (<= 512) of device memory to the host through pinned staging; synchronises.
int read_back(const void* dev, size_t bytes, int64_t* host_out, hipStream_t s);
// n copies with one synchronisation, packed into host_out at 8-byte-aligned offsets (at most
// 512 bytes in all).
int read_back_n(int n, const void* const* dev, const size_t* bytes, int64_t* host_out,
                hipStream_t s);
// Two such copies (each <= 256 bytes) with one synchronisation.
int read_back2(const void* dev0, size_t bytes0, int64_t* host0, const void* dev1, size_t bytes1,
               int64_t* host1, hipStream_t s);

size_t sort_ws_bytes(int64_t n);

// Multi-key stable sort; *perm_out points into the workspace and holds n uint32 row ids.
int sort_perm_u32(const tq_sortkey_t* keys, int32_t nkeys, int64_t n, void* ws, size_t ws_bytes,
                  hipStream_t s, uint32_t** perm_out);

}  // namespace tq
