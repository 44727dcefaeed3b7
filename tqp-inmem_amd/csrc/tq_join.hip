// Equi-join index and probe.
//
// Build.  One pass over the build table counts the selected rows of every 4096-row tile and
// reduces the order-transformed keys' OR, AND, min and max into per-tile partials, which one
// small kernel combines (no contended atomics).  The host reads the row count and those four
// numbers (one synchronisation) and picks the index layout:
//   * dense: a directory indexed by key - min, when the keys span at most 4*nb + 65536 values;
//   * hash: open addressing with linear probing over a power-of-two slot array of at least
//     twice the selected rows, multiplicative (Fibonacci) hashing.
// The build first assumes unique keys: every selected row claims its key's entry (a plain store
// into the dense directory, whose occupied entries are then counted; a compare-and-swap on the
// hash slot's key, which flags a key claimed twice).  The check is read back together with the
// first probe's pair count, so the build itself synchronises once.  Without duplicates an entry holds its build row and the build is done,
// with no sort.  With duplicates the selected (key, row) pairs are compacted in row order and
// radix-sorted by key (stable, so each key's rows stay ascending); an entry then holds the
// first sorted position of its key and the number of rows with that key.
//
// Probe.  Two passes over 4096-row probe tiles.  The count pass looks up every selected probe
// row and writes the number of pairs per tile; an exclusive scan gives each tile's first output
// position; the write pass repeats the lookups and writes the pairs at their positions, probe
// row ascending then build row ascending, gathering the requested build and probe columns
// straight into the outputs.  A thread owns four runs of four consecutive rows, so mask and key
// loads are vector loads, and the directory loads of all its rows are issued before any is
// used.
#include "tq_common.h"
#include "tq_sort.h"

namespace tq {

constexpr uint32_t kNoPos = 0xFFFFFFFFu;
constexpr int kRuns = kItems / 4;  // runs of 4 consecutive rows per thread and tile

struct Slot {
  uint64_t key;
  uint32_t a;    // build row (unique) or first sorted position
  uint32_t cnt;  // rows with this key
};

struct JoinView {
  const void* table;
  const uint32_t* bits;  // dense layouts: presence bit per directory entry
  const uint32_t* rows;
  uint64_t base;
  uint64_t empty;
  int64_t range;
  int32_t hash_bits;
};

__host__ __device__ inline uint64_t slot_of(uint64_t t, int bits) {
  return (t * 0x9E3779B97F4A7C15ull) >> (64 - bits);
}

template <typename K>
__host__ __device__ __forceinline__ K order_key(int64_t v) {
  return sizeof(K) == 4 ? (K)((uint32_t)(int32_t)v ^ 0x80000000u)
                        : (K)((uint64_t)v ^ 0x8000000000000000ull);
}

// A probe value can only match 32-bit build keys when it lies in the int32 range.
template <typename K>
__device__ __forceinline__ bool probe_fits(int64_t v) {
  return sizeof(K) == 8 || (v >= INT32_MIN && v <= INT32_MAX);
}

// Number of build rows with key t (0 when !valid) and, through *a, the entry's row/position.
// Dense layouts test the presence bitmap first (range/8 bytes, cache resident) and read the
// directory only for present keys, and only when the caller needs the row (NEED_ROW) or the
// count of a repeated key.
template <typename K, bool HASH, bool UNIQUE, bool NEED_ROW>
__device__ __forceinline__ uint32_t lookup(const JoinView& J, K t, bool valid, uint32_t* a) {
  *a = 0;
  if (!HASH) {
    const uint64_t off = (uint64_t)t - J.base;  // below the base wraps to a huge value
    valid = valid && off < (uint64_t)J.range;
    const uint32_t word = J.bits[valid ? (off >> 5) : 0];
    const bool present = valid && ((word >> (off & 31)) & 1u);
    if (UNIQUE) {
      if (NEED_ROW && present) *a = ((const uint32_t*)J.table)[off];
      return present ? 1u : 0u;
    }
    if (!present) return 0;
    const uint2 e = ((const uint2*)J.table)[off];
    *a = e.x;
    return e.y;
  }
  if (!valid) return 0;
  const Slot* S = (const Slot*)J.table;
  const uint64_t mask = (uint64_t)J.range - 1;
  for (uint64_t h = slot_of((uint64_t)t, J.hash_bits);; h = (h + 1) & mask) {
    const Slot s = S[h];
    if (s.key == (uint64_t)t) {
      *a = s.a;
      return UNIQUE ? 1u : s.cnt;
    }
    if (s.key == J.empty) return 0;
  }
}

// Selection bits and values of a thread's 16 rows of a tile: run q covers rows
// tile*4096 + q*1024 + 4*threadIdx.x + (0..3).
__device__ __forceinline__ void load_tile_rows(const void* keys, int dt, const uint8_t* mask,
                                               int64_t n, int64_t tile, uint32_t* sel,
                                               int64_t v[kItems]) {
  uint64_t raw[kItems];
  load_raw16(keys, dt, tile * kTileRows, n, raw);
  *sel = load_sel16(mask, tile * kTileRows, n);
#pragma unroll
  for (int j = 0; j < kItems; ++j) v[j] = (int64_t)raw[j];
}

__device__ __forceinline__ int64_t tile_row(int64_t tile, int j) {
  return tile * kTileRows + (int64_t)(j >> 2) * (4 * kBlock) + 4 * threadIdx.x + (j & 3);
}

// ------------------------------------------------------------------ build: statistics
// Per tile: selected rows -> tile_counts[tile]; OR, AND, min, max of the selected transformed
// keys -> part[4*tile ..] (min > max when the tile selects nothing).
template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_stats(const void* __restrict__ bkeys, int dt,
                                                       const uint8_t* __restrict__ mask,
                                                       int64_t n, uint32_t* __restrict__ tile_counts,
                                                       uint64_t* __restrict__ part) {
  __shared__ uint64_t sh[5][kWaves];
  const int64_t tile = blockIdx.x;
  uint32_t sel;
  int64_t v[kItems];
  load_tile_rows(bkeys, dt, mask, n, tile, &sel, v);
  uint64_t so = 0, sa = ~0ull, slo = ~0ull, shi = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if ((sel >> j) & 1u) {
      const uint64_t k = (uint64_t)order_key<K>(v[j]);
      so |= k;
      sa &= k;
      slo = k < slo ? k : slo;
      shi = k > shi ? k : shi;
    }
  }
  uint64_t c = (uint64_t)__popc(sel);
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    c += __shfl_xor(c, d, 64);
    so |= __shfl_xor(so, d, 64);
    sa &= __shfl_xor(sa, d, 64);
    const uint64_t a2 = __shfl_xor(slo, d, 64), b2 = __shfl_xor(shi, d, 64);
    slo = a2 < slo ? a2 : slo;
    shi = b2 > shi ? b2 : shi;
  }
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sh[0][w] = c; sh[1][w] = so; sh[2][w] = sa; sh[3][w] = slo; sh[4][w] = shi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 1; i < kWaves; ++i) {
      c += sh[0][i]; so |= sh[1][i]; sa &= sh[2][i];
      slo = sh[3][i] < slo ? sh[3][i] : slo;
      shi = sh[4][i] > shi ? sh[4][i] : shi;
    }
    tile_counts[tile] = (uint32_t)c;
    part[4 * tile + 0] = so;
    part[4 * tile + 1] = sa;
    part[4 * tile + 2] = slo;
    part[4 * tile + 3] = shi;
  }
}

// One workgroup: stats = {selected rows, OR, AND, min, max}; clears the duplicate flag and the
// occupied-entry counters.
constexpr int kOccSlots = 64;   // occupied-entry counters, spread so their atomics rarely meet
constexpr int kOccStride = 32;  // u32 words between counters: one 128-byte line each

__global__ __launch_bounds__(kBlock) void k_join_stats_final(const uint32_t* __restrict__ counts,
                                                             const uint64_t* __restrict__ part,
                                                             int64_t ntiles,
                                                             uint64_t* __restrict__ stats,
                                                             uint32_t* __restrict__ dup,
                                                             uint32_t* __restrict__ occ) {
  __shared__ uint64_t sh[5][kWaves];
  if (threadIdx.x < kOccSlots) occ[threadIdx.x * kOccStride] = 0;
  uint64_t c = 0, so = 0, sa = ~0ull, slo = ~0ull, shi = 0;
  for (int64_t i = threadIdx.x; i < ntiles; i += kBlock) {
    c += counts[i];
    if (counts[i]) {
      so |= part[4 * i];
      sa &= part[4 * i + 1];
      slo = part[4 * i + 2] < slo ? part[4 * i + 2] : slo;
      shi = part[4 * i + 3] > shi ? part[4 * i + 3] : shi;
    }
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    c += __shfl_xor(c, d, 64);
    so |= __shfl_xor(so, d, 64);
    sa &= __shfl_xor(sa, d, 64);
    const uint64_t a2 = __shfl_xor(slo, d, 64), b2 = __shfl_xor(shi, d, 64);
    slo = a2 < slo ? a2 : slo;
    shi = b2 > shi ? b2 : shi;
  }
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sh[0][w] = c; sh[1][w] = so; sh[2][w] = sa; sh[3][w] = slo; sh[4][w] = shi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 1; i < kWaves; ++i) {
      c += sh[0][i]; so |= sh[1][i]; sa &= sh[2][i];
      slo = sh[3][i] < slo ? sh[3][i] : slo;
      shi = sh[4][i] > shi ? sh[4][i] : shi;
    }
    stats[0] = c; stats[1] = so; stats[2] = sa; stats[3] = slo; stats[4] = shi;
    *dup = 0;
  }
}

// ------------------------------------------------------------------ build: unique claim
// Also zeroes the uniqueness-check words (flags[0..5]) and the occupied-entry counters when
// given: a build with host-known key bounds has no statistics pass to do it.  A dense
// directory is filled four entries per thread (16-byte stores).
template <bool HASH>
__global__ __launch_bounds__(kBlock) void k_join_clear(void* __restrict__ table, int64_t n,
                                                       uint64_t empty, uint32_t* __restrict__ flags,
                                                       uint32_t* __restrict__ occ) {
  if (flags && blockIdx.x == 0) {
    if (threadIdx.x < 6) flags[threadIdx.x] = 0;
    if (threadIdx.x < kOccSlots) occ[threadIdx.x * kOccStride] = 0;
  }
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  if (HASH) {
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
      Slot s;
      s.key = empty; s.a = 0; s.cnt = 0;
      ((Slot*)table)[i] = s;
    }
    return;
  }
  const int64_t n4 = n >> 2;  // the directory is 256-byte aligned
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n4; i += stride)
    ((uint4*)table)[i] = make_uint4(kNoPos, kNoPos, kNoPos, kNoPos);
  for (int64_t i = 4 * n4 + (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
    ((uint32_t*)table)[i] = kNoPos;
}

// Every selected build row claims its key's entry; in a hash table a key claimed twice sets
// *dup.  Dense directories take plain stores (one of several rows with the same key wins);
// k_join_bits then counts the occupied entries, which equals the selected row count exactly
// when the keys are unique.
template <typename K, bool HASH>
__global__ __launch_bounds__(kBlock) void k_join_claim(const void* __restrict__ bkeys, int dt,
                                                       const uint8_t* __restrict__ mask,
                                                       int64_t n, JoinView J,
                                                       uint32_t* __restrict__ dup,
                                                       uint32_t* __restrict__ tile_counts,
                                                       uint64_t kmin, uint64_t kspan,
                                                       uint32_t* __restrict__ oob) {
  const int64_t tile = blockIdx.x;
  uint32_t sel;
  int64_t v[kItems];
  load_tile_rows(bkeys, dt, mask, n, tile, &sel, v);
  if (tile_counts) {  // selected rows of the tile
    __shared__ uint32_t wc[kWaves];
    const uint32_t c = wave_sum((uint32_t)__popc(sel));
    if (lane_id() == 0) wc[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t b = 0;
#pragma unroll
      for (int i = 0; i < kWaves; ++i) b += wc[i];
      tile_counts[tile] = b;
    }
  }
  bool seen = false, out = false;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if (!((sel >> j) & 1u)) continue;
    const uint64_t t = (uint64_t)order_key<K>(v[j]);
    const uint32_t row = (uint32_t)tile_row(tile, j);
    if (oob && t - kmin > kspan) {  // outside the bounds the caller promised: not claimed
      out = true;
      continue;
    }
    if (!HASH) {
      ((uint32_t*)J.table)[t - J.base] = row;
    } else {
      Slot* S = (Slot*)J.table;
      const uint64_t m = (uint64_t)J.range - 1;
      for (uint64_t h = slot_of(t, J.hash_bits);; h = (h + 1) & m) {
        const unsigned long long old =
            atomicCAS((unsigned long long*)&S[h].key, (unsigned long long)J.empty,
                      (unsigned long long)t);
        if (old == J.empty) {
          S[h].a = row;
          S[h].cnt = 1;
          break;
        }
        if (old == t) {
          seen = true;
          break;
        }
      }
    }
  }
  if (seen) *dup = 1;
  if (out) *oob = 1;
}

// Presence bitmap of a dense directory: bit i of bits[w] is set when entry 32*w + i holds a key.
// A thread tests four consecutive entries (one 16-byte load for unique directories), so a wave
// covers 256 entries = 8 bitmap words, each assembled from 8 lanes.  With occ, the number of
// occupied entries is added to occ[block % kOccSlots]: after the plain-store claim, the keys
// were unique exactly when that number equals the selected row count.
template <bool UNIQUE>
__global__ __launch_bounds__(kBlock) void k_join_bits(const void* __restrict__ table,
                                                      int64_t range, uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ occ) {
  const int64_t stride = (int64_t)gridDim.x * kBlock * 4;
  const int l = lane_id();
  uint32_t c = 0;
  for (int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 4; (e & ~255ll) < range;
       e += stride) {
    uint32_t nib = 0;
    if (UNIQUE && e + 3 < range) {
      const uint4 q = *(const uint4*)((const uint32_t*)table + e);
      nib = (q.x != kNoPos ? 1u : 0u) | (q.y != kNoPos ? 2u : 0u) | (q.z != kNoPos ? 4u : 0u) |
            (q.w != kNoPos ? 8u : 0u);
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t x = e + i;
        const bool p = x < range && (UNIQUE ? ((const uint32_t*)table)[x] != kNoPos
                                            : ((const uint2*)table)[x].y != 0u);
        nib |= (p ? 1u : 0u) << i;
      }
    }
    c += (uint32_t)__popc(nib);
    // word (e / 32) takes the nibbles of lanes 8*(l/8) .. 8*(l/8)+7
    uint32_t word = nib << (4 * (l & 7));
    word |= __shfl_xor(word, 1, 64);
    word |= __shfl_xor(word, 2, 64);
    word |= __shfl_xor(word, 4, 64);
    if ((l & 7) == 0 && e < range) bits[e >> 5] = word;
  }
  if (!occ) return;
  __shared__ uint32_t wc[kWaves];
  c = wave_sum(c);
  if (l == 0) wc[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) b += wc[i];
    // one add per workgroup, on a counter in its own 128-byte line
    if (b) atomicAdd(&occ[(blockIdx.x % kOccSlots) * kOccStride], b);
  }
}

// The uniqueness check's numbers next to the duplicate flag (one workgroup): the occupied-entry
// sum (dense directories) and, after a ranged build, the selected rows (sum of tile counts).
// The build's uniqueness check, summed where the first probe reads it: the spread occupied-entry
// counters (dense layouts) into out[0] and, for a ranged build, the per-tile selected counts into
// out[1].  Run by one workgroup: its own kernel, or the first workgroup of the count pass.
struct BuildCheck {
  const uint32_t* occ;          // null: no occupied-entry counters
  const uint32_t* tile_counts;  // null: the build counted its rows already
  int64_t ntiles;
  int64_t* out;                 // null: nothing to check
};

__device__ __forceinline__ void build_check(const BuildCheck& B) {
  __shared__ int64_t sh[kWaves];
  int64_t v = (B.occ && threadIdx.x < kOccSlots) ? B.occ[threadIdx.x * kOccStride] : 0;
  v = wave_sum(v);  // the counters are in the first wave
  if (threadIdx.x == 0) B.out[0] = v;
  if (!B.tile_counts) return;
  int64_t c = 0;
  for (int64_t i = threadIdx.x; i < B.ntiles; i += kBlock) c += B.tile_counts[i];
  c = wave_sum(c);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t t = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) t += sh[i];
    B.out[1] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_join_check(BuildCheck B) { build_check(B); }

// ------------------------------------------------------------------ build: duplicate keys
// Compact the selected (key, row) pairs in row order.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_pack(const void* __restrict__ bkeys, int dt,
                                                      const uint8_t* __restrict__ mask, int64_t n,
                                                      const int64_t* __restrict__ tile_off,
                                                      K* __restrict__ keys,
                                                      uint32_t* __restrict__ rows) {
  __shared__ uint32_t sh[kWaves];
  const int64_t tile = blockIdx.x;
  uint32_t sel;
  int64_t v[kItems];
  load_tile_rows(bkeys, dt, mask, n, tile, &sel, v);
  int64_t base = tile_off[tile];
#pragma unroll
  for (int q = 0; q < kRuns; ++q) {
    const uint32_t s4 = (sel >> (4 * q)) & 15u;
    uint32_t tot;
    const uint32_t ex = block_exclusive_sum((uint32_t)__popc(s4), sh, &tot);
    int64_t o = base + ex;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if ((s4 >> i) & 1u) {
        keys[o] = order_key<K>(v[4 * q + i]);
        rows[o] = (uint32_t)tile_row(tile, 4 * q + i);
        ++o;
      }
    }
    base += tot;
  }
}

// Entries of the sorted keys: at the first position i of each key, count its rows (galloping
// then binary search for the end of the run) and record {i, count}.
template <typename K, bool HASH>
__global__ __launch_bounds__(kBlock) void k_join_runs(const K* __restrict__ keys, int64_t nsel,
                                                      JoinView J) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nsel; i += stride) {
    const K t = keys[i];
    if (i > 0 && keys[i - 1] == t) continue;
    int64_t lo = i, step = 1;  // keys[lo] == t
    while (lo + step < nsel && keys[lo + step] == t) {
      lo += step;
      step <<= 1;
    }
    int64_t hi = lo + step < nsel ? lo + step : nsel;  // keys[hi] != t or hi == nsel
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (keys[mid] == t) lo = mid; else hi = mid;
    }
    const uint32_t cnt = (uint32_t)(hi - i);
    if (!HASH) {
      ((uint2*)J.table)[(uint64_t)t - J.base] = make_uint2((uint32_t)i, cnt);
    } else {
      Slot* S = (Slot*)J.table;
      const uint64_t m = (uint64_t)J.range - 1;
      for (uint64_t h = slot_of((uint64_t)t, J.hash_bits);; h = (h + 1) & m) {
        const unsigned long long old =
            atomicCAS((unsigned long long*)&S[h].key, (unsigned long long)J.empty,
                      (unsigned long long)t);
        if (old == J.empty) {
          S[h].a = (uint32_t)i;
          S[h].cnt = cnt;
          break;
        }
      }
    }
  }
}

// ------------------------------------------------------------------ probe
template <typename K, bool HASH, bool UNIQUE, bool NEED_ROW>
__device__ __forceinline__ void probe_rows(const JoinView& J, const void* pkeys, int pdt,
                                           const uint8_t* pmask, int64_t np, int64_t tile,
                                           uint32_t c[kItems], uint32_t a[kItems]) {
  uint32_t sel;
  int64_t v[kItems];
  load_tile_rows(pkeys, pdt, pmask, np, tile, &sel, v);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const bool ok = ((sel >> j) & 1u) && probe_fits<K>(v[j]);
    c[j] = lookup<K, HASH, UNIQUE, NEED_ROW>(J, order_key<K>(v[j]), ok, &a[j]);
  }
}

// Pairs per probe tile.  With unique build keys a probe row has at most one pair, and the
// thread's 16 match bits go to hits[tile * 256 + t] for the write pass.
template <typename K, bool HASH, bool UNIQUE>
__global__ __launch_bounds__(kBlock) void k_join_count(JoinView J, const void* __restrict__ pkeys,
                                                       int pdt, const uint8_t* __restrict__ pmask,
                                                       int64_t np, int64_t ntiles,
                                                       uint32_t* __restrict__ tile_counts,
                                                       uint16_t* __restrict__ hits,
                                                       BuildCheck check) {
  __shared__ uint32_t sh[kWaves];
  if (check.out && blockIdx.x == 0) {  // a pending build's check rides on the first workgroup
    build_check(check);
    __syncthreads();
  }
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t c[kItems], a[kItems];
    probe_rows<K, HASH, UNIQUE, false>(J, pkeys, pdt, pmask, np, tile, c, a);
    uint32_t cnt = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) cnt += c[j];
    if (UNIQUE) {
      uint32_t h = 0;
#pragma unroll
      for (int j = 0; j < kItems; ++j) h |= c[j] << j;
      hits[tile * kBlock + threadIdx.x] = (uint16_t)h;
    }
    cnt = wave_sum(cnt);
    if (lane_id() == 0) sh[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += sh[w];
      tile_counts[tile] = s;
    }
    __syncthreads();
  }
}

// Output column x of the pairs listed in LDS: element i of the tile's output (at obase + i)
// takes src row from_build[i] (build side) or tile0 + from_probe[i] (probe side).  Consecutive
// threads write consecutive outputs.  Thread t owns elements t, t + kBlock, ... (R rounds, R *
// kBlock >= total) and issues all R loads before its first store, so R loads per thread are in
// flight instead of one.  The code is straight-line: lanes past the list's end load the last
// listed pair again and store that same value to the last output, a duplicate of the same
// bytes, so no branch separates the loads and the stores.
template <typename T, bool PROBE, int R>
__device__ __forceinline__ void gather_rounds(T* __restrict__ d, const T* __restrict__ s,
                                              uint32_t total, const uint32_t* from_build,
                                              const uint16_t* from_probe, int64_t tile0) {
  T v[R];
  uint32_t at[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t i = r * kBlock + threadIdx.x;
    at[r] = i < total ? i : total - 1;
    const int64_t from = PROBE ? tile0 + from_probe[at[r]] : (int64_t)from_build[at[r]];
    v[r] = s[from];
  }
#pragma unroll
  for (int r = 0; r < R; ++r) d[at[r]] = v[r];
}

template <typename T, bool PROBE>
__device__ __forceinline__ void gather_list(void* dst, const void* src, int64_t obase,
                                            uint32_t total, const uint32_t* from_build,
                                            const uint16_t* from_probe, int64_t tile0) {
  static_assert(kItems * kBlock == kTileRows, "a tile lists at most kItems rounds of pairs");
  if (total == 0) return;
  T* d = (T*)dst + obase;
  const T* s = (const T*)src;
  if (total <= 4 * kBlock)
    gather_rounds<T, PROBE, 4>(d, s, total, from_build, from_probe, tile0);
  else
    gather_rounds<T, PROBE, kItems>(d, s, total, from_build, from_probe, tile0);
}

template <bool PROBE>
__device__ __forceinline__ void gather_col(void* dst, const void* src, int bytes, int64_t obase,
                                           uint32_t total, const uint32_t* fb, const uint16_t* fp,
                                           int64_t tile0) {
  switch (bytes) {
    case 1: gather_list<uint8_t, PROBE>(dst, src, obase, total, fb, fp, tile0); break;
    case 2: gather_list<uint16_t, PROBE>(dst, src, obase, total, fb, fp, tile0); break;
    case 4: gather_list<uint32_t, PROBE>(dst, src, obase, total, fb, fp, tile0); break;
    case 8: gather_list<uint64_t, PROBE>(dst, src, obase, total, fb, fp, tile0); break;
    default:
      for (uint32_t i = threadIdx.x; i < total; i += kBlock)
        copy_row(dst, obase + i, src, PROBE ? tile0 + fp[i] : (int64_t)fb[i], bytes);
  }
}

// Write pass for unique build keys: the count pass's match bits say which probe rows have their
// (single) pair, so the probe keys and mask are not read again; the build row is looked up
// only when build columns or build row ids are written (NEED_ROW).  The tile's pairs are listed
// in output order in LDS, then each output column is written with consecutive threads on
// consecutive outputs.
template <typename K, bool HASH, bool NEED_ROW>
__global__ __launch_bounds__(kBlock) void k_join_write_unique(
    JoinView J, const void* __restrict__ pkeys, int pdt, int64_t ntiles,
    const uint16_t* __restrict__ hits, const int64_t* __restrict__ tile_off, ColPack L,
    ColPack Rc, int64_t* __restrict__ build_idx, int64_t* __restrict__ probe_idx) {
  __shared__ uint32_t sh[kWaves];
  __shared__ uint32_t fb[kTileRows];  // build row of the tile's i-th pair
  __shared__ uint16_t fp[kTileRows];  // probe row - tile0 of the i-th pair
  const int t = threadIdx.x;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t h16 = hits[tile * kBlock + t];
    uint32_t a[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      a[j] = 0;
      if (NEED_ROW && ((h16 >> j) & 1u))
        lookup<K, HASH, true, true>(J, order_key<K>(load_int(pkeys, pdt, tile_row(tile, j))), true,
                                    &a[j]);
    }
    const int64_t base = tile_off[tile];
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < kRuns; ++q) {
      const uint32_t h4 = (h16 >> (4 * q)) & 15u;
      uint32_t tot;
      const uint32_t ex = block_exclusive_sum((uint32_t)__popc(h4), sh, &tot);
      uint32_t o = run + ex;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if ((h4 >> i) & 1u) {
          fb[o] = a[4 * q + i];
          fp[o] = (uint16_t)(q * (4 * kBlock) + 4 * t + i);
          ++o;
        }
      }
      run += tot;
    }
    __syncthreads();
    const int64_t tile0 = tile * kTileRows;
    for (int x = 0; x < L.n; ++x)
      gather_col<false>(L.dst[x], L.src[x], L.bytes[x], base, run, fb, fp, tile0);
    for (int x = 0; x < Rc.n; ++x)
      gather_col<true>(Rc.dst[x], Rc.src[x], Rc.bytes[x], base, run, fb, fp, tile0);
    for (uint32_t i = t; i < run; i += kBlock) {
      if (build_idx) build_idx[base + i] = fb[i];
      if (probe_idx) probe_idx[base + i] = tile0 + fp[i];
    }
    __syncthreads();
  }
}

// Write pass for repeated build keys.  The lookups are repeated, then the tile's pairs are
// spread over the whole workgroup in output order: each probe row's first output (within the
// tile) and bucket start go to LDS, and thread t writes the tile's outputs t, t + 256, ...,
// finding the probe row of each by a search over the first outputs (forward from the row of
// its previous output).  A probe row whose key repeats many times in the build is expanded by
// all 256 threads with consecutive stores, not by one lane.  Output order: probe rows
// ascending, each row's build rows in bucket order.
// With bypos, L's sources are the build columns already in bucket order (k_join_bucket_cols),
// indexed by sorted position: a probe row's pairs read one short run of each.
template <typename K, bool HASH, bool UNIQUE, bool NEED_ROW>
__global__ __launch_bounds__(kBlock) void k_join_write(
    JoinView J, const void* __restrict__ pkeys, int pdt, const uint8_t* __restrict__ pmask,
    int64_t np, int64_t ntiles, const int64_t* __restrict__ tile_off, ColPack L, ColPack Rc,
    int64_t* __restrict__ build_idx, int64_t* __restrict__ probe_idx, int bypos) {
  __shared__ uint32_t sh[kWaves];
  __shared__ uint32_t first[kTileRows + 1];  // row's first output in the tile; [4096] = total
  __shared__ uint32_t bucket[kTileRows];     // row's first position in J.rows
  const int t = threadIdx.x;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t c[kItems], a[kItems];
    probe_rows<K, HASH, UNIQUE, true>(J, pkeys, pdt, pmask, np, tile, c, a);
    uint32_t base = 0;
#pragma unroll
    for (int q = 0; q < kRuns; ++q) {
      const uint32_t tc = c[4 * q] + c[4 * q + 1] + c[4 * q + 2] + c[4 * q + 3];
      uint32_t tot;
      uint32_t o = base + block_exclusive_sum(tc, sh, &tot);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int lr = q * (4 * kBlock) + 4 * t + i;  // tile_row(tile, 4q + i) - tile * 4096
        first[lr] = o;
        bucket[lr] = a[4 * q + i];
        o += c[4 * q + i];
      }
      base += tot;
    }
    if (t == 0) first[kTileRows] = base;
    __syncthreads();
    const int64_t obase = tile_off[tile];
    const int64_t row0 = tile * kTileRows;
    int r = 0;  // first[r] <= o for every output o this thread still writes
    for (uint32_t o = t; o < base; o += kBlock) {
      // the row whose outputs [first[r], first[r + 1]) hold o: gallop from r, then bisect
      int lo = r, step = 1;
      while (lo + step < kTileRows && first[lo + step] <= o) {
        lo += step;
        step <<= 1;
      }
      int hi = lo + step < kTileRows ? lo + step : kTileRows;  // first[hi] > o
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (first[mid] <= o) lo = mid; else hi = mid;
      }
      r = lo;
      const uint32_t pos = bucket[r] + (o - first[r]);
      const uint32_t b = (!bypos || build_idx) ? J.rows[pos] : 0u;
      const int64_t li = bypos ? (int64_t)pos : (int64_t)b;
      const int64_t pr = row0 + r;
      const int64_t out = obase + o;
      for (int x = 0; x < L.n; ++x) copy_row(L.dst[x], out, L.src[x], li, L.bytes[x]);
      for (int x = 0; x < Rc.n; ++x) copy_row(Rc.dst[x], out, Rc.src[x], pr, Rc.bytes[x]);
      if (build_idx) build_idx[out] = b;
      if (probe_idx) probe_idx[out] = pr;
    }
    __syncthreads();  // the next tile rewrites first[] and bucket[]
  }
}

// Build-side columns of a repeated-key join in bucket order: B.dst[x][p] = L.src[x][rows[p]]
// for the nsel sorted positions.  One pass over the selected build rows per probe, after which
// the write pass reads a probe row's build values as one short run per column instead of one
// random line per pair.
__global__ __launch_bounds__(kBlock) void k_join_bucket_cols(const uint32_t* __restrict__ rows,
                                                             int64_t nsel, ColPack L, ColPack B) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x; p < nsel; p += stride) {
    const uint32_t r = rows[p];
    for (int x = 0; x < L.n; ++x) copy_row(B.dst[x], p, L.src[x], r, L.bytes[x]);
  }
}

// ------------------------------------------------------------------ workspace layouts
static int64_t dense_cap(int64_t nb) { return 4 * nb + 65536; }
static int hash_bits_for(int64_t n) {
  int b = 10;
  while (((int64_t)1 << b) < 2 * n) ++b;
  return b;
}

struct BuildWs {
  uint32_t* tile_counts;
  int64_t* tile_off;
  uint64_t* stats;   // selected rows, OR, AND, min, max
  uint32_t* dup;
  uint32_t* occ;     // kOccSlots counters of occupied dense entries
  int64_t* total;
  int64_t* scan_ws;
  uint64_t* part;
  void* k0;
  void* k1;
  uint32_t* v0;
  uint32_t* v1;
  void* radix;
  void* table;
  uint32_t* bits;
  size_t bytes;
};

static BuildWs carve_build(void* base, int64_t nb) {
  if (nb < 1) nb = 1;
  const int64_t nt = ceil_div(nb, kTileRows);
  const size_t dense_bytes = (size_t)dense_cap(nb) * 8;
  const size_t hash_bytes = ((size_t)1 << hash_bits_for(nb)) * sizeof(Slot);
  char* p = (char*)base;
  BuildWs w;
  w.tile_counts = (uint32_t*)p; p += align_up((size_t)nt * 4, 256);
  w.tile_off = (int64_t*)p; p += align_up((size_t)nt * 8, 256);
  w.stats = (uint64_t*)p;
  w.dup = (uint32_t*)(p + 64);
  w.total = (int64_t*)(p + 128); p += 256;
  w.occ = (uint32_t*)p; p += kOccSlots * kOccStride * 4;
  w.scan_ws = (int64_t*)p; p += tq_scan_ws_bytes(nt);
  w.part = (uint64_t*)p; p += align_up((size_t)nt * 32, 256);
  w.k0 = p; p += align_up((size_t)nb * 8, 256);
  w.k1 = p; p += align_up((size_t)nb * 8, 256);
  w.v0 = (uint32_t*)p; p += align_up((size_t)nb * 4, 256);
  w.v1 = (uint32_t*)p; p += align_up((size_t)nb * 4, 256);
  w.radix = p; p += radix_ws_bytes(nb);
  w.table = p; p += align_up(dense_bytes > hash_bytes ? dense_bytes : hash_bytes, 256);
  w.bits = (uint32_t*)p; p += align_up((size_t)ceil_div(dense_cap(nb), 32) * 4, 256);
  w.bytes = (size_t)(p - (char*)base);
  return w;
}

struct ProbeWs {
  uint32_t* tile_counts;
  int64_t* tile_off;
  int64_t* total;
  int64_t* scan_ws;
  uint16_t* hits;  // unique build keys: per (tile, thread) match bits of the thread's 16 rows
  size_t bytes;
};

static ProbeWs carve_probe(void* base, int64_t np) {
  if (np < 1) np = 1;
  const int64_t nt = ceil_div(np, kTileRows);
  char* p = (char*)base;
  ProbeWs w;
  w.tile_counts = (uint32_t*)p; p += align_up((size_t)nt * 4, 256);
  w.tile_off = (int64_t*)p; p += align_up((size_t)nt * 8, 256);
  w.total = (int64_t*)p; p += 256;
  w.scan_ws = (int64_t*)p; p += tq_scan_ws_bytes(nt);
  w.hits = (uint16_t*)p; p += align_up((size_t)nt * kBlock * 2, 256);
  w.bytes = (size_t)(p - (char*)base);
  return w;
}

static JoinView view_of(const tq_join_state_t* st) {
  JoinView J;
  J.table = st->table;
  J.bits = st->bits;
  J.rows = st->rows;
  J.base = st->base;
  J.empty = st->empty_key;
  J.range = st->range;
  J.hash_bits = st->hash_bits;
  return J;
}



// Sorted layout for repeated keys: compact the selected (key, row) pairs in row order (the
// statistics pass left the per-tile counts in the workspace), sort them stably by key and record
// {first position, count} per key.
template <typename K>
static int build_multi(tq_join_state_t* st, BuildWs& w, hipStream_t s) {
  const int64_t nb = st->nb, nsel = st->nsel;
  const int64_t nt = ceil_div(nb, kTileRows);
  const bool hash = st->layout == TQ_JOIN_HASH;
  st->unique = 0;
  TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
  TQ_TRY(launch("k_join_pack", k_join_pack<K>, (unsigned)nt, kBlock, 0, s, st->bkeys,
                (int)st->bdtype, st->bmask, nb, (const int64_t*)w.tile_off, (K*)w.k0, w.v0));
  int lo = 0, hi = 0;
  radix_bit_range(st->key_or, st->key_and, &lo, &hi);
  bool alt;
  TQ_TRY(radix_sort_pairs<K>((K*)w.k0, w.v0, (K*)w.k1, w.v1, nsel, lo, hi, w.radix, s, &alt));
  const K* keys = (const K*)(alt ? w.k1 : w.k0);
  st->rows = alt ? w.v1 : w.v0;
  JoinView J = view_of(st);
  const unsigned cg = stream_grid(st->range);
  if (hash) {
    TQ_TRY(launch("k_join_clear", k_join_clear<true>, cg, kBlock, 0, s, w.table, st->range,
                  st->empty_key, (uint32_t*)nullptr, (uint32_t*)nullptr));
    TQ_TRY(launch("k_join_runs", (k_join_runs<K, true>), stream_grid(nsel), kBlock, 0, s, keys,
                  nsel, J));
  } else {
    TQ_TRY(fill_u64(w.table, 0ull, st->range, s));  // {first, count} = {0, 0}: absent
    TQ_TRY(launch("k_join_runs", (k_join_runs<K, false>), stream_grid(nsel), kBlock, 0, s, keys,
                  nsel, J));
    st->bits = w.bits;
    TQ_TRY(launch("k_join_bits", k_join_bits<false>, stream_grid(ceil_div(st->range, 4)), kBlock, 0, s,
                  (const void*)w.table, st->range, w.bits, (uint32_t*)nullptr));
  }
  return TQ_OK;
}

// Layout for selected transformed keys within [kmin, kmax]; a hash table gets at least twice
// cap_rows slots.
template <typename K>
static int set_layout(tq_join_state_t* st, uint64_t kmin, uint64_t kmax, int64_t cap_rows,
                      int64_t nb) {
  const bool hash = kmax - kmin >= (uint64_t)dense_cap(nb);
  st->layout = hash ? TQ_JOIN_HASH : TQ_JOIN_DENSE;
  if (hash) {
    st->hash_bits = hash_bits_for(cap_rows);
    st->range = (int64_t)1 << st->hash_bits;
    st->base = 0;
    if (sizeof(K) == 4 || kmax != ~0ull) {
      st->empty_key = sizeof(K) == 4 ? ~0ull : kmax + 1;
    } else if (kmin != 0) {
      st->empty_key = kmin - 1;
    } else {
      return set_error(TQ_EUNSUPPORTED,
                       "join build keys include both INT64_MIN and INT64_MAX with a sparse "
                       "range: no free key value marks an empty hash slot");
    }
  } else {
    st->base = kmin;
    st->range = (int64_t)(kmax - kmin) + 1;
  }
  return TQ_OK;
}

// Statistics pass: selected rows, OR/AND and range of their keys (one synchronisation).
template <typename K>
static int build_stats(const void* bkeys, int32_t dt, const uint8_t* bmask, int64_t nb,
                       BuildWs& w, tq_join_state_t* st, hipStream_t s, uint64_t* kmin,
                       uint64_t* kmax) {
  const int64_t nt = ceil_div(nb, kTileRows);
  TQ_TRY(launch("k_join_stats", k_join_stats<K>, (unsigned)nt, kBlock, 0, s, bkeys, dt, bmask,
                nb, w.tile_counts, w.part));
  TQ_TRY(launch("k_join_stats_final", k_join_stats_final, 1, kBlock, 0, s,
                (const uint32_t*)w.tile_counts, (const uint64_t*)w.part, nt, w.stats, w.dup,
                w.occ));
  int64_t hv[5];
  TQ_TRY(read_back(w.stats, 40, hv, s));
  st->nsel = hv[0];
  st->key_or = (uint64_t)hv[1];
  st->key_and = (uint64_t)hv[2];
  *kmin = (uint64_t)hv[3];
  *kmax = (uint64_t)hv[4];
  return TQ_OK;
}

// Claim pass assuming unique keys; the check is read with the first pair count.  A ranged
// build (key bounds from the caller, no statistics pass) also clears the check words, counts
// the selected rows and flags keys outside the bounds.
template <typename K>
static int claim_unique(const void* bkeys, int32_t dt, const uint8_t* bmask, int64_t nb,
                        BuildWs& w, tq_join_state_t* st, bool ranged, uint64_t kmin,
                        uint64_t kmax, hipStream_t s) {
  const int64_t nt = ceil_div(nb, kTileRows);
  st->table = w.table;
  JoinView J = view_of(st);
  const unsigned cg = stream_grid(st->range);
  uint32_t* clear_flags = ranged ? w.dup : nullptr;
  uint32_t* tcounts = ranged ? w.tile_counts : nullptr;
  uint32_t* oob = ranged ? w.dup + 1 : nullptr;
  const unsigned cg4 = stream_grid(ceil_div(st->range, 4));
  const bool hash = st->layout == TQ_JOIN_HASH;
  if (hash) {
    TQ_TRY(launch("k_join_clear", k_join_clear<true>, cg, kBlock, 0, s, w.table, st->range,
                  st->empty_key, clear_flags, w.occ));
    TQ_TRY(launch("k_join_claim", (k_join_claim<K, true>), (unsigned)nt, kBlock, 0, s, bkeys, dt,
                  bmask, nb, J, w.dup, tcounts, kmin, kmax - kmin, oob));
  } else {
    TQ_TRY(launch("k_join_clear", k_join_clear<false>, cg4, kBlock, 0, s, w.table, st->range,
                  st->empty_key, clear_flags, w.occ));
    TQ_TRY(launch("k_join_claim", (k_join_claim<K, false>), (unsigned)nt, kBlock, 0, s, bkeys, dt,
                  bmask, nb, J, w.dup, tcounts, kmin, kmax - kmin, oob));
    st->bits = w.bits;
    TQ_TRY(launch("k_join_bits", k_join_bits<true>, cg4, kBlock, 0, s, (const void*)w.table,
                  st->range, w.bits, w.occ));
  }
  st->unique = 1;
  st->pending = 1;
  st->bdtype = dt;
  st->bkeys = bkeys;
  st->bmask = bmask;
  st->nb = nb;
  return TQ_OK;
}

template <typename K>
static int build_impl(const void* bkeys, int32_t dt, const uint8_t* bmask, int64_t nb, BuildWs& w,
                      tq_join_state_t* st, hipStream_t s) {
  uint64_t kmin, kmax;
  TQ_TRY(build_stats<K>(bkeys, dt, bmask, nb, w, st, s, &kmin, &kmax));
  if (st->nsel == 0) {
    st->layout = TQ_JOIN_EMPTY;
    return TQ_OK;
  }
  TQ_TRY(set_layout<K>(st, kmin, kmax, st->nsel, nb));
  return claim_unique<K>(bkeys, dt, bmask, nb, w, st, false, 0, 0, s);
}

// Build with caller-given bounds lo <= every build key <= hi: no statistics pass, so no
// synchronisation; the selected row count is settled by the first pair count (nsel = -1 until
// then).
template <typename K>
static int build_ranged(const void* bkeys, int32_t dt, const uint8_t* bmask, int64_t nb,
                        int64_t lo, int64_t hi, BuildWs& w, tq_join_state_t* st, hipStream_t s) {
  const uint64_t kmin = (uint64_t)order_key<K>(lo), kmax = (uint64_t)order_key<K>(hi);
  st->nsel = -1;
  st->key_or = 0;
  st->key_and = ~0ull;
  TQ_TRY(set_layout<K>(st, kmin, kmax, nb, nb));
  return claim_unique<K>(bkeys, dt, bmask, nb, w, st, true, kmin, kmax, s);
}

// The build's statistics again (after a ranged build's check failed), then the layout for
// repeated keys or, when the check failed on the bounds, the unique claim.
template <typename K>
static int rebuild_from_stats(tq_join_state_t* st, BuildWs& w, bool repeated, hipStream_t s) {
  uint64_t kmin, kmax;
  TQ_TRY(build_stats<K>(st->bkeys, st->bdtype, st->bmask, st->nb, w, st, s, &kmin, &kmax));
  if (st->nsel == 0) {
    st->layout = TQ_JOIN_EMPTY;
    return TQ_OK;
  }
  TQ_TRY(set_layout<K>(st, kmin, kmax, st->nsel, st->nb));
  if (repeated) return build_multi<K>(st, w, s);
  return claim_unique<K>(st->bkeys, st->bdtype, st->bmask, st->nb, w, st, false, 0, 0, s);
}

template <typename K, bool HASH, bool UNIQUE>
static int count_impl(const tq_join_state_t* st, const void* pkeys, int32_t pdt,
                      const uint8_t* pmask, int64_t np, int64_t nt, ProbeWs& w,
                      const BuildCheck& check, hipStream_t s) {
  return launch("k_join_count", (k_join_count<K, HASH, UNIQUE>),
                resident_grid(k_join_count<K, HASH, UNIQUE>, 0, nt), kBlock, 0, s,
                view_of(st), pkeys, pdt, pmask, np, nt, w.tile_counts,
                UNIQUE ? w.hits : (uint16_t*)nullptr, check);
}

// The check words of a pending build (dup u32, out-of-bounds u32, occupied entries i64,
// selected rows i64 for a ranged build), summed by the workgroup that runs the check.
static BuildCheck pending_check(const tq_join_state_t* st) {
  BuildCheck c{};
  if (!st->pending) return c;
  BuildWs bw = carve_build(st->ws, st->nb);
  const bool dense = st->layout == TQ_JOIN_DENSE;
  if (!dense && st->nsel >= 0) return c;  // hash, not ranged: the duplicate flag says it all
  c.occ = dense ? (const uint32_t*)bw.occ : nullptr;
  c.tile_counts = st->nsel < 0 ? (const uint32_t*)bw.tile_counts : nullptr;
  c.ntiles = ceil_div(st->nb, kTileRows);
  c.out = (int64_t*)(bw.dup + 2);
  return c;
}

template <typename K, bool HASH, bool UNIQUE>
static int write_impl(const tq_join_state_t* st, const void* pkeys, int32_t pdt,
                      const uint8_t* pmask, int64_t np, int64_t nt, const ProbeWs& w,
                      const ColPack& L, const ColPack& R, int64_t* bidx, int64_t* pidx,
                      hipStream_t s) {
  const JoinView J = view_of(st);
  const int64_t* off = (const int64_t*)w.tile_off;
  if (UNIQUE && (L.n > 0 || bidx))
    return launch("k_join_write", (k_join_write_unique<K, HASH, true>),
                  resident_grid(k_join_write_unique<K, HASH, true>, 0, nt), kBlock, 0, s, J,
                  pkeys, pdt, nt, (const uint16_t*)w.hits, off, L, R, bidx, pidx);
  if (UNIQUE)
    return launch("k_join_write", (k_join_write_unique<K, HASH, false>),
                  resident_grid(k_join_write_unique<K, HASH, false>, 0, nt), kBlock, 0, s, J,
                  pkeys, pdt, nt, (const uint16_t*)w.hits, off, L, R, bidx, pidx);
  // repeated keys: the build columns go to bucket order first when they fit the build
  // workspace's sort-key buffers (free once the index is built: only the sorted rows and the
  // directory are read after it)
  ColPack Lw = L;
  int bypos = 0;
  if (L.n > 0 && st->nsel > 0) {
    const BuildWs bw = carve_build(st->ws, st->nb);
    const size_t avail = (size_t)((char*)bw.v0 - (char*)bw.k0);
    size_t need = 0;
    for (int x = 0; x < L.n; ++x) need += align_up((size_t)st->nsel * L.bytes[x], 256);
    if (need <= avail) {
      ColPack B = L;
      char* p = (char*)bw.k0;
      for (int x = 0; x < L.n; ++x) {
        B.dst[x] = p;
        Lw.src[x] = p;
        p += align_up((size_t)st->nsel * L.bytes[x], 256);
      }
      TQ_TRY(launch("k_join_bucket_cols", k_join_bucket_cols, stream_grid(st->nsel), kBlock, 0,
                    s, (const uint32_t*)st->rows, st->nsel, L, B));
      bypos = 1;
    }
  }
  return launch("k_join_write", (k_join_write<K, HASH, false, true>),
                resident_grid(k_join_write<K, HASH, false, true>, 0, nt), kBlock, 0, s, J, pkeys,
                pdt, pmask, np, nt, off, Lw, R, bidx, pidx, bypos);
}

// Calls F<K, HASH, UNIQUE>::run(args...) for the state's key width and layout.
template <template <typename, bool, bool> class F, typename... A>
static int dispatch(const tq_join_state_t* st, A&&... a) {
  const bool h = st->layout == TQ_JOIN_HASH, u = st->unique != 0;
  if (st->key_bits == 32) {
    if (h) return u ? F<uint32_t, true, true>::run(a...) : F<uint32_t, true, false>::run(a...);
    return u ? F<uint32_t, false, true>::run(a...) : F<uint32_t, false, false>::run(a...);
  }
  if (h) return u ? F<uint64_t, true, true>::run(a...) : F<uint64_t, true, false>::run(a...);
  return u ? F<uint64_t, false, true>::run(a...) : F<uint64_t, false, false>::run(a...);
}

template <typename K, bool H, bool U>
struct CountF {
  template <typename... A>
  static int run(A&&... a) { return count_impl<K, H, U>(a...); }
};
template <typename K, bool H, bool U>
struct WriteF {
  template <typename... A>
  static int run(A&&... a) { return write_impl<K, H, U>(a...); }
};

}  // namespace tq

using namespace tq;

extern "C" size_t tq_join_build_ws_bytes(int64_t nb) { return carve_build(nullptr, nb).bytes; }
extern "C" size_t tq_join_probe_ws_bytes(int64_t np) { return carve_probe(nullptr, np).bytes; }

extern "C" int tq_join_build(const void* bkeys, int32_t bdtype, const uint8_t* bmask, int64_t nb,
                             void* ws, size_t ws_bytes, tq_join_state_t* state, void* stream) {
  if (!state) return set_error(TQ_EINVAL, "join build: null state");
  if (bdtype < TQ_BOOL || bdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", bdtype);
  if (nb >= ((int64_t)1 << 32) - 1) return set_error(TQ_EINVAL, "build side too large");
  BuildWs w = carve_build(ws, nb);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "join build workspace too small");
  *state = tq_join_state_t{};
  state->key_bits = bdtype == TQ_I64 ? 64 : 32;
  state->layout = TQ_JOIN_EMPTY;
  if (nb <= 0) return TQ_OK;
  hipStream_t s = (hipStream_t)stream;
  state->ws = ws;
  return bdtype == TQ_I64 ? build_impl<uint64_t>(bkeys, bdtype, bmask, nb, w, state, s)
                          : build_impl<uint32_t>(bkeys, bdtype, bmask, nb, w, state, s);
}

extern "C" int tq_join_build_ranged(const void* bkeys, int32_t bdtype, const uint8_t* bmask,
                                    int64_t nb, int64_t key_lo, int64_t key_hi, void* ws,
                                    size_t ws_bytes, tq_join_state_t* state, void* stream) {
  if (!state) return set_error(TQ_EINVAL, "join build: null state");
  if (bdtype < TQ_BOOL || bdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", bdtype);
  if (nb >= ((int64_t)1 << 32) - 1) return set_error(TQ_EINVAL, "build side too large");
  if (key_lo > key_hi) return set_error(TQ_EINVAL, "join build: key bounds %lld > %lld",
                                        (long long)key_lo, (long long)key_hi);
  BuildWs w = carve_build(ws, nb);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "join build workspace too small");
  *state = tq_join_state_t{};
  state->key_bits = bdtype == TQ_I64 ? 64 : 32;
  state->layout = TQ_JOIN_EMPTY;
  if (nb <= 0) return TQ_OK;
  hipStream_t s = (hipStream_t)stream;
  state->ws = ws;
  if (bdtype != TQ_I64) {  // bounds beyond the 32-bit key range cannot hold a key
    key_lo = key_lo < INT32_MIN ? INT32_MIN : key_lo;
    key_hi = key_hi > INT32_MAX ? INT32_MAX : key_hi;
    if (key_lo > key_hi) return build_impl<uint32_t>(bkeys, bdtype, bmask, nb, w, state, s);
    return build_ranged<uint32_t>(bkeys, bdtype, bmask, nb, key_lo, key_hi, w, state, s);
  }
  return build_ranged<uint64_t>(bkeys, bdtype, bmask, nb, key_lo, key_hi, w, state, s);
}

extern "C" int tq_join_count(tq_join_state_t* st, const void* pkeys, int32_t pdtype,
                             const uint8_t* pmask, int64_t np, void* pws, int64_t* total_host,
                             void* stream) {
  if (!st) return set_error(TQ_EINVAL, "join count: null state");
  if (pdtype < TQ_BOOL || pdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", pdtype);
  hipStream_t s = (hipStream_t)stream;
  ProbeWs w = carve_probe(pws, np);
  const int64_t nt = np > 0 ? ceil_div(np, kTileRows) : 0;
  if (st->pending) {
    // settle the build's uniqueness with the same synchronisation as this count
    BuildWs bw = carve_build(st->ws, st->nb);
    const bool dense = st->layout == TQ_JOIN_DENSE;
    const BuildCheck check = pending_check(st);
    // dup (u32), out-of-bounds (u32), occupied-entry sum (i64), selected rows (i64, ranged)
    const void* flag = (const void*)bw.dup;
    int64_t total = 0, fl[3] = {};
    if (np > 0) {
      TQ_TRY(dispatch<CountF>(st, st, pkeys, pdtype, pmask, np, nt, w, check, s));
      TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
      TQ_TRY(read_back2(w.total, 8, &total, flag, 24, fl, s));
    } else {
      if (check.out) TQ_TRY(launch("k_join_check", k_join_check, 1, kBlock, 0, s, check));
      TQ_TRY(read_back(flag, 24, fl, s));
    }
    st->pending = 0;
    const bool ranged = st->nsel < 0;
    const bool is64 = st->key_bits == 64;
    if (ranged) {
      st->nsel = fl[2];
      if ((fl[0] >> 32) != 0) {
        // a key outside the bounds given to the build: build again from the keys' statistics
        TQ_TRY(is64 ? rebuild_from_stats<uint64_t>(st, bw, false, s)
                    : rebuild_from_stats<uint32_t>(st, bw, false, s));
        return tq_join_count(st, pkeys, pdtype, pmask, np, pws, total_host, stream);
      }
      if (st->nsel == 0) st->layout = TQ_JOIN_EMPTY;  // nothing claimed: the count was 0
    }
    // dense: unique exactly when every selected row occupies its own directory entry
    const bool dup = st->nsel > 0 && (dense ? fl[1] != st->nsel : (fl[0] & 0xFFFFFFFF) != 0);
    if (!dup) {
      if (np <= 0) return tq_join_count(st, pkeys, pdtype, pmask, np, pws, total_host, stream);
      if (total_host) *total_host = total;
      return TQ_OK;
    }
    if (ranged) {
      // repeated keys need the statistics (per-tile counts, key bits) for the sorted layout
      TQ_TRY(is64 ? rebuild_from_stats<uint64_t>(st, bw, true, s)
                  : rebuild_from_stats<uint32_t>(st, bw, true, s));
    } else {
      TQ_TRY(is64 ? build_multi<uint64_t>(st, bw, s) : build_multi<uint32_t>(st, bw, s));
    }
  }
  if (np <= 0 || st->nsel <= 0 || st->layout == TQ_JOIN_EMPTY) {
    TQ_TRY(fill_u32(w.tile_counts, 0, nt, s));
    TQ_TRY(fill_u64(w.tile_off, 0, nt, s));
    TQ_TRY(fill_u64(w.total, 0, 1, s));
    if (total_host) *total_host = 0;
    return TQ_OK;
  }
  TQ_TRY(dispatch<CountF>(st, st, pkeys, pdtype, pmask, np, nt, w, BuildCheck{}, s));
  TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
  if (total_host) TQ_TRY(read_back(w.total, 8, total_host, s));
  return TQ_OK;
}

extern "C" int tq_join_probe_deferred(const tq_join_state_t* st, const void* pkeys,
                                      int32_t pdtype, const uint8_t* pmask, int64_t np, void* pws,
                                      int32_t nleft, const void* const* lsrc,
                                      const int32_t* lbytes, void* const* ldst, int32_t nright,
                                      const void* const* rsrc, const int32_t* rbytes,
                                      void* const* rdst, const int64_t** total_dev,
                                      const void** check_dev, int32_t* applied, void* stream) {
  if (!st || !total_dev || !check_dev || !applied)
    return set_error(TQ_EINVAL, "join probe (deferred): null argument");
  if (pdtype < TQ_BOOL || pdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", pdtype);
  *applied = 0;
  *total_dev = nullptr;
  *check_dev = nullptr;
  if (np <= 0 || !st->unique || st->layout == TQ_JOIN_EMPTY || st->nsel == 0) return TQ_OK;
  ColPack L, R;
  TQ_TRY(make_colpack(&L, nleft, lsrc, lbytes, nullptr, ldst));
  TQ_TRY(make_colpack(&R, nright, rsrc, rbytes, nullptr, rdst));
  hipStream_t s = (hipStream_t)stream;
  ProbeWs w = carve_probe(pws, np);
  const int64_t nt = ceil_div(np, kTileRows);
  if (st->pending)  // the same check words tq_join_count reads, left for the caller's read-back
    *check_dev = (const void*)carve_build(st->ws, st->nb).dup;
  TQ_TRY(dispatch<CountF>(st, st, pkeys, pdtype, pmask, np, nt, w, pending_check(st), s));
  TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
  TQ_TRY(dispatch<WriteF>(st, st, pkeys, pdtype, pmask, np, nt, w, L, R, (int64_t*)nullptr,
                          (int64_t*)nullptr, s));
  *total_dev = w.total;
  *applied = 1;
  return TQ_OK;
}

extern "C" int tq_join_settle(tq_join_state_t* st, const int64_t* words, int32_t* ok) {
  if (!st || !ok) return set_error(TQ_EINVAL, "join settle: null argument");
  *ok = 1;
  if (!st->pending) return TQ_OK;
  if (!words) return set_error(TQ_EINVAL, "join settle: the build's check words are needed");
  const bool ranged = st->nsel < 0;
  const bool dense = st->layout == TQ_JOIN_DENSE;
  const int64_t nsel = ranged ? words[2] : st->nsel;
  const uint64_t w0 = (uint64_t)words[0];
  const bool oob = ranged && (w0 >> 32) != 0;
  const bool dup = nsel > 0 && (dense ? words[1] != nsel : (w0 & 0xFFFFFFFFull) != 0);
  if (oob || dup) {
    *ok = 0;  // the state stays pending: tq_join_count settles it and builds again
    return TQ_OK;
  }
  st->nsel = nsel;
  st->pending = 0;
  if (nsel == 0) st->layout = TQ_JOIN_EMPTY;
  return TQ_OK;
}

extern "C" int tq_join_write(const tq_join_state_t* st, const void* pkeys, int32_t pdtype,
                             const uint8_t* pmask, int64_t np, const void* pws, int32_t nleft,
                             const void* const* lsrc, const int32_t* lbytes, void* const* ldst,
                             int32_t nright, const void* const* rsrc, const int32_t* rbytes,
                             void* const* rdst, int64_t* build_idx, int64_t* probe_idx,
                             void* stream) {
  if (!st) return set_error(TQ_EINVAL, "join write: null state");
  if (st->pending) return set_error(TQ_EINVAL, "join write: run tq_join_count first");
  ColPack L, R;
  TQ_TRY(make_colpack(&L, nleft, lsrc, lbytes, nullptr, ldst));
  TQ_TRY(make_colpack(&R, nright, rsrc, rbytes, nullptr, rdst));
  if (np <= 0 || st->nsel <= 0 || st->layout == TQ_JOIN_EMPTY) return TQ_OK;
  hipStream_t s = (hipStream_t)stream;
  ProbeWs w = carve_probe((void*)pws, np);
  const int64_t nt = ceil_div(np, kTileRows);
  return dispatch<WriteF>(st, st, pkeys, pdtype, pmask, np, nt, w, L, R, build_idx, probe_idx, s);
}
