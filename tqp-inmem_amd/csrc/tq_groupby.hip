// Group-by aggregation on packed keys.
//
// Keys are packed in mixed radix from their [min, max] ranges, so the packed value is the key
// tuple's rank in lexicographic order and every table below is addressed by it:
//   DENSE_LDS    - the whole key space fits a per-workgroup LDS array; each workgroup
//                  accumulates its rows there and merges occupied slots into the global array
//                  with one atomic per slot and aggregate;
//   DENSE_GLOBAL - slots of one global array addressed directly by the packed key;
//   HASH         - a global open-addressing table (linear probing, CAS on the key word).
// With the LDS_FRONT flag, a per-workgroup LDS open-addressing table (half-full limit) absorbs
// repeated keys first and only keys it cannot hold go to the global table, which keeps hot
// groups (skewed keys) off global atomics.
// Dense tables come out in key order by compacting occupied slots; hash tables are ordered by a
// radix sort of their packed keys.  Sums use float64 atomics (hardware global/LDS f64 add), so
// their summation order varies between runs; counts and integer sums are exact.
#include <map>
#include <mutex>

#include "tq_agg.h"
#include "tq_common.h"
#include "tq_expr.h"
#include "tq_sort.h"

namespace tq {

__device__ __forceinline__ int64_t gb_insert(const GbTable& T, uint64_t p) {
  const uint64_t mask = (uint64_t)T.slots - 1;
  uint64_t h = mix(p) >> (64 - T.log2);
  while (true) {
    const unsigned long long k = T.hkey[h];
    if (k == p) return (int64_t)h;
    if (k == kEmpty) {
      const unsigned long long old = atomicCAS(&T.hkey[h], kEmpty, (unsigned long long)p);
      if (old == kEmpty || old == p) return (int64_t)h;
    }
    h = (h + 1) & mask;
  }
}

__device__ __forceinline__ int64_t global_slot(int base_mode, const GbTable& T, uint64_t p) {
  return base_mode == TQ_GB_HASH ? gb_insert(T, p) : (int64_t)p;
}

// ------------------------------------------------------------------ bounds
struct KeyCols {
  const void* ptr[TQ_GB_MAX_KEYS];
  int32_t dtype[TQ_GB_MAX_KEYS];
  int32_t nkeys;
};

constexpr int kBoundWords = 1 + 2 * TQ_GB_MAX_KEYS;  // per tile: count, then (min, max) per key

// Per 4096-row tile: selected rows and each key's min and max -> part[tile * kBoundWords ..].
// A thread reads four runs of four consecutive rows (vector loads, all in flight together).
__global__ __launch_bounds__(kBlock) void k_gb_bounds(KeyCols K, const uint8_t* __restrict__ mask,
                                                      int64_t n, int64_t* __restrict__ part) {
  __shared__ int64_t sh[kBoundWords][kWaves];
  const int64_t tile = blockIdx.x;
  const int t = threadIdx.x;
  const uint32_t sel = load_sel16(mask, tile * kTileRows, n);
  int64_t cnt = __popc(sel);
  int64_t lo[TQ_GB_MAX_KEYS], hi[TQ_GB_MAX_KEYS];
#pragma unroll
  for (int k = 0; k < TQ_GB_MAX_KEYS; ++k) {
    lo[k] = INT64_MAX;
    hi[k] = INT64_MIN;
    if (k >= K.nkeys) continue;
    uint64_t raw[kItems];
    load_raw16(K.ptr[k], K.dtype[k], tile * kTileRows, n, raw);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int64_t vj = (int64_t)raw[j];
      if ((sel >> j) & 1u) {
        lo[k] = vj < lo[k] ? vj : lo[k];
        hi[k] = vj > hi[k] ? vj : hi[k];
      }
    }
  }
  cnt = wave_sum(cnt);
#pragma unroll
  for (int k = 0; k < TQ_GB_MAX_KEYS; ++k) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const int64_t x = __shfl_xor(lo[k], d, 64), y = __shfl_xor(hi[k], d, 64);
      lo[k] = x < lo[k] ? x : lo[k];
      hi[k] = y > hi[k] ? y : hi[k];
    }
  }
  const int w = t >> 6;
  if (lane_id() == 0) {
    sh[0][w] = cnt;
#pragma unroll
    for (int k = 0; k < TQ_GB_MAX_KEYS; ++k) {
      sh[1 + 2 * k][w] = lo[k];
      sh[2 + 2 * k][w] = hi[k];
    }
  }
  __syncthreads();
  if (t < kBoundWords) {
    int64_t r = sh[t][0];
    for (int i = 1; i < kWaves; ++i) {
      const int64_t x = sh[t][i];
      r = t == 0 ? r + x : ((t & 1) ? (x < r ? x : r) : (x > r ? x : r));
    }
    part[tile * kBoundWords + t] = r;
  }
}

// One workgroup: out = {min0, max0, ..., min_{k-1}, max_{k-1}, selected rows}.
__global__ __launch_bounds__(kBlock) void k_gb_bounds_final(const int64_t* __restrict__ part,
                                                            int64_t ntiles, int nkeys,
                                                            int64_t* __restrict__ out) {
  __shared__ int64_t sh[kBoundWords][kWaves];
  int64_t acc[kBoundWords];
#pragma unroll
  for (int i = 0; i < kBoundWords; ++i) acc[i] = i == 0 ? 0 : ((i & 1) ? INT64_MAX : INT64_MIN);
  for (int64_t tl = threadIdx.x; tl < ntiles; tl += kBlock) {
#pragma unroll
    for (int i = 0; i < kBoundWords; ++i) {
      const int64_t x = part[tl * kBoundWords + i];
      acc[i] = i == 0 ? acc[i] + x : ((i & 1) ? (x < acc[i] ? x : acc[i]) : (x > acc[i] ? x : acc[i]));
    }
  }
#pragma unroll
  for (int i = 0; i < kBoundWords; ++i) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
      const int64_t x = __shfl_xor(acc[i], d, 64);
      acc[i] = i == 0 ? acc[i] + x : ((i & 1) ? (x < acc[i] ? x : acc[i]) : (x > acc[i] ? x : acc[i]));
    }
  }
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
#pragma unroll
    for (int i = 0; i < kBoundWords; ++i) sh[i][w] = acc[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int v = 1; v < kWaves; ++v) {
#pragma unroll
      for (int i = 0; i < kBoundWords; ++i) {
        const int64_t x = sh[i][v];
        acc[i] = i == 0 ? acc[i] + x : ((i & 1) ? (x < acc[i] ? x : acc[i]) : (x > acc[i] ? x : acc[i]));
      }
    }
#pragma unroll
    for (int k = 0; k < TQ_GB_MAX_KEYS; ++k) {
      if (k < nkeys) {
        out[2 * k] = acc[1 + 2 * k];
        out[2 * k + 1] = acc[2 + 2 * k];
      }
    }
    out[2 * nkeys] = acc[0];
  }
}

// ------------------------------------------------------------------ accumulate
__global__ __launch_bounds__(kBlock) void k_gb_clear(GbTable T, GbSpec G, int hashed) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t s = (int64_t)blockIdx.x * kBlock + threadIdx.x; s < T.slots; s += stride) {
    T.cnt[s] = 0;
    if (hashed) T.hkey[s] = kEmpty;
    for (int j = 0; j < G.naggs; ++j) T.acc[(int64_t)j * T.slots + s] = G.init[j];
  }
}

__device__ __forceinline__ uint64_t pack_key(const GbSpec& G, const uint64_t* stage, int j) {
  uint64_t p = 0;
  for (int k = 0; k < G.nkeys; ++k) {
    const int64_t v = (int64_t)stage[(G.key_slot[k] * kSub + j) * kBlock + threadIdx.x];
    p += key_digit(G, k, v) * (uint64_t)G.key_stride[k];
  }
  return p;
}

// BASE == DENSE_LDS: the workgroup's table is the whole key space in LDS (with LANES one copy
// per lane: entry s of lane l at s * 64 + l, so the rows of a few hot groups never meet on an
// address); at the end the workgroup writes its table to part[blockIdx.x] and k_gb_merge
// combines the partials, so no global atomics are contended.  FRONT: an LDS hash table of cap
// entries (filled to 3/4) absorbs repeated keys; rows whose key finds no place go to the
// global table, and the LDS entries are added into it at the end.
template <int BASE, bool FRONT, bool LANES>
__global__ __launch_bounds__(kBlock) void k_gb_accum(GbSpec G, tq_program_t P, tq_colset_t C,
                                                     const uint8_t* __restrict__ mask, int64_t n,
                                                     GbTable T, int32_t cap, int32_t stage_words,
                                                     uint64_t* __restrict__ part) {
  extern __shared__ uint64_t smem[];
  uint64_t* stage = smem;
  uint64_t* lacc = smem + stage_words;           // naggs * cap accumulators
  unsigned long long* lkey = (unsigned long long*)(lacc + (int64_t)G.naggs * cap);
  uint32_t* lcnt = (uint32_t*)(lkey + (FRONT ? cap : 0));
  __shared__ uint32_t fill;
  const bool use_lds = BASE == TQ_GB_DENSE_LDS || FRONT;
  const int t = threadIdx.x;
  const int lane = lane_id();
  if (use_lds) {
    for (int s = t; s < cap; s += kBlock) {
      lcnt[s] = 0;
      if (FRONT) lkey[s] = kEmpty;
      for (int a = 0; a < G.naggs; ++a) lacc[(int64_t)a * cap + s] = G.init[a];
    }
    if (t == 0) fill = 0;
    __syncthreads();
  }
  const int64_t nsub = ceil_div(n, kSubRows);
  const uint32_t limit = (uint32_t)cap / 4 * 3;
  for (int64_t sub = blockIdx.x; sub < nsub; sub += gridDim.x) {
    const int64_t sub0 = sub * kSubRows;
    bool live[kSub];
    select_rows(mask, sub0, n, live);
    stage_loads(P, C, sub0, live, stage);
    // the table slot of every live row: LDS (dense-LDS tables and the hash front; inl[j]) or
    // global.  The accumulators are then addressed through lacc or T.acc themselves, never a
    // pointer that may be either: an LDS-or-global pointer compiles to flat atomics, which are
    // far slower on LDS than its own atomics.
    bool inl[kSub];
    int64_t aslot[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      inl[j] = BASE == TQ_GB_DENSE_LDS;
      aslot[j] = 0;
      if (!live[j]) continue;
      const uint64_t p = pack_key(G, stage, j);
      if (BASE == TQ_GB_DENSE_LDS) {
        aslot[j] = LANES ? (int64_t)p * 64 + lane : (int64_t)p;
        atomicAdd(&lcnt[aslot[j]], 1u);
        continue;
      }
      if (FRONT) {
        const uint32_t m = (uint32_t)cap - 1;
        uint32_t h = (uint32_t)(mix(p) >> 40) & m;
        int64_t ls = -1;
        for (int probe = 0; probe < cap; ++probe) {
          const unsigned long long k = lkey[h];
          if (k == p) { ls = h; break; }
          if (k == kEmpty) {
            if (fill >= limit) break;
            const unsigned long long old = atomicCAS(&lkey[h], kEmpty, (unsigned long long)p);
            if (old == kEmpty) { atomicAdd(&fill, 1u); ls = h; break; }
            if (old == p) { ls = h; break; }
          }
          h = (h + 1) & m;
        }
        if (ls >= 0) {
          inl[j] = true;
          aslot[j] = ls;
          atomicAdd(&lcnt[ls], 1u);
          continue;
        }
      }
      aslot[j] = global_slot(BASE, T, p);
      atomicAdd(&T.cnt[aslot[j]], 1u);
    }
    // every aggregate argument for the sub-tile's rows at once, then one atomic per live row
    int64_t rows[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) rows[j] = sub0 + (int64_t)j * kBlock + t;
    // one switch on the accumulator kind per aggregate, its four rows inside (a switch per row
    // costs the scalar unit, which every wave of the CU shares, more than the rows' memory)
#define TQ_ADD(J, A, V, K)                                                              \
  if (live[J]) {                                                                        \
    if (BASE == TQ_GB_DENSE_LDS || (FRONT && inl[J]))                                   \
      acc_apply(lacc + (int64_t)(A) * cap + aslot[J], K, (V));                          \
    else                                                                                \
      acc_apply(T.acc + (int64_t)(A) * T.slots + aslot[J], K, (V));                     \
  }
#define TQ_ADD4(A, V, K) TQ_ADD(0, A, V[0], K) TQ_ADD(1, A, V[1], K) TQ_ADD(2, A, V[2], K) \
  TQ_ADD(3, A, V[3], K)
#define TQ_APPLY(A, V)                                           \
  switch (G.kind[A]) {                                           \
    case ACC_FSUM: TQ_ADD4(A, V, ACC_FSUM) break;                \
    case ACC_ISUM: TQ_ADD4(A, V, ACC_ISUM) break;                \
    case ACC_COUNT: TQ_ADD4(A, V, ACC_COUNT) break;              \
    case ACC_FMIN: TQ_ADD4(A, V, ACC_FMIN) break;                \
    case ACC_FMAX: TQ_ADD4(A, V, ACC_FMAX) break;                \
    case ACC_IMIN: TQ_ADD4(A, V, ACC_IMIN) break;                \
    default: TQ_ADD4(A, V, ACC_IMAX) break;                      \
  }
    if (G.D.direct) {
      // the arguments straight from the staged loads (no interpreter)
      for (int a = 0; a < G.naggs; ++a) {
        uint64_t v[kSub];
        direct_args(G.D, G.kind[a], a, stage, v);
        TQ_APPLY(a, v)
      }
    } else {
      RegsN R{};
      int begin = 0;
      for (int a = 0; a < G.naggs; ++a) {
        const int end = P.seg_end[a];
        run_rows(P, C, begin, end, rows, live, stage, R);
        begin = end;
        const V4 V = R.get(P.seg_result[a]);
        uint64_t v[kSub];
#pragma unroll
        for (int j = 0; j < kSub; ++j) v[j] = to_acc(G.kind[a], G.arg_dtype[a], V.v[j]);
        TQ_APPLY(a, v)
      }
    }
#undef TQ_APPLY
#undef TQ_ADD4
#undef TQ_ADD
  }
  if (BASE == TQ_GB_DENSE_LDS) {
    // this workgroup's table -> part[blockIdx.x][slot][count, aggregates...]
    __syncthreads();
    const int64_t S = LANES ? cap / 64 : cap;
    const int nf = 1 + G.naggs;
    uint64_t* mine = part + (int64_t)blockIdx.x * S * nf;
    if (LANES) {
      // wave w combines the 64 lane copies of slots w, w + 4, ...
      for (int64_t s = t >> 6; s < S; s += kWaves) {
        uint64_t c = lcnt[s * 64 + lane];
        c = wave_sum(c);
        if (lane == 0) mine[s * nf] = c;
        for (int a = 0; a < G.naggs; ++a) {
          uint64_t v = lacc[(int64_t)a * cap + s * 64 + lane];
          const int kind = G.kind[a];
#pragma unroll
          for (int d = 32; d >= 1; d >>= 1) v = acc_combine(kind, v, __shfl_xor(v, d, 64));
          if (lane == 0) mine[s * nf + 1 + a] = v;
        }
      }
    } else {
      for (int64_t s = t; s < S; s += kBlock) {
        mine[s * nf] = lcnt[s];
        for (int a = 0; a < G.naggs; ++a) mine[s * nf + 1 + a] = lacc[(int64_t)a * cap + s];
      }
    }
  } else if (FRONT) {
    __syncthreads();
    for (int s = t; s < cap; s += kBlock) {
      const uint32_t c = lcnt[s];
      if (c == 0) continue;
      const uint64_t p = (uint64_t)lkey[s];
      const int64_t gs = global_slot(BASE, T, p);
      atomicAdd(&T.cnt[gs], c);
      for (int a = 0; a < G.naggs; ++a)
        acc_apply(&T.acc[(int64_t)a * T.slots + gs], G.kind[a], lacc[(int64_t)a * cap + s]);
    }
  }
}

// Dense-LDS partials of nblocks workgroups combined into the table (count, then aggregates):
// one wave per field, its lanes splitting the partials and combining them by shuffles.
__global__ __launch_bounds__(kBlock) void k_gb_merge(GbSpec G, const uint64_t* __restrict__ part,
                                                     int32_t nblocks, GbTable T) {
  const int nf = 1 + G.naggs;
  const int64_t S = T.slots, total = S * nf;
  const int l = lane_id();
  const int64_t nwaves = (int64_t)gridDim.x * kWaves;
  for (int64_t e = (int64_t)blockIdx.x * kWaves + (threadIdx.x >> 6); e < total; e += nwaves) {
    const int64_t s = e / nf;
    const int f = (int)(e - s * nf);
    const int kind = f == 0 ? ACC_COUNT : G.kind[f - 1];
    const uint64_t init = f == 0 ? 0ull : G.init[f - 1];
    uint64_t v = init;
    for (int b0 = l; b0 < nblocks; b0 += 4 * 64) {
      uint64_t x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int b = b0 + u * 64;
        x[u] = b < nblocks ? part[(int64_t)b * total + e] : init;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) v = acc_combine(kind, v, x[u]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = acc_combine(kind, v, __shfl_xor(v, d, 64));
    if (l == 0) {
      if (f == 0) T.cnt[s] = (uint32_t)v;
      else T.acc[(int64_t)(f - 1) * S + s] = v;
    }
  }
}

// ------------------------------------------------------------------ count / write
__global__ __launch_bounds__(kBlock) void k_gb_occupied(const uint32_t* __restrict__ cnt,
                                                        int64_t slots, int force0,
                                                        uint32_t* __restrict__ tile_counts) {
  __shared__ uint32_t sh[kWaves];
  const int64_t base = (int64_t)blockIdx.x * kTileRows + threadIdx.x;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t s = base + (int64_t)j * kBlock;
    if (s < slots) c += (cnt[s] > 0) || (force0 && s == 0);
  }
  c = wave_sum(c);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s += sh[w];
    tile_counts[blockIdx.x] = s;
  }
}

struct OutCols {
  void* key[TQ_GB_MAX_KEYS];
  void* agg[TQ_GB_MAX_AGGS];
  int32_t op[TQ_GB_MAX_AGGS];
};

// Key k of packed key p (32-bit division when the packed space allows it); a dictionary digit
// is looked up.
__device__ __forceinline__ int64_t key_of(const GbSpec& G, int k, uint64_t p) {
  uint64_t q;
  if (G.nkeys == 1)
    q = p;
  else if (p <= 0xFFFFFFFFull && G.key_stride[k] <= 0xFFFFFFFFll && G.key_range[k] <= 0xFFFFFFFFll)
    q = ((uint32_t)p / (uint32_t)G.key_stride[k]) % (uint32_t)G.key_range[k];
  else
    q = (p / (uint64_t)G.key_stride[k]) % (uint64_t)G.key_range[k];
  return G.dict_val[k] ? G.dict_val[k][q] : (int64_t)q + G.key_min[k];
}

__device__ __forceinline__ void write_group(const GbSpec& G, const OutCols& O, const GbTable& T,
                                            int64_t o, int64_t slot, uint64_t p) {
  for (int k = 0; k < G.nkeys; ++k) store_int(O.key[k], G.key_dtype[k], o, key_of(G, k, p));
  const uint32_t rows = T.cnt[slot];
  for (int a = 0; a < G.naggs; ++a)
    write_agg(O.agg[a], o, O.op[a], G.kind[a], G.arg_dtype[a], T.acc[(int64_t)a * T.slots + slot], rows);
}

// Dense tables: groups are the occupied slots in slot (= key) order.  A thread ranks its 16
// slots of a 4096-slot tile first, then writes each output column with all 16 loads in flight.
__global__ __launch_bounds__(kBlock) void k_gb_write_dense(GbSpec G, OutCols O, GbTable T,
                                                           const int64_t* __restrict__ tile_off,
                                                           int force0) {
  __shared__ uint32_t wcnt[kItems][kWaves];
  const int t = threadIdx.x, w = t >> 6, l = lane_id();
  const int64_t tile0 = (int64_t)blockIdx.x * kTileRows;
  uint32_t cnt[kItems], rel[kItems];
  uint32_t occ = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t s = tile0 + (int64_t)j * kBlock + t;
    cnt[j] = s < T.slots ? T.cnt[s] : 0u;
    const bool f = s < T.slots && (cnt[j] > 0 || (force0 && s == 0));
    const uint64_t bal = ballot(f);
    if (l == 0) wcnt[j][w] = (uint32_t)__popcll(bal);
    rel[j] = (uint32_t)rank_below(bal);
    occ |= (uint32_t)f << j;
  }
  __syncthreads();
  uint32_t run = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      const uint32_t c = wcnt[j][v];
      tot += c;
      before += (v < w) ? c : 0;
    }
    rel[j] += run + before;
    run += tot;
  }
  const int64_t base = tile_off[blockIdx.x];
  for (int k = 0; k < G.nkeys; ++k) {
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if ((occ >> j) & 1u)
        store_int(O.key[k], G.key_dtype[k], base + rel[j],
                  key_of(G, k, (uint64_t)(tile0 + (int64_t)j * kBlock + t)));
  }
  if (T.slots <= kBlock) {
    // a table of at most one slot per thread: every aggregate's accumulator is loaded before
    // the first is written (a loop of load-then-write per aggregate waits once per aggregate)
    uint64_t va[TQ_GB_MAX_AGGS];
#pragma unroll
    for (int a = 0; a < TQ_GB_MAX_AGGS; ++a)
      va[a] = (a < G.naggs && (occ & 1u)) ? T.acc[(int64_t)a * T.slots + t] : 0ull;
    if (occ & 1u) {
#pragma unroll
      for (int a = 0; a < TQ_GB_MAX_AGGS; ++a)
        if (a < G.naggs)
          write_agg(O.agg[a], base + rel[0], O.op[a], G.kind[a], G.arg_dtype[a], va[a], cnt[0]);
    }
    return;
  }
  for (int a = 0; a < G.naggs; ++a) {
    const uint64_t* acc = T.acc + (int64_t)a * T.slots;
    uint64_t v[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      v[j] = ((occ >> j) & 1u) ? acc[tile0 + (int64_t)j * kBlock + t] : 0ull;
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if ((occ >> j) & 1u)
        write_agg(O.agg[a], base + rel[j], O.op[a], G.kind[a], G.arg_dtype[a], v[j], cnt[j]);
  }
}

// ------------------------------------------------------------------ partitioned dense tables
// The key space [0, S) is cut into partitions of ps = 2^pb slots.  A scatter tile is tsub
// sub-tiles of 1024 rows; hist[b * ptiles + tile] counts the tile's selected rows in partition
// b, and its exclusive scan (partition-major) is where the tile's rows of partition b go, so
// the scatter needs no global atomics.

// Packed keys of rows sub0 + j*256 + t (j < kSub; rows not live read row 0 and are ignored):
// per key column the dtype is dispatched once and the kSub loads are issued together.
template <typename T>
__device__ __forceinline__ void add_key4(const GbSpec& G, int k, const void* col, int64_t sub0,
                                         const bool (&live)[kSub], uint64_t (&pk)[kSub]) {
  const T* p = (const T*)col;
  T x[kSub];
#pragma unroll
  for (int j = 0; j < kSub; ++j) x[j] = p[live[j] ? sub0 + (int64_t)j * kBlock + threadIdx.x : 0];
  const uint64_t stride = (uint64_t)G.key_stride[k];
#pragma unroll
  for (int j = 0; j < kSub; ++j)  // rows that are not live read row 0, which may lie outside
    pk[j] += live[j] ? key_digit(G, k, (int64_t)x[j]) * stride : 0;  // the key's range
}

__device__ __forceinline__ void packed_keys4(const KeyCols& K, const GbSpec& G, int64_t sub0,
                                             const bool (&live)[kSub], uint64_t (&pk)[kSub]) {
#pragma unroll
  for (int j = 0; j < kSub; ++j) pk[j] = 0;
  for (int k = 0; k < G.nkeys; ++k) {
    switch (K.dtype[k]) {
      case TQ_BOOL: add_key4<uint8_t>(G, k, K.ptr[k], sub0, live, pk); break;  // 0 or 1
      case TQ_I8: add_key4<int8_t>(G, k, K.ptr[k], sub0, live, pk); break;
      case TQ_I16: add_key4<int16_t>(G, k, K.ptr[k], sub0, live, pk); break;
      case TQ_I32: add_key4<int32_t>(G, k, K.ptr[k], sub0, live, pk); break;
      default: add_key4<int64_t>(G, k, K.ptr[k], sub0, live, pk); break;
    }
  }
}

// Workgroup 0 also clears the look-back words of the fused reduce (flags[0..nparts + 1)).
// Rows and histogram stride of a partitioned pass.  With n_dev, a producer earlier on the stream
// left the row count in device memory and the rows are the first min(n, *n_dev), n being the
// capacity the launches were sized for: only the tiles holding rows take part, and the
// histogram is laid out with that many tiles per partition, so its first nparts * tiles entries
// are the whole histogram (the scan runs over the capacity; what follows is never used).
__device__ __forceinline__ int64_t part_rows(int64_t n, const int64_t* n_dev) {
  if (!n_dev) return n;
  const int64_t m = *n_dev;
  return m < n ? (m > 0 ? m : 0) : n;
}
__device__ __forceinline__ int64_t part_stride(int64_t rows, const int64_t* n_dev, int tsub,
                                               int64_t ptiles) {
  return n_dev ? ceil_div(rows, (int64_t)tsub * kSubRows) : ptiles;
}

__global__ __launch_bounds__(kBlock) void k_gbp_hist(KeyCols K, GbSpec G,
                                                     const uint8_t* __restrict__ mask, int64_t n,
                                                     const int64_t* __restrict__ n_dev,
                                                     int pb, int nparts, int tsub, int64_t ptiles,
                                                     uint32_t* __restrict__ hist,
                                                     uint64_t* __restrict__ flags) {
  extern __shared__ uint32_t lh[];
  const int t = threadIdx.x;
  n = part_rows(n, n_dev);
  const int64_t pt = part_stride(n, n_dev, tsub, ptiles);
  if (flags && blockIdx.x == 0)
    for (int i = t; i <= nparts; i += kBlock) flags[i] = 0;
  const int64_t tile = blockIdx.x;
  if (tile >= pt) return;
  for (int i = t; i < nparts; i += kBlock) lh[i] = 0;
  __syncthreads();
  for (int sub = 0; sub < tsub; ++sub) {
    const int64_t sub0 = (tile * tsub + sub) * kSubRows;
    bool live[kSub];
    select_rows(mask, sub0, n, live);
    uint64_t pk[kSub];
    packed_keys4(K, G, sub0, live, pk);
#pragma unroll
    for (int j = 0; j < kSub; ++j)
      if (live[j]) atomicAdd(&lh[pk[j] >> pb], 1u);
  }
  __syncthreads();
  for (int i = t; i < nparts; i += kBlock) hist[(int64_t)i * pt + tile] = lh[i];
}

// Scatter a tile's selected rows: each takes the next position of its partition's range and
// stores one record there, {slot within the partition, aggregate arguments (accumulator
// domain)...} in 1 + naggs words (one 16-byte store for a single aggregate), so a row costs
// one scattered write rather than one per field.
__global__ __launch_bounds__(kBlock) void k_gbp_scatter(
    GbSpec G, tq_program_t P, tq_colset_t C, const uint8_t* __restrict__ mask, int64_t n,
    const int64_t* __restrict__ n_dev, int pb, int nparts, int tsub, int64_t ptiles,
    const uint32_t* __restrict__ hoff, uint64_t* __restrict__ rec, int32_t stage_words) {
  extern __shared__ uint64_t smem[];
  uint64_t* stage = smem;
  uint32_t* lh = (uint32_t*)(smem + stage_words);
  const int t = threadIdx.x;
  const int64_t tile = blockIdx.x;
  n = part_rows(n, n_dev);  // rows and stride as in k_gbp_hist
  const int64_t pt = part_stride(n, n_dev, tsub, ptiles);
  if (tile >= pt) return;
  for (int i = t; i < nparts; i += kBlock) lh[i] = hoff[(int64_t)i * pt + tile];
  __syncthreads();
  const uint64_t pmask = ((uint64_t)1 << pb) - 1;
  for (int sub = 0; sub < tsub; ++sub) {
    const int64_t sub0 = (tile * tsub + sub) * kSubRows;
    bool live[kSub];
    int64_t rows[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) rows[j] = sub0 + (int64_t)j * kBlock + t;
    select_rows(mask, sub0, n, live);
    stage_loads(P, C, sub0, live, stage);
    uint32_t pos[kSub];
    uint64_t slot[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      pos[j] = 0;
      slot[j] = 0;
      if (!live[j]) continue;
      const uint64_t pk = pack_key(G, stage, j);
      pos[j] = atomicAdd(&lh[pk >> pb], 1u);
      slot[j] = pk & pmask;
    }
    const int rw = 1 + G.naggs;
    RegsN R{};
    int begin = 0;
    for (int a = 0; a < G.naggs; ++a) {
      uint64_t v[kSub];
      if (G.D.direct) {
        direct_args(G.D, G.kind[a], a, stage, v);  // the arguments straight from the staged loads
      } else {
        const int end = P.seg_end[a];
        run_rows(P, C, begin, end, rows, live, stage, R);
        begin = end;
        const V4 V = R.get(P.seg_result[a]);
#pragma unroll
        for (int j = 0; j < kSub; ++j) v[j] = to_acc(G.kind[a], G.arg_dtype[a], V.v[j]);
      }
#pragma unroll
      for (int j = 0; j < kSub; ++j) {
        if (!live[j]) continue;
        if (G.naggs == 1)
          *(ulonglong2*)(rec + (int64_t)pos[j] * 2) = make_ulonglong2(slot[j], v[j]);
        else
          rec[(int64_t)pos[j] * rw + 1 + a] = v[j];
      }
    }
    if (G.naggs != 1) {
#pragma unroll
      for (int j = 0; j < kSub; ++j)
        if (live[j]) rec[(int64_t)pos[j] * rw] = slot[j];
    }
  }
}

// Look-back word of a partition: (groups << 2) | 1 once its own group count is known (0 until
// then).
struct LookBack {
  uint64_t* flags;   // [nparts] look-back words, then [nparts] = error flag
  int64_t* ngroups;  // total groups (written by the last partition)
  int64_t out_cap;   // rows of the outputs
  uint32_t spin_limit;
};

// The words carry only their own values (no other data is published through them), so relaxed
// device-scope atomics suffice: they reach every XCD without the cache write-back and
// invalidation that release/acquire ordering costs.
__device__ __forceinline__ uint64_t ld_word64(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word64(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup per partition: reduce its rows into an LDS table, then compact its occupied
// slots in slot order.  Unfused: they go to [b << pb, ...) of gslot/gcnt/gacc and their number
// to grp_cnt[b], for a scan and k_gbp_write.  FUSED: the workgroup publishes its group count
// and adds up those of the partitions before it to find its output offset (workgroups start in
// index order, so every one waited on is running or done; the wait is bounded and a timeout
// sets the error word, after which the host takes the unfused path), then writes the final key
// and aggregate columns itself.
template <bool FUSED>
__global__ __launch_bounds__(kBlock) void k_gbp_reduce(
    GbSpec G, int pb, int nparts, int64_t ptiles, int64_t n, const int64_t* __restrict__ n_dev,
    int tsub, const uint32_t* __restrict__ hoff,
    const uint32_t* __restrict__ htotal, const uint64_t* __restrict__ rec,
    uint32_t* __restrict__ grp_cnt, uint32_t* __restrict__ gslot, uint32_t* __restrict__ gcnt,
    uint64_t* __restrict__ gacc, OutCols O, LookBack LB) {
  extern __shared__ uint64_t smem[];
  __shared__ uint32_t wocc[kItems][kWaves];
  __shared__ int64_t s_off;
  const int ps = 1 << pb;
  uint64_t* lacc = smem;                                   // naggs * ps
  uint32_t* lcnt = (uint32_t*)(smem + (int64_t)G.naggs * ps);
  const int t = threadIdx.x;
  const int64_t b = blockIdx.x;
  for (int i = t; i < ps; i += kBlock) {
    lcnt[i] = 0;
    for (int a = 0; a < G.naggs; ++a) lacc[a * ps + i] = G.init[a];
  }
  __syncthreads();
  // partition b's records: from its first histogram entry to the next partition's (the grand
  // total after the last partition); no rows, no records
  const int64_t pt = part_stride(part_rows(n, n_dev), n_dev, tsub, ptiles);
  const uint32_t r0 = pt > 0 ? hoff[b * pt] : 0u;
  const uint32_t r1 = pt > 0 ? (b + 1 < nparts ? hoff[(b + 1) * pt] : *htotal) : 0u;
  constexpr int U = 8;  // records per thread whose loads are issued together
  const int rw = 1 + G.naggs;
  for (uint32_t base = r0 + t; base < r1; base += U * kBlock) {
    // records past r1 read record r0 (the partition has rows when the loop runs) and are dropped
    uint32_t sl[U];
    if (G.naggs == 1) {
      ulonglong2 x[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t r = base + u * kBlock;
        x[u] = *(const ulonglong2*)(rec + (int64_t)(r < r1 ? r : r0) * 2);
      }
      const int kind = G.kind[0];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (base + u * kBlock >= r1) continue;
        sl[u] = (uint32_t)x[u].x;
        atomicAdd(&lcnt[sl[u]], 1u);
        acc_apply(&lacc[sl[u]], kind, x[u].y);
      }
      continue;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t r = base + u * kBlock;
      sl[u] = (uint32_t)rec[(int64_t)(r < r1 ? r : r0) * rw];
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (base + u * kBlock < r1) atomicAdd(&lcnt[sl[u]], 1u);
    for (int a = 0; a < G.naggs; ++a) {
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t r = base + u * kBlock;
        v[u] = rec[(int64_t)(r < r1 ? r : r0) * rw + 1 + a];
      }
      const int kind = G.kind[a];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (base + u * kBlock < r1) acc_apply(&lacc[a * ps + sl[u]], kind, v[u]);
    }
  }
  __syncthreads();
  // stable compaction of the occupied slots: slot i = j*256 + t, ranked by wave ballots and one
  // exchange of the per-(j, wave) counts
  const int64_t gbase = b << pb;
  const int64_t gstride = (int64_t)nparts << pb;
  const int w = t >> 6;
  uint64_t bal[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int i = j * kBlock + t;
    bal[j] = ballot(i < ps && lcnt[i] > 0);
    if (lane_id() == 0) wocc[j][w] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  int64_t obase = 0;
  if (FUSED) {
    if (t < 64) {
      // wave 0: publish this partition's group count, then add up the counts of every partition
      // before it, 64 per load round (all partitions reduce at once, so nothing waits on a
      // chain of offsets)
      const int l = lane_id();
      uint64_t own = 0;
#pragma unroll
      for (int j = 0; j < kItems; ++j)
#pragma unroll
        for (int v = 0; v < kWaves; ++v) own += wocc[j][v];
      if (l == 0) st_word64(LB.flags + b, (own << 2) | 1ull);
      uint64_t before = 0;
      bool ok = true;
      for (int64_t q0 = 0; q0 < b && ok; q0 += 64) {
        const int64_t q = q0 + l;
        uint64_t f = 1ull;  // lanes past b - 1 add nothing
        uint32_t spins = 0;
        while (true) {
          if (q < b) f = ld_word64(LB.flags + q);
          if (ballot(f == 0) == 0) break;
          if (++spins > LB.spin_limit) {
            ok = false;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        before += wave_sum(q < b ? (f >> 2) : 0ull);
      }
      if (l == 0) {
        if (!ok) st_word64(LB.flags + nparts, 1ull);
        if (b == nparts - 1) *LB.ngroups = (int64_t)(before + own);
        s_off = ok ? (int64_t)before : -1;
      }
    }
    __syncthreads();
    obase = s_off;
  }
  uint32_t run = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      const uint32_t c = wocc[j][v];
      tot += c;
      before += (v < w) ? c : 0u;
    }
    if ((bal[j] >> lane_id()) & 1u) {
      const int i = j * kBlock + t;
      const uint32_t local = run + before + rank_below(bal[j]);
      if (FUSED) {
        const int64_t o = obase + local;
        if (obase >= 0 && o < LB.out_cap) {
          const uint64_t pk = (uint64_t)gbase + (uint64_t)i;
          for (int k = 0; k < G.nkeys; ++k) store_int(O.key[k], G.key_dtype[k], o, key_of(G, k, pk));
          for (int a = 0; a < G.naggs; ++a)
            write_agg(O.agg[a], o, O.op[a], G.kind[a], G.arg_dtype[a], lacc[a * ps + i], lcnt[i]);
        } else if (obase >= 0) {
          st_word64(LB.flags + nparts, 1ull);  // more groups than the outputs hold
        }
      } else {
        const int64_t o = gbase + local;
        gslot[o] = (uint32_t)i;
        gcnt[o] = lcnt[i];
        for (int a = 0; a < G.naggs; ++a) gacc[(int64_t)a * gstride + o] = lacc[a * ps + i];
      }
    }
    run += tot;
  }
  if (!FUSED && t == 0) grp_cnt[b] = run;
}

// Write partition b's groups at grp_off[b]...: keys decoded from the packed key, aggregates
// finished in their output dtypes.
__global__ __launch_bounds__(kBlock) void k_gbp_write(GbSpec G, OutCols O, int pb, int nparts,
                                                      const uint32_t* __restrict__ grp_cnt,
                                                      const int64_t* __restrict__ grp_off,
                                                      const uint32_t* __restrict__ gslot,
                                                      const uint32_t* __restrict__ gcnt,
                                                      const uint64_t* __restrict__ gacc) {
  const int64_t b = blockIdx.x;
  const uint32_t m = grp_cnt[b];
  const int64_t o0 = grp_off[b];
  const int64_t gbase = b << pb;
  const int64_t gstride = (int64_t)nparts << pb;
  constexpr int U = 4;  // groups per thread whose loads are issued together
  for (uint32_t i0 = threadIdx.x; i0 < m; i0 += U * kBlock) {
    uint32_t sl[U], rows[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * kBlock;
      sl[u] = i < m ? gslot[gbase + i] : 0u;
      rows[u] = i < m ? gcnt[gbase + i] : 0u;
    }
    for (int k = 0; k < G.nkeys; ++k) {
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t i = i0 + u * kBlock;
        if (i < m) store_int(O.key[k], G.key_dtype[k], o0 + i, key_of(G, k, (uint64_t)gbase + sl[u]));
      }
    }
    for (int a = 0; a < G.naggs; ++a) {
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t i = i0 + u * kBlock;
        v[u] = i < m ? gacc[(int64_t)a * gstride + gbase + i] : 0ull;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t i = i0 + u * kBlock;
        if (i < m) write_agg(O.agg[a], o0 + i, O.op[a], G.kind[a], G.arg_dtype[a], v[u], rows[u]);
      }
    }
  }
}

// Hash tables, step 1: compact (packed key, slot) of occupied slots.
__global__ __launch_bounds__(kBlock) void k_gb_compact_hash(GbTable T,
                                                            const int64_t* __restrict__ tile_off,
                                                            uint64_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals) {
  __shared__ uint32_t wcnt[kItems][kWaves];
  const int t = threadIdx.x, w = t >> 6, l = lane_id();
  const int64_t tile0 = (int64_t)blockIdx.x * kTileRows;
  bool f[kItems];
  uint64_t bal[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t s = tile0 + (int64_t)j * kBlock + t;
    f[j] = s < T.slots && T.cnt[s] > 0;
    bal[j] = ballot(f[j]);
    if (l == 0) wcnt[j][w] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  int64_t run = tile_off[blockIdx.x];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      const uint32_t c = wcnt[j][v];
      tot += c;
      before += (v < w) ? c : 0;
    }
    if (f[j]) {
      const int64_t s = tile0 + (int64_t)j * kBlock + t;
      const int64_t o = run + before + rank_below(bal[j]);
      keys[o] = T.hkey[s];
      vals[o] = (uint32_t)s;
    }
    run += tot;
  }
}

// Hash tables, step 3: write groups in sorted packed-key order.
__global__ __launch_bounds__(kBlock) void k_gb_write_sorted(GbSpec G, OutCols O, GbTable T,
                                                            const uint64_t* __restrict__ keys,
                                                            const uint32_t* __restrict__ vals,
                                                            int64_t ngroups) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < ngroups; i += stride)
    write_group(G, O, T, i, vals[i], keys[i]);
}

// ------------------------------------------------------------------ host side
struct GbWs {
  GbTable T;
  uint32_t* tile_counts;
  int64_t* tile_off;
  int64_t* ngroups;
  int64_t* scan_ws;
  uint64_t* k0;
  uint64_t* k1;
  uint32_t* v0;
  uint32_t* v1;
  void* stats;
  void* radix;
  // dense-LDS mode: per-workgroup tables, combined by k_gb_merge
  uint64_t* part;             // part[block][slot][count, aggregates...]
  int32_t pblocks;            // workgroups of the accumulate launch
  // partitioned dense mode
  int32_t part_bits, nparts;  // partitions of 2^part_bits slots
  int32_t tsub;               // 1024-row sub-tiles per scatter tile
  int64_t ptiles;             // scatter tiles over the input rows
  uint32_t* hist;             // rows of tile t in partition b at hist[b * ptiles + t]
  uint32_t* hoff;             // exclusive scan of hist: first scattered position of (b, t)
  uint32_t* htotal;
  uint32_t* hscan_ws;
  uint64_t* rec;              // per scattered row: {slot within the partition, arguments...}
  int64_t rows_cap;
  uint64_t* flags;            // fused reduce: look-back words [nparts], error word
  uint32_t* grp_cnt;          // groups per partition
  int64_t* grp_off;           // exclusive scan of grp_cnt
  int64_t* grp_scan_ws;
  uint32_t* gslot;            // partition b's groups in slot order at [b << part_bits, ...)
  uint32_t* gcnt;
  uint64_t* gacc;             // gacc[a * (nparts << part_bits) + i]
  size_t bytes;
};

static int ilog2_ceil(int64_t v) {
  int l = 0;
  while (((int64_t)1 << l) < v) ++l;
  return l;
}

// Partitions of the partitioned dense mode: the largest power-of-two slot count whose LDS table
// (u32 count + one u64 per aggregate per slot) fits kPartLds.
constexpr size_t kPartLds = 48 * 1024;
constexpr int kPartMax = 2048;
constexpr int64_t kPartRows = 65536;  // average input rows per partition at most
static int part_bits_for(int naggs) {
  const size_t per = 4 + 8 * (size_t)(naggs > 0 ? naggs : 1);
  int b = 0;
  while (((size_t)2 << b) * per <= kPartLds && ((int64_t)2 << b) <= kTileRows) ++b;
  return b;  // 2^b <= 4096 slots: the reduce kernel ranks 16 slots per thread
}

// Sub-tiles per scatter tile: enough that the per-tile partition histogram (nparts x tiles)
// stays within the two-level scan's 2^24 elements.
static int part_tsub(int64_t nrows, int nparts) {
  const int64_t subs = ceil_div(nrows > 0 ? nrows : 1, kSubRows);
  int t = 1;
  while (ceil_div(subs, t) * nparts > ((int64_t)1 << 24)) t *= 2;
  return t;
}

// Workgroups of a dense-LDS accumulate launch: one per 1024-row sub-tile, at most 512, and few
// enough that their partial tables stay within 32 MB.
static int32_t dense_lds_grid(int64_t nrows, int64_t slots, int naggs) {
  int64_t g = ceil_div(nrows > 0 ? nrows : 1, kSubRows);
  // one round of resident workgroups (a dense-LDS table with its stage takes up to 80 KB: two
  // per CU), each over several sub-tiles, so the table set-up and the partial write are paid
  // once per workgroup and no second round trails
  if (g > 512) g = 512;
  const int64_t per = slots * (1 + naggs) * 8;
  const int64_t cap = ((int64_t)32 << 20) / (per > 0 ? per : 1);
  if (g > cap) g = cap;
  return (int32_t)(g < 1 ? 1 : g);
}

static GbWs carve_gb(void* base, int mode, int64_t slots, int naggs, int64_t rows_cap,
                     int64_t nrows) {
  const int bm = mode & 7;
  char* p = (char*)base;
  const int64_t nt = ceil_div(slots, kTileRows);
  GbWs w{};
  w.T.slots = slots;
  w.T.log2 = ilog2_ceil(slots);
  if (bm == TQ_GB_DENSE_PART) {
    const int na = naggs > 0 ? naggs : 1;
    w.part_bits = part_bits_for(naggs);
    w.nparts = (int32_t)ceil_div(slots, (int64_t)1 << w.part_bits);
    w.tsub = part_tsub(nrows, w.nparts);
    w.ptiles = ceil_div(nrows > 0 ? nrows : 1, (int64_t)kSubRows * w.tsub);
    const int64_t nh = (int64_t)w.nparts * w.ptiles;
    const int64_t gcap = (int64_t)w.nparts << w.part_bits;
    w.rows_cap = rows_cap > 0 ? rows_cap : 1;
    w.hist = (uint32_t*)p; p += align_up((size_t)nh * 4, 256);
    w.hoff = (uint32_t*)p; p += align_up((size_t)nh * 4, 256);
    w.htotal = (uint32_t*)p; p += 256;
    w.hscan_ws = (uint32_t*)p; p += tq_scan_ws_bytes(nh);
    w.rec = (uint64_t*)p; p += align_up((size_t)w.rows_cap * 8 * (1 + na), 256);
    w.flags = (uint64_t*)p; p += align_up((size_t)(w.nparts + 1) * 8, 256);
    w.grp_cnt = (uint32_t*)p; p += align_up((size_t)w.nparts * 4, 256);
    w.grp_off = (int64_t*)p; p += align_up((size_t)w.nparts * 8, 256);
    w.ngroups = (int64_t*)p; p += 256;
    w.grp_scan_ws = (int64_t*)p; p += tq_scan_ws_bytes(w.nparts);
    w.gslot = (uint32_t*)p; p += align_up((size_t)gcap * 4, 256);
    w.gcnt = (uint32_t*)p; p += align_up((size_t)gcap * 4, 256);
    w.gacc = (uint64_t*)p; p += align_up((size_t)gcap * 8 * na, 256);
    w.bytes = (size_t)(p - (char*)base);
    return w;
  }
  if (bm == TQ_GB_DENSE_LDS) {
    w.pblocks = dense_lds_grid(nrows, slots, naggs);
    w.part = (uint64_t*)p;
    p += align_up((size_t)w.pblocks * slots * (1 + naggs) * 8, 256);
  }
  w.T.cnt = (uint32_t*)p; p += align_up((size_t)slots * 4, 256);
  w.T.hkey = (unsigned long long*)p; p += bm == TQ_GB_HASH ? align_up((size_t)slots * 8, 256) : 0;
  w.T.acc = (uint64_t*)p; p += align_up((size_t)slots * 8 * (naggs > 0 ? naggs : 1), 256);
  w.tile_counts = (uint32_t*)p; p += align_up((size_t)nt * 4, 256);
  w.tile_off = (int64_t*)p; p += align_up((size_t)nt * 8, 256);
  w.ngroups = (int64_t*)p; p += 256;
  w.scan_ws = (int64_t*)p; p += tq_scan_ws_bytes(nt);
  if (bm == TQ_GB_HASH) {
    w.k0 = (uint64_t*)p; p += align_up((size_t)slots * 8, 256);
    w.k1 = (uint64_t*)p; p += align_up((size_t)slots * 8, 256);
    w.v0 = (uint32_t*)p; p += align_up((size_t)slots * 4, 256);
    w.v1 = (uint32_t*)p; p += align_up((size_t)slots * 4, 256);
    w.stats = p; p += 256;
    w.radix = p; p += radix_ws_bytes(slots);
  }
  w.bytes = (size_t)(p - (char*)base);
  return w;
}

int agg_kind_init(int op, int dt, int32_t* kind, uint64_t* init) {
  const bool isf = dt == TQ_F64 || dt == TQ_F32;
  *init = 0;
  switch (op) {
    case TQ_AGG_SUM:
    case TQ_AGG_AVG: *kind = isf ? ACC_FSUM : ACC_ISUM; break;
    case TQ_AGG_COUNT: *kind = ACC_COUNT; break;
    case TQ_AGG_MIN:
      *kind = isf ? ACC_FMIN : ACC_IMIN;
      if (dt == TQ_F64) { double v = 1.7976931348623157e308; memcpy(init, &v, 8); }
      else if (dt == TQ_F32) { double v = 3.4028234663852886e38; memcpy(init, &v, 8); }
      else *init = (uint64_t)(dt == TQ_I8 ? 127 : dt == TQ_I16 ? 32767 : dt == TQ_I32 ? 2147483647
                              : dt == TQ_BOOL ? 1 : INT64_MAX);
      break;
    case TQ_AGG_MAX: *kind = isf ? ACC_FMAX : ACC_IMAX; break;
    default: return set_error(TQ_EUNSUPPORTED, "aggregate op %d", op);
  }
  return TQ_OK;
}

// A key dictionary over a value range of `range` values: presence bits and exclusive ranks per
// 32-value word, then the distinct values (room for all of them).
struct DictLayout {
  size_t rank_off, val_off, bytes;
};
static DictLayout dict_layout(int64_t range) {
  const size_t w = (size_t)ceil_div(range > 0 ? range : 1, 32);
  DictLayout L;
  L.rank_off = w * 4;
  L.val_off = align_up(w * 8, 8);
  L.bytes = align_up(L.val_off + (size_t)(range > 0 ? range : 1) * 8, 256);
  return L;
}

static int make_spec(const tq_groupby_desc_t* d, const tq_program_t* prog, GbSpec* G) {
  *G = GbSpec{};
  if (d->nkeys < 0 || d->nkeys > TQ_GB_MAX_KEYS) return set_error(TQ_EINVAL, "group-by: %d keys", d->nkeys);
  if (d->naggs < 0 || d->naggs > TQ_GB_MAX_AGGS) return set_error(TQ_EINVAL, "group-by: %d aggregates", d->naggs);
  G->nkeys = d->nkeys;
  G->naggs = d->naggs;
  for (int k = 0; k < d->nkeys; ++k) {
    G->key_slot[k] = k;  // key i is load slot i of the program
    G->key_min[k] = d->key_min[k];
    G->key_stride[k] = d->key_stride[k] > 0 ? d->key_stride[k] : 1;
    G->key_range[k] = d->key_range[k] > 0 ? d->key_range[k] : 1;
    G->key_dtype[k] = d->key_dtype[k];
    if (d->key_dict[k]) {
      const DictLayout L = dict_layout(d->key_dict_range[k]);
      const char* base = (const char*)d->key_dict[k];
      G->dict_bits[k] = (const uint32_t*)base;
      G->dict_rank[k] = (const uint32_t*)(base + L.rank_off);
      G->dict_val[k] = (const int64_t*)(base + L.val_off);
    }
    if (d->key_dtype[k] < TQ_BOOL || d->key_dtype[k] > TQ_I64)
      return set_error(TQ_EUNSUPPORTED, "group-by key %d must be an integer column", k);
    if (prog && (prog->nload <= k || prog->load_dtype[k] != d->key_dtype[k]))
      return set_error(TQ_EINVAL, "group-by: load slot %d must hold key %d", k, k);
  }
  for (int a = 0; a < d->naggs; ++a) {
    G->arg_dtype[a] = d->agg_dtype[a];
    TQ_TRY(agg_kind_init(d->agg_op[a], d->agg_dtype[a], &G->kind[a], &G->init[a]));
  }
  if (prog && prog->nseg < d->naggs) return set_error(TQ_EINVAL, "group-by: program has %d segments for %d aggregates", prog->nseg, d->naggs);
  if (prog) match_direct(*prog, G->naggs, G->kind, G->arg_dtype, &G->D);
  return TQ_OK;
}

// Launch with `lds` bytes of dynamic LDS; above 64 KB the kernel is first allowed that much
// (gfx950 gives a workgroup up to 160 KB).
// The allowance is set once per kernel (per device: a kernel's attribute is per device, and the
// cache below is keyed by both), raised when a larger table is asked for.
template <typename... P, typename... A>
static int launch_lds(const char* name, void (*kernel)(P...), unsigned grid, size_t lds,
                      hipStream_t s, A&&... a) {
  if (lds > 64 * 1024) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, size_t> allowed;
    int dev = 0;
    TQ_TRY_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    size_t& have = allowed[{dev, (const void*)kernel}];
    if (have < lds) {
      const hipError_t e = hipFuncSetAttribute(
          (const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig);
      if (e != hipSuccess)
        return set_error(TQ_EHIP, "%s: %zu bytes of LDS: %s", name, lds, hipGetErrorString(e));
      have = kLdsBig;
    }
  }
  return launch(name, kernel, grid, kBlock, lds, s, std::forward<A>(a)...);
}

static size_t lds_stage_bytes(int32_t nload) { return (size_t)(nload > 0 ? nload : 1) * kSubRows * 8; }

static int32_t lds_cap(int mode, int32_t nload, int naggs, int64_t slots) {
  if ((mode & 7) == TQ_GB_DENSE_LDS) return (int32_t)(slots * ((mode & TQ_GB_LDS_LANES) ? 64 : 1));
  if (!(mode & TQ_GB_LDS_FRONT)) return 0;
  const size_t avail = kLdsBig - lds_stage_bytes(nload);
  const size_t per = 12 + 8 * (size_t)naggs;
  int32_t cap = 1;
  while ((size_t)cap * 2 * per <= avail && cap < 8192) cap *= 2;
  return cap;
}

// LDS bytes of an accumulate launch.
static size_t accum_lds(int mode, int32_t nload, int naggs, int32_t cap) {
  size_t lds = lds_stage_bytes(nload);
  if ((mode & 7) == TQ_GB_DENSE_LDS) lds += (size_t)cap * (8 * naggs + 4);
  else if (mode & TQ_GB_LDS_FRONT) lds += (size_t)cap * (8 * naggs + 12);
  return lds;
}

}  // namespace tq

using namespace tq;

extern "C" size_t tq_groupby_bounds_ws_bytes(int64_t n) {
  return align_up((size_t)ceil_div(n > 0 ? n : 1, kTileRows) * kBoundWords * sizeof(int64_t), 256);
}

extern "C" int tq_groupby_bounds(const tq_groupby_desc_t* d, const uint8_t* mask, int64_t n,
                                 int64_t* out, void* ws, void* stream) {
  if (!d || d->nkeys < 0 || d->nkeys > TQ_GB_MAX_KEYS)
    return set_error(TQ_EINVAL, "group-by bounds: bad descriptor");
  if (!out || (n > 0 && !ws)) return set_error(TQ_EINVAL, "group-by bounds: null argument");
  hipStream_t s = (hipStream_t)stream;
  KeyCols K{};
  K.nkeys = d->nkeys;
  for (int k = 0; k < d->nkeys; ++k) {
    K.ptr[k] = d->key_ptr[k];
    K.dtype[k] = d->key_dtype[k];
    if (K.dtype[k] < TQ_BOOL || K.dtype[k] > TQ_I64)
      return set_error(TQ_EUNSUPPORTED, "group-by key %d must be an integer column", k);
  }
  const int64_t nt = n > 0 ? ceil_div(n, kTileRows) : 0;
  if (nt > 0)
    TQ_TRY(launch("k_gb_bounds", k_gb_bounds, (unsigned)nt, kBlock, 0, s, K, mask, n,
                  (int64_t*)ws));
  return launch("k_gb_bounds_final", k_gb_bounds_final, 1, kBlock, 0, s, (const int64_t*)ws, nt,
                (int)d->nkeys, out);
}

// ------------------------------------------------------------------ key dictionaries
struct DictPtrs {
  uint32_t* bits[TQ_GB_MAX_KEYS];
  uint32_t* rank[TQ_GB_MAX_KEYS];
  int64_t* val[TQ_GB_MAX_KEYS];
  int64_t lo[TQ_GB_MAX_KEYS];
  int32_t words[TQ_GB_MAX_KEYS];
  int32_t lds_off[TQ_GB_MAX_KEYS + 1];  // word offsets of the keys' bitmaps in LDS; [nkeys]: total
  int64_t* ndistinct;                     // [nkeys]
};
constexpr int kDictLdsWords = 16384;  // 64 KB of LDS bitmaps per workgroup

// Presence bits of every key's selected values: each workgroup marks a grid-stride share of the
// 4096-row tiles in LDS copies of the bitmaps, then ORs its non-zero words into the global ones.
// Selected values lie in [lo, lo + 32 * words) (the bounds were taken over the same rows).
__global__ __launch_bounds__(kBlock) void k_gb_presence(KeyCols K, const uint8_t* __restrict__ mask,
                                                        int64_t n, DictPtrs D) {
  extern __shared__ uint32_t lb[];
  const int t = threadIdx.x;
  const int total = D.lds_off[K.nkeys];
  for (int i = t; i < total; i += kBlock) lb[i] = 0;
  __syncthreads();
  const int64_t nt = ceil_div(n, kTileRows);
  for (int64_t tile = blockIdx.x; tile < nt; tile += gridDim.x) {
    const int64_t tile0 = tile * kTileRows;
    const uint32_t sel = load_sel16(mask, tile0, n);
    for (int k = 0; k < K.nkeys; ++k) {
      uint64_t v[kItems];
      load_raw16(K.ptr[k], K.dtype[k], tile0, n, v);
      uint32_t* bk = lb + D.lds_off[k];
#pragma unroll
      for (int j = 0; j < kItems; ++j) {
        if ((sel >> j) & 1u) {
          const uint64_t off = (uint64_t)((int64_t)v[j] - D.lo[k]);
          atomicOr(&bk[off >> 5], 1u << (off & 31));
        }
      }
    }
  }
  __syncthreads();
  for (int k = 0; k < K.nkeys; ++k) {
    const int w0 = D.lds_off[k], nw = D.lds_off[k + 1] - w0;
    for (int i = t; i < nw; i += kBlock) {
      const uint32_t w = lb[w0 + i];
      if (w) atomicOr(&D.bits[k][i], w);
    }
  }
}

// One workgroup per key: exclusive ranks of the presence words and the distinct values listed
// in ascending order; the count goes to ndistinct[k].
__global__ __launch_bounds__(kBlock) void k_gb_dict_rank(DictPtrs D) {
  __shared__ uint32_t sh[kWaves];
  const int k = blockIdx.x, t = threadIdx.x;
  const int nw = D.words[k];
  uint32_t run = 0;
  for (int w0 = 0; w0 < nw; w0 += kBlock) {
    const int i = w0 + t;
    const uint32_t bits = i < nw ? D.bits[k][i] : 0u;
    uint32_t tot;
    const uint32_t r = run + block_exclusive_sum((uint32_t)__popc(bits), sh, &tot);
    if (i < nw) {
      D.rank[k][i] = r;
      uint32_t b = bits, o = r;
      while (b) {
        const int bit = __ffs(b) - 1;
        D.val[k][o++] = D.lo[k] + (int64_t)i * 32 + bit;
        b &= b - 1;
      }
    }
    run += tot;
  }
  if (t == 0) D.ndistinct[k] = run;
}

extern "C" size_t tq_groupby_dict_ws_bytes(const tq_groupby_desc_t* d) {
  if (!d || d->nkeys < 0 || d->nkeys > TQ_GB_MAX_KEYS) return 0;
  size_t b = 256;  // the distinct counts
  for (int k = 0; k < d->nkeys; ++k) b += dict_layout(d->key_range[k]).bytes;
  return b;
}

extern "C" int tq_groupby_dict(const tq_groupby_desc_t* d, const uint8_t* mask, int64_t n,
                               void* ws, size_t ws_bytes, void** dict_out, int64_t* ndistinct_host,
                               void* stream) {
  if (!d || d->nkeys < 1 || d->nkeys > TQ_GB_MAX_KEYS || !dict_out || !ndistinct_host)
    return set_error(TQ_EINVAL, "group-by dictionary: bad argument");
  if (ws_bytes < tq_groupby_dict_ws_bytes(d))
    return set_error(TQ_ECAPACITY, "group-by dictionary: workspace too small");
  hipStream_t s = (hipStream_t)stream;
  KeyCols K{};
  K.nkeys = d->nkeys;
  DictPtrs D{};
  char* p = (char*)ws;
  D.ndistinct = (int64_t*)p;
  p += 256;
  int lds = 0;
  for (int k = 0; k < d->nkeys; ++k) {
    const int64_t range = d->key_range[k];
    if (range < 1 || range > TQ_GB_DICT_MAX_RANGE)
      return set_error(TQ_EINVAL, "group-by dictionary: key %d spans %lld values", k,
                       (long long)range);
    if (d->key_dtype[k] < TQ_BOOL || d->key_dtype[k] > TQ_I64)
      return set_error(TQ_EUNSUPPORTED, "group-by key %d must be an integer column", k);
    K.ptr[k] = d->key_ptr[k];
    K.dtype[k] = d->key_dtype[k];
    const DictLayout L = dict_layout(range);
    D.bits[k] = (uint32_t*)p;
    D.rank[k] = (uint32_t*)(p + L.rank_off);
    D.val[k] = (int64_t*)(p + L.val_off);
    D.lo[k] = d->key_min[k];
    D.words[k] = (int32_t)ceil_div(range, 32);
    D.lds_off[k] = lds;
    lds += D.words[k];
    dict_out[k] = (void*)p;
    p += L.bytes;
  }
  D.lds_off[d->nkeys] = lds;
  if (lds > kDictLdsWords)
    return set_error(TQ_EINVAL, "group-by dictionary: %d bitmap words exceed LDS", lds);
  TQ_TRY(fill_u64(ws, 0, (int64_t)((p - (char*)ws) / 8), s));
  if (n > 0) {
    const int64_t nt = ceil_div(n, kTileRows);
    TQ_TRY(launch("k_gb_presence", k_gb_presence, (unsigned)(nt < 1024 ? nt : 1024), kBlock,
                  (size_t)lds * 4, s, K, mask, n, D));
    TQ_TRY(launch("k_gb_dict_rank", k_gb_dict_rank, (unsigned)d->nkeys, kBlock, 0, s, D));
  }
  return read_back(D.ndistinct, (size_t)d->nkeys * 8, ndistinct_host, s);
}

extern "C" int tq_groupby_plan(tq_groupby_desc_t* d, int64_t nsel, int32_t nload, int32_t* mode,
                               int64_t* slots, size_t* ws_bytes) {
  if (!d || d->nkeys < 0 || d->nkeys > TQ_GB_MAX_KEYS || d->naggs < 0 || d->naggs > TQ_GB_MAX_AGGS)
    return set_error(TQ_EINVAL, "group-by plan: bad descriptor");
  // key space = product of ranges (a range <= 0 means no rows: one empty slot)
  __int128 space = 1;
  for (int k = d->nkeys - 1; k >= 0; --k) {
    if (d->key_range[k] <= 0) d->key_range[k] = 1;
    d->key_stride[k] = (int64_t)space;
    space *= d->key_range[k];
    if (space > ((__int128)1 << 62))
      return set_error(TQ_EUNSUPPORTED, "group-by key space exceeds 62 bits");
  }
  const int64_t S = (int64_t)space;
  const int64_t ns = nsel > 0 ? nsel : 0;
  d->nsel = ns;
  const size_t stage = lds_stage_bytes(nload);
  const size_t dense_per = 4 + 8 * (size_t)d->naggs;
  // a key space shrunk by dictionaries may use up to half of a CU's LDS per workgroup (two
  // workgroups per CU), where the dense-LDS table otherwise stays within kLdsBudget
  bool dict = false;
  for (int k = 0; k < d->nkeys; ++k) dict = dict || d->key_dict[k] != nullptr;
  const size_t dense_budget = dict ? (size_t)80 * 1024 : (size_t)kLdsBudget;
  int32_t m;
  int64_t sl;
  if (S <= 64 && (size_t)S * 64 * dense_per + stage <= (size_t)kLdsBig) {
    // a handful of groups: one table copy per lane keeps their rows off shared LDS addresses
    m = TQ_GB_DENSE_LDS | TQ_GB_LDS_LANES;
    sl = S;
  } else if ((size_t)S * dense_per + stage <= (size_t)dense_budget) {
    m = TQ_GB_DENSE_LDS;
    sl = S;
  } else {
    const int64_t distinct = ns < S ? ns : S;
    const int64_t nparts = ceil_div(S, (int64_t)1 << part_bits_for(d->naggs));
    const bool dense_ok = S <= ((int64_t)1 << 27) && S <= (4 * ns > 65536 ? 4 * ns : 65536);
    // partitions are reduced one workgroup each, so each must get a modest share of the rows
    if (dense_ok && nparts <= kPartMax && stage + (size_t)nparts * 4 <= (size_t)kLdsBudget &&
        d->nrows > 0 && d->nrows < ((int64_t)1 << 32) && d->nrows <= nparts * kPartRows) {
      m = TQ_GB_DENSE_PART;
      sl = S;
    } else if (dense_ok) {
      m = TQ_GB_DENSE_GLOBAL;
      sl = S;
    } else {
      m = TQ_GB_HASH;
      sl = (int64_t)1 << ilog2_ceil(2 * (distinct > 2048 ? distinct : 2048));
    }
    const size_t per = 12 + 8 * (size_t)d->naggs;
    const bool room = stage + 1024 * per <= (size_t)kLdsBig;
    if (m != TQ_GB_DENSE_PART && room && distinct > 0 && ns >= 4 * distinct) m |= TQ_GB_LDS_FRONT;
  }
  *mode = m;
  *slots = sl;
  *ws_bytes = carve_gb(nullptr, m, sl, d->naggs, ns, d->nrows).bytes;
  return TQ_OK;
}

// Accumulate; with fused (partitioned mode only), the partition reduce writes the final groups
// into *fused through the look-back words.
static int accumulate_impl(const tq_groupby_desc_t* d, const tq_program_t* prog,
                           const tq_colset_t* cols, const uint8_t* mask, int64_t n, int32_t mode,
                           int64_t slots, GbWs& w, hipStream_t s, const OutCols* fused,
                           LookBack lb, const int64_t* n_dev = nullptr) {
  GbSpec G;
  TQ_TRY(make_spec(d, prog, &G));
  if (prog->nload > TQ_MAX_LOADS || prog->n_instr > TQ_MAX_INSTR) return set_error(TQ_EINVAL, "group-by: malformed program");
  const int base = mode & 7;
  const bool front = (mode & TQ_GB_LDS_FRONT) != 0;
  if (base == TQ_GB_DENSE_PART) {
    const int np = w.nparts;
    const int64_t nh = (int64_t)np * w.ptiles;
    if (n <= 0 || d->nsel <= 0) return fill_u32(w.grp_cnt, 0u, np, s);
    if (n != d->nrows) return set_error(TQ_EINVAL, "group-by: %lld rows, planned for %lld",
                                        (long long)n, (long long)d->nrows);
    KeyCols K{};
    K.nkeys = d->nkeys;
    for (int k = 0; k < d->nkeys; ++k) {
      K.ptr[k] = d->key_ptr[k];
      K.dtype[k] = d->key_dtype[k];
    }
    const int32_t stage_words = (int32_t)(lds_stage_bytes(prog->nload) / 8);
    const size_t lds_scatter = lds_stage_bytes(prog->nload) + (size_t)np * 4;
    const size_t lds_reduce = ((size_t)1 << w.part_bits) * (4 + 8 * (size_t)d->naggs);
    if (lds_scatter > (size_t)kLdsBudget || lds_reduce > (size_t)kLdsBudget)
      return set_error(TQ_ECAPACITY, "group-by: partition tables exceed LDS");
    TQ_TRY(launch("k_gbp_hist", k_gbp_hist, (unsigned)w.ptiles, kBlock, (size_t)np * 4, s, K, G,
                  mask, n, n_dev, w.part_bits, np, w.tsub, w.ptiles, w.hist,
                  fused ? w.flags : (uint64_t*)nullptr));
    // with a device row count, only the histogram's first np * (tiles holding rows) entries
    ScanLen hl{};
    if (n_dev) hl = ScanLen{n_dev, n, (int64_t)w.tsub * kSubRows, (int64_t)np};
    TQ_TRY((scan_exclusive<uint32_t, uint32_t>(w.hist, nh, w.hoff, w.htotal, w.hscan_ws, s, hl)));
    TQ_TRY(launch("k_gbp_scatter", k_gbp_scatter, (unsigned)w.ptiles, kBlock, lds_scatter, s, G,
                  *prog, *cols, mask, n, n_dev, w.part_bits, np, w.tsub, w.ptiles,
                  (const uint32_t*)w.hoff, w.rec, stage_words));
    if (fused)
      return launch("k_gbp_reduce", k_gbp_reduce<true>, (unsigned)np, kBlock, lds_reduce, s, G,
                    w.part_bits, np, w.ptiles, n, n_dev, w.tsub, (const uint32_t*)w.hoff,
                    (const uint32_t*)w.htotal, (const uint64_t*)w.rec, w.grp_cnt, w.gslot, w.gcnt,
                    w.gacc, *fused, lb);
    return launch("k_gbp_reduce", k_gbp_reduce<false>, (unsigned)np, kBlock, lds_reduce, s, G,
                  w.part_bits, np, w.ptiles, n, n_dev, w.tsub, (const uint32_t*)w.hoff, (const uint32_t*)w.htotal,
                  (const uint64_t*)w.rec, w.grp_cnt, w.gslot, w.gcnt, w.gacc, OutCols{}, lb);
  }
  const bool dense_lds = base == TQ_GB_DENSE_LDS;
  const bool lanes = dense_lds && (mode & TQ_GB_LDS_LANES) != 0;
  // a dense-LDS table is written whole by the merge of the workgroups' partials
  if (!dense_lds || n <= 0)
    TQ_TRY(launch("k_gb_clear", k_gb_clear, stream_grid(slots), kBlock, 0, s, w.T, G, base == TQ_GB_HASH));
  if (n <= 0) return TQ_OK;
  const int32_t cap = lds_cap(mode, prog->nload, d->naggs, slots);
  const int32_t stage_words = (int32_t)(lds_stage_bytes(prog->nload) / 8);
  const size_t lds = accum_lds(mode, prog->nload, d->naggs, cap);
  if (lds > (size_t)kLdsBig) return set_error(TQ_ECAPACITY, "group-by: LDS table too large");
  if (dense_lds && n != d->nrows)
    return set_error(TQ_EINVAL, "group-by: %lld rows, planned for %lld", (long long)n,
                     (long long)d->nrows);
  // global and hash tables: one resident round of workgroups striding over the sub-tiles
  const int64_t g = ceil_div(n, kSubRows);
#define TQ_ACCUM(BASE, FRONT, LANES, CAP)                                                     \
  launch_lds("k_gb_accum", (k_gb_accum<BASE, FRONT, LANES>),                                 \
             dense_lds ? (unsigned)w.pblocks                                                  \
                       : resident_grid(k_gb_accum<BASE, FRONT, LANES>, lds, g),               \
             lds, s, G, *prog, *cols, mask, n, w.T, CAP, stage_words, w.part)
  if (dense_lds) {
    TQ_TRY(lanes ? TQ_ACCUM(TQ_GB_DENSE_LDS, false, true, cap)
                 : TQ_ACCUM(TQ_GB_DENSE_LDS, false, false, cap));
    int64_t mg = ceil_div(slots * (1 + d->naggs), kWaves);
    if (mg > 4096) mg = 4096;
    return launch("k_gb_merge", k_gb_merge, (unsigned)mg, kBlock, 0, s, G,
                  (const uint64_t*)w.part, w.pblocks, w.T);
  }
  if (base == TQ_GB_DENSE_GLOBAL && front) return TQ_ACCUM(TQ_GB_DENSE_GLOBAL, true, false, cap);
  if (base == TQ_GB_DENSE_GLOBAL) return TQ_ACCUM(TQ_GB_DENSE_GLOBAL, false, false, 1);
  if (front) return TQ_ACCUM(TQ_GB_HASH, true, false, cap);
  return TQ_ACCUM(TQ_GB_HASH, false, false, 1);
#undef TQ_ACCUM
}

extern "C" int tq_groupby_accumulate(const tq_groupby_desc_t* d, const tq_program_t* prog,
                                     const tq_colset_t* cols, const uint8_t* mask, int64_t n,
                                     int32_t mode, int64_t slots, void* ws, size_t ws_bytes,
                                     void* stream) {
  if (!d || !prog || !cols) return set_error(TQ_EINVAL, "group-by accumulate: null argument");
  GbWs w = carve_gb(ws, mode, slots, d->naggs, d->nsel, d->nrows);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "group-by workspace too small");
  return accumulate_impl(d, prog, cols, mask, n, mode, slots, w, (hipStream_t)stream, nullptr,
                         LookBack{});
}

extern "C" int tq_groupby_count(const tq_groupby_desc_t* d, int32_t mode, int64_t slots, void* ws,
                                int64_t* ngroups_host, void* stream) {
  if (!d) return set_error(TQ_EINVAL, "group-by count: null descriptor");
  GbWs w = carve_gb(ws, mode, slots, d->naggs, d->nsel, d->nrows);
  hipStream_t s = (hipStream_t)stream;
  if ((mode & 7) == TQ_GB_DENSE_PART) {
    TQ_TRY((scan_exclusive<uint32_t, int64_t>(w.grp_cnt, w.nparts, w.grp_off, w.ngroups,
                                              w.grp_scan_ws, s)));
    if (ngroups_host) TQ_TRY(read_back(w.ngroups, 8, ngroups_host, s));
    return TQ_OK;
  }
  const int64_t nt = ceil_div(slots, kTileRows);
  TQ_TRY(launch("k_gb_occupied", k_gb_occupied, (unsigned)nt, kBlock, 0, s, w.T.cnt, slots, d->nkeys == 0, w.tile_counts));
  TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.ngroups, w.scan_ws, s));
  if (ngroups_host) TQ_TRY(read_back(w.ngroups, 8, ngroups_host, s));
  return TQ_OK;
}

extern "C" int tq_groupby_write(const tq_groupby_desc_t* d, int32_t mode, int64_t slots, void* ws,
                                int64_t ngroups, void* const* key_out, void* const* agg_out,
                                void* stream) {
  if (!d) return set_error(TQ_EINVAL, "group-by write: null descriptor");
  GbSpec G;
  TQ_TRY(make_spec(d, nullptr, &G));
  GbWs w = carve_gb(ws, mode, slots, d->naggs, d->nsel, d->nrows);
  hipStream_t s = (hipStream_t)stream;
  if (ngroups <= 0) return TQ_OK;
  OutCols O{};
  for (int k = 0; k < d->nkeys; ++k) O.key[k] = key_out[k];
  for (int a = 0; a < d->naggs; ++a) {
    O.agg[a] = agg_out[a];
    O.op[a] = d->agg_op[a];
  }
  if ((mode & 7) == TQ_GB_DENSE_PART)
    return launch("k_gbp_write", k_gbp_write, (unsigned)w.nparts, kBlock, 0, s, G, O, w.part_bits,
                  w.nparts, (const uint32_t*)w.grp_cnt, (const int64_t*)w.grp_off,
                  (const uint32_t*)w.gslot, (const uint32_t*)w.gcnt, (const uint64_t*)w.gacc);
  const int64_t nt = ceil_div(slots, kTileRows);
  if ((mode & 7) != TQ_GB_HASH) {
    return launch("k_gb_write_dense", k_gb_write_dense, (unsigned)nt, kBlock, 0, s, G, O, w.T, w.tile_off, d->nkeys == 0);
  }
  TQ_TRY(launch("k_gb_compact_hash", k_gb_compact_hash, (unsigned)nt, kBlock, 0, s, w.T, w.tile_off, w.k0, w.v0));
  int lo, hi;
  TQ_TRY(key_bit_range<uint64_t>(w.k0, ngroups, w.stats, s, &lo, &hi));
  bool alt;
  TQ_TRY(radix_sort_pairs<uint64_t>(w.k0, w.v0, w.k1, w.v1, ngroups, lo, hi, w.radix, s, &alt));
  return launch("k_gb_write_sorted", k_gb_write_sorted, stream_grid(ngroups), kBlock, 0, s, G, O, w.T, alt ? w.k1 : w.k0, alt ? w.v1 : w.v0, ngroups);
}

extern "C" int tq_groupby_run_deferred(const tq_groupby_desc_t* d, const tq_program_t* prog,
                                       const tq_colset_t* cols, const uint8_t* mask, int64_t n,
                                       const int64_t* n_dev, int32_t mode, int64_t slots,
                                       void* ws, size_t ws_bytes, void* const* key_out,
                                       void* const* agg_out, int64_t out_cap,
                                       const int64_t** ngroups_dev, const int64_t** err_dev,
                                       void* stream) {
  if (!d || !prog || !cols || !n_dev || !ngroups_dev || !err_dev)
    return set_error(TQ_EINVAL, "group-by run (deferred): null argument");
  if ((mode & 7) != TQ_GB_DENSE_PART || n <= 0 || d->nsel <= 0 || n != d->nrows)
    return set_error(TQ_EINVAL, "group-by run (deferred): needs the partitioned mode planned for "
                                "the %lld-row capacity", (long long)n);
  GbWs w = carve_gb(ws, mode, slots, d->naggs, d->nsel, d->nrows);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "group-by workspace too small");
  OutCols O{};
  for (int k = 0; k < d->nkeys; ++k) O.key[k] = key_out[k];
  for (int a = 0; a < d->naggs; ++a) {
    O.agg[a] = agg_out[a];
    O.op[a] = d->agg_op[a];
  }
  const LookBack lb{w.flags, w.ngroups, out_cap, 1u << 20};
  TQ_TRY(accumulate_impl(d, prog, cols, mask, n, mode, slots, w, (hipStream_t)stream, &O, lb,
                         n_dev));
  *ngroups_dev = w.ngroups;
  *err_dev = (const int64_t*)(w.flags + w.nparts);
  return TQ_OK;
}

extern "C" int tq_groupby_run(const tq_groupby_desc_t* d, const tq_program_t* prog,
                              const tq_colset_t* cols, const uint8_t* mask, int64_t n,
                              int32_t mode, int64_t slots, void* ws, size_t ws_bytes,
                              void* const* key_out, void* const* agg_out, int64_t out_cap,
                              int64_t* ngroups_host, void* stream) {
  if (!d || !prog || !cols || !ngroups_host)
    return set_error(TQ_EINVAL, "group-by run: null argument");
  GbWs w = carve_gb(ws, mode, slots, d->naggs, d->nsel, d->nrows);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "group-by workspace too small");
  hipStream_t s = (hipStream_t)stream;
  if ((mode & 7) == TQ_GB_DENSE_PART && n > 0 && d->nsel > 0) {
    OutCols O{};
    for (int k = 0; k < d->nkeys; ++k) O.key[k] = key_out[k];
    for (int a = 0; a < d->naggs; ++a) {
      O.agg[a] = agg_out[a];
      O.op[a] = d->agg_op[a];
    }
    const LookBack lb{w.flags, w.ngroups, out_cap, 1u << 20};
    TQ_TRY(accumulate_impl(d, prog, cols, mask, n, mode, slots, w, s, &O, lb));
    int64_t g = 0, err = 0;
    TQ_TRY(read_back2(w.ngroups, 8, &g, w.flags + w.nparts, 8, &err, s));
    if (!err) {
      *ngroups_host = g;
      return TQ_OK;
    }
    // the look-back gave up: reduce again without it (the scattered records are intact)
    GbSpec G;
    TQ_TRY(make_spec(d, prog, &G));
    const size_t lds_reduce = ((size_t)1 << w.part_bits) * (4 + 8 * (size_t)d->naggs);
    TQ_TRY(launch("k_gbp_reduce", k_gbp_reduce<false>, (unsigned)w.nparts, kBlock, lds_reduce, s,
                  G, w.part_bits, w.nparts, w.ptiles, n, (const int64_t*)nullptr, w.tsub,
                  (const uint32_t*)w.hoff, (const uint32_t*)w.htotal, (const uint64_t*)w.rec,
                  w.grp_cnt, w.gslot, w.gcnt, w.gacc, OutCols{}, LookBack{}));
  } else {
    TQ_TRY(accumulate_impl(d, prog, cols, mask, n, mode, slots, w, s, nullptr, LookBack{}));
  }
  int64_t g = 0;
  TQ_TRY(tq_groupby_count(d, mode, slots, ws, &g, stream));
  if (g > out_cap) return set_error(TQ_ECAPACITY, "group-by: %lld groups, outputs hold %lld",
                                    (long long)g, (long long)out_cap);
  *ngroups_host = g;
  return tq_groupby_write(d, mode, slots, ws, g, key_out, agg_out, stream);
}
