// Equi-join fused with the group-by of its output on the join key ("groupjoin").
//
// When a join's pairs only feed an aggregate that groups by the join key, and the build keys are
// unique, every group is one build key and every pair one probe row: the aggregate is a
// reduction of probe-row values into the build keys' slots.  Three launches, no pair columns, no
// global atomics on the row path:
//
//  1. build: each selected build row stamps its key's byte, pres[key - lo] = epoch.  The bytes
//     are never cleared: a call's epoch (1..255) differs from every stamp left since the byte
//     array was last zeroed, so only this call's keys read as present.  Per 4096-row tile the
//     selected rows are counted (and a key outside [lo, hi] is flagged in the count's top bit).
//  2. probe: a workgroup takes a chunk of 8192 probe rows.  It loads every key and mask byte of
//     the chunk, then tests every selected row's presence byte (all loads in flight before any
//     is used), counts the matches per key slice (slices of 2^sb consecutive keys) and lists
//     them in LDS; the slice counts' exclusive scan places the chunk's records, slice by slice,
//     in the chunk's own region of the record array (no global scan, no atomics).  Then the
//     matched rows' aggregate arguments are gathered and evaluated by the expression program,
//     and each match becomes one record {slot, argument values...}.
//  3. reduce: a workgroup per key slice gathers its records from every chunk (the chunks' slice
//     offsets are a row of the offset table), reduces them into an LDS table, compacts the
//     occupied slots in key order and writes the final key and aggregate columns at the offset
//     it gets from the group counts of the slices before it (decoupled look-back with a
//     wall-clock bound; a timeout sets the error word and the host takes the per-operator
//     path).  Each slice also counts its keys' stamps, so the last slice, which reads every
//     slice's look-back word, holds the distinct build keys and the pairs for the uniqueness
//     check.
//
// Groups come out in ascending key order.  f64 sums are added in LDS in varying order.
#include <map>
#include <mutex>

#include "tq_agg.h"
#include "tq_common.h"
#include "tq_conj.h"
#include "tq_expr.h"

namespace tq {

constexpr int kGjTiles = 2;                     // probe tiles per chunk
constexpr int kGjChunk = kGjTiles * kTileRows;  // probe rows (and record capacity) per chunk
constexpr int kGjMaxSlices = 2048;
constexpr int kGjBatch = 1024;                  // chunks gathered per reduce round
constexpr size_t kGjTableLds = 48 * 1024;       // reduce table: count + aggregates per slot

// Words the host reads back (int64, in the workspace): groups, pairs (matched probe rows),
// selected build rows, distinct selected build keys, error bits.
enum { GJ_GROUPS = 0, GJ_PAIRS, GJ_NSEL, GJ_NSET, GJ_ERR, GJ_NWORDS = 8 };
enum { GJ_ERR_BOUNDS = 1, GJ_ERR_WAIT = 2, GJ_ERR_CAP = 4 };
constexpr uint32_t kOobBit = 0x80000000u;       // in a build tile count: a key outside [lo, hi]

// A slice's look-back word: bit 0 published, then its groups, its stamped keys (each at most
// 2^sb <= 4096 < 2^14) and its records (the pairs whose key lies in the slice).
__host__ __device__ inline uint64_t lb_word(uint64_t groups, uint64_t keys, uint64_t recs) {
  return 1ull | (groups << 1) | (keys << 15) | (recs << 29);
}
__device__ __forceinline__ uint64_t lb_groups(uint64_t w) { return (w >> 1) & 0x3FFF; }
__device__ __forceinline__ uint64_t lb_keys(uint64_t w) { return (w >> 15) & 0x3FFF; }
__device__ __forceinline__ uint64_t lb_recs(uint64_t w) { return w >> 29; }

struct GjSpec {
  int32_t naggs;
  int32_t kind[TQ_GB_MAX_AGGS];
  int32_t arg_dtype[TQ_GB_MAX_AGGS];
  int32_t op[TQ_GB_MAX_AGGS];
  uint64_t init[TQ_GB_MAX_AGGS];
  DirectArgs D;  // the arguments evaluated straight from the gathered loads, when they allow it
  // Accumulated aggregates: a count of a constant that is not the null sentinel equals the
  // slot's pair count, which the reduce keeps anyway, so it has no record word and no table
  // column (word[a] = -1); the others are record word / table column word[a] < nacc, and
  // acc_agg[w] is the aggregate of word w.
  int32_t nacc;
  int32_t word[TQ_GB_MAX_AGGS];
  int32_t acc_agg[TQ_GB_MAX_AGGS];
};

struct GjOut {
  void* key;
  int32_t key_dtype;
  void* agg[TQ_GB_MAX_AGGS];
  int64_t cap;
};

// ------------------------------------------------------------------ 1. build
// Workgroup 0 also zeroes the reduce's look-back words (flags[0..nflags)) and the words the
// host reads back (the later kernels of the call only add to or OR into them).  A fused filter
// (BC.nslot > 0) selects rows in the tile's own loads, and its mask is never written.
__global__ __launch_bounds__(kBlock) void k_gj_build(const void* __restrict__ bkeys, int dt,
                                                     const uint8_t* __restrict__ bmask, Conj BC,
                                                     int64_t nb, int64_t lo, uint64_t range,
                                                     uint8_t* __restrict__ pres, uint32_t epoch,
                                                     uint32_t* __restrict__ tile_counts,
                                                     uint64_t* __restrict__ flags, int nflags,
                                                     int64_t* __restrict__ words) {
  const int t = threadIdx.x;
  if (blockIdx.x == 0) {
    for (int i = t; i < nflags; i += kBlock) flags[i] = 0;
    if (t < GJ_NWORDS) words[t] = 0;
  }
  const int64_t tile = blockIdx.x;
  const int64_t tile0 = tile * kTileRows;
  if (tile0 >= nb) return;  // nb == 0: the launch only clears the words
  uint64_t raw[kItems];
  uint32_t sel;
  switch (tile0 + kTileRows <= nb
              ? head_case(dt, bkeys, bmask, BC.nslot, BC.slot_col[0], BC.slot_dtype[0]) : -1) {
#define TQ_HEAD(FW, M) \
  sel = tile_head<FW, M>(bkeys, bmask, BC.slot_col[0], BC, 0, BC.slot_dtype[0], tile0, raw)
    case 0: TQ_HEAD(0, false); break;
    case 1: TQ_HEAD(0, true); break;
    case 2: TQ_HEAD(1, false); break;
    case 3: TQ_HEAD(1, true); break;
    case 4: TQ_HEAD(2, false); break;
    case 5: TQ_HEAD(2, true); break;
    case 8: TQ_HEAD(4, false); break;
    case 9: TQ_HEAD(4, true); break;
#undef TQ_HEAD
    default:
      load_raw16(bkeys, dt, tile0, nb, raw);
      sel = load_sel16(bmask, tile0, nb);
      if (BC.nslot > 0) sel = conj_sel16(BC, tile0, nb, sel);
      break;
  }
  bool oob = false;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if (!((sel >> j) & 1u)) continue;
    const uint64_t off = raw[j] - (uint64_t)lo;  // below lo wraps to a huge offset
    if (off < range)
      pres[off] = (uint8_t)epoch;
    else
      oob = true;
  }
  __shared__ uint32_t wc[kWaves];
  const uint32_t c = wave_sum((uint32_t)__popc(sel)) | (ballot(oob) ? kOobBit : 0u);
  if (lane_id() == 0) wc[t >> 6] = c;
  __syncthreads();
  if (t == 0) {
    uint32_t n = 0, f = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) {
      n += wc[i] & ~kOobBit;
      f |= wc[i] & kOobBit;
    }
    tile_counts[tile] = n | f;
  }
}

// ------------------------------------------------------------------ 2. probe
// Dynamic LDS: stage (nload * kSubRows u64), hist[S] u32, the chunk's matches in slice order
// u32[kGjChunk] as (local row << sb) | slot within the slice.
// loff[c * (S + 1) + s] = first record of slice s in chunk c's region (s < S),
// loff[c * (S + 1) + S] = the chunk's matches: a chunk writes its own row, whole lines (a
// slice-major table had every line written by 32 chunks on different XCDs, each XCD writing
// back its partial line).  A record is 1 + naggs words: {slot within the slice, accumulator-domain
// values}; a chunk's records are written in slice order by consecutive threads.
__global__ __launch_bounds__(kBlock, 3) void k_gj_probe(
    GjSpec G, tq_program_t P, tq_colset_t C, const void* __restrict__ pkeys, int pdt,
    const uint8_t* __restrict__ pmask, Conj PC, int64_t np, int64_t lo, uint64_t range,
    const uint8_t* __restrict__ pres, uint32_t epoch, int sb, int S, int64_t W,
    uint32_t* __restrict__ loff, uint64_t* __restrict__ rec, int32_t stage_words) {
  extern __shared__ uint64_t smem[];
  uint64_t* stage = smem;
  uint32_t* hist = (uint32_t*)(smem + stage_words);
  uint32_t* srt = hist + ((S + 1) & ~1);
  __shared__ uint32_t sh[kWaves];
  const int t = threadIdx.x;
  const int64_t c = blockIdx.x;
  const int64_t base = c * kGjChunk;
  const uint32_t smask = (1u << sb) - 1u;
  for (int i = t; i < S; i += kBlock) hist[i] = 0;
  // every key, mask byte and filter value of the chunk, then every presence byte, before any
  // is used; a row's key offset is kNoKey when the row is unselected or its key is outside
  // [lo, hi] (offsets fit 32 bits: the range is at most kGjMaxSlices << 12 keys)
  uint32_t off[kGjTiles][kItems];
  uint32_t sel[kGjTiles];
  uint8_t b[kGjTiles][kItems];  // presence bytes
  const bool p32 = PC.nslot == 0 ||
                   (PC.nslot == 1 && (PC.slot_dtype[0] <= TQ_I32 || PC.slot_dtype[0] == TQ_F32));
  const int head = base + kGjChunk <= np
                       ? head_case(pdt, pkeys, pmask, PC.nslot, PC.slot_col[0], PC.slot_dtype[0])
                       : -1;
  if (head >= 0) {
#define TQ_HEAD(FW, M)                                                                  \
  probe_head<kGjTiles, FW, M>(pkeys, pmask, PC.slot_col[0], PC, 0, PC.slot_dtype[0], base, lo, \
                              range, pres, off, b, sel)
    switch (head) {
      case 0: TQ_HEAD(0, false); break;
      case 1: TQ_HEAD(0, true); break;
      case 2: TQ_HEAD(1, false); break;
      case 3: TQ_HEAD(1, true); break;
      case 4: TQ_HEAD(2, false); break;
      case 5: TQ_HEAD(2, true); break;
      case 8: TQ_HEAD(4, false); break;
      default: TQ_HEAD(4, true); break;
    }
#undef TQ_HEAD
  } else if (pdt <= TQ_I32 && p32) {
    // keys and the filter's column in 32 bits (fewer registers, all four loads in flight)
    uint32_t kv[kGjTiles][kItems];
#pragma unroll
    for (int q = 0; q < kGjTiles; ++q) {
      const int64_t tile0 = base + (int64_t)q * kTileRows;
      load_raw16_32(pkeys, pdt, tile0, np, kv[q]);  // rows past np read row 0 (unselected)
      sel[q] = load_sel16(pmask, tile0, np);
    }
    uint32_t pv[kGjTiles][kItems];
    if (PC.nslot == 1) {
#pragma unroll
      for (int q = 0; q < kGjTiles; ++q)
        load_raw16_32(PC.slot_col[0], PC.slot_dtype[0], base + (int64_t)q * kTileRows, np, pv[q]);
    }
    // a tile's presence lookups go out as soon as its own loads are in, the next tile's loads
    // still in flight
#pragma unroll
    for (int q = 0; q < kGjTiles; ++q) {
      if (PC.nslot == 1) sel[q] = conj_apply32(PC, 0, pv[q], sel[q]);
#pragma unroll
      for (int j = 0; j < kItems; ++j) {
        const uint64_t o = (uint64_t)(int64_t)(int32_t)kv[q][j] - (uint64_t)lo;
        off[q][j] = ((sel[q] >> j) & 1u) && o < range ? (uint32_t)o : kNoKey;
        b[q][j] = off[q][j] != kNoKey ? pres[off[q][j]] : (uint8_t)0;
      }
    }
  } else {
    // 8-byte keys or filter columns, or filters over several columns: a tile at a time
#pragma unroll
    for (int q = 0; q < kGjTiles; ++q) {
      const int64_t tile0 = base + (int64_t)q * kTileRows;
      uint64_t raw[kItems];
      load_raw16(pkeys, pdt, tile0, np, raw);
      sel[q] = load_sel16(pmask, tile0, np);
      if (PC.nslot > 0) sel[q] = conj_sel16(PC, tile0, np, sel[q]);
#pragma unroll
      for (int j = 0; j < kItems; ++j) {
        const uint64_t o = raw[j] - (uint64_t)lo;
        off[q][j] = ((sel[q] >> j) & 1u) && o < range ? (uint32_t)o : kNoKey;
        b[q][j] = off[q][j] != kNoKey ? pres[off[q][j]] : (uint8_t)0;
      }
    }
  }
  uint32_t hit[kGjTiles];
#pragma unroll
  for (int q = 0; q < kGjTiles; ++q) {
    hit[q] = 0;
#pragma unroll
    for (int j = 0; j < kItems; ++j) hit[q] |= (b[q][j] == (uint8_t)epoch ? 1u : 0u) << j;
  }
  __syncthreads();  // hist cleared
#pragma unroll
  for (int q = 0; q < kGjTiles; ++q)
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if ((hit[q] >> j) & 1u) atomicAdd(&hist[off[q][j] >> sb], 1u);
  __syncthreads();
  // exclusive scan of the slice counts (in place: hist becomes the slices' write cursors)
  uint32_t M;
  {
    constexpr int kPer = kGjMaxSlices / kBlock;
    uint32_t v[kPer], sum = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = t * kPer + k;
      v[k] = i < S ? hist[i] : 0u;
      sum += v[k];
    }
    uint32_t run = block_exclusive_sum(sum, sh, &M);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = t * kPer + k;
      if (i < S) {
        hist[i] = run;
        loff[c * (S + 1) + i] = run;
      }
      run += v[k];
    }
    if (t == 0) loff[c * (S + 1) + S] = M;
  }
  __syncthreads();
  // the matches in slice order
#pragma unroll
  for (int q = 0; q < kGjTiles; ++q)
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if ((hit[q] >> j) & 1u) {
        const uint32_t slot = off[q][j];
        const uint32_t lrow = (uint32_t)(q * kTileRows + (j >> 2) * (4 * kBlock) + 4 * t + (j & 3));
        srt[atomicAdd(&hist[slot >> sb], 1u)] = (lrow << sb) | (slot & smask);
      }
  __syncthreads();
  const int rw = 1 + G.nacc;
  uint64_t* crec = rec + base * rw;
  for (uint32_t e0 = 0; e0 < M; e0 += kSubRows) {
    bool live[kSub];
    int64_t rows[kSub];
    uint32_t slot[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      const uint32_t e = e0 + j * kBlock + t;
      live[j] = e < M;
      const uint32_t v = live[j] ? srt[e] : 0u;
      rows[j] = live[j] ? base + (v >> sb) : 0;
      slot[j] = v & smask;
    }
    gather_loads(P, C, rows, live, stage);
    RegsN R{};
    int begin = 0;
    for (int a = 0; a < G.naggs; ++a) {
      const int wd = G.word[a];
      if (wd < 0) continue;  // a count the reduce takes from the pair counts (direct form only)
      uint64_t va[kSub];
      if (G.D.direct) {
        direct_args(G.D, G.kind[a], a, stage, va);
      } else {
        const int end = P.seg_end[a];
        run_rows(P, C, begin, end, rows, live, stage, R);
        begin = end;
        const V4 V = R.get(P.seg_result[a]);
#pragma unroll
        for (int j = 0; j < kSub; ++j) va[j] = to_acc(G.kind[a], G.arg_dtype[a], V.v[j]);
      }
#pragma unroll
      for (int j = 0; j < kSub; ++j) {
        if (!live[j]) continue;
        const uint64_t v = va[j];
        const int64_t e = e0 + j * kBlock + t;
        if (G.nacc == 1)
          *(ulonglong2*)(crec + e * 2) = make_ulonglong2(slot[j], v);
        else
          crec[e * rw + 1 + wd] = v;
      }
    }
    if (G.nacc != 1) {
#pragma unroll
      for (int j = 0; j < kSub; ++j)
        if (live[j]) crec[(int64_t)(e0 + j * kBlock + t) * rw] = slot[j];
    }
  }
}

// ------------------------------------------------------------------ 3. reduce
__device__ __forceinline__ uint64_t gj_ld(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gj_st(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct GjCheck {
  const uint8_t* pres;      // presence bytes of this call's keys
  uint32_t epoch;
  uint64_t range;
  const uint32_t* btile;    // build tile counts (top bit: a key outside the bounds)
  int64_t nbt;
  int64_t* words;           // GJ_* words
  int64_t spin_ticks;       // look-back bound in wall-clock ticks
};

// Stamped keys of slice s: presence bytes [s << sb, (s + 1) << sb) that hold the epoch.  Thread
// t tests the 16 bytes at t * 16 with one load (2^sb <= 4096 = 16 * 256 bytes; the byte array
// is padded to 256 bytes): slice_bytes issues the load, slice_keys counts.
__device__ __forceinline__ uint4 slice_bytes(const GjCheck& K, int s, int sb) {
  const uint64_t at = ((uint64_t)s << sb) + (uint64_t)threadIdx.x * 16;
  if ((int)threadIdx.x * 16 >= (1 << sb) || at >= K.range) return make_uint4(0, 0, 0, 0);
  return *(const uint4*)(K.pres + at);
}
__device__ __forceinline__ uint32_t slice_keys(const GjCheck& K, int s, int sb, uint4 q) {
  const uint64_t at = ((uint64_t)s << sb) + (uint64_t)threadIdx.x * 16;
  if ((int)threadIdx.x * 16 >= (1 << sb) || at >= K.range) return 0;
  const uint32_t rep = K.epoch * 0x01010101u;
  const uint32_t d[4] = {q.x, q.y, q.z, q.w};
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t x = d[k] ^ rep;  // bytes equal to the epoch become zero
#pragma unroll
    for (int b = 0; b < 4; ++b)
      c += (((x >> (8 * b)) & 0xFFu) == 0u && at + 4 * k + b < K.range) ? 1u : 0u;
  }
  return c;
}

// A reduce workgroup has kRB threads (8 waves: a slice's records, slots and writes spread over
// twice the waves of a 256-thread group, and two such groups fit a CU's LDS).
constexpr int kRB = 512;
constexpr int kRWaves = kRB / 64;
constexpr int kRItems = 4096 / kRB;  // slots per thread (a slice has at most 4096)

// Dynamic LDS: acc[nacc][ps] u64, cnt[ps] u32.
// With xper > 0 the grid is 8 * xper workgroups and workgroup b (which the dispatcher places on
// XCD b % 8) reduces slice (b % 8) * xper + b / 8: consecutive slices run on one XCD, so the
// record lines one slice fetches beyond its own run (the neighbouring slices' records of the same
// chunk) are found in that XCD's L2 by the next slice.  The host takes this mapping only when
// every workgroup of the grid is resident at once, which the look-back (a slice waits for every
// slice below it, whatever its workgroup index) then needs.
__global__ __launch_bounds__(kRB) void k_gj_reduce(GjSpec G, int sb, int S, int64_t W,
                                                      const uint32_t* __restrict__ loff,
                                                      const uint64_t* __restrict__ rec,
                                                      int64_t lo, GjOut O,
                                                      uint64_t* __restrict__ flags, GjCheck K,
                                                      int xper) {
  extern __shared__ uint64_t smem[];
  __shared__ uint32_t wocc[kRItems][kRWaves];
  __shared__ uint32_t s_keys, s_nrec;
  __shared__ int64_t s_off;
  const int ps = 1 << sb;
  uint64_t* lacc = smem;
  uint32_t* lcnt = (uint32_t*)(smem + (int64_t)G.nacc * ps);
  const int t = threadIdx.x;
  const int s = xper > 0 ? (int)(blockIdx.x % 8) * xper + (int)(blockIdx.x / 8) : (int)blockIdx.x;
  if (s >= S) return;  // the remapped grid's unused places
  const int rw = 1 + G.nacc;
  // the presence bytes and the first batch's chunk offsets are in flight while the table clears
  const uint4 kq = slice_bytes(K, s, sb);
  constexpr int kPer = kGjBatch / kRB;
  uint32_t la[kPer], lb[kPer];
  auto load_batch = [&](int64_t wb) {
    const int nbw = (int)(W - wb < kGjBatch ? W - wb : kGjBatch);
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = t * kPer + k;
      la[k] = i < nbw ? loff[(wb + i) * (S + 1) + s] : 0u;
      lb[k] = i < nbw ? loff[(wb + i) * (S + 1) + s + 1] : 0u;
    }
  };
  load_batch(0);
  for (int i = t; i < ps; i += kRB) {
    lcnt[i] = 0;
    for (int w = 0; w < G.nacc; ++w) lacc[w * ps + i] = G.init[G.acc_agg[w]];
  }
  if (t == 0) {
    s_keys = 0;
    s_nrec = 0;
  }
  __syncthreads();
  {
    const uint32_t kc = wave_sum(slice_keys(K, s, sb, kq));
    if (lane_id() == 0 && kc) atomicAdd(&s_keys, kc);
  }
  const uint64_t smask = (uint64_t)ps - 1;
  for (int64_t wb = 0; wb < W; wb += kGjBatch) {
    const int nbw = (int)(W - wb < kGjBatch ? W - wb : kGjBatch);
    // thread t expands the runs of slice s in its own chunks t * kPer + k of the batch
    uint32_t cnt[kPer], mine = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      cnt[k] = t * kPer + k < nbw ? lb[k] - la[k] : 0u;
      mine += cnt[k];
    }
    {
      const uint32_t wn = wave_sum(mine);
      if (lane_id() == 0 && wn) atomicAdd(&s_nrec, wn);
    }
    constexpr int U = 8;  // records per thread whose loads are issued together
    for (uint32_t p0 = 0; p0 < mine; p0 += U) {
      int64_t at[U];
      bool ok[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        // record p0 + u of the thread's runs: run k, position q in it
        uint32_t q = p0 + u;
        ok[u] = q < mine;
        int k = 0;
#pragma unroll
        for (int kk = 0; kk + 1 < kPer; ++kk)
          if (k == kk && q >= cnt[kk]) {
            q -= cnt[kk];
            k = kk + 1;
          }
        uint32_t first = la[0];
#pragma unroll
        for (int kk = 1; kk < kPer; ++kk)
          if (k == kk) first = la[kk];
        at[u] = ok[u] ? ((wb + t * kPer + k) * kGjChunk + first + q) : 0;
      }
      if (G.nacc == 1) {
        ulonglong2 x[U];
#pragma unroll
        for (int u = 0; u < U; ++u)
          x[u] = ok[u] ? *(const ulonglong2*)(rec + at[u] * 2) : make_ulonglong2(0, 0);
        const int kind = G.kind[G.acc_agg[0]];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          if (!ok[u]) continue;
          const uint32_t sl = (uint32_t)(x[u].x & smask);
          atomicAdd(&lcnt[sl], 1u);
          acc_apply(&lacc[sl], kind, x[u].y);
        }
      } else {
        uint32_t sl[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sl[u] = ok[u] ? (uint32_t)(rec[at[u] * rw] & smask) : 0u;
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (ok[u]) atomicAdd(&lcnt[sl[u]], 1u);
        for (int w = 0; w < G.nacc; ++w) {
          uint64_t v[U];
#pragma unroll
          for (int u = 0; u < U; ++u) v[u] = ok[u] ? rec[at[u] * rw + 1 + w] : 0;
          const int kind = G.kind[G.acc_agg[w]];
#pragma unroll
          for (int u = 0; u < U; ++u)
            if (ok[u]) acc_apply(&lacc[w * ps + sl[u]], kind, v[u]);
        }
      }
    }
    __syncthreads();
    if (wb + kGjBatch < W) load_batch(wb + kGjBatch);
  }
  const uint32_t nrec = s_nrec;  // this slice's records
  // occupied slots in slot order: slot i = j * kRB + t, ranked by wave ballots
  const int w = t >> 6;
  uint64_t bal[kRItems];
#pragma unroll
  for (int j = 0; j < kRItems; ++j) {
    const int i = j * kRB + t;
    bal[j] = ballot(i < ps && lcnt[i] > 0);
    if (lane_id() == 0) wocc[j][w] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  // publish this slice's counts, then add up those of the slices below it, two per thread and
  // round (every slice waited on belongs to a workgroup that is resident or done)
  uint64_t own = 0;
#pragma unroll
  for (int j = 0; j < kRItems; ++j)
#pragma unroll
    for (int v = 0; v < kRWaves; ++v) own += wocc[j][v];
  if (t == 0) gj_st(flags + s, lb_word(own, s_keys, nrec));
  {
    uint64_t before = 0, keys = 0, recs = 0;
    bool okw = true;
    const int64_t t0 = wall_clock64();
    for (int q0 = 0; q0 < s && okw; q0 += 2 * kRB) {
      const int qa = q0 + t, qb = q0 + kRB + t;
      uint64_t fa = 1ull, fb = 1ull;  // threads past s - 1 add nothing
      while (true) {
        if (qa < s) fa = gj_ld(flags + qa);
        if (qb < s) fb = gj_ld(flags + qb);
        if (ballot(fa == 0 || fb == 0) == 0) break;
        if (wall_clock64() - t0 > K.spin_ticks) {
          okw = false;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
      if (qa < s) {
        before += lb_groups(fa);
        keys += lb_keys(fa);
        recs += lb_recs(fa);
      }
      if (qb < s) {
        before += lb_groups(fb);
        keys += lb_keys(fb);
        recs += lb_recs(fb);
      }
    }
    __shared__ uint64_t lbs[3][kRWaves];
    __shared__ uint32_t lbok[kRWaves];
    before = wave_sum(before);
    keys = wave_sum(keys);
    recs = wave_sum(recs);
    if (lane_id() == 0) {
      lbs[0][w] = before;
      lbs[1][w] = keys;
      lbs[2][w] = recs;
      lbok[w] = okw ? 1u : 0u;
    }
    __syncthreads();
    if (t == 0) {
      uint64_t b = 0, k = 0, r = 0;
      bool ok = true;
#pragma unroll
      for (int v = 0; v < kRWaves; ++v) {
        b += lbs[0][v];
        k += lbs[1][v];
        r += lbs[2][v];
        ok = ok && lbok[v];
      }
      if (!ok) atomicOr((unsigned long long*)&K.words[GJ_ERR], (unsigned long long)GJ_ERR_WAIT);
      if (s == S - 1) {
        K.words[GJ_GROUPS] = (int64_t)(b + own);
        K.words[GJ_NSET] = (int64_t)(k + s_keys);
        K.words[GJ_PAIRS] = (int64_t)(r + nrec);
      }
      s_off = ok ? (int64_t)b : -1;
    }
  }
  __syncthreads();
  const int64_t obase = s_off;
  const int64_t kbase = lo + ((int64_t)s << sb);
  // output row of slot j * kRB + t (-1: unoccupied, or no valid offset)
  int64_t opos[kRItems];
  {
    uint32_t run = 0;
    bool over = false;
#pragma unroll
    for (int j = 0; j < kRItems; ++j) {
      uint32_t before = 0, tot = 0;
#pragma unroll
      for (int v = 0; v < kRWaves; ++v) {
        const uint32_t cc = wocc[j][v];
        tot += cc;
        before += (v < w) ? cc : 0u;
      }
      const int64_t o = obase + run + before + rank_below(bal[j]);
      const bool mine = obase >= 0 && ((bal[j] >> lane_id()) & 1u);
      over = over || (mine && o >= O.cap);  // more groups than the outputs hold
      opos[j] = mine && o < O.cap ? o : -1;
      run += tot;
    }
    if (over) atomicOr((unsigned long long*)&K.words[GJ_ERR], (unsigned long long)GJ_ERR_CAP);
  }
  // one tight loop per output column, the dtype dispatched once per column
#define TQ_PUT(T, VAL)                                     \
  _Pragma("unroll") for (int j = 0; j < kRItems; ++j) {    \
    if (opos[j] < 0) continue;                             \
    const int i = j * kRB + t;                             \
    ((T*)dst)[opos[j]] = (VAL);                            \
  }
  {
    void* dst = O.key;
    switch (O.key_dtype) {
      case TQ_I32: TQ_PUT(int32_t, (int32_t)(kbase + i)) break;
      case TQ_I64: TQ_PUT(int64_t, kbase + i) break;
      case TQ_I16: TQ_PUT(int16_t, (int16_t)(kbase + i)) break;
      default: TQ_PUT(int8_t, (int8_t)(kbase + i)) break;
    }
  }
  for (int a = 0; a < G.naggs; ++a) {
    void* dst = O.agg[a];
    if (G.word[a] < 0) {
      TQ_PUT(int64_t, (int64_t)lcnt[i])       // a count of a constant: the slot's pairs
      continue;
    }
    const uint64_t* acc = lacc + (int64_t)G.word[a] * ps;
    if ((G.op[a] == TQ_AGG_SUM || G.op[a] == TQ_AGG_MIN || G.op[a] == TQ_AGG_MAX) &&
        G.arg_dtype[a] == TQ_F64) {
      TQ_PUT(double, as_f64(acc[i]))          // f64 accumulators are the f64 results
    } else if (G.op[a] == TQ_AGG_COUNT ||
               (G.op[a] == TQ_AGG_SUM && G.kind[a] == ACC_ISUM)) {
      TQ_PUT(int64_t, (int64_t)acc[i])        // counts and exact integer sums
    } else {
      const int op = G.op[a], kind = G.kind[a], dt = G.arg_dtype[a];
#pragma unroll
      for (int j = 0; j < kRItems; ++j) {
        if (opos[j] < 0) continue;
        const int i = j * kRB + t;
        write_agg(dst, opos[j], op, kind, dt, acc[i], lcnt[i]);
      }
    }
  }
#undef TQ_PUT
  if (s == 0) {
    // slice 0 (which waits on no other) also sums the build's tile counts: the selected build
    // rows, and whether a key fell outside the bounds
    int64_t nsel = 0;
    uint32_t oob = 0;
    for (int64_t i = t; i < K.nbt; i += kRB) {
      const uint32_t v = K.btile[i];
      nsel += v & ~kOobBit;
      oob |= v & kOobBit;
    }
    nsel = wave_sum(nsel);
    const bool any_oob = ballot(oob != 0) != 0;
    __shared__ int64_t red[kRWaves];
    __shared__ uint32_t roob[kRWaves];
    if (lane_id() == 0) {
      red[w] = nsel;
      roob[w] = any_oob ? 1u : 0u;
    }
    __syncthreads();
    if (t == 0) {
      int64_t a = 0;
      uint32_t e = 0;
#pragma unroll
      for (int i = 0; i < kRWaves; ++i) {
        a += red[i];
        e |= roob[i];
      }
      K.words[GJ_NSEL] = a;
      if (e) atomicOr((unsigned long long*)&K.words[GJ_ERR], (unsigned long long)GJ_ERR_BOUNDS);
    }
  }
}

// ------------------------------------------------------------------ host side
struct GjPlan {
  int sb, S;
  int64_t range, nbt, W;
  int32_t stage_words;
  size_t off_btile, off_loff, off_flags, off_words, off_rec, bytes;
};

static int gj_plan(int64_t nb, int64_t np, int64_t range, int32_t naggs, int32_t nload,
                   GjPlan* p) {
  if (naggs < 1 || naggs > TQ_GB_MAX_AGGS) return set_error(TQ_EINVAL, "groupjoin: %d aggregates", naggs);
  if (range < 1) return set_error(TQ_EINVAL, "groupjoin: empty key range");
  const size_t per = 4 + 8 * (size_t)naggs;
  int sb = 0;
  while (((size_t)2 << sb) * per <= kGjTableLds && ((int64_t)2 << sb) <= kTileRows) ++sb;
  const int64_t S = ceil_div(range, (int64_t)1 << sb);
  if (S > kGjMaxSlices)
    return set_error(TQ_EUNSUPPORTED, "groupjoin: key range %lld needs %lld slices (at most %d)",
                     (long long)range, (long long)S, kGjMaxSlices);
  if (np > ((int64_t)1 << 31) - 1 || nb > ((int64_t)1 << 31) - 1)
    return set_error(TQ_EUNSUPPORTED, "groupjoin: more than 2^31 - 1 rows");
  p->sb = sb;
  p->S = (int)S;
  p->range = range;
  p->nbt = ceil_div(nb > 0 ? nb : 1, kTileRows);
  p->W = ceil_div(np > 0 ? np : 1, kGjChunk);
  p->stage_words = (int32_t)(nload > 0 ? nload : 1) * kSubRows;
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = align_up(o + bytes, 256);
    return at;
  };
  p->off_words = take(GJ_NWORDS * sizeof(int64_t));
  p->off_flags = take((size_t)S * sizeof(uint64_t));
  p->off_btile = take((size_t)p->nbt * sizeof(uint32_t));
  p->off_loff = take((size_t)(S + 1) * p->W * sizeof(uint32_t));
  p->off_rec = take((size_t)p->W * kGjChunk * (1 + naggs) * sizeof(uint64_t));
  p->bytes = o;
  return TQ_OK;
}

static int gj_check_dtype(int dt, const char* what) {
  if (dt < TQ_BOOL || dt > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "groupjoin: %s must be an integer column", what);
  return TQ_OK;
}

template <typename... P, typename... A>
static int gj_launch_lds(const char* name, void (*kernel)(P...), unsigned grid, int block,
                         size_t lds, hipStream_t s, A&&... a) {
  if (lds > (size_t)kLdsBig) return set_error(TQ_ECAPACITY, "%s: %zu bytes of LDS", name, lds);
  if (lds > 64 * 1024) {
    static std::mutex mu;
    static std::map<std::pair<int, const void*>, bool> allowed;
    int dev = 0;
    TQ_TRY_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    bool& have = allowed[{dev, (const void*)kernel}];
    if (!have) {
      const hipError_t e = hipFuncSetAttribute(
          (const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBig);
      if (e != hipSuccess)
        return set_error(TQ_EHIP, "%s: %zu bytes of LDS: %s", name, lds, hipGetErrorString(e));
      have = true;
    }
  }
  return launch(name, kernel, grid, block, lds, s, std::forward<A>(a)...);
}

// A fused filter's conjunction (none: nslot = 0).
static int gj_pred(const tq_program_t* P, const tq_colset_t* C, const char* side, Conj* out) {
  *out = Conj{};
  if (!P) return TQ_OK;
  if (!C || P->nseg < 1 || P->n_instr > TQ_MAX_INSTR || P->seg_end[0] > P->n_instr ||
      P->nload > TQ_MAX_LOADS || !match_conj(*P, *C, out))
    return set_error(TQ_EUNSUPPORTED,
                     "groupjoin: the %s predicate is not a conjunction of comparisons", side);
  return TQ_OK;
}

// The direct argument forms and the accumulated aggregates of G (kind, op, arg_dtype, naggs
// set): a count(c), c a constant other than the null sentinel (accumulator value 1 per pair),
// needs no record word and no table column.
static void gj_acc_words(const tq_program_t& prog, GjSpec* G) {
  match_direct(prog, G->naggs, G->kind, G->arg_dtype, &G->D);
  G->nacc = 0;
  for (int a = 0; a < G->naggs; ++a) {
    const bool pairs = G->D.direct && G->op[a] == TQ_AGG_COUNT && G->D.nf[a] == 0 &&
                       G->D.dconst[a] == 1ull;
    G->word[a] = pairs ? -1 : G->nacc;
    if (!pairs) G->acc_agg[G->nacc++] = a;
  }
}

static int64_t gj_spin_ticks() {
  // a look-back wait gives up after 100 ms of wall-clock time
  static int64_t ticks = 0;
  if (ticks == 0) {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0) {
      (void)hipGetLastError();
      khz = 100000;  // 100 MHz
    }
    ticks = (int64_t)khz * 100;
  }
  return ticks;
}

}  // namespace tq

using namespace tq;

extern "C" size_t tq_groupjoin_ws_bytes(int64_t nb, int64_t np, int64_t key_range, int32_t naggs,
                                        int32_t nload) {
  GjPlan p;
  if (gj_plan(nb, np, key_range, naggs, nload, &p) != TQ_OK) return 0;
  return p.bytes;
}

extern "C" int32_t tq_groupjoin_record_words(const tq_program_t* prog, int32_t naggs,
                                             const int32_t* agg_op, const int32_t* agg_dtype) {
  if (!prog || !agg_op || !agg_dtype || naggs < 1 || naggs > TQ_GB_MAX_AGGS ||
      prog->nseg != naggs)
    return -1;
  GjSpec G{};
  G.naggs = naggs;
  for (int a = 0; a < naggs; ++a) {
    if (agg_kind_init(agg_op[a], agg_dtype[a], &G.kind[a], &G.init[a]) != TQ_OK) return -1;
    G.arg_dtype[a] = agg_dtype[a];
    G.op[a] = agg_op[a];
  }
  gj_acc_words(*prog, &G);
  return 1 + G.nacc;
}

extern "C" size_t tq_groupjoin_pres_bytes(int64_t key_range) {
  return align_up((size_t)(key_range > 0 ? key_range : 1), 256);
}

namespace tq {

// The launches of one call.  The result words go to words_ext when given, else into ws (their
// address is returned through *words).
static int gj_run(const tq_groupjoin_desc_t* d, uint32_t epoch, const tq_program_t* prog,
                  const tq_colset_t* cols, void* pres, size_t pres_bytes, void* ws,
                  size_t ws_bytes, int64_t* words_ext, int64_t** words, void* stream) {
  if (!d || !words) return set_error(TQ_EINVAL, "groupjoin: null argument");
  const int ph = d->phases == 0 ? 3 : d->phases;  // bit 0: the build, bit 1: probe and reduce
  if (ph < 1 || ph > 3) return set_error(TQ_EINVAL, "groupjoin: phases %d", d->phases);
  if ((ph & 2) && (!prog || !cols)) return set_error(TQ_EINVAL, "groupjoin: null program");
  if (d->nb < 0 || d->np < 0) return set_error(TQ_EINVAL, "groupjoin: negative row count");
  if ((d->nb > 0 && !d->bkey) || (d->np > 0 && !d->pkey))
    return set_error(TQ_EINVAL, "groupjoin: null key column");
  TQ_TRY(gj_check_dtype(d->bkey_dtype, "the build key"));
  TQ_TRY(gj_check_dtype(d->pkey_dtype, "the probe key"));
  TQ_TRY(gj_check_dtype(d->out_key_dtype, "the group key output"));
  Conj BC{}, PC{};
  if (ph & 1) TQ_TRY(gj_pred(d->bpred, d->bpred_cols, "build", &BC));
  if (ph & 2) TQ_TRY(gj_pred(d->ppred, d->ppred_cols, "probe", &PC));
  if (d->key_hi < d->key_lo) return set_error(TQ_EINVAL, "groupjoin: key bounds hi < lo");
  const uint64_t urange = (uint64_t)d->key_hi - (uint64_t)d->key_lo + 1;
  if (urange == 0 || urange > ((uint64_t)1 << 40))
    return set_error(TQ_EUNSUPPORTED, "groupjoin: key range too wide");
  const int64_t range = (int64_t)urange;
  if ((ph & 2) && prog->nseg != d->naggs)
    return set_error(TQ_EINVAL, "groupjoin: %d program segments for %d aggregates", prog->nseg,
                     d->naggs);
  if (epoch < 1 || epoch > 255) return set_error(TQ_EINVAL, "groupjoin: epoch %u", epoch);
  GjPlan p;
  TQ_TRY(gj_plan(d->nb, d->np, range, d->naggs, (ph & 2) ? prog->nload : 1, &p));
  if (!ws || ws_bytes < p.bytes)
    return set_error(TQ_ECAPACITY, "groupjoin: workspace of %zu bytes, %zu needed", ws_bytes, p.bytes);
  if (!pres || pres_bytes < tq_groupjoin_pres_bytes(range))
    return set_error(TQ_ECAPACITY, "groupjoin: presence array of %zu bytes, %zu needed",
                     pres_bytes, tq_groupjoin_pres_bytes(range));
  GjSpec G{};
  GjOut O{};
  if (ph & 2) {
    if (d->out_cap > 0 && !d->out_key) return set_error(TQ_EINVAL, "groupjoin: null key output");
    G.naggs = d->naggs;
    for (int a = 0; a < d->naggs; ++a) {
      TQ_TRY(agg_kind_init(d->agg_op[a], d->agg_dtype[a], &G.kind[a], &G.init[a]));
      G.arg_dtype[a] = d->agg_dtype[a];
      G.op[a] = d->agg_op[a];
      O.agg[a] = d->out_agg[a];
      if (d->out_cap > 0 && !O.agg[a])
        return set_error(TQ_EINVAL, "groupjoin: null aggregate output");
    }
    gj_acc_words(*prog, &G);
    O.key = d->out_key;
    O.key_dtype = d->out_key_dtype;
    O.cap = d->out_cap;
  }
  hipStream_t s = (hipStream_t)stream;
  char* w = (char*)ws;
  int64_t* wd = words_ext ? words_ext : (int64_t*)(w + p.off_words);
  uint64_t* flags = (uint64_t*)(w + p.off_flags);
  uint32_t* btile = (uint32_t*)(w + p.off_btile);
  uint32_t* loff = (uint32_t*)(w + p.off_loff);
  uint64_t* rec = (uint64_t*)(w + p.off_rec);
  *words = wd;
  const unsigned bgrid = (unsigned)(d->nb > 0 ? p.nbt : 1);
  if (ph & 1)
    TQ_TRY(launch("k_gj_build", k_gj_build, bgrid, kBlock, 0, s, d->bkey, (int)d->bkey_dtype,
                  d->bmask, BC, d->nb, d->key_lo, (uint64_t)range, (uint8_t*)pres, epoch,
                  btile, flags, p.S, wd));
  if (!(ph & 2)) return TQ_OK;
  if (d->np > 0) {
    const size_t lds = (size_t)p.stage_words * 8 + (size_t)((p.S + 1) & ~1) * 4 +
                       (size_t)kGjChunk * 4;
    TQ_TRY(gj_launch_lds("k_gj_probe", k_gj_probe, (unsigned)p.W, kBlock, lds, s, G, *prog, *cols,
                         d->pkey, (int)d->pkey_dtype, d->pmask, PC, d->np, d->key_lo,
                         (uint64_t)range, (const uint8_t*)pres, epoch, p.sb, p.S,
                         p.W, loff, rec, p.stage_words));
  } else {
    // no probe rows: every chunk's slice offsets are zero
    TQ_TRY(fill_u32(loff, 0u, (int64_t)(p.S + 1) * p.W, s));
  }
  GjCheck K{(const uint8_t*)pres, epoch, (uint64_t)range, btile,
            d->nb > 0 ? p.nbt : 0, wd, gj_spin_ticks()};
  const size_t rlds = ((size_t)8 * G.nacc + 4) << p.sb;
  // consecutive slices on one XCD when the remapped grid is resident all at once
  const int xper = (int)ceil_div(p.S, 8);
  const bool remap = p.S > 8 && rlds <= 64 * 1024 &&
                     (int64_t)8 * xper <= resident_blocks((const void*)k_gj_reduce, rlds, kRB);
  const unsigned rgrid = remap ? (unsigned)(8 * xper) : (unsigned)p.S;
  return gj_launch_lds("k_gj_reduce", k_gj_reduce, rgrid, kRB, rlds, s, G, p.sb, p.S, p.W,
                       (const uint32_t*)loff, (const uint64_t*)rec, d->key_lo, O, flags, K,
                       remap ? xper : 0);
}

}  // namespace tq

extern "C" int tq_groupjoin_run(const tq_groupjoin_desc_t* d, const tq_program_t* prog,
                                const tq_colset_t* cols, void* pres, size_t pres_bytes, void* ws,
                                size_t ws_bytes, int64_t** words, void* stream) {
  return gj_run(d, d ? d->epoch : 0u, prog, cols, pres, pres_bytes, ws, ws_bytes, nullptr, words,
                stream);
}

// ------------------------------------------------------------------ without allocation
// Callers that cannot allocate (graph capture, serving loops, harnesses that forbid it) run the
// fused call on this library's own scratch: a file-scope device array, one copy per GPU, held
// until the process ends.  Its front holds the presence bytes (their size only grows, zeroed
// when it grows and when the epochs are used up), the rest the workspace.  Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU (different streams, or host threads)
// race and give wrong results.
namespace tq {

constexpr size_t kGjScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kGjScratchBytes];

struct GjScratch {
  uint8_t* ptr = nullptr;  // device address of this GPU's copy of `scratch`
  size_t pres = 0;         // bytes at the front kept for presence bytes
  uint32_t epoch = 255;    // last epoch used (255: zero before the next call)
};
static std::mutex g_scratch_mu;
static std::map<int, GjScratch> g_scratch;

}  // namespace tq

extern "C" int tq_groupjoin_run_noalloc(const tq_groupjoin_desc_t* d, const tq_program_t* prog,
                                        const tq_colset_t* cols, int64_t* words_out,
                                        void* stream) {
  if (!d || !prog || !cols || !words_out)
    return set_error(TQ_EINVAL, "groupjoin (no allocation): null argument");
  if (d->phases != 0 && d->phases != 3)
    return set_error(TQ_EINVAL, "groupjoin (no allocation): runs whole calls (phases 0)");
  if (d->key_hi < d->key_lo)
    return set_error(TQ_EINVAL, "groupjoin (no allocation): key bounds hi < lo");
  const uint64_t urange = (uint64_t)d->key_hi - (uint64_t)d->key_lo + 1;
  if (urange == 0 || urange > ((uint64_t)1 << 40))
    return set_error(TQ_EUNSUPPORTED, "groupjoin (no allocation): key range too wide");
  GjPlan p;
  TQ_TRY(gj_plan(d->nb, d->np, (int64_t)urange, d->naggs, prog->nload, &p));
  const size_t pres_b = tq_groupjoin_pres_bytes((int64_t)urange);
  int dev = 0;
  TQ_TRY_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  GjScratch& sc = g_scratch[dev];
  const size_t front = pres_b > sc.pres ? pres_b : sc.pres;
  if (front + p.bytes > kGjScratchBytes)
    return set_error(TQ_ECAPACITY,
                     "groupjoin (no allocation): %zu bytes of scratch needed (presence %zu, "
                     "workspace %zu), the library holds %zu",
                     front + p.bytes, front, p.bytes, kGjScratchBytes);
  if (!sc.ptr) {
    // this GPU's copy of the array: its device address, taken once with this device current
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(TQ_EHIP, "groupjoin (no allocation): scratch address: %s",
                       hipGetErrorString(e));
    sc.ptr = (uint8_t*)ptr;
  }
  hipStream_t s = (hipStream_t)stream;
  if (front > sc.pres || sc.epoch >= 255) {
    // presence bytes grown over workspace bytes, or every epoch used: start again from zero
    TQ_TRY(fill_u32(sc.ptr, 0u, (int64_t)(front / 4), s));
    sc.pres = front;
    sc.epoch = 0;
  }
  ++sc.epoch;
  int64_t* unused = nullptr;
  return gj_run(d, sc.epoch, prog, cols, sc.ptr, sc.pres, sc.ptr + sc.pres,
                kGjScratchBytes - sc.pres, words_out, &unused, stream);
}
