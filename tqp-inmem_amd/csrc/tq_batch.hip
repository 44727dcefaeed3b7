// Batched filter -> join -> group-by pipelines in one persistent launch.
//
// Workgroups pop (pipeline, tile) items from a device queue (one atomic per item) in phase-major
// order: clear the pipeline's group table, insert selected build rows, probe selected rows, reduce
// the probe's records (records pipelines), count groups per table tile, write groups.  An item of
// phase k waits until every item of its own pipeline's last earlier phase that has items has
// retired (per-pipeline completion counters with agent-scope release/acquire; the wait is bounded
// and reports an error instead of hanging).  Items are popped in order by running workgroups, so
// the items waited on are already held by resident workgroups and the queue always drains.
//
// The probe accumulates in one of three ways.  A table that fits LDS is accumulated there per item
// and merged once.  A records pipeline (a join pipeline whose table does not fit) writes, per
// probe chunk, one record {slot within its key slice, accumulator values} per pair, grouped by
// key slice in the chunk's own region, and a reduce item per slice folds that slice's records
// from every chunk in LDS and writes the slice's cnt/acc slots: no global atomics and no table
// clearing beyond the directory.  The remaining pipelines accumulate with global atomics.
//
// The group table of a pipeline is addressed by build key - kmin: it is the join directory
// (build row per key) and the aggregation table (pair count and accumulators per key), so the
// probe phase does one directory load per selected probe row and groups come out in key order
// without sorting.  Predicates and aggregate arguments run on the expression interpreter with the
// same LDS staging as the per-operator kernels.
#include "tq_agg.h"
#include <cstring>
#include <map>
#include <mutex>

#include "tq_common.h"
#include "tq_conj.h"
#include "tq_expr.h"

namespace tq {

static_assert(sizeof(Conj) <= sizeof(((tq_batch_pipe_t*)nullptr)->prep[0]),
              "a prepared predicate fits its descriptor field");

__device__ __forceinline__ const Conj& prepared(const tq_batch_pipe_t& B, int side) {
  return *(const Conj*)B.prep[side];
}


// The prepared predicate (terms over the program's load slots) for row j of the staged sub-tile.
__device__ __forceinline__ bool stage_conj(const Conj& C, const uint64_t* stage, int j) {
  bool ok = true;
  for (int k = 0; k < C.nterm; ++k)
    ok = ok && term_holds(C.term_op[k], C.term_imm[k],
                          stage[(C.term_slot[k] * kSub + j) * kBlock + threadIdx.x]);
  return ok;
}

// Row j of the sub-tile passes the build or probe predicate: the prepared terms, else the
// interpreter.
__device__ __forceinline__ bool passes(const tq_batch_pipe_t& B, int side,
                                       const tq_program_t& P, const tq_colset_t& cols, int j,
                                       int64_t row, const uint64_t* stage) {
  if (B.fast[side]) return stage_conj(prepared(B, side), stage, j);
  Regs R{};
  run_program(P, cols, 0, P.seg_end[0], j, row, true, stage, R);
  return R.get(P.seg_result[0]) != 0;
}

struct PhaseStarts {
  int64_t s[TQ_BATCH_PHASES + 1];
};

struct BatchWs {
  unsigned long long* head;
  uint32_t* done;   // [npipes * phases]
  uint32_t* err;
};

__device__ __forceinline__ uint32_t load_acquire(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// A poll of a phase counter: relaxed (an acquire load also invalidates the CU's caches on every
// iteration, which with hundreds of workgroups polling slows the workgroups they wait on); the
// waiter takes one acquire fence when the count is reached.
__device__ __forceinline__ uint32_t load_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ int64_t slot_of(const tq_batch_pipe_t& B, int64_t key) {
  const int64_t s = key - B.kmin;
  return (s >= 0 && s < B.range) ? s : -1;
}

__device__ __forceinline__ int acc_kind(int op, int dt) {
  const bool isf = dt == TQ_F64 || dt == TQ_F32;
  switch (op) {
    case TQ_AGG_SUM: case TQ_AGG_AVG: return isf ? ACC_FSUM : ACC_ISUM;
    case TQ_AGG_COUNT: return ACC_COUNT;
    case TQ_AGG_MIN: return isf ? ACC_FMIN : ACC_IMIN;
    default: return isf ? ACC_FMAX : ACC_IMAX;
  }
}

// Initial accumulator bits (the device twin of agg_kind_init).
__device__ __forceinline__ uint64_t acc_init(int op, int dt) {
  if (op != TQ_AGG_MIN) return 0;
  if (dt == TQ_F64) return f64_bits(1.7976931348623157e308);
  if (dt == TQ_F32) return f64_bits(3.4028234663852886e38);
  return (uint64_t)(dt == TQ_I8 ? 127 : dt == TQ_I16 ? 32767 : dt == TQ_I32 ? 2147483647
                    : dt == TQ_BOOL ? 1 : INT64_MAX);
}

__device__ uint32_t block_sum_u32(uint32_t v, uint32_t* sh) {
  v = wave_sum(v);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) t += sh[w];
  __syncthreads();
  return t;
}

// Phase 0: clear table slots [s0, s1) (a records pipeline's reduce writes every cnt/acc slot
// itself, so only its presence bytes are cleared); the first tile also zeroes the key counts.
__device__ void phase_clear(const tq_batch_pipe_t& B, int64_t s0, int64_t s1) {
  if (s0 == 0 && B.keycount && threadIdx.x < 2) B.keycount[threadIdx.x] = 0;
  if (B.pres) {
    for (int64_t s = s0 + threadIdx.x; s < s1; s += kBlock) B.pres[s] = 0;
  }
  if (B.records) return;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += kBlock) B.cnt[s] = 0;
  for (int a = 0; a < B.naggs; ++a) {
    const uint64_t init = acc_init(B.agg_op[a], B.agg_dtype[a]);
    uint64_t* acc = B.acc + (int64_t)a * B.range;
    for (int64_t s = s0 + threadIdx.x; s < s1; s += kBlock) acc[s] = init;
  }
}

// Phase 1: mark the keys of the selected build rows [r0, r1) present (plain byte stores) and
// add the item's selected rows to keycount[0]; phase 4 adds the distinct present keys to
// keycount[1], so a repeated key makes them differ.
__device__ void phase_build(const tq_batch_pipe_t& B, int64_t r0, int64_t r1, uint64_t* stage,
                            uint32_t* sh) {
  const int t = threadIdx.x;
  const tq_program_t& P = B.bprog;
  uint32_t nsel = 0;
  for (int64_t sub0 = r0; sub0 < r1; sub0 += kSubRows) {
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < r1;
    stage_loads_by_slot(P, B.bcols, sub0, live, stage);
#pragma unroll 1
    for (int j = 0; j < kSub; ++j) {
      if (!live[j]) continue;
      const int64_t row = sub0 + (int64_t)j * kBlock + t;
      if (!passes(B, 0, P, B.bcols, j, row, stage)) continue;
      const int64_t key = (int64_t)stage[(0 * kSub + j) * kBlock + t];
      const int64_t s = slot_of(B, key);
      if (s < 0) {
        atomicOr((unsigned int*)B.flags, 2u);  // key outside the declared range
        continue;
      }
      B.pres[s] = 1;
      ++nsel;
    }
  }
  nsel = block_sum_u32(nsel, sh);
  if (t == 0 && nsel) atomicAdd((unsigned long long*)&B.keycount[0], (unsigned long long)nsel);
}

// Packed group key of a group-by pipeline from its staged key slots.
__device__ __forceinline__ int64_t pack_gkey(const tq_batch_pipe_t& B, const uint64_t* stage, int j) {
  int64_t p = 0;
  for (int k = 0; k < B.nkeys; ++k) {
    const int64_t v = (int64_t)stage[(k * kSub + j) * kBlock + threadIdx.x];
    p += (v - B.gkey_min[k]) * B.gkey_stride[k];
  }
  return p;
}

// Phase 2: probe / scan selected rows [r0, r1) and accumulate each matched row into its group.
// With lds_table, the item accumulates into an LDS copy of the (small) table and merges it into
// the global table once at the end, so few hot groups cost one global atomic per slot per item.
__device__ void phase_probe(const tq_batch_pipe_t& B, int64_t r0, int64_t r1, uint64_t* stage,
                            uint64_t* ltab) {
  const int t = threadIdx.x;
  const tq_program_t& P = B.pprog;
  const tq_program_t& A = B.aprog;
  const bool local = B.lds_table != 0;
  uint32_t* lcnt = (uint32_t*)(ltab + (int64_t)B.naggs * B.range);
  if (local) {
    for (int64_t s = t; s < B.range; s += kBlock) {
      lcnt[s] = 0;
      for (int a = 0; a < B.naggs; ++a)
        ltab[(int64_t)a * B.range + s] = acc_init(B.agg_op[a], B.agg_dtype[a]);
    }
    __syncthreads();
  }
  uint32_t* cnt = local ? lcnt : B.cnt;
  uint64_t* acc = local ? ltab : B.acc;
  for (int64_t sub0 = r0; sub0 < r1; sub0 += kSubRows) {
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < r1;
    stage_loads_by_slot(P, B.pcols, sub0, live, stage);
    bool hit[kSub];
    int64_t slot[kSub];
#pragma unroll 1
    for (int j = 0; j < kSub; ++j) {
      hit[j] = false;
      slot[j] = -1;
      if (!live[j]) continue;
      const int64_t row = sub0 + (int64_t)j * kBlock + t;
      if (!passes(B, 1, P, B.pcols, j, row, stage)) continue;
      if (B.mode == 1) {
        hit[j] = true;
        slot[j] = pack_gkey(B, stage, j);
        continue;
      }
      slot[j] = slot_of(B, (int64_t)stage[(0 * kSub + j) * kBlock + t]);
    }
    if (B.mode == 0) {
      // the presence bytes of the four rows, all loads in flight together
      uint8_t b[kSub];
#pragma unroll
      for (int j = 0; j < kSub; ++j) b[j] = slot[j] >= 0 ? B.pres[slot[j]] : (uint8_t)0;
#pragma unroll
      for (int j = 0; j < kSub; ++j) hit[j] = b[j] != 0;
    }
    // every thread has read its own staged words; the aggregate stage reuses the buffer
    stage_loads_by_slot(A, B.acols, sub0, hit, stage);
#pragma unroll 1
    for (int j = 0; j < kSub; ++j) {
      if (!hit[j]) continue;
      const int64_t row = sub0 + (int64_t)j * kBlock + t;
      const int64_t s = slot[j];
      atomicAdd(&cnt[s], 1u);
      Regs R{};
      int begin = 0;
      for (int a = 0; a < B.naggs; ++a) {
        const int end = A.seg_end[a];
        run_program(A, B.acols, begin, end, j, row, true, stage, R);
        begin = end;
        const int kind = acc_kind(B.agg_op[a], B.agg_dtype[a]);
        const uint64_t v = to_acc(kind, B.agg_dtype[a], R.get(A.seg_result[a]));
        acc_apply(&acc[(int64_t)a * B.range + s], kind, v);
      }
    }
  }
  if (local) {
    __syncthreads();
    for (int64_t s = t; s < B.range; s += kBlock) {
      const uint32_t c = lcnt[s];
      if (c == 0) continue;
      atomicAdd(&B.cnt[s], c);
      for (int a = 0; a < B.naggs; ++a)
        acc_apply(&B.acc[(int64_t)a * B.range + s], acc_kind(B.agg_op[a], B.agg_dtype[a]),
                  ltab[(int64_t)a * B.range + s]);
    }
    __syncthreads();
  }
}

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
static_assert(TQ_BATCH_MAX_SLICES % kBlock == 0, "slice counts scan in whole rows per thread");

// Phase 2 of a records pipeline: probe chunk c (TQ_BATCH_CHUNK rows).  Pass 1 evaluates the
// predicate, looks each selected row's key up in the directory, keeps one hit bit per row (a
// wave ballot per 64 rows) and counts the pairs per key slice; the counts' exclusive scan is
// the chunk's row of loff and the slices' write cursors.  Pass 2 reloads the key of each paired
// row, evaluates its aggregate arguments (in row order) and writes its record at its slice's
// cursor.  lbuf: the hit bits (u64 per 64 rows), then the slice counts.
__device__ void phase_probe_rec(const tq_batch_pipe_t& B, int64_t c, uint64_t* stage,
                                uint32_t* lbuf, uint32_t* sh) {
  const int t = threadIdx.x, l = lane_id(), w = t >> 6;
  const tq_program_t& P = B.pprog;
  const tq_program_t& A = B.aprog;
  const int64_t r0 = c * TQ_BATCH_CHUNK;
  const int64_t r1 = min(r0 + (int64_t)TQ_BATCH_CHUNK, B.np);
  const int S = B.nslices, sb = B.sb;
  const int64_t W = B.ntiles[2];
  uint64_t* hm = (uint64_t*)lbuf;
  uint32_t* hist = lbuf + TQ_BATCH_CHUNK / 32;
  for (int i = t; i < S; i += kBlock) hist[i] = 0;
  __syncthreads();
  for (int64_t sub0 = r0; sub0 < r1; sub0 += kSubRows) {
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < r1;
    stage_loads_by_slot(P, B.pcols, sub0, live, stage);
#pragma unroll 1
    for (int j = 0; j < kSub; ++j) {
      bool hit = false;
      if (live[j]) {
        const int64_t row = sub0 + (int64_t)j * kBlock + t;
        if (passes(B, 1, P, B.pcols, j, row, stage)) {
          const int64_t s = slot_of(B, (int64_t)stage[(0 * kSub + j) * kBlock + t]);
          if (s >= 0 && B.pres[s] != 0) {
            hit = true;
            atomicAdd(&hist[(uint32_t)s >> sb], 1u);
          }
        }
      }
      // rows sub0 - r0 + j * kBlock + 64 w + [0, 64): one word of hit bits per wave
      const uint64_t bits = ballot(hit);
      if (l == 0) hm[(sub0 - r0 + (int64_t)j * kBlock) / 64 + w] = bits;
    }
  }
  __syncthreads();
  // exclusive scan of the slice counts (hist becomes the slices' write cursors)
  constexpr int kPer = TQ_BATCH_MAX_SLICES / kBlock;
  uint32_t v[kPer], sum = 0;
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = t * kPer + k;
    v[k] = i < S ? hist[i] : 0u;
    sum += v[k];
  }
  uint32_t M;
  uint32_t run = block_exclusive_sum(sum, sh, &M);
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int i = t * kPer + k;
    if (i < S) {
      hist[i] = run;
      B.loff[(int64_t)i * W + c] = run;
    }
    run += v[k];
  }
  if (t == 0) B.loff[(int64_t)S * W + c] = M;
  __syncthreads();
  const int rw = 1 + B.naggs;
  const uint32_t smask = (1u << sb) - 1u;
  uint64_t* crec = B.rec + r0 * rw;
  for (int64_t sub0 = r0; sub0 < r1; sub0 += kSubRows) {
    bool hit[kSub];
    bool any = false;
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      hit[j] = (hm[(sub0 - r0 + (int64_t)j * kBlock) / 64 + w] >> l) & 1ull;
      any |= hit[j];
    }
    if (!__any(any)) continue;  // no pair in the wave: its loads are skipped (no LDS barrier here)
    // the key (load slot 0 of the predicate) of each paired row, then the aggregate arguments
    stage_loads_by_slot(P, B.pcols, sub0, hit, stage);
    uint32_t sv[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j)
      sv[j] = hit[j] ? (uint32_t)slot_of(B, (int64_t)stage[(0 * kSub + j) * kBlock + t]) : 0u;
    stage_loads_by_slot(A, B.acols, sub0, hit, stage);
#pragma unroll 1
    for (int j = 0; j < kSub; ++j) {
      if (!hit[j]) continue;
      const int64_t row = sub0 + (int64_t)j * kBlock + t;
      const uint32_t pos = atomicAdd(&hist[sv[j] >> sb], 1u);
      uint64_t* r = crec + (int64_t)pos * rw;
      r[0] = sv[j] & smask;
      Regs R{};
      int begin = 0;
      for (int a = 0; a < B.naggs; ++a) {
        const int end = A.seg_end[a];
        run_program(A, B.acols, begin, end, j, row, true, stage, R);
        begin = end;
        r[1 + a] = to_acc(acc_kind(B.agg_op[a], B.agg_dtype[a]), B.agg_dtype[a],
                          R.get(A.seg_result[a]));
      }
    }
  }
}

// Phase 3 of a records pipeline: key slice s.  Its records in every chunk (loff column s) are
// folded into an LDS table of the slice's 2^sb slots, which then overwrites the slice's cnt/acc
// slots (coalesced, no atomics).  lbuf: cnt[2^sb], then acc[naggs][2^sb].
__device__ void phase_reduce(const tq_batch_pipe_t& B, int s, uint32_t* lbuf) {
  const int t = threadIdx.x;
  const int64_t ns = (int64_t)1 << B.sb;
  const int64_t g0 = (int64_t)s * ns;
  const int64_t n = min(ns, B.range - g0);
  uint32_t* lcnt = lbuf;
  uint64_t* lacc = (uint64_t*)(lbuf + ns);
  for (int64_t i = t; i < n; i += kBlock) {
    lcnt[i] = 0;
    for (int a = 0; a < B.naggs; ++a) lacc[a * ns + i] = acc_init(B.agg_op[a], B.agg_dtype[a]);
  }
  __syncthreads();
  const int64_t W = B.ntiles[2];
  const int rw = 1 + B.naggs;
  for (int64_t c = t; c < W; c += kBlock) {
    const uint32_t e0 = B.loff[(int64_t)s * W + c];
    const uint32_t e1 = B.loff[(int64_t)(s + 1) * W + c];
    const uint64_t* cr = B.rec + c * TQ_BATCH_CHUNK * rw;
    for (uint32_t e = e0; e < e1; ++e) {
      const uint64_t* r = cr + (int64_t)e * rw;
      const uint32_t slot = (uint32_t)r[0];
      atomicAdd(&lcnt[slot], 1u);
      for (int a = 0; a < B.naggs; ++a)
        acc_apply(&lacc[a * ns + slot], acc_kind(B.agg_op[a], B.agg_dtype[a]), r[1 + a]);
    }
  }
  __syncthreads();
  for (int64_t i = t; i < n; i += kBlock) {
    B.cnt[g0 + i] = lcnt[i];
    for (int a = 0; a < B.naggs; ++a) B.acc[(int64_t)a * B.range + g0 + i] = lacc[a * ns + i];
  }
}


// Phase 3: occupied slots of [s0, s1) -> tile_counts[tile].
__device__ void phase_count(const tq_batch_pipe_t& B, int64_t s0, int64_t s1, int tile, uint32_t* sh) {
  const bool force0 = B.mode == 1 && B.nkeys == 0;
  uint32_t c = 0, k = 0;
  for (int64_t s = s0 + threadIdx.x; s < s1; s += kBlock) {
    c += (B.cnt[s] > 0) || (force0 && s == 0);
    if (B.pres) k += B.pres[s] != 0;
  }
  c = block_sum_u32(c, sh);
  if (threadIdx.x == 0) B.tile_counts[tile] = c;
  if (B.pres) {
    k = block_sum_u32(k, sh);
    if (threadIdx.x == 0 && k) atomicAdd((unsigned long long*)&B.keycount[1], (unsigned long long)k);
  }
}

// Phase 4: write the groups of [s0, s1) after the groups of earlier tiles.
__device__ void phase_write(const tq_batch_pipe_t& B, int64_t s0, int64_t s1, int tile,
                            uint32_t* sh, uint32_t* wsum) {
  uint32_t pre = 0;
  for (int i = threadIdx.x; i < tile; i += kBlock) pre += B.tile_counts[i];
  int64_t run = block_sum_u32(pre, sh);
  const int t = threadIdx.x, w = t >> 6;
  const bool force0 = B.mode == 1 && B.nkeys == 0;
  for (int64_t base = s0; base < s1; base += kBlock) {
    const int64_t s = base + t;
    const bool f = s < s1 && ((B.cnt[s] > 0) || (force0 && s == 0));
    const uint64_t bal = ballot(f);
    if (lane_id() == 0) wsum[w] = (uint32_t)__popcll(bal);
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      tot += wsum[v];
      before += v < w ? wsum[v] : 0;
    }
    if (f) {
      const int64_t o = run + before + rank_below(bal);
      if (B.mode == 1) {
        for (int k = 0; k < B.nkeys; ++k)
          store_int(B.gkey_out[k], B.gkey_dtype[k], o,
                    (s / B.gkey_stride[k]) % B.gkey_range[k] + B.gkey_min[k]);
      } else {
        store_int(B.key_out, B.key_dtype, o, B.kmin + s);
      }
      const uint32_t rows = B.cnt[s];
      for (int a = 0; a < B.naggs; ++a)
        write_agg(B.agg_out[a], o, B.agg_op[a], acc_kind(B.agg_op[a], B.agg_dtype[a]),
                  B.agg_dtype[a], B.acc[(int64_t)a * B.range + s], rows);
    }
    run += tot;
    __syncthreads();
  }
  if (tile == B.ntiles[5] - 1 && threadIdx.x == 0) *B.ngroups = run;
}

// Completions a workgroup has not yet published: the phase they belong to, and (pipeline,
// items) runs.  They are published -- one release fence (the write-back of the XCD's L2) for all
// of them, then the counters -- when the workgroup takes an item of another phase, when the list
// is full, and when it leaves the queue.  A workgroup's unpublished completions are therefore
// always of the phase of the item it is working on, and no workgroup waits on a phase that is not
// earlier than its own item's, so every awaited completion is published or held by a workgroup
// that is working, not waiting.
constexpr int kPubMax = 32;
struct PubList {
  int n;
  int phase;
  int32_t pipe[kPubMax];
  uint32_t count[kPubMax];
};

__device__ void publish(PubList& L, uint32_t* done) {
  __syncthreads();  // every wave's stores have reached the L2
  if (threadIdx.x == 0 && L.n > 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    for (int i = 0; i < L.n; ++i)
      __hip_atomic_fetch_add(done + (int64_t)L.pipe[i] * TQ_BATCH_PHASES + L.phase, L.count[i],
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    L.n = 0;
  }
  __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_batch(const tq_batch_pipe_t* __restrict__ pipes,
                                                  const int32_t* __restrict__ items,
                                                  PhaseStarts ps, int32_t tile_rows, BatchWs ws,
                                                  int64_t wait_ticks, int32_t stage_words) {
  extern __shared__ uint64_t stage[];
  uint64_t* ltab = stage + stage_words;
  __shared__ int64_t s_item;
  __shared__ int s_abort;
  __shared__ uint32_t sh[kWaves];
  __shared__ uint32_t wsum[kWaves];
  __shared__ PubList pub;
  __shared__ int s_acq_pipe, s_acq_phase;  // the dependency this workgroup last acquired
  const int64_t nitems = ps.s[TQ_BATCH_PHASES];
  if (threadIdx.x == 0) {
    pub.n = 0;
    pub.phase = -1;
    s_acq_pipe = -1;
    s_acq_phase = -1;
  }
  while (true) {
    if (threadIdx.x == 0) {
      s_item = (int64_t)atomicAdd(ws.head, 1ull);
      s_abort = __hip_atomic_load(ws.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    }
    __syncthreads();
    const int64_t it = s_item;
    const bool abort = s_abort;
    __syncthreads();
    if (it >= nitems || abort) break;
    const int p = __builtin_amdgcn_readfirstlane(items[2 * it]);
    const int tile = __builtin_amdgcn_readfirstlane(items[2 * it + 1]);
    int phase = 0;
    while (phase < TQ_BATCH_PHASES - 1 && it >= ps.s[phase + 1]) ++phase;
    const tq_batch_pipe_t& B = pipes[p];
    // completions of an earlier phase go out before this item waits on anything
    if (pub.n > 0 && pub.phase != phase) publish(pub, ws.done);
    int prev = phase - 1;  // the pipeline's last earlier phase that has items
    while (prev >= 0 && B.ntiles[prev] == 0) --prev;
    if (prev >= 0 && !(s_acq_pipe == p && s_acq_phase == prev)) {
      if (threadIdx.x == 0) {
        // wait for that phase, at most wait_ticks of wall-clock time
        const uint32_t need = (uint32_t)B.ntiles[prev];
        const uint32_t* d = ws.done + (int64_t)p * TQ_BATCH_PHASES + prev;
        const int64_t t0 = wall_clock64();
        bool late = false;
        while (load_relaxed(d) < need) {
          __builtin_amdgcn_s_sleep(2);
          if (wall_clock64() - t0 > wait_ticks) {
            atomicOr(ws.err, 1u);
            late = true;
            break;
          }
        }
        s_abort = late;
      }
      __syncthreads();
      if (s_abort) break;
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __syncthreads();  // every thread has read s_acq_* above
      if (threadIdx.x == 0) {
        s_acq_pipe = p;
        s_acq_phase = prev;
      }
    }
    const int64_t lo = (int64_t)tile * tile_rows;
    switch (phase) {
      case 0: phase_clear(B, lo, min(lo + tile_rows, B.range)); break;
      case 1: phase_build(B, lo, min(lo + tile_rows, B.nb), stage, sh); break;
      case 2:
        if (B.records) phase_probe_rec(B, tile, stage, (uint32_t*)ltab, sh);
        else phase_probe(B, lo, min(lo + tile_rows, B.np), stage, ltab);
        break;
      case 3: phase_reduce(B, tile, (uint32_t*)ltab); break;
      case 4: phase_count(B, lo, min(lo + tile_rows, B.range), tile, sh); break;
      default: phase_write(B, lo, min(lo + tile_rows, B.range), tile, sh, wsum); break;
    }
    __syncthreads();
    // every wave reads the list length before thread 0 appends to it: a wave that read it after
    // the append would call publish (and its barriers) alone and fall out of step with the
    // workgroup's other waves
    const bool full = pub.n == kPubMax;
    __syncthreads();
    if (full) publish(pub, ws.done);
    if (threadIdx.x == 0) {
      if (pub.n > 0 && pub.pipe[pub.n - 1] == p) {
        ++pub.count[pub.n - 1];
      } else {
        pub.pipe[pub.n] = p;
        pub.count[pub.n] = 1;
        ++pub.n;
      }
      pub.phase = phase;
    }
  }
  publish(pub, ws.done);
}

static BatchWs carve_batch(void* base, int32_t npipes) {
  char* p = (char*)base;
  BatchWs w;
  w.head = (unsigned long long*)p; p += 256;
  w.err = (uint32_t*)p; p += 256;
  w.done = (uint32_t*)p;
  return w;
}

}  // namespace tq

using namespace tq;

extern "C" int tq_batch_prepare(tq_batch_pipe_t* pipes, int32_t npipes) {
  if (npipes < 0 || (npipes > 0 && !pipes)) return set_error(TQ_EINVAL, "batch prepare: bad arguments");
  int n = 0;
  for (int i = 0; i < npipes; ++i) {
    tq_batch_pipe_t& B = pipes[i];
    for (int side = 0; side < 2; ++side) {
      B.fast[side] = 0;
      memset(B.prep[side], 0, sizeof(B.prep[side]));
      if (B.mode != 0 || !B.pres) continue;
      const tq_program_t& P = side ? B.pprog : B.bprog;
      const tq_colset_t& cols = side ? B.pcols : B.bcols;
      // load slot 0 is the key, read as an integer whatever the predicate reads
      if (P.nload < 1 || P.load_dtype[0] < TQ_I8 || P.load_dtype[0] > TQ_I64 ||
          cols.width[P.load_col[0]] != 1)
        continue;
      Conj C{};
      if (!match_conj(P, cols, &C, false)) continue;
      memcpy(B.prep[side], &C, sizeof(C));
      B.fast[side] = 1;
      ++n;
    }
  }
  return n;
}

extern "C" size_t tq_batch_ws_bytes(int32_t npipes) {
  return 512 + align_up((size_t)(npipes > 0 ? npipes : 1) * TQ_BATCH_PHASES * 4, 256);
}

extern "C" int tq_batch_run(const tq_batch_pipe_t* pipes, int32_t npipes, const int32_t* items,
                            const int64_t* phase_start, int32_t tile_rows, int32_t grid,
                            int32_t stage_loads, int32_t lds_table_bytes, void* ws,
                            uint32_t wait_us, void* stream) {
  if (!pipes || !items || !phase_start || !ws) return set_error(TQ_EINVAL, "batch: null argument");
  if (tile_rows <= 0 || tile_rows % kSubRows) return set_error(TQ_EINVAL, "batch: tile_rows %d not a multiple of %d", tile_rows, kSubRows);
  if (stage_loads < 1 || stage_loads > TQ_MAX_LOADS) return set_error(TQ_EINVAL, "batch: stage_loads %d", stage_loads);
  if (grid < 1) return set_error(TQ_EINVAL, "batch: grid %d", grid);
  PhaseStarts ps;
  for (int i = 0; i <= TQ_BATCH_PHASES; ++i) {
    ps.s[i] = phase_start[i];
    if (i > 0 && ps.s[i] < ps.s[i - 1]) return set_error(TQ_EINVAL, "batch: phase starts not ascending");
  }
  hipStream_t s = (hipStream_t)stream;
  BatchWs w = carve_batch(ws, npipes);
  TQ_TRY(fill_u64(w.head, 0, 1, s));
  TQ_TRY(fill_u32(w.err, 0, 1, s));
  TQ_TRY(fill_u32(w.done, 0, (int64_t)npipes * TQ_BATCH_PHASES, s));
  if (ps.s[TQ_BATCH_PHASES] == 0) return TQ_OK;
  if (lds_table_bytes < 0 || lds_table_bytes > 48 * 1024)
    return set_error(TQ_EINVAL, "batch: LDS table of %d bytes", lds_table_bytes);
  const int32_t stage_words = stage_loads * kSubRows;
  const size_t lds = (size_t)stage_words * 8 + align_up((size_t)lds_table_bytes, 8);
  if (lds > 64 * 1024) return set_error(TQ_EINVAL, "batch: %zu bytes of LDS", lds);
  // the wall clock's rate, once per device (kHz: ticks per millisecond)
  static std::mutex mu;
  static std::map<int, int64_t> khz_of;
  int dev = 0;
  TQ_TRY_HIP(hipGetDevice(&dev));
  int64_t khz;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = khz_of.find(dev);
    if (it == khz_of.end()) {
      int r = 0;
      if (hipDeviceGetAttribute(&r, hipDeviceAttributeWallClockRate, dev) != hipSuccess || r <= 0) {
        (void)hipGetLastError();
        r = 100000;  // 100 MHz
      }
      it = khz_of.emplace(dev, (int64_t)r).first;
    }
    khz = it->second;
  }
  const int64_t wait_ticks = (int64_t)wait_us * khz / 1000;
  return launch("k_batch", k_batch, (unsigned)grid, kBlock, lds, s, pipes, items, ps, tile_rows, w,
                wait_ticks, stage_words);
}
