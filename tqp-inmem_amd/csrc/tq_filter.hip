// Predicate filter, tile scans, stream compaction and multi-column gather.
//
// Tiles: a workgroup of 256 threads owns 4096 consecutive rows; thread t handles rows
// tile0 + j*256 + t (j < 16), so every wave-level load touches 64 consecutive rows.  The filter
// writes one mask byte per row and the tile's popcount; compaction ranks selected rows inside
// the tile with wave ballots + mbcnt and places the tile at its exclusive-scan offset, so the
// output keeps input order exactly as `tensor[mask]` does.
#include "tq_common.h"
#include "tq_conj.h"
#include "tq_expr.h"

namespace tq {

int make_colpack(ColPack* cp, int32_t ncols, const void* const* src, const int32_t* bytes,
                 const int32_t* fkind, void* const* dst) {
  if (ncols < 0 || ncols > TQ_MAX_COLS)
    return set_error(TQ_EINVAL, "column count %d outside [0, %d]", ncols, TQ_MAX_COLS);
  *cp = ColPack{};
  cp->n = ncols;
  for (int c = 0; c < ncols; ++c) {
    if (bytes[c] <= 0 || bytes[c] > 4096)
      return set_error(TQ_EINVAL, "column %d: row size %d bytes", c, bytes[c]);
    cp->src[c] = src[c];
    cp->dst[c] = dst[c];
    cp->bytes[c] = bytes[c];
    cp->fkind[c] = fkind ? fkind[c] : 0;
  }
  return TQ_OK;
}

// ------------------------------------------------------------------ filter
__global__ __launch_bounds__(kBlock) void k_filter(tq_program_t P, tq_colset_t C, int64_t n,
                                                   uint8_t* __restrict__ mask,
                                                   uint32_t* __restrict__ tile_counts) {
  extern __shared__ uint64_t stage[];
  __shared__ uint32_t sh[kWaves];
  const int t = threadIdx.x;
  const int end = P.seg_end[0];
  const int res = P.seg_result[0] & (TQ_NREG - 1);
  uint32_t cnt = 0;
  for (int sub = 0; sub < kTileRows / kSubRows; ++sub) {
    const int64_t sub0 = (int64_t)blockIdx.x * kTileRows + (int64_t)sub * kSubRows;
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < n;
    stage_loads(P, C, sub0, live, stage);
    int64_t rows[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) rows[j] = sub0 + (int64_t)j * kBlock + t;
    RegsN R{};
    run_rows(P, C, 0, end, rows, live, stage, R);
    const V4 O = R.get(res);
    const uint64_t* o = O.v;
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      if (live[j]) {
        const uint8_t v = o[j] != 0;
        mask[rows[j]] = v;
        cnt += v;
      }
    }
  }
  if (tile_counts) {
    cnt = wave_sum(cnt);
    if (lane_id() == 0) sh[t >> 6] = cnt;
    __syncthreads();
    if (t == 0) {
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += sh[w];
      tile_counts[blockIdx.x] = s;
    }
  }
}

// ------------------------------------------------------------------ comparison conjunctions
// Most predicates are conjunctions of column-versus-constant comparisons (tq_conj.h):
// tq_filter_mask runs them through k_filter_conj instead of the interpreter, and writes
// a run's four mask bytes as one 32-bit store.
__global__ __launch_bounds__(kBlock) void k_filter_conj(Conj C, int64_t n,
                                                        uint8_t* __restrict__ mask,
                                                        uint32_t* __restrict__ tile_counts) {
  __shared__ uint32_t sh[kWaves];
  const int64_t tile = blockIdx.x;
  const int t = threadIdx.x;
  uint32_t acc = load_sel16(nullptr, tile * kTileRows, n);
  acc = conj_sel16(C, tile * kTileRows, n, acc);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t r = tile * kTileRows + (int64_t)q * (4 * kBlock) + 4 * t;
    const uint32_t b4 = (acc >> (4 * q)) & 15u;
    if (r + 3 < n && ((uintptr_t)(mask + r) & 3) == 0) {
      *(uint32_t*)(mask + r) = (b4 & 1u) | ((b4 & 2u) << 7) | ((b4 & 4u) << 14) | ((b4 & 8u) << 21);
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (r + i < n) mask[r + i] = (b4 >> i) & 1u;
    }
  }
  if (tile_counts) {
    uint32_t cnt = wave_sum((uint32_t)__popc(acc));
    if (lane_id() == 0) sh[t >> 6] = cnt;
    __syncthreads();
    if (t == 0) {
      uint32_t s2 = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s2 += sh[w];
      tile_counts[tile] = s2;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_eval(tq_program_t P, tq_colset_t C, int64_t n,
                                                 void* __restrict__ out, int out_bytes) {
  extern __shared__ uint64_t stage[];
  const int t = threadIdx.x;
  const int end = P.seg_end[0];
  const int res = P.seg_result[0];
  for (int64_t sub = blockIdx.x; sub * kSubRows < n; sub += gridDim.x) {
    const int64_t sub0 = sub * kSubRows;
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < n;
    stage_loads(P, C, sub0, live, stage);
    int64_t rows[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) rows[j] = sub0 + (int64_t)j * kBlock + t;
    RegsN R{};
    run_rows(P, C, 0, end, rows, live, stage, R);
    const V4 O = R.get(res);
    const uint64_t* o = O.v;
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      const int64_t row = rows[j];
      if (live[j]) {
        const uint64_t v = o[j];
        switch (out_bytes) {
          case 1: ((uint8_t*)out)[row] = (uint8_t)v; break;
          case 2: ((uint16_t*)out)[row] = (uint16_t)v; break;
          case 4: ((uint32_t*)out)[row] = (uint32_t)v; break;
          default: ((uint64_t*)out)[row] = v; break;
        }
      }
    }
  }
}

// Every segment of a program per row, segment a into out[a] (one staging of the loads for all;
// a thread's stage entries are its own, as in k_eval).
struct OutSet {
  void* out[TQ_MAX_SEGS];
  int32_t bytes[TQ_MAX_SEGS];
};

__device__ __forceinline__ void store_low(void* out, int bytes, int64_t row, uint64_t v) {
  switch (bytes) {
    case 1: ((uint8_t*)out)[row] = (uint8_t)v; break;
    case 2: ((uint16_t*)out)[row] = (uint16_t)v; break;
    case 4: ((uint32_t*)out)[row] = (uint32_t)v; break;
    default: ((uint64_t*)out)[row] = v; break;
  }
}

__global__ __launch_bounds__(kBlock) void k_eval_multi(tq_program_t P, tq_colset_t C, int64_t n,
                                                       OutSet O) {
  extern __shared__ uint64_t stage[];
  const int t = threadIdx.x;
  for (int64_t sub = blockIdx.x; sub * kSubRows < n; sub += gridDim.x) {
    const int64_t sub0 = sub * kSubRows;
    bool live[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) live[j] = sub0 + (int64_t)j * kBlock + t < n;
    stage_loads(P, C, sub0, live, stage);
    int64_t rows[kSub];
#pragma unroll
    for (int j = 0; j < kSub; ++j) rows[j] = sub0 + (int64_t)j * kBlock + t;
    RegsN R{};
    int begin = 0;
    for (int a = 0; a < P.nseg; ++a) {
      const int end = P.seg_end[a];
      run_rows(P, C, begin, end, rows, live, stage, R);
      begin = end;
      const V4 V = R.get(P.seg_result[a]);
#pragma unroll
      for (int j = 0; j < kSub; ++j)
        if (live[j]) store_low(O.out[a], O.bytes[a], rows[j], V.v[j]);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_mask_counts(const uint8_t* __restrict__ mask,
                                                        int64_t n,
                                                        uint32_t* __restrict__ tile_counts) {
  __shared__ uint32_t sh[kWaves];
  bool f[kItems];
  select_rows(mask, (int64_t)blockIdx.x * kTileRows, n, f);
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) cnt += f[j] ? 1u : 0u;
  cnt = wave_sum(cnt);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += sh[w];
    tile_counts[blockIdx.x] = t;
  }
}

// ------------------------------------------------------------------ scans
// A scan chunk is 4096 u32 inputs: thread t owns the 16 consecutive inputs [16t, 16t + 16),
// read with four 16-byte loads and written back with 16-byte stores.
constexpr int kScanChunk = kBlock * kItems;

__device__ __forceinline__ void load16_u32(const uint32_t* in, int64_t b0, int64_t n,
                                           uint32_t (&v)[kItems]) {
  if (b0 + kItems <= n && ((uintptr_t)(in + b0) & 15) == 0) {
#pragma unroll
    for (int q = 0; q < kItems / 4; ++q) {
      const uint4 x = *(const uint4*)(in + b0 + 4 * q);
      v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
    }
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j) v[j] = b0 + j < n ? in[b0 + j] : 0u;
  }
}

template <typename TOut>
__device__ __forceinline__ void store16(TOut* out, int64_t b0, int64_t n, const TOut (&v)[kItems]) {
  if (b0 + kItems <= n && ((uintptr_t)(out + b0) & 15) == 0) {
    if (sizeof(TOut) == 4) {
#pragma unroll
      for (int q = 0; q < kItems / 4; ++q)
        *(uint4*)(out + b0 + 4 * q) = make_uint4((uint32_t)v[4 * q], (uint32_t)v[4 * q + 1],
                                                 (uint32_t)v[4 * q + 2], (uint32_t)v[4 * q + 3]);
    } else {
#pragma unroll
      for (int q = 0; q < kItems / 2; ++q)
        *(longlong2*)(out + b0 + 2 * q) = make_longlong2((long long)v[2 * q], (long long)v[2 * q + 1]);
    }
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if (b0 + j < n) out[b0 + j] = v[j];
  }
}

template <typename TOut>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const uint32_t* __restrict__ in, int64_t n,
                                                       TOut* __restrict__ sums, ScanLen len) {
  __shared__ TOut sh[kWaves];
  n = len.of(n);
  if ((int64_t)blockIdx.x * kScanChunk >= n) return;  // past a device length: never read
  const int64_t b0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  load16_u32(in, b0, n, v);
  TOut s = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) s += (TOut)v[j];
  s = wave_sum(s);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    TOut t = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) t += sh[w];
    sums[blockIdx.x] = t;
  }
}

// Exclusive scan of one chunk; with chunk_sums (the sums of all chunks), the chunk first adds
// up the sums of the chunks before it.  The last chunk writes the grand total to *total when
// total is non-null.
template <typename TOut>
__global__ __launch_bounds__(kBlock) void k_scan_chunk(const uint32_t* __restrict__ in, int64_t n,
                                                       TOut* __restrict__ out,
                                                       const TOut* __restrict__ chunk_sums,
                                                       TOut* __restrict__ total, ScanLen len) {
  __shared__ TOut sh[kWaves];
  n = len.of(n);
  if ((int64_t)blockIdx.x * kScanChunk >= n) {  // past a device length: never read
    if (n == 0 && total && blockIdx.x == 0 && threadIdx.x == 0) *total = 0;
    return;
  }
  const int64_t last = (n - 1) / kScanChunk;  // the last chunk with elements
  const int64_t b0 = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kItems;
  TOut prior = 0;
  if (chunk_sums) {  // at most kScanChunk chunks: 16 sums per thread
    TOut p = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += kBlock) p += chunk_sums[i];
    TOut tp;
    block_exclusive_sum(p, sh, &tp);
    prior = tp;
  }
  uint32_t v[kItems];
  load16_u32(in, b0, n, v);
  TOut s = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) s += (TOut)v[j];
  TOut tot;
  TOut run = prior + block_exclusive_sum(s, sh, &tot);
  TOut o[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    o[j] = run;
    run += (TOut)v[j];
  }
  store16(out, b0, n, o);
  if (total && (int64_t)blockIdx.x == last && threadIdx.x == 0) *total = prior + tot;
}

template <typename TIn, typename TOut>
int scan_exclusive(const TIn* in, int64_t n, TOut* out, TOut* total, TOut* ws,
                   hipStream_t stream, ScanLen len) {
  const int64_t nchunks = ceil_div(n, kScanChunk);
  if (n <= 0) {
    if (total) return sizeof(TOut) == 8 ? fill_u64(total, 0, 1, stream) : fill_u32(total, 0, 1, stream);
    return TQ_OK;
  }
  static_assert(sizeof(TIn) == 4, "scans read u32 counts");
  if (nchunks == 1) {
    return launch("k_scan_chunk", (k_scan_chunk<TOut>), 1, kBlock, 0, stream, (const uint32_t*)in,
                  n, out, (const TOut*)nullptr, total, len);
  }
  if (nchunks > kScanChunk) return set_error(TQ_EINVAL, "scan of %lld elements too large", (long long)n);
  TQ_TRY(launch("k_chunk_sums", (k_chunk_sums<TOut>), (unsigned)nchunks, kBlock, 0, stream,
                (const uint32_t*)in, n, ws, len));
  return launch("k_scan_chunk", (k_scan_chunk<TOut>), (unsigned)nchunks, kBlock, 0, stream,
                (const uint32_t*)in, n, out, (const TOut*)ws, total, len);
}

template int scan_exclusive<uint32_t, uint32_t>(const uint32_t*, int64_t, uint32_t*, uint32_t*,
                                                uint32_t*, hipStream_t, ScanLen);
template int scan_exclusive<uint32_t, int64_t>(const uint32_t*, int64_t, int64_t*, int64_t*,
                                               int64_t*, hipStream_t, ScanLen);

// ------------------------------------------------------------------ compaction
__global__ __launch_bounds__(kBlock) void k_compact(const uint8_t* __restrict__ mask, int64_t n,
                                                    const int64_t* __restrict__ tile_off,
                                                    ColPack cp) {
  __shared__ uint32_t wcnt[kItems][kWaves];
  const int t = threadIdx.x, w = t >> 6, l = lane_id();
  const int64_t tile0 = (int64_t)blockIdx.x * kTileRows;
  bool f[kItems];
  uint64_t bal[kItems];
  select_rows(mask, tile0, n, f);
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    bal[j] = ballot(f[j]);
    if (l == 0) wcnt[j][w] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  int64_t pos[kItems];
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
    pos[j] = run + before + rank_below(bal[j]);
    run += tot;
  }
  for (int c = 0; c < cp.n; ++c) {
    const void* src = cp.src[c];
    void* dst = cp.dst[c];
    const int bytes = cp.bytes[c];
    if (src == nullptr) {  // row numbers instead of a column
#pragma unroll
      for (int j = 0; j < kItems; ++j)
        if (f[j]) ((int64_t*)dst)[pos[j]] = tile0 + (int64_t)j * kBlock + t;
      continue;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j)
      if (f[j]) copy_row(dst, pos[j], src, tile0 + (int64_t)j * kBlock + t, bytes);
  }
}

// ------------------------------------------------------------------ gather
__global__ __launch_bounds__(kBlock) void k_gather(const int64_t* __restrict__ idx, int64_t m,
                                                   ColPack cp, int null_fill) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) {
    const int64_t s = idx[i];
    if (null_fill && s == -64) {
      for (int c = 0; c < cp.n; ++c) write_null(cp.dst[c], i, cp.bytes[c], cp.fkind[c]);
    } else {
      for (int c = 0; c < cp.n; ++c) copy_row(cp.dst[c], i, cp.src[c], s, cp.bytes[c]);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ p, T v, int64_t count) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) p[i] = v;
}

int fill_u32(void* p, uint32_t v, int64_t count, hipStream_t s) {
  if (count <= 0) return TQ_OK;
  return launch("k_fill", k_fill<uint32_t>, stream_grid(count), kBlock, 0, s, (uint32_t*)p, v, count);
}

int fill_u64(void* p, uint64_t v, int64_t count, hipStream_t s) {
  if (count <= 0) return TQ_OK;
  return launch("k_fill", k_fill<uint64_t>, stream_grid(count), kBlock, 0, s, (uint64_t*)p, v, count);
}

__global__ __launch_bounds__(kBlock) void k_mark(const int64_t* __restrict__ idx, int64_t m,
                                                uint8_t* __restrict__ flags, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) {
    const int64_t r = idx[i];
    if (r >= 0 && r < n) flags[r] = 1;
  }
}

unsigned stream_grid(int64_t m) {
  int64_t g = ceil_div(m, kBlock);
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

}  // namespace tq

using namespace tq;

extern "C" int tq_filter_mask(const tq_program_t* prog, const tq_colset_t* cols, int64_t n,
                              uint8_t* mask_out, uint32_t* tile_counts, void* stream) {
  if (!prog || !cols || (n > 0 && !mask_out)) return set_error(TQ_EINVAL, "filter: null argument");
  if (prog->nseg < 1 || prog->n_instr > TQ_MAX_INSTR || prog->seg_end[0] > prog->n_instr)
    return set_error(TQ_EINVAL, "filter: malformed program");
  if (n <= 0) return TQ_OK;
  if (prog->nload > TQ_MAX_LOADS) return set_error(TQ_EINVAL, "filter: too many load slots");
  const int64_t tiles = ceil_div(n, kTileRows);
  Conj C;
  if (match_conj(*prog, *cols, &C))
    return launch("k_filter_conj", k_filter_conj, (unsigned)tiles, kBlock, 0,
                  (hipStream_t)stream, C, n, mask_out, tile_counts);
  return launch("k_filter", k_filter, (unsigned)tiles, kBlock, stage_bytes(*prog),
                (hipStream_t)stream, *prog, *cols, n, mask_out, tile_counts);
}

extern "C" int tq_filter_is_conjunction(const tq_program_t* prog, const tq_colset_t* cols) {
  Conj C;
  if (!prog || !cols || prog->nseg < 1 || prog->n_instr > TQ_MAX_INSTR ||
      prog->seg_end[0] > prog->n_instr || prog->nload > TQ_MAX_LOADS)
    return 0;
  return match_conj(*prog, *cols, &C) ? 1 : 0;
}

extern "C" int tq_eval_column(const tq_program_t* prog, const tq_colset_t* cols, int64_t n,
                              void* out, int32_t out_bytes, void* stream) {
  if (!prog || !cols || (n > 0 && !out)) return set_error(TQ_EINVAL, "eval: null argument");
  if (prog->nseg < 1 || prog->n_instr > TQ_MAX_INSTR || prog->nload > TQ_MAX_LOADS)
    return set_error(TQ_EINVAL, "eval: malformed program");
  if (out_bytes != 1 && out_bytes != 2 && out_bytes != 4 && out_bytes != 8)
    return set_error(TQ_EINVAL, "eval: output width %d", out_bytes);
  if (n <= 0) return TQ_OK;
  int64_t g = ceil_div(n, kSubRows);
  if (g > 4096) g = 4096;
  return launch("k_eval", k_eval, (unsigned)g, kBlock, stage_bytes(*prog), (hipStream_t)stream, *prog, *cols, n, out, out_bytes);
}

extern "C" int tq_eval_columns(const tq_program_t* prog, const tq_colset_t* cols, int64_t n,
                               void* const* outs, const int32_t* out_bytes, void* stream) {
  if (!prog || !cols || !outs || !out_bytes) return set_error(TQ_EINVAL, "eval: null argument");
  if (prog->nseg < 1 || prog->nseg > TQ_MAX_SEGS || prog->n_instr > TQ_MAX_INSTR ||
      prog->nload > TQ_MAX_LOADS)
    return set_error(TQ_EINVAL, "eval: malformed program");
  OutSet O{};
  for (int a = 0; a < prog->nseg; ++a) {
    const int b = out_bytes[a];
    if (b != 1 && b != 2 && b != 4 && b != 8)
      return set_error(TQ_EINVAL, "eval: output %d width %d", a, b);
    if (n > 0 && !outs[a]) return set_error(TQ_EINVAL, "eval: null output %d", a);
    O.out[a] = outs[a];
    O.bytes[a] = b;
  }
  if (n <= 0) return TQ_OK;
  int64_t g = ceil_div(n, kSubRows);
  if (g > 4096) g = 4096;
  return launch("k_eval_multi", k_eval_multi, (unsigned)g, kBlock, stage_bytes(*prog),
                (hipStream_t)stream, *prog, *cols, n, O);
}

extern "C" int tq_mask_tile_counts(const uint8_t* mask, int64_t n, uint32_t* tile_counts,
                                   void* stream) {
  if (n <= 0) return TQ_OK;
  return launch("k_mask_counts", k_mask_counts, (unsigned)ceil_div(n, kTileRows), kBlock, 0, (hipStream_t)stream, mask, n, tile_counts);
}

extern "C" size_t tq_scan_ws_bytes(int64_t ntiles) {
  return align_up((size_t)(ceil_div(ntiles, kScanChunk) + 1) * sizeof(int64_t), 256);
}

extern "C" int tq_scan_tiles(const uint32_t* tile_counts, int64_t ntiles, int64_t* tile_offsets,
                             int64_t* total, void* ws, void* stream) {
  return scan_exclusive<uint32_t, int64_t>(tile_counts, ntiles, tile_offsets, total,
                                           (int64_t*)ws, (hipStream_t)stream);
}

extern "C" int tq_compact(const uint8_t* mask, int64_t n, const int64_t* tile_offsets,
                          int32_t ncols, const void* const* src, const int32_t* elem_bytes,
                          void* const* dst, void* stream) {
  ColPack cp;
  TQ_TRY(make_colpack(&cp, ncols, src, elem_bytes, nullptr, dst));
  if (n <= 0 || ncols == 0) return TQ_OK;
  return launch("k_compact", k_compact, (unsigned)ceil_div(n, kTileRows), kBlock, 0, (hipStream_t)stream, mask, n, tile_offsets, cp);
}

extern "C" int tq_mark_rows(const int64_t* idx, int64_t m, uint8_t* flags, int64_t n,
                             void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (n <= 0) return TQ_OK;
  TQ_TRY(launch("k_fill", k_fill<uint8_t>, stream_grid(n), kBlock, 0, s, flags, (uint8_t)0, n));
  if (m <= 0) return TQ_OK;
  return launch("k_mark", k_mark, stream_grid(m), kBlock, 0, s, idx, m, flags, n);
}

extern "C" int tq_gather(const int64_t* idx, int64_t m, int32_t ncols, const void* const* src,
                         const int32_t* elem_bytes, const int32_t* float_kind, void* const* dst,
                         int32_t null_fill, void* stream) {
  ColPack cp;
  TQ_TRY(make_colpack(&cp, ncols, src, elem_bytes, float_kind, dst));
  if (m <= 0 || ncols == 0) return TQ_OK;
  return launch("k_gather", k_gather, stream_grid(m), kBlock, 0, (hipStream_t)stream, idx, m, cp, null_fill);
}
