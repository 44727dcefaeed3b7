// LSD radix sort of (key, row id) pairs and the multi-key sort / top-k built on it.
//
// One pass sorts by an 8-bit digit in three launches:
//   hist    - per 4096-key tile, a digit histogram (per-wave LDS counters), stored digit-major
//             so that one exclusive scan yields every (digit, tile) output offset;
//   scan    - device-wide exclusive scan of the 256 x tiles histogram;
//   scatter - each tile ranks its keys stably: wave w owns the contiguous quarter
//             [w*1024, (w+1)*1024) of the tile and walks it 64 keys at a time; lanes holding
//             the same digit find each other with 8 ballots, the lowest of them bumps that
//             digit's per-wave counter, and each lane's rank is the counter plus its mbcnt
//             within the peer mask.  The tile is then reordered in LDS and written out in runs
//             of equal digits, so global writes are coalesced.
// Passes cover only the bit range in which keys differ (found by one OR/AND reduction), and
// keys are order-preserving transforms of the typed columns, so unsigned order = value order.
#include "tq_common.h"
#include "tq_sort.h"

namespace tq {

constexpr int kBins = 256;
constexpr int kSortItems = 16;
constexpr int kSortTile = kBlock * kSortItems;  // 4096
constexpr int kWaveSeg = kSortTile / kWaves;    // 1024 keys per wave

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const K* __restrict__ keys, int64_t n,
                                                       int shift, uint32_t* __restrict__ hist,
                                                       int64_t ntiles) {
  __shared__ uint32_t cnt[kWaves][kBins];
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < kWaves * kBins; i += kBlock) (&cnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  K k[kSortItems];  // every load in flight before the first atomic (rows past n read row 0)
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = tile0 + (int64_t)j * kBlock + t;
    k[j] = keys[e < n ? e : 0];
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = tile0 + (int64_t)j * kBlock + t;
    if (e < n) atomicAdd(&cnt[w][(uint32_t)(k[j] >> shift) & (kBins - 1)], 1u);
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int v = 0; v < kWaves; ++v) s += cnt[v][t];
  hist[(int64_t)t * ntiles + blockIdx.x] = s;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const K* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n, int shift,
    const uint32_t* __restrict__ offs, int64_t ntiles, K* __restrict__ kout,
    uint32_t* __restrict__ vout) {
  __shared__ uint32_t wcnt[kWaves][kBins];
  __shared__ uint32_t lstart[kBins];
  __shared__ uint32_t gbase[kBins];
  __shared__ uint32_t shs[kWaves];
  __shared__ K skey[kSortTile];
  __shared__ uint32_t sval[kSortTile];
  const int t = threadIdx.x, w = t >> 6;
  for (int i = t; i < kWaves * kBins; i += kBlock) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const int64_t tile0 = (int64_t)blockIdx.x * kSortTile;
  const int64_t wave0 = tile0 + (int64_t)w * kWaveSeg;
  const int l = lane_id();
  K key[kSortItems];
  uint32_t val[kSortItems], dig[kSortItems], rnk[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {  // every load in flight first (rows past n read row 0)
    const int64_t e = wave0 + (int64_t)j * 64 + l;
    key[j] = kin[e < n ? e : 0];
    val[j] = vin[e < n ? e : 0];
  }
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = wave0 + (int64_t)j * 64 + l;
    const bool valid = e < n;
    if (!valid) {
      key[j] = (K)0;
      val[j] = 0u;
    }
    const uint32_t d = (uint32_t)(key[j] >> shift) & (kBins - 1);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1;
      const uint64_t m = ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    const int below = rank_below(peers);
    uint32_t base = 0;
    if (valid) base = wcnt[w][d];
    if (valid && below == 0) wcnt[w][d] = base + (uint32_t)__popcll(peers);
    dig[j] = d;
    rnk[j] = base + below;
  }
  __syncthreads();
  {  // thread t owns digit t: exclusive prefix over waves, then over digits
    uint32_t run = 0;
#pragma unroll
    for (int v = 0; v < kWaves; ++v) {
      const uint32_t c = wcnt[v][t];
      wcnt[v][t] = run;
      run += c;
    }
    uint32_t tot;
    lstart[t] = block_exclusive_sum(run, shs, &tot);
    gbase[t] = offs[(int64_t)t * ntiles + blockIdx.x];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t e = wave0 + (int64_t)j * 64 + l;
    if (e < n) {
      const uint32_t lp = lstart[dig[j]] + wcnt[w][dig[j]] + rnk[j];
      skey[lp] = key[j];
      sval[lp] = val[j];
    }
  }
  __syncthreads();
  const int64_t rem = n - tile0;
  const int nvalid = rem < kSortTile ? (int)rem : kSortTile;
  for (int i = t; i < nvalid; i += kBlock) {
    const K k = skey[i];
    const uint32_t d = (uint32_t)(k >> shift) & (kBins - 1);
    const uint32_t gp = gbase[d] + (uint32_t)i - lstart[d];
    kout[gp] = k;
    vout[gp] = sval[i];
  }
}

// OR, AND, min and max of n keys: per-thread, then per-workgroup through LDS, then one atomic
// per statistic and workgroup (the grid is kept small, so the atomics do not contend).
template <typename K>
__global__ __launch_bounds__(kBlock) void k_key_bits(const K* __restrict__ keys, int64_t n,
                                                     unsigned long long* __restrict__ stats) {
  __shared__ uint64_t sh[4][kWaves];
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  uint64_t o = 0, a = ~0ull, lo = ~0ull, hi = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint64_t k = (uint64_t)keys[i];
    o |= k;
    a &= k;
    lo = k < lo ? k : lo;
    hi = k > hi ? k : hi;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    o |= __shfl_xor(o, d, 64);
    a &= __shfl_xor(a, d, 64);
    const uint64_t l2 = __shfl_xor(lo, d, 64), h2 = __shfl_xor(hi, d, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
  }
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sh[0][w] = o; sh[1][w] = a; sh[2][w] = lo; sh[3][w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 1; i < kWaves; ++i) {
      o |= sh[0][i];
      a &= sh[1][i];
      lo = sh[2][i] < lo ? sh[2][i] : lo;
      hi = sh[3][i] > hi ? sh[3][i] : hi;
    }
    if (lo <= hi) {
      atomicOr(&stats[0], (unsigned long long)o);
      atomicAnd(&stats[1], (unsigned long long)a);
      atomicMin(&stats[2], (unsigned long long)lo);
      atomicMax(&stats[3], (unsigned long long)hi);
    }
  }
}

// ------------------------------------------------------------------ host driver
size_t radix_ws_bytes(int64_t n) {
  const int64_t ntiles = ceil_div(n > 0 ? n : 1, kSortTile);
  const int64_t h = kBins * ntiles;
  return align_up((size_t)h * 4, 256) * 2 + align_up((size_t)(ceil_div(h, kSortTile) + 1) * 4, 256) + 256;
}

template <typename K>
int radix_sort_pairs(K* keys, uint32_t* vals, K* keys_alt, uint32_t* vals_alt, int64_t n,
                     int lo_bit, int hi_bit, void* ws, hipStream_t s, bool* in_alt) {
  *in_alt = false;
  if (n <= 1 || hi_bit <= lo_bit) return TQ_OK;
  const int64_t ntiles = ceil_div(n, kSortTile);
  const int64_t h = kBins * ntiles;
  uint32_t* hist = (uint32_t*)ws;
  uint32_t* offs = (uint32_t*)((char*)ws + align_up((size_t)h * 4, 256));
  uint32_t* sws = (uint32_t*)((char*)offs + align_up((size_t)h * 4, 256));
  K *ki = keys, *ko = keys_alt;
  uint32_t *vi = vals, *vo = vals_alt;
  for (int shift = lo_bit; shift < hi_bit; shift += 8) {
    TQ_TRY(launch("k_radix_hist", k_radix_hist<K>, (unsigned)ntiles, kBlock, 0, s, ki, n, shift, hist, ntiles));
    TQ_TRY((scan_exclusive<uint32_t, uint32_t>(hist, h, offs, nullptr, sws, s)));
    TQ_TRY(launch("k_radix_scatter", k_radix_scatter<K>, (unsigned)ntiles, kBlock, 0, s, ki, vi, n, shift, offs, ntiles, ko, vo));
    K* tk = ki; ki = ko; ko = tk;
    uint32_t* tv = vi; vi = vo; vo = tv;
    *in_alt = !*in_alt;
  }
  return TQ_OK;
}

template int radix_sort_pairs<uint32_t>(uint32_t*, uint32_t*, uint32_t*, uint32_t*, int64_t, int,
                                        int, void*, hipStream_t, bool*);
template int radix_sort_pairs<uint64_t>(uint64_t*, uint32_t*, uint64_t*, uint32_t*, int64_t, int,
                                        int, void*, hipStream_t, bool*);

// Pinned host staging for the few values a host decision needs.
int64_t* host_scratch() {
  static thread_local int64_t* p = nullptr;
  if (!p) {
    if (hipHostMalloc((void**)&p, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess) p = nullptr;
  }
  return p;
}

int read_back2(const void* dev0, size_t bytes0, int64_t* host0, const void* dev1, size_t bytes1,
               int64_t* host1, hipStream_t s) {
  if (bytes0 > 256 || bytes1 > 256) return set_error(TQ_EINVAL, "read_back2: copy too large");
  const void* dev[2] = {dev0, dev1};
  const size_t bytes[2] = {bytes0, bytes1};
  int64_t h[64];
  TQ_TRY(read_back_n(2, dev, bytes, h, s));
  memcpy(host0, h, bytes0);
  memcpy(host1, (const char*)h + align_up(bytes0, 8), bytes1);
  return TQ_OK;
}

// Up to 8 runs of 8-byte words gathered by one small kernel straight into host memory.
constexpr int kMaxWordRuns = 8;
struct WordRuns {
  const uint64_t* src[kMaxWordRuns];
  int32_t nw[kMaxWordRuns];
  int32_t off[kMaxWordRuns];
  int32_t n;
};

__global__ __launch_bounds__(64) void k_collect_words(WordRuns W, uint64_t* __restrict__ dst) {
  for (int p = 0; p < W.n; ++p)
    for (int w = threadIdx.x; w < W.nw[p]; w += 64) dst[W.off[p] + w] = W.src[p][w];
}

// Pinned, coherent host words that kernels write directly (device view in *dev).
static int64_t* host_words(uint64_t** dev) {
  static thread_local int64_t* p = nullptr;
  static thread_local uint64_t* d = nullptr;
  if (!p) {
    if (hipHostMalloc((void**)&p, 64 * sizeof(int64_t), hipHostMallocCoherent) != hipSuccess) {
      p = nullptr;
      return nullptr;
    }
    if (hipHostGetDevicePointer((void**)&d, p, 0) != hipSuccess) d = nullptr;
  }
  *dev = d;
  return d ? p : nullptr;
}

int read_back_n(int n, const void* const* dev, const size_t* bytes, int64_t* host_out,
                hipStream_t s) {
  size_t off = 0;
  bool words = n <= kMaxWordRuns;
  for (int i = 0; i < n; ++i) {
    off += align_up(bytes[i], 8);
    words = words && bytes[i] % 8 == 0 && ((uintptr_t)dev[i] & 7) == 0;
  }
  if (off > 512) return set_error(TQ_EINVAL, "read_back_n: %zu bytes", off);
  uint64_t* hd = nullptr;
  int64_t* hw = words ? host_words(&hd) : nullptr;
  if (hw) {
    // one launch instead of one copy per run
    WordRuns W{};
    W.n = n;
    int32_t o = 0;
    for (int i = 0; i < n; ++i) {
      W.src[i] = (const uint64_t*)dev[i];
      W.nw[i] = (int32_t)(bytes[i] / 8);
      W.off[i] = o;
      o += W.nw[i];
    }
    TQ_TRY(launch("k_collect_words", k_collect_words, 1, 64, 0, s, W, hd));
    TQ_TRY_HIP(hipStreamSynchronize(s));
    memcpy(host_out, hw, off);
    return TQ_OK;
  }
  int64_t* h = host_scratch();
  if (!h) return set_error(TQ_EHIP, "pinned staging allocation failed");
  off = 0;
  for (int i = 0; i < n; ++i) {
    TQ_TRY_HIP(hipMemcpyAsync((char*)h + off, dev[i], bytes[i], hipMemcpyDeviceToHost, s));
    off += align_up(bytes[i], 8);
  }
  TQ_TRY_HIP(hipStreamSynchronize(s));
  memcpy(host_out, h, off);
  return TQ_OK;
}

int read_back(const void* dev, size_t bytes, int64_t* host_out, hipStream_t s) {
  int64_t* h = host_scratch();
  if (!h) return set_error(TQ_EHIP, "pinned staging allocation failed");
  TQ_TRY_HIP(hipMemcpyAsync(h, dev, bytes, hipMemcpyDeviceToHost, s));
  TQ_TRY_HIP(hipStreamSynchronize(s));
  memcpy(host_out, h, bytes);
  return TQ_OK;
}

template <typename K>
int key_bit_range(const K* keys, int64_t n, void* stats_dev, hipStream_t s, int* lo, int* hi,
                  uint64_t* kmin, uint64_t* kmax) {
  *lo = *hi = 0;
  if (n <= 0) return TQ_OK;
  unsigned long long* st = (unsigned long long*)stats_dev;
  TQ_TRY(fill_u64(st, 0ull, 1, s));
  TQ_TRY(fill_u64(st + 1, ~0ull, 2, s));
  TQ_TRY(fill_u64(st + 3, 0ull, 1, s));
  int64_t g = ceil_div(n, kBlock * 8);
  if (g > 512) g = 512;
  TQ_TRY(launch("k_key_bits", k_key_bits<K>, (unsigned)g, kBlock, 0, s, keys, n, st));
  int64_t hv[4];
  TQ_TRY(read_back(st, 32, hv, s));
  if (kmin) *kmin = (uint64_t)hv[2];
  if (kmax) *kmax = (uint64_t)hv[3];
  radix_bit_range((uint64_t)hv[0], (uint64_t)hv[1], lo, hi);
  return TQ_OK;
}

void radix_bit_range(uint64_t key_or, uint64_t key_and, int* lo, int* hi) {
  const uint64_t diff = key_or ^ key_and;
  *lo = *hi = 0;
  if (diff == 0) return;
  const int low = __builtin_ctzll(diff), high = 64 - __builtin_clzll(diff);
  *lo = (low / 8) * 8;
  *hi = high;
}

template int key_bit_range<uint32_t>(const uint32_t*, int64_t, void*, hipStream_t, int*, int*,
                                     uint64_t*, uint64_t*);
template int key_bit_range<uint64_t>(const uint64_t*, int64_t, void*, hipStream_t, int*, int*,
                                     uint64_t*, uint64_t*);

// ------------------------------------------------------------------ multi-key sort
// A field is one typed scalar of a key (a 2-D key contributes one int8 field per column).
struct Field {
  const void* ptr;
  int32_t dtype;
  int32_t width;  // row stride in elements (2-D keys), 1 otherwise
  int32_t col;    // column within a 2-D key
  int32_t desc;
  int32_t bits;
  int32_t shift;  // position inside the packed key
};

struct FieldPack {
  Field f[16];
  int32_t n;
};

__device__ __forceinline__ uint64_t field_bits(const Field& f, int64_t r) {
  uint64_t v;
  switch (f.dtype) {
    case TQ_BOOL: v = ((const uint8_t*)f.ptr)[r] != 0; break;
    case TQ_I8: v = (uint8_t)(((const int8_t*)f.ptr)[r * f.width + f.col]) ^ 0x80u; break;
    case TQ_I16: v = (uint16_t)((const int16_t*)f.ptr)[r] ^ 0x8000u; break;
    case TQ_I32: v = (uint32_t)((const int32_t*)f.ptr)[r] ^ 0x80000000u; break;
    case TQ_I64: v = (uint64_t)((const int64_t*)f.ptr)[r] ^ 0x8000000000000000ull; break;
    case TQ_F32: v = f32_order_bits(((const float*)f.ptr)[r]); break;
    default: v = f64_order_bits(((const double*)f.ptr)[r]); break;
  }
  if (f.desc) v = ~v;
  if (f.bits < 64) v &= (1ull << f.bits) - 1;
  return v << f.shift;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_pack_keys(FieldPack fp, const uint32_t* __restrict__ perm,
                                                      int64_t n, K* __restrict__ keys,
                                                      uint32_t* __restrict__ vals) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    const uint32_t r = perm ? perm[i] : (uint32_t)i;
    uint64_t k = 0;
    for (int f = 0; f < fp.n; ++f) k |= field_bits(fp.f[f], r);
    keys[i] = (K)k;
    vals[i] = r;
  }
}

// ------------------------------------------------------------------ top-k
// The k smallest (key, row) pairs of a tile of up to 4096 candidates, without sorting the tile:
// a radix select in LDS finds the k-th smallest pair T, digit by digit from the most
// significant (8 key bytes, then 4 row bytes; each pass histograms the candidates whose higher
// digits equal T's digits found so far, and a pass whose bin holds exactly the remaining rank
// ends the search).  Every real pair <= T is kept, at most k; the tile's k output slots are
// filled up with sentinels (all-ones key and row).  Rounds of this kernel shrink any candidate
// set to one tile without a host synchronisation; the last round also ranks its k pairs and
// writes their rows, smallest first, as the permutation.  Rows are distinct, so ordering by
// (key, row) is the stable order of the keys.
constexpr int kTopTile = 4096;
constexpr int kTopPer = kTopTile / kBlock;

__device__ __forceinline__ uint32_t digit_of(uint64_t key, uint32_t row, int d) {
  return d >= 4 ? (uint32_t)(key >> (8 * (d - 4))) & 0xFFu : (row >> (8 * d)) & 0xFFu;
}

__global__ __launch_bounds__(kBlock) void k_topk_select(const uint64_t* __restrict__ kin,
                                                        const uint32_t* __restrict__ vin,
                                                        int64_t n, int k,
                                                        uint64_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout,
                                                        int64_t* __restrict__ perm_out,
                                                        int64_t nout) {
  __shared__ uint32_t hist[256];
  __shared__ uint32_t scan_sh[kWaves];
  __shared__ uint32_t found[3];  // digit, remaining rank, bin count
  __shared__ uint32_t nsel;
  __shared__ uint64_t fk[256];
  __shared__ uint32_t fv[256];
  const int t = threadIdx.x, l = lane_id();
  const int64_t base = (int64_t)blockIdx.x * kTopTile;
  uint64_t key[kTopPer];
  uint32_t row[kTopPer];
#pragma unroll
  for (int j = 0; j < kTopPer; ++j) {
    const int64_t i = base + (int64_t)j * kBlock + t;
    const bool v = i < n;
    key[j] = v ? kin[i] : ~0ull;
    row[j] = v ? (vin ? vin[i] : (uint32_t)i) : 0xFFFFFFFFu;
  }
  uint64_t pk = 0, mk = 0;  // digits of T found so far (key part) and their mask
  uint32_t pr = 0, mr = 0;  // (row part)
  uint32_t kk = (uint32_t)k;
  for (int d = 11; d >= 0; --d) {
    hist[t] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTopPer; ++j) {
      const bool cand = (key[j] & mk) == pk && (row[j] & mr) == pr;
      const uint32_t dg = digit_of(key[j], row[j], d);
      // lanes holding the same digit add once, through their lowest lane; waves without a
      // candidate (most of them once the first digits are fixed) skip the matching
      uint64_t peers = ballot(cand);
      if (peers == 0) continue;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const uint64_t m = ballot(cand && ((dg >> b) & 1u));
        peers &= ((dg >> b) & 1u) ? m : ~m;
      }
      if (cand && (int)__builtin_ctzll(peers) == l) atomicAdd(&hist[dg], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    const uint32_t h = hist[t];
    uint32_t tot;
    const uint32_t ex = block_exclusive_sum(h, scan_sh, &tot);
    if (ex < kk && kk <= ex + h) {
      found[0] = (uint32_t)t;
      found[1] = kk - ex;
      found[2] = h;
    }
    __syncthreads();
    const uint32_t dg = found[0];
    kk = found[1];
    const bool whole = found[2] == kk;
    if (d >= 4) {
      pk |= (uint64_t)dg << (8 * (d - 4));
      mk |= 0xFFull << (8 * (d - 4));
    } else {
      pr |= dg << (8 * d);
      mr |= 0xFFu << (8 * d);
    }
    __syncthreads();
    if (whole) break;
  }
  // keep the real pairs whose digits so far are <= T's
  if (t == 0) nsel = 0;
  __syncthreads();
  const int64_t obase = (int64_t)blockIdx.x * k;
#pragma unroll
  for (int j = 0; j < kTopPer; ++j) {
    const uint64_t km = key[j] & mk;
    const bool keep = row[j] != 0xFFFFFFFFu &&
                      (km < pk || (km == pk && (row[j] & mr) <= pr));
    if (keep) {
      const uint32_t slot = atomicAdd(&nsel, 1u);
      if (perm_out) {
        fk[slot] = key[j];
        fv[slot] = row[j];
      } else {
        kout[obase + slot] = key[j];
        vout[obase + slot] = row[j];
      }
    }
  }
  __syncthreads();
  const uint32_t m = nsel;
  if (!perm_out) {
    for (int i = (int)m + t; i < k; i += kBlock) {
      kout[obase + i] = ~0ull;
      vout[obase + i] = 0xFFFFFFFFu;
    }
    return;
  }
  // last round: rank the kept pairs by (key, row)
  for (int i = t; i < (int)m; i += kBlock) {
    const uint64_t a = fk[i];
    const uint32_t va = fv[i];
    int rank = 0;
    for (int j = 0; j < (int)m; ++j) rank += (fk[j] < a || (fk[j] == a && fv[j] < va)) ? 1 : 0;
    if (rank < nout) perm_out[rank] = va;
  }
}

// ------------------------------------------------------------------ small top-k (k <= 64)
// A workgroup reduces 4096 candidates (16 per thread) to its k smallest (key, row) pairs, in
// order, without sorting them:
//   1. every thread takes the smallest of its 16, and each wave sorts its 64 minima across its
//      lanes (bitonic network of shuffles);
//   2. T = the k-th smallest of the 256 minima: lane l < k of a wave ranks its minimum by l plus
//      a binary search in each other wave's first k.  k minima are <= T, so the workgroup's k
//      smallest are all <= T, and only the (at most k) threads whose minimum is <= T hold pairs
//      <= T: at most 16k survivors;
//   3. the survivors go to LDS (one atomic per wave) and each one's rank among them, by
//      counting, is its output slot.
// A round turns m candidates into k per workgroup; the last round (one workgroup) writes the
// rows of its first nout pairs as the permutation.  Rows are distinct, so (key, row) order is
// the stable key order and every rank is unique.
constexpr int kSmallK = 64;
constexpr int kListPer = 16;                    // candidates per thread and round
constexpr int kListTile = kBlock * kListPer;    // candidates per workgroup and round
constexpr int kSurvMax = kSmallK * kListPer;    // survivors of step 2

__device__ __forceinline__ bool pair_less(uint64_t ka, uint32_t ra, uint64_t kb, uint32_t rb) {
  return ka < kb || (ka == kb && ra < rb);
}

// Order bits of one key field for rows base + j*kBlock (j < kListPer), ORed into key[]: the
// dtype is dispatched once, then all loads are issued before any is used (rows past the end
// read row base instead and are dropped by the caller).
template <typename T, typename Bits>
__device__ __forceinline__ void or_field(const Field& F, int64_t base, uint32_t valid,
                                         uint64_t key[kListPer], Bits bits) {
  const T* p = (const T*)F.ptr;
  T x[kListPer];
#pragma unroll
  for (int j = 0; j < kListPer; ++j) x[j] = p[((valid >> j) & 1u) ? base + (int64_t)j * kBlock : base];
#pragma unroll
  for (int j = 0; j < kListPer; ++j) {
    uint64_t v = bits(x[j]);
    if (F.desc) v = ~v;
    if (F.bits < 64) v &= (1ull << F.bits) - 1;
    key[j] |= v << F.shift;
  }
}

// Sort key bits of rows base + j*kBlock (j < kListPer) from the key fields (1-D keys); rows at
// or past n are flagged invalid through *valid.
__device__ __forceinline__ void field_keys(const FieldPack& fp, int64_t base, int64_t n,
                                           uint64_t key[kListPer], uint32_t* valid) {
  uint32_t v = 0;
#pragma unroll
  for (int j = 0; j < kListPer; ++j) {
    key[j] = 0;
    v |= (base + (int64_t)j * kBlock < n ? 1u : 0u) << j;
  }
  *valid = v;
  if (!v) return;
  for (int f = 0; f < fp.n; ++f) {
    const Field F = fp.f[f];
    switch (F.dtype) {
      case TQ_BOOL:
        or_field<uint8_t>(F, base, v, key, [](uint8_t x) { return (uint64_t)(x != 0); });
        break;
      case TQ_I8:
        or_field<int8_t>(F, base, v, key, [](int8_t x) { return (uint64_t)((uint8_t)x ^ 0x80u); });
        break;
      case TQ_I16:
        or_field<int16_t>(F, base, v, key,
                          [](int16_t x) { return (uint64_t)((uint16_t)x ^ 0x8000u); });
        break;
      case TQ_I32:
        or_field<int32_t>(F, base, v, key,
                          [](int32_t x) { return (uint64_t)((uint32_t)x ^ 0x80000000u); });
        break;
      case TQ_I64:
        or_field<int64_t>(F, base, v, key,
                          [](int64_t x) { return (uint64_t)x ^ 0x8000000000000000ull; });
        break;
      case TQ_F32:
        or_field<float>(F, base, v, key, [](float x) { return (uint64_t)f32_order_bits(x); });
        break;
      default:
        or_field<double>(F, base, v, key, [](double x) { return f64_order_bits(x); });
        break;
    }
  }
}

// Row r of every column of gp into row i of its output (the last round's gather).
__device__ __forceinline__ void gather_row(const ColPack& gp, int64_t i, int64_t r) {
  for (int c = 0; c < gp.n; ++c) copy_row(gp.dst[c], i, gp.src[c], r, gp.bytes[c]);
}

// One round.  Source: the key fields of rows 0..m-1 (kin == null) or m (key, row) pairs; with
// n_dev (key fields only) the rows are the first min(m, *n_dev), a count a producer earlier on
// the stream left in device memory.  Writes the workgroup's k smallest in order to
// kout/vout[blockIdx.x*k ..] (sentinels when it has fewer), or, in the last round (perm_out !=
// null, one workgroup), the rows of the first nout as the permutation, and those rows of the
// gp columns into row 0..nout-1 of their outputs; permutation entries past the candidates
// (only when n_dev cut the rows below nout) get row 0.
__global__ __launch_bounds__(kBlock) void k_topk_block(FieldPack fp, const uint64_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin, int64_t m,
                                                       const int64_t* __restrict__ n_dev,
                                                       int k, uint64_t* __restrict__ kout,
                                                       uint32_t* __restrict__ vout,
                                                       int64_t* __restrict__ perm_out, int64_t nout,
                                                       ColPack gp) {
  // (key, row) pairs as 16-byte LDS entries: one broadcast read per pair in the counting loops
  __shared__ ulonglong2 mins[kBlock];
  __shared__ ulonglong2 surv[kSurvMax];
  __shared__ uint64_t tk;
  __shared__ uint32_t tr, ns;
  const int t = threadIdx.x;
  const int64_t base = (int64_t)blockIdx.x * kListTile + t;
  if (n_dev && !kin) {
    const int64_t c = *n_dev;
    m = c < m ? (c > 0 ? c : 0) : m;
  }
  if ((int64_t)blockIdx.x * kListTile >= m) {
    // no candidate in this workgroup (a grid sized for a capacity): only the sentinels, or
    // the zero permutation, that the full pass would write with no survivors
    if (perm_out) {
      for (int64_t i = t; i < nout; i += kBlock) {
        perm_out[i] = 0;
        gather_row(gp, i, 0);
      }
    } else {
      for (int i = t; i < k; i += kBlock) {
        kout[(int64_t)blockIdx.x * k + i] = ~0ull;
        vout[(int64_t)blockIdx.x * k + i] = 0xFFFFFFFFu;
      }
    }
    return;
  }
  uint64_t key[kListPer];
  uint32_t row[kListPer];
  uint32_t valid = 0;
  if (kin) {
#pragma unroll
    for (int j = 0; j < kListPer; ++j) {
      const int64_t i = base + (int64_t)j * kBlock;
      const bool ok = i < m;
      valid |= (ok ? 1u : 0u) << j;
      key[j] = ok ? kin[i] : ~0ull;
      row[j] = ok ? vin[i] : 0xFFFFFFFFu;
    }
  } else {
    field_keys(fp, base, m, key, &valid);
#pragma unroll
    for (int j = 0; j < kListPer; ++j) row[j] = (uint32_t)(base + (int64_t)j * kBlock);
  }
#pragma unroll
  for (int j = 0; j < kListPer; ++j) {  // invalid candidates become sentinels
    if (!((valid >> j) & 1u)) {
      key[j] = ~0ull;
      row[j] = 0xFFFFFFFFu;
    }
  }
  // 1. the thread's minimum, then the wave's 64 minima sorted across lanes
  const int l = lane_id(), w = t >> 6;
  uint64_t ak = key[0];
  uint32_t ar = row[0];
#pragma unroll
  for (int j = 1; j < kListPer; ++j) {
    if (pair_less(key[j], row[j], ak, ar)) {
      ak = key[j];
      ar = row[j];
    }
  }
  uint64_t sk_ = ak;
  uint32_t sr_ = ar;
#pragma unroll
  for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
    for (int j = size >> 1; j > 0; j >>= 1) {
      const uint64_t ok = __shfl_xor(sk_, j, 64);
      const uint32_t orr = __shfl_xor(sr_, j, 64);
      const bool up = (l & size) == 0;          // this pair of the stage sorts ascending
      const bool low = (l & j) == 0;            // this lane keeps the smaller of the pair
      const bool other_less = pair_less(ok, orr, sk_, sr_);
      if (other_less == (low == up)) {
        sk_ = ok;
        sr_ = orr;
      }
    }
  }
  mins[t] = make_ulonglong2(sk_, sr_);  // mins[w * 64 + l]: wave w's l-th smallest
  if (t == 0) {
    tk = ~0ull;
    tr = 0xFFFFFFFFu;
    ns = 0;
  }
  __syncthreads();
  // 2. T (stays the sentinel when fewer than k threads hold a real candidate)
  if (l < k && sr_ != 0xFFFFFFFFu) {
    int rank = l;
#pragma unroll
    for (int o = 1; o < kWaves; ++o) {
      const ulonglong2* L = mins + ((w + o) & (kWaves - 1)) * 64;
      int lo = 0, hi = k;  // entries of L[0, k) less than (sk_, sr_)
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        const ulonglong2 e = L[mid];
        if (pair_less(e.x, (uint32_t)e.y, sk_, sr_)) lo = mid + 1;
        else hi = mid;
      }
      rank += lo;
    }
    if (rank == k - 1) {
      tk = sk_;
      tr = sr_;
    }
  }
  __syncthreads();
  const uint64_t Tk = tk;
  const uint32_t Tr = tr;
  // 3. survivors <= T into LDS (a lane's survivors after those of lower lanes; one atomic per
  //    wave), then each one's rank among them
  uint32_t keep = 0;
#pragma unroll
  for (int j = 0; j < kListPer; ++j)
    keep |= (row[j] != 0xFFFFFFFFu && !pair_less(Tk, Tr, key[j], row[j]) ? 1u : 0u) << j;
  const uint32_t c = (uint32_t)__popc(keep);
  const uint32_t inc = wave_inclusive_sum(c);
  uint32_t wbase = 0;
  if (l == 63 && inc) wbase = atomicAdd(&ns, inc);
  wbase = __shfl(wbase, 63, 64);
  uint32_t p = wbase + inc - c;
#pragma unroll
  for (int j = 0; j < kListPer; ++j) {
    if (((keep >> j) & 1u) && p < kSurvMax) surv[p++] = make_ulonglong2(key[j], row[j]);
  }
  __syncthreads();
  const int s = (int)ns < kSurvMax ? (int)ns : kSurvMax;  // at most 16k by construction
  for (int i = t; i < s; i += kBlock) {
    const ulonglong2 x = surv[i];
    const uint64_t xk = x.x;
    const uint32_t xr = (uint32_t)x.y;
    int rank = 0;
#pragma unroll 8
    for (int q = 0; q < s; ++q) {
      const ulonglong2 e = surv[q];
      rank += pair_less(e.x, (uint32_t)e.y, xk, xr) ? 1 : 0;
    }
    if (perm_out) {
      if (rank < nout) {
        perm_out[rank] = xr;
        gather_row(gp, rank, xr);
      }
    } else if (rank < k) {
      kout[(int64_t)blockIdx.x * k + rank] = xk;
      vout[(int64_t)blockIdx.x * k + rank] = xr;
    }
  }
  if (!perm_out) {
    for (int i = s + t; i < k; i += kBlock) {
      kout[(int64_t)blockIdx.x * k + i] = ~0ull;
      vout[(int64_t)blockIdx.x * k + i] = 0xFFFFFFFFu;
    }
  } else {
    for (int64_t i = s + t; i < nout; i += kBlock) {
      perm_out[i] = 0;
      gather_row(gp, i, 0);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_perm_out(const uint32_t* __restrict__ v, int64_t m,
                                                     int64_t* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += stride) out[i] = v[i];
}

static int bits_of(int dtype) {
  switch (dtype) {
    case TQ_BOOL: case TQ_I8: return 8;
    case TQ_I16: return 16;
    case TQ_I32: case TQ_F32: return 32;
    default: return 64;
  }
}

struct SortWs {
  uint64_t* k0;
  uint64_t* k1;
  uint32_t* v0;
  uint32_t* v1;
  void* stats;
  void* radix;
};

static SortWs carve_sort_ws(void* ws, int64_t n) {
  char* p = (char*)ws;
  SortWs w;
  const size_t kb = align_up((size_t)n * 8, 256), vb = align_up((size_t)n * 4, 256);
  w.k0 = (uint64_t*)p; p += kb;
  w.k1 = (uint64_t*)p; p += kb;
  w.v0 = (uint32_t*)p; p += vb;
  w.v1 = (uint32_t*)p; p += vb;
  w.stats = p; p += 256;
  w.radix = p;
  return w;
}

size_t sort_ws_bytes(int64_t n) {
  if (n < 1) n = 1;
  return 2 * align_up((size_t)n * 8, 256) + 2 * align_up((size_t)n * 4, 256) + 256 +
         radix_ws_bytes(n);
}

// Sort one packed group of fields stably, starting from permutation `perm` (null = identity);
// leaves the new permutation in *perm_out (one of the workspace value buffers).
static int sort_group(const FieldPack& fp, int total_bits, const uint32_t* perm, int64_t n,
                      SortWs& w, hipStream_t s, uint32_t** perm_out) {
  unsigned g = stream_grid(n);
  if (total_bits <= 32) {
    uint32_t* k0 = (uint32_t*)w.k0;
    uint32_t* k1 = (uint32_t*)w.k1;
    uint32_t* v0 = perm == w.v0 ? w.v1 : w.v0;
    uint32_t* v1 = perm == w.v0 ? w.v0 : w.v1;
    TQ_TRY(launch("k_pack_keys", k_pack_keys<uint32_t>, g, kBlock, 0, s, fp, perm, n, k0, v0));
    int lo, hi;
    TQ_TRY(key_bit_range<uint32_t>(k0, n, w.stats, s, &lo, &hi));
    bool alt;
    TQ_TRY(radix_sort_pairs<uint32_t>(k0, v0, k1, v1, n, lo, hi, w.radix, s, &alt));
    *perm_out = alt ? v1 : v0;
  } else {
    uint32_t* v0 = perm == w.v0 ? w.v1 : w.v0;
    uint32_t* v1 = perm == w.v0 ? w.v0 : w.v1;
    TQ_TRY(launch("k_pack_keys", k_pack_keys<uint64_t>, g, kBlock, 0, s, fp, perm, n, w.k0, v0));
    int lo, hi;
    TQ_TRY(key_bit_range<uint64_t>(w.k0, n, w.stats, s, &lo, &hi));
    bool alt;
    TQ_TRY(radix_sort_pairs<uint64_t>(w.k0, v0, w.k1, v1, n, lo, hi, w.radix, s, &alt));
    *perm_out = alt ? v1 : v0;
  }
  return TQ_OK;
}

int sort_perm_u32(const tq_sortkey_t* keys, int32_t nkeys, int64_t n, void* ws, size_t ws_bytes,
                  hipStream_t s, uint32_t** perm_out) {
  if (ws_bytes < sort_ws_bytes(n)) return set_error(TQ_ECAPACITY, "sort workspace too small");
  if (n >= (int64_t)1 << 32) return set_error(TQ_EINVAL, "sort of %lld rows exceeds 2^32", (long long)n);
  // Expand keys into fields, most significant first.
  Field fields[64];
  int nf = 0;
  for (int k = 0; k < nkeys; ++k) {
    const tq_sortkey_t& sk = keys[k];
    const int w = sk.width > 0 ? sk.width : 1;
    if (w > 1 && sk.dtype != TQ_I8 && sk.dtype != TQ_BOOL)
      return set_error(TQ_EUNSUPPORTED, "2-D sort keys must be int8");
    for (int c = 0; c < w; ++c) {
      if (nf >= 64) return set_error(TQ_EUNSUPPORTED, "more than 64 sort key fields");
      fields[nf++] = Field{sk.ptr, sk.dtype, w, c, sk.descending ? 1 : 0, bits_of(sk.dtype), 0};
    }
  }
  SortWs w = carve_sort_ws(ws, n);
  uint32_t* perm = nullptr;
  // Group fields from the least significant end into packed keys of <= 64 bits; sort the
  // groups least significant first (LSD across groups, each group a stable sort).
  int end = nf;
  while (end > 0) {
    int begin = end, bits = 0;
    while (begin > 0 && bits + fields[begin - 1].bits <= 64 && (end - begin) < 16) {
      bits += fields[begin - 1].bits;
      --begin;
    }
    FieldPack fp{};
    fp.n = end - begin;
    int shift = 0;
    for (int f = end - 1; f >= begin; --f) {
      fp.f[f - begin] = fields[f];
      fp.f[f - begin].shift = shift;
      shift += fields[f].bits;
    }
    uint32_t* np;
    TQ_TRY(sort_group(fp, bits, perm, n, w, s, &np));
    perm = np;
    end = begin;
  }
  if (!perm) {  // no keys: identity
    FieldPack fp{};
    TQ_TRY(launch("k_pack_keys", k_pack_keys<uint32_t>, stream_grid(n), kBlock, 0, s, fp, nullptr, n, (uint32_t*)w.k0, w.v0));
    perm = w.v0;
  }
  *perm_out = perm;
  return TQ_OK;
}

}  // namespace tq

using namespace tq;

extern "C" size_t tq_sort_ws_bytes(int64_t n) { return sort_ws_bytes(n); }

extern "C" int tq_read_words(int32_t n, const void* const* dev, const size_t* bytes,
                             int64_t* host_out, void* stream) {
  if (n < 0 || (n > 0 && (!dev || !bytes || !host_out)))
    return set_error(TQ_EINVAL, "read words: bad argument");
  return read_back_n(n, dev, bytes, host_out, (hipStream_t)stream);
}

// Reads the host collects later.  Per thread and device, a ring of TQ_WORD_SLOTS pinned,
// coherent host slots of 64 words: a post gathers the words into a slot with one
// k_collect_words launch and records the slot's event after it on the stream; a wait blocks on
// that event and copies the slot out.  The host queues further work on the stream in between.
constexpr int kWordSlots = 16;
constexpr int kRingDevices = 16;
struct WordRing {
  int64_t* host = nullptr;
  uint64_t* dev = nullptr;
  hipEvent_t ev[kWordSlots] = {};
  size_t bytes[kWordSlots] = {};
  bool pending[kWordSlots] = {};
};
static thread_local WordRing rings[kRingDevices];

static int word_ring(WordRing** out) {
  int d = 0;
  TQ_TRY_HIP(hipGetDevice(&d));
  if (d < 0 || d >= kRingDevices) return set_error(TQ_EUNSUPPORTED, "device %d has no word ring", d);
  WordRing& r = rings[d];
  if (!r.host) {
    int64_t* h = nullptr;
    TQ_TRY_HIP(hipHostMalloc((void**)&h, kWordSlots * 64 * sizeof(int64_t), hipHostMallocCoherent));
    uint64_t* dp = nullptr;
    if (hipHostGetDevicePointer((void**)&dp, h, 0) != hipSuccess || !dp) {
      (void)hipHostFree(h);
      return set_error(TQ_EHIP, "word ring: no device view of the host slots");
    }
    for (int i = 0; i < kWordSlots; ++i) {
      if (!r.ev[i]) TQ_TRY_HIP(hipEventCreateWithFlags(&r.ev[i], hipEventDisableTiming));
    }
    r.dev = dp;
    r.host = h;
  }
  *out = &r;
  return TQ_OK;
}

extern "C" int32_t tq_word_slots(void) { return kWordSlots; }

extern "C" int tq_words_post(int32_t n, const void* const* dev, const size_t* bytes, int32_t slot,
                             void* stream) {
  if (n < 1 || n > kMaxWordRuns || !dev || !bytes)
    return set_error(TQ_EINVAL, "words post: %d runs (1..%d)", n, kMaxWordRuns);
  if (slot < 0 || slot >= kWordSlots) return set_error(TQ_EINVAL, "words post: slot %d", slot);
  WordRing* r = nullptr;
  TQ_TRY(word_ring(&r));
  if (r->pending[slot]) return set_error(TQ_EINVAL, "words post: slot %d is still pending", slot);
  WordRuns W{};
  W.n = n;
  int32_t o = 0;
  for (int i = 0; i < n; ++i) {
    if (bytes[i] % 8 != 0 || ((uintptr_t)dev[i] & 7) != 0)
      return set_error(TQ_EINVAL, "words post: run %d is not whole aligned words", i);
    W.src[i] = (const uint64_t*)dev[i];
    W.nw[i] = (int32_t)(bytes[i] / 8);
    W.off[i] = o;
    o += W.nw[i];
    if (o > 64) return set_error(TQ_EINVAL, "words post: more than 64 words");
  }
  hipStream_t s = (hipStream_t)stream;
  TQ_TRY(launch("k_collect_words", k_collect_words, 1, 64, 0, s, W, r->dev + (size_t)slot * 64));
  TQ_TRY_HIP(hipEventRecord(r->ev[slot], s));
  r->bytes[slot] = (size_t)o * 8;
  r->pending[slot] = true;
  return TQ_OK;
}

extern "C" int tq_words_wait(int32_t slot, int64_t* host_out) {
  if (slot < 0 || slot >= kWordSlots || !host_out)
    return set_error(TQ_EINVAL, "words wait: slot %d", slot);
  WordRing* r = nullptr;
  TQ_TRY(word_ring(&r));
  if (!r->pending[slot]) return set_error(TQ_EINVAL, "words wait: slot %d was not posted", slot);
  r->pending[slot] = false;
  TQ_TRY_HIP(hipEventSynchronize(r->ev[slot]));
  memcpy(host_out, r->host + (size_t)slot * 64, r->bytes[slot]);
  return TQ_OK;
}

// Top-k for a limit of at most 256 over keys that pack into one 64-bit sort key, with no host
// synchronisation: rounds of per-workgroup threshold selection for a limit of at most 64, else
// packing and rounds of per-tile radix selection.
static int topk_perm(const tq_sortkey_t* keys, int32_t nkeys, int64_t n, int64_t limit,
                     int64_t* perm_out, void* ws, size_t ws_bytes, hipStream_t s, bool* done,
                     const int64_t* n_dev = nullptr, const ColPack* gather = nullptr) {
  *done = false;
  ColPack gp{};
  if (gather) gp = *gather;
  FieldPack fp{};
  int bits = 0;
  for (int k = 0; k < nkeys; ++k) {
    if (keys[k].width > 1) return TQ_OK;
    bits += bits_of(keys[k].dtype);
  }
  if (bits > 64 || nkeys > 16 || nkeys < 1) return TQ_OK;
  if (n >= (int64_t)0xFFFFFFFF) return TQ_OK;  // rows must stay below the sentinel row
  if (ws_bytes < sort_ws_bytes(n)) return set_error(TQ_ECAPACITY, "sort workspace too small");
  int shift = 0;
  for (int k = nkeys - 1; k >= 0; --k) {
    fp.f[k] = Field{keys[k].ptr, keys[k].dtype, 1, 0, keys[k].descending ? 1 : 0,
                    bits_of(keys[k].dtype), shift};
    shift += bits_of(keys[k].dtype);
  }
  fp.n = nkeys;
  SortWs w = carve_sort_ws(ws, n);
  const int k = (int)limit;
  const int64_t out = limit < n ? limit : n;
  if (k >= 1 && k <= kSmallK) {
    // rounds of per-workgroup selection straight from the key columns; each keeps k per 4096
    const uint64_t* ki = nullptr;
    const uint32_t* vi = nullptr;
    uint64_t* ko = w.k0;
    uint32_t* vo = w.v0;
    int64_t m = n;
    while (true) {
      const int64_t blocks = ceil_div(m, kListTile);
      const bool last = blocks == 1;
      TQ_TRY(launch("k_topk_block", k_topk_block, (unsigned)blocks, kBlock, 0, s, fp, ki, vi, m,
                    ki ? (const int64_t*)nullptr : n_dev, k, ko, vo,
                    last ? perm_out : (int64_t*)nullptr, out, gp));
      if (last) break;
      m = blocks * k;
      ki = ko;
      vi = vo;
      ko = ki == w.k0 ? w.k1 : w.k0;
      vo = vi == w.v0 ? w.v1 : w.v0;
    }
    *done = true;
    return TQ_OK;
  }
  // a device row count is read, and columns gathered, by the small-limit rounds only
  if (n_dev || gp.n) return TQ_OK;
  TQ_TRY(launch("k_pack_keys", k_pack_keys<uint64_t>, stream_grid(n), kBlock, 0, s, fp, nullptr, n, w.k0, w.v0));
  uint64_t *ki = w.k0, *ko = w.k1;
  uint32_t *vi = nullptr, *vo = w.v1;
  int64_t m = n;
  while (true) {
    const int64_t tiles = ceil_div(m, kTopTile);
    const bool last = tiles == 1;
    TQ_TRY(launch("k_topk_select", k_topk_select, (unsigned)tiles, kBlock, 0, s,
                  (const uint64_t*)ki, (const uint32_t*)vi, m, k, ko, vo,
                  last ? perm_out : (int64_t*)nullptr, out));
    if (last) break;
    m = tiles * k;
    ki = ko;  // this round's output is the next round's input
    vi = vo;
    ko = ki == w.k0 ? w.k1 : w.k0;
    vo = vi == w.v0 ? w.v1 : w.v0;
  }
  *done = true;
  return TQ_OK;
}

extern "C" int tq_sort_perm_deferred(const tq_sortkey_t* keys, int32_t nkeys, int64_t n,
                                     const int64_t* n_dev, int64_t limit, int64_t* perm_out,
                                     void* ws, size_t ws_bytes, int32_t* applied, void* stream) {
  if (!applied || !n_dev) return set_error(TQ_EINVAL, "sort (deferred): null argument");
  *applied = 0;
  if (n <= 0 || limit < 1 || limit > kSmallK) return TQ_OK;
  if (nkeys < 1 || !keys) return set_error(TQ_EINVAL, "sort: bad key list");
  bool done = false;
  TQ_TRY(topk_perm(keys, nkeys, n, limit, perm_out, ws, ws_bytes, (hipStream_t)stream, &done,
                   n_dev));
  *applied = done ? 1 : 0;
  return TQ_OK;
}

extern "C" int tq_topk_gather(const tq_sortkey_t* keys, int32_t nkeys, int64_t n,
                              const int64_t* n_dev, int64_t limit, int32_t ncols,
                              const void* const* src, const int32_t* row_bytes,
                              void* const* dst, int64_t* perm_out, void* ws, size_t ws_bytes,
                              int32_t* applied, void* stream) {
  if (!applied || !perm_out) return set_error(TQ_EINVAL, "top-k gather: null argument");
  *applied = 0;
  if (n <= 0 || limit < 1 || limit > kSmallK) return TQ_OK;
  if (nkeys < 1 || !keys) return set_error(TQ_EINVAL, "top-k gather: bad key list");
  if (ncols < 0 || ncols > TQ_MAX_COLS || (ncols > 0 && (!src || !row_bytes || !dst)))
    return set_error(TQ_EINVAL, "top-k gather: %d columns (0..%d)", ncols, TQ_MAX_COLS);
  ColPack gp{};
  gp.n = ncols;
  for (int c = 0; c < ncols; ++c) {
    if (!src[c] || !dst[c] || row_bytes[c] < 1)
      return set_error(TQ_EINVAL, "top-k gather: column %d", c);
    gp.src[c] = src[c];
    gp.dst[c] = dst[c];
    gp.bytes[c] = row_bytes[c];
  }
  bool done = false;
  TQ_TRY(topk_perm(keys, nkeys, n, limit, perm_out, ws, ws_bytes, (hipStream_t)stream, &done,
                   n_dev, &gp));
  *applied = done ? 1 : 0;
  return TQ_OK;
}

namespace tq {

// ------------------------------------------------------------------ small sorts
// Up to kSmallSort rows whose key fields pack into one 64-bit key sort in one workgroup by
// rank: the packed keys go to LDS, and row i's position is the number of rows whose key is
// smaller, plus the earlier rows with an equal key (so equal keys keep their input order: the
// stable sort the radix passes give).  Every lane reads the same key at each step (an LDS
// broadcast) and the count needs no barrier.  One launch, where the radix path takes its
// passes, scans and a read-back of the key bit range.
constexpr int kSmallSort = 2048;
constexpr int kSmallBlock = 1024;

__global__ __launch_bounds__(kSmallBlock) void k_sort_small(FieldPack fp, int64_t n, int64_t m,
                                                            int64_t* __restrict__ perm_out) {
  __shared__ uint64_t key[kSmallSort];
  const int t = threadIdx.x;
  const int nn = (int)n;
  for (int i = t; i < nn; i += kSmallBlock) {
    uint64_t k = 0;
    for (int f = 0; f < fp.n; ++f) k |= field_bits(fp.f[f], i);
    key[i] = k;
  }
  __syncthreads();
  for (int i = t; i < nn; i += kSmallBlock) {
    const uint64_t ki = key[i];
    uint32_t r = 0;
    int j = 0;
    for (; j + 4 <= nn; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint64_t kj = key[j + u];
        r += (kj < ki || (kj == ki && j + u < i)) ? 1u : 0u;
      }
    }
    for (; j < nn; ++j) {
      const uint64_t kj = key[j];
      r += (kj < ki || (kj == ki && j < i)) ? 1u : 0u;
    }
    if (r < m) perm_out[r] = i;
  }
}

// The small path, when it applies (one packed key of at most 64 bits, n <= kSmallSort): true
// in *done after launching it.
static int sort_small(const tq_sortkey_t* keys, int32_t nkeys, int64_t n, int64_t m,
                      int64_t* perm_out, hipStream_t s, bool* done) {
  *done = false;
  if (n > kSmallSort || nkeys < 1) return TQ_OK;
  FieldPack fp{};
  int bits = 0;
  for (int k = 0; k < nkeys; ++k) {
    const int w = keys[k].width > 0 ? keys[k].width : 1;
    if (w > 1 && keys[k].dtype != TQ_I8 && keys[k].dtype != TQ_BOOL) return TQ_OK;
    bits += w * bits_of(keys[k].dtype);
    if (bits > 64 || fp.n + w > 16) return TQ_OK;
    for (int c = 0; c < w; ++c)
      fp.f[fp.n++] = Field{keys[k].ptr, keys[k].dtype, w, c, keys[k].descending ? 1 : 0,
                           bits_of(keys[k].dtype), 0};
  }
  // most significant field first: shifts from the least significant end
  int shift = 0;
  for (int f = fp.n - 1; f >= 0; --f) {
    fp.f[f].shift = shift;
    shift += fp.f[f].bits;
  }
  TQ_TRY(launch("k_sort_small", k_sort_small, 1, kSmallBlock, 0, s, fp, n, m, perm_out));
  *done = true;
  return TQ_OK;
}

}  // namespace tq

extern "C" int tq_sort_perm(const tq_sortkey_t* keys, int32_t nkeys, int64_t n, int64_t limit,
                            int64_t* perm_out, void* ws, size_t ws_bytes, void* stream) {
  if (n <= 0) return TQ_OK;
  if (nkeys < 0 || (nkeys > 0 && !keys)) return set_error(TQ_EINVAL, "sort: bad key list");
  hipStream_t s = (hipStream_t)stream;
  if (limit >= 0 && limit <= 256 && limit < n && n > kTopTile) {
    bool done;
    TQ_TRY(topk_perm(keys, nkeys, n, limit, perm_out, ws, ws_bytes, s, &done));
    if (done) return TQ_OK;
  }
  const int64_t m = (limit >= 0 && limit < n) ? limit : n;
  {
    bool done;
    TQ_TRY(sort_small(keys, nkeys, n, m, perm_out, s, &done));
    if (done) return TQ_OK;
  }
  uint32_t* perm;
  TQ_TRY(sort_perm_u32(keys, nkeys, n, ws, ws_bytes, s, &perm));
  return launch("k_perm_out", k_perm_out, stream_grid(m), kBlock, 0, s, perm, m, perm_out);
}

// ------------------------------------------------------------------ dense ranks of sorted rows
// Rows in a stable sort order (perm) get the rank of their key among the distinct keys: the
// number of key changes up to their position, so equal keys share a rank and ranks ascend
// with the order.  Floats compare as the sort orders them (-0.0 equals +0.0, every NaN equals
// every other NaN); a 2-D key compares its row bytes.  A group-by over keys that do not pack
// into integer ranges (strings, floats) groups by these ranks instead.
namespace tq {

struct RankKeys {
  const void* ptr[TQ_GB_MAX_KEYS];
  int32_t dtype[TQ_GB_MAX_KEYS];
  int32_t width[TQ_GB_MAX_KEYS];
  int32_t n;
};

__device__ __forceinline__ bool rank_rows_equal(const RankKeys& K, int64_t a, int64_t b) {
  for (int k = 0; k < K.n; ++k) {
    const int dt = K.dtype[k];
    const int w = K.width[k];
    if (w > 1) {
      const int8_t* p = (const int8_t*)K.ptr[k];
      for (int c = 0; c < w; ++c)
        if (p[a * w + c] != p[b * w + c]) return false;
    } else if (dt == TQ_F64) {
      const double x = ((const double*)K.ptr[k])[a], y = ((const double*)K.ptr[k])[b];
      if (!(x == y || (x != x && y != y))) return false;
    } else if (dt == TQ_F32) {
      const float x = ((const float*)K.ptr[k])[a], y = ((const float*)K.ptr[k])[b];
      if (!(x == y || (x != x && y != y))) return false;
    } else if (load_int(K.ptr[k], dt, a) != load_int(K.ptr[k], dt, b)) {
      return false;
    }
  }
  return true;
}

// Sorted positions tile0 + 16 * t + j (j < 16) of a 4096-position tile: bit j set when the
// key differs from the previous position's (position 0 always starts a rank).
__device__ __forceinline__ uint32_t rank_starts(const RankKeys& K, const int64_t* perm, int64_t n,
                                                int64_t tile0) {
  const int64_t first = tile0 + 16 * (int64_t)threadIdx.x;
  uint32_t bits = 0;
  if (first >= n) return 0;
  int64_t prev = first > 0 ? perm[first - 1] : -1;
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = first + j;
    if (i >= n) break;
    const int64_t row = perm[i];
    if (prev < 0 || !rank_rows_equal(K, prev, row)) bits |= 1u << j;
    prev = row;
  }
  return bits;
}

__global__ __launch_bounds__(kBlock) void k_rank_count(RankKeys K, const int64_t* __restrict__ perm,
                                                       int64_t n, uint32_t* __restrict__ counts) {
  __shared__ uint32_t wc[kWaves];
  const uint32_t c = wave_sum((uint32_t)__popc(rank_starts(K, perm, n, (int64_t)blockIdx.x * kTileRows)));
  if (lane_id() == 0) wc[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b = 0;
#pragma unroll
    for (int i = 0; i < kWaves; ++i) b += wc[i];
    counts[blockIdx.x] = b;
  }
}

// rank_out[perm[i]] = rank of position i; rep_out[r] = the first row (in sort order) of rank r.
__global__ __launch_bounds__(kBlock) void k_rank_write(RankKeys K, const int64_t* __restrict__ perm,
                                                       int64_t n, const int64_t* __restrict__ tile_off,
                                                       int32_t* __restrict__ rank_out,
                                                       int64_t* __restrict__ rep_out) {
  __shared__ uint32_t sh[kWaves];
  const int64_t tile0 = (int64_t)blockIdx.x * kTileRows;
  const uint32_t bits = rank_starts(K, perm, n, tile0);
  uint32_t tot;
  const uint32_t before = block_exclusive_sum((uint32_t)__popc(bits), sh, &tot);
  int64_t r = tile_off[blockIdx.x] + before - 1;  // rank of the position before this thread's
  const int64_t first = tile0 + 16 * (int64_t)threadIdx.x;
  for (int j = 0; j < kItems; ++j) {
    const int64_t i = first + j;
    if (i >= n) break;
    const int64_t row = perm[i];
    if ((bits >> j) & 1u) {
      ++r;
      rep_out[r] = row;
    }
    rank_out[row] = (int32_t)r;
  }
}

}  // namespace tq

extern "C" size_t tq_rank_ws_bytes(int64_t n) {
  const int64_t nt = ceil_div(n > 0 ? n : 1, kTileRows);
  return align_up((size_t)nt * 4, 256) + align_up((size_t)nt * 8, 256) + tq_scan_ws_bytes(nt);
}

extern "C" int tq_rank_sorted(const tq_sortkey_t* keys, int32_t nkeys, int64_t n,
                              const int64_t* perm, int32_t* rank_out, int64_t* rep_out,
                              int64_t* nranks, void* ws, size_t ws_bytes, void* stream) {
  if (!keys || nkeys < 1 || nkeys > TQ_GB_MAX_KEYS || !nranks)
    return set_error(TQ_EINVAL, "rank: bad keys");
  if (n < 0 || n >= ((int64_t)1 << 31)) return set_error(TQ_EUNSUPPORTED, "rank: %lld rows", (long long)n);
  if (n > 0 && (!perm || !rank_out || !rep_out || !ws))
    return set_error(TQ_EINVAL, "rank: null argument");
  if (ws_bytes < tq_rank_ws_bytes(n)) return set_error(TQ_ECAPACITY, "rank: workspace too small");
  RankKeys K{};
  K.n = nkeys;
  for (int k = 0; k < nkeys; ++k) {
    if (!keys[k].ptr && n > 0) return set_error(TQ_EINVAL, "rank: null key column");
    K.ptr[k] = keys[k].ptr;
    K.dtype[k] = keys[k].dtype;
    K.width[k] = keys[k].width > 0 ? keys[k].width : 1;
    if (K.width[k] > 1 && K.dtype[k] != TQ_I8)
      return set_error(TQ_EUNSUPPORTED, "rank: 2-D keys must be int8 strings");
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t nt = ceil_div(n > 0 ? n : 1, kTileRows);
  char* w = (char*)ws;
  uint32_t* counts = (uint32_t*)w;
  int64_t* off = (int64_t*)(w + align_up((size_t)nt * 4, 256));
  void* sws = w + align_up((size_t)nt * 4, 256) + align_up((size_t)nt * 8, 256);
  if (n == 0) return fill_u64(nranks, 0, 1, s);
  TQ_TRY(launch("k_rank_count", k_rank_count, (unsigned)nt, kBlock, 0, s, K, perm, n, counts));
  TQ_TRY(tq_scan_tiles(counts, nt, off, nranks, sws, s));
  return launch("k_rank_write", k_rank_write, (unsigned)nt, kBlock, 0, s, K, perm, n,
                (const int64_t*)off, rank_out, rep_out);
}
