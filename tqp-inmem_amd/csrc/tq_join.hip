// Sort-based equi-join: build = compact + radix-sort the selected build keys (row ids ride
// along, so each equal-key bucket lists build rows ascending); probe = two passes over probe
// tiles (count, then write at exclusive-scan offsets) that emit pairs in probe-row order.
//
// Probe lookups: when the selected build keys are dense (span <= 4*nsel + 65536 values) a
// directory indexed by key - min key gives the first sorted position in one load.  Otherwise
// the sorted keys are searched in two levels: every workgroup first copies an evenly spaced
// sample of the sorted keys into LDS (the last key of each block of `step` keys), finds the
// block by binary search in LDS, and only then searches that block in global memory, so a
// lookup touches one or two cache lines of the key array instead of log2(nsel).
#include "tq_common.h"
#include "tq_sort.h"

namespace tq {

constexpr int kSampleBytes = 32768;

struct JoinView {
  const void* keys;
  const uint32_t* rows;
  const void* sample;
  const uint32_t* dir;
  uint64_t dir_base;
  int64_t dir_range;
  int64_t nsel;
  int64_t step;
  int32_t nsample;
};

constexpr uint32_t kNoPos = 0xFFFFFFFFu;

template <typename K>
__device__ __forceinline__ bool probe_key(const void* p, int dt, int64_t r, K* out) {
  const int64_t v = load_int(p, dt, r);
  if (sizeof(K) == 4) {
    if (v < INT32_MIN || v > INT32_MAX) return false;
    *out = (K)((uint32_t)(int32_t)v ^ 0x80000000u);
  } else {
    *out = (K)((uint64_t)v ^ 0x8000000000000000ull);
  }
  return true;
}

// First index in [0, nsel) whose key is >= t (strict=false) or > t (strict=true), nsel if none.
template <typename K>
__device__ __forceinline__ int64_t search(const JoinView& J, const K* __restrict__ ssample,
                                          const K* __restrict__ keys, K t, bool strict) {
  int lo = 0, hi = J.nsample;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const K s = ssample[mid];
    if (strict ? (s <= t) : (s < t)) lo = mid + 1; else hi = mid;
  }
  if (lo == J.nsample) return J.nsel;
  int64_t b = (int64_t)lo * J.step;
  int64_t e = b + J.step;
  if (e > J.nsel) e = J.nsel;
  while (b < e) {
    const int64_t mid = (b + e) >> 1;
    const K k = keys[mid];
    if (strict ? (k <= t) : (k < t)) b = mid + 1; else e = mid;
  }
  return b;
}

// Bucket of key t: first position and length (0 when absent).
template <typename K>
__device__ __forceinline__ uint32_t bucket(const JoinView& J, const K* __restrict__ ssample, K t,
                                           int64_t* first) {
  const K* keys = (const K*)J.keys;
  if (J.dir) {
    const uint64_t off = (uint64_t)t - J.dir_base;  // wraps to a huge value below the base
    if (off >= (uint64_t)J.dir_range) return 0;
    const uint32_t p = J.dir[off];
    *first = p;
    if (p == kNoPos) return 0;
    if ((int64_t)p + 1 >= J.nsel || keys[p + 1] != t) return 1;
    return (uint32_t)(search<K>(J, ssample, keys, t, true) - p);
  }
  const int64_t p = search<K>(J, ssample, keys, t, false);
  *first = p;
  if (p >= J.nsel || keys[p] != t) return 0;
  if (p + 1 >= J.nsel || keys[p + 1] != t) return 1;
  return (uint32_t)(search<K>(J, ssample, keys, t, true) - p);
}

template <typename K>
__device__ __forceinline__ void load_sample(const JoinView& J, K* ssample) {
  const K* src = (const K*)J.sample;
  for (int i = threadIdx.x; i < J.nsample; i += kBlock) ssample[i] = src[i];
  __syncthreads();
}

// ------------------------------------------------------------------ build
// Also reduces OR, AND, min and max of the packed keys into stats[0..3], so the host learns the
// sort's bit range and the directory's span together with the selected-row count.
template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_pack(const void* __restrict__ bkeys, int dt,
                                                      const uint8_t* __restrict__ mask, int64_t n,
                                                      const int64_t* __restrict__ tile_off,
                                                      K* __restrict__ keys,
                                                      uint32_t* __restrict__ rows,
                                                      unsigned long long* __restrict__ stats) {
  uint64_t so = 0, sa = ~0ull, slo = ~0ull, shi = 0;
  __shared__ uint32_t wcnt[kItems][kWaves];
  const int t = threadIdx.x, w = t >> 6, l = lane_id();
  const int64_t tile0 = (int64_t)blockIdx.x * kTileRows;
  bool f[kItems];
  uint64_t bal[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const int64_t r = tile0 + (int64_t)j * kBlock + t;
    f[j] = (r < n) && (!mask || mask[r] != 0);
    bal[j] = ballot(f[j]);
    if (l == 0) wcnt[j][w] = (uint32_t)__popcll(bal[j]);
  }
  __syncthreads();
  int64_t run = mask ? tile_off[blockIdx.x] : tile0;
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
      const int64_t r = tile0 + (int64_t)j * kBlock + t;
      const int64_t pos = run + before + rank_below(bal[j]);
      const int64_t v = load_int(bkeys, dt, r);
      const K k = sizeof(K) == 4 ? (K)((uint32_t)(int32_t)v ^ 0x80000000u)
                                 : (K)((uint64_t)v ^ 0x8000000000000000ull);
      keys[pos] = k;
      rows[pos] = (uint32_t)r;
      so |= (uint64_t)k;
      sa &= (uint64_t)k;
      slo = (uint64_t)k < slo ? (uint64_t)k : slo;
      shi = (uint64_t)k > shi ? (uint64_t)k : shi;
    }
    run += tot;
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    so |= __shfl_xor(so, d, 64);
    sa &= __shfl_xor(sa, d, 64);
    const uint64_t a2 = __shfl_xor(slo, d, 64), b2 = __shfl_xor(shi, d, 64);
    slo = a2 < slo ? a2 : slo;
    shi = b2 > shi ? b2 : shi;
  }
  if (l == 0 && shi >= slo) {
    atomicOr(&stats[0], (unsigned long long)so);
    atomicAnd(&stats[1], (unsigned long long)sa);
    atomicMin(&stats[2], (unsigned long long)slo);
    atomicMax(&stats[3], (unsigned long long)shi);
  }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_sample(const K* __restrict__ keys, int64_t nsel,
                                                        int64_t step, int32_t nsample,
                                                        K* __restrict__ sample) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nsample) {
    int64_t e = (int64_t)(i + 1) * step;
    if (e > nsel) e = nsel;
    sample[i] = keys[e - 1];
  }
}

// Dense directory: dir[key - base] = first sorted position of key (dir pre-filled with kNoPos).
template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_dir(const K* __restrict__ keys, int64_t nsel,
                                                     uint64_t base, uint32_t* __restrict__ dir) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nsel; i += stride) {
    const K k = keys[i];
    if (i == 0 || keys[i - 1] != k) dir[(uint64_t)k - base] = (uint32_t)i;
  }
}

// ------------------------------------------------------------------ probe
template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_count(JoinView J, const void* __restrict__ pkeys,
                                                       int pdt, const uint8_t* __restrict__ pmask,
                                                       int64_t np, int64_t ntiles,
                                                       uint32_t* __restrict__ tile_counts) {
  extern __shared__ unsigned char smem[];
  K* ssample = (K*)smem;
  __shared__ uint32_t sh[kWaves];
  load_sample<K>(J, ssample);
  const int t = threadIdx.x;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t cnt = 0;
    for (int j = 0; j < kItems; ++j) {
      const int64_t r = tile * kTileRows + (int64_t)j * kBlock + t;
      K key;
      if (r < np && (!pmask || pmask[r] != 0) && probe_key<K>(pkeys, pdt, r, &key)) {
        int64_t first;
        cnt += bucket<K>(J, ssample, key, &first);
      }
    }
    cnt = wave_sum(cnt);
    if (lane_id() == 0) sh[t >> 6] = cnt;
    __syncthreads();
    if (t == 0) {
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < kWaves; ++w) s += sh[w];
      tile_counts[tile] = s;
    }
    __syncthreads();
  }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_join_write(
    JoinView J, const void* __restrict__ pkeys, int pdt, const uint8_t* __restrict__ pmask,
    int64_t np, int64_t ntiles, const int64_t* __restrict__ tile_off, ColPack L, ColPack Rc,
    int64_t* __restrict__ build_idx, int64_t* __restrict__ probe_idx) {
  extern __shared__ unsigned char smem[];
  K* ssample = (K*)smem;
  __shared__ uint32_t wtot[kItems][kWaves];
  load_sample<K>(J, ssample);
  const int t = threadIdx.x, w = t >> 6, l = lane_id();
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint32_t cnt[kItems], incl[kItems];
    int64_t first[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int64_t r = tile * kTileRows + (int64_t)j * kBlock + t;
      K key;
      cnt[j] = 0;
      first[j] = 0;
      if (r < np && (!pmask || pmask[r] != 0) && probe_key<K>(pkeys, pdt, r, &key))
        cnt[j] = bucket<K>(J, ssample, key, &first[j]);
      incl[j] = wave_inclusive_sum(cnt[j]);
      if (l == 63) wtot[j][w] = incl[j];
    }
    __syncthreads();
    int64_t run = tile_off[tile];
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      uint32_t before = 0, tot = 0;
#pragma unroll
      for (int v = 0; v < kWaves; ++v) {
        const uint32_t c = wtot[j][v];
        tot += c;
        before += (v < w) ? c : 0;
      }
      if (cnt[j]) {
        const int64_t r = tile * kTileRows + (int64_t)j * kBlock + t;
        const int64_t o0 = run + before + incl[j] - cnt[j];
        for (uint32_t k = 0; k < cnt[j]; ++k) {
          const int64_t o = o0 + k;
          const uint32_t b = J.rows[first[j] + k];
          for (int c = 0; c < L.n; ++c) copy_row(L.dst[c], o, L.src[c], b, L.bytes[c]);
          for (int c = 0; c < Rc.n; ++c) copy_row(Rc.dst[c], o, Rc.src[c], r, Rc.bytes[c]);
          if (build_idx) build_idx[o] = b;
          if (probe_idx) probe_idx[o] = r;
        }
      }
      run += tot;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ workspace layouts
struct BuildWs {
  uint32_t* tile_counts;
  int64_t* tile_off;
  int64_t* total;
  int64_t* scan_ws;
  void* k0;
  void* k1;
  uint32_t* v0;
  uint32_t* v1;
  void* stats;
  void* radix;
  void* sample;
  uint32_t* dir;
  size_t bytes;
};

static int64_t dir_cap(int64_t nsel) { return 4 * nsel + 65536; }

static BuildWs carve_build(void* base, int64_t nb) {
  if (nb < 1) nb = 1;
  const int64_t nt = ceil_div(nb, kTileRows);
  char* p = (char*)base;
  BuildWs w;
  w.tile_counts = (uint32_t*)p; p += align_up((size_t)nt * 4, 256);
  w.tile_off = (int64_t*)p; p += align_up((size_t)nt * 8, 256);
  w.total = (int64_t*)p;          // total and the 4 key statistics are read back together
  w.stats = p + 8; p += 256;
  w.scan_ws = (int64_t*)p; p += tq_scan_ws_bytes(nt);
  w.k0 = p; p += align_up((size_t)nb * 8, 256);
  w.k1 = p; p += align_up((size_t)nb * 8, 256);
  w.v0 = (uint32_t*)p; p += align_up((size_t)nb * 4, 256);
  w.v1 = (uint32_t*)p; p += align_up((size_t)nb * 4, 256);
  w.radix = p; p += radix_ws_bytes(nb);
  w.sample = p; p += kSampleBytes;
  w.dir = (uint32_t*)p; p += align_up((size_t)dir_cap(nb) * 4, 256);
  w.bytes = (size_t)(p - (char*)base);
  return w;
}

struct ProbeWs {
  uint32_t* tile_counts;
  int64_t* tile_off;
  int64_t* total;
  int64_t* scan_ws;
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
  w.bytes = (size_t)(p - (char*)base);
  return w;
}

static JoinView view_of(const tq_join_state_t* st) {
  JoinView J;
  J.keys = st->keys;
  J.rows = st->rows;
  J.sample = st->sample;
  J.dir = st->dir;
  J.dir_base = st->dir_base;
  J.dir_range = st->dir_range;
  J.nsel = st->nsel;
  J.step = st->step;
  J.nsample = st->nsample;
  return J;
}

static unsigned probe_grid(int64_t ntiles) {
  return (unsigned)(ntiles < 1024 ? (ntiles > 0 ? ntiles : 1) : 1024);
}

template <typename K>
static int build_impl(const void* bkeys, int32_t dt, const uint8_t* bmask, int64_t nb, BuildWs& w,
                      tq_join_state_t* st, hipStream_t s) {
  const int64_t nt = ceil_div(nb, kTileRows);
  int64_t nsel = nb;
  if (bmask) {
    TQ_TRY(tq_mask_tile_counts(bmask, nb, w.tile_counts, s));
    TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
  }
  unsigned long long* kst = (unsigned long long*)w.stats;
  TQ_TRY(fill_u64(kst, 0ull, 1, s));
  TQ_TRY(fill_u64(kst + 1, ~0ull, 2, s));
  TQ_TRY(fill_u64(kst + 3, 0ull, 1, s));
  k_join_pack<K><<<(unsigned)nt, kBlock, 0, s>>>(bkeys, dt, bmask, nb, w.tile_off, (K*)w.k0, w.v0,
                                                kst);
  TQ_TRY(check_launch("join pack"));
  int64_t hv[5];
  if (!bmask) TQ_TRY(fill_u64(w.total, (uint64_t)nb, 1, s));
  TQ_TRY(read_back(w.total, 40, hv, s));   // nsel, OR, AND, min, max in one synchronisation
  nsel = hv[0];
  const uint64_t kmin = (uint64_t)hv[3], kmax = (uint64_t)hv[4];
  int lo = 0, hi = 0;
  radix_bit_range((uint64_t)hv[1], (uint64_t)hv[2], &lo, &hi);
  bool alt;
  TQ_TRY(radix_sort_pairs<K>((K*)w.k0, w.v0, (K*)w.k1, w.v1, nsel, lo, hi, w.radix, s, &alt));
  const K* keys = (const K*)(alt ? w.k1 : w.k0);
  const uint32_t* rows = alt ? w.v1 : w.v0;
  const int64_t cap = kSampleBytes / (int64_t)sizeof(K);
  const int32_t nsample = (int32_t)(nsel < cap ? nsel : cap);
  const int64_t step = nsample > 0 ? ceil_div(nsel, nsample) : 1;
  const int32_t nsample_eff = nsample > 0 ? (int32_t)ceil_div(nsel, step) : 0;
  if (nsample_eff > 0) {
    k_join_sample<K><<<(unsigned)ceil_div(nsample_eff, kBlock), kBlock, 0, s>>>(
        keys, nsel, step, nsample_eff, (K*)w.sample);
    TQ_TRY(check_launch("join sample"));
  }
  st->dir = nullptr;
  st->dir_base = 0;
  st->dir_range = 0;
  if (nsel > 0 && kmax - kmin < (uint64_t)dir_cap(nsel)) {
    const int64_t range = (int64_t)(kmax - kmin) + 1;
    TQ_TRY(fill_u32(w.dir, kNoPos, range, s));
    k_join_dir<K><<<stream_grid(nsel), kBlock, 0, s>>>(keys, nsel, kmin, w.dir);
    TQ_TRY(check_launch("join directory"));
    st->dir = w.dir;
    st->dir_base = kmin;
    st->dir_range = range;
  }
  st->nsel = nsel;
  st->key_bits = (int32_t)(8 * sizeof(K));
  st->nsample = nsample_eff;
  st->step = step;
  st->keys = keys;
  st->rows = rows;
  st->sample = w.sample;
  return TQ_OK;
}

}  // namespace tq

using namespace tq;

extern "C" size_t tq_join_build_ws_bytes(int64_t nb) { return carve_build(nullptr, nb).bytes; }
extern "C" size_t tq_join_probe_ws_bytes(int64_t np) { return carve_probe(nullptr, np).bytes; }

extern "C" int tq_join_build(const void* bkeys, int32_t bdtype, const uint8_t* bmask, int64_t nb,
                             void* ws, size_t ws_bytes, tq_join_state_t* state, void* stream) {
  if (!state) return set_error(TQ_EINVAL, "join build: null state");
  if (bdtype < TQ_BOOL || bdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", bdtype);
  if (nb >= ((int64_t)1 << 32)) return set_error(TQ_EINVAL, "build side too large");
  BuildWs w = carve_build(ws, nb);
  if (ws_bytes < w.bytes) return set_error(TQ_ECAPACITY, "join build workspace too small");
  *state = tq_join_state_t{};
  state->key_bits = bdtype == TQ_I64 ? 64 : 32;
  if (nb <= 0) return TQ_OK;
  hipStream_t s = (hipStream_t)stream;
  return bdtype == TQ_I64 ? build_impl<uint64_t>(bkeys, bdtype, bmask, nb, w, state, s)
                          : build_impl<uint32_t>(bkeys, bdtype, bmask, nb, w, state, s);
}

extern "C" int tq_join_count(const tq_join_state_t* st, const void* pkeys, int32_t pdtype,
                             const uint8_t* pmask, int64_t np, void* pws, int64_t* total_host,
                             void* stream) {
  if (!st) return set_error(TQ_EINVAL, "join count: null state");
  if (pdtype < TQ_BOOL || pdtype > TQ_I64)
    return set_error(TQ_EUNSUPPORTED, "join keys must be integer columns (dtype %d)", pdtype);
  hipStream_t s = (hipStream_t)stream;
  ProbeWs w = carve_probe(pws, np);
  const int64_t nt = np > 0 ? ceil_div(np, kTileRows) : 0;
  if (np <= 0 || st->nsel <= 0) {
    TQ_TRY(fill_u32(w.tile_counts, 0, nt, s));
    TQ_TRY(fill_u64(w.tile_off, 0, nt, s));
    TQ_TRY(fill_u64(w.total, 0, 1, s));
    if (total_host) *total_host = 0;
    return TQ_OK;
  }
  JoinView J = view_of(st);
  const size_t lds = (size_t)st->nsample * (st->key_bits / 8);
  if (st->key_bits == 32)
    k_join_count<uint32_t><<<probe_grid(nt), kBlock, lds, s>>>(J, pkeys, pdtype, pmask, np, nt,
                                                                w.tile_counts);
  else
    k_join_count<uint64_t><<<probe_grid(nt), kBlock, lds, s>>>(J, pkeys, pdtype, pmask, np, nt,
                                                                w.tile_counts);
  TQ_TRY(check_launch("join count"));
  TQ_TRY(tq_scan_tiles(w.tile_counts, nt, w.tile_off, w.total, w.scan_ws, s));
  if (total_host) TQ_TRY(read_back(w.total, 8, total_host, s));
  return TQ_OK;
}

extern "C" int tq_join_write(const tq_join_state_t* st, const void* pkeys, int32_t pdtype,
                             const uint8_t* pmask, int64_t np, const void* pws, int32_t nleft,
                             const void* const* lsrc, const int32_t* lbytes, void* const* ldst,
                             int32_t nright, const void* const* rsrc, const int32_t* rbytes,
                             void* const* rdst, int64_t* build_idx, int64_t* probe_idx,
                             void* stream) {
  if (!st) return set_error(TQ_EINVAL, "join write: null state");
  ColPack L, R;
  TQ_TRY(make_colpack(&L, nleft, lsrc, lbytes, nullptr, ldst));
  TQ_TRY(make_colpack(&R, nright, rsrc, rbytes, nullptr, rdst));
  if (np <= 0 || st->nsel <= 0) return TQ_OK;
  hipStream_t s = (hipStream_t)stream;
  ProbeWs w = carve_probe((void*)pws, np);
  const int64_t nt = ceil_div(np, kTileRows);
  JoinView J = view_of(st);
  const size_t lds = (size_t)st->nsample * (st->key_bits / 8);
  if (st->key_bits == 32)
    k_join_write<uint32_t><<<probe_grid(nt), kBlock, lds, s>>>(J, pkeys, pdtype, pmask, np, nt,
                                                                w.tile_off, L, R, build_idx,
                                                                probe_idx);
  else
    k_join_write<uint64_t><<<probe_grid(nt), kBlock, lds, s>>>(J, pkeys, pdtype, pmask, np, nt,
                                                                w.tile_off, L, R, build_idx,
                                                                probe_idx);
  return check_launch("join write");
}
