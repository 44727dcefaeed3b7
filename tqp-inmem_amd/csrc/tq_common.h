// Shared device/host helpers for the gfx950 kernels: error reporting, launch geometry and
// wave64 primitives.  Every kernel in this library is memory-bound; the helpers keep loads
// coalesced (64 consecutive rows per wave instruction) and reductions inside LDS.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <tuple>
#include <utility>

#include "../../include/tqhip.h"

namespace tq {

// ------------------------------------------------------------------ geometry
constexpr int kBlock = 256;                // threads per workgroup (4 wave64s)
constexpr int kWaves = kBlock / 64;
constexpr int kItems = 16;                 // rows per thread in a filter/compaction tile
constexpr int kTileRows = kBlock * kItems; // 4096 rows per tile

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// ------------------------------------------------------------------ errors
int set_error(int code, const char* fmt, ...);
int check_launch(const char* what);

#define TQ_TRY_HIP(expr)                                                            \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return ::tq::set_error(TQ_EHIP, "%s: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

#define TQ_TRY(expr)          \
  do {                        \
    int r_ = (expr);          \
    if (r_ != TQ_OK) return r_; \
  } while (0)

// ------------------------------------------------------------------ launches
// Every kernel is launched through `launch`.  While kernel timing is on (tq_timing), the
// dispatch carries a start and a stop event that are stamped when the kernel itself starts and
// ends, so a recorded duration is the kernel's alone, excluding host gaps, as a kernel trace
// measures it.  Timing is meant for eager launches, not for graph capture.
bool timing_on();
void timing_slot(const char* name, hipEvent_t* e0, hipEvent_t* e1);

template <typename Tup, size_t... I>
inline void arg_ptrs(Tup& t, void** out, std::index_sequence<I...>) {
  ((out[I] = (void*)&std::get<I>(t)), ...);
}

template <typename... P, typename... A>
int launch(const char* name, void (*kernel)(P...), dim3 grid, dim3 block, size_t lds,
           hipStream_t s, A&&... a) {
  static_assert(sizeof...(P) == sizeof...(A), "kernel argument count");
  std::tuple<P...> t(std::forward<A>(a)...);
  void* argv[sizeof...(P) > 0 ? sizeof...(P) : 1];
  arg_ptrs(t, argv, std::index_sequence_for<P...>{});
  hipError_t e;
  if (timing_on()) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    timing_slot(name, &e0, &e1);
    e = hipExtLaunchKernel((const void*)kernel, grid, block, argv, lds, s, e0, e1, 0);
  } else {
    e = hipLaunchKernel((const void*)kernel, grid, block, argv, lds, s);
  }
  if (e != hipSuccess) return set_error(TQ_EHIP, "%s launch: %s", name, hipGetErrorString(e));
  return TQ_OK;
}

// ------------------------------------------------------------------ wave64 primitives
__device__ __forceinline__ int lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
// number of set bits of m at lanes strictly below this lane
__device__ __forceinline__ int rank_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    T o = __shfl_up(v, d, 64);
    if (l >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

// Exclusive prefix sum across the 256 threads of a block; returns the block total in *total.
// `sh` must hold kWaves elements of T.  Contains __syncthreads.
template <typename T>
__device__ __forceinline__ T block_exclusive_sum(T v, T* sh, T* total) {
  const int l = lane_id();
  const int w = threadIdx.x >> 6;
  T inc = wave_inclusive_sum(v);
  if (l == 63) sh[w] = inc;
  __syncthreads();
  T base = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kWaves; ++i) {
    T s = sh[i];
    if (i < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + inc - v;
}

// Order-preserving maps of typed values to unsigned integers (used by sort and group-by).
__device__ __forceinline__ uint64_t f64_order_bits(double x) {
  // torch.sort compares values: -0.0 ties +0.0 and every NaN ranks above +inf.
  if (x == 0.0) x = 0.0;
  uint64_t b = __double_as_longlong(x);
  if (x != x) b = 0x7ff8000000000000ull;
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ uint32_t f32_order_bits(float x) {
  if (x == 0.0f) x = 0.0f;
  uint32_t b = __float_as_uint(x);
  if (x != x) b = 0x7fc00000u;
  return (b >> 31) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ int64_t load_int(const void* p, int dtype, int64_t i) {
  switch (dtype) {
    case TQ_BOOL: return ((const uint8_t*)p)[i] != 0;
    case TQ_I8: return ((const int8_t*)p)[i];
    case TQ_I16: return ((const int16_t*)p)[i];
    case TQ_I32: return ((const int32_t*)p)[i];
    default: return ((const int64_t*)p)[i];
  }
}

// Selection of rows base + j*kBlock + threadIdx.x (j < N): the row exists and, with a mask,
// its byte is set.  The mask bytes are loaded unconditionally (a row past n reads row 0, so
// n > 0 is required), so all N loads are in flight together rather than one per branch.
template <int N>
__device__ __forceinline__ void select_rows(const uint8_t* mask, int64_t base, int64_t n,
                                            bool (&live)[N]) {
  uint8_t m[N];
#pragma unroll
  for (int j = 0; j < N; ++j) m[j] = 1;
  if (mask) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
      const int64_t r = base + (int64_t)j * blockDim.x + threadIdx.x;
      m[j] = mask[r < n ? r : 0];
    }
  }
#pragma unroll
  for (int j = 0; j < N; ++j)
    live[j] = base + (int64_t)j * blockDim.x + threadIdx.x < n && m[j] != 0;
}

// ------------------------------------------------------------------ tile row loads
// A thread's 16 rows of a 4096-row tile: item j is row tile0 + (j/4)*1024 + 4*threadIdx.x +
// j%4 (four runs of four consecutive rows).  The loaders issue every load before using any:
// a full tile with aligned columns takes one vector load per run, any other tile clamped
// scalar loads (rows past n read row 0, so n > 0 is required).  Branches depend only on the
// tile, the pointer and the dtype, so they are uniform across the workgroup.
__device__ __forceinline__ int64_t run_row(int64_t tile0, int j) {
  return tile0 + (int64_t)(j >> 2) * (4 * kBlock) + 4 * threadIdx.x + (j & 3);
}

template <typename T, typename O, typename Conv>
__device__ __forceinline__ void load_rows16(const T* p, int64_t tile0, int64_t n,
                                            O (&v)[kItems], Conv conv) {
  T x[kItems];
  const bool full = tile0 + kTileRows <= n;
  const bool vec = (sizeof(T) <= 8) && ((uintptr_t)p % (sizeof(T) * 4 < 16 ? sizeof(T) * 4 : 16)) == 0;
  if (full && vec) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const T* r = p + run_row(tile0, 4 * q);
      if constexpr (sizeof(T) == 8) {
        const longlong2 a = *(const longlong2*)r, b = *(const longlong2*)(r + 2);
        x[4 * q] = __builtin_bit_cast(T, a.x); x[4 * q + 1] = __builtin_bit_cast(T, a.y);
        x[4 * q + 2] = __builtin_bit_cast(T, b.x); x[4 * q + 3] = __builtin_bit_cast(T, b.y);
      } else if constexpr (sizeof(T) == 4) {
        const uint4 a = *(const uint4*)r;
        x[4 * q] = __builtin_bit_cast(T, a.x); x[4 * q + 1] = __builtin_bit_cast(T, a.y);
        x[4 * q + 2] = __builtin_bit_cast(T, a.z); x[4 * q + 3] = __builtin_bit_cast(T, a.w);
      } else {
        // 1- and 2-byte values: one 4- or 8-byte load of the run
        T tmp[4];
        if constexpr (sizeof(T) == 2) {
          const uint2 a = *(const uint2*)r;
          __builtin_memcpy(tmp, &a, 8);
        } else {
          const uint32_t a = *(const uint32_t*)r;
          __builtin_memcpy(tmp, &a, 4);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) x[4 * q + i] = tmp[i];
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const int64_t r = run_row(tile0, j);
      x[j] = p[r < n ? r : 0];
    }
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) v[j] = run_row(tile0, j) < n ? (O)conv(x[j]) : (O)0;
}

// Column values of the thread's 16 tile rows as the expression programs hold them: integers
// sign-extended, bool as 0/1, f64 as its bits, f32 as its bits in the low half; rows past n
// read as 0.
__device__ __forceinline__ void load_raw16(const void* p, int dt, int64_t tile0, int64_t n,
                                           uint64_t (&v)[kItems]) {
  switch (dt) {
    case TQ_BOOL:
      load_rows16((const uint8_t*)p, tile0, n, v, [](uint8_t x) { return (uint64_t)(x != 0); });
      break;
    case TQ_I8:
      load_rows16((const int8_t*)p, tile0, n, v, [](int8_t x) { return (uint64_t)(int64_t)x; });
      break;
    case TQ_I16:
      load_rows16((const int16_t*)p, tile0, n, v, [](int16_t x) { return (uint64_t)(int64_t)x; });
      break;
    case TQ_I32:
      load_rows16((const int32_t*)p, tile0, n, v, [](int32_t x) { return (uint64_t)(int64_t)x; });
      break;
    case TQ_F32:
      load_rows16((const uint32_t*)p, tile0, n, v, [](uint32_t x) { return (uint64_t)x; });
      break;
    default:
      load_rows16((const uint64_t*)p, tile0, n, v, [](uint64_t x) { return x; });
      break;
  }
}

// The same in 32 bits, for columns of at most 4 bytes (TQ_BOOL..TQ_I32, TQ_F32): integers
// sign-extended to 32 bits, bool as 0/1, f32 as its bits.
__device__ __forceinline__ void load_raw16_32(const void* p, int dt, int64_t tile0, int64_t n,
                                              uint32_t (&v)[kItems]) {
  switch (dt) {
    case TQ_BOOL:
      load_rows16((const uint8_t*)p, tile0, n, v, [](uint8_t x) { return (uint32_t)(x != 0); });
      break;
    case TQ_I8:
      load_rows16((const int8_t*)p, tile0, n, v, [](int8_t x) { return (uint32_t)(int32_t)x; });
      break;
    case TQ_I16:
      load_rows16((const int16_t*)p, tile0, n, v, [](int16_t x) { return (uint32_t)(int32_t)x; });
      break;
    default:  // TQ_I32, TQ_F32
      load_rows16((const uint32_t*)p, tile0, n, v, [](uint32_t x) { return x; });
      break;
  }
}

// Selection bits of the thread's 16 tile rows (bit j: the row exists and, with a mask, its
// byte is set).
__device__ __forceinline__ uint32_t load_sel16(const uint8_t* mask, int64_t tile0, int64_t n) {
  uint32_t exist = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) exist |= (run_row(tile0, j) < n ? 1u : 0u) << j;
  if (!mask) return exist;
  uint64_t m[kItems];
  load_rows16(mask, tile0, n, m, [](uint8_t x) { return (uint64_t)(x != 0); });
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) s |= (m[j] ? 1u : 0u) << j;
  return s & exist;
}

// ------------------------------------------------------------------ issue-only tile loads
// The loaders above convert each value right after loading it, inside a switch on the dtype, so
// a kernel that calls them for several columns waits for one column's loads before issuing the
// next column's.  For a full tile of a column aligned to its run (four values), issue_runs only
// issues the four run loads of the thread (4 bytes per run for 1-byte values, 8 for 2-byte, 16
// for 4-byte) and touches nothing it loaded; run_value32 converts afterwards.  A kernel issues
// every column's runs, then converts, so all its loads are in flight together.
template <int W> struct RunRaw;
template <> struct RunRaw<1> { uint32_t r[4]; };
template <> struct RunRaw<2> { uint2 r[4]; };
template <> struct RunRaw<4> { uint4 r[4]; };

template <int W>
__device__ __forceinline__ void issue_runs(const void* p, int64_t tile0, RunRaw<W>& out) {
  const uint8_t* b = (const uint8_t*)p;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t row = run_row(tile0, 4 * q);
    if constexpr (W == 1) out.r[q] = *(const uint32_t*)(b + row);
    else if constexpr (W == 2) out.r[q] = *(const uint2*)(b + 2 * row);
    else out.r[q] = *(const uint4*)(b + 4 * row);
  }
}

// Value j (of the thread's 16 tile rows) in load_raw16_32's convention: bool as 0/1, 1- and
// 2-byte integers sign-extended, 4-byte values as their bits.  j is a compile-time index after
// unrolling, so the run and lane selects fold away.
template <int W>
__device__ __forceinline__ uint32_t run_value32(const RunRaw<W>& x, int j, int dt) {
  const int q = j >> 2, i = j & 3;
  if constexpr (W == 1) {
    const uint32_t byte = (x.r[q] >> (8 * i)) & 0xFFu;
    return dt == TQ_BOOL ? (uint32_t)(byte != 0) : (uint32_t)(int32_t)(int8_t)byte;
  } else if constexpr (W == 2) {
    const uint32_t w = (i < 2 ? x.r[q].x : x.r[q].y) >> (16 * (i & 1));
    return (uint32_t)(int32_t)(int16_t)(w & 0xFFFFu);
  } else {
    return i == 0 ? x.r[q].x : i == 1 ? x.r[q].y : i == 2 ? x.r[q].z : x.r[q].w;
  }
}

// Bytes per value of a dtype code, on either side.
__host__ __device__ __forceinline__ int value_bytes(int dtype) {
  return dtype <= TQ_I8 ? 1 : dtype == TQ_I16 ? 2 : (dtype == TQ_I32 || dtype == TQ_F32) ? 4 : 8;
}

// A column whose full tiles issue_runs can read: W-byte values with the run (4 * W bytes) aligned.
__device__ __forceinline__ bool runs_ok(const void* p, int w) {
  return w <= 4 && ((uintptr_t)p % (uintptr_t)(4 * w)) == 0;
}

inline int dtype_bytes(int dtype) {
  switch (dtype) {
    case TQ_BOOL: case TQ_I8: return 1;
    case TQ_I16: return 2;
    case TQ_I32: case TQ_F32: return 4;
    default: return 8;
  }
}

// Copy one row of `bytes` bytes (power-of-two widths take a single load/store).
__device__ __forceinline__ void copy_row(void* dst, int64_t di, const void* src, int64_t si,
                                         int bytes) {
  switch (bytes) {
    case 1: ((uint8_t*)dst)[di] = ((const uint8_t*)src)[si]; break;
    case 2: ((uint16_t*)dst)[di] = ((const uint16_t*)src)[si]; break;
    case 4: ((uint32_t*)dst)[di] = ((const uint32_t*)src)[si]; break;
    case 8: ((uint64_t*)dst)[di] = ((const uint64_t*)src)[si]; break;
    default: {
      const uint8_t* s = (const uint8_t*)src + si * bytes;
      uint8_t* d = (uint8_t*)dst + di * bytes;
      for (int b = 0; b < bytes; ++b) d[b] = s[b];
    }
  }
}

// Column pointer bundles passed by value as kernel arguments.
struct ColPack {
  const void* src[TQ_MAX_COLS];
  void* dst[TQ_MAX_COLS];
  int32_t bytes[TQ_MAX_COLS];
  int32_t fkind[TQ_MAX_COLS];  // 4 = f32, 8 = f64, 0 = integer/bool/bytes (null-fill type)
  int32_t n;
};

// Write the NULL sentinel (-64) into row di of a column of the given width/kind.
__device__ __forceinline__ void write_null(void* dst, int64_t di, int bytes, int fkind) {
  if (fkind == 8) {
    ((double*)dst)[di] = -64.0;
  } else if (fkind == 4) {
    ((float*)dst)[di] = -64.0f;
  } else {
    switch (bytes) {
      case 1: ((int8_t*)dst)[di] = -64; break;
      case 2: ((int16_t*)dst)[di] = -64; break;
      case 4: ((int32_t*)dst)[di] = -64; break;
      case 8: ((int64_t*)dst)[di] = -64; break;
      default: {
        int8_t* d = (int8_t*)dst + di * bytes;
        for (int b = 0; b < bytes; ++b) d[b] = -64;
      }
    }
  }
}

// Graph-capture-safe fills (a kernel, never hipMemsetAsync): count elements of value v.
int fill_u32(void* p, uint32_t v, int64_t count, hipStream_t s);
int fill_u64(void* p, uint64_t v, int64_t count, hipStream_t s);
unsigned stream_grid(int64_t m);

// Workgroups of `kernel` (kBlock threads, `lds` bytes of dynamic LDS) resident on the current
// device at once: occupancy per CU times the CUs (cached per device, kernel and LDS size).
int64_t resident_blocks(const void* kernel, size_t lds, int block = kBlock);

// Grid of a kernel whose workgroups stride over `items` work items: one resident round, so no
// workgroup starts only after others have finished all of theirs.
template <typename... P>
unsigned resident_grid(void (*kernel)(P...), size_t lds, int64_t items) {
  const int64_t r = resident_blocks((const void*)kernel, lds);
  const int64_t g = items < r ? items : r;
  return (unsigned)(g < 1 ? 1 : g);
}
// Length of a scan whose element count a producer earlier on the stream left in device memory:
// with n_dev, the elements are the first mult * ceil(min(*n_dev, cap) / per) of the n the launch
// was sized for (a histogram of per-tile counts whose tile count follows a device row count).
struct ScanLen {
  const int64_t* n_dev = nullptr;
  int64_t cap = 0, per = 1, mult = 1;
  __device__ __forceinline__ int64_t of(int64_t n) const {
    if (!n_dev) return n;
    int64_t r = *n_dev;
    r = r < 0 ? 0 : (r < cap ? r : cap);
    const int64_t e = mult * ((r + per - 1) / per);
    return e < n ? e : n;
  }
};

template <typename TIn, typename TOut>
int scan_exclusive(const TIn* in, int64_t n, TOut* out, TOut* total, TOut* ws, hipStream_t stream,
                   ScanLen len = ScanLen{});

// Fill a host-side ColPack from caller arrays; returns TQ_EINVAL text on bad widths.
int make_colpack(ColPack* cp, int32_t ncols, const void* const* src, const int32_t* bytes,
                 const int32_t* fkind, void* const* dst);

}  // namespace tq
