// Aggregate accumulators shared by the group-by and batched-pipeline kernels: argument values
// are converted into an accumulator domain (f64 sums/min/max, exact int64 sums, counts) and
// combined with device-scope atomics, which also merge LDS partials into global tables.
#pragma once

#include "tq_common.h"
#include "tq_expr.h"

namespace tq {

enum AccKind { ACC_FSUM = 0, ACC_ISUM, ACC_COUNT, ACC_FMIN, ACC_FMAX, ACC_IMIN, ACC_IMAX };

constexpr unsigned long long kEmpty = ~0ull;
constexpr int kLdsBudget = 64 * 1024;
constexpr int kLdsBig = 152 * 1024;  // gfx950 gives a workgroup up to 160 KB of LDS on request

struct GbSpec {
  int32_t nkeys, naggs;
  int32_t key_slot[TQ_GB_MAX_KEYS];  // load slot holding key i
  int64_t key_min[TQ_GB_MAX_KEYS];
  int64_t key_stride[TQ_GB_MAX_KEYS];
  int64_t key_range[TQ_GB_MAX_KEYS];
  int32_t key_dtype[TQ_GB_MAX_KEYS];
  int32_t kind[TQ_GB_MAX_AGGS];
  int32_t arg_dtype[TQ_GB_MAX_AGGS];
  uint64_t init[TQ_GB_MAX_AGGS];
  // key dictionaries (null: the digit is value - key_min): presence bits and exclusive ranks
  // per 32-value word over the value range, and the distinct values in ascending order
  const uint32_t* dict_bits[TQ_GB_MAX_KEYS];
  const uint32_t* dict_rank[TQ_GB_MAX_KEYS];
  const int64_t* dict_val[TQ_GB_MAX_KEYS];
};

// Packed-key digit of key k's value v: its offset in the key's range, or its rank among the
// key's distinct values when the key has a dictionary.
__device__ __forceinline__ uint64_t key_digit(const GbSpec& G, int k, int64_t v) {
  const uint64_t off = (uint64_t)(v - G.key_min[k]);
  if (!G.dict_bits[k]) return off;
  const uint32_t w = (uint32_t)(off >> 5), b = (uint32_t)(off & 31);
  return G.dict_rank[k][w] + (uint32_t)__popc(G.dict_bits[k][w] & ((1u << b) - 1u));
}

struct GbTable {
  uint32_t* cnt;
  unsigned long long* hkey;
  uint64_t* acc;  // acc[j * slots + slot]
  int64_t slots;
  int32_t log2;
};

__device__ __forceinline__ uint64_t mix(uint64_t p) { return p * 0x9E3779B97F4A7C15ull; }

// Convert an argument value (program register bits) into its accumulator domain.
__device__ __forceinline__ uint64_t to_acc(int kind, int dt, uint64_t v) {
  const bool isf = dt == TQ_F64 || dt == TQ_F32;
  const double d = dt == TQ_F64 ? as_f64(v) : (double)as_f32(v);
  switch (kind) {
    case ACC_FSUM: case ACC_FMIN: case ACC_FMAX: return f64_bits(isf ? d : (double)(int64_t)v);
    case ACC_COUNT: return isf ? (uint64_t)(d != -64.0) : (uint64_t)((int64_t)v != -64);
    default: return v;  // integer domain
  }
}

// Two accumulators of one kind combined (no atomics).
__device__ __forceinline__ uint64_t acc_combine(int kind, uint64_t a, uint64_t b) {
  switch (kind) {
    case ACC_FSUM: return f64_bits(as_f64(a) + as_f64(b));
    case ACC_ISUM:
    case ACC_COUNT: return a + b;
    case ACC_FMIN: return as_f64(b) < as_f64(a) ? b : a;
    case ACC_FMAX: return as_f64(b) > as_f64(a) ? b : a;
    case ACC_IMIN: return (int64_t)b < (int64_t)a ? b : a;
    default: return (int64_t)b > (int64_t)a ? b : a;
  }
}

__device__ __forceinline__ void acc_apply(uint64_t* a, int kind, uint64_t v) {
  switch (kind) {
    case ACC_FSUM: atomicAdd((double*)a, as_f64(v)); break;
    case ACC_ISUM:
    case ACC_COUNT: atomicAdd((unsigned long long*)a, (unsigned long long)v); break;
    case ACC_FMIN: atomicMin((double*)a, as_f64(v)); break;
    case ACC_FMAX: atomicMax((double*)a, as_f64(v)); break;
    case ACC_IMIN: atomicMin((long long*)a, (long long)v); break;
    default: atomicMax((long long*)a, (long long)v); break;
  }
}

__device__ __forceinline__ void store_int(void* p, int dt, int64_t i, int64_t v) {
  switch (dt) {
    case TQ_BOOL: ((uint8_t*)p)[i] = v != 0; break;
    case TQ_I8: ((int8_t*)p)[i] = (int8_t)v; break;
    case TQ_I16: ((int16_t*)p)[i] = (int16_t)v; break;
    case TQ_I32: ((int32_t*)p)[i] = (int32_t)v; break;
    default: ((int64_t*)p)[i] = v; break;
  }
}

// Accumulator kind and initial bits of an aggregate (host side).
int agg_kind_init(int op, int dtype, int32_t* kind, uint64_t* init);

// Final value of aggregate slot: writes row o of out in the output dtype of (op, dtype).
__device__ __forceinline__ void write_agg(void* out, int64_t o, int op, int kind, int dt,
                                          uint64_t acc, uint32_t rows) {
  switch (op) {
    case TQ_AGG_AVG: {
      const double sum = kind == ACC_FSUM ? as_f64(acc) : (double)(int64_t)acc;
      const double r = sum / (double)rows;
      if (dt == TQ_F32) ((float*)out)[o] = (float)r; else ((double*)out)[o] = r;
      break;
    }
    case TQ_AGG_COUNT: ((int64_t*)out)[o] = (int64_t)acc; break;
    case TQ_AGG_SUM:
      if (kind == ACC_ISUM) ((int64_t*)out)[o] = (int64_t)acc;
      else if (dt == TQ_F32) ((float*)out)[o] = (float)as_f64(acc);
      else ((double*)out)[o] = as_f64(acc);
      break;
    default:  // MIN / MAX in the argument dtype
      if (dt == TQ_F64) ((double*)out)[o] = as_f64(acc);
      else if (dt == TQ_F32) ((float*)out)[o] = (float)as_f64(acc);
      else store_int(out, dt, o, (int64_t)acc);
      break;
  }
}

}  // namespace tq
