// Aggregate accumulators shared by the group-by and batched-pipeline kernels: argument values
// are converted into an accumulator domain (f64 sums/min/max, exact int64 sums, counts) and
// combined with device-scope atomics, which also merge LDS partials into global tables.
#pragma once

#include <cstring>

#include "tq_common.h"
#include "tq_expr.h"

namespace tq {

enum AccKind { ACC_FSUM = 0, ACC_ISUM, ACC_COUNT, ACC_FMIN, ACC_FMAX, ACC_IMIN, ACC_IMAX };

constexpr unsigned long long kEmpty = ~0ull;
constexpr int kLdsBudget = 64 * 1024;
constexpr int kLdsBig = 152 * 1024;  // gfx950 gives a workgroup up to 160 KB of LDS on request

enum DirectForm { DF_X = 0, DF_ONE_MINUS, DF_ONE_PLUS };
constexpr int kDirectFactors = 4;

// Direct arguments (direct != 0, set by match_direct): every aggregate's argument is either a
// constant, whose accumulator bits are dconst[a] (nf[a] == 0), or the left-to-right product of
// nf[a] factors over f64 load slots (factor i reads slot fslot[a][i] as x and is x, 1 - x or
// 1 + x by fform[a][i]), which the kernels evaluate from the staged loads without the
// instruction interpreter.
struct DirectArgs {
  int32_t direct;
  int8_t nf[TQ_GB_MAX_AGGS];
  int8_t fslot[TQ_GB_MAX_AGGS][kDirectFactors];
  int8_t fform[TQ_GB_MAX_AGGS][kDirectFactors];
  uint64_t dconst[TQ_GB_MAX_AGGS];
};

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
  DirectArgs D;
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

// The accumulator-domain values of aggregate a for the thread's kSub staged rows, evaluated
// directly (G.direct): the same operations in the same order as the argument's program, so the
// values are bit-identical to the interpreter's.
__device__ __forceinline__ void direct_args(const DirectArgs& D, int kind, int a,
                                            const uint64_t* stage, uint64_t (&v)[kSub]) {
  const int nf = D.nf[a];
  if (nf == 0) {
#pragma unroll
    for (int j = 0; j < kSub; ++j) v[j] = D.dconst[a];
    return;
  }
  double p[kSub];
  for (int i = 0; i < nf; ++i) {
    const int s = D.fslot[a][i], form = D.fform[a][i];
#pragma unroll
    for (int j = 0; j < kSub; ++j) {
      const double x = as_f64(stage[(s * kSub + j) * kBlock + threadIdx.x]);
      const double f = form == DF_X ? x : form == DF_ONE_MINUS ? 1.0 - x : 1.0 + x;
      p[j] = i == 0 ? f : p[j] * f;
    }
  }
  // the accumulator domain chosen once for the rows
  switch (kind) {
    case ACC_COUNT:
#pragma unroll
      for (int j = 0; j < kSub; ++j) v[j] = (uint64_t)(p[j] != -64.0);
      break;
    case ACC_ISUM: case ACC_IMIN: case ACC_IMAX:
#pragma unroll
      for (int j = 0; j < kSub; ++j) v[j] = to_acc(kind, TQ_F64, f64_bits(p[j]));
      break;
    default:
#pragma unroll
      for (int j = 0; j < kSub; ++j) v[j] = f64_bits(p[j]);
      break;
  }
}

// Host twin of to_acc (for a constant argument's accumulator bits).
inline uint64_t to_acc_host(int kind, int dt, uint64_t v) {
  const bool isf = dt == TQ_F64 || dt == TQ_F32;
  double d;
  if (dt == TQ_F64) {
    memcpy(&d, &v, 8);
  } else if (dt == TQ_F32) {
    float f;
    const uint32_t lo = (uint32_t)v;
    memcpy(&f, &lo, 4);
    d = (double)f;
  } else {
    d = 0.0;
  }
  uint64_t out;
  switch (kind) {
    case ACC_FSUM: case ACC_FMIN: case ACC_FMAX: {
      const double x = isf ? d : (double)(int64_t)v;
      memcpy(&out, &x, 8);
      return out;
    }
    case ACC_COUNT: return isf ? (uint64_t)(d != -64.0) : (uint64_t)((int64_t)v != -64);
    default: return v;
  }
}

// Recognise aggregate arguments that direct_args can evaluate: per segment of the program, a
// constant, an f64 column, 1 - column, 1 + column, or a left-deep product of up to
// kDirectFactors such factors (what TPC-H's sums over price, discount and tax compile to).
// Follows the registers the way match_conj does; anything else leaves G->direct at 0.
inline void match_direct(const tq_program_t& P, int naggs, const int32_t* kind,
                         const int32_t* arg_dtype, DirectArgs* D) {
  enum { UNK, COL, CST, PROD };
  struct RV {
    int kind;
    int slot;
    uint64_t bits;
    int nf;
    int8_t fs[kDirectFactors], ff[kDirectFactors];
  };
  uint64_t one;
  {
    const double d = 1.0;
    memcpy(&one, &d, 8);
  }
  D->direct = 0;
  if (P.nseg < naggs) return;
  for (int a = 0; a < naggs; ++a) {
    RV R[TQ_NREG];
    for (int i = 0; i < TQ_NREG; ++i) R[i] = RV{UNK, 0, 0, 0, {0}, {0}};
    const int begin = a == 0 ? 0 : P.seg_end[a - 1], end = P.seg_end[a];
    if (begin < 0 || end > P.n_instr || end > TQ_MAX_INSTR || begin > end) return;
    auto factor = [&](const RV& x) {  // a column as a one-factor product
      if (x.kind != COL) return x;
      RV f{PROD, 0, 0, 1, {(int8_t)x.slot}, {(int8_t)DF_X}};
      return f;
    };
    for (int pc = begin; pc < end; ++pc) {
      const tq_instr_t& I = P.code[pc];
      const int d = I.dst & (TQ_NREG - 1);
      const RV A = R[I.a & (TQ_NREG - 1)], B = R[I.b & (TQ_NREG - 1)];
      RV out{UNK, 0, 0, 0, {0}, {0}};
      switch (I.op) {
        case TQ_OP_NOP: continue;
        case TQ_OP_LOAD:
          if (I.a < P.nload && P.load_dtype[I.a] == TQ_F64) out = RV{COL, I.a, 0, 0, {0}, {0}};
          break;
        case TQ_OP_CONST: out = RV{CST, 0, (uint64_t)I.imm, 0, {0}, {0}}; break;
        case TQ_OP_SUB_F64:
          if (A.kind == CST && A.bits == one && B.kind == COL)
            out = RV{PROD, 0, 0, 1, {(int8_t)B.slot}, {(int8_t)DF_ONE_MINUS}};
          break;
        case TQ_OP_ADD_F64:
          if (A.kind == CST && A.bits == one && B.kind == COL)
            out = RV{PROD, 0, 0, 1, {(int8_t)B.slot}, {(int8_t)DF_ONE_PLUS}};
          else if (A.kind == COL && B.kind == CST && B.bits == one)
            out = RV{PROD, 0, 0, 1, {(int8_t)A.slot}, {(int8_t)DF_ONE_PLUS}};
          break;
        case TQ_OP_MUL_F64: {
          const RV L = factor(A), Rt = factor(B);
          if (L.kind == PROD && Rt.kind == PROD && Rt.nf == 1 && L.nf < kDirectFactors) {
            out = L;
            out.fs[out.nf] = Rt.fs[0];
            out.ff[out.nf] = Rt.ff[0];
            ++out.nf;
          }
          break;
        }
        default: break;
      }
      R[d] = out;
    }
    const RV res = factor(R[P.seg_result[a] & (TQ_NREG - 1)]);
    if (res.kind == CST) {
      D->nf[a] = 0;
      D->dconst[a] = to_acc_host(kind[a], arg_dtype[a], res.bits);
    } else if (res.kind == PROD && arg_dtype[a] == TQ_F64) {
      D->nf[a] = (int8_t)res.nf;
      for (int i = 0; i < res.nf; ++i) {
        D->fslot[a][i] = res.fs[i];
        D->fform[a][i] = res.ff[i];
      }
    } else {
      return;
    }
  }
  D->direct = 1;
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
