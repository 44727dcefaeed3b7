// Per-row evaluation of host-compiled expression programs (predicates and aggregate arguments).
//
// Two phases per sub-tile of kSubRows rows (thread t owns rows sub0 + j*256 + t, j < kSub):
//  1. stage: for each load slot of the program, every thread issues its kSub independent,
//     coalesced loads and parks the values (widened to 64 bits) in LDS.  All loads of a slot
//     are in flight together, which is what keeps HBM busy.
//  2. interpret: for each of its rows, the thread runs the program on an 8-register file.  The
//     program sits in the kernel-argument segment, so instruction fetches are scalar loads and
//     the opcode switch is a uniform branch; a single row's register file is small enough to
//     stay in VGPRs (indexed moves), so nothing spills to scratch.
// Each thread only reads back LDS words it wrote itself, so no barrier separates the phases.
//
// Arithmetic matches the ATen kernels the reference runs (one rounding per operation, no FMA
// contraction: the library is built with -ffp-contract=off).
#pragma once

#include "tq_common.h"

namespace tq {

constexpr int kSub = 4;                  // rows per thread per staged sub-tile
constexpr int kSubRows = kBlock * kSub;  // 1024 rows

__device__ __forceinline__ double as_f64(uint64_t b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ uint64_t f64_bits(double d) { return (uint64_t)__double_as_longlong(d); }
__device__ __forceinline__ float as_f32(uint64_t b) { return __uint_as_float((uint32_t)b); }
__device__ __forceinline__ uint64_t f32_bits(float f) { return (uint64_t)__float_as_uint(f); }

__device__ __forceinline__ const void* col_ptr(const tq_colset_t& C, int i) {
  const void* p = C.ptr[0];
#pragma unroll
  for (int j = 1; j < TQ_MAX_COLS; ++j) p = (i == j) ? C.ptr[j] : p;
  return p;
}
__device__ __forceinline__ int col_width(const tq_colset_t& C, int i) {
  int w = C.width[0];
#pragma unroll
  for (int j = 1; j < TQ_MAX_COLS; ++j) w = (i == j) ? C.width[j] : w;
  return w;
}

// Day count since 1990-01-01 -> calendar year, as torch.bucketize(right=True) over the
// cumulative year lengths 1990..1999 indexes year_range (conversion.py:10-11).
__device__ __forceinline__ int64_t year_of_day(int64_t d) {
  return 1990 + (d >= 365) + (d >= 730) + (d >= 1096) + (d >= 1461) + (d >= 1826) +
         (d >= 2191) + (d >= 2557) + (d >= 2922) + (d >= 3287) + (d >= 3652);
}

// Bytes of LDS the stage needs for a program.
inline size_t stage_bytes(const tq_program_t& P) {
  return (size_t)(P.nload > 0 ? P.nload : 1) * kSubRows * sizeof(uint64_t);
}

// One load form for every column width: the aligned 8-byte word that holds value r of a w-byte
// column (w = 1, 2, 4 or 8).  A width switch around loads of different sizes makes the compiler
// merge their results into one register, which waits for each load before the next is issued;
// one load form has no merge, so every slot's loads are in flight together.  An aligned 8-byte
// word lies inside the 4 KB page of the value it holds, so it reads no page the value's own load
// would not.  word_bits extracts the value afterwards.
__device__ __forceinline__ uint64_t load_word(const void* col, int w, int64_t r) {
  // pointer arithmetic from col (not an integer round trip) keeps the global address space, so
  // the load is a global load, counted only by vmcnt
  const uint8_t* p = (const uint8_t*)col + r * (int64_t)w;
  return *(const uint64_t*)(p - ((uintptr_t)p & 7));
}
__device__ __forceinline__ uint64_t word_bits(uint64_t word, const void* col, int w, int64_t r) {
  if (w == 8) return word;
  const uintptr_t a = (uintptr_t)col + (uintptr_t)(r * (int64_t)w);
  return (word >> (8 * (a & 7))) & ((1ull << (8 * w)) - 1);
}

// Raw bits of a dt-typed value -> the program's convention: integers sign-extended to 64 bits,
// bool as 0/1, f64 as its bits, f32 as its bits in the low half.
__device__ __forceinline__ uint64_t bits_to_value(int dt, uint64_t b) {
  switch (dt) {
    case TQ_BOOL: return (uint64_t)((b & 0xFFu) != 0);
    case TQ_I8: return (uint64_t)(int64_t)(int8_t)(uint8_t)b;
    case TQ_I16: return (uint64_t)(int64_t)(int16_t)(uint16_t)b;
    case TQ_I32: return (uint64_t)(int64_t)(int32_t)(uint32_t)b;
    case TQ_F32: return (uint64_t)(uint32_t)b;
    default: return b;
  }
}

// Every load slot's values of rows[j] (j < kSub) into the stage: all slots' loads are issued
// before any is converted (a per-slot load-then-store would wait for each slot in turn), then
// each value is converted and parked.  Rows with live[j] == false are zeroed (the caller points
// them at a row that exists).
__device__ __forceinline__ void stage_rows(const tq_program_t& P, const tq_colset_t& C,
                                           const int64_t (&rows)[kSub], const bool (&live)[kSub],
                                           uint64_t* __restrict__ stage) {
  const int t = threadIdx.x;
  static_assert(TQ_MAX_LOADS == 8, "one line per load slot below");
  const int n = P.nload;
  // straight-line per slot, so every index into raw is a constant whatever the inliner does
  uint64_t raw[TQ_MAX_LOADS][kSub];
#define TQ_ISSUE(S)                                                                 \
  if (S < n) {                                                                      \
    const void* col = col_ptr(C, P.load_col[S]);                                    \
    const int w = value_bytes(P.load_dtype[S]);                                     \
    raw[S][0] = load_word(col, w, rows[0]);                                         \
    raw[S][1] = load_word(col, w, rows[1]);                                         \
    raw[S][2] = load_word(col, w, rows[2]);                                         \
    raw[S][3] = load_word(col, w, rows[3]);                                         \
  }
  TQ_ISSUE(0) TQ_ISSUE(1) TQ_ISSUE(2) TQ_ISSUE(3) TQ_ISSUE(4) TQ_ISSUE(5) TQ_ISSUE(6) TQ_ISSUE(7)
#undef TQ_ISSUE
  // the dtype is switched on once per slot, outside the rows
#define TQ_PARK1(S, J, DT, W)                                                              \
  stage[(S * kSub + J) * kBlock + t] =                                                     \
      live[J] ? bits_to_value(DT, word_bits(raw[S][J], col, W, rows[J])) : 0;
#define TQ_PARKDT(S, DT, W) \
  TQ_PARK1(S, 0, DT, W) TQ_PARK1(S, 1, DT, W) TQ_PARK1(S, 2, DT, W) TQ_PARK1(S, 3, DT, W)
#define TQ_PARK(S)                                                                         \
  if (S < n) {                                                                             \
    const void* col = col_ptr(C, P.load_col[S]);                                           \
    switch (P.load_dtype[S]) {                                                             \
      case TQ_BOOL: TQ_PARKDT(S, TQ_BOOL, 1) break;                                        \
      case TQ_I8: TQ_PARKDT(S, TQ_I8, 1) break;                                            \
      case TQ_I16: TQ_PARKDT(S, TQ_I16, 2) break;                                          \
      case TQ_I32: TQ_PARKDT(S, TQ_I32, 4) break;                                          \
      case TQ_F32: TQ_PARKDT(S, TQ_F32, 4) break;                                          \
      case TQ_I64: TQ_PARKDT(S, TQ_I64, 8) break;                                          \
      default: TQ_PARKDT(S, TQ_F64, 8) break;                                              \
    }                                                                                      \
  }
  TQ_PARK(0) TQ_PARK(1) TQ_PARK(2) TQ_PARK(3) TQ_PARK(4) TQ_PARK(5) TQ_PARK(6) TQ_PARK(7)
#undef TQ_PARK
#undef TQ_PARKDT
#undef TQ_PARK1
}

// Phase 1: stage every load slot for the sub-tile starting at row sub0.  Rows with
// live[j] == false (past the end, or unselected by a mask) read row 0 of the column (which
// exists whenever a sub-tile does) and are zeroed, so the loads are unconditional.
__device__ __forceinline__ void stage_loads(const tq_program_t& P, const tq_colset_t& C,
                                            int64_t sub0, const bool (&live)[kSub],
                                            uint64_t* __restrict__ stage) {
  int64_t rows[kSub];
#pragma unroll
  for (int j = 0; j < kSub; ++j) rows[j] = live[j] ? sub0 + (int64_t)j * kBlock + threadIdx.x : 0;
  stage_rows(P, C, rows, live, stage);
}

// The same staging one load slot at a time (each slot's loads, then its values parked, then the
// next slot's loads).  The batched kernel stages this way: its phase functions were compiled
// and validated with it, and with stage_rows they compile as out-of-line calls.
__device__ __forceinline__ void stage_loads_by_slot(const tq_program_t& P, const tq_colset_t& C,
                                                    int64_t sub0, const bool (&live)[kSub],
                                                    uint64_t* __restrict__ stage) {
  const int t = threadIdx.x;
  for (int s = 0; s < P.nload; ++s) {
    const void* col = col_ptr(C, P.load_col[s]);
    const int dt = P.load_dtype[s];
    uint64_t v[kSub];
#define TQ_STAGE(T, EXPR)                                          \
  {                                                                \
    const T* p = (const T*)col;                                    \
    _Pragma("unroll") for (int j = 0; j < kSub; ++j) {             \
      const int64_t r = live[j] ? sub0 + (int64_t)j * kBlock + t : 0; \
      v[j] = (EXPR);                                               \
    }                                                              \
    _Pragma("unroll") for (int j = 0; j < kSub; ++j)               \
      if (!live[j]) v[j] = 0;                                      \
  }
    switch (dt) {
      case TQ_BOOL: TQ_STAGE(uint8_t, (uint64_t)(p[r] != 0)) break;
      case TQ_I8: TQ_STAGE(int8_t, (uint64_t)(int64_t)p[r]) break;
      case TQ_I16: TQ_STAGE(int16_t, (uint64_t)(int64_t)p[r]) break;
      case TQ_I32: TQ_STAGE(int32_t, (uint64_t)(int64_t)p[r]) break;
      case TQ_F32: TQ_STAGE(uint32_t, (uint64_t)p[r]) break;
      default: TQ_STAGE(uint64_t, p[r]) break;
    }
#undef TQ_STAGE
#pragma unroll
    for (int j = 0; j < kSub; ++j) stage[(s * kSub + j) * kBlock + t] = v[j];
  }
}

// '%w1%w2%' over one row's string bytes sp[0, w): the positions scanned as like_mask does
// (operators/like.py:15-25): at each pos, w1 seen at or before pos, then w2 at pos + len1.
__device__ __forceinline__ bool str_like2(const int8_t* sp, int w, const uint8_t* pool,
                                          int64_t imm) {
  const int off1 = (int)(imm & 0xffff), len1 = (int)((imm >> 16) & 0xffff);
  const int off2 = (int)((imm >> 32) & 0xffff), len2 = (int)((imm >> 48) & 0xffff);
  bool appear = false;
  for (int pos = 0; pos + len1 + len2 <= w; ++pos) {
    if (!appear) {
      bool m = true;
      for (int c = 0; c < len1 && m; ++c) m = sp[pos + c] == (int8_t)pool[off1 + c];
      appear = m;
    }
    if (appear) {
      bool m = true;
      for (int c = 0; c < len2 && m; ++c) m = sp[pos + len1 + c] == (int8_t)pool[off2 + c];
      if (m) return true;
    }
  }
  return false;
}

// pool[off:off+len] at some position of one row's string bytes sp[0, w) (operators/like.py:3-11).
__device__ __forceinline__ bool str_contains(const int8_t* sp, int w, const uint8_t* pool,
                                             int64_t imm) {
  const int off = (int)(imm & 0xffffffff), len = (int)(imm >> 32);
  for (int pos = 0; pos + len <= w; ++pos) {
    bool m = true;
    for (int c = 0; c < len && m; ++c) m = sp[pos + c] == (int8_t)pool[off + c];
    if (m) return true;
  }
  return false;
}

// One row's eight registers, kept as named scalars and selected by switch: an indexable array
// would be placed in private scratch memory by the compiler.
struct Regs {
  uint64_t r0, r1, r2, r3, r4, r5, r6, r7;
  __device__ __forceinline__ uint64_t get(int i) const {
    switch (i & 7) {
      case 0: return r0;
      case 1: return r1;
      case 2: return r2;
      case 3: return r3;
      case 4: return r4;
      case 5: return r5;
      case 6: return r6;
      default: return r7;
    }
  }
  __device__ __forceinline__ void put(int i, uint64_t v) {
    switch (i & 7) {
      case 0: r0 = v; break;
      case 1: r1 = v; break;
      case 2: r2 = v; break;
      case 3: r3 = v; break;
      case 4: r4 = v; break;
      case 5: r5 = v; break;
      case 6: r6 = v; break;
      default: r7 = v; break;
    }
  }
};

// Phase 2: run instructions [begin, end) for row j of the sub-tile (global row `row`).
__device__ __forceinline__ void run_program(const tq_program_t& P, const tq_colset_t& C, int begin,
                                            int end, int j, int64_t row, bool valid,
                                            const uint64_t* __restrict__ stage, Regs& R) {
  const int t = threadIdx.x;
  for (int pc = begin; pc < end; ++pc) {
    const tq_instr_t I = P.code[pc];
    const uint64_t a = R.get(I.a);
    const uint64_t b = R.get(I.b);
    uint64_t o;
    switch (I.op) {
      case TQ_OP_LOAD: o = stage[(I.a * kSub + j) * kBlock + t]; break;
      case TQ_OP_CONST: o = (uint64_t)I.imm; break;
      case TQ_OP_STR_LIKE2:
      case TQ_OP_STR_CONTAINS: {
        const int stride = col_width(C, I.a);
        const int8_t* sp = (const int8_t*)col_ptr(C, I.a) + row * (int64_t)stride;
        const int w = I.c ? (int)I.c : stride;
        o = valid && (I.op == TQ_OP_STR_LIKE2 ? str_like2(sp, w, P.strpool, I.imm)
                                              : str_contains(sp, w, P.strpool, I.imm));
        break;
      }
      case TQ_OP_STR_EQ:
      case TQ_OP_STR_PREFIX: {
        const int8_t* p = (const int8_t*)col_ptr(C, I.a);
        const int stride = col_width(C, I.a);
        const int w = I.c ? (int)I.c : stride;
        const int off = (int)(I.imm & 0xffffffff);
        const int len = (int)(I.imm >> 32);
        const int upto = (I.op == TQ_OP_STR_EQ) ? w : len;
        bool eq = valid;
        if (eq) {
          const int8_t* s = p + row * (int64_t)stride;
          for (int c = 0; c < upto; ++c) {
            const int8_t want = c < len ? (int8_t)P.strpool[off + c] : (int8_t)0;
            eq = eq && (s[c] == want);
          }
        }
        o = eq;
        break;
      }
      case TQ_OP_CVT_I64_F64: o = f64_bits((double)(int64_t)a); break;
      case TQ_OP_CVT_I64_F32: o = f32_bits((float)(int64_t)a); break;
      case TQ_OP_CVT_F32_F64: o = f64_bits((double)as_f32(a)); break;
      case TQ_OP_CVT_F64_F32: o = f32_bits((float)as_f64(a)); break;
      case TQ_OP_WRAP_I8: o = (uint64_t)(int64_t)(int8_t)a; break;
      case TQ_OP_WRAP_I16: o = (uint64_t)(int64_t)(int16_t)a; break;
      case TQ_OP_WRAP_I32: o = (uint64_t)(int64_t)(int32_t)a; break;
      case TQ_OP_ADD_I64: o = a + b; break;
      case TQ_OP_SUB_I64: o = a - b; break;
      case TQ_OP_MUL_I64: o = a * b; break;
      case TQ_OP_ADD_F64: o = f64_bits(as_f64(a) + as_f64(b)); break;
      case TQ_OP_SUB_F64: o = f64_bits(as_f64(a) - as_f64(b)); break;
      case TQ_OP_MUL_F64: o = f64_bits(as_f64(a) * as_f64(b)); break;
      case TQ_OP_DIV_F64: o = f64_bits(as_f64(a) / as_f64(b)); break;
      case TQ_OP_ADD_F32: o = f32_bits(as_f32(a) + as_f32(b)); break;
      case TQ_OP_SUB_F32: o = f32_bits(as_f32(a) - as_f32(b)); break;
      case TQ_OP_MUL_F32: o = f32_bits(as_f32(a) * as_f32(b)); break;
      case TQ_OP_DIV_F32: o = f32_bits(as_f32(a) / as_f32(b)); break;
      case TQ_OP_LT_I64: o = (int64_t)a < (int64_t)b; break;
      case TQ_OP_LE_I64: o = (int64_t)a <= (int64_t)b; break;
      case TQ_OP_GT_I64: o = (int64_t)a > (int64_t)b; break;
      case TQ_OP_GE_I64: o = (int64_t)a >= (int64_t)b; break;
      case TQ_OP_EQ_I64: o = a == b; break;
      case TQ_OP_NE_I64: o = a != b; break;
      case TQ_OP_LT_F64: o = as_f64(a) < as_f64(b); break;
      case TQ_OP_LE_F64: o = as_f64(a) <= as_f64(b); break;
      case TQ_OP_GT_F64: o = as_f64(a) > as_f64(b); break;
      case TQ_OP_GE_F64: o = as_f64(a) >= as_f64(b); break;
      case TQ_OP_EQ_F64: o = as_f64(a) == as_f64(b); break;
      case TQ_OP_NE_F64: o = as_f64(a) != as_f64(b); break;
      case TQ_OP_LT_F32: o = as_f32(a) < as_f32(b); break;
      case TQ_OP_LE_F32: o = as_f32(a) <= as_f32(b); break;
      case TQ_OP_GT_F32: o = as_f32(a) > as_f32(b); break;
      case TQ_OP_GE_F32: o = as_f32(a) >= as_f32(b); break;
      case TQ_OP_EQ_F32: o = as_f32(a) == as_f32(b); break;
      case TQ_OP_NE_F32: o = as_f32(a) != as_f32(b); break;
      case TQ_OP_ISCLOSE_F64: {
        // torch.isclose: a == b, or |a-b| finite and <= atol + |rtol * b| with rtol = 1e-5
        const double x = as_f64(a), y = as_f64(b);
        const double err = fabs(x - y);
        const double allowed = __longlong_as_double(I.imm) + fabs(1e-5 * y);
        o = (x == y) || (isfinite(err) && err <= allowed);
        break;
      }
      case TQ_OP_AND: o = (a != 0) && (b != 0); break;
      case TQ_OP_OR: o = (a != 0) || (b != 0); break;
      case TQ_OP_NOT: o = a == 0; break;
      case TQ_OP_WHERE: o = R.get(I.c) ? a : b; break;
      case TQ_OP_YEAR: o = (uint64_t)year_of_day((int64_t)a); break;
      default: o = 0; break;
    }
    R.put(I.dst, o);
  }
}

// Gather the program's load slots at rows[j] (live rows only) into the stage, in the layout
// stage_loads uses, so run_rows can evaluate them.  Rows that are not live read row 0.
__device__ __forceinline__ void gather_loads(const tq_program_t& P, const tq_colset_t& C,
                                             const int64_t (&rows)[kSub],
                                             const bool (&live)[kSub],
                                             uint64_t* __restrict__ stage) {
  int64_t r[kSub];
#pragma unroll
  for (int j = 0; j < kSub; ++j) r[j] = live[j] ? rows[j] : 0;
  stage_rows(P, C, r, live, stage);
}

// The kSub rows of a thread's sub-tile together: one Regs per row (each a set of named scalars
// selected by switch), reached only with compile-time row indices, so nothing goes to scratch.
struct V4 {
  uint64_t v[kSub];
};
// One row's eight registers as an 8-lane vector: a register index that is uniform across the
// wave selects an element with the hardware's register-indexed moves, not through scratch.
static_assert(kSub == 4, "RegsN holds one register set per sub-tile row");
// Registers are read and written with selects on the (wave-uniform) register index: plain
// VALU work, where a switch would let the compiler turn the register file into an indexed
// scratch array.
struct RegRow {
  uint64_t r0, r1, r2, r3, r4, r5, r6, r7;
  // a mux tree on the index bits
  __device__ __forceinline__ uint64_t get(int i) const {
    const bool b0 = i & 1, b1 = i & 2, b2 = i & 4;
    const uint64_t x01 = b0 ? r1 : r0, x23 = b0 ? r3 : r2;
    const uint64_t x45 = b0 ? r5 : r4, x67 = b0 ? r7 : r6;
    const uint64_t x03 = b1 ? x23 : x01, x47 = b1 ? x67 : x45;
    return b2 ? x47 : x03;
  }
  __device__ __forceinline__ void put(int i, uint64_t v) {
    const uint32_t hot = 1u << (i & 7);
    r0 = (hot & 1u) ? v : r0;
    r1 = (hot & 2u) ? v : r1;
    r2 = (hot & 4u) ? v : r2;
    r3 = (hot & 8u) ? v : r3;
    r4 = (hot & 16u) ? v : r4;
    r5 = (hot & 32u) ? v : r5;
    r6 = (hot & 64u) ? v : r6;
    r7 = (hot & 128u) ? v : r7;
  }
};

struct RegsN {
  RegRow q0, q1, q2, q3;
  __device__ __forceinline__ V4 get(int i) const {
    const int k = i & 7;
    V4 x;
    x.v[0] = q0.get(k);
    x.v[1] = q1.get(k);
    x.v[2] = q2.get(k);
    x.v[3] = q3.get(k);
    return x;
  }
  __device__ __forceinline__ void put(int i, const V4& x) {
    const int k = i & 7;
    q0.put(k, x.v[0]);
    q1.put(k, x.v[1]);
    q2.put(k, x.v[2]);
    q3.put(k, x.v[3]);
  }
};

// Instructions [begin, end) for all kSub rows of the sub-tile at once: each instruction is
// fetched and dispatched once and then runs as kSub independent operations.  row[j] is the
// global row of sub-tile row j; rows with valid[j] == false compute garbage that callers drop.
__device__ __forceinline__ void run_rows(const tq_program_t& P, const tq_colset_t& C, int begin,
                                         int end, const int64_t (&row)[kSub],
                                         const bool (&valid)[kSub],
                                         const uint64_t* __restrict__ stage, RegsN& R) {
  const int t = threadIdx.x;
  for (int pc = begin; pc < end; ++pc) {
    const tq_instr_t I = P.code[pc];
    const V4 A = R.get(I.a), B = R.get(I.b);
    const uint64_t(&a)[kSub] = A.v;
    const uint64_t(&b)[kSub] = B.v;
    V4 O;
    uint64_t(&o)[kSub] = O.v;
#define TQ_EACH(EXPR) _Pragma("unroll") for (int j = 0; j < kSub; ++j) o[j] = (uint64_t)(EXPR);
    switch (I.op) {
      case TQ_OP_LOAD: TQ_EACH(stage[(I.a * kSub + j) * kBlock + t]) break;
      case TQ_OP_CONST: TQ_EACH(I.imm) break;
      case TQ_OP_STR_LIKE2:
      case TQ_OP_STR_CONTAINS: {
        const int8_t* p = (const int8_t*)col_ptr(C, I.a);
        const int stride = col_width(C, I.a);
        const int w = I.c ? (int)I.c : stride;
        const bool like = I.op == TQ_OP_STR_LIKE2;
#pragma unroll
        for (int j = 0; j < kSub; ++j) {
          const int8_t* sp = p + row[j] * (int64_t)stride;
          o[j] = valid[j] && (like ? str_like2(sp, w, P.strpool, I.imm)
                                   : str_contains(sp, w, P.strpool, I.imm));
        }
        break;
      }
      case TQ_OP_STR_EQ:
      case TQ_OP_STR_PREFIX: {
        const int8_t* p = (const int8_t*)col_ptr(C, I.a);
        const int stride = col_width(C, I.a);
        const int w = I.c ? (int)I.c : stride;
        const int off = (int)(I.imm & 0xffffffff);
        const int len = (int)(I.imm >> 32);
        const int upto = (I.op == TQ_OP_STR_EQ) ? w : len;
#pragma unroll
        for (int j = 0; j < kSub; ++j) {
          bool eq = valid[j];
          if (eq) {
            const int8_t* sp = p + row[j] * (int64_t)stride;
            for (int c = 0; c < upto; ++c) {
              const int8_t want = c < len ? (int8_t)P.strpool[off + c] : (int8_t)0;
              eq = eq && (sp[c] == want);
            }
          }
          o[j] = eq;
        }
        break;
      }
      case TQ_OP_CVT_I64_F64: TQ_EACH(f64_bits((double)(int64_t)a[j])) break;
      case TQ_OP_CVT_I64_F32: TQ_EACH(f32_bits((float)(int64_t)a[j])) break;
      case TQ_OP_CVT_F32_F64: TQ_EACH(f64_bits((double)as_f32(a[j]))) break;
      case TQ_OP_CVT_F64_F32: TQ_EACH(f32_bits((float)as_f64(a[j]))) break;
      case TQ_OP_WRAP_I8: TQ_EACH((int64_t)(int8_t)a[j]) break;
      case TQ_OP_WRAP_I16: TQ_EACH((int64_t)(int16_t)a[j]) break;
      case TQ_OP_WRAP_I32: TQ_EACH((int64_t)(int32_t)a[j]) break;
      case TQ_OP_ADD_I64: TQ_EACH(a[j] + b[j]) break;
      case TQ_OP_SUB_I64: TQ_EACH(a[j] - b[j]) break;
      case TQ_OP_MUL_I64: TQ_EACH(a[j] * b[j]) break;
      case TQ_OP_ADD_F64: TQ_EACH(f64_bits(as_f64(a[j]) + as_f64(b[j]))) break;
      case TQ_OP_SUB_F64: TQ_EACH(f64_bits(as_f64(a[j]) - as_f64(b[j]))) break;
      case TQ_OP_MUL_F64: TQ_EACH(f64_bits(as_f64(a[j]) * as_f64(b[j]))) break;
      case TQ_OP_DIV_F64: TQ_EACH(f64_bits(as_f64(a[j]) / as_f64(b[j]))) break;
      case TQ_OP_ADD_F32: TQ_EACH(f32_bits(as_f32(a[j]) + as_f32(b[j]))) break;
      case TQ_OP_SUB_F32: TQ_EACH(f32_bits(as_f32(a[j]) - as_f32(b[j]))) break;
      case TQ_OP_MUL_F32: TQ_EACH(f32_bits(as_f32(a[j]) * as_f32(b[j]))) break;
      case TQ_OP_DIV_F32: TQ_EACH(f32_bits(as_f32(a[j]) / as_f32(b[j]))) break;
      case TQ_OP_LT_I64: TQ_EACH((int64_t)a[j] < (int64_t)b[j]) break;
      case TQ_OP_LE_I64: TQ_EACH((int64_t)a[j] <= (int64_t)b[j]) break;
      case TQ_OP_GT_I64: TQ_EACH((int64_t)a[j] > (int64_t)b[j]) break;
      case TQ_OP_GE_I64: TQ_EACH((int64_t)a[j] >= (int64_t)b[j]) break;
      case TQ_OP_EQ_I64: TQ_EACH(a[j] == b[j]) break;
      case TQ_OP_NE_I64: TQ_EACH(a[j] != b[j]) break;
      case TQ_OP_LT_F64: TQ_EACH(as_f64(a[j]) < as_f64(b[j])) break;
      case TQ_OP_LE_F64: TQ_EACH(as_f64(a[j]) <= as_f64(b[j])) break;
      case TQ_OP_GT_F64: TQ_EACH(as_f64(a[j]) > as_f64(b[j])) break;
      case TQ_OP_GE_F64: TQ_EACH(as_f64(a[j]) >= as_f64(b[j])) break;
      case TQ_OP_EQ_F64: TQ_EACH(as_f64(a[j]) == as_f64(b[j])) break;
      case TQ_OP_NE_F64: TQ_EACH(as_f64(a[j]) != as_f64(b[j])) break;
      case TQ_OP_LT_F32: TQ_EACH(as_f32(a[j]) < as_f32(b[j])) break;
      case TQ_OP_LE_F32: TQ_EACH(as_f32(a[j]) <= as_f32(b[j])) break;
      case TQ_OP_GT_F32: TQ_EACH(as_f32(a[j]) > as_f32(b[j])) break;
      case TQ_OP_GE_F32: TQ_EACH(as_f32(a[j]) >= as_f32(b[j])) break;
      case TQ_OP_EQ_F32: TQ_EACH(as_f32(a[j]) == as_f32(b[j])) break;
      case TQ_OP_NE_F32: TQ_EACH(as_f32(a[j]) != as_f32(b[j])) break;
      case TQ_OP_ISCLOSE_F64: {
        // torch.isclose: a == b, or |a-b| finite and <= atol + |rtol * b| with rtol = 1e-5
        const double atol = __longlong_as_double(I.imm);
#pragma unroll
        for (int j = 0; j < kSub; ++j) {
          const double x = as_f64(a[j]), y = as_f64(b[j]);
          const double err = fabs(x - y);
          o[j] = (x == y) || (isfinite(err) && err <= atol + fabs(1e-5 * y));
        }
        break;
      }
      case TQ_OP_AND: TQ_EACH((a[j] != 0) && (b[j] != 0)) break;
      case TQ_OP_OR: TQ_EACH((a[j] != 0) || (b[j] != 0)) break;
      case TQ_OP_NOT: TQ_EACH(a[j] == 0) break;
      case TQ_OP_WHERE: {
        const V4 Cv = R.get(I.c);
        TQ_EACH(Cv.v[j] ? a[j] : b[j])
        break;
      }
      case TQ_OP_YEAR: TQ_EACH(year_of_day((int64_t)a[j])) break;
      default: TQ_EACH(0) break;
    }
#undef TQ_EACH
    R.put(I.dst, O);
  }
}

}  // namespace tq
