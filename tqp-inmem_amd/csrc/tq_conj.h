// Conjunctions of column-versus-constant comparisons: their device evaluation over a tile and
// their recognition in a predicate program (the filter and the fused groupjoin use both).
#pragma once
#include "tq_common.h"
#include "tq_expr.h"

namespace tq {

// Most predicates are conjunctions of column-versus-constant comparisons (isnotnull compiles to
// a constant true).  match_conj recognises such a program by following its registers; a kernel
// then tests a tile per load slot: a thread reads its 16 rows (four runs of four consecutive
// rows, all loads in flight together) and tests them against every term on that slot.
constexpr int kMaxTerms = 8;

struct Conj {
  int32_t nterm;
  int32_t nslot;
  int32_t slot_dtype[TQ_MAX_LOADS];
  const void* slot_col[TQ_MAX_LOADS];
  int32_t term_slot[kMaxTerms];
  int32_t term_op[kMaxTerms];
  int64_t term_imm[kMaxTerms];
};

template <typename T>
__device__ __forceinline__ uint32_t cmp_bits(int op, const T (&x)[kItems], T c) {
  uint32_t b = 0;
  switch (op) {
#define TQ_CMP(CASE, EXPR)                                                  \
  case CASE:                                                                \
    _Pragma("unroll") for (int j = 0; j < kItems; ++j) b |= (uint32_t)(EXPR) << j; \
    break;
    TQ_CMP(0, x[j] < c)
    TQ_CMP(1, x[j] <= c)
    TQ_CMP(2, x[j] > c)
    TQ_CMP(3, x[j] >= c)
    TQ_CMP(4, x[j] == c)
    default: TQ_CMP(5, x[j] != c)
#undef TQ_CMP
  }
  return b;
}

// The terms on load slot s applied to that slot's 16 values of a tile, ANDed into acc.
__device__ __forceinline__ uint32_t conj_apply(const Conj& C, int s, const uint64_t (&v)[kItems],
                                               uint32_t acc) {
  for (int k = 0; k < C.nterm; ++k) {
    if (C.term_slot[k] != s) continue;
    const int op = C.term_op[k];
    const int64_t imm = C.term_imm[k];
    if (op < 6) {
      int64_t x[kItems];
#pragma unroll
      for (int j = 0; j < kItems; ++j) x[j] = (int64_t)v[j];
      acc &= cmp_bits<int64_t>(op, x, imm);
    } else if (op < 12) {
      double x[kItems];
#pragma unroll
      for (int j = 0; j < kItems; ++j) x[j] = __longlong_as_double((long long)v[j]);
      acc &= cmp_bits<double>(op - 6, x, __longlong_as_double(imm));
    } else {
      float x[kItems];
#pragma unroll
      for (int j = 0; j < kItems; ++j) x[j] = __uint_as_float((uint32_t)v[j]);
      acc &= cmp_bits<float>(op - 12, x, __uint_as_float((uint32_t)imm));
    }
  }
  return acc;
}

// conj_apply over values loaded in 32 bits (load_raw16_32: slot dtypes of at most 4 bytes, so
// no f64 term reads them).
__device__ __forceinline__ uint32_t conj_apply32(const Conj& C, int s, const uint32_t (&v)[kItems],
                                                 uint32_t acc) {
  for (int k = 0; k < C.nterm; ++k) {
    if (C.term_slot[k] != s) continue;
    const int op = C.term_op[k];
    const int64_t imm = C.term_imm[k];
    if (op < 6) {
      int64_t x[kItems];
#pragma unroll
      for (int j = 0; j < kItems; ++j) x[j] = (int64_t)(int32_t)v[j];
      acc &= cmp_bits<int64_t>(op, x, imm);
    } else {
      float x[kItems];
#pragma unroll
      for (int j = 0; j < kItems; ++j) x[j] = __uint_as_float(v[j]);
      acc &= cmp_bits<float>(op - 12, x, __uint_as_float((uint32_t)imm));
    }
  }
  return acc;
}

// One term against one staged value (the stage's conventions: integers sign-extended to 64 bits,
// f64 as its bits, f32 as its bits in the low half).
__device__ __forceinline__ bool term_holds(int op, int64_t imm, uint64_t v) {
  if (op < 6) {
    const int64_t x = (int64_t)v;
    switch (op) {
      case 0: return x < imm;
      case 1: return x <= imm;
      case 2: return x > imm;
      case 3: return x >= imm;
      case 4: return x == imm;
      default: return x != imm;
    }
  }
  if (op < 12) {
    const double x = __longlong_as_double((long long)v), c = __longlong_as_double(imm);
    switch (op - 6) {
      case 0: return x < c;
      case 1: return x <= c;
      case 2: return x > c;
      case 3: return x >= c;
      case 4: return x == c;
      default: return x != c;
    }
  }
  const float x = __uint_as_float((uint32_t)v), c = __uint_as_float((uint32_t)imm);
  switch (op - 12) {
    case 0: return x < c;
    case 1: return x <= c;
    case 2: return x > c;
    case 3: return x >= c;
    case 4: return x == c;
    default: return x != c;
  }
}

// The rows of a 4096-row tile (thread t: the 16 rows run_row(tile0, j)) that pass every term,
// as bits of `acc` (rows past n must already be clear in it).
__device__ __forceinline__ uint32_t conj_sel16(const Conj& C, int64_t tile0, int64_t n,
                                               uint32_t acc) {
  for (int s = 0; s < C.nslot; ++s) {
    uint64_t v[kItems];
    load_raw16(C.slot_col[s], C.slot_dtype[s], tile0, n, v);
    acc = conj_apply(C, s, v, acc);
  }
  return acc;
}

// Recognise a conjunction of column-versus-constant comparisons in segment 0 of a program.
// Compare opcodes map to 0..5 (lt, le, gt, ge, eq, ne) plus 6 for f64 and 12 for f32.
inline int cmp_code(int op) {
  if (op >= TQ_OP_LT_I64 && op <= TQ_OP_NE_I64) return op - TQ_OP_LT_I64;
  if (op >= TQ_OP_LT_F64 && op <= TQ_OP_NE_F64) return 6 + op - TQ_OP_LT_F64;
  if (op >= TQ_OP_LT_F32 && op <= TQ_OP_NE_F32) return 12 + op - TQ_OP_LT_F32;
  return -1;
}

// With remap == false the terms keep the program's own load slots (term_slot indexes P's load
// table) and nslot / slot_* are left empty: for kernels that evaluate the terms over the
// program's staged loads.
inline bool match_conj(const tq_program_t& P, const tq_colset_t& cols, Conj* C,
                       bool remap = true) {
  enum { UNK, CONSTV, LOADV, TERMS };
  struct RS { int kind; int64_t imm; int slot; uint32_t terms; };
  RS R[TQ_NREG];
  for (int i = 0; i < TQ_NREG; ++i) R[i] = RS{UNK, 0, 0, 0};
  Conj all{};
  for (int pc = 0; pc < P.seg_end[0]; ++pc) {
    const tq_instr_t& I = P.code[pc];
    const int d = I.dst & (TQ_NREG - 1);
    const RS a = R[I.a & (TQ_NREG - 1)], b = R[I.b & (TQ_NREG - 1)];
    const int cc = cmp_code(I.op);
    if (I.op == TQ_OP_NOP) continue;
    if (I.op == TQ_OP_CONST) {
      R[d] = RS{CONSTV, I.imm, 0, 0};
    } else if (I.op == TQ_OP_LOAD) {
      if (I.a >= P.nload) return false;
      R[d] = RS{LOADV, 0, I.a, 0};
    } else if (cc >= 0) {
      int code = cc, slot;
      int64_t imm;
      if (a.kind == LOADV && b.kind == CONSTV) {
        slot = a.slot;
        imm = b.imm;
      } else if (a.kind == CONSTV && b.kind == LOADV) {
        static const int flip[6] = {2, 3, 0, 1, 4, 5};  // c < x  <=>  x > c, ...
        code = (cc / 6) * 6 + flip[cc % 6];
        slot = b.slot;
        imm = a.imm;
      } else {
        return false;
      }
      const int dt = P.load_dtype[slot];
      const bool ok = cc < 6 ? (dt <= TQ_I64) : (cc < 12 ? dt == TQ_F64 : dt == TQ_F32);
      if (!ok || all.nterm == kMaxTerms || cols.width[P.load_col[slot]] != 1) return false;
      all.term_slot[all.nterm] = slot;
      all.term_op[all.nterm] = code;
      all.term_imm[all.nterm] = imm;
      R[d] = RS{TERMS, 0, 0, 1u << all.nterm};
      ++all.nterm;
    } else if (I.op == TQ_OP_AND) {
      // a constant operand must be true (nonzero); the result holds both operands' terms
      if ((a.kind != TERMS && !(a.kind == CONSTV && a.imm != 0)) ||
          (b.kind != TERMS && !(b.kind == CONSTV && b.imm != 0)))
        return false;
      const uint32_t tm = (a.kind == TERMS ? a.terms : 0u) | (b.kind == TERMS ? b.terms : 0u);
      R[d] = tm ? RS{TERMS, 0, 0, tm} : RS{CONSTV, 1, 0, 0};
    } else {
      return false;
    }
  }
  const RS res = R[P.seg_result[0] & (TQ_NREG - 1)];
  if (!(res.kind == TERMS || (res.kind == CONSTV && res.imm != 0))) return false;
  // keep the terms the result is the conjunction of
  *C = Conj{};
  for (int k = 0; k < all.nterm; ++k) {
    if (res.kind != TERMS || !((res.terms >> k) & 1u)) continue;
    C->term_slot[C->nterm] = all.term_slot[k];
    C->term_op[C->nterm] = all.term_op[k];
    C->term_imm[C->nterm] = all.term_imm[k];
    ++C->nterm;
  }
  if (!remap) return true;
  // load only the slots some kept term reads, renumbered densely
  int dense[TQ_MAX_LOADS];
  for (int sl = 0; sl < TQ_MAX_LOADS; ++sl) dense[sl] = -1;
  for (int k = 0; k < C->nterm; ++k) {
    const int sl = C->term_slot[k];
    if (dense[sl] < 0) {
      dense[sl] = C->nslot;
      C->slot_dtype[C->nslot] = P.load_dtype[sl];
      C->slot_col[C->nslot] = cols.ptr[P.load_col[sl]];
      ++C->nslot;
    }
    C->term_slot[k] = dense[sl];
  }
  return true;
}


}  // namespace tq
