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

// ------------------------------------------------------------------ full-tile heads
// A full tile with 4-byte keys, a conjunction whose terms read at most one column of at most 4
// bytes, and aligned columns loads through issue_runs: the keys, the filter values and the mask
// bytes of the tile (of every tile a kernel takes at once) are all issued before the first is
// used.  FW is the filter column's width (0: no filter column), MASK whether a mask column is
// read; the terms applied are C's terms on load slot cs, to the values of column fcol (dtype
// fdt).  C stays where the caller keeps it (kernel arguments, a descriptor in global memory), so
// its terms are read as uniform scalars.
constexpr uint32_t kNoKey = 0xFFFFFFFFu;  // a row with no key offset to look up

template <int FW, bool MASK>
__device__ __forceinline__ uint32_t conj_sel_runs(const RunRaw<(FW > 0 ? FW : 1)>& fr,
                                                  const RunRaw<1>& mr, const Conj& C, int cs,
                                                  int fdt) {
  uint32_t s = 0xFFFFu;
  if constexpr (MASK) {
#pragma unroll
    for (int j = 0; j < kItems; ++j) s &= ~((run_value32<1>(mr, j, TQ_BOOL) == 0u ? 1u : 0u) << j);
  }
  if constexpr (FW > 0) {
    uint32_t pv[kItems];
#pragma unroll
    for (int j = 0; j < kItems; ++j) pv[j] = run_value32<FW>(fr, j, fdt);
    s = conj_apply32(C, cs, pv, s);
  }
  return s;
}

// One full tile: the selection bits, and the keys sign-extended to 64 bits in raw.
template <int FW, bool MASK>
__device__ __forceinline__ uint32_t tile_head(const void* keys, const uint8_t* mask,
                                              const void* fcol, const Conj& C, int cs, int fdt,
                                              int64_t tile0, uint64_t (&raw)[kItems]) {
  RunRaw<4> kr;
  RunRaw<(FW > 0 ? FW : 1)> fr;
  RunRaw<1> mr;
  issue_runs<4>(keys, tile0, kr);
  if constexpr (FW > 0) issue_runs<FW>(fcol, tile0, fr);
  if constexpr (MASK) issue_runs<1>(mask, tile0, mr);
  const uint32_t s = conj_sel_runs<FW, MASK>(fr, mr, C, cs, fdt);
#pragma unroll
  for (int j = 0; j < kItems; ++j)
    raw[j] = (uint64_t)(int64_t)(int32_t)run_value32<4>(kr, j, TQ_I32);
  return s;
}

// T consecutive full tiles from row base: every load issued tile by tile, then per tile the
// selection, the key offsets (key - lo, kNoKey when unselected or outside [lo, lo + range);
// range < 2^32 - 1) and the presence-byte lookups (a tile's lookups go out while the next
// tile's loads are still in flight).
template <int T, int FW, bool MASK>
__device__ __forceinline__ void probe_head(const void* keys, const uint8_t* mask,
                                           const void* fcol, const Conj& C, int cs, int fdt,
                                           int64_t base, int64_t lo, uint64_t range,
                                           const uint8_t* __restrict__ pres,
                                           uint32_t (&off)[T][kItems], uint8_t (&b)[T][kItems],
                                           uint32_t (&sel)[T]) {
  RunRaw<4> kr[T];
  RunRaw<(FW > 0 ? FW : 1)> fr[T];
  RunRaw<1> mr[T];
#pragma unroll
  for (int q = 0; q < T; ++q) {
    const int64_t tile0 = base + (int64_t)q * kTileRows;
    issue_runs<4>(keys, tile0, kr[q]);
    if constexpr (FW > 0) issue_runs<FW>(fcol, tile0, fr[q]);
    if constexpr (MASK) issue_runs<1>(mask, tile0, mr[q]);
  }
#pragma unroll
  for (int q = 0; q < T; ++q) {
    sel[q] = conj_sel_runs<FW, MASK>(fr[q], mr[q], C, cs, fdt);
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint64_t o =
          (uint64_t)(int64_t)(int32_t)run_value32<4>(kr[q], j, TQ_I32) - (uint64_t)lo;
      off[q][j] = ((sel[q] >> j) & 1u) && o < range ? (uint32_t)o : kNoKey;
      b[q][j] = off[q][j] != kNoKey ? pres[off[q][j]] : (uint8_t)0;
    }
  }
}

// Which head a full tile takes: -1 for the general loaders, else FW * 2 + MASK.  nfilt: 0 for
// no filter column, 1 for the column fcol of dtype fdt.
__device__ __forceinline__ int head_case(int kdt, const void* keys, const uint8_t* mask,
                                         int nfilt, const void* fcol, int fdt) {
  if (kdt != TQ_I32 || !runs_ok(keys, 4) || nfilt > 1) return -1;
  if (mask && !runs_ok(mask, 1)) return -1;
  int fw = 0;
  if (nfilt == 1) {
    fw = value_bytes(fdt);
    if (fw > 4 || !runs_ok(fcol, fw)) return -1;
  }
  return fw * 2 + (mask ? 1 : 0);
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
