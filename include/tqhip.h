/*
 * tqhip.h — C ABI of the MI355X (gfx950) kernels behind the relational hot path:
 * predicate filter + stream compaction, LSD radix sort / top-k, sort-based equi-join and
 * group-by aggregation.
 *
 * The reference engine has no native layer: each entry point below replaces a chain of ATen
 * calls inside its Python operators.  The replaced call sites are cited on each declaration
 * (paths are relative to the reference tree).
 *
 * Conventions
 *  - Every entry point returns TQ_OK (0) or a negative TQ_E* code; tq_last_error() returns the
 *    message of the calling thread's last failure.
 *  - All pointers are device pointers unless the parameter says "host".  Caller allocates every
 *    output and workspace buffer; the library never allocates or frees device memory.
 *  - `stream` is a hipStream_t passed as void*; NULL means the legacy default stream.  Calls are
 *    asynchronous on that stream; results are valid once the stream has been synchronised.
 *  - Row counts are int64_t; a single launch handles up to 2^31-1 rows.
 */
#ifndef TQHIP_H
#define TQHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status codes */
#define TQ_OK 0
#define TQ_EINVAL (-1)       /* bad argument: Python raises ValueError                  */
#define TQ_EUNSUPPORTED (-2) /* unsupported operation/type: Python raises NotImplemented */
#define TQ_EHIP (-3)         /* HIP runtime / launch failure: RuntimeError              */
#define TQ_ECAPACITY (-4)    /* workspace or table too small: MemoryError               */

/* ---------------------------------------------------------------- element types */
#define TQ_BOOL 0
#define TQ_I8 1
#define TQ_I16 2
#define TQ_I32 3
#define TQ_I64 4
#define TQ_F32 5
#define TQ_F64 6

/* ---------------------------------------------------------------- expression programs
 * A predicate or aggregate argument is compiled on the host into a register program that one
 * kernel evaluates per row.  Registers hold 64 raw bits; each opcode fixes how they are read:
 * integers as int64, booleans as 0/1 int64, f64 as a double, f32 as a float in the low half.
 * The host compiler resolves torch's type promotion, so the program computes in exactly the
 * dtype the reference's ATen ops would (e.g. int32 column vs float literal compares in f32).
 */
#define TQ_MAX_INSTR 64
#define TQ_MAX_COLS 16
#define TQ_NREG 8
#define TQ_STRPOOL 128
#define TQ_MAX_SEGS 12
#define TQ_MAX_LOADS 8

enum tq_opcode {
  TQ_OP_NOP = 0,
  /* dst <- value of load slot a for this row (see tq_program_t.load_*) */
  TQ_OP_LOAD,
  TQ_OP_CONST, /* dst <- imm (raw bits) */
  /* conversions */
  TQ_OP_CVT_I64_F64, TQ_OP_CVT_I64_F32, TQ_OP_CVT_F32_F64, TQ_OP_CVT_F64_F32,
  /* narrow an int64 result back to torch's result dtype (two's-complement wrap) */
  TQ_OP_WRAP_I8, TQ_OP_WRAP_I16, TQ_OP_WRAP_I32,
  /* arithmetic */
  TQ_OP_ADD_I64, TQ_OP_SUB_I64, TQ_OP_MUL_I64,
  TQ_OP_ADD_F64, TQ_OP_SUB_F64, TQ_OP_MUL_F64, TQ_OP_DIV_F64,
  TQ_OP_ADD_F32, TQ_OP_SUB_F32, TQ_OP_MUL_F32, TQ_OP_DIV_F32,
  /* comparisons -> bool */
  TQ_OP_LT_I64, TQ_OP_LE_I64, TQ_OP_GT_I64, TQ_OP_GE_I64, TQ_OP_EQ_I64, TQ_OP_NE_I64,
  TQ_OP_LT_F64, TQ_OP_LE_F64, TQ_OP_GT_F64, TQ_OP_GE_F64, TQ_OP_EQ_F64, TQ_OP_NE_F64,
  TQ_OP_LT_F32, TQ_OP_LE_F32, TQ_OP_GT_F32, TQ_OP_GE_F32, TQ_OP_EQ_F32, TQ_OP_NE_F32,
  /* torch.isclose(a, b, rtol=1e-5, atol=imm as f64) */
  TQ_OP_ISCLOSE_F64,
  /* boolean */
  TQ_OP_AND, TQ_OP_OR, TQ_OP_NOT,
  /* dst <- reg[c] ? reg[a] : reg[b] (raw bits) */
  TQ_OP_WHERE,
  /* dst <- all(string column a [row, 0:len] == pool[off:off+len]) (imm = off | len << 32);
     EQ also requires the column's remaining bytes to be 0 (torch pads literals with 0).
     For every string opcode, c != 0 narrows the row to its first c bytes (a substring column:
     the column set entry points at the substring's first byte and keeps the full row width as
     its row stride) */
  TQ_OP_STR_EQ, TQ_OP_STR_PREFIX,
  /* dst <- calendar year of int64 day count since 1990-01-01 (bucketize on year edges) */
  TQ_OP_YEAR,
  /* dst <- string column a matches '%w1%w2%': w1 occurs at some position q and w2 at some
     position >= q + len(w1), both inside the column's width (padding bytes included), as the
     reference's like_mask scans (operators/like.py:15-25, filter.py:121-127);
     imm = off1 | len1 << 16 | off2 << 32 | len2 << 48 in the string pool */
  TQ_OP_STR_LIKE2,
  /* dst <- string column a holds pool[off:off+len] at some position q with q + len <= width,
     padding bytes included, as like_contains scans (operators/like.py:3-11; filter.py:56-62,
     which also takes EndsWith this way); imm = off | len << 32 */
  TQ_OP_STR_CONTAINS,
  TQ_OP_COUNT_
};

typedef struct {
  uint8_t op, dst, a, b, c, pad0, pad1, pad2;
  int64_t imm;
} tq_instr_t;

/* A program is a bank of `nseg` segments run in order; segment s ends before instruction
 * seg_end[s] and leaves its value in register seg_result[s].  Registers persist across
 * segments, so later segments may reuse values of earlier ones.  A predicate has one segment;
 * a group-by passes one segment per aggregate argument.
 * Column reads are declared up front in the load table: slot i reads column load_col[i] of
 * type load_dtype[i] (TQ_BOOL..TQ_F64) for every row of a tile before any instruction runs, so
 * all of a tile's loads are in flight together; TQ_OP_LOAD then reads the slot. */
typedef struct {
  int32_t n_instr;
  int32_t nseg;
  int16_t seg_end[TQ_MAX_SEGS];
  uint8_t seg_result[TQ_MAX_SEGS];
  int32_t nload;
  uint8_t load_col[TQ_MAX_LOADS];
  uint8_t load_dtype[TQ_MAX_LOADS];
  tq_instr_t code[TQ_MAX_INSTR];
  uint8_t strpool[TQ_STRPOOL];
} tq_program_t;

typedef struct {
  const void* ptr[TQ_MAX_COLS]; /* device column base pointers */
  int32_t width[TQ_MAX_COLS];   /* bytes per row for 2-D int8 string columns, else 1 */
} tq_colset_t;

/* ---------------------------------------------------------------- housekeeping */
int tq_version(void);
const char* tq_last_error(void);
/* Rows covered by one filter/compaction tile (the unit of tile_counts / tile_offsets). */
int tq_tile_rows(void);

/* Kernel timing.  The reference times operators with torch.cuda.synchronize() and wall clock
 * around each operator (main.py:273-277); these calls give the per-kernel durations behind such
 * a time.  While on, every kernel launch of this library carries a start and a stop event that
 * are stamped when that kernel starts and ends on the device (hipExtLaunchKernel), so host gaps
 * between launches are excluded.  tq_timing_count returns the number of launches logged so far;
 * after the caller has synchronised the streams involved, tq_timing_read copies up to cap
 * records in launch order (name is a static string naming the kernel) and clears the log; at
 * most 65536 launches are logged between reads, later ones run untimed and are counted in
 * *dropped.  Not for launches under graph capture. */
typedef struct {
  const char* name;
  float ms;
  int32_t pad;
} tq_ktime_t;
int tq_timing(int32_t on);
int64_t tq_timing_count(void);
int tq_timing_read(tq_ktime_t* out, int64_t cap, int64_t* n, int64_t* dropped);

/* ---------------------------------------------------------------- filter + compaction */

/* Evaluate a predicate program over n rows into a bool mask (1 byte/row) and, if tile_counts is
 * non-NULL, the number of true rows of each tile of tq_tile_rows() rows.
 * Replaces operators/filter.py:14-162 (`evaluate`: one ATen kernel per tree node,
 * torch.le/lt/ge/gt/eq/isclose, logical_and/or, ones for isnotnull). */
int tq_filter_mask(const tq_program_t* prog /* host */, const tq_colset_t* cols /* host */,
                   int64_t n, uint8_t* mask_out, uint32_t* tile_counts, void* stream);

/* 1 when segment 0 of a predicate program is a conjunction of column-versus-constant
 * comparisons (at most 8, over integer, f32 or f64 columns, isnotnull counting as true), the
 * form tq_filter_mask runs without its interpreter and tq_groupjoin_run can fuse; else 0.
 * Host-side only (launches nothing).  A query of this library, no reference counterpart. */
int tq_filter_is_conjunction(const tq_program_t* prog /* host */,
                             const tq_colset_t* cols /* host */);

/* Evaluate segment 0 of a program per row into a column of out_bytes-wide values: the low
 * bytes of the result register (bool/int8: 1, int16: 2, int32/f32: 4, int64/f64: 8).
 * Replaces the elementwise ATen chains of operators/aggregate.py:197-233 when they combine
 * aggregate results (e.g. 0.2 * avg(x)) and of operators/project.py:26-40. */
int tq_eval_column(const tq_program_t* prog /* host */, const tq_colset_t* cols /* host */,
                   int64_t n, void* out, int32_t out_bytes, void* stream);

/* Every segment of a program per row, segment a into outs[a] (device) as out_bytes[a]-wide
 * values, the loads staged once for all segments.  Replaces the per-expression `evaluate`
 * calls of operators/project.py:26-40 (each an elementwise ATen chain, then a cast through
 * float64). */
int tq_eval_columns(const tq_program_t* prog /* host */, const tq_colset_t* cols /* host */,
                    int64_t n, void* const* outs /* host array */,
                    const int32_t* out_bytes /* host array */, void* stream);

/* Per-tile popcount of an existing bool mask (for masks not made by tq_filter_mask). */
int tq_mask_tile_counts(const uint8_t* mask, int64_t n, uint32_t* tile_counts, void* stream);

/* Exclusive scan of per-tile counts: tile_offsets[t] = sum of counts before t; total[0] = sum.
 * ws must hold tq_scan_ws_bytes(ntiles) bytes. */
size_t tq_scan_ws_bytes(int64_t ntiles);
int tq_scan_tiles(const uint32_t* tile_counts, int64_t ntiles, int64_t* tile_offsets,
                  int64_t* total, void* ws, void* stream);

/* Order-preserving stream compaction of up to TQ_MAX_COLS columns by one mask:
 * dst[c][tile_offsets[t] + rank] = src[c][row] for every selected row.
 * elem_bytes[c] is the row size of column c (1/2/4/8, or the width of a 2-D int8 column).
 * A NULL src[c] makes dst[c] an int64 column of the selected row numbers (torch.nonzero).
 * Replaces IO/vortex_pipeline.py:58 (`tensor[mask]`: nonzero + index_select per column). */
int tq_compact(const uint8_t* mask, int64_t n, const int64_t* tile_offsets, int32_t ncols,
               const void* const* src /* host array */, const int32_t* elem_bytes /* host */,
               void* const* dst /* host array */, void* stream);

/* Multi-column row gather: dst[c][i] = src[c][idx[i]] for i < m.  idx is int64.  When
 * null_fill is non-zero an index equal to -64 yields the value -64 in the column's type
 * instead; float_kind[c] (host, may be NULL without null_fill) is 4 for f32, 8 for f64 and 0
 * for integer/bool columns.
 * Replaces operators/hashjoin.py:161-175 and operators/sort.py:44-54 (per-column gathers) and
 * conversion.py:52-57 (index_with_null). */
int tq_gather(const int64_t* idx, int64_t m, int32_t ncols, const void* const* src /* host */,
              const int32_t* elem_bytes /* host */, const int32_t* float_kind /* host */,
              void* const* dst /* host */, int32_t null_fill, void* stream);

/* Row flags from an index list: flags[0:n] = 0, then flags[idx[i]] = 1 for i < m.
 * Replaces operators/hashjoin.py:118-141 (torch.unique of matched probe rows, ones + index_put
 * of the semi/anti/outer finishing). */
int tq_mark_rows(const int64_t* idx, int64_t m, uint8_t* flags, int64_t n, void* stream);

/* ---------------------------------------------------------------- sort / top-k */
typedef struct {
  const void* ptr;    /* device column */
  int32_t dtype;      /* TQ_I8..TQ_F64 or TQ_BOOL */
  int32_t width;      /* columns of a 2-D key (compared lexicographically), else 1 */
  int32_t descending; /* 0 ascending, 1 descending */
  int32_t pad;
} tq_sortkey_t;

/* Workspace for sorting n rows. */
size_t tq_sort_ws_bytes(int64_t n);

/* Stable multi-key sort permutation: perm_out[0:min(n,limit)] lists row ids in the order a
 * stable sort by key 0, then key 1, ... gives (ties keep input order).  Float keys follow
 * torch.sort: -0.0 ties +0.0 and NaN sorts above +inf.  limit < 0 means all rows; a limit
 * below n returns the first `limit` rows of that order (top-k).
 * Replaces operators/sort.py:20-42 (per-key torch.sort(stable=True) passes, unique(dim=0) for
 * 2-D keys, perm[:limit]) and the build-side torch.sort of operators/hashjoin.py:43. */
int tq_sort_perm(const tq_sortkey_t* keys /* host */, int32_t nkeys, int64_t n, int64_t limit,
                 int64_t* perm_out, void* ws, size_t ws_bytes, void* stream);

/* tq_sort_perm with a limit of 1..64 over rows whose count a producer earlier on the stream left
 * in device memory: the rows are the first min(n, *n_dev) of the key columns, n being the
 * capacity; perm_out holds min(limit, n) entries, of which the first min(limit, rows) are the
 * permutation and the rest row 0.  Applies (*applied = 1) when the keys pack into 64 bits;
 * otherwise nothing is launched.  Nothing is read back.  Same reference interface as
 * tq_sort_perm (operators/sort.py:6-58). */
int tq_sort_perm_deferred(const tq_sortkey_t* keys /* host */, int32_t nkeys, int64_t n,
                          const int64_t* n_dev, int64_t limit, int64_t* perm_out, void* ws,
                          size_t ws_bytes, int32_t* applied /* host out */, void* stream);

/* A take-ordered in one call: the top-k of tq_sort_perm_deferred (n_dev may be null: all n
 * rows) and, in its last round's kernel, the permuted rows of ncols columns (row_bytes[c] bytes
 * per row) into dst[c][0 .. min(limit, n)), entries past the candidates taking row 0.  Applies
 * (*applied = 1) for a limit of 1..64 over keys that pack into 64 bits; otherwise nothing is
 * launched.  Replaces operators/sort.py:20-54 (the limit's perm[:limit] and the per-column
 * gathers) for the small limits of a take-ordered. */
int tq_topk_gather(const tq_sortkey_t* keys /* host */, int32_t nkeys, int64_t n,
                   const int64_t* n_dev, int64_t limit, int32_t ncols,
                   const void* const* src /* host */, const int32_t* row_bytes /* host */,
                   void* const* dst /* host */, int64_t* perm_out, void* ws, size_t ws_bytes,
                   int32_t* applied /* host out */, void* stream);

/* Dense ranks of rows in a sort order: perm (a permutation of 0..n-1, e.g. tq_sort_perm's by
 * the same keys) lists the rows sorted; rank_out[row] receives the number of distinct keys
 * before the row's key in that order (equal keys share a rank: floats as the sort compares
 * them, -0.0 equal to +0.0 and NaN to NaN; 2-D int8 keys by their row bytes), rep_out[r] the
 * first row of rank r in the order, and *nranks (device) the number of distinct keys.
 * rep_out holds n entries.  ws: tq_rank_ws_bytes(n).  A group-by over keys that do not pack
 * into integer ranges (string rows, floats) groups by these ranks and maps them back through
 * rep_out.  Replaces the key stacking and torch.unique(hash) of operators/aggregate.py:302-327
 * for such keys. */
size_t tq_rank_ws_bytes(int64_t n);
int tq_rank_sorted(const tq_sortkey_t* keys /* host */, int32_t nkeys, int64_t n,
                   const int64_t* perm, int32_t* rank_out, int64_t* rep_out, int64_t* nranks,
                   void* ws, size_t ws_bytes, void* stream);

/* Copy n small device words (bytes[i] each, 512 bytes in all at most) to host_out, packed at
 * 8-byte-aligned offsets, with one synchronisation of the stream.  Replaces the `.item()` /
 * `.cpu()` reads the reference makes of device scalars (e.g. operators/hashjoin.py:76). */
int tq_read_words(int32_t n, const void* const* dev /* host */, const size_t* bytes /* host */,
                  int64_t* host_out /* host out */, void* stream);

/* The same read in two halves, so that the host can queue the next query's work before it
 * collects this one's counts.  Replaces the same reads (e.g. operators/hashjoin.py:76), which the
 * reference makes synchronously.  Each thread has, per device, tq_word_slots() pinned host
 * slots of 64 words.  tq_words_post gathers the n runs (whole aligned 8-byte words, 64 in all
 * at most) into `slot` with one launch on `stream` and records the slot's event after it;
 * tq_words_wait blocks on that event and copies the slot's words to host_out.  A slot is posted
 * again only after its wait (TQ_EINVAL otherwise). */
int32_t tq_word_slots(void);
int tq_words_post(int32_t n, const void* const* dev /* host */, const size_t* bytes /* host */,
                  int32_t slot, void* stream);
int tq_words_wait(int32_t slot, int64_t* host_out /* host out, 64 words */);

/* ---------------------------------------------------------------- equi-join */
/* Host-side handle of a built join index (filled by tq_join_build, read by the probe passes).
 * Keys are order-transformed (sign bit flipped; 32-bit for build keys of up to 32 bits, else
 * 64-bit).  The index maps a key to its build rows in one of two layouts:
 *   TQ_JOIN_DENSE: table[key - base] for range consecutive key values;
 *   TQ_JOIN_HASH:  2^hash_bits slots {u64 key, u32 a, u32 count} of linear probing from
 *                  (key * 0x9E3779B97F4A7C15) >> (64 - hash_bits); key == empty_key is free.
 * With unique == 1 every selected key occurs once and an entry holds the build row itself
 * (dense: u32 row, 0xFFFFFFFF = absent).  Otherwise rows[] lists the selected build rows
 * grouped by key, ascending within a key, and an entry holds {first position, count} (dense:
 * two u32, count 0 = absent). */
#define TQ_JOIN_EMPTY 0
#define TQ_JOIN_DENSE 1
#define TQ_JOIN_HASH 2
typedef struct {
  int64_t nsel;          /* selected build rows */
  int32_t key_bits;      /* 32 or 64: width of the transformed keys */
  int32_t layout;        /* TQ_JOIN_EMPTY / TQ_JOIN_DENSE / TQ_JOIN_HASH */
  int32_t unique;        /* 1: entries hold build rows; 0: entries index rows[] */
  int32_t hash_bits;     /* hash layout: log2 of the slot count */
  uint64_t base;         /* dense layout: transformed key of table[0] */
  uint64_t empty_key;    /* hash layout: key value of a free slot */
  int64_t range;         /* dense: table entries; hash: slots */
  const void* table;     /* device directory or slots */
  const uint32_t* bits;  /* dense layout: presence bit per table entry (bit e%32 of word e/32) */
  const uint32_t* rows;  /* device: build rows grouped by key (unique == 0) */
  /* Build inputs, kept while `pending`: the build assumed unique keys without waiting for its
   * duplicate check; the first tq_join_count reads the check together with its pair count and,
   * for repeated keys, rebuilds the index (sorted layout) before counting again. */
  int32_t pending;
  int32_t bdtype;
  const void* bkeys;
  const uint8_t* bmask;
  int64_t nb;
  void* ws;
  uint64_t key_or, key_and; /* OR / AND of the selected transformed keys */
} tq_join_state_t;

/* Workspace for building over nb build rows (room for either layout and for sorting). */
size_t tq_join_build_ws_bytes(int64_t nb);

/* Build the index over the selected build rows (bmask may be NULL).  Keys are integer columns
 * (TQ_BOOL..TQ_I64).  Synchronises the stream once (the row count and key range pick the
 * layout) and fills *state; bkeys, bmask and ws must stay valid until the probe passes are
 * done.  Replaces operators/hashjoin.py:41-52 (arange, torch.sort, unique_consecutive,
 * cumsum) and the build-side mask compaction of IO/vortex_pipeline.py:58. */
int tq_join_build(const void* bkeys, int32_t bdtype, const uint8_t* bmask, int64_t nb, void* ws,
                  size_t ws_bytes, tq_join_state_t* state /* host out */, void* stream);

/* tq_join_build for a caller that knows bounds of the build key column (key_lo <= every key
 * <= key_hi, e.g. column statistics recorded when the table was loaded): the bounds pick the
 * layout, so the build launches without synchronising.  state->nsel is -1 until the first
 * tq_join_count settles it.  A key outside the bounds is never trusted: that count detects it
 * and builds again as tq_join_build does.  Same reference interface as tq_join_build
 * (operators/hashjoin.py:41-52). */
int tq_join_build_ranged(const void* bkeys, int32_t bdtype, const uint8_t* bmask, int64_t nb,
                         int64_t key_lo, int64_t key_hi, void* ws, size_t ws_bytes,
                         tq_join_state_t* state /* host out */, void* stream);

/* Workspace for probing np rows. */
size_t tq_join_probe_ws_bytes(int64_t np);

/* Probe, count pass: number of (build, probe) pairs with equal keys among selected probe rows
 * (pmask may be NULL), per probe tile and in total.  The total is left in device memory inside
 * pws and, when total_host is non-NULL, also returned there (which synchronises the stream).
 * While the state is `pending`, the call always synchronises: it settles the build's
 * uniqueness (updating *state) before its count is final.
 * Replaces operators/hashjoin.py:57-98 (searchsorted, eq, nonzero, lenR.max(), masked_select,
 * repeat_interleave) up to the pair count. */
int tq_join_count(tq_join_state_t* state /* host, in/out */, const void* pkeys, int32_t pdtype,
                  const uint8_t* pmask, int64_t np, void* pws, int64_t* total_host,
                  void* stream);

/* Probe, write pass: emits every pair in canonical order (probe row ascending, then build row
 * ascending) and materialises output columns directly: left column c gets lsrc[c][build row],
 * right column c gets rsrc[c][probe row].  build_idx / probe_idx (int64, may be NULL) receive
 * the pair row ids.  Replaces operators/hashjoin.py:100-113 (key re-check, pair selection) and
 * operators/hashjoin.py:161-172 (output gathers). */
int tq_join_write(const tq_join_state_t* state /* host */, const void* pkeys, int32_t pdtype,
                  const uint8_t* pmask, int64_t np, const void* pws, int32_t nleft,
                  const void* const* lsrc /* host */, const int32_t* lbytes /* host */,
                  void* const* ldst /* host */, int32_t nright, const void* const* rsrc /* host */,
                  const int32_t* rbytes /* host */, void* const* rdst /* host */,
                  int64_t* build_idx, int64_t* probe_idx, void* stream);

/* Count, scan and write in one call with nothing read back, for a caller that goes on
 * launching work that consumes the pairs before it synchronises.  Applies (*applied = 1) when
 * the state's entries hold build rows (a unique build, settled or still pending) and np > 0;
 * otherwise *applied = 0 and nothing is launched.  The outputs must hold np rows (a probe row
 * has at most one pair); the pair count m is left in device memory at *total_dev and the first
 * m output rows are the pairs, as tq_join_write writes them.  While the state is pending,
 * *check_dev points at the build's 24 check bytes (read them back with the count and pass them
 * to tq_join_settle); the pairs are only valid once tq_join_settle accepts them.  Same
 * reference interface as tq_join_count + tq_join_write (operators/hashjoin.py:57-98). */
int tq_join_probe_deferred(const tq_join_state_t* state /* host */, const void* pkeys,
                           int32_t pdtype, const uint8_t* pmask, int64_t np, void* pws,
                           int32_t nleft, const void* const* lsrc, const int32_t* lbytes,
                           void* const* ldst, int32_t nright, const void* const* rsrc,
                           const int32_t* rbytes, void* const* rdst,
                           const int64_t** total_dev /* host out */,
                           const void** check_dev /* host out */, int32_t* applied /* host out */,
                           void* stream);

/* Settle a pending build from its check words as read back after tq_join_probe_deferred:
 * *ok = 1 when the build's keys were unique (and inside the bounds of a ranged build), which
 * makes the deferred pairs final and the state settled; *ok = 0 leaves the state pending, and
 * the caller discards the deferred pairs and probes with tq_join_count / tq_join_write, which
 * build again.  Host only.  No counterpart in the reference (its join sorts the build side
 * every time, operators/hashjoin.py:41-52). */
int tq_join_settle(tq_join_state_t* state /* host, in/out */, const int64_t* words /* host */,
                   int32_t* ok /* host out */);

/* ---------------------------------------------------------------- group-by aggregate */
#define TQ_AGG_SUM 0   /* f64 sum for f64 args, f32 for f32, exact int64 for integer/bool args */
#define TQ_AGG_COUNT 1 /* int64 count of rows whose argument != -64                            */
#define TQ_AGG_AVG 2   /* f64 sum / row count                                                  */
#define TQ_AGG_MIN 3   /* argument dtype; init = dtype max                                     */
#define TQ_AGG_MAX 4   /* argument dtype; init = 0, as the reference does                      */

#define TQ_GB_MAX_KEYS 8
#define TQ_GB_MAX_AGGS TQ_MAX_SEGS

/* Table modes (chosen by tq_groupby_plan).  Slots are addressed by the packed key. */
#define TQ_GB_DENSE_LDS 0    /* whole key space in one LDS array per workgroup, then merged  */
#define TQ_GB_DENSE_GLOBAL 1 /* key space addressed directly in one global array            */
#define TQ_GB_HASH 2         /* global open-addressing hash table on the packed key         */
#define TQ_GB_DENSE_PART 3   /* dense key space in partitions of 2^k slots: rows are scattered
                                by partition, one workgroup reduces each partition in LDS and
                                keeps its groups in key order (no global atomics)             */
#define TQ_GB_LDS_FRONT 8    /* flag: an LDS hash table absorbs repeated keys first         */
#define TQ_GB_LDS_LANES 16   /* flag (DENSE_LDS): one table copy per lane, so rows of a few hot
                              * groups never meet on an LDS address                         */

typedef struct {
  int32_t nkeys;
  int32_t naggs;
  int32_t key_dtype[TQ_GB_MAX_KEYS]; /* TQ_BOOL..TQ_I64                                    */
  const void* key_ptr[TQ_GB_MAX_KEYS];
  int64_t key_min[TQ_GB_MAX_KEYS];    /* from tq_groupby_bounds                            */
  int64_t key_range[TQ_GB_MAX_KEYS];  /* max - min + 1                                    */
  int64_t key_stride[TQ_GB_MAX_KEYS]; /* mixed-radix weights, filled by tq_groupby_plan    */
  int32_t agg_op[TQ_GB_MAX_AGGS];     /* TQ_AGG_*                                          */
  int32_t agg_dtype[TQ_GB_MAX_AGGS];  /* dtype of the argument segment's value            */
  int64_t nsel;                       /* selected rows; set by tq_groupby_plan             */
  int64_t nrows;                      /* rows the accumulate pass reads (set before plan)  */
  /* Key dictionaries (all NULL: none).  key_dict[k] non-NULL: key k's packed digit is the rank
   * of its value among the distinct selected values, which tq_groupby_dict found over the
   * value range [key_min, key_min + key_dict_range - 1]; key_range is then their count. */
  const void* key_dict[TQ_GB_MAX_KEYS];
  int64_t key_dict_range[TQ_GB_MAX_KEYS];
} tq_groupby_desc_t;

/* Pass 1: per-key-column min and max over selected rows (mask may be NULL) and the selected row
 * count.  out (device): int64[2*nkeys + 1] = {min0, max0, min1, max1, ..., nsel} (min > max
 * when no row is selected).  ws: tq_groupby_bounds_ws_bytes(n) bytes of per-tile partials.
 * Replaces the key stacking of operators/aggregate.py:297-312 (unsqueeze/torch.cat of the
 * key columns that precedes the hashed grouping). */
size_t tq_groupby_bounds_ws_bytes(int64_t n);
int tq_groupby_bounds(const tq_groupby_desc_t* desc /* host */, const uint8_t* mask, int64_t n,
                      int64_t* out, void* ws, void* stream);

/* Key dictionaries, for keys whose value range is much wider than their distinct values (e.g.
 * attribute codes spaced apart): over the selected rows, one pass marks every present value of
 * every key in a presence bitmap over [key_min, key_min + key_range - 1] (ranges as
 * tq_groupby_bounds gives them, each at most TQ_GB_DICT_MAX_RANGE), then each key's bitmap is
 * ranked and its distinct values listed in ascending order.  dict_out[k] receives key k's
 * dictionary (inside ws) and ndistinct_host[k] its distinct count (one synchronisation); the
 * caller that adopts them sets key_dict[k] = dict_out[k], key_dict_range[k] = key_range[k] and
 * key_range[k] = ndistinct_host[k] before tq_groupby_plan.  Groups and their order are
 * unchanged; only the packed key space shrinks.  No counterpart in the reference (its group ids
 * come from a hash sort, operators/aggregate.py:320-327). */
#define TQ_GB_DICT_MAX_RANGE (1 << 16)
size_t tq_groupby_dict_ws_bytes(const tq_groupby_desc_t* desc /* host */);
int tq_groupby_dict(const tq_groupby_desc_t* desc /* host */, const uint8_t* mask, int64_t n,
                    void* ws, size_t ws_bytes, void** dict_out /* host out, [nkeys] */,
                    int64_t* ndistinct_host /* host out, [nkeys] */, void* stream);

/* Host planning from the bounds: fills key_stride, picks the table mode and slot count and
 * returns the workspace size.  The packed key is sum((k_i - min_i) * stride_i), so packed order
 * is lexicographic key order and packed values are < product of ranges.
 * Fails with TQ_EUNSUPPORTED if the key space does not fit 62 bits. */
int tq_groupby_plan(tq_groupby_desc_t* desc /* host, in/out */, int64_t nsel, int32_t nload,
                    int32_t* mode, int64_t* slots, size_t* ws_bytes);

/* Pass 2: clear the table and accumulate every selected row.  Segment j of `prog` computes
 * aggregate j's argument per row; the key columns are read directly.
 * Replaces operators/aggregate.py:302-327 (key stacking, linear hash, torch.unique) and
 * operators/aggregate.py:127-193 (bincount / 2-step scatter_reduce sums, counts, averages). */
int tq_groupby_accumulate(const tq_groupby_desc_t* desc /* host */,
                          const tq_program_t* prog /* host */, const tq_colset_t* cols /* host */,
                          const uint8_t* mask, int64_t n, int32_t mode, int64_t slots, void* ws,
                          size_t ws_bytes, void* stream);

/* Pass 3: number the occupied slots.  ngroups (device int64) receives the group count; when
 * ngroups_host is non-NULL it is also returned there (synchronising the stream). */
int tq_groupby_count(const tq_groupby_desc_t* desc /* host */, int32_t mode, int64_t slots,
                     void* ws, int64_t* ngroups_host, void* stream);

/* Pass 4: write the ngroups groups in ascending key order: key_out[i] gets key i (in its dtype)
 * and agg_out[j] aggregate j (types as listed at TQ_AGG_*).  Hash tables are ordered by a
 * radix sort of the packed keys (synchronises the stream). */
int tq_groupby_write(const tq_groupby_desc_t* desc /* host */, int32_t mode, int64_t slots,
                     void* ws, int64_t ngroups, void* const* key_out /* host */,
                     void* const* agg_out /* host */, void* stream);

/* Passes 2-4 in one call with one synchronisation: the caller allocates the outputs for
 * out_cap groups (min(selected rows, slots), at least 1, always suffices) and receives the
 * group count in *ngroups_host; groups as tq_groupby_write writes them.  In the partitioned
 * mode each partition's workgroup adds up the group counts the partitions before it publish
 * and writes its groups at that offset itself, so no separate numbering or write pass runs.
 * Same reference interface as passes 2-4 (operators/aggregate.py:127-193, 302-327). */
int tq_groupby_run(const tq_groupby_desc_t* desc /* host */, const tq_program_t* prog /* host */,
                   const tq_colset_t* cols /* host */, const uint8_t* mask, int64_t n,
                   int32_t mode, int64_t slots, void* ws, size_t ws_bytes,
                   void* const* key_out /* host */, void* const* agg_out /* host */,
                   int64_t out_cap, int64_t* ngroups_host /* host out */, void* stream);

/* tq_groupby_run over rows whose count a producer earlier on the stream left in device memory:
 * the rows are the first min(n, *n_dev) of the columns, n being the capacity the call is
 * planned for (desc->nrows == n, partitioned mode only).  Nothing is read back: the group
 * count is left at *ngroups_dev and a look-back failure flag at *err_dev (non-zero: the
 * outputs are not valid; run tq_groupby_run instead).  Same reference interface as
 * tq_groupby_run (operators/aggregate.py:127-193, 302-327). */
int tq_groupby_run_deferred(const tq_groupby_desc_t* desc /* host */,
                            const tq_program_t* prog /* host */, const tq_colset_t* cols /* host */,
                            const uint8_t* mask, int64_t n, const int64_t* n_dev, int32_t mode,
                            int64_t slots, void* ws, size_t ws_bytes,
                            void* const* key_out /* host */, void* const* agg_out /* host */,
                            int64_t out_cap, const int64_t** ngroups_dev /* host out */,
                            const int64_t** err_dev /* host out */, void* stream);

/* ---------------------------------------------------------------- join fused with group-by */
/* An inner equi-join whose pairs feed only a group-by on the join key, with unique selected
 * build keys, computed without materialising the pairs: every group is a selected build key
 * that some selected probe row matches, and each matching probe row contributes its aggregate
 * arguments (segment j of `prog`, evaluated over probe columns at the probe row) once.  Groups
 * come out in ascending key order: out_key[g] = the key (in out_key_dtype), out_agg[j][g] as
 * listed at TQ_AGG_*.  Three launches, no synchronisation.
 *
 * Key bounds: every build key (selected or not) lies in [key_lo, key_hi]; the slots are
 * key - key_lo.  `pres` is a byte array of tq_groupjoin_pres_bytes(hi - lo + 1) bytes that the
 * caller zeroes once and passes again with a new epoch (1..255) each call: a byte holds the
 * epoch of the last call that saw its key, so it never needs clearing between calls; when the
 * epochs are used up the caller zeroes it again.  Calls sharing a `pres` array must be ordered
 * (one stream).
 *
 * *words (host out) receives the device address of int64 result words, valid once the stream
 * has synchronised: [0] groups, [1] pairs (matching probe rows), [2] selected build rows, [3]
 * distinct selected build keys, [4] error bits (1: a build key outside the bounds, 2: a
 * look-back wait timed out, 4: more groups than out_cap).  The outputs are valid only when
 * words[4] == 0 and words[2] == words[3] (unique keys) and words[1] > 0; otherwise the caller
 * runs the join and the group-by separately.  Fails with TQ_EUNSUPPORTED when the key range
 * needs more than 2048 reduce slices (2^12 keys each for one aggregate).
 * Replaces operators/hashjoin.py:24-177 (join_kernel_new: sort, searchsorted, bucket expansion
 * and gathers of the pair columns) followed by operators/aggregate.py:282-364 (agg_wrapper's
 * hashed grouping and the per-aggregate bincount/scatter_reduce of aggregate.py:127-193) when
 * the aggregate's only key is the join key. */
typedef struct {
  const void* bkey;          /* build key column (TQ_BOOL..TQ_I64), nb rows                  */
  int32_t bkey_dtype;
  const uint8_t* bmask;      /* selected build rows (NULL: all)                              */
  int64_t nb;
  const void* pkey;          /* probe key column, np rows                                    */
  int32_t pkey_dtype;
  const uint8_t* pmask;      /* selected probe rows (NULL: all)                              */
  int64_t np;
  int64_t key_lo, key_hi;    /* bounds of the build key column                               */
  uint32_t epoch;            /* 1..255, see above                                            */
  int32_t naggs;
  int32_t agg_op[TQ_GB_MAX_AGGS];
  int32_t agg_dtype[TQ_GB_MAX_AGGS];  /* dtype of each argument segment's value             */
  void* out_key;
  int32_t out_key_dtype;     /* TQ_BOOL..TQ_I64                                              */
  void* out_agg[TQ_GB_MAX_AGGS];
  int64_t out_cap;           /* rows of the outputs: min(hi - lo + 1, np) always suffices    */
  /* Filters fused into the call (NULL: none): a build row is selected when bmask selects it
   * and bpred holds for it, a probe row likewise with pmask and ppred.  Each predicate is a
   * program of one segment over columns of nb (build) or np (probe) rows that
   * tq_filter_is_conjunction accepts; the call fails with TQ_EUNSUPPORTED otherwise.  They
   * replace the tqp_filter calls (operators/filter.py:164-205) whose masks the join would
   * read, so the masks are never written. */
  const tq_program_t* bpred;  /* host */
  const tq_colset_t* bpred_cols;
  const tq_program_t* ppred;  /* host */
  const tq_colset_t* ppred_cols;
  /* 0 (or 3): the whole call.  1: only the build's launch (prog and cols may be NULL; the
   * aggregate and output fields are not read), so it runs while the caller prepares the rest;
   * 2: the rest, after a phase-1 call on the same stream with the same desc, pres and ws. */
  int32_t phases;
} tq_groupjoin_desc_t;

size_t tq_groupjoin_pres_bytes(int64_t key_range);
/* Workspace bytes (0 when the shape is unsupported); nload = prog->nload. */
size_t tq_groupjoin_ws_bytes(int64_t nb, int64_t np, int64_t key_range, int32_t naggs,
                             int32_t nload);
/* Words per probe record of a call with this program and these aggregates (host only, for byte
 * accounting): 1 (the slot) plus one per accumulated aggregate; a count of a constant other than
 * the null sentinel is taken from the slots' pair counts and has no word.  -1 for bad input.
 * No counterpart in the reference: it describes the fused kernels' intermediate records. */
int32_t tq_groupjoin_record_words(const tq_program_t* prog /* host */, int32_t naggs,
                                  const int32_t* agg_op /* host */,
                                  const int32_t* agg_dtype /* host */);
int tq_groupjoin_run(const tq_groupjoin_desc_t* desc /* host */, const tq_program_t* prog /* host */,
                     const tq_colset_t* cols /* host */, void* pres, size_t pres_bytes, void* ws,
                     size_t ws_bytes, int64_t** words /* host out */, void* stream);

/* tq_groupjoin_run without any allocation, for callers that cannot allocate device memory
 * (code under graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The workspace and the presence bytes live in a scratch array
 * the library holds on every GPU it runs on (64 MiB, until the process ends); desc->epoch is
 * ignored (the library keeps the epochs of its scratch).  The caller passes the inputs and
 * outputs as for tq_groupjoin_run, plus words_out (device, 8 int64) for the result words.
 * Fails with TQ_ECAPACITY, launching nothing, when the presence bytes and the workspace
 * (tq_groupjoin_pres_bytes + tq_groupjoin_ws_bytes) exceed the scratch.  Calls that overlap on
 * one GPU share its scratch and race. */
int tq_groupjoin_run_noalloc(const tq_groupjoin_desc_t* desc /* host */,
                             const tq_program_t* prog /* host */,
                             const tq_colset_t* cols /* host */, int64_t* words_out,
                             void* stream);

/* ---------------------------------------------------------------- batched pipelines */
/* Many filter -> join -> group-by pipelines in one persistent launch.  A join pipeline (mode 0)
 * joins its selected build rows (unique build keys) with its selected probe rows on key equality
 * and groups the pairs by the build key, reducing aggregate arguments computed over probe
 * columns.  A group-by pipeline (mode 1) has no build side: it groups its selected rows by up to
 * 8 integer key columns packed in mixed radix (as tq_groupby_accumulate does).
 * The group table is addressed by build key - kmin over the build key's [min, max] range, so it
 * is the join directory and the aggregation table at once, and groups come out in key order.
 * Work is a device queue of (pipeline, tile) items in six phases -- clear the table, insert
 * build rows, probe, reduce records, count groups per tile, write groups -- where an item waits
 * (bounded wait) for its pipeline's last earlier phase that has items.  A pipeline accumulates
 * in one of three ways: into an LDS copy of a small table (lds_table), with global atomics, or
 * (records) without atomics as the fused groupjoin does: each TQ_BATCH_CHUNK-row probe chunk writes one
 * record {slot within its key slice, accumulator values} per pair into the chunk's own region,
 * slice by slice (slice offsets in loff), and a reduce item per key slice folds the slice's
 * records from every chunk in LDS and writes the slice's cnt/acc slots.  No host
 * synchronisation occurs; the launch sequence can be captured into a graph.  A selected build
 * key outside [kmin, kmin + range) sets flags[0] bit 1, and a repeated one makes the two
 * keycount words differ (the caller reruns such a pipeline through the per-operator path).
 * Replaces, per pipeline, the whole chain tqp_filter (operators/filter.py:164-205) x2 ->
 * join_vortex (operators/hashjoin.py:179-248) -> tqp_hash_aggregate (operators/aggregate.py:
 * 236-393) grouped by the join key. */
#define TQ_BATCH_PHASES 6
#define TQ_BATCH_CHUNK 32768   /* probe rows per item of a records pipeline          */
#define TQ_BATCH_MAX_SLICES 2048

typedef struct {
  tq_program_t bprog;  /* build predicate; load slot 0 must be the build key column  */
  tq_colset_t bcols;
  tq_program_t pprog;  /* probe predicate; load slot 0 must be the probe key column;  */
  tq_colset_t pcols;   /*   group-by pipelines: slots 0..nkeys-1 are the key columns   */
  tq_program_t aprog;  /* aggregate arguments over probe columns, one segment each   */
  tq_colset_t acols;
  int64_t nb, np;      /* build / probe rows                                          */
  int64_t kmin, range; /* table slot = key - kmin (join) or packed key (group-by)     */
  int32_t naggs;
  int32_t key_dtype;   /* dtype of the build key (and of the group key output)        */
  int32_t agg_op[TQ_MAX_SEGS];
  int32_t agg_dtype[TQ_MAX_SEGS];
  uint8_t* pres;       /* join pipelines: device [range]: 1 = a selected build key      */
  uint32_t* cnt;       /* device [range]: rows per group                               */
  uint64_t* acc;       /* device [naggs * range]: accumulators                         */
  uint32_t* tile_counts; /* device [ntiles[4]]                                         */
  void* key_out;       /* device [min(nb, range)] group keys (join pipelines)          */
  void* agg_out[TQ_MAX_SEGS];
  int64_t* ngroups;    /* device [1]                                                   */
  int32_t* flags;      /* device [1]                                                   */
  int32_t ntiles[TQ_BATCH_PHASES];
  int32_t mode;        /* 0: join pipeline; 1: group-by pipeline (no build side)        */
  int32_t nkeys;       /* group-by pipelines: key columns (0 = one global group)        */
  int32_t lds_table;   /* 1: probe items accumulate in an LDS copy of the table first   */
  int64_t gkey_min[TQ_GB_MAX_KEYS];    /* group-by pipelines: mixed-radix packing      */
  int64_t gkey_stride[TQ_GB_MAX_KEYS];
  int64_t gkey_range[TQ_GB_MAX_KEYS];
  int32_t gkey_dtype[TQ_GB_MAX_KEYS];
  void* gkey_out[TQ_GB_MAX_KEYS];
  int32_t records;     /* 1: the probe writes records and phase 3 reduces them            */
  int32_t sb;          /* records: a key slice is 2^sb table slots                        */
  int32_t nslices;     /* records: ceil(range / 2^sb), at most TQ_BATCH_MAX_SLICES          */
  int32_t pad_;
  uint64_t* rec;       /* records: device [np * (1 + naggs)] words                         */
  uint32_t* loff;      /* records: device [(nslices + 1) * ntiles[2]]: slice s of chunk c
                        * starts at record loff[s * ntiles[2] + c] of the chunk's region;
                        * row nslices holds the chunk's records                           */
  int64_t* keycount;   /* join pipelines: device [2]: selected build rows, and distinct
                        * selected build keys (they differ when a build key repeats)       */
  int32_t fast[2];     /* set by tq_batch_prepare: the build / probe predicate runs as a
                        * conjunction of column-versus-constant terms over whole tiles     */
  uint64_t prep[2][32]; /* set by tq_batch_prepare (opaque)                                */
} tq_batch_pipe_t;

/* Fill the fast[] and prep[] fields of npipes host descriptors (before they are copied to the
 * device): a join pipeline's build or probe predicate that is a conjunction of column-versus-
 * constant comparisons is then evaluated over whole 4096-row tiles with all loads in flight, its
 * key looked up in the presence bits, and its aggregate arguments evaluated for the joined rows
 * only.  Returns the number of predicates prepared (>= 0) or a TQ_E* code. */
int tq_batch_prepare(tq_batch_pipe_t* pipes /* host */, int32_t npipes);

/* Counters the launch needs (queue head, per-pipeline phase completion, error word). */
size_t tq_batch_ws_bytes(int32_t npipes);

/* Run the batch.  pipes: device array of npipes descriptors; items: device int32 pairs
 * (pipeline, tile), phase-major, with phase_start[0..6] (host) the first item of each phase;
 * tile_rows: rows (or table slots) per item, a multiple of 1024 (a records pipeline's probe
 * items are TQ_BATCH_CHUNK rows); grid: persistent workgroups; stage_loads: the largest nload of
 * any program; lds_table_bytes: LDS bytes beside the stage, at least the largest LDS table
 * (range * (4 + 8 * naggs)) of the lds_table pipelines and, with a records pipeline, the
 * records scratch (TQ_BATCH_CHUNK / 32 + TQ_BATCH_MAX_SLICES words) and its slice tables
 * (2^sb * (4 + 8 * naggs)); at most 48 KB.  ws is cleared by the call
 * (kernels only).  An item's wait for its pipeline's previous phase polls with relaxed loads
 * and gives up after wait_us microseconds of wall-clock time (the device's constant-rate
 * clock): ws error word 1 then means a wait timed out (the launch drains without finishing and
 * the caller reruns the pipelines per operator). */
int tq_batch_run(const tq_batch_pipe_t* pipes, int32_t npipes, const int32_t* items,
                 const int64_t* phase_start /* host, 7 */, int32_t tile_rows, int32_t grid,
                 int32_t stage_loads, int32_t lds_table_bytes, void* ws, uint32_t wait_us,
                 void* stream);

#ifdef __cplusplus
}
#endif
#endif /* TQHIP_H */
