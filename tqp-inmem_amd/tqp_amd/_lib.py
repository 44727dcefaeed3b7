"""ctypes binding of libtqhip.so (the gfx950 kernels of the relational hot path).

The structures below mirror include/tqhip.h field for field.  `call()` turns a negative status
into the Python exception the reference raises for the same kind of failure (ValueError,
NotImplementedError, RuntimeError, MemoryError).  There is no fallback: importing this module
without the built library, or calling it without a GPU, raises.
"""
import ctypes
import os

import torch

from . import memory

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtqhip.so")

# ---------------------------------------------------------------- constants (tqhip.h)
TQ_OK, TQ_EINVAL, TQ_EUNSUPPORTED, TQ_EHIP, TQ_ECAPACITY = 0, -1, -2, -3, -4
TQ_BOOL, TQ_I8, TQ_I16, TQ_I32, TQ_I64, TQ_F32, TQ_F64 = range(7)

TQ_MAX_INSTR = 64
TQ_MAX_COLS = 16
TQ_NREG = 8
TQ_STRPOOL = 128
TQ_MAX_SEGS = 12
TQ_MAX_LOADS = 8

OPCODES = [
    "NOP", "LOAD", "CONST",
    "CVT_I64_F64", "CVT_I64_F32", "CVT_F32_F64", "CVT_F64_F32",
    "WRAP_I8", "WRAP_I16", "WRAP_I32",
    "ADD_I64", "SUB_I64", "MUL_I64",
    "ADD_F64", "SUB_F64", "MUL_F64", "DIV_F64",
    "ADD_F32", "SUB_F32", "MUL_F32", "DIV_F32",
    "LT_I64", "LE_I64", "GT_I64", "GE_I64", "EQ_I64", "NE_I64",
    "LT_F64", "LE_F64", "GT_F64", "GE_F64", "EQ_F64", "NE_F64",
    "LT_F32", "LE_F32", "GT_F32", "GE_F32", "EQ_F32", "NE_F32",
    "ISCLOSE_F64",
    "AND", "OR", "NOT",
    "WHERE",
    "STR_EQ", "STR_PREFIX",
    "YEAR",
    "STR_LIKE2",
    "STR_CONTAINS",
]
OP = {name: i for i, name in enumerate(OPCODES)}

TQ_AGG_SUM, TQ_AGG_COUNT, TQ_AGG_AVG, TQ_AGG_MIN, TQ_AGG_MAX = range(5)
TQ_GB_MAX_KEYS = 8
TQ_GB_DENSE_LDS, TQ_GB_DENSE_GLOBAL, TQ_GB_HASH, TQ_GB_DENSE_PART, TQ_GB_LDS_FRONT = 0, 1, 2, 3, 8
TQ_GB_DICT_MAX_RANGE = 1 << 16
TQ_GB_LDS_LANES = 16
TQ_JOIN_EMPTY, TQ_JOIN_DENSE, TQ_JOIN_HASH = 0, 1, 2

DTYPE_CODE = {
    torch.bool: TQ_BOOL, torch.int8: TQ_I8, torch.int16: TQ_I16, torch.int32: TQ_I32,
    torch.int64: TQ_I64, torch.float32: TQ_F32, torch.float64: TQ_F64,
}
CODE_DTYPE = {v: k for k, v in DTYPE_CODE.items()}


# ---------------------------------------------------------------- structures
class Instr(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint8), ("dst", ctypes.c_uint8), ("a", ctypes.c_uint8),
                ("b", ctypes.c_uint8), ("c", ctypes.c_uint8), ("pad", ctypes.c_uint8 * 3),
                ("imm", ctypes.c_int64)]


class Program(ctypes.Structure):
    _fields_ = [("n_instr", ctypes.c_int32), ("nseg", ctypes.c_int32),
                ("seg_end", ctypes.c_int16 * TQ_MAX_SEGS),
                ("seg_result", ctypes.c_uint8 * TQ_MAX_SEGS),
                ("nload", ctypes.c_int32),
                ("load_col", ctypes.c_uint8 * TQ_MAX_LOADS),
                ("load_dtype", ctypes.c_uint8 * TQ_MAX_LOADS),
                ("code", Instr * TQ_MAX_INSTR),
                ("strpool", ctypes.c_uint8 * TQ_STRPOOL)]


class ColSet(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p * TQ_MAX_COLS), ("width", ctypes.c_int32 * TQ_MAX_COLS)]


class KTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("ms", ctypes.c_float), ("pad", ctypes.c_int32)]


class SortKey(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("dtype", ctypes.c_int32), ("width", ctypes.c_int32),
                ("descending", ctypes.c_int32), ("pad", ctypes.c_int32)]


class JoinState(ctypes.Structure):
    _fields_ = [("nsel", ctypes.c_int64), ("key_bits", ctypes.c_int32),
                ("layout", ctypes.c_int32), ("unique", ctypes.c_int32),
                ("hash_bits", ctypes.c_int32), ("base", ctypes.c_uint64),
                ("empty_key", ctypes.c_uint64), ("range", ctypes.c_int64),
                ("table", ctypes.c_void_p), ("bits", ctypes.c_void_p), ("rows", ctypes.c_void_p),
                ("pending", ctypes.c_int32), ("bdtype", ctypes.c_int32),
                ("bkeys", ctypes.c_void_p), ("bmask", ctypes.c_void_p), ("nb", ctypes.c_int64),
                ("ws", ctypes.c_void_p), ("key_or", ctypes.c_uint64), ("key_and", ctypes.c_uint64)]


class GroupByDesc(ctypes.Structure):
    _fields_ = [("nkeys", ctypes.c_int32), ("naggs", ctypes.c_int32),
                ("key_dtype", ctypes.c_int32 * TQ_GB_MAX_KEYS),
                ("key_ptr", ctypes.c_void_p * TQ_GB_MAX_KEYS),
                ("key_min", ctypes.c_int64 * TQ_GB_MAX_KEYS),
                ("key_range", ctypes.c_int64 * TQ_GB_MAX_KEYS),
                ("key_stride", ctypes.c_int64 * TQ_GB_MAX_KEYS),
                ("agg_op", ctypes.c_int32 * TQ_MAX_SEGS),
                ("agg_dtype", ctypes.c_int32 * TQ_MAX_SEGS),
                ("nsel", ctypes.c_int64), ("nrows", ctypes.c_int64),
                ("key_dict", ctypes.c_void_p * TQ_GB_MAX_KEYS),
                ("key_dict_range", ctypes.c_int64 * TQ_GB_MAX_KEYS)]


class GroupJoinDesc(ctypes.Structure):
    _fields_ = [("bkey", ctypes.c_void_p), ("bkey_dtype", ctypes.c_int32),
                ("bmask", ctypes.c_void_p), ("nb", ctypes.c_int64),
                ("pkey", ctypes.c_void_p), ("pkey_dtype", ctypes.c_int32),
                ("pmask", ctypes.c_void_p), ("np", ctypes.c_int64),
                ("key_lo", ctypes.c_int64), ("key_hi", ctypes.c_int64),
                ("epoch", ctypes.c_uint32), ("naggs", ctypes.c_int32),
                ("agg_op", ctypes.c_int32 * TQ_MAX_SEGS),
                ("agg_dtype", ctypes.c_int32 * TQ_MAX_SEGS),
                ("out_key", ctypes.c_void_p), ("out_key_dtype", ctypes.c_int32),
                ("out_agg", ctypes.c_void_p * TQ_MAX_SEGS), ("out_cap", ctypes.c_int64),
                ("bpred", ctypes.c_void_p), ("bpred_cols", ctypes.c_void_p),
                ("ppred", ctypes.c_void_p), ("ppred_cols", ctypes.c_void_p),
                ("phases", ctypes.c_int32)]


GJ_GROUPS, GJ_PAIRS, GJ_NSEL, GJ_NSET, GJ_ERR = range(5)


# ---------------------------------------------------------------- loading
_lib = None

P, I32, I64, SZ = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_size_t
_SIGNATURES = {
    "tq_version": ([], I32),
    "tq_last_error": ([], ctypes.c_char_p),
    "tq_tile_rows": ([], I32),
    "tq_timing": ([I32], I32),
    "tq_timing_count": ([], I64),
    "tq_timing_read": ([P, I64, P, P], I32),
    "tq_filter_mask": ([P, P, I64, P, P, P], I32),
    "tq_filter_is_conjunction": ([P, P], I32),
    "tq_eval_columns": ([P, P, I64, P, P, P], I32),
    "tq_mask_tile_counts": ([P, I64, P, P], I32),
    "tq_scan_ws_bytes": ([I64], SZ),
    "tq_scan_tiles": ([P, I64, P, P, P, P], I32),
    "tq_compact": ([P, I64, P, I32, P, P, P, P], I32),
    "tq_gather": ([P, I64, I32, P, P, P, P, I32, P], I32),
    "tq_mark_rows": ([P, I64, P, I64, P], I32),
    "tq_eval_column": ([P, P, I64, P, I32, P], I32),
    "tq_sort_ws_bytes": ([I64], SZ),
    "tq_sort_perm": ([P, I32, I64, I64, P, P, SZ, P], I32),
    "tq_sort_perm_deferred": ([P, I32, I64, P, I64, P, P, SZ, P, P], I32),
    "tq_batch_prepare": ([P, I32], I32),
    "tq_topk_gather": ([P, I32, I64, P, I64, I32, P, P, P, P, P, SZ, P, P], I32),
    "tq_join_build_ws_bytes": ([I64], SZ),
    "tq_join_build": ([P, I32, P, I64, P, SZ, P, P], I32),
    "tq_join_build_ranged": ([P, I32, P, I64, I64, I64, P, SZ, P, P], I32),
    "tq_join_probe_ws_bytes": ([I64], SZ),
    "tq_join_count": ([P, P, I32, P, I64, P, P, P], I32),
    "tq_join_write": ([P, P, I32, P, I64, P, I32, P, P, P, I32, P, P, P, P, P, P], I32),
    "tq_groupby_bounds_ws_bytes": ([I64], SZ),
    "tq_groupby_bounds": ([P, P, I64, P, P, P], I32),
    "tq_groupby_plan": ([P, I64, I32, P, P, P], I32),
    "tq_groupby_dict_ws_bytes": ([P], SZ),
    "tq_groupby_dict": ([P, P, I64, P, SZ, P, P, P], I32),
    "tq_groupby_accumulate": ([P, P, P, P, I64, I32, I64, P, SZ, P], I32),
    "tq_groupby_count": ([P, I32, I64, P, P, P], I32),
    "tq_groupby_write": ([P, I32, I64, P, I64, P, P, P], I32),
    "tq_groupby_run": ([P, P, P, P, I64, I32, I64, P, SZ, P, P, I64, P, P], I32),
    "tq_groupby_run_deferred": ([P, P, P, P, I64, P, I32, I64, P, SZ, P, P, I64, P, P, P], I32),
    "tq_join_probe_deferred": ([P, P, I32, P, I64, P, I32, P, P, P, I32, P, P, P, P, P, P, P],
                               I32),
    "tq_join_settle": ([P, P, P], I32),
    "tq_read_words": ([I32, P, P, P, P], I32),
    "tq_word_slots": ([], I32),
    "tq_words_post": ([I32, P, P, I32, P], I32),
    "tq_words_wait": ([I32, P], I32),
    "tq_rank_ws_bytes": ([I64], SZ),
    "tq_rank_sorted": ([P, I32, I64, P, P, P, P, P, SZ, P], I32),
    "tq_groupjoin_pres_bytes": ([I64], SZ),
    "tq_groupjoin_ws_bytes": ([I64, I64, I64, I32, I32], SZ),
    "tq_groupjoin_record_words": ([P, I32, P, P], I32),
    "tq_groupjoin_run": ([P, P, P, P, SZ, P, SZ, P, P], I32),
    "tq_groupjoin_run_noalloc": ([P, P, P, P, P], I32),
    "tq_batch_ws_bytes": ([I32], SZ),
    "tq_batch_run": ([P, I32, P, P, I32, I32, I32, I32, P, ctypes.c_uint32, P], I32),
}


def load():
    """Load the library (once).  Raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


_ERRORS = {TQ_EINVAL: ValueError, TQ_EUNSUPPORTED: NotImplementedError, TQ_EHIP: RuntimeError,
           TQ_ECAPACITY: MemoryError}

# Kernel timing.  While TRACE is a list, the library stamps every kernel launch on the device
# (tq_timing) and every call appends (entry point, principal kernel, algorithmic bytes, first,
# end): launches first..end-1 of the library's timing log are the ones this call made.
TRACE = None


def call(name, *args, alg_bytes=None, kernel=None):
    lib = load()
    fn = getattr(lib, name)
    if TRACE is None:
        rc = fn(*args)
    else:
        c0 = lib.tq_timing_count()
        rc = fn(*args)
        c1 = lib.tq_timing_count()
        if isinstance(kernel, dict):    # {kernel: algorithmic bytes of this call}
            for k, b in kernel.items():
                TRACE.append((name, k, b, c0, c1))
        elif isinstance(kernel, tuple):   # several principal kernels, bytes listed alike
            for k, b in zip(kernel, alg_bytes):
                TRACE.append((name, k, b, c0, c1))
        else:
            TRACE.append((name, kernel, alg_bytes, c0, c1))
    if rc != TQ_OK:
        msg = lib.tq_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)(f"{name}: {msg}")
    return rc


def trace_len():
    """Entries in the call log so far (None while not timing)."""
    return None if TRACE is None else len(TRACE)


def trace_set_bytes(at, kernel, alg_bytes):
    """Set the algorithmic bytes of the logged entry for `kernel` at or after index `at` (for a
    call whose bytes depend on a count read back later)."""
    if TRACE is None or at is None:
        return
    for i in range(at, len(TRACE)):
        e = TRACE[i]
        if e[1] == kernel:
            TRACE[i] = (e[0], e[1], alg_bytes, e[3], e[4])
            return


def start_trace():
    """Clear the timing log and start timing kernel launches."""
    global TRACE
    lib = load()
    torch.cuda.synchronize()
    _read_log(lib)
    TRACE = []
    lib.tq_timing(1)


def stop_trace():
    """Stop timing; returns (calls, launches) where launches is [(kernel name, ms)] in launch
    order and calls index into it.  Synchronises the device first."""
    global TRACE
    lib = load()
    torch.cuda.synchronize()
    lib.tq_timing(0)
    calls, TRACE = TRACE, None
    return calls, _read_log(lib)


def _read_log(lib):
    cap = max(int(lib.tq_timing_count()), 1)
    buf = (KTime * cap)()
    n, dropped = ctypes.c_int64(), ctypes.c_int64()
    rc = lib.tq_timing_read(buf, cap, ctypes.byref(n), ctypes.byref(dropped))
    if rc != TQ_OK:
        raise RuntimeError(lib.tq_last_error().decode(errors="replace"))
    if dropped.value:
        raise RuntimeError(f"kernel timing log overflowed ({dropped.value} launches untimed)")
    return [(buf[i].name.decode(), float(buf[i].ms)) for i in range(n.value)]


def size_call(name, *args) -> int:
    return int(getattr(load(), name)(*args))


# ---------------------------------------------------------------- helpers
def stream_handle():
    """torch's current stream on the current device, as a raw hipStream_t (the C calls below
    skip torch.cuda.current_stream's Python-level device lookup)."""
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


def ptr(t: torch.Tensor):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def require_gpu(t: torch.Tensor, what="column"):
    if not t.is_cuda:
        raise RuntimeError(f"{what} must be a GPU tensor (the HIP path has no CPU fallback)")


def dtype_code(dt) -> int:
    try:
        return DTYPE_CODE[dt]
    except KeyError:
        raise NotImplementedError(f"unsupported column dtype {dt}")


def row_bytes(t: torch.Tensor) -> int:
    """Bytes per row: element size, times the width of a 2-D column."""
    w = t.shape[1] if t.dim() == 2 else 1
    return t.element_size() * w


def ptr_array(tensors):
    arr = (ctypes.c_void_p * max(len(tensors), 1))()
    for i, t in enumerate(tensors):
        arr[i] = t.data_ptr()
    return arr


def int_array(vals):
    arr = (ctypes.c_int32 * max(len(vals), 1))()
    for i, v in enumerate(vals):
        arr[i] = int(v)
    return arr


_WORKSPACES = {}


def workspace(nbytes: int, device, role=None) -> torch.Tensor:
    """Raw device workspace.  With a role, the buffer is kept per (role, device, stream) and
    handed out again to the next call of that role on the same stream, which runs after the
    previous one's kernels in stream order; a holder must not keep using its buffer once
    another call of its role has been issued.  While a device arena is active on the thread,
    the workspace comes from it."""
    nbytes = max(int(nbytes), 256)
    arena = memory.active()
    if arena is not None:
        return arena.workspace(nbytes, role)
    if role is None:
        return torch.empty(nbytes, dtype=torch.uint8, device=device)
    # the buffer is kept per device and per the current device's current stream
    cur = torch._C._cuda_getDevice()
    key = (role, device, cur, torch._C._cuda_getCurrentRawStream(cur))
    ent = _WORKSPACES.get(key)
    if ent is None or ent[1] < nbytes:
        buf = torch.empty(nbytes + nbytes // 4, dtype=torch.uint8, device=device)
        ent = (buf, buf.numel(), {})
        _WORKSPACES[key] = ent
    views = ent[2]             # views of the buffer by size, made once
    v = views.get(nbytes)
    if v is None:
        if len(views) > 64:
            views.clear()
        v = views[nbytes] = ent[0][:nbytes]
    return v
