"""Tensor-level wrappers over the C ABI: each takes and returns torch GPU tensors.

These are the seams where the reference calls ATen (filter.evaluate, the `tensor[mask]` of the
pipeline, join_kernel_new, agg_wrapper, the sort passes); the operator modules call them.
Outputs are allocated here with torch's caching allocator, workspaces likewise; the library only
launches kernels on the current stream.
"""
import ctypes
import functools
import threading
import weakref

import torch

from . import _lib as L
from . import expr as E
from . import memory


def _dev(t):
    return t.device


# ---------------------------------------------------------------- filter / compaction
def tile_rows() -> int:
    return int(L.load().tq_tile_rows())


def ntiles(n: int) -> int:
    tr = tile_rows()
    return (n + tr - 1) // tr


def filter_program(tree, tensors, subquery_result=None):
    """The predicate tree compiled and bound to its columns (cached per tree)."""
    return E.cached_program(("filter", repr(tree), repr(subquery_result)), tensors,
                            lambda pb: E.compile_predicate(pb, tree, tensors, subquery_result))


def filter_mask(tree, tensors, subquery_result=None, n=None):
    """Bool mask of a predicate tree (operators/filter.py:14-162) plus per-tile counts."""
    b = filter_program(tree, tensors, subquery_result)
    rows = b.n if b.n is not None else n
    if rows is None:
        rows = next(iter(tensors.values())).tensor.shape[0]
    dev = b.col_tensors[0].device if b.col_tensors else next(iter(tensors.values())).tensor.device
    mask = memory.empty(rows, dtype=torch.bool, device=dev)
    counts = memory.empty(max(ntiles(rows), 1), dtype=torch.int32, device=dev)
    if not b.col_tensors:
        # no column referenced (e.g. only isnotnull): the kernel's column set is never read,
        # point it at the mask
        b.cols.ptr[0] = mask.data_ptr()
        b.cols.width[0] = 1
    # read each loaded column once, write one mask byte per row
    alg = rows * (1 + b.load_bytes())
    L.call("tq_filter_mask", E.program_ptr(b), E.colset_ptr(b), rows, L.ptr(mask), L.ptr(counts),
           L.stream_handle(), kernel={"k_filter_conj": alg, "k_filter": alg})
    return mask, counts


def mask_tile_counts(mask: torch.Tensor):
    n = mask.shape[0]
    counts = memory.empty(max(ntiles(n), 1), dtype=torch.int32, device=mask.device)
    L.call("tq_mask_tile_counts", L.ptr(mask), n, L.ptr(counts), L.stream_handle())
    return counts


def scan_tiles(counts: torch.Tensor, nt: int):
    """Exclusive scan of tile counts -> (offsets int64[nt], total int64[1]) on the device."""
    dev = counts.device
    offs = memory.empty(max(nt, 1), dtype=torch.int64, device=dev)
    total = memory.empty(1, dtype=torch.int64, device=dev)
    ws = L.workspace(L.size_call("tq_scan_ws_bytes", nt), dev, "scan")
    L.call("tq_scan_tiles", L.ptr(counts), nt, L.ptr(offs), L.ptr(total), L.ptr(ws),
           L.stream_handle())
    return offs, total


def compact(mask: torch.Tensor, columns, tile_counts=None):
    """`[c[mask] for c in columns]` in one launch per 16 columns (vortex_pipeline.py:58)."""
    L.require_gpu(mask, "mask")
    if mask.dtype != torch.bool:
        raise ValueError("compaction mask must be bool")
    n = mask.shape[0]
    for c in columns:
        L.require_gpu(c)
        if c.shape[0] != n:
            raise IndexError(f"mask of {n} rows applied to a column of {c.shape[0]} rows")
    nt = ntiles(n)
    if tile_counts is None:
        tile_counts = mask_tile_counts(mask)
    offs, total = scan_tiles(tile_counts, nt)
    m = int(total.item()) if n > 0 else 0
    outs = [memory.empty((m,) + tuple(c.shape[1:]), dtype=c.dtype, device=c.device) for c in columns]
    for i in range(0, len(columns), L.TQ_MAX_COLS):
        src = [c.contiguous() for c in columns[i:i + L.TQ_MAX_COLS]]
        dst = outs[i:i + L.TQ_MAX_COLS]
        if m == 0:
            continue
        rb = sum(L.row_bytes(c) for c in src)
        L.call("tq_compact", L.ptr(mask), n, L.ptr(offs), len(src), L.ptr_array(src),
               L.int_array([L.row_bytes(c) for c in src]), L.ptr_array(dst), L.stream_handle(),
               alg_bytes=n + 2 * m * rb, kernel="k_compact")
    return outs


def gather(idx: torch.Tensor, columns, null_fill=False):
    """`[c[idx] for c in columns]`; with null_fill, idx == -64 yields -64 (index_with_null)."""
    m = idx.shape[0]
    idx = idx.to(torch.int64).contiguous()
    outs = [memory.empty((m,) + tuple(c.shape[1:]), dtype=c.dtype, device=c.device) for c in columns]
    if m == 0:
        return outs
    for i in range(0, len(columns), L.TQ_MAX_COLS):
        src = [c.contiguous() for c in columns[i:i + L.TQ_MAX_COLS]]
        fk = [8 if c.dtype == torch.float64 else 4 if c.dtype == torch.float32 else 0 for c in src]
        rb = sum(L.row_bytes(c) for c in src)
        # the row ids, each gathered row read once and written once
        L.call("tq_gather", L.ptr(idx), m, len(src), L.ptr_array(src),
               L.int_array([L.row_bytes(c) for c in src]), L.int_array(fk),
               L.ptr_array(outs[i:i + L.TQ_MAX_COLS]), 1 if null_fill else 0, L.stream_handle(),
               kernel={"k_gather": m * (8 + 2 * rb)})
    return outs


def eval_programs(b: E.ProgramBuilder, n: int, out_dtypes, device):
    """Every segment of b per row into its own column (dtype out_dtypes[a]), one launch."""
    outs = [memory.empty(n, dtype=dt, device=device) for dt in out_dtypes]
    if n > 0:
        if not b.col_tensors:
            b.column(outs[0])
        ptrs = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
        widths = L.int_array([o.element_size() for o in outs])
        rb = sum(o.element_size() for o in outs)
        L.call("tq_eval_columns", E.program_ptr(b), E.colset_ptr(b), n, ptrs, widths,
               L.stream_handle(), kernel={"k_eval_multi": n * (b.load_bytes() + rb)})
    return outs


def eval_program(b: E.ProgramBuilder, n: int, out_dtype, device):
    out = memory.empty(n, dtype=out_dtype, device=device)
    if n > 0:
        if not b.col_tensors:
            b.column(out)
        L.call("tq_eval_column", E.program_ptr(b), E.colset_ptr(b), n, L.ptr(out),
               out.element_size(), L.stream_handle())
    return out


# ---------------------------------------------------------------- sort
@functools.lru_cache(maxsize=256)
def _sort_ws_bytes(n: int) -> int:
    return L.size_call("tq_sort_ws_bytes", n)


def sort_perm(keys, n: int, limit=None, device=None):
    """Stable multi-key sort permutation (int64[min(n, limit)]).  keys: list of
    (tensor, descending) from most to least significant; 2-D int8 keys compare row-wise.

    Under a deferral whose group-by left its group count on the device, a sort of those groups
    with a limit of at most 64 reads the count there too (the permutation then has
    min(n, limit) entries, of which the first min(groups, limit) are valid); any other sort of
    them reads the count back first and sorts that many rows."""
    df = _deferred()
    pending = (df is not None and df.group is not None and df.G is None
               and n == df.group["cap"])
    if pending:
        outs = {t.data_ptr() for t in df.group["outs"]}
        if not all(t.data_ptr() in outs for t, _ in keys):
            pending = False
        elif limit is None or limit < 1 or limit > 64:
            n = df.resolve()[1]
            pending = False
    m = n if limit is None or limit < 0 or limit > n else int(limit)
    dev = device if device is not None else keys[0][0].device
    perm = memory.empty(m, dtype=torch.int64, device=dev)
    if n == 0 or m == 0:
        return perm
    arr = (L.SortKey * max(len(keys), 1))()
    hold = []
    for i, (t, desc) in enumerate(keys):
        L.require_gpu(t, "sort key")
        if t.dim() == 2 and t.shape[1] == 1:
            t = t.reshape(-1)
        t = t.contiguous()
        hold.append(t)
        arr[i].ptr = t.data_ptr()
        arr[i].dtype = L.dtype_code(t.dtype)
        arr[i].width = t.shape[1] if t.dim() == 2 else 1
        arr[i].descending = 1 if desc else 0
    ws_bytes = _sort_ws_bytes(n)
    ws = L.workspace(ws_bytes, dev, "sort")
    kb = sum(L.row_bytes(t) for t in hold)
    # top-k rounds read every key once (later rounds read the few survivors); a packing pass
    # reads the keys and writes one 12-byte (key, row) pair per row
    sort_bytes = {"k_topk_block": n * kb, "k_topk_select": n * 12, "k_pack_keys": n * (kb + 12)}
    if pending:
        applied = ctypes.c_int32()
        at = L.trace_len()
        L.call("tq_sort_perm_deferred", arr, len(keys), n, ctypes.c_void_p(df.group["ngroups"]),
               int(limit), L.ptr(perm), L.ptr(ws), ws_bytes, ctypes.byref(applied),
               L.stream_handle(), kernel=sort_bytes)
        if applied.value:
            df.sorted = {"cap": m, "limit": int(limit), "trace": at, "kb": kb}
            return perm
        return sort_perm(keys, df.resolve()[1], limit, device)
    L.call("tq_sort_perm", arr, len(keys), n, -1 if limit is None else int(limit), L.ptr(perm),
           L.ptr(ws), ws_bytes, L.stream_handle(), kernel=sort_bytes)
    return perm


def take_ordered_deferred(keys, cols, n: int, limit: int, device):
    """A take-ordered of at most 64 rows over the outputs of the active deferral's group-by
    (group count still on the device): the sort and the gathers of `cols` in one call
    (tq_topk_gather), the gathers inside the top-k's last kernel.  Returns the gathered
    columns (min(n, limit) rows, the first min(groups, limit) valid), or None when this form
    does not apply (then nothing was launched)."""
    df = _deferred()
    if df is None or df.group is None or df.G is not None or n != df.group["cap"] \
            or not 1 <= limit <= 64 or len(cols) > L.TQ_MAX_COLS:
        return None
    outs = {t.data_ptr() for t in df.group["outs"]}
    arr = (L.SortKey * len(keys))()
    kb = 0
    for i, (t, desc) in enumerate(keys):
        if t.dim() == 2 and t.shape[1] == 1:
            t = t.reshape(-1)
        if t.dim() != 1 or not t.is_contiguous() or t.data_ptr() not in outs:
            return None
        arr[i].ptr = t.data_ptr()
        arr[i].dtype = L.dtype_code(t.dtype)
        arr[i].width = 1
        arr[i].descending = 1 if desc else 0
        kb += t.element_size()
    m = min(n, limit)
    nc = len(cols)
    src = (ctypes.c_void_p * max(nc, 1))()
    dst = (ctypes.c_void_p * max(nc, 1))()
    rb = (ctypes.c_int32 * max(nc, 1))()
    gathered = []
    for i, c in enumerate(cols):
        L.require_gpu(c, "sort input")
        if not c.is_contiguous():
            return None
        o = memory.empty((m,) + tuple(c.shape[1:]), dtype=c.dtype, device=device)
        gathered.append(o)
        src[i], dst[i], rb[i] = c.data_ptr(), o.data_ptr(), L.row_bytes(c)
    perm = L.workspace(8 * m, device, "topk_perm")
    ws_bytes = _sort_ws_bytes(n)
    ws = L.workspace(ws_bytes, device, "sort")
    applied = ctypes.c_int32()
    at = L.trace_len()
    L.call("tq_topk_gather", arr, len(keys), n, ctypes.c_void_p(df.group["ngroups"]), limit, nc,
           src, rb, dst, L.ptr(perm), L.ptr(ws), ws_bytes, ctypes.byref(applied),
           L.stream_handle(), kernel={"k_topk_block": n * kb})
    if not applied.value:
        return None
    df.sorted = {"cap": m, "limit": int(limit), "trace": at, "kb": kb}
    return gathered


# ---------------------------------------------------------------- value bounds
# Host-known value bounds of integer columns: a join writes its key outputs from keys whose
# range its build already measured, so a group-by over those columns can plan without its own
# bounds pass and synchronisation.  An entry is only trusted while the tensor's in-place
# version counter is unchanged.  Tensors created under torch.inference_mode() (the reference's
# main.py runs every query inside it) have no version counter, so an in-place write to them
# could not be noticed: no bounds are recorded or trusted for them, and the consumers measure
# the bounds on the device instead.
_BOUNDS = {}   # id(tensor) -> (weak reference, version, lo, hi); tensors compare elementwise


def note_bounds(t: torch.Tensor, lo: int, hi: int):
    if t.is_inference():
        return
    k = id(t)
    _BOUNDS[k] = (weakref.ref(t, lambda _r, k=k: _BOUNDS.pop(k, None)), t._version, int(lo), int(hi))


def known_bounds(t: torch.Tensor):
    """(lo, hi) recorded for t, or None."""
    e = _BOUNDS.get(id(t))
    if e is None or e[0]() is not t or t.is_inference() or e[1] != t._version:
        return None
    return e[2], e[3]


def _untransform(bits: int, key_bits: int) -> int:
    if key_bits == 32:
        v = (bits ^ 0x80000000) & 0xFFFFFFFF
        return v - (1 << 32) if v >= 1 << 31 else v
    v = (bits ^ 0x8000000000000000) & 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= 1 << 63 else v


# ---------------------------------------------------------------- join
class JoinIndex:
    """A built join index over the selected rows of a build key column."""

    def __init__(self, bkey: torch.Tensor, bmask=None):
        L.require_gpu(bkey, "join key")
        self.bkey = bkey.contiguous()
        self.bmask = bmask
        nb = self.bkey.shape[0]
        ws_bytes = L.size_call("tq_join_build_ws_bytes", nb)
        self.ws = L.workspace(ws_bytes, bkey.device, "join_build")
        self.state = L.JoinState()
        self.deferred = False   # pairs written by a deferred probe, not yet settled
        bm = L.ptr(bmask) if bmask is not None else None
        kb = known_bounds(self.bkey)
        at = L.trace_len()
        names = {"k_join_stats": 0, "k_join_clear": 0, "k_join_claim": 0, "k_join_bits": 0}
        if kb is not None:
            # bounds of the whole key column pick the layout: no statistics pass, no wait
            L.call("tq_join_build_ranged", L.ptr(self.bkey), L.dtype_code(self.bkey.dtype), bm, nb,
                   kb[0], kb[1], L.ptr(self.ws), ws_bytes, ctypes.byref(self.state),
                   L.stream_handle(), kernel=names)
        else:
            L.call("tq_join_build", L.ptr(self.bkey), L.dtype_code(self.bkey.dtype), bm, nb,
                   L.ptr(self.ws), ws_bytes, ctypes.byref(self.state), L.stream_handle(),
                   kernel=names)
        if at is not None:
            # keys and mask read by the statistics and claim passes; the directory (4 bytes an
            # entry, 16 in a hash table) cleared, then read to build its presence bitmap
            rb = nb * (self.bkey.element_size() + (1 if bmask is not None else 0))
            ent = int(self.state.range) * (16 if self.state.layout == L.TQ_JOIN_HASH else 4)
            for k, v in {"k_join_stats": rb, "k_join_claim": rb, "k_join_clear": ent,
                         "k_join_bits": ent + int(self.state.range) // 8}.items():
                L.trace_set_bytes(at, k, v)

    @property
    def nsel(self):
        """Selected build rows (-1 after a bounded build until the first probe count)."""
        return int(self.state.nsel)

    def key_bounds(self):
        """Bounds of every key that can have a pair: the dense directory's key range (the
        selected build keys' measured range, or the bounds a ranged build was given), else
        None."""
        st = self.state
        if st.nsel == 0 or st.layout != L.TQ_JOIN_DENSE:
            return None
        return (_untransform(st.base, st.key_bits),
                _untransform(st.base + st.range - 1, st.key_bits))

    def probe(self, pkey: torch.Tensor, pmask, left_cols, right_cols, want_index=False):
        """All (build, probe) pairs with equal keys, probe row ascending then build row
        ascending.  Returns (left outputs, right outputs, build_idx, probe_idx)."""
        L.require_gpu(pkey, "join key")
        pkey = pkey.contiguous()
        np_ = pkey.shape[0]
        dev = pkey.device
        if len(left_cols) > L.TQ_MAX_COLS or len(right_cols) > L.TQ_MAX_COLS:
            raise NotImplementedError("join with more than 16 columns per side")
        pws = L.workspace(L.size_call("tq_join_probe_ws_bytes", np_), dev, "join_probe")
        total = ctypes.c_int64(0)
        pm = L.ptr(pmask) if pmask is not None else None
        s = L.stream_handle()
        # everything the write pass needs but the pair count is prepared before the count's
        # synchronisation, so the device waits as little as possible for the write launch
        lsrc = [c.contiguous() for c in left_cols]
        rsrc = [c.contiguous() for c in right_cols]
        lbytes = [L.row_bytes(c) for c in lsrc]
        rbytes = [L.row_bytes(c) for c in rsrc]
        lsp, rsp = L.ptr_array(lsrc), L.ptr_array(rsrc)
        lba, rba = L.int_array(lbytes), L.int_array(rbytes)
        # the join's boundary bytes less what the build and count passes read: every other
        # gathered input column read once over its table, every output written once
        seen = {self.bkey.data_ptr(), pkey.data_ptr()}
        reads = 0
        for c, rb in zip(lsrc + rsrc, lbytes + rbytes):
            if c.data_ptr() not in seen:
                seen.add(c.data_ptr())
                reads += c.shape[0] * rb
        per_pair = sum(lbytes) + sum(rbytes) + (16 if want_index else 0)
        count_bytes = np_ * (pkey.element_size() + (1 if pmask is not None else 0))
        d = _deferred()
        if d is not None and d.join is None and not want_index and np_ > 0:
            # outputs for every probe row (a unique build gives a probe row one pair at most);
            # the pair count stays on the device until the deferral resolves
            lout = [memory.empty((np_,) + tuple(c.shape[1:]), dtype=c.dtype, device=dev)
                    for c in lsrc]
            rout = [memory.empty((np_,) + tuple(c.shape[1:]), dtype=c.dtype, device=dev)
                    for c in rsrc]
            tot, chk, applied = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int32()
            at = L.trace_len()
            L.call("tq_join_probe_deferred", ctypes.byref(self.state), L.ptr(pkey),
                   L.dtype_code(pkey.dtype), pm, np_, L.ptr(pws), len(lsrc), lsp, lba,
                   L.ptr_array(lout), len(rsrc), rsp, rba, L.ptr_array(rout), ctypes.byref(tot),
                   ctypes.byref(chk), ctypes.byref(applied), s,
                   alg_bytes=(count_bytes, 0), kernel=("k_join_count", "k_join_write"))
            if applied.value:
                self.deferred = True
                d.join = {"index": self, "total": tot.value, "check": chk.value,
                          "outs": lout + rout, "cap": np_, "ws": pws, "trace": at,
                          "write_bytes": (reads, per_pair)}
                return lout, rout, None, None
        L.call("tq_join_count", ctypes.byref(self.state), L.ptr(pkey), L.dtype_code(pkey.dtype), pm,
               np_, L.ptr(pws), ctypes.byref(total), s, alg_bytes=count_bytes,
               kernel="k_join_count")
        m = int(total.value)
        lout = [memory.empty((m,) + tuple(c.shape[1:]), dtype=c.dtype, device=dev) for c in lsrc]
        rout = [memory.empty((m,) + tuple(c.shape[1:]), dtype=c.dtype, device=dev) for c in rsrc]
        bidx = memory.empty(m, dtype=torch.int64, device=dev) if want_index else None
        pidx = memory.empty(m, dtype=torch.int64, device=dev) if want_index else None
        if m > 0:
            L.call("tq_join_write", ctypes.byref(self.state), L.ptr(pkey), L.dtype_code(pkey.dtype),
                   pm, np_, L.ptr(pws), len(lsrc), lsp, lba, L.ptr_array(lout), len(rsrc), rsp, rba,
                   L.ptr_array(rout), L.ptr(bidx), L.ptr(pidx), s, alg_bytes=reads + m * per_pair,
                   kernel="k_join_write")
        return lout, rout, bidx, pidx


# ---------------------------------------------------------------- deferred row counts
class _WordSlots(threading.local):
    """The free word slots (tq_words_post) of this thread, per device."""

    def __init__(self):
        self.free = {}


_SLOTS = _WordSlots()


def post_words(ptrs, sizes, kernel_bytes=None):
    """Queue a read of small device words behind the work already on the current stream
    (tq_words_post); returns the ticket that `wait_words` collects."""
    dev = torch.cuda.current_device()
    free = _SLOTS.free.get(dev)
    if free is None:
        free = _SLOTS.free[dev] = list(range(L.load().tq_word_slots() - 1, -1, -1))
    if not free:
        raise RuntimeError("more word reads in flight on this thread than it has slots "
                           f"({L.load().tq_word_slots()})")
    slot = free.pop()
    n = len(ptrs)
    try:
        L.call("tq_words_post", n, (ctypes.c_void_p * n)(*ptrs), (ctypes.c_size_t * n)(*sizes),
               slot, L.stream_handle(),
               kernel={"k_collect_words": kernel_bytes if kernel_bytes is not None
                       else 2 * sum(sizes)})
    except BaseException:
        free.append(slot)
        raise
    return dev, slot


def wait_words(ticket):
    """The words of a `post_words` ticket (a ctypes array of 64 int64), once they are in."""
    dev, slot = ticket
    host = (ctypes.c_int64 * 64)()
    try:
        if dev == torch.cuda.current_device():
            L.call("tq_words_wait", slot, host)
        else:
            with torch.cuda.device(dev):
                L.call("tq_words_wait", slot, host)
    finally:
        _SLOTS.free[dev].append(slot)
    return host


class DeferredJoinFailed(Exception):
    """A deferred join's build keys were not unique (or left the bounds its build was given):
    its pairs are void and the operators run again without deferral."""


class Deferred:
    """A unique-key join whose pairs feed a group-by (and a top-k of its groups), run with one
    synchronisation.

    While a Deferred is the thread's active one (`set_deferred`), JoinIndex.probe writes its
    pairs into outputs sized for every probe row and leaves the pair count m on the device; a
    partitioned group-by whose columns are those outputs reads m there and leaves its group
    count G on the device; a top-k of those groups reads G there (`sort_perm`).  `resolve`
    then reads m, G and the build's uniqueness check back together; a consumer that needs m on
    the host sooner calls `resolve_join`, which reads m and the check."""

    def __init__(self):
        self.join = None
        self.group = None
        self.sorted = None
        self.m = None
        self.G = None

    def _read(self, with_group):
        j = self.join
        ptrs, sizes = [j["total"]], [8]
        if j["check"]:
            ptrs.append(j["check"])
            sizes.append(24)
        if with_group:
            ptrs += [self.group["ngroups"], self.group["err"]]
            sizes += [8, 8]
        host = (ctypes.c_int64 * 64)()
        L.call("tq_read_words", len(ptrs), (ctypes.c_void_p * len(ptrs))(*ptrs),
               (ctypes.c_size_t * len(ptrs))(*sizes), host, L.stream_handle(),
               kernel={"k_collect_words": 2 * sum(sizes)})
        m = int(host[0])
        ok = ctypes.c_int32(1)
        if j["check"]:
            L.call("tq_join_settle", ctypes.byref(j["index"].state),
                   (ctypes.c_int64 * 3)(host[1], host[2], host[3]), ctypes.byref(ok))
        if not ok.value or m <= 0:
            # repeated or out-of-bounds build keys void the pairs; no pairs at all take the
            # plain path, which reproduces the reference's empty-side results
            raise DeferredJoinFailed()
        reads, per_pair = j["write_bytes"]
        L.trace_set_bytes(j["trace"], "k_join_write", reads + m * per_pair)
        j["index"].deferred = False
        self.m = m
        k = 4 if j["check"] else 1
        return host, k

    def resolve_join(self):
        """Settle the pairs alone (for a consumer that needs m on the host)."""
        if self.m is None:
            self._read(False)
        return self.m

    def resolve(self):
        """Read everything back (once); returns (m, G), G None when no group-by was
        deferred."""
        if self.group is None:
            return self.resolve_join(), None
        if self.G is not None:
            return self.m, self.G
        host, k = self._read(True)
        g, err = int(host[k]), int(host[k + 1])
        if err:
            raise DeferredJoinFailed()   # the fused reduce gave up: run the plain path
        self.G = g
        for kern, nbytes in self.group["bytes"](self.m, g).items():
            L.trace_set_bytes(self.group["trace"], kern, nbytes)
        self.sorted_bytes(g)
        return self.m, g

    def sorted_bytes(self, g):
        """A deferred top-k was traced with the capacity as its row count: state the bytes of
        the g rows it read."""
        if self.sorted is not None and self.sorted.get("trace") is not None:
            L.trace_set_bytes(self.sorted["trace"], "k_topk_block", g * self.sorted["kb"])


_TLS = threading.local()


def set_deferred(df):
    """Make `df` (a Deferred, or None) the calling thread's active deferral."""
    _TLS.df = df


def _deferred():
    return getattr(_TLS, "df", None)


# ---------------------------------------------------------------- group-by
class GroupBy:
    """Group rows by integer key columns and reduce aggregate arguments.

    `aggs` is a list of (op, tree) where op is one of 'sum', 'avg', 'count', 'min', 'max' and
    tree the argument's row expression (aggregate.py dialect) over `tensors`."""

    OPS = {'sum': L.TQ_AGG_SUM, 'count': L.TQ_AGG_COUNT, 'avg': L.TQ_AGG_AVG,
           'min': L.TQ_AGG_MIN, 'max': L.TQ_AGG_MAX}

    def __init__(self, keys, aggs, tensors, mask=None, n=None):
        self.keys = []
        # keys that do not pack into an integer range (2-D string rows, floats) are grouped by
        # their dense rank in sort order: position -> (the key column, first row of each rank)
        self.ranked = {}
        for k in keys:
            L.require_gpu(k, "group key")
            if k.dim() == 2 and k.shape[1] == 1:
                k = k.reshape(-1)
            if k.dim() == 2 or k.dtype.is_floating_point:
                df = _deferred()
                if df is not None and df.join is not None and df.m is None:
                    n = df.resolve_join()   # ranks are taken over the real rows
                rows = k.shape[0] if n is None else n
                rank, first = dense_ranks(k, rows)
                self.ranked[len(self.keys)] = (k, first)
                k = rank
            self.keys.append(k.contiguous())
        if len(self.keys) > L.TQ_GB_MAX_KEYS:
            raise NotImplementedError("more than 8 group keys")
        self.mask = mask

        def compile_fn(b):
            for k in self.keys:          # key columns take load slots 0..nkeys-1
                b.load_slot(k)
            dtypes = []
            for op, tree in aggs:
                v = E.compile_row_expr(b, tree, tensors)
                if v.kind == "lit":
                    v = b.const(v.value, E.literal_dtype(v.value))
                if v.kind != "reg":
                    raise NotImplementedError("string-valued aggregate argument")
                dtypes.append(v.dtype)
                b.end_segment(v)
            return dtypes

        # the program depends on which columns are the keys, the aggregate trees and dtypes
        ids = {id(v.tensor): cid for cid, v in tensors.items()}
        key_cids = tuple(ids.get(id(k)) for k in self.keys)
        b = E.cached_program(("groupby", key_cids, repr(list(aggs))), tensors, compile_fn)
        self.arg_dtypes = list(b.extra)
        self.builder = b
        rows = n
        if rows is None:
            rows = self.keys[0].shape[0] if self.keys else (b.n if b.n is not None else 0)
        self.n = rows
        if not b.col_tensors:
            # no column read (e.g. count(*) without keys): the column set is never read
            self._dummy = memory.zeros(max(rows, 1), dtype=torch.int8,
                                       device=mask.device if mask is not None else 'cuda')
            b.col_tensors = [self._dummy]
            b.cols.ptr[0] = self._dummy.data_ptr()
            b.cols.width[0] = 1
        self.ops = [self.OPS[op] for op, _ in aggs]

    # a dense LDS table of at most this many bytes (with the staging rows) runs two workgroups
    # per CU; above it the group-by takes a global or hash table
    DENSE_LDS_BYTES = 64 * 1024
    DICT_LDS_BYTES = 80 * 1024

    def _dictionaries_pay(self, d, nsel, nload):
        """Key dictionaries are worth a presence pass and a read-back when the raw key space is
        too large for a dense LDS table, every key's value range is narrow enough for a bitmap,
        and there are enough rows to amortise the pass."""
        if not self.keys or nsel < (1 << 20):
            return False
        ranges = [d.key_range[i] for i in range(len(self.keys))]
        if any(r > L.TQ_GB_DICT_MAX_RANGE for r in ranges) or \
                sum((r + 31) // 32 for r in ranges) > 16384:
            return False
        space = 1
        for r in ranges:
            space *= r
        per = 4 + 8 * len(self.ops)
        return space * per + max(nload, 1) * 8192 > self.DENSE_LDS_BYTES

    def _adopt_dictionaries(self, d, mptr, s, dev, nload):
        """Build the keys' dictionaries and adopt them when the shrunk key space fits a dense
        LDS table."""
        wsb = L.size_call("tq_groupby_dict_ws_bytes", ctypes.byref(d))
        ws = L.workspace(wsb, dev, "gb_dict")
        out = (ctypes.c_void_p * L.TQ_GB_MAX_KEYS)()
        nd = (ctypes.c_int64 * L.TQ_GB_MAX_KEYS)()
        L.call("tq_groupby_dict", ctypes.byref(d), mptr, self.n, L.ptr(ws), wsb, out, nd, s)
        k = len(self.keys)
        space = 1
        for i in range(k):
            space *= max(int(nd[i]), 1)
        if space * (4 + 8 * len(self.ops)) + max(nload, 1) * 8192 > self.DICT_LDS_BYTES:
            return
        for i in range(k):
            d.key_dict[i] = out[i]
            d.key_dict_range[i] = d.key_range[i]
            d.key_range[i] = max(int(nd[i]), 1)
        self._dict_ws = ws

    def _reads_pairs(self, df):
        """Whether every column this group-by reads is an output of the deferred join (sized
        for the join's capacity, unmasked)."""
        if self.mask is not None or self.n != df.join["cap"]:
            return False
        outs = {t.data_ptr() for t in df.join["outs"]}
        cols = list(self.keys) + [t for t in self.builder.col_tensors if t is not getattr(self, "_dummy", None)]
        return bool(cols) and all(t.data_ptr() in outs for t in cols)

    def run(self):
        """Returns (key columns, aggregate columns), groups in ascending key order."""
        b = self.builder
        dev = b.col_tensors[0].device
        d = L.GroupByDesc()
        d.nkeys = len(self.keys)
        d.naggs = len(self.ops)
        d.nrows = self.n
        for i, k in enumerate(self.keys):
            d.key_dtype[i] = L.dtype_code(k.dtype)
            d.key_ptr[i] = k.data_ptr()
        for j, (op, dt) in enumerate(zip(self.ops, self.arg_dtypes)):
            d.agg_op[j] = op
            d.agg_dtype[j] = L.dtype_code(dt)
        s = L.stream_handle()
        mptr = L.ptr(self.mask) if self.mask is not None else None
        df = _deferred()
        pend = df is not None and df.join is not None and df.m is None
        if pend and not self._reads_pairs(df):
            if self.n == df.join["cap"]:
                raise DeferredJoinFailed()   # reads the pairs alongside other columns
            pend = False
        known = [known_bounds(k) for k in self.keys] if self.mask is None else [None]
        if pend and not (self.keys and all(x is not None for x in known)):
            self.n = df.resolve_join()    # the bounds pass needs the real row count
            return self.run()
        if self.keys and all(b is not None for b in known):
            # every row is selected and the key ranges are already known: no bounds pass
            hb = [v for b in known for v in b] + [self.n]
        else:
            bounds = memory.empty(2 * len(self.keys) + 1, dtype=torch.int64, device=dev)
            bws = L.workspace(L.size_call("tq_groupby_bounds_ws_bytes", self.n), dev, "gb_bounds")
            L.call("tq_groupby_bounds", ctypes.byref(d), mptr, self.n, L.ptr(bounds), L.ptr(bws), s,
                   kernel={"k_gb_bounds": self.n * (sum(k.element_size() for k in self.keys)
                                                    + (1 if self.mask is not None else 0))})
            hb = bounds.cpu().tolist()
        nsel = hb[-1]
        self.nsel = nsel
        for i in range(len(self.keys)):
            lo, hi = hb[2 * i], hb[2 * i + 1]
            d.key_min[i] = lo if nsel > 0 else 0
            d.key_range[i] = (hi - lo + 1) if nsel > 0 else 1
        if not pend and self._dictionaries_pay(d, nsel, b.nload):
            self._adopt_dictionaries(d, mptr, s, dev, b.nload)
        mode = ctypes.c_int32()
        slots = ctypes.c_int64()
        wsb = ctypes.c_size_t()
        L.call("tq_groupby_plan", ctypes.byref(d), nsel, b.nload, ctypes.byref(mode),
               ctypes.byref(slots), ctypes.byref(wsb))
        self.mode, self.slots = mode.value, slots.value
        if pend and (self.mode & 7) != L.TQ_GB_DENSE_PART:
            # only the partitioned mode reads its row count on the device: settle the pairs
            # and plan again for the real count
            self.n = df.resolve_join()
            return self.run()
        ws = L.workspace(wsb.value, dev, "groupby")
        # every selected row's loaded columns (keys included) and the mask byte of every row
        alg = nsel * b.load_bytes() + (self.n if self.mask is not None else 0)
        out_dtypes = []
        for op, dt in zip(self.ops, self.arg_dtypes):
            if op == L.TQ_AGG_COUNT:
                odt = torch.int64
            elif op == L.TQ_AGG_AVG:
                odt = torch.float32 if dt == torch.float32 else torch.float64
            elif op == L.TQ_AGG_SUM:
                odt = dt if dt.is_floating_point else torch.int64
            else:
                odt = dt
            out_dtypes.append(odt)
        kbytes = sum(k.element_size() for k in self.keys)
        rec_bytes = 8 * (1 + len(self.ops))
        out_bytes = kbytes + sum(dt.itemsize for dt in out_dtypes)

        def gb_bytes(rows, groups):
            # the accumulate pass reads every selected row's loaded columns; the partitioned
            # mode splits that into a histogram of the keys, a scatter of one record per row and
            # the partition reduce that reads the records and writes the groups
            return {"k_gb_accum": rows * b.load_bytes() + (self.n if self.mask is not None else 0),
                    "k_gbp_hist": rows * kbytes + (self.n if self.mask is not None else 0),
                    "k_gbp_scatter": rows * (b.load_bytes() + rec_bytes),
                    "k_gbp_reduce": rows * rec_bytes + groups * out_bytes,
                    "k_gb_write_dense": groups * out_bytes}
        # outputs sized for the most groups there can be; the groups written are a prefix
        cap = max(1, min(nsel, self.slots))
        key_full = [memory.empty(cap, dtype=k.dtype, device=dev) for k in self.keys]
        agg_full = [memory.empty(cap, dtype=odt, device=dev) for odt in out_dtypes]
        if pend:
            ngd, errd = ctypes.c_void_p(), ctypes.c_void_p()
            at = L.trace_len()
            L.call("tq_groupby_run_deferred", ctypes.byref(d), E.program_ptr(b), E.colset_ptr(b),
                   None, self.n, ctypes.c_void_p(df.join["total"]), mode, slots, L.ptr(ws),
                   wsb.value, L.ptr_array(key_full), L.ptr_array(agg_full), cap,
                   ctypes.byref(ngd), ctypes.byref(errd), s, kernel=gb_bytes(0, 0))
            df.group = {"ngroups": ngd.value, "err": errd.value, "outs": key_full + agg_full,
                        "cap": cap, "trace": at, "bytes": gb_bytes, "ws": ws}
            return key_full, agg_full
        ng = ctypes.c_int64()
        at = L.trace_len()
        L.call("tq_groupby_run", ctypes.byref(d), E.program_ptr(b), E.colset_ptr(b), mptr, self.n,
               mode, slots, L.ptr(ws), wsb.value, L.ptr_array(key_full), L.ptr_array(agg_full), cap,
               ctypes.byref(ng), s, kernel=gb_bytes(nsel, 0))
        G = int(ng.value)
        for kern, nbytes in gb_bytes(nsel, G).items():
            L.trace_set_bytes(at, kern, nbytes)
        key_out = self._key_values([o[:G] for o in key_full])
        agg_out = [o[:G] for o in agg_full]
        return key_out, agg_out

    def _key_values(self, key_out):
        """Ranked keys' groups back to their values: each group's rank picks its first row."""
        for pos, (col, first) in self.ranked.items():
            rows = gather(key_out[pos], [first])[0]
            key_out[pos] = gather(rows, [col])[0]
        return key_out


def dense_ranks(t: torch.Tensor, n: int):
    """Dense ranks of the first n rows of a key column in ascending sort order (string rows
    byte-wise, floats as torch.sort orders them): (int32 rank per row, int64 first row of each
    rank).  Reads the number of ranks back (one synchronisation) and records it as the rank
    column's bounds."""
    t = t[:n].contiguous()
    dev = t.device
    perm = sort_perm([(t, False)], n, device=dev)
    rank = memory.empty(max(n, 1), dtype=torch.int32, device=dev)
    first = memory.empty(max(n, 1), dtype=torch.int64, device=dev)
    nr = memory.empty(1, dtype=torch.int64, device=dev)
    key = (L.SortKey * 1)()
    key[0].ptr = t.data_ptr()
    key[0].dtype = L.dtype_code(t.dtype)
    key[0].width = t.shape[1] if t.dim() == 2 else 1
    wsb = L.size_call("tq_rank_ws_bytes", n)
    ws = L.workspace(wsb, dev, "rank")
    L.call("tq_rank_sorted", key, 1, n, L.ptr(perm), L.ptr(rank), L.ptr(first), L.ptr(nr),
           L.ptr(ws), wsb, L.stream_handle())
    d = int(nr.item())
    rank = rank[:n]
    note_bounds(rank, 0, max(d - 1, 0))
    return rank, first


def mark_rows(idx: torch.Tensor, n: int, device):
    """bool[n] with True at every row listed in idx."""
    flags = memory.empty(n, dtype=torch.bool, device=device)
    if n > 0:
        L.call("tq_mark_rows", L.ptr(idx) if idx.numel() else None, idx.numel(), L.ptr(flags), n,
               L.stream_handle())
    return flags


def mask_and_not(keep, drop, n: int, device):
    """keep & ~drop over bool columns (keep may be None = all rows)."""
    b = E.ProgramBuilder()
    d = b.load_column(drop)
    b.emit("NOT", dst=d.reg, a=d.reg)
    if keep is not None:
        k = b.load_column(keep)
        b.emit("AND", dst=d.reg, a=d.reg, b=k.reg)
    b.end_segment(d)
    b.finish()
    return eval_program(b, n, torch.bool, device)


def mask_and(a: torch.Tensor, b: torch.Tensor):
    """a & b over bool columns."""
    pb = E.ProgramBuilder()
    x = pb.load_column(a)
    y = pb.load_column(b)
    pb.emit("AND", dst=x.reg, a=x.reg, b=y.reg)
    pb.end_segment(x)
    pb.finish()
    return eval_program(pb, a.shape[0], torch.bool, a.device)


def compact_rows(mask: torch.Tensor):
    """int64 row numbers of the set rows of a mask, ascending (torch.nonzero of a 1-D mask)."""
    n = mask.shape[0]
    offs, total = scan_tiles(mask_tile_counts(mask), ntiles(n))
    m = int(total.item()) if n > 0 else 0
    out = memory.empty(m, dtype=torch.int64, device=mask.device)
    if m > 0:
        src = (ctypes.c_void_p * 1)()
        src[0] = None
        L.call("tq_compact", L.ptr(mask), n, L.ptr(offs), 1, src, L.int_array([8]),
               L.ptr_array([out]), L.stream_handle())
    return out


def count_true(mask: torch.Tensor) -> int:
    counts = mask_tile_counts(mask)
    _, total = scan_tiles(counts, ntiles(mask.shape[0]))
    return int(total.item()) if mask.shape[0] > 0 else 0
