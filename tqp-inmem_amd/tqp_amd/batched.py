"""Batched form: many filter -> join -> group-by pipelines in one persistent launch.

`BatchedEngine` takes pipelines in the reference's plan-argument form (a Filter on the build
table, a Filter on the probe table, an inner single-key Join, an Aggregate grouped by the build
join key with aggregates over probe columns -- the shape of TPC-H Q3's core and of the config-1
workload), compiles their predicates and aggregate arguments into register programs, lays out
one group table per pipeline (addressed by build key - min key; it is the join directory and the
aggregation table at once), and runs all of them with `tq_batch_run`: a persistent kernel that
drains a device queue of (pipeline, tile) items.  A join pipeline whose table does not fit LDS
aggregates through records (one per pair, grouped by key slice per probe chunk) and a reduce
per key slice, as the fused groupjoin does, rather than with global atomics.  `run()` issues only kernel launches, so it can
be captured into a graph and replayed; `results()` synchronises once and slices the outputs.

Outputs per pipeline are what `tqp_hash_aggregate` leaves in the tensor group: the group key
column [G, 1] in the build key's dtype and one column per aggregate, groups ascending by key.
Pipelines the batch cannot run (sparse build keys, duplicate selected build keys) are reported
by `results()` so the caller can run them through the per-operator path.
"""
import ctypes
import json
import math
import os
import time

import torch

from . import _lib as L
from . import expr as E
from . import memory
from .conversion import get_id, right_most
from .operators.aggregate import _is_key, _strip_alias, _unwrap

PHASES = 6
CHUNK = 32768             # probe rows per item of a records pipeline
MAX_SLICES = 2048
REDUCE_LDS = 32 * 1024    # a records pipeline's slice table: 2^sb * (4 + 8 * naggs) bytes
LDS_EXTRA_MAX = 48 * 1024


class BatchPipe(ctypes.Structure):
    _fields_ = [("bprog", L.Program), ("bcols", L.ColSet),
                ("pprog", L.Program), ("pcols", L.ColSet),
                ("aprog", L.Program), ("acols", L.ColSet),
                ("nb", ctypes.c_int64), ("np", ctypes.c_int64),
                ("kmin", ctypes.c_int64), ("range", ctypes.c_int64),
                ("naggs", ctypes.c_int32), ("key_dtype", ctypes.c_int32),
                ("agg_op", ctypes.c_int32 * L.TQ_MAX_SEGS),
                ("agg_dtype", ctypes.c_int32 * L.TQ_MAX_SEGS),
                ("pres", ctypes.c_void_p), ("cnt", ctypes.c_void_p), ("acc", ctypes.c_void_p),
                ("tile_counts", ctypes.c_void_p), ("key_out", ctypes.c_void_p),
                ("agg_out", ctypes.c_void_p * L.TQ_MAX_SEGS),
                ("ngroups", ctypes.c_void_p), ("flags", ctypes.c_void_p),
                ("ntiles", ctypes.c_int32 * PHASES), ("mode", ctypes.c_int32),
                ("nkeys", ctypes.c_int32), ("lds_table", ctypes.c_int32),
                ("gkey_min", ctypes.c_int64 * L.TQ_GB_MAX_KEYS),
                ("gkey_stride", ctypes.c_int64 * L.TQ_GB_MAX_KEYS),
                ("gkey_range", ctypes.c_int64 * L.TQ_GB_MAX_KEYS),
                ("gkey_dtype", ctypes.c_int32 * L.TQ_GB_MAX_KEYS),
                ("gkey_out", ctypes.c_void_p * L.TQ_GB_MAX_KEYS),
                ("records", ctypes.c_int32), ("sb", ctypes.c_int32),
                ("nslices", ctypes.c_int32), ("pad_", ctypes.c_int32),
                ("rec", ctypes.c_void_p), ("loff", ctypes.c_void_p),
                ("keycount", ctypes.c_void_p), ("fast", ctypes.c_int32 * 2),
                ("prep", (ctypes.c_uint64 * 32) * 2)]

LDS_TABLE_MAX = 32 * 1024


class PipelineSpec:
    """One batched join pipeline: build/probe columns (id -> tensor), predicates, join keys,
    aggregates [(output id, op, argument tree)] and the output id of the group key."""
    mode = 0

    def __init__(self, build_cols, probe_cols, build_key, probe_key, build_pred, probe_pred,
                 aggs, key_out_id):
        self.build_cols, self.probe_cols = build_cols, probe_cols
        self.build_key, self.probe_key = build_key, probe_key
        self.build_pred, self.probe_pred = build_pred, probe_pred
        self.aggs = aggs
        self.key_out_id = key_out_id


class GroupSpec:
    """One batched group-by pipeline (no join): columns, predicate, key column ids (output ids
    equal to the key ids, as the reference's key results are written back), aggregates."""
    mode = 1

    def __init__(self, cols, pred, keys, aggs, key_out_ids):
        self.probe_cols, self.probe_pred = cols, pred
        self.keys, self.aggs, self.key_out_ids = keys, aggs, key_out_ids


class _Cols(dict):
    """{id: object with .tensor} view over plain tensors for the expression compiler."""

    class _V:
        __slots__ = ("tensor",)

        def __init__(self, t):
            self.tensor = t

    def __init__(self, tensors):
        super().__init__({k: self._V(v) for k, v in tensors.items()})


def match_plan(plan, tensor_group):
    """A PipelineSpec for a (Filter build, Filter probe, Join, Aggregate) plan, or None."""
    ops = [a for _, a in plan]
    names = [a['name'] for a in ops]
    if names[:4] != ['Filter', 'Filter', 'Join', 'Aggregate']:
        return None
    fb, fp, j, ag = ops[:4]
    if j.get('join type') != 'Inner' or len(j['left keys']) != 1 or 'join condition' in j:
        return None
    bkey, pkey = j['left keys'][0], j['right keys'][0]
    masks = j.get('mask map', {})
    lg, rg = set(j['left group']), set(j['right group'])
    if fb['out mask id'] not in masks.values() or fp['out mask id'] not in masks.values():
        return None
    if any(masks.get(c) != fb['out mask id'] for c in j['materialized in'] if c in lg and c in masks):
        return None
    if set(fb['materialized in']) - lg or set(fp['materialized in']) - rg:
        return None
    if [get_id(k) for k in ag['keys']] != [bkey]:
        return None
    aggs, key_out = [], None
    for tree in ag['results']:
        tree = _unwrap(tree)
        if _is_key(tree):
            key_out = get_id(right_most(tree))
            continue
        body = _strip_alias(tree[0])
        if not (isinstance(body, list) and len(body) == 2 and body[0] in E.AGG_FUNCS):
            return None
        found = set()
        _collect_ids(body[1], found)
        if found - rg:
            return None
        aggs.append((get_id(right_most(tree)), body[0], body[1]))
    bcols = {c: tensor_group[c].tensor for c in lg if c in tensor_group}
    pcols = {c: tensor_group[c].tensor for c in rg if c in tensor_group}
    return PipelineSpec(bcols, pcols, bkey, pkey, fb['filter tree'], fp['filter tree'], aggs, key_out)


def match_groupby_plan(plan, tensor_group):
    """A GroupSpec for a (Filter, Aggregate over the filtered columns[, Sort by the group keys
    ascending]) plan, or None."""
    ops = [a for _, a in plan]
    names = [a['name'] for a in ops]
    if names[:2] != ['Filter', 'Aggregate'] or len(ops) > 3:
        return None
    f, ag = ops[0], ops[1]
    mid = f['out mask id']
    mm = ag.get('mask map', {})
    if set(mm.values()) != {mid} or set(mm) != set(ag['input ids']) - {mid}:
        return None
    keys = [get_id(k) for k in ag['keys']]
    key_out, aggs = [], []
    for tree in ag['results']:
        tree = _unwrap(tree)
        if _is_key(tree):
            key_out.append((get_id(tree if isinstance(tree, str) else tree[0]), get_id(right_most(tree))))
            continue
        body = _strip_alias(tree[0])
        if not (isinstance(body, list) and len(body) == 2 and body[0] in E.AGG_FUNCS):
            return None
        aggs.append((get_id(right_most(tree)), body[0], body[1]))
    if [k for k, _ in key_out] != keys:
        return None
    if len(ops) == 3:
        so = ops[2]
        if so['name'] != 'Sort' or so['key ids'] != [o for _, o in key_out] or any(so['key orders']):
            return None
    cols = {c: tensor_group[c].tensor for c in set(ag['input ids']) | set(f['materialized in'])
            if c in tensor_group and c != mid}
    for k in keys:
        if cols[k].dtype.is_floating_point:
            return None
    return GroupSpec(cols, f['filter tree'], keys, aggs, [o for _, o in key_out])


def _collect_ids(tree, out):
    if isinstance(tree, str):
        if '#' in tree and tree.split('#')[-1].replace('L', '').isdigit():
            out.add(get_id(tree))
    elif isinstance(tree, list):
        for e in tree:
            _collect_ids(e, out)


def _key_bounds(t: torch.Tensor):
    """(min, max) of an integer column on the device (one bounds launch)."""
    d = L.GroupByDesc()
    d.nkeys = 1
    d.key_dtype[0] = L.dtype_code(t.dtype)
    d.key_ptr[0] = t.data_ptr()
    out = memory.empty(3, dtype=torch.int64, device=t.device)
    ws = L.workspace(L.size_call("tq_groupby_bounds_ws_bytes", t.shape[0]), t.device, "gb_bounds")
    L.call("tq_groupby_bounds", ctypes.byref(d), None, t.shape[0], L.ptr(out), L.ptr(ws),
           L.stream_handle())
    lo, hi, _ = out.cpu().tolist()
    return lo, hi


class BatchedEngine:
    """Prepare once, then run() (graph-capturable) and results()."""

    def __init__(self, specs, tile_rows=4096, grid=None, wait_us=100_000, key_bounds=None,
                 records=False, fast=True, lds_table_max=None, reduce_lds=None,
                 fused_min_rows=None, fused_streams=3):
        self.specs = list(specs)
        # the fused calls are spread round-robin over this many side streams (default
        # priority), forked from and joined back into the launching stream
        self.fused_streams = max(int(fused_streams), 1)
        self._sides = None
        # join pipelines with at least this many probe rows (and conjunctive filters) run as
        # fused groupjoin launches recorded by run() instead of items of the persistent kernel
        self.fused_min_rows = fused_min_rows
        self.records = records    # True: large join tables accumulate through records
        # LDS capacities: a pipeline whose whole table (4 + 8 * aggregates bytes a key) fits
        # lds_table_max is accumulated in LDS; a records pipeline's slice table fits reduce_lds
        self.lds_table_max = int(lds_table_max or LDS_TABLE_MAX)
        self.reduce_lds = int(reduce_lds or REDUCE_LDS)
        self.fast = fast          # False: every predicate runs on the per-row interpreter
        self.tile_rows = int(tile_rows)
        self.device = torch.device("cuda:0")
        props = torch.cuda.get_device_properties(self.device)
        self.grid = grid or props.multi_processor_count * 4
        self.wait_us = wait_us
        self.key_bounds = key_bounds or {}
        self.rejected = []        # indices of specs the batch cannot run
        self._prepare()

    # ---------------------------------------------------------------- set-up
    def _bounds(self, t):
        key = (t.data_ptr(), t.shape[0], t.dtype)
        kb = self.key_bounds.get(key)
        if kb is None:
            kb = _key_bounds(t) if t.shape[0] > 0 else (0, -1)
            self.key_bounds[key] = kb
        return kb

    def _plan_join(self, i, sp):
        bk = sp.build_cols[sp.build_key].reshape(-1)
        pk = sp.probe_cols[sp.probe_key].reshape(-1)
        if bk.dtype.is_floating_point or pk.dtype.is_floating_point:
            return None
        nb, np_ = bk.shape[0], pk.shape[0]
        lo, hi = self._bounds(bk)
        rng = max(hi - lo + 1, 1)
        if rng > 4 * nb + 65536:
            return None
        bb = E.ProgramBuilder()
        bb.load_slot(bk)
        E.compile_predicate(bb, sp.build_pred, _Cols(sp.build_cols))
        pb = E.ProgramBuilder()
        pb.load_slot(pk)
        E.compile_predicate(pb, sp.probe_pred, _Cols(sp.probe_cols))
        return dict(i=i, bb=bb, pb=pb, lo=lo, rng=rng, cap=min(nb, rng), nb=nb, np=np_,
                    key_dtype=bk.dtype, gkeys=None)

    def _conjunction(self, tree, cols):
        """A predicate program over `cols` that the groupjoin kernels evaluate in their own
        loads (a conjunction of column-versus-constant comparisons), else None."""
        b = E.ProgramBuilder()
        E.compile_predicate(b, tree, _Cols(cols))
        b.finish()
        if not b.col_tensors or not L.load().tq_filter_is_conjunction(E.program_ptr(b),
                                                                      E.colset_ptr(b)):
            return None
        return b

    def _plan_fused(self, i, sp):
        """Set up join pipeline i as one fused groupjoin call (tq_groupjoin_run: build, probe
        and reduce launches) when it is large enough and its parts fit the fused kernels; False
        leaves it to the persistent kernel."""
        bk = sp.build_cols[sp.build_key].reshape(-1)
        pk = sp.probe_cols[sp.probe_key].reshape(-1)
        nb, np_ = bk.shape[0], pk.shape[0]
        if np_ < self.fused_min_rows or nb == 0 or not bk.is_contiguous() or \
                not pk.is_contiguous():
            return False
        if bk.dtype.is_floating_point or pk.dtype.is_floating_point or \
                torch.bool in (bk.dtype, pk.dtype):
            return False
        lo, hi = self._bounds(bk)
        rng = hi - lo + 1
        if rng < 1 or rng > 4 * nb + 65536:
            return False
        bpred = self._conjunction(sp.build_pred, sp.build_cols)
        ppred = self._conjunction(sp.probe_pred, sp.probe_cols)
        if bpred is None or ppred is None or bpred.n != nb or ppred.n != np_:
            return False
        ab = E.ProgramBuilder()
        arg_dtypes = []
        for _, op, tree in sp.aggs:
            v = E.compile_row_expr(ab, tree, _Cols(sp.probe_cols))
            if v.kind == "lit":
                v = ab.const(v.value, E.literal_dtype(v.value))
            if v.kind != "reg":
                return False
            arg_dtypes.append(v.dtype)
            ab.end_segment(v)
        ab.finish()
        if not ab.col_tensors:
            ab.column(pk)
        naggs = len(sp.aggs)
        wsb = L.size_call("tq_groupjoin_ws_bytes", nb, np_, rng, naggs, ab.nload)
        if wsb == 0:
            return False
        ops = {'sum': L.TQ_AGG_SUM, 'count': L.TQ_AGG_COUNT, 'avg': L.TQ_AGG_AVG,
               'min': L.TQ_AGG_MIN, 'max': L.TQ_AGG_MAX}
        dev = self.device
        cap = max(1, min(rng, np_))
        key_out = memory.empty(cap, dtype=bk.dtype, device=dev)
        agg_outs = []
        d = L.GroupJoinDesc()
        d.bkey, d.bkey_dtype, d.nb = bk.data_ptr(), L.dtype_code(bk.dtype), nb
        d.pkey, d.pkey_dtype, d.np = pk.data_ptr(), L.dtype_code(pk.dtype), np_
        d.key_lo, d.key_hi, d.naggs = lo, hi, naggs
        d.bpred, d.bpred_cols = ctypes.addressof(bpred.prog), ctypes.addressof(bpred.cols)
        d.ppred, d.ppred_cols = ctypes.addressof(ppred.prog), ctypes.addressof(ppred.cols)
        for a, ((_, op, _), dt) in enumerate(zip(sp.aggs, arg_dtypes)):
            d.agg_op[a] = ops[op]
            d.agg_dtype[a] = L.dtype_code(dt)
            if op == 'count':
                odt = torch.int64
            elif op == 'avg':
                odt = torch.float32 if dt == torch.float32 else torch.float64
            elif op == 'sum':
                odt = dt if dt.is_floating_point else torch.int64
            else:
                odt = dt
            t = memory.empty(cap, dtype=odt, device=dev)
            agg_outs.append(t)
            d.out_agg[a] = t.data_ptr()
        d.out_key, d.out_key_dtype, d.out_cap = key_out.data_ptr(), L.dtype_code(bk.dtype), cap
        d.phases = 0
        # every call's build stamps its keys with its own epoch in the shared presence bytes
        d.epoch = len(self.fused) % 255 + 1
        ws = memory.empty(wsb, dtype=torch.uint8, device=dev)
        self.fused.append(dict(i=i, d=d, keep=(bk, pk, bpred, ppred, ab), ws=ws,
                               keys_out=[key_out], agg_outs=agg_outs,
                               pres_bytes=L.size_call("tq_groupjoin_pres_bytes", rng),
                               words=None))
        return True

    def _plan_group(self, i, sp):
        cols = [sp.probe_cols[k].reshape(-1) for k in sp.keys]
        n = next(iter(sp.probe_cols.values())).shape[0]
        mins, ranges = [], []
        for c in cols:
            lo, hi = self._bounds(c)
            mins.append(lo)
            ranges.append(max(hi - lo + 1, 1))
        space = 1
        for r in ranges:
            space *= r
        if space > min(1 << 27, 4 * n + 65536):
            return None
        strides = []
        acc = 1
        for r in reversed(ranges):
            strides.append(acc)
            acc *= r
        strides.reverse()
        pb = E.ProgramBuilder()
        for c in cols:
            pb.load_slot(c)
        E.compile_predicate(pb, sp.probe_pred, _Cols(sp.probe_cols))
        bb = E.ProgramBuilder()
        bb.const(True, torch.bool)
        bb.end_segment(E.Value("reg", reg=0, dtype=torch.bool))
        bb.finish()
        bb.column(cols[0] if cols else next(iter(sp.probe_cols.values())))
        return dict(i=i, bb=bb, pb=pb, lo=0, rng=space, cap=max(1, min(n, space)), nb=0, np=n,
                    key_dtype=torch.int64, gkeys=(cols, mins, ranges, strides))

    def _prepare(self):
        dev = self.device
        T = self.tile_rows
        self.outputs = []
        self.builders = []        # keep compiled programs (and their tensors) alive
        stage_loads = 1
        lds_table = 0
        layout = []
        self.fused = []
        for i, sp in enumerate(self.specs):
            if self.fused_min_rows is not None and sp.mode == 0 and self._plan_fused(i, sp):
                continue
            lay = self._plan_join(i, sp) if sp.mode == 0 else self._plan_group(i, sp)
            if lay is None:
                self.rejected.append(i)
                continue
            ab = E.ProgramBuilder()
            arg_dtypes = []
            for _, op, tree in sp.aggs:
                v = E.compile_row_expr(ab, tree, _Cols(sp.probe_cols))
                if v.kind == "lit":
                    v = ab.const(v.value, E.literal_dtype(v.value))
                arg_dtypes.append(v.dtype)
                ab.end_segment(v)
            ab.finish()
            if not ab.col_tensors:
                ab.column(next(iter(sp.probe_cols.values())))
            lay.update(ab=ab, arg_dtypes=arg_dtypes)
            naggs = len(sp.aggs)
            tbytes = lay["rng"] * (4 + 8 * naggs)
            lay["lds"] = 1 if tbytes <= self.lds_table_max else 0
            lay["records"] = 0
            if lay["lds"]:
                lds_table = max(lds_table, tbytes)
            elif sp.mode == 0 and self.records:
                # records: the largest slice whose LDS table fits, if the slices are few enough
                sb = 12
                while sb > 6 and (1 << sb) * (4 + 8 * naggs) > self.reduce_lds:
                    sb -= 1
                nsl = _ceil(lay["rng"], 1 << sb)
                if (1 << sb) * (4 + 8 * naggs) <= self.reduce_lds and nsl <= MAX_SLICES:
                    lay.update(records=1, sb=sb, nslices=nsl)
                    lds_table = max(lds_table, (CHUNK // 32 + MAX_SLICES) * 4,
                                    (1 << sb) * (4 + 8 * naggs))
            stage_loads = max(stage_loads, lay["bb"].nload, lay["pb"].nload, ab.nload)
            layout.append(lay)
        budget = min(LDS_EXTRA_MAX, 64 * 1024 - stage_loads * 8 * 1024)
        if lds_table > budget:
            # the records scratch does not fit beside the stage: atomics for those pipelines,
            # and whole tables that do not fit beside it accumulate in global memory
            for lay in layout:
                lay["records"] = 0
                if lay["lds"] and lay["rng"] * (4 + 8 * len(self.specs[lay["i"]].aggs)) > budget:
                    lay["lds"] = 0
            lds_table = max([lay["rng"] * (4 + 8 * len(self.specs[lay["i"]].aggs))
                             for lay in layout if lay["lds"]] or [0])
        self.stage_loads = stage_loads
        self.lds_table_bytes = lds_table
        # one table per pipeline (cleared by the launch's first phase), and a records
        # pipeline's records and slice offsets
        self.tables = [memory.empty(max(_table_bytes(lay["rng"], lay["ab"].nseg, T)[-1], 256),
                                    dtype=torch.uint8, device=dev) for lay in layout]
        self.records = [memory.empty(max(_records_bytes(lay)[-1], 256), dtype=torch.uint8,
                                     device=dev) if lay["records"] else None for lay in layout]
        self.ngroups = memory.empty(max(len(layout), 1), dtype=torch.int64, device=dev)
        self.flags = memory.zeros(max(len(layout), 1), dtype=torch.int32, device=dev)
        # per join pipeline: selected build rows, distinct selected build keys
        self.keycount = memory.zeros(2 * max(len(layout), 1), dtype=torch.int64, device=dev)
        items = [[] for _ in range(PHASES)]
        arr = (BatchPipe * max(len(layout), 1))()
        ops = {'sum': L.TQ_AGG_SUM, 'count': L.TQ_AGG_COUNT, 'avg': L.TQ_AGG_AVG,
               'min': L.TQ_AGG_MIN, 'max': L.TQ_AGG_MAX}
        for k, (lay, table, recbuf) in enumerate(zip(layout, self.tables, self.records)):
            sp = self.specs[lay["i"]]
            d = arr[k]
            d.bprog, d.bcols = lay["bb"].prog, lay["bb"].cols
            d.pprog, d.pcols = lay["pb"].prog, lay["pb"].cols
            d.aprog, d.acols = lay["ab"].prog, lay["ab"].cols
            d.nb, d.np, d.kmin, d.range = lay["nb"], lay["np"], lay["lo"], lay["rng"]
            d.naggs = len(sp.aggs)
            d.mode = sp.mode
            d.lds_table = lay["lds"]
            d.key_dtype = L.dtype_code(lay["key_dtype"])
            cnt_off, acc_off, tc_off, _ = _table_bytes(lay["rng"], d.naggs, T)
            p = table.data_ptr()
            d.cnt, d.acc, d.tile_counts = p + cnt_off, p + acc_off, p + tc_off
            if sp.mode == 0:
                d.pres = p
                d.keycount = self.keycount.data_ptr() + 16 * k
            cap = lay["cap"]
            keys_out = []
            if sp.mode == 0:
                key_out = memory.empty(cap, dtype=lay["key_dtype"], device=dev)
                d.key_out = key_out.data_ptr()
                keys_out.append(key_out)
            else:
                cols, mins, ranges, strides = lay["gkeys"]
                d.nkeys = len(cols)
                for j, c in enumerate(cols):
                    t = memory.empty(cap, dtype=c.dtype, device=dev)
                    keys_out.append(t)
                    d.gkey_out[j] = t.data_ptr()
                    d.gkey_min[j], d.gkey_range[j], d.gkey_stride[j] = mins[j], ranges[j], strides[j]
                    d.gkey_dtype[j] = L.dtype_code(c.dtype)
            agg_outs = []
            for a, ((oid, op, _), dt) in enumerate(zip(sp.aggs, lay["arg_dtypes"])):
                d.agg_op[a] = ops[op]
                d.agg_dtype[a] = L.dtype_code(dt)
                if op == 'count':
                    odt = torch.int64
                elif op == 'avg':
                    odt = torch.float32 if dt == torch.float32 else torch.float64
                elif op == 'sum':
                    odt = dt if dt.is_floating_point else torch.int64
                else:
                    odt = dt
                t = memory.empty(cap, dtype=odt, device=dev)
                agg_outs.append(t)
                d.agg_out[a] = t.data_ptr()
            d.ngroups = self.ngroups.data_ptr() + 8 * k
            d.flags = self.flags.data_ptr() + 4 * k
            rng, nb, np_ = lay["rng"], lay["nb"], lay["np"]
            if lay["records"]:
                d.records, d.sb, d.nslices = 1, lay["sb"], lay["nslices"]
                loff_off, _ = _records_bytes(lay)
                d.rec, d.loff = recbuf.data_ptr(), recbuf.data_ptr() + loff_off
                nt = [_ceil(rng, T), _ceil(nb, T), _ceil(np_, CHUNK), lay["nslices"],
                      _ceil(rng, T), _ceil(rng, T)]
            else:
                nt = [_ceil(rng, T), _ceil(nb, T), _ceil(np_, T), 0, _ceil(rng, T),
                      _ceil(rng, T)]
            for ph in range(PHASES):
                d.ntiles[ph] = nt[ph]
                items[ph].append((nt[ph], k))
            self.builders.append((lay["bb"], lay["pb"], lay["ab"]))
            self.outputs.append((lay["i"], keys_out, agg_outs))
        flat = []
        starts = [0]
        for ph in range(PHASES):
            # larger pipelines first inside a phase keeps the queue's tail short
            for ntile, k in sorted(items[ph], key=lambda x: -x[0]):
                flat.extend((k, t) for t in range(ntile))
            starts.append(len(flat))
        # conjunctive build / probe predicates run over whole tiles (fills fast[] and prep[])
        self.fast_predicates = 0
        if self.fast:
            n = L.load().tq_batch_prepare(arr, len(layout))
            if n < 0:
                raise RuntimeError("tq_batch_prepare: " + L.load().tq_last_error().decode())
            self.fast_predicates = n
        self.phase_start = (ctypes.c_int64 * (PHASES + 1))(*starts)
        self.nitems = len(flat)
        it = torch.tensor(flat if flat else [(0, 0)], dtype=torch.int32).reshape(-1)
        self.items = memory.to_device(it, dev)
        self.desc = memory.to_device(torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8), dev)
        self._desc_host = arr
        self.ws = memory.empty(max(L.size_call("tq_batch_ws_bytes", len(layout)), 256),
                               dtype=torch.uint8, device=dev)
        self.npipes = len(layout)
        # presence bytes of the fused calls, one array per side stream: the calls of one stream
        # run one after another and stamp their keys with their own epochs (the arrays are
        # cleared by run() before the first and every 255th call, so no call finds another's
        # stamps under its own epoch); calls on different streams overlap in time, and keys of
        # one would overwrite the other's stamps in a shared array
        nb_pres = max([f["pres_bytes"] for f in self.fused] or [256])
        self.fused_pres = [memory.empty(nb_pres, dtype=torch.uint8, device=dev)
                           for _ in range(self.fused_streams)] if self.fused else []

    # ---------------------------------------------------------------- execution
    def run(self):
        """Launch the batch on the current stream (kernel launches only): the persistent
        kernel's pipelines, then each fused pipeline's groupjoin call."""
        s = L.stream_handle()
        main = torch.cuda.current_stream()
        if self.fused:
            if self._sides is None:
                self._sides = _side_streams(self.device, self.fused_streams)
            self._fork(main)
        if self.npipes:
            L.call("tq_batch_run", L.ptr(self.desc), self.npipes, L.ptr(self.items),
                   self.phase_start, self.tile_rows, self.grid, self.stage_loads,
                   self.lds_table_bytes, L.ptr(self.ws), ctypes.c_uint32(self.wait_us), s)
        for j, f in enumerate(self.fused):
            if j and j % 255 == 0:
                # epochs start over: every call so far is done before the bytes are cleared
                self._join(main)
                self._fork(main)
            ab = f["keep"][4]
            words = ctypes.c_void_p()
            side = j % len(self._sides)
            pres = self.fused_pres[side]
            with torch.cuda.stream(self._sides[side]):
                L.call("tq_groupjoin_run", ctypes.byref(f["d"]), E.program_ptr(ab),
                       E.colset_ptr(ab), L.ptr(pres), pres.numel(), L.ptr(f["ws"]),
                       f["ws"].numel(), ctypes.byref(words), L.stream_handle())
            f["words"] = words.value
        if self.fused:
            self._join(main)

    def _fork(self, main):
        """Clear the fused calls' presence bytes on `main`, then let the side streams start
        after it."""
        for pres in self.fused_pres:
            pres.zero_()
        ev = torch.cuda.Event()
        ev.record(main)
        for st in self._sides:
            st.wait_event(ev)

    def _join(self, main):
        """`main` continues after everything queued on the side streams."""
        for st in self._sides:
            ev = torch.cuda.Event()
            ev.record(st)
            main.wait_event(ev)

    def results(self, sync=True):
        """Returns ({spec index: {out id: tensor}}, [spec indices to rerun]).  With sync=False
        the caller has already waited for the launch (e.g. on an event of its stream)."""
        if sync:
            torch.cuda.synchronize()
        err = int(self.ws[256:260].view(torch.int32).item()) if self.npipes else 0
        if err:
            raise RuntimeError("batched launch aborted: a phase wait exceeded its time limit")
        ng = self.ngroups.cpu().tolist()
        fl = self.flags.cpu().tolist()
        kc = self.keycount.cpu().tolist()
        out, rerun = {}, list(self.rejected)
        for k, (i, keys_out, agg_outs) in enumerate(self.outputs):
            sp = self.specs[i]
            if fl[k] or (sp.mode == 0 and kc[2 * k] != kc[2 * k + 1]):
                # a build key outside the bounds, or a repeated one
                rerun.append(i)
                continue
            G = ng[k]
            ids = [sp.key_out_id] if sp.mode == 0 else sp.key_out_ids
            res = {oid: t[:G].reshape(-1, 1) for oid, t in zip(ids, keys_out)}
            for (oid, _, _), t in zip(sp.aggs, agg_outs):
                res[oid] = t[:G]
            out[i] = res
        if self.fused:
            # each fused call's words (groups, pairs, selected build rows, distinct selected
            # build keys, error bits) sit in its own workspace: one gather and one copy
            rows = []
            for f in self.fused:
                if f["words"] is None:
                    raise RuntimeError("batched results() before run()")
                off = f["words"] - f["ws"].data_ptr()
                if off < 0 or off % 8 or off + 40 > f["ws"].numel():
                    raise RuntimeError("groupjoin words outside the call's workspace")
                rows.append(f["ws"][off:off + 40].view(torch.int64))
            for f, (G, m, nsel, nset, err) in zip(self.fused, torch.stack(rows).cpu().tolist()):
                i = f["i"]
                if err or nsel != nset or m <= 0:
                    # a repeated or out-of-bounds build key, a look-back that gave up, or no
                    # pairs: the per-operator path gives the reference's result
                    rerun.append(i)
                    continue
                sp = self.specs[i]
                res = {sp.key_out_id: f["keys_out"][0][:G].reshape(-1, 1)}
                for (oid, _, _), t in zip(sp.aggs, f["agg_outs"]):
                    res[oid] = t[:G]
                out[i] = res
        return out, sorted(rerun)

    def reset_flags(self):
        self.flags.zero_()


_SIDES = {}


def _side_streams(device, n):
    """n side streams of `device` (default priority), made once and shared by every engine: a
    process holds few streams, as the device has few hardware queues to map them on."""
    key = (torch.device(device).index or 0, n)
    st = _SIDES.get(key)
    if st is None:
        st = _SIDES[key] = [torch.cuda.Stream(device=device) for _ in range(n)]
    return st


def _table_bytes(rng, naggs, tile):
    """Offsets of cnt, acc and tile_counts in one pipeline's table (the presence bytes first),
    and its total size."""
    cnt_off = _align(rng)
    acc_off = cnt_off + _align(rng * 4)
    tc_off = acc_off + _align(rng * 8 * max(naggs, 1))
    return cnt_off, acc_off, tc_off, tc_off + _align(_ceil(rng, tile) * 4)


def _records_bytes(lay):
    """Offset of the slice offsets in a records pipeline's buffer, and its total size: records
    of up to one per probe row (1 + naggs words each), then (nslices + 1) x chunks u32."""
    rw = 1 + lay["ab"].nseg
    loff_off = _align(max(lay["np"], 1) * rw * 8)
    return loff_off, loff_off + _align((lay["nslices"] + 1) * max(_ceil(lay["np"], CHUNK), 1) * 4)


def _ceil(a, b):
    return (a + b - 1) // b


def _align(n, a=256):
    return (n + a - 1) // a * a


# ==================================================================== autotuner
TUNE_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def autotune(specs, candidates=None, repeats=3, cache=TUNE_CACHE, key=None,
             lds_candidates=None, fused_candidates=None):
    """Pick tile rows, persistent grid size and the accumulation of large tables (global atomics
    or records with slice reduces) for a batch by timing candidates on it, then, for that
    choice, the LDS capacities (whole-table limit, records slice table) from `lds_candidates`,
    then the probe-row threshold from which join pipelines run as fused groupjoin calls
    (`fused_candidates`; None, no fused calls, is always a candidate).
    Results are cached in tuning.json under `key` (device name + workload signature).  The
    batched path has no radix pass (tables are addressed by key - min), so there are no radix
    bits to choose."""
    dev = torch.device("cuda:0")
    props = torch.cuda.get_device_properties(dev)
    key = key or f"{props.name}|{len(specs)}"
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
    if key in table:
        return table[key]
    cu = props.multi_processor_count
    candidates = candidates or [(t, cu * g, r) for t in (4096, 8192) for g in (2, 4, 8)
                                for r in (False, True)]
    best, timings = None, {}
    bounds = {}
    for cand in candidates:
        tile, grid = cand[0], cand[1]
        records = cand[2] if len(cand) > 2 else True
        eng = BatchedEngine(specs, tile_rows=tile, grid=grid, key_bounds=bounds, records=records)
        eng.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            eng.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / repeats
        timings[f"{tile}x{grid}{'r' if records else 'a'}"] = dt
        if best is None or dt < best[0]:
            best = (dt, tile, grid, records)
        del eng
    lds_candidates = lds_candidates or [(t, r) for t in (16 * 1024, 32 * 1024, 48 * 1024)
                                        for r in ((16 * 1024, 32 * 1024) if best[3] else
                                                  (REDUCE_LDS,))]
    best_lds = (best[0], LDS_TABLE_MAX, REDUCE_LDS)
    for tmax, rlds in lds_candidates:
        if (tmax, rlds) == (LDS_TABLE_MAX, REDUCE_LDS):
            continue          # timed above
        eng = BatchedEngine(specs, tile_rows=best[1], grid=best[2], key_bounds=bounds,
                            records=best[3], lds_table_max=tmax, reduce_lds=rlds)
        eng.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            eng.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / repeats
        timings[f"{best[1]}x{best[2]}{'r' if best[3] else 'a'} lds{tmax >> 10}k/{rlds >> 10}k"] = dt
        if dt < best_lds[0]:
            best_lds = (dt, tmax, rlds)
        del eng
    # then which join pipelines run as fused groupjoin calls (none, or those with at least
    # this many probe rows), with the persistent grid as tuned or halved (its workgroups hold
    # the CUs until its queue drains, so a smaller grid leaves room for the fused calls)
    best_fused = (best_lds[0], None, best[2])
    fms = fused_candidates if fused_candidates is not None else \
        (2_000_000, 1_000_000, 500_000, 250_000)
    for fm in fms:
        for grid in (best[2], max(best[2] // 2, 1)):
            eng = BatchedEngine(specs, tile_rows=best[1], grid=grid, key_bounds=bounds,
                                records=best[3], lds_table_max=best_lds[1],
                                reduce_lds=best_lds[2], fused_min_rows=fm)
            eng.run()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(repeats):
                eng.run()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / repeats
            timings[f"fused >= {fm} probe rows ({len(eng.fused)} pipelines), grid {grid}"] = dt
            if dt < best_fused[0]:
                best_fused = (dt, fm, grid)
            del eng
    choice = {"tile_rows": best[1], "grid": best_fused[2], "records": best[3],
              "lds_table_max": best_lds[1], "reduce_lds": best_lds[2],
              "fused_min_rows": best_fused[1], "seconds": best_fused[0], "timings": timings}
    table[key] = choice
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except OSError:
            pass
    return choice
