"""Batched form: many filter -> join -> group-by pipelines in one persistent launch.

`BatchedEngine` takes pipelines in the reference's plan-argument form (a Filter on the build
table, a Filter on the probe table, an inner single-key Join, an Aggregate grouped by the build
join key with aggregates over probe columns -- the shape of TPC-H Q3's core and of the config-1
workload), compiles their predicates and aggregate arguments into register programs, lays out
one group table per pipeline (addressed by build key - min key; it is the join directory and the
aggregation table at once), and runs all of them with `tq_batch_run`: a persistent kernel that
drains a device queue of (pipeline, tile) items.  `run()` issues only kernel launches, so it can
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
from .conversion import get_id, right_most
from .operators.aggregate import _is_key, _strip_alias, _unwrap

PHASES = 5


class BatchPipe(ctypes.Structure):
    _fields_ = [("bprog", L.Program), ("bcols", L.ColSet),
                ("pprog", L.Program), ("pcols", L.ColSet),
                ("aprog", L.Program), ("acols", L.ColSet),
                ("nb", ctypes.c_int64), ("np", ctypes.c_int64),
                ("kmin", ctypes.c_int64), ("range", ctypes.c_int64),
                ("naggs", ctypes.c_int32), ("key_dtype", ctypes.c_int32),
                ("agg_op", ctypes.c_int32 * L.TQ_MAX_SEGS),
                ("agg_dtype", ctypes.c_int32 * L.TQ_MAX_SEGS),
                ("dir", ctypes.c_void_p), ("cnt", ctypes.c_void_p), ("acc", ctypes.c_void_p),
                ("tile_counts", ctypes.c_void_p), ("key_out", ctypes.c_void_p),
                ("agg_out", ctypes.c_void_p * L.TQ_MAX_SEGS),
                ("ngroups", ctypes.c_void_p), ("flags", ctypes.c_void_p),
                ("ntiles", ctypes.c_int32 * PHASES), ("pad", ctypes.c_int32)]


class PipelineSpec:
    """One batched pipeline: build/probe columns (id -> tensor), predicates, join keys,
    aggregates [(output id, op, argument tree)] and the output ids of the group key."""

    def __init__(self, build_cols, probe_cols, build_key, probe_key, build_pred, probe_pred,
                 aggs, key_out_id):
        self.build_cols, self.probe_cols = build_cols, probe_cols
        self.build_key, self.probe_key = build_key, probe_key
        self.build_pred, self.probe_pred = build_pred, probe_pred
        self.aggs = aggs
        self.key_out_id = key_out_id


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
    out = torch.empty(3, dtype=torch.int64, device=t.device)
    L.call("tq_groupby_bounds", ctypes.byref(d), None, t.shape[0], L.ptr(out), L.stream_handle())
    lo, hi, _ = out.cpu().tolist()
    return lo, hi


class BatchedEngine:
    """Prepare once, then run() (graph-capturable) and results()."""

    def __init__(self, specs, tile_rows=4096, grid=None, spin_limit=1 << 22, key_bounds=None):
        self.specs = list(specs)
        self.tile_rows = int(tile_rows)
        self.device = torch.device("cuda:0")
        props = torch.cuda.get_device_properties(self.device)
        self.grid = grid or props.multi_processor_count * 4
        self.spin_limit = spin_limit
        self.key_bounds = key_bounds or {}
        self.rejected = []        # indices of specs the batch cannot run
        self._prepare()

    # ---------------------------------------------------------------- set-up
    def _prepare(self):
        dev = self.device
        T = self.tile_rows
        self.pipes = []
        self.builders = []        # keep compiled programs (and their tensors) alive
        self.outputs = []
        self.index = []           # spec index of each batched pipe
        stage_loads = 1
        arena_bytes = 0
        layout = []
        for i, sp in enumerate(self.specs):
            bk = sp.build_cols[sp.build_key].reshape(-1)
            pk = sp.probe_cols[sp.probe_key].reshape(-1)
            if bk.dtype.is_floating_point or pk.dtype.is_floating_point:
                self.rejected.append(i)
                continue
            nb, np_ = bk.shape[0], pk.shape[0]
            bkey_id = (bk.data_ptr(), nb, bk.dtype)
            kb = self.key_bounds.get(bkey_id)
            if kb is None:
                kb = _key_bounds(bk) if nb > 0 else (0, -1)
                self.key_bounds[bkey_id] = kb
            lo, hi = kb
            rng = max(hi - lo + 1, 1)
            if rng > 4 * nb + 65536:
                self.rejected.append(i)
                continue
            bb = E.ProgramBuilder()
            bb.load_slot(bk)
            E.compile_predicate(bb, sp.build_pred, _Cols(sp.build_cols))
            pb = E.ProgramBuilder()
            pb.load_slot(pk)
            E.compile_predicate(pb, sp.probe_pred, _Cols(sp.probe_cols))
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
                ab.column(pk)
            stage_loads = max(stage_loads, bb.nload, pb.nload, ab.nload)
            cap = min(nb, rng)
            layout.append((i, bb, pb, ab, arg_dtypes, lo, rng, cap, nb, np_))
        self.stage_loads = stage_loads
        # one arena for all tables: dir + cnt (u32) and accumulators (8 B each)
        offs = []
        for (_, _, _, ab, _, _, rng, _, _, _) in layout:
            offs.append(arena_bytes)
            arena_bytes += _table_bytes(rng, ab.nseg, T)[-1]
        self.arena = torch.empty(max(arena_bytes, 256), dtype=torch.uint8, device=dev)
        self.ngroups = torch.empty(max(len(layout), 1), dtype=torch.int64, device=dev)
        self.flags = torch.zeros(max(len(layout), 1), dtype=torch.int32, device=dev)
        base = self.arena.data_ptr()
        items = [[] for _ in range(PHASES)]
        arr = (BatchPipe * max(len(layout), 1))()
        for k, ((i, bb, pb, ab, arg_dtypes, lo, rng, cap, nb, np_), off) in enumerate(zip(layout, offs)):
            sp = self.specs[i]
            d = arr[k]
            d.bprog, d.bcols = bb.prog, bb.cols
            d.pprog, d.pcols = pb.prog, pb.cols
            d.aprog, d.acols = ab.prog, ab.cols
            d.nb, d.np, d.kmin, d.range = nb, np_, lo, rng
            d.naggs = len(sp.aggs)
            bk = sp.build_cols[sp.build_key]
            d.key_dtype = L.dtype_code(bk.dtype)
            ops = {'sum': L.TQ_AGG_SUM, 'count': L.TQ_AGG_COUNT, 'avg': L.TQ_AGG_AVG,
                   'min': L.TQ_AGG_MIN, 'max': L.TQ_AGG_MAX}
            cnt_off, acc_off, tc_off, _ = _table_bytes(rng, d.naggs, T)
            p = base + off
            d.dir = p
            d.cnt = p + cnt_off
            d.acc = p + acc_off
            d.tile_counts = p + tc_off
            key_out = torch.empty(cap, dtype=bk.dtype, device=dev)
            agg_outs = []
            for a, ((oid, op, _), dt) in enumerate(zip(sp.aggs, arg_dtypes)):
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
                t = torch.empty(cap, dtype=odt, device=dev)
                agg_outs.append(t)
                d.agg_out[a] = t.data_ptr()
            d.key_out = key_out.data_ptr()
            d.ngroups = self.ngroups.data_ptr() + 8 * k
            d.flags = self.flags.data_ptr() + 4 * k
            nt = [_ceil(rng, T), _ceil(nb, T), _ceil(np_, T), _ceil(rng, T), _ceil(rng, T)]
            for ph in range(PHASES):
                d.ntiles[ph] = nt[ph]
            self.pipes.append(k)
            self.builders.append((bb, pb, ab))
            self.outputs.append((i, key_out, agg_outs))
            self.index.append(i)
            for ph in range(PHASES):
                items[ph].append((nt[ph], k))
        flat = []
        starts = [0]
        for ph in range(PHASES):
            # larger pipelines first inside a phase keeps the queue's tail short
            for ntile, k in sorted(items[ph], key=lambda x: -x[0]):
                flat.extend((k, t) for t in range(ntile))
            starts.append(len(flat))
        self.phase_start = (ctypes.c_int64 * (PHASES + 1))(*starts)
        self.nitems = len(flat)
        it = torch.tensor(flat if flat else [(0, 0)], dtype=torch.int32).reshape(-1)
        self.items = it.to(dev)
        raw = bytes(arr)
        self.desc = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(dev)
        self._desc_host = arr
        self.ws = torch.empty(max(L.size_call("tq_batch_ws_bytes", len(layout)), 256),
                              dtype=torch.uint8, device=dev)
        self.npipes = len(layout)

    # ---------------------------------------------------------------- execution
    def run(self):
        """Launch the batch on the current stream (kernel launches only)."""
        if self.npipes == 0:
            return
        L.call("tq_batch_run", L.ptr(self.desc), self.npipes, L.ptr(self.items), self.phase_start,
               self.tile_rows, self.grid, self.stage_loads, L.ptr(self.ws),
               ctypes.c_uint32(self.spin_limit), L.stream_handle())

    def results(self):
        """Synchronise; returns ({spec index: {out id: tensor}}, [spec indices to rerun])."""
        torch.cuda.synchronize()
        err = int(self.ws[256:260].view(torch.int32).item()) if self.npipes else 0
        if err:
            raise RuntimeError("batched launch aborted: a phase wait exceeded its spin limit")
        ng = self.ngroups.cpu().tolist()
        fl = self.flags.cpu().tolist()
        out, rerun = {}, list(self.rejected)
        for k, (i, key_out, agg_outs) in enumerate(self.outputs):
            if fl[k]:
                rerun.append(i)
                continue
            G = ng[k]
            sp = self.specs[i]
            res = {sp.key_out_id: key_out[:G].reshape(-1, 1)}
            for (oid, _, _), t in zip(sp.aggs, agg_outs):
                res[oid] = t[:G]
            out[i] = res
        return out, sorted(rerun)

    def reset_flags(self):
        self.flags.zero_()


def _table_bytes(rng, naggs, tile):
    """Offsets of cnt, acc and tile_counts in one pipeline's table, and its total size."""
    cnt_off = _align(rng * 4)
    acc_off = cnt_off + _align(rng * 4)
    tc_off = acc_off + _align(rng * 8 * max(naggs, 1))
    return cnt_off, acc_off, tc_off, tc_off + _align(_ceil(rng, tile) * 4)


def _ceil(a, b):
    return (a + b - 1) // b


def _align(n, a=256):
    return (n + a - 1) // a * a


# ==================================================================== autotuner
TUNE_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def autotune(specs, candidates=None, repeats=3, cache=TUNE_CACHE, key=None):
    """Pick tile rows and persistent grid size for a batch by timing candidates on it.
    Results are cached in tuning.json under `key` (device name + workload signature)."""
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
    candidates = candidates or [(t, cu * g) for t in (2048, 4096, 8192) for g in (2, 4, 8)]
    best, timings = None, {}
    bounds = {}
    for tile, grid in candidates:
        eng = BatchedEngine(specs, tile_rows=tile, grid=grid, key_bounds=bounds)
        eng.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            eng.run()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / repeats
        timings[f"{tile}x{grid}"] = dt
        if best is None or dt < best[0]:
            best = (dt, tile, grid)
        del eng
    choice = {"tile_rows": best[1], "grid": best[2], "seconds": best[0], "timings": timings}
    table[key] = choice
    if cache:
        try:
            with open(cache, "w") as f:
                json.dump(table, f, indent=1)
        except OSError:
            pass
    return choice
