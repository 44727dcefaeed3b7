"""Plan executor for the hot-path operators (the dispatch of main.py:123-279).

`execute` walks a list of (op index, args) in order and calls the operator the reference's
`exec` calls for each plan node name: Filter -> tqp_filter, Join -> join_vortex (two-key joins
keep one key and turn the other into an equality condition, Spark join types map to the
reference's inner / right-semi / right-outer / right-anti), Aggregate -> tqp_hash_aggregate,
Sort -> tqp_sort, 'take ordered' -> tqp_sort with the limit.  Like the reference it
synchronises the device after every operator and records per-operator wall time.  Scan, Project
and output nodes are outside the hot path: tables arrive in `tensor_group` already resident.
"""
import time

import torch

from . import groupjoin, kernels, memory
from .operators.aggregate import (forget_parses, result_ids, results_are_direct,
                                  tqp_hash_aggregate)
from .operators.filter import tqp_filter
from .operators.hashjoin import join_vortex
from .operators.project import tqp_project
from .operators.scan import tqp_scan
from .operators.sort import tqp_sort

JOIN_TYPE = {'Inner': 'inner', 'LeftSemi': 'right-semi', 'LeftOuter': 'right-outer',
             'LeftAnti': 'right-anti'}


def _run_join(tensor_group, args, u, cpu_mem_pool, gpu_mem_pool, gpu_enable):
    left_keys, right_keys = args['left keys'], args['right keys']
    condition_list = []
    if 'join condition' in args:
        condition_list.append(args['join condition'])
    join_left_key, join_right_key = left_keys[0], right_keys[0]
    if len(left_keys) == 2:
        if tensor_group[left_keys[0]].tensor.is_floating_point():
            assert not tensor_group[left_keys[1]].tensor.is_floating_point(), \
                "Do not support join on two float-type keys"
            join_left_key, join_right_key = left_keys[1], right_keys[1]
            condition_list.append([f'#{left_keys[0]}', '=', f'#{right_keys[0]}'])
        else:
            condition_list.append([f'#{left_keys[1]}', '=', f'#{right_keys[1]}'])
    assert args['join type'] in JOIN_TYPE, f"Unsupport Join Type {args['join type']}"
    return join_vortex(gpu_enable, tensor_group, args, join_left_key, join_right_key,
                       condition_list, JOIN_TYPE[args['join type']],
                       cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool, name=f"join {u}")


def _deferrable(plan, i):
    """How many nodes from plan[i] run with one synchronisation: 2 when plan[i] is an inner
    single-key Join without extra conditions whose pairs only feed the Aggregate at plan[i + 1]
    (unmasked inputs, every result a key or one aggregate call), 3 when a 'take ordered' with a
    limit of at most 64 over that aggregate's outputs follows; else 0."""
    if i + 1 >= len(plan):
        return 0
    ja, aa = plan[i][1], plan[i + 1][1]
    if ja['name'] != 'Join' or aa['name'] != 'Aggregate' or aa.get('is partial'):
        return 0
    if ja.get('join type') != 'Inner' or 'join condition' in ja or len(ja['left keys']) != 1:
        return 0
    ids = aa['input ids']
    if any(k in ids for k in aa.get('mask map', {})):
        return 0
    if not (set(ids) <= set(ja['materialized in']) and results_are_direct(aa['results'])):
        return 0
    if i + 2 < len(plan):
        sa = plan[i + 2][1]
        outs = set(result_ids(aa['results']))
        if sa['name'] == 'take ordered' and 1 <= sa.get('limit', 0) <= 64 \
                and set(sa['input ids']) <= outs and set(sa['key ids']) <= outs:
            return 3
    return 2


def _run_node(tensor_group, u, args, cpu_mem_pool, gpu_mem_pool, gpu_enable):
    name = args['name']
    if name == 'Join':
        return _run_join(tensor_group, args, u, cpu_mem_pool, gpu_mem_pool, gpu_enable)
    if name == 'Aggregate':
        return tqp_hash_aggregate(gpu_enable, tensor_group, args, cpu_mem_pool=cpu_mem_pool,
                                  gpu_mem_pool=gpu_mem_pool, name=f"aggregation {u}")
    return tqp_sort(gpu_enable, tensor_group, args, with_limit=True, name=f"sort lim {u}")


def _run_deferred(plan, i, count, tensor_group, cpu_mem_pool, gpu_mem_pool, gpu_enable,
                  op_times):
    """plan[i:i + count] (Join, Aggregate[, take ordered]) with one synchronisation.  The join
    leaves its pair count on the device, the group-by reads it there and leaves its group count,
    the top-k reads that (kernels.Deferred); the counts and the build's uniqueness check are read
    back together at the end, and the outputs sized for the capacities are cut to the real
    counts.  When the check fails (repeated or out-of-bounds build keys, no pairs, a fused reduce
    that gave up) the nodes run again as usual, from the state before the join."""
    nodes = plan[i:i + count]
    saved_group = dict(tensor_group)
    ja = nodes[0][1]
    saved_mask_map = dict(ja['mask map']) if 'mask map' in ja else None
    saved_ids = list(nodes[1][1]['input ids'])
    df = kernels.Deferred()
    stages = []          # tensor_group after each node
    times = []
    rets = []
    kernels.set_deferred(df)
    try:
        for u, args in nodes:
            t0 = time.perf_counter()
            rets.append(_run_node(tensor_group, u, args, cpu_mem_pool, gpu_mem_pool, gpu_enable))
            stages.append(dict(tensor_group))
            times.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        m, g = df.resolve() if df.join is not None else (None, None)
        ok = True
    except kernels.DeferredJoinFailed:
        ok = False
    finally:
        kernels.set_deferred(None)
    if ok:
        if m is not None:
            _cut_to_counts(tensor_group, saved_group, stages, df, m, g)
        times[-1] += time.perf_counter() - t0
        if count == 3:
            # the sort's row count before its limit: the group count, which the sort saw as
            # a capacity when the group-by was deferred
            num_rows = g if g is not None else rets[-1]
        else:
            num_rows = next(iter(tensor_group.values())).tensor.size(0)
    else:
        tensor_group.clear()
        tensor_group.update(saved_group)
        if saved_mask_map is not None:
            ja['mask map'] = saved_mask_map
        nodes[1][1]['input ids'][:] = saved_ids
        times = []
        for u, args in nodes:
            t0 = time.perf_counter()
            num_rows = _run_node(tensor_group, u, args, cpu_mem_pool, gpu_mem_pool, gpu_enable)
            times.append(time.perf_counter() - t0)
    if op_times is not None:
        for (u, args), dt in zip(nodes, times):
            op_times.setdefault(f"{args['name']} {u}", []).append(dt)
    return num_rows


def _run_groupjoin(plan, i, count, tensor_group, cpu_mem_pool, gpu_mem_pool, gpu_enable,
                   op_times, defer=False):
    """plan[i:i + count] (Join, Aggregate[, take ordered]) with the join and the aggregate fused
    (groupjoin.py) when the aggregate groups by the join key alone; returns None, with nothing
    changed, when they cannot be fused.  The counts are read back once at the end; a void
    fused result (repeated or out-of-bounds build keys, no pairs, a look-back that gave up)
    reruns the nodes per operator from the state before the join.  With `defer`, the kernels
    and the read of the counts are queued and a `_Finish` is returned, whose call collects the
    counts, settles the outputs and returns the row count."""
    (uj, ja), (ua, aa) = plan[i], plan[i + 1]
    p = groupjoin.plan(tensor_group, ja, aa)
    if p is None:
        return None
    saved_group = dict(tensor_group)
    saved_mask_map = dict(ja['mask map']) if 'mask map' in ja else None
    saved_ids = list(aa['input ids'])

    def run_join(tg, args):
        return _run_join(tg, args, uj, cpu_mem_pool, gpu_mem_pool, gpu_enable)

    t0 = time.perf_counter()
    try:
        df = groupjoin.run(tensor_group, ja, aa, p, run_join)
    except kernels.DeferredJoinFailed:
        return None
    times = [time.perf_counter() - t0]
    after_agg = dict(tensor_group)
    state = {"ok": True, "rows_sorted": None}
    try:
        if count == 3:
            us, sa = plan[i + 2]
            t1 = time.perf_counter()
            kernels.set_deferred(df)
            try:
                state["rows_sorted"] = tqp_sort(gpu_enable, tensor_group, sa, with_limit=True,
                                                name=f"sort lim {us}")
            finally:
                kernels.set_deferred(None)
            times.append(time.perf_counter() - t1)
        df.post()
    except kernels.DeferredJoinFailed:
        state["ok"] = False
    nodes = plan[i:i + count]

    def finish():
        ok = state["ok"]
        if ok:
            t2 = time.perf_counter()
            try:
                m, g = df.resolve()
            except kernels.DeferredJoinFailed:
                ok = False
            times[-1] += time.perf_counter() - t2
        if ok:
            for v in df.lazy:
                v.rows = m
            for k, v in tensor_group.items():
                if after_agg.get(k) is v and k in df.written:
                    v.tensor = v.tensor[:g]                       # written by the aggregate
                elif count == 3 and saved_group.get(k) is not v and after_agg.get(k) is not v \
                        and df.sorted is not None and v.tensor.shape[0] == df.sorted["cap"]:
                    v.tensor = v.tensor[:min(g, df.sorted["limit"])]   # written by the top-k
            if count == 3:
                rows_sorted = state["rows_sorted"]
                num_rows = g if rows_sorted == df.group["cap"] else rows_sorted
            else:
                first = next(iter(tensor_group.values()))
                num_rows = first.rows if isinstance(first, groupjoin.LazyJoinColumn) \
                    and not first.materialized else first.tensor.size(0)
            if op_times is not None:
                op_times.setdefault(f"Join+Aggregate {uj}+{ua}", []).append(times[0])
                if count == 3:
                    op_times.setdefault(f"take ordered {plan[i + 2][0]}", []).append(times[1])
            return num_rows
        tensor_group.clear()
        tensor_group.update(saved_group)
        if saved_mask_map is not None:
            ja['mask map'] = saved_mask_map
        aa['input ids'][:] = saved_ids
        num_rows = None
        for u, args in nodes:
            t0 = time.perf_counter()
            num_rows = _run_node(tensor_group, u, args, cpu_mem_pool, gpu_mem_pool, gpu_enable)
            if op_times is not None:
                op_times.setdefault(f"{args['name']} {u}", []).append(time.perf_counter() - t0)
        return num_rows

    return _Finish(finish) if defer else finish()


class _Finish:
    """The rest of a fused group whose kernels are queued: called once, it returns the row
    count."""

    def __init__(self, fn):
        self.fn = fn


class PendingQuery:
    """A plan whose last kernels are queued on the stream (`execute(..., wait=False)`).
    `result()` collects their counts, settles `tensor_group` and returns the row count that
    `execute` returns; the tensor group is not to be read before it.  Work queued on the same
    stream afterwards (the next query's) runs after this query's kernels."""

    def __init__(self, finish, arena):
        self._finish = finish
        self._arena = arena
        self._rows = None
        self._done = False

    def done(self):
        return self._done

    def result(self):
        if not self._done:
            with memory.use(self._arena):
                self._rows = self._finish.fn()
            self._done = True
        return self._rows


def _cut_to_counts(tensor_group, before, stages, df, m, g):
    """Replace every column the deferred nodes left at a capacity size by its real rows: the
    first m (pairs), g (groups) or min(g, limit) (top rows)."""
    def cut(v, n):
        if n == v.tensor.shape[0]:
            return
        t = v.tensor[:n]
        kb = kernels.known_bounds(v.tensor)
        if kb is not None:
            kernels.note_bounds(t, *kb)
        v.tensor = t

    join_outs = {t.data_ptr() for t in df.join["outs"]}
    group_outs = {t.data_ptr() for t in df.group["outs"]} if df.group is not None else set()
    for k, v in tensor_group.items():
        if before.get(k) is v:
            continue
        t = v.tensor
        if stages[0].get(k) is v:            # written by the join
            if t.data_ptr() in join_outs and t.shape[0] == df.join["cap"]:
                cut(v, m)
        elif stages[1].get(k) is v:          # written by the aggregate
            if g is not None and t.data_ptr() in group_outs and t.shape[0] == df.group["cap"]:
                cut(v, g)
        elif df.sorted is not None and g is not None and t.shape[0] == df.sorted["cap"]:
            cut(v, min(g, df.sorted["limit"]))   # written by the top-k


def execute(plan, tensor_group, subquery_result=None, cpu_mem_pool=None, gpu_mem_pool=None,
            gpu_enable=True, op_times=None, sync=True, fuse=True, config_json=None, schema=None,
            scale_factor=None, wait=True):
    """Run `plan` (list of (u, args)) over `tensor_group` in place; returns the last row count.

    Without per-operator synchronisation (sync=False), a Join whose pairs only feed the next
    Aggregate runs together with it (one synchronisation instead of two); when the aggregate
    groups by the join key alone and `fuse` is set, the two run as one fused groupjoin whose
    pairs are never written (the join's output columns are computed if something reads them).
    The results are the same.  With a `GPUMemory` as `gpu_mem_pool`, operator outputs and
    workspaces are placed in its slab; otherwise in the thread's active device arena, if any,
    else in torch's allocator.  Scan nodes read their tables' column files as main.py:138-149
    does, with config_json, schema (the tables' schemas by name) and scale_factor.

    With wait=False (and sync=False) the call returns a `PendingQuery` as soon as the plan's
    kernels are queued, when the plan ends with a fused join and aggregate (and top-k): the
    host prepares the next query while the device runs this one, and `result()` returns the
    row count.  Other plans complete inside the call, and their PendingQuery is already done."""
    forget_parses()
    arena = getattr(gpu_mem_pool, 'arena', None)
    arena = arena if arena is not None else memory.active()
    with memory.use(arena):
        out = _execute(plan, tensor_group, subquery_result, cpu_mem_pool, gpu_mem_pool,
                       gpu_enable, op_times, sync, fuse, (config_json, schema, scale_factor),
                       defer=not wait and not sync)
    if isinstance(out, _Finish):
        return PendingQuery(out, arena)
    if wait:
        return out
    q = PendingQuery(None, arena)
    q._rows, q._done = out, True
    return q


def _execute(plan, tensor_group, subquery_result, cpu_mem_pool, gpu_mem_pool, gpu_enable,
             op_times, sync, fuse, scan_ctx, defer=False):
    num_rows = None
    i = 0
    while i < len(plan):
        u, args = plan[i]
        name = args['name']
        count = _deferrable(plan, i) if not sync and gpu_enable else 0
        if count:
            last = defer and i + count == len(plan)
            num_rows = _run_groupjoin(plan, i, count, tensor_group, cpu_mem_pool, gpu_mem_pool,
                                      gpu_enable, op_times, defer=last) if fuse else None
            if isinstance(num_rows, _Finish):
                return num_rows
            if num_rows is None:
                num_rows = _run_deferred(plan, i, count, tensor_group, cpu_mem_pool,
                                         gpu_mem_pool, gpu_enable, op_times)
            i += count
            continue
        i += 1
        t0 = time.perf_counter()
        if name == 'Scan':
            config_json, schema, sf = scan_ctx
            num_rows = tqp_scan(gpu_enable, tensor_group, args, config_json,
                                schema[args['table name']], SF=sf, mem_pool=gpu_mem_pool,
                                name=f"scan {u}")
        elif name == 'Filter':
            # unsynchronised: the mask is computed when first read, or never when a fused
            # consumer evaluates the predicate in its own kernels
            num_rows = tqp_filter(gpu_enable, tensor_group, args, subquery_result,
                                  cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool,
                                  name=f"filter {u}", lazy=not sync)
        elif name == 'Aggregate':
            num_rows = tqp_hash_aggregate(gpu_enable, tensor_group, args,
                                          cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool,
                                          name=f"aggregation {u}")
        elif name == 'Project':
            num_rows = tqp_project(gpu_enable, tensor_group, args, cpu_mem_pool=cpu_mem_pool,
                                   gpu_mem_pool=gpu_mem_pool, name=f"project {u}")
        elif name == 'Sort':
            num_rows = tqp_sort(gpu_enable, tensor_group, args, with_limit=False, name=f"sort {u}")
        elif name == 'take ordered':
            num_rows = tqp_sort(gpu_enable, tensor_group, args, with_limit=True,
                                name=f"sort lim {u}")
        elif name == 'Join':
            num_rows = _run_join(tensor_group, args, u, cpu_mem_pool, gpu_mem_pool, gpu_enable)
        elif name.endswith('Exchange'):
            raise AssertionError("Exchange is eliminated")
        else:
            raise Exception(f"Unsupported Operator: {name}")
        if sync and gpu_enable:
            torch.cuda.synchronize()
        if op_times is not None:
            op_times.setdefault(f"{name} {u}", []).append(time.perf_counter() - t0)
    return num_rows
