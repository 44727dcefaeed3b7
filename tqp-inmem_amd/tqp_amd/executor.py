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

from . import kernels
from .operators.aggregate import results_are_direct, tqp_hash_aggregate
from .operators.filter import tqp_filter
from .operators.hashjoin import join_vortex
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


def _pairs_feed_aggregate(plan, i):
    """Whether plan[i] is an inner single-key Join without extra conditions whose pairs only
    feed the Aggregate at plan[i + 1] (unmasked inputs, every result a key or one aggregate
    call): the two can run with one synchronisation."""
    if i + 1 >= len(plan):
        return False
    ja, aa = plan[i][1], plan[i + 1][1]
    if ja['name'] != 'Join' or aa['name'] != 'Aggregate' or aa.get('is partial'):
        return False
    if ja.get('join type') != 'Inner' or 'join condition' in ja or len(ja['left keys']) != 1:
        return False
    ids = aa['input ids']
    if any(k in ids for k in aa.get('mask map', {})):
        return False
    return set(ids) <= set(ja['materialized in']) and results_are_direct(aa['results'])


def _join_then_aggregate(plan, i, tensor_group, cpu_mem_pool, gpu_mem_pool, gpu_enable,
                         op_times):
    """plan[i] (Join) and plan[i + 1] (Aggregate) with one synchronisation.  The join leaves its
    pair count on the device and the group-by reads it there (kernels.Deferred); the counts and
    the build's uniqueness check are read back together at the end, and the outputs sized for
    the capacities are cut to the real counts.  When the check fails (repeated or out-of-bounds
    build keys, no pairs, a fused reduce that gave up) both operators run again as usual, from
    the state before the join."""
    (uj, ja), (ua, aa) = plan[i], plan[i + 1]
    saved_group = dict(tensor_group)
    saved_mask_map = dict(ja['mask map']) if 'mask map' in ja else None
    saved_ids = list(aa['input ids'])
    df = kernels.Deferred()
    t0 = time.perf_counter()
    kernels.set_deferred(df)
    try:
        _run_join(tensor_group, ja, uj, cpu_mem_pool, gpu_mem_pool, gpu_enable)
        after_join = dict(tensor_group)
        t1 = time.perf_counter()
        tqp_hash_aggregate(gpu_enable, tensor_group, aa, cpu_mem_pool=cpu_mem_pool,
                           gpu_mem_pool=gpu_mem_pool, name=f"aggregation {ua}")
        m, g = df.resolve() if df.join is not None else (None, None)
        ok = True
    except kernels.DeferredJoinFailed:
        ok = False
    finally:
        kernels.set_deferred(None)
    if ok:
        if m is not None:
            _cut_to_counts(tensor_group, saved_group, after_join, df, m, g)
        t2 = time.perf_counter()
    else:
        tensor_group.clear()
        tensor_group.update(saved_group)
        if saved_mask_map is not None:
            ja['mask map'] = saved_mask_map
        aa['input ids'][:] = saved_ids
        t0 = time.perf_counter()
        _run_join(tensor_group, ja, uj, cpu_mem_pool, gpu_mem_pool, gpu_enable)
        t1 = time.perf_counter()
        tqp_hash_aggregate(gpu_enable, tensor_group, aa, cpu_mem_pool=cpu_mem_pool,
                           gpu_mem_pool=gpu_mem_pool, name=f"aggregation {ua}")
        t2 = time.perf_counter()
    if op_times is not None:
        op_times.setdefault(f"Join {uj}", []).append(t1 - t0)
        op_times.setdefault(f"Aggregate {ua}", []).append(t2 - t1)
    return next(iter(tensor_group.values())).tensor.size(0)


def _cut_to_counts(tensor_group, before, after_join, df, m, g):
    """Replace every column the two operators left at a capacity size by its first m (pairs)
    or g (groups) rows."""
    def cut(v, n):
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
        if after_join.get(k) is v:
            if t.data_ptr() in join_outs and t.shape[0] == df.join["cap"]:
                cut(v, m)
        elif g is not None and t.data_ptr() in group_outs and t.shape[0] == df.group["cap"]:
            cut(v, g)


def execute(plan, tensor_group, subquery_result=None, cpu_mem_pool=None, gpu_mem_pool=None,
            gpu_enable=True, op_times=None, sync=True):
    """Run `plan` (list of (u, args)) over `tensor_group` in place; returns the last row count.

    Without per-operator synchronisation (sync=False), a Join whose pairs only feed the next
    Aggregate runs together with it (one synchronisation instead of two); the results are the
    same."""
    num_rows = None
    i = 0
    while i < len(plan):
        u, args = plan[i]
        name = args['name']
        if not sync and gpu_enable and _pairs_feed_aggregate(plan, i):
            num_rows = _join_then_aggregate(plan, i, tensor_group, cpu_mem_pool, gpu_mem_pool,
                                            gpu_enable, op_times)
            i += 2
            continue
        i += 1
        t0 = time.perf_counter()
        if name == 'Filter':
            num_rows = tqp_filter(gpu_enable, tensor_group, args, subquery_result,
                                  cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool,
                                  name=f"filter {u}")
        elif name == 'Aggregate':
            num_rows = tqp_hash_aggregate(gpu_enable, tensor_group, args,
                                          cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool,
                                          name=f"aggregation {u}")
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
