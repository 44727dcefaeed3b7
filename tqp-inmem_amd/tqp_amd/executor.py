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

from .operators.aggregate import tqp_hash_aggregate
from .operators.filter import tqp_filter
from .operators.hashjoin import join_vortex
from .operators.sort import tqp_sort

JOIN_TYPE = {'Inner': 'inner', 'LeftSemi': 'right-semi', 'LeftOuter': 'right-outer',
             'LeftAnti': 'right-anti'}


def execute(plan, tensor_group, subquery_result=None, cpu_mem_pool=None, gpu_mem_pool=None,
            gpu_enable=True, op_times=None, sync=True):
    """Run `plan` (list of (u, args)) over `tensor_group` in place; returns the last row count."""
    num_rows = None
    for u, args in plan:
        name = args['name']
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
            num_rows = join_vortex(gpu_enable, tensor_group, args, join_left_key, join_right_key,
                                   condition_list, JOIN_TYPE[args['join type']],
                                   cpu_mem_pool=cpu_mem_pool, gpu_mem_pool=gpu_mem_pool,
                                   name=f"join {u}")
        elif name.endswith('Exchange'):
            raise AssertionError("Exchange is eliminated")
        else:
            raise Exception(f"Unsupported Operator: {name}")
        if sync and gpu_enable:
            torch.cuda.synchronize()
        if op_times is not None:
            op_times.setdefault(f"{name} {u}", []).append(time.perf_counter() - t0)
    return num_rows
