"""Group-by aggregate operator (operators/aggregate.py:236-393 of the reference).

`tqp_hash_aggregate` keeps the reference's signature, argument schema and side effects
(`args['input ids']` gains the mask ids in place; key outputs are 2-D [G, 1] columns of the key's
dtype and type name; aggregate outputs are 1-D with the type name of their dtype; the return
value is the row count of the group's first entry).

Grouping is exact: groups are the distinct key tuples, emitted in ascending lexicographic key
order.  The reference numbers groups by the rank of a linear hash with random coefficients
(aggregate.py:320-327), so its group order is arbitrary and distinct keys can collide into one
group; ascending key order is the canonical order of its output, and collisions are not
reproduced.

All aggregates of one operator run in one accumulate launch that reads the uncompacted input
columns and the filter mask (no compaction), evaluates every aggregate argument per row with the
expression kernel, and reduces into an LDS or global table (see the group-by kernels).  Results
that combine aggregates (e.g. `0.2 * avg(x)`) are evaluated over the group rows with the same
expression kernel.
"""
import torch

from .. import constants, kernels, memory
from .. import expr as E
from ..conversion import get_id, right_most
from ..IO.vortex_pipeline import InMemoryPipeline
from ..utility.logger import message_logger, perf_logger
from ..variable import Variable, torch_to_type


def get_result_type(tree, transfer_set, transfer_types):
    """Declared dtype of a result expression (aggregate.py:23-44)."""
    if isinstance(tree, str):
        if '#' in tree and not tree.startswith('Brand#') and tree != '#':
            if get_id(tree) in transfer_set and \
                    transfer_types[transfer_set.index(get_id(tree))] == constants.float_dtype:
                return constants.float_dtype
            return constants.int_dtype
        if tree.isdigit():
            return constants.int_dtype
        if E.is_float(tree):
            return constants.float_dtype
    if isinstance(tree, list):
        if tree and tree[0] == 'CASE':
            return get_result_type(tree[4], transfer_set, transfer_types)
        for e in tree:
            s = get_result_type(e, transfer_set, transfer_types)
            if s == constants.float_dtype:
                return constants.float_dtype
            if s == torch.int64:
                return torch.int64
        return constants.int_dtype


def _is_key(tree):
    return isinstance(tree, str) or (len(tree) == 3 and tree[1] == 'AS' and isinstance(tree[0], str))


def _unwrap(tree):
    while isinstance(tree, list) and len(tree) == 1:
        tree = tree[0]
    return tree


def _strip_alias(tree):
    """[expr, '#', id] and single-element lists evaluate to expr (aggregate.py:69-74)."""
    while isinstance(tree, list) and (len(tree) == 1 or (len(tree) == 3 and tree[1] == '#')):
        tree = tree[0]
    return tree


def _zero_row_aggregates(keys, calls, gb):
    """Aggregates over an input whose rows are all masked out, as the reference computes them.

    With grouping keys there are no groups, and the reference fails on `cnt_rows / cnt_groups`
    (aggregate.py:128) with ZeroDivisionError.  Without keys there is one group: `avg` fails
    when it divides the integer bincount of an empty index in place (aggregate.py:175-177);
    `sum` and `count` give the int64 bincount of an empty index, [0], for any argument dtype
    (aggregate.py:190); `min` gives the dtype's maximum and `max` gives 0 (aggregate.py:154-166).
    """
    if keys:
        raise ZeroDivisionError("aggregate with grouping keys over zero selected rows "
                                "(the group count is zero)")
    if any(c[0] == 'avg' for c in calls):
        raise RuntimeError("avg over zero selected rows: the integer row count cannot "
                           "receive the quotient in place")
    dev = gb.builder.col_tensors[0].device
    out = []
    for c, dt in zip(calls, gb.arg_dtypes):
        if c[0] in ('sum', 'count'):
            out.append(memory.zeros(1, dtype=torch.int64, device=dev))
        elif c[0] == 'min':
            mx = torch.finfo(dt).max if dt.is_floating_point else torch.iinfo(dt).max
            out.append(memory.full((1,), mx, dtype=dt, device=dev))
        else:
            out.append(memory.zeros(1, dtype=dt, device=dev))
    return out


# Parsed result lists, keyed by their text: the plan arguments of an operator are plain nested
# lists that the executor hands over afresh for every query, so the parse is reused by value.
_PARSED = {}


_PARSED_BY_ID = {}   # id(result list) -> (the list, its parse): one query's repeated lookups


def forget_parses():
    """Drop the by-identity memo (the executor calls this as a query starts, so a result list
    edited in place between queries is parsed again)."""
    _PARSED_BY_ID.clear()


def _parse_results(results):
    """(calls, plans) of a result list: the aggregate call nodes in evaluation order, and per
    result ('key', key id) or ('agg', tree).  The returned trees are a cached parse, shared by
    every call with equal text."""
    ent = _PARSED_BY_ID.get(id(results))
    if ent is not None and ent[0] is results and ent[2] == len(results):
        return ent[1]
    key = repr(results)
    hit = _PARSED.get(key)
    if hit is not None:
        _remember(results, hit)
        return hit
    calls = []
    plans = []
    for tree in results:
        tree = _unwrap(tree)
        out_id = get_id(right_most(tree))
        if _is_key(tree):
            plans.append((out_id, 'key', get_id(tree if isinstance(tree, str) else tree[0])))
        else:
            body = tree[0]
            found = E.find_aggregates(body, [])
            if not found:
                raise ValueError(f"aggregate result without an aggregate function: {tree}")
            calls.extend(found)
            plans.append((out_id, 'agg', body))
    if len(_PARSED) > 256:
        _PARSED.clear()
    _PARSED[key] = (calls, plans)
    _remember(results, (calls, plans))
    return calls, plans


def _remember(results, parse):
    # the list is held with its parse (so its id is not reused) and its length checked on a
    # hit; the executor reads a result list without editing it
    if len(_PARSED_BY_ID) >= 64:
        _PARSED_BY_ID.clear()
    _PARSED_BY_ID[id(results)] = (results, parse, len(results))


def results_are_direct(results):
    """Whether every result is a grouping key or one aggregate call as it stands (no
    expression combines aggregates, which would be evaluated over the group rows)."""
    calls, plans = _parse_results(results)
    ids = {id(c) for c in calls}
    return all(kind == 'key' or id(_strip_alias(what)) in ids for _, kind, what in plans)


def result_ids(results):
    """Output ids of a result list, keys and aggregates, in order."""
    return [out_id for out_id, _, _ in _parse_results(results)[1]]


def agg_group(columns, key_ids, results, mask=None):
    """Group `columns` by `key_ids` and evaluate each result tree (aggregate.py:282-364).
    Returns {output id: tensor}."""
    keys = [columns[k].tensor for k in key_ids]
    n = next(iter(columns.values())).tensor.shape[0]
    calls, plans = _parse_results(results)
    gb = kernels.GroupBy(keys, [(c[0], c[1]) for c in calls], columns, mask=mask, n=n)
    key_out, agg_out = gb.run()
    if gb.nsel == 0 and calls:
        agg_out = _zero_row_aggregates(keys, calls, gb)
    by_call = {id(c): t for c, t in zip(calls, agg_out)}
    out = {}
    for out_id, kind, what in plans:
        if kind == 'key':
            pos = key_ids.index(what)
            kv = key_out[pos]
            # keys come out 2-D, [G, 1] or [G, width] for string rows (aggregate.py:343-346)
            out[out_id] = (kv if kv.dim() == 2 else kv.reshape(-1, 1)).to(columns[what].tensor.dtype)
            continue
        body = _strip_alias(what)
        if id(body) in by_call:
            out[out_id] = by_call[id(body)]
            continue
        b = E.ProgramBuilder()
        v = E.compile_row_expr(b, what, columns, agg_ref=by_call)
        if v.kind == "lit":
            v = b.const(v.value, E.literal_dtype(v.value))
        b.end_segment(v)
        b.finish()
        G = key_out[0].shape[0] if key_out else 1
        out[out_id] = kernels.eval_program(b, G, v.dtype, agg_out[0].device)
    return out


def evaluate(device_name, tensors, tree, cnt_rows, cnt_groups, inverted_index):
    """One result tree over rows already numbered into groups (aggregate.py:58-233): a row
    expression gives one value per row; a tree with aggregate calls gives one value per group,
    group g being the rows with inverted_index == g (0 <= g < cnt_groups, each present).

    The aggregate calls of the tree reduce together in one group-by launch keyed by the group
    index (a dense key range, so no hashing), and arithmetic over their results runs as one
    expression launch over the groups.  agg_group is the operator's own path (it numbers the
    groups itself, exactly)."""
    if not gpu_index_ok(inverted_index, cnt_rows):
        raise ValueError("inverted_index must hold one group id per row")
    calls = E.find_aggregates(tree, [])
    body = _strip_alias(_unwrap(tree))
    if not calls:
        b = E.ProgramBuilder()
        v = E.compile_row_expr(b, body, tensors)
        if v.kind == "lit":
            return v.value
        if v.kind != "reg":
            raise NotImplementedError("string-valued row expression")
        b.end_segment(v)
        b.finish()
        return kernels.eval_program(b, cnt_rows, v.dtype, inverted_index.device)
    idx = inverted_index if inverted_index.dtype == torch.int64 else inverted_index.to(torch.int64)
    kernels.note_bounds(idx, 0, max(int(cnt_groups) - 1, 0))
    gb = kernels.GroupBy([idx], [(c[0], c[1]) for c in calls], tensors, n=cnt_rows)
    key_out, agg_out = gb.run()
    if key_out[0].shape[0] != cnt_groups:
        raise ValueError(f"{key_out[0].shape[0]} of {cnt_groups} groups hold rows")
    by_call = {id(c): t for c, t in zip(calls, agg_out)}
    if id(body) in by_call:
        return by_call[id(body)]
    b = E.ProgramBuilder()
    v = E.compile_row_expr(b, body, tensors, agg_ref=by_call)
    if v.kind == "lit":
        v = b.const(v.value, E.literal_dtype(v.value))
    b.end_segment(v)
    b.finish()
    return kernels.eval_program(b, cnt_groups, v.dtype, agg_out[0].device)


def gpu_index_ok(idx, n):
    return torch.is_tensor(idx) and idx.is_cuda and idx.dim() == 1 and idx.shape[0] == n \
        and not idx.dtype.is_floating_point


_STATIC = {}


def tqp_hash_aggregate(gpu_enable, tensor_group, args, gpu_mem_pool, cpu_mem_pool,
                       name="aggregation"):
    perf_logger().start(f"{name}")
    if not gpu_enable:
        raise NotImplementedError("tqp_amd executes operators on the GPU only (gpu_enable=False has no CPU path)")
    assert not args['is partial']

    transfer_set = args['input ids']
    transfer_types = []
    for i in transfer_set:
        transfer_types.extend(tensor_group[i].get_type_torch())

    agg_mask_map = {}
    if 'mask map' in args:
        for k, v in args['mask map'].items():
            if k in transfer_set:
                agg_mask_map[k] = v
                if v not in transfer_set:
                    transfer_set.append(v)

    # the result list's keys and declared dtypes depend on its text and the input column types
    ck = (repr(args['results']), tuple(transfer_set), tuple(transfer_types),
          tuple(tensor_group[i].tensor_type for i in transfer_set))
    static = _STATIC.get(ck)
    if static is None:
        key_ids, key_ids_new, key_types, key_types_name = [], [], [], []
        agg_types = []
        for tree in args['results']:
            tree = _unwrap(tree)
            if _is_key(tree):
                key_ids.append(get_id(tree if isinstance(tree, str) else tree[0]))
                key_ids_new.append(get_id(tree if isinstance(tree, str) else tree[-1]))
                key_types.extend(tensor_group[key_ids[-1]].get_type_torch())
                key_types_name.append(tensor_group[key_ids[-1]].tensor_type)
            else:
                agg_types.append(get_result_type(tree, transfer_set, transfer_types))
        if len(_STATIC) > 256:
            _STATIC.clear()
        static = _STATIC[ck] = (key_ids, key_ids_new, key_types, key_types_name, agg_types)
    key_ids, key_ids_new, key_types, key_types_name, agg_types = static
    agg_ids = args['out agg ids']

    def agg_wrapper(columns):
        masks = {id(v.mask): v.mask for v in columns.values() if getattr(v, 'mask', None) is not None}
        if len(masks) > 1 or (masks and any(getattr(v, 'mask', None) is None for v in columns.values())):
            for v in columns.values():   # mixed masks: compact each column by its own mask
                if getattr(v, 'mask', None) is not None:
                    v.tensor = kernels.compact(v.mask, [v.tensor])[0]
                    v.mask = None
            mask = None
        else:
            mask = next(iter(masks.values())) if masks else None
        key_ids_d = [get_id(key) for key in args['keys']] if args['keys'] else []
        out = agg_group(columns, key_ids_d, args['results'], mask=mask)
        res = [out[i] for i in key_ids_new] + [out[i] for i in agg_ids]
        return [r.contiguous() for r in res]

    agg_pipeline = InMemoryPipeline(
        {i: tensor_group[i] for i in transfer_set},
        key_types + agg_types,
        agg_wrapper,
        mask_map=agg_mask_map,
        gpu_mem_pool=gpu_mem_pool,
        cpu_mem_pool=cpu_mem_pool,
        chunk_size=2_000_000_000,
        name=f"{name} aggregation pipe",
        fuse_masks=True,
    )
    agg_pipeline.do_exchange(20_000_000)
    result = agg_pipeline.get_result()
    message_logger().debug(result)

    rows_keys = [result[i].tensor for i in range(len(key_types))]
    rows_vals = [result[i].tensor for i in range(len(key_types), len(key_types) + len(agg_types))]
    for i in range(len(rows_vals)):
        tensor_group[agg_ids[i]] = Variable(rows_vals[i], torch_to_type(rows_vals[i].dtype))
    for i in range(len(rows_keys)):
        tensor_group[key_ids_new[i]] = Variable(rows_keys[i], key_types_name[i])
    perf_logger().stop(f"{name}")
    return list(tensor_group.values())[0].tensor.size(0)
