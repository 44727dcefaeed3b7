"""Sort / take-ordered operator (operators/sort.py:6-58 of the reference).

`tqp_sort` keeps the reference's signature and semantics: rows are ordered by `key ids` with one
`key orders` flag (descending) per key, ties keep input order (the reference applies one
stable torch.sort pass per key, least significant first), 2-D int8 (string) keys compare
row-wise, `limit` keeps the first rows, every `input ids` column is gathered, and the return
value is the row count before the limit.  Keys above 8 GB raise RuntimeError as in the
reference.  The permutation comes from the HIP radix sort (all keys packed into as few 64-bit
sort keys as fit) and the gathers run as one HIP launch.
"""
from .. import kernels
from ..utility.logger import perf_logger
from ..variable import Variable


def tqp_sort(gpu_enable, tensor_group, args, with_limit, name="sort"):
    perf_logger().start(name)
    if not gpu_enable:
        raise NotImplementedError("tqp_amd executes operators on the GPU only (gpu_enable=False has no CPU path)")
    limit = args['limit'] if with_limit else None

    input_ids = args['input ids']
    rows = tensor_group[input_ids[0]].tensor.shape[0]
    key_ids, key_descs = args['key ids'], args['key orders']

    k0 = tensor_group[key_ids[0]].tensor
    if k0.numel() * k0.element_size() >= 8_000_000_000:
        raise RuntimeError("sort input exceeds in-memory limit")
    keys = [(tensor_group[k].tensor, bool(d)) for k, d in zip(key_ids, key_descs)]
    cols = [tensor_group[k].tensor for k in input_ids]
    outs = None
    if limit is not None and 1 <= limit <= 64:
        # over a deferred group-by's outputs: top-k and gathers in one launch sequence
        outs = kernels.take_ordered_deferred(keys, cols, rows, int(limit), k0.device)
    if outs is None:
        perm = kernels.sort_perm(keys, rows, limit, device=k0.device)
        outs = kernels.gather(perm, cols)
    for key_id, t in zip(input_ids, outs):
        tensor_group[key_id] = Variable(t, tensor_group[key_id].tensor_type)
    perf_logger().stop(name)
    return rows
