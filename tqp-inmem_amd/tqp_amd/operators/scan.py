"""Scan operator (operators/scan.py:8-54 of the reference): column files into the tensor group.

`tqp_scan` keeps the reference's signature and side effects: for each output column it reads
`SF{SF}-tensor-{COLUMN}.pth` from `config_json['tensors path']` (a TorchScript archive whose first
parameter is the column, as the reference's files are; a file holding a plain tensor is read with
`torch.load(weights_only=True)`), names its type from the table schema, marks the primary key
column sorted, and casts it to the engine's dtype (dates int32, decimals float64, text int8, the
rest int64).  It returns the row count.

On the GPU the ingress is arranged for the device instead of column by column:
- a column moves over PCIe once, from pinned host memory, asynchronously on a copy stream, while
  the next file is read from disk;
- a cast that widens (int32 -> int64, float32 -> float64) runs on the device after the copy, so
  the narrower bytes cross the bus; a cast that narrows runs on the host before it;
- the destination is the query's device memory: `mem_pool` when it is a `GPUMemory` slab (the
  reference copies into its pool the same way, scan.py:45-49), else torch's allocator;
- each integer column's min and max ride along as column statistics (`kernels.note_bounds`), the
  way a columnar file's footer carries them, so joins and group-bys over scanned keys plan
  without a bounds pass.
"""
import os
import warnings

import torch

from .. import constants, kernels
from ..utility.logger import message_logger, perf_logger
from ..variable import Variable, VariableState

TYPE_MAPPING = {
    "identifier": "int",
    "integer": "int",
    "decimal": "float",
    "date": "date",
    "variable text": "string",
    "fixed text": "string",
}


def read_column(path):
    """The column tensor of one file (on the host)."""
    try:
        with warnings.catch_warnings():
            warnings.simplefilter("ignore", DeprecationWarning)   # torch.jit's own notice
            return list(torch.jit.load(path, map_location='cpu').parameters())[0].detach()
    except RuntimeError:
        t = torch.load(path, map_location='cpu', weights_only=True)
        if not torch.is_tensor(t):
            raise ValueError(f"{path}: neither a TorchScript archive nor a tensor")
        return t


def target_dtype(t: torch.Tensor, type_name: str):
    """scan.py:32-40: the engine dtype of a column."""
    if type_name == 'date':
        return constants.date_dtype
    if torch.is_floating_point(t):
        return constants.float_dtype
    if type_name == 'string':
        return constants.string_dtype
    assert t.dim() == 1
    return constants.int_dtype


def tqp_scan(gpu_enable, o_tensor_group, args, config_json, schema_table, SF, mem_pool,
             name="scan"):
    tensor_path = config_json.get('tensors path')
    col_names, col_ids = args['output names'], args['output ids']
    rows = 0
    pending = []          # (column id, device tensor, device dtype) copied but not yet cast
    copy_stream = torch.cuda.Stream() if gpu_enable else None
    for col_name, col_id in zip(col_names, col_ids):
        col_name = col_name.upper()
        with perf_logger().time(f"{name} read from disk"):
            path = os.path.join(tensor_path, f'SF{SF}-tensor-{col_name}.pth')
            t = read_column(path)
        type_name = TYPE_MAPPING[schema_table[col_name].split(',')[0]]
        var = Variable(t, type_name)
        if col_name == schema_table['PRIMARY_KEY']:
            var.is_sorted = True
        rows = t.shape[0]
        dt = target_dtype(t, type_name)
        if not gpu_enable:
            var.tensor = t.to(dt)
            o_tensor_group[col_id] = var
            message_logger().info("%s %s", col_name, var)
            continue
        # narrowing casts on the host, widening ones on the device
        narrow = t.element_size() > torch.empty((), dtype=dt).element_size()
        src = t.to(dt) if narrow or t.dtype == torch.bool else t
        bounds = None
        if dt in (torch.int32, torch.int64, torch.int16) and src.dim() == 1 and src.numel() > 0:
            lo, hi = torch.aminmax(src)
            bounds = (int(lo), int(hi))
        with perf_logger().time(f"{name} pin memory"):
            staged = torch.empty(src.shape, dtype=src.dtype, pin_memory=True)
            staged.copy_(src)
            with torch.cuda.stream(copy_stream):
                raw = torch.empty(src.shape, dtype=src.dtype, device='cuda')
                raw.copy_(staged, non_blocking=True)
                raw.record_stream(copy_stream)
        pending.append((col_id, var, raw, dt, staged, bounds))
    if not gpu_enable:
        return rows
    torch.cuda.current_stream().wait_stream(copy_stream)
    for col_id, var, raw, dt, staged, bounds in pending:
        out = _place(raw, dt, mem_pool)
        if bounds is not None:
            kernels.note_bounds(out, *bounds)
        var.tensor = out
        var.state = VariableState.GPU
        if mem_pool is not None and getattr(mem_pool, 'arena', None) is not None:
            var.backing_mem_ptr = None   # the slab is freed as a whole (free_all)
        o_tensor_group[col_id] = var
        message_logger().info("%s", var)
    torch.cuda.synchronize()
    return rows


def _place(raw: torch.Tensor, dt, mem_pool):
    """The column in its engine dtype, in the pool's slab when there is one."""
    arena = getattr(mem_pool, 'arena', None)
    if arena is None:
        return raw if raw.dtype == dt else raw.to(dt)
    _, out = mem_pool.malloc(raw.numel(), dt)
    out = out.view(raw.shape)
    out.copy_(raw)          # the widening cast, when there is one, on the device
    return out
