"""Operator wrapper that applies late-materialised filter masks (IO/vortex_pipeline.py:17-73).

The reference compacts every masked input column with `tensor[mask]` (a nonzero + index_select
and a host sync per column, vortex_pipeline.py:58), calls the operator closure, and copies each
output into the GPU memory pool.  Here:

* compaction runs as one HIP launch per mask (all columns sharing a mask move together, in input
  order, with the filter's tile counts reused when present);
* with `fuse_masks=True` nothing is compacted: each masked column reaches the operator as its
  full tensor with `Variable.mask` set, and the operator's kernels skip unselected rows (the
  join and group-by kernels do this);
* outputs stay in the tensors the kernels wrote (the query's device arena, a `GPUMemory` slab,
  when there is one, else torch's caching allocator), so the pool copy is gone and
  `memory_ptr` is None: the arena is freed as a whole after the query.

Everything else follows the reference: mask columns are dropped unless `pass_mask`, all-empty
inputs short-circuit to empty float32 tensors, a 0-d result is wrapped in a list.
"""
from dataclasses import dataclass
from typing import Dict, List

import torch

from .. import kernels
from ..utility.logger import perf_logger
from ..variable import Variable, VariableState


@dataclass
class ColumnMetaData:
    memory_ptr: int
    tensor: torch.Tensor
    cur_ind: int
    placement: VariableState


class InMemoryPipeline:
    def __init__(self, input_columns: Dict[int, Variable], output_columns_type: List[torch.dtype],
                 operator, cpu_mem_pool, gpu_mem_pool=None, div=None, chunk_size=20_000_000_000,
                 pass_mask=False, mask_map: Dict[int, int] = None, name="InMemPipe",
                 fuse_masks=False):
        mask_map = mask_map or {}
        self._input_cols = input_columns
        self._out = []
        self._output_columns_type = list(output_columns_type)
        self._operator = operator
        self._name = name
        self._gpu_mem_pool = gpu_mem_pool
        self._pass_mask = pass_mask
        self._fuse = fuse_masks
        self._mask_map = {}
        cols = list(self._input_cols.keys())
        for k, v in mask_map.items():
            assert v in cols, "mask col not passed in"
            self._mask_map[cols.index(k)] = cols.index(v)
        for col in cols:
            var = self._input_cols[col]
            if not var.tensor.is_cuda:
                raise RuntimeError(f"column {col} is not GPU-resident")
            var.tensor = var.tensor.contiguous()

    def do_exchange(self, granularity: int):
        perf_logger().start(f"{self._name}")
        outbuf = self.safe_apply_op([v.tensor for v in self._input_cols.values()])
        perf_logger().stop(f"{self._name}")
        for t in outbuf:
            self._out.append(ColumnMetaData(None, t, -1, VariableState.GPU))

    def safe_apply_op(self, input_tensors: List[torch.Tensor]):
        if all(t.numel() == 0 for t in input_tensors):
            return [torch.tensor([], device='cuda') for _ in self._output_columns_type]
        items = list(self._input_cols.items())
        mask_positions = set(self._mask_map.values())
        input_dict = {}
        # group masked columns by their mask so each mask costs one compaction launch
        by_mask = {}
        for i, (col_name, old_var) in enumerate(items):
            if i in mask_positions:
                if self._pass_mask:
                    input_dict[col_name] = Variable(input_tensors[i], old_var.tensor_type)
                continue
            if i in self._mask_map:
                by_mask.setdefault(self._mask_map[i], []).append(i)
            else:
                input_dict[col_name] = Variable(input_tensors[i], old_var.tensor_type)
        for mpos, members in by_mask.items():
            mvar = items[mpos][1]
            mask = input_tensors[mpos]
            if self._fuse:
                for i in members:
                    v = Variable(input_tensors[i], items[i][1].tensor_type)
                    v.mask = mask
                    input_dict[items[i][0]] = v
            else:
                outs = kernels.compact(mask, [input_tensors[i] for i in members],
                                       tile_counts=getattr(mvar, 'tile_counts', None))
                for i, t in zip(members, outs):
                    input_dict[items[i][0]] = Variable(t, items[i][1].tensor_type)
        # keep the reference's input order
        ordered = {k: input_dict[k] for k, _ in items if k in input_dict}
        res = self._operator(ordered)
        if isinstance(res, torch.Tensor) and res.ndim == 0:
            res = [res]
        return list(res)

    def get_result(self):
        return self._out
