"""Column wrapper carried between operators (variable.py:5-61 of the reference).

A `Variable` is a tensor plus its SQL type name ('int', 'float', 'date', 'string', 'bool'), the
pool block that backs it (or None when torch's caching allocator owns it), its placement and a
sortedness hint.  Operators replace `tensor_group[id]` entries with new Variables.

Two attributes have no counterpart in the reference:
* `tile_counts`: when a filter mask was produced by the HIP filter kernel, the per-tile popcounts
  it wrote ride along, so a consumer that compacts by this mask skips its counting pass;
* `mask`: set by a mask-fusing pipeline on an uncompacted input column; the consuming kernel
  reads only rows whose mask byte is set, instead of a compacted copy.
"""
from enum import Enum

import torch

from . import constants


class VariableState(str, Enum):
    CPU = 'cpu resident'
    GPU = 'gpu resident'


class Variable:
    def __init__(self, tensor: torch.Tensor, tensor_type: str, backing_mem_ptr=None,
                 var_state: VariableState = VariableState.CPU, is_sorted: bool = False):
        self.tensor = tensor
        self.tensor_type = tensor_type
        self.state = var_state
        self.backing_mem_ptr = backing_mem_ptr
        self.is_sorted = is_sorted
        self.tile_counts = None
        self.mask = None

    def __repr__(self):
        return (f"Variable (tensor size = {self.tensor.shape} ({self.tensor.dtype}), "
                f"tensor type = '{self.tensor_type}')")

    def __iter__(self):
        yield self.tensor
        yield self.tensor_type

    def __getitem__(self, index):
        return Variable(self.tensor[index], self.tensor_type, self.backing_mem_ptr)

    def free_underlying_mem(self, cpu_pool, gpu_pool):
        if self.backing_mem_ptr:
            if self.state == VariableState.GPU:
                gpu_pool.free(self.backing_mem_ptr)
            else:
                cpu_pool.free(self.backing_mem_ptr)

    def normalize(self):
        """Squeeze to 1-D (variable.py:33-38)."""
        if self.tensor.dim() == 2:
            self.tensor = self.tensor.squeeze()
        if self.tensor.dim() == 0:
            self.tensor = self.tensor.unsqueeze(0)
        assert self.tensor.dim() == 1

    def get_type_torch(self):
        return [type_to_torch(self.tensor_type)]


def torch_to_type(dtype):
    if dtype == constants.float_dtype:
        return 'float'
    if dtype == constants.date_dtype:
        return 'date'
    if dtype == constants.int_dtype:
        return 'int'
    return 'string'


def type_to_torch(dtype):
    if dtype == 'float':
        return constants.float_dtype
    if dtype == 'int':
        return constants.int_dtype
    if dtype == 'date':
        return constants.date_dtype
    return constants.string_dtype
