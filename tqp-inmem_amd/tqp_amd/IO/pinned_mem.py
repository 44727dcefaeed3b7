"""Memory-pool objects the executor hands to every operator.

The reference (IO/pinned_mem.py:10-136) carves operator outputs out of a 40 GB device slab with a
Python first-fit over 40 MB blocks and copies every output into it.  Here operator outputs are
allocated by torch's caching allocator (the HIP kernels write them in place, so the extra copy is
gone) and carry `backing_mem_ptr=None`.  The pools keep the reference's accounting surface
(`malloc_like`, `free`, `free_all`, utilisation) so executor code that passes them around and
frees them per query runs unchanged; `MemoryError` on exhaustion is kept.
"""
import torch


class MemoryPool:
    def __init__(self, capacity_gb: float, block_size_mb: float, device='cpu'):
        self.pool_size_bytes = int(capacity_gb * 1000 ** 3)
        self.block_size_bytes = int(block_size_mb * 1000 ** 2)
        if self.pool_size_bytes < self.block_size_bytes:
            raise ValueError("Pool size must be larger than the block size.")
        self._num_blocks = self.pool_size_bytes // self.block_size_bytes
        self._allocs = {}
        self._next = 0
        self._used = 0
        self._device = device

    def _reserve(self, nbytes: int) -> int:
        blocks = max((nbytes + self.block_size_bytes - 1) // self.block_size_bytes, 1)
        if self._used + blocks > self._num_blocks:
            raise MemoryError(f"pool cannot allocate {blocks} blocks")
        ptr = self._next
        self._next += 1
        self._allocs[ptr] = blocks
        self._used += blocks
        return ptr

    def malloc_like(self, subject: torch.Tensor):
        ptr = self._reserve(subject.numel() * subject.element_size())
        return ptr, torch.empty_like(subject, device=self._device)

    def malloc(self, num_elements, dtype):
        t = torch.empty(num_elements, dtype=dtype, device=self._device)
        return self._reserve(t.numel() * t.element_size()), t

    def free(self, ptr: int):
        blocks = self._allocs.pop(ptr, None)
        if blocks is None:
            raise ValueError(f"{ptr} not found in pool")
        self._used -= blocks

    def free_all(self):
        self._allocs.clear()
        self._used = 0

    def get_utilization(self):
        return self._used / self._num_blocks


class PinnedMemory(MemoryPool):
    def __init__(self, capacity_gb: float, block_size_mb: float):
        super().__init__(capacity_gb, block_size_mb, device='cpu')


class GPUMemory(MemoryPool):
    DEVICE = 'cuda:0'

    def __init__(self, capacity_gb: float, block_size_mb: float):
        super().__init__(capacity_gb, block_size_mb, device=self.DEVICE)
