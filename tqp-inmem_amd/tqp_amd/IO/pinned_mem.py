"""Memory-pool objects the executor hands to every operator (IO/pinned_mem.py:10-136).

`GPUMemory` is a real device slab: a `memory.DeviceArena` of `capacity_gb` with a first-fit
sub-allocator (256-byte grain instead of the reference's 40 MB blocks).  `executor.execute`
makes it the active arena for the query, so the kernels write operator outputs and workspaces
into the slab directly; the reference's per-output copy into the pool
(IO/vortex_pipeline.py:44-49) is gone.  `free_all` after each query returns everything, as the
reference's main loop does (main.py:293-294); exhaustion raises MemoryError.

`PinnedMemory` keeps the reference's accounting surface only: host tables are not staged
through it on this path (the tables are resident in HBM before the query runs).
"""
import torch

from .. import memory


class MemoryPool:
    """Block accounting with the reference's surface (`malloc_like`, `malloc`, `free`,
    `free_all`, `get_utilization`)."""

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


class GPUMemory:
    """A device slab of `capacity_gb` (decimal GB) sub-allocated first-fit."""
    DEVICE = 'cuda:0'

    def __init__(self, capacity_gb: float, block_size_mb: float, device=None):
        nbytes = int(capacity_gb * 1000 ** 3)
        if nbytes < int(block_size_mb * 1000 ** 2):
            raise ValueError("Pool size must be larger than the block size.")
        self.arena = memory.DeviceArena(nbytes, device or self.DEVICE)
        self.pool_size_bytes = self.arena.nbytes

    def malloc_like(self, subject: torch.Tensor):
        """(handle, tensor shaped like `subject`) in the slab."""
        nbytes = subject.numel() * subject.element_size()
        ptr = self.arena.alloc(nbytes)
        t = self.arena.view(ptr, nbytes).view(subject.dtype).view(subject.shape)
        return ptr, t

    def malloc(self, num_elements, dtype):
        nbytes = int(num_elements) * torch.empty((), dtype=dtype).element_size()
        ptr = self.arena.alloc(nbytes)
        return ptr, self.arena.view(ptr, nbytes).view(dtype)

    def free(self, ptr: int):
        self.arena.free(ptr)

    def free_all(self):
        self.arena.free_all()

    def get_utilization(self):
        return self.arena.used_bytes() / self.arena.nbytes
