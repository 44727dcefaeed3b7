"""Device memory for operator outputs and workspaces.

By default outputs and workspaces come from torch's caching allocator.  While a `DeviceArena`
is active on the calling thread (`use(arena)`), they come from the arena instead: one
pre-sized device slab with a first-fit sub-allocator over 256-byte-aligned byte ranges.  This
is the reference's GPU pool (IO/pinned_mem.py:10-123: a first-fit over 40 MB blocks of a 40 GB
slab, created at main.py:44-57) at a finer grain and without its per-output copy
(IO/vortex_pipeline.py:44-49): the kernels write their outputs in the slab directly.

An arena hands a byte range of its slab to a child arena (`reserve`), which sub-allocates inside
it; the served form reserves each admitted request's memory this way, so admission counts the
slab's real free bytes and a request's operators cannot use more than it was given (they raise
MemoryError).  Freed ranges coalesce with their free neighbours.

Reuse is in host order: a range freed by the host may be handed out at once, which is safe for
work ordered on one stream (a later kernel that writes the range runs after the earlier ones
that read it).  Work on several streams frees a range only after waiting for its kernels, as
the served form does when it releases a request after the request's completion event.
Tensors handed out are views of the slab, valid until their range is freed.
"""
import bisect
import threading

import torch

ALIGN = 256


def _nbytes(shape, dtype):
    n = 1
    for s in shape:
        n *= int(s)
    return n * torch.empty((), dtype=dtype).element_size()


def _shape(shape):
    return tuple(int(s) for s in (shape if isinstance(shape, (tuple, list, torch.Size)) else (shape,)))


def _stream_key():
    if not torch.cuda.is_available():
        return 0
    return torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice())


class DeviceArena:
    """First-fit sub-allocator over bytes [base, base + nbytes) of a device slab."""

    def __init__(self, nbytes: int, device='cuda:0', slab=None, base=0, parent=None):
        nbytes = int(nbytes) // ALIGN * ALIGN
        if nbytes <= 0:
            raise ValueError("an arena needs at least 256 bytes")
        if slab is None:
            slab = torch.empty(nbytes, dtype=torch.uint8, device=device)
        self.slab = slab
        self.device = slab.device
        self.base = base
        self.nbytes = nbytes
        self.parent = parent
        self.released = False
        self.peak = 0                  # most bytes in use at once
        self._starts = [base]          # free ranges, sorted by start
        self._lens = [nbytes]
        self._allocs = {}              # start -> length
        self._free = nbytes
        self._lock = threading.Lock()
        self._roles = {}               # (role, stream) -> (start, length)

    # ---------------------------------------------------------------- byte ranges
    def alloc(self, nbytes: int) -> int:
        """Start of a free range of at least nbytes (rounded up to 256); MemoryError if none."""
        need = max(ALIGN, (int(nbytes) + ALIGN - 1) // ALIGN * ALIGN)
        with self._lock:
            if self.released:
                raise RuntimeError("allocation from a released arena")
            for i, ln in enumerate(self._lens):
                if ln >= need:
                    st = self._starts[i]
                    if ln == need:
                        del self._starts[i], self._lens[i]
                    else:
                        self._starts[i] = st + need
                        self._lens[i] = ln - need
                    self._allocs[st] = need
                    self._free -= need
                    self.peak = max(self.peak, self.nbytes - self._free)
                    return st
            free, largest = self._free, max(self._lens, default=0)
        raise MemoryError(f"device arena cannot allocate {need} bytes "
                          f"(free {free}, largest free range {largest})")

    def free(self, start: int):
        with self._lock:
            ln = self._allocs.pop(start, None)
            if ln is None:
                raise ValueError(f"{start} is not an allocation of this arena")
            self._free += ln
            i = bisect.bisect_left(self._starts, start)
            if i < len(self._starts) and start + ln == self._starts[i]:     # merge right
                ln += self._lens[i]
                del self._starts[i], self._lens[i]
            if i > 0 and self._starts[i - 1] + self._lens[i - 1] == start:  # merge left
                self._lens[i - 1] += ln
            else:
                self._starts.insert(i, start)
                self._lens.insert(i, ln)

    def free_all(self):
        with self._lock:
            self._starts, self._lens = [self.base], [self.nbytes]
            self._allocs.clear()
            self._roles.clear()
            self._free = self.nbytes

    def free_bytes(self) -> int:
        return self._free

    def largest_free(self) -> int:
        with self._lock:
            return max(self._lens, default=0)

    def used_bytes(self) -> int:
        return self.nbytes - self._free

    def free_ranges(self):
        with self._lock:
            return list(zip(self._starts, self._lens))

    def view(self, start: int, nbytes: int) -> torch.Tensor:
        """Bytes [start, start + nbytes) of the slab as a tensor of its own: set_ on the slab's
        storage rather than a slice, so the tensor has its own version counter and a write
        elsewhere in the slab does not look like a write to it (column statistics noted on a
        tensor stay valid until the tensor itself is written)."""
        t = torch.empty(0, dtype=torch.uint8, device=self.slab.device)
        return t.set_(self.slab.untyped_storage(), self.slab.storage_offset() + int(start),
                      (int(nbytes),))

    # ---------------------------------------------------------------- tensors
    def empty(self, shape, dtype) -> torch.Tensor:
        shape = _shape(shape)
        nbytes = _nbytes(shape, dtype)
        return self.view(self.alloc(nbytes), nbytes).view(dtype).view(shape)

    def workspace(self, nbytes: int, role) -> torch.Tensor:
        """A workspace of `role`, handed out again to the next call of the role on the same
        stream (which runs after the previous call's kernels), grown when a call needs more."""
        nbytes = max(int(nbytes), ALIGN)
        if role is None:
            return self.view(self.alloc(nbytes), nbytes)
        key = (role, _stream_key())
        have = self._roles.get(key)
        if have is None or have[1] < nbytes:
            if have is not None:
                self.free(have[0])
                del self._roles[key]
            size = nbytes + nbytes // 4
            have = (self.alloc(size), size)
            self._roles[key] = have
        return self.view(have[0], nbytes)

    # ---------------------------------------------------------------- child arenas
    def reserve(self, nbytes: int) -> "DeviceArena":
        """A child arena over a free range of this one (MemoryError if none is large enough);
        `release()` gives the range back."""
        st = self.alloc(nbytes)
        return DeviceArena(self._allocs[st], slab=self.slab, base=st, parent=self)

    def release(self):
        if self.parent is None:
            raise ValueError("only a reserved child arena is released")
        with self._lock:
            if self.released:
                raise ValueError("arena released twice")
            self.released = True
        self.parent.free(self.base)


class ArenaSet:
    """Allocates from several arenas, first fit in order: the memory a batch of requests
    brings, one reservation per request."""

    def __init__(self, arenas):
        self.arenas = list(arenas)
        self._roles = {}

    def _first(self, fn):
        err = None
        for a in self.arenas:
            try:
                return fn(a)
            except MemoryError as e:
                err = e
        raise err if err is not None else MemoryError("no arena to allocate from")

    def empty(self, shape, dtype) -> torch.Tensor:
        return self._first(lambda a: a.empty(shape, dtype))

    def workspace(self, nbytes: int, role) -> torch.Tensor:
        nbytes = max(int(nbytes), ALIGN)
        if role is None:
            return self._first(lambda a: a.view(a.alloc(nbytes), nbytes))
        key = (role, _stream_key())
        have = self._roles.get(key)
        if have is None or have[2] < nbytes:
            if have is not None:
                have[0].free(have[1])
                del self._roles[key]
            size = nbytes + nbytes // 4
            have = self._first(lambda a: (a, a.alloc(size), size))
            self._roles[key] = have
        return have[0].view(have[1], nbytes)


# ---------------------------------------------------------------- the thread's active arena
_TLS = threading.local()


def active():
    """The calling thread's active arena (or arena set), or None for torch's allocator."""
    return getattr(_TLS, "arena", None)


class use:
    """Context manager: allocate outputs and workspaces from `arena` on this thread
    (None: from torch)."""

    def __init__(self, arena):
        self.arena = arena

    def __enter__(self):
        self.prev = active()
        _TLS.arena = self.arena
        return self.arena

    def __exit__(self, *exc):
        _TLS.arena = self.prev
        return False


def empty(shape, dtype, device) -> torch.Tensor:
    """An uninitialised device tensor from the active arena, else from torch."""
    a = active()
    if a is None:
        return torch.empty(shape, dtype=dtype, device=device)
    return a.empty(shape, dtype)


def zeros(shape, dtype, device) -> torch.Tensor:
    return empty(shape, dtype, device).zero_()


def full(shape, value, dtype, device) -> torch.Tensor:
    return empty(shape, dtype, device).fill_(value)


def to_device(host: torch.Tensor, device) -> torch.Tensor:
    """A device copy of a host tensor, placed in the active arena when there is one."""
    if active() is None:
        return host.to(device)
    out = empty(host.shape, host.dtype, device)
    out.copy_(host)
    return out


def cat(tensors, device) -> torch.Tensor:
    """Concatenation of 1-D device tensors of one dtype."""
    if active() is None:
        return torch.cat(tensors)
    n = sum(int(t.shape[0]) for t in tensors)
    out = empty((n,), tensors[0].dtype, device)
    o = 0
    for t in tensors:
        out[o:o + t.shape[0]].copy_(t)
        o += t.shape[0]
    return out
