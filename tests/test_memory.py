"""Device arena allocator on the CPU (its slab on the host): first fit, coalescing, child
reservations, workspaces by role, and the thread's active arena -- checked against a brute-force
model of which bytes are taken."""
import threading

import numpy as np
import pytest
import torch

from tqp_amd import memory


def _model_check(a, taken):
    """The arena's free ranges are exactly the untaken bytes of the model, coalesced."""
    free = ~taken
    ranges = []
    i = 0
    n = free.size
    while i < n:
        if free[i]:
            j = i
            while j < n and free[j]:
                j += 1
            ranges.append((a.base + i, j - i))
            i = j
        else:
            i += 1
    assert a.free_ranges() == ranges
    assert a.free_bytes() == int(free.sum())


def test_random_alloc_free_matches_byte_model():
    g = np.random.default_rng(0)
    a = memory.DeviceArena(1 << 16, device='cpu')
    taken = np.zeros(a.nbytes, dtype=bool)
    live = {}
    for step in range(3000):
        if live and (g.random() < 0.45 or len(live) > 40):
            st = list(live)[g.integers(len(live))]
            a.free(st)
            taken[st:st + live.pop(st)] = False
        else:
            want = int(g.integers(1, 4000))
            need = (want + 255) // 256 * 256
            # first fit: the lowest start of an untaken run of `need` bytes
            expect = None
            run = 0
            for i in range(0, a.nbytes, 256):
                run = run + 256 if not taken[i] else 0
                if run >= need:
                    expect = i + 256 - need
                    break
            if expect is None:
                with pytest.raises(MemoryError):
                    a.alloc(want)
                continue
            st = a.alloc(want)
            assert st == expect and st % 256 == 0
            assert not taken[st:st + need].any()
            taken[st:st + need] = True
            live[st] = need
        if step % 50 == 0:
            _model_check(a, taken)
    for st in list(live):
        a.free(st)
    assert a.free_ranges() == [(0, a.nbytes)]


def test_coalescing_both_sides_and_errors():
    a = memory.DeviceArena(4096, device='cpu')
    x, y, z = a.alloc(256), a.alloc(512), a.alloc(256)
    assert (x, y, z) == (0, 256, 768)
    a.free(x)
    a.free(z)
    assert a.free_ranges() == [(0, 256), (768, 4096 - 768)]
    a.free(y)                      # joins the ranges on both sides
    assert a.free_ranges() == [(0, 4096)]
    with pytest.raises(ValueError):
        a.free(y)
    with pytest.raises(MemoryError):
        a.alloc(4097)
    assert a.peak == 1024


def test_tensors_are_views_of_the_slab():
    a = memory.DeviceArena(8192, device='cpu')
    t = a.empty((10, 3), torch.float64)
    u = a.empty(7, torch.int32)
    assert t.shape == (10, 3) and t.dtype == torch.float64 and u.shape == (7,)
    base = a.slab.data_ptr()
    assert t.data_ptr() == base and u.data_ptr() == base + 256      # 240 bytes round up
    t.fill_(1.5)
    assert torch.all(a.slab[:240].view(torch.float64) == 1.5)


def test_noted_bounds_survive_writes_elsewhere_in_the_slab():
    """Each arena tensor has its own version counter: bounds noted on it stay known while other
    tensors of the slab are written, and are dropped when it is written itself."""
    from tqp_amd import kernels
    a = memory.DeviceArena(1 << 16, device='cpu')
    c = memory.DeviceArena(1 << 14, device='cpu', slab=a.slab[4096:4096 + (1 << 14)])
    for arena in (a, c):
        t = arena.empty(100, torch.int64)
        t.copy_(torch.arange(100))
        kernels.note_bounds(t, 0, 99)
        u = arena.empty(50, torch.int64)
        u.fill_(7)
        arena.slab[-8:].zero_()
        assert kernels.known_bounds(t) == (0, 99)
        assert t.data_ptr() == arena.slab.data_ptr() and torch.equal(t, torch.arange(100))
        t[0] = 5
        assert kernels.known_bounds(t) is None


def test_child_reservations_and_release():
    root = memory.DeviceArena(10_000, device='cpu')
    c1 = root.reserve(3000)
    c2 = root.reserve(3000)
    assert c1.nbytes == 3072 and c2.base == 3072
    assert root.free_bytes() == root.nbytes - 6144
    with pytest.raises(MemoryError):
        root.reserve(4000)         # 9984 - 6144 = 3840 left
    t = c1.empty(100, torch.int64)
    assert t.data_ptr() == root.slab.data_ptr() + c1.base
    with pytest.raises(MemoryError):
        c1.empty(400, torch.int64)  # 3200 > 3072 - 1024: the child enforces its range
    c1.release()
    with pytest.raises(RuntimeError):
        c1.alloc(1)
    with pytest.raises(ValueError):
        c1.release()
    assert root.reserve(3000).base == 0     # the released range is reused first fit


def test_workspace_roles_reuse_and_grow():
    a = memory.DeviceArena(1 << 16, device='cpu')
    w1 = a.workspace(1000, "sort")
    w2 = a.workspace(900, "sort")
    assert w1.data_ptr() == w2.data_ptr() and w2.numel() == 900
    w3 = a.workspace(5000, "sort")          # grows: the old range is freed, a new one taken
    assert w3.numel() == 5000
    assert a.used_bytes() == (5000 + 1250 + 255) // 256 * 256
    a.free_all()
    assert a.used_bytes() == 0 and a.workspace(10, "sort").numel() == 256


def test_arena_set_first_fit_across_reservations():
    root = memory.DeviceArena(1 << 14, device='cpu')
    cs = [root.reserve(1024), root.reserve(4096)]
    s = memory.ArenaSet(cs)
    t = s.empty(200, torch.int32)           # 800 bytes: first child
    u = s.empty(300, torch.int32)           # 1200: does not fit the first's rest, second child
    assert cs[0].used_bytes() == 1024 and cs[1].used_bytes() == 1280
    w = s.workspace(2000, "batch")
    assert w.numel() == 2000 and cs[1].used_bytes() == 1280 + 2560
    with pytest.raises(MemoryError):
        s.empty(1000, torch.int32)
    del t, u


def test_active_arena_is_per_thread_and_nests():
    a = memory.DeviceArena(1 << 14, device='cpu')
    b = memory.DeviceArena(1 << 14, device='cpu')
    assert memory.active() is None
    seen = {}

    def other():
        seen["other"] = memory.active()

    with memory.use(a):
        x = memory.empty(10, torch.int64, 'cpu')
        assert x.data_ptr() == a.slab.data_ptr()
        with memory.use(b):
            assert memory.active() is b
            y = memory.zeros(4, torch.float32, 'cpu')
            assert y.data_ptr() == b.slab.data_ptr() and torch.all(y == 0)
        assert memory.active() is a
        th = threading.Thread(target=other)
        th.start()
        th.join()
        c = memory.cat([torch.arange(3), torch.arange(5)], 'cpu')
        assert c.tolist() == [0, 1, 2, 0, 1, 2, 3, 4]
        assert a.used_bytes() == 512
    assert seen["other"] is None
    assert memory.active() is None
    z = memory.empty(10, torch.int64, 'cpu')      # torch's allocator again
    assert not (a.slab.data_ptr() <= z.data_ptr() < a.slab.data_ptr() + a.nbytes)
