"""Served-form scheduler on the CPU with a fake backend: admission by reservations in an arena
(the device arena's allocator over a host slab), continuous batching into concurrent launches,
and the host-forceable failures (oversized request, malformed request, a failed execution) --
no GPU involved."""
import time

import pytest

from tqp_amd import memory, serving


class FakeBackend:
    def __init__(self, batch_s=0.02, single_s=0.01, fp=100, fail_rid=None, budget=10**6):
        self.batch_s, self.single_s, self.fp = batch_s, single_s, fp
        self.fail_rid = fail_rid
        self.arena = memory.DeviceArena(budget, device='cpu')
        self.held = {}
        self.max_live_bytes = 0
        self.batches = []
        self.concurrent = 0
        self.max_concurrent = 0

    def footprint(self, r):
        return self.fp * (10 if r.shape == 2 else 1)

    def capacity(self):
        return self.arena.nbytes

    def reserve(self, r):
        try:
            self.held[r.rid] = self.arena.reserve(r.footprint)
        except MemoryError:
            return False
        self.max_live_bytes = max(self.max_live_bytes, self.arena.used_bytes())
        return True

    def release(self, reqs):
        for r in reqs:
            self.held.pop(r.rid).release()

    def batchable(self, r):
        return r.shape in (0, 1)

    def _start(self, reqs):
        self.concurrent += 1
        self.max_concurrent = max(self.max_concurrent, self.concurrent)

    def submit_batch(self, reqs):
        self._start(reqs)
        self.batches.append([r.rid for r in reqs])
        return [time.perf_counter() + self.batch_s, reqs]

    def submit_single(self, r):
        self._start([r])
        return [time.perf_counter() + self.single_s, [r]]

    def poll(self, h):
        return time.perf_counter() >= h[0]

    def collect(self, h):
        self.concurrent -= 1
        return {r.rid: (RuntimeError("boom") if r.rid == self.fail_rid else ("ok", r.rid)) for r in h[1]}


def test_all_requests_complete_with_latency():
    reqs = serving.make_requests(200, rate=2000, seed=1)
    be = FakeBackend(budget=10_000 * 256 // 100)
    done = serving.Server(be, max_inflight=4).run(reqs)
    assert sorted(r.rid for r in done) == list(range(200))
    assert all(r.error is None and r.result == ("ok", r.rid) for r in done)
    assert all(r.done >= r.admitted >= r.arrival - 1e-9 for r in done)
    st = serving.latency_stats(done)
    assert 0 < st["p50_ms"] <= st["p99_ms"]


def test_admission_respects_budget_and_overlaps_batches():
    reqs = serving.make_requests(300, rate=5000, seed=2)
    be = FakeBackend(batch_s=0.01, single_s=0.005, fp=256, budget=6400)
    done = serving.Server(be, max_inflight=4).run(reqs)
    assert be.max_live_bytes <= 6400
    assert sum(1 for r in done if r.refused) > 0    # some requests waited for memory
    assert all(r.error is None for r in done) and be.arena.used_bytes() == 0
    assert be.max_concurrent >= 2            # launches overlap
    # requests that arrived during a running batch went into later batches
    assert len(be.batches) > 1 and all(len(b) >= 1 for b in be.batches)


def test_oversized_and_malformed_requests_fail_alone():
    reqs = serving.make_requests(20, rate=1000, seed=3)
    reqs[3].params = {"bogus": 1}
    big = [r for r in reqs if r.shape == 2]
    be = FakeBackend(fp=256, budget=1024)
    done = serving.Server(be).run(reqs)   # shape 2 needs 2560 > 1024
    by = {r.rid: r for r in done}
    assert isinstance(by[3].error, ValueError)
    for r in big:
        if r.rid != 3:
            assert isinstance(by[r.rid].error, MemoryError)
    ok = [r for r in done if r.error is None]
    assert ok and all(r.result == ("ok", r.rid) for r in ok)


def test_failed_execution_is_reported_per_request():
    reqs = serving.make_requests(30, rate=1000, seed=4)
    be = FakeBackend(fail_rid=7)
    done = {r.rid: r for r in serving.Server(be).run(reqs)}
    assert isinstance(done[7].error, RuntimeError)
    assert all(done[i].error is None for i in done if i != 7)


def test_poisson_arrivals_rate():
    t = serving.poisson_arrivals(1000, 20000, seed=5)
    assert abs(20000 / t[-1] - 1000) < 50
