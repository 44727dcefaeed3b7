"""Served form on the GPU (config 4 shape, reduced star): mixed requests arriving faster than
they finish are admitted while earlier launches run, their launches overlap on the device
streams, and every result equals the same request run alone through the per-operator path
(keys and counts bit-exact, sums within 1e-12)."""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


class _Recording:
    """Wraps a backend and records, per launch, when it was submitted and when the server saw
    it done."""

    def __init__(self, be):
        self.be = be
        self.spans = []          # [submit time, done time]
        self._open = {}

    def __getattr__(self, k):
        return getattr(self.be, k)

    def _track(self, h):
        self._open[id(h)] = len(self.spans)
        self.spans.append([time.perf_counter(), None])
        return h

    def submit_batch(self, reqs):
        return self._track(self.be.submit_batch(reqs))

    def submit_single(self, r):
        return self._track(self.be.submit_single(r))

    def poll(self, h):
        done = self.be.poll(h)
        if done and self.spans[self._open[id(h)]][1] is None:
            self.spans[self._open[id(h)]][1] = time.perf_counter()
        return done


def test_served_mix_matches_alone_and_overlaps(device):
    from tqp_amd import serving, serving_bench
    be = serving.GpuBackend(device, c2_scale=0.05)
    serving.Server(be).run(serving.make_requests(10, 1000, seed=5))      # warm-up
    rec = _Recording(be)
    reqs = serving.make_requests(72, 3000, seed=11)
    print("served: 72 requests at 3000/s", flush=True)
    done = serving.Server(rec, max_inflight=4).run(reqs)
    assert len(done) == 72
    failed = [r for r in done if r.error is not None]
    assert not failed, [(r.rid, r.shape, repr(r.error)) for r in failed]
    assert {r.shape for r in done} == {0, 1, 2}
    for r in done:
        alone = serving_bench._alone(be, r)
        assert serving_bench._same(r.result, alone), f"request {r.rid} (shape {r.shape})"
    # launches overlapped: some launch was submitted while another was still in flight
    spans = sorted(s for s in rec.spans if s[1] is not None)
    overlap = sum(1 for a, b in zip(spans, spans[1:]) if b[0] < a[1])
    print("served: launches", len(spans), "overlapping pairs", overlap, flush=True)
    assert len(spans) >= 2 and overlap >= 1
    stats = serving.latency_stats(done)
    assert stats["p50_ms"] is not None and stats["p99_ms"] >= stats["p50_ms"]
