"""Served form: a request loop over HBM-resident base tables with continuous batching.

Requests are parameterised pipelines of three shapes:
- shape 0: filter + 6-group aggregate over the 1M-row fact table;
- shape 1: filter -> join -> aggregate over the 1.5M x 6M tables;
- shape 2: the 3-join star over 60M rows.

The base tables stay resident; a request only carries its constants.  The loop:

* **Admission.** Admits queued requests in arrival order while the estimated intermediate memory
  of everything in flight stays within the HBM budget and an execution slot is free.  The estimate
  is `footprint_bytes`, an upper bound on masks, join outputs and group tables.  A request larger
  than the whole budget is rejected with MemoryError.
* **Continuous batching.** Every pass it moves all admitted shape-0 and shape-1 requests into one
  persistent batched launch (`batched.BatchedEngine`: group-by pipelines and join pipelines).  Requests that arrive while that launch runs are
  admitted as soon as slots and memory free up, into the next launch on another stream.  Batches
  on different streams overlap on the device.
* **Other shapes.** Shape-2 requests (three joins) run through the per-operator path on a worker
  thread with its own stream, overlapping the batches.
* **Results and latency.** Completion is detected with stream events, without blocking.
  Results are stored per request; latency is completion time minus arrival time.

`Server` is agnostic of how work executes: `GpuBackend` runs it on the device, and the CPU tests
use a fake backend to check admission, batching and failure handling without a GPU.
"""
import collections
import copy
import math
import threading
import time

import numpy as np


class Request:
    __slots__ = ("rid", "shape", "params", "arrival", "admitted", "done", "result", "error",
                 "footprint")

    def __init__(self, rid, shape, params, arrival):
        self.rid, self.shape, self.params, self.arrival = rid, shape, params, arrival
        self.admitted = self.done = None
        self.result = self.error = None
        self.footprint = 0


def poisson_arrivals(rate, n, seed):
    """n arrival times (seconds from 0) of a Poisson process of `rate` per second."""
    g = np.random.default_rng(seed)
    return np.cumsum(g.exponential(1.0 / rate, n))


def make_requests(n, rate, seed, mix=(0.4, 0.4, 0.2)):
    """Open-loop request stream: shapes drawn with probabilities `mix`, randomized constants."""
    g = np.random.default_rng(seed + 1)
    arrivals = poisson_arrivals(rate, n, seed)
    shapes = g.choice(3, size=n, p=list(mix))
    reqs = []
    for i in range(n):
        s = int(shapes[i])
        if s == 0:
            params = {"ship_day_max": int(g.integers(1800, 2557))}
        elif s == 1:
            params = {"segment": int(g.integers(0, 5)), "ship_day_min": int(g.integers(0, 2557))}
        else:
            params = {}
        reqs.append(Request(i, s, params, float(arrivals[i])))
    return reqs


def validate(req):
    if req.shape not in (0, 1, 2):
        raise ValueError(f"request {req.rid}: unknown shape {req.shape}")
    keys = {0: {"ship_day_max"}, 1: {"segment", "ship_day_min"}, 2: set()}[req.shape]
    if set(req.params) != keys:
        raise ValueError(f"request {req.rid}: parameters {sorted(req.params)} != {sorted(keys)}")


class Server:
    """The admission / continuous-batching loop.  `backend` provides:
    footprint(req) -> bytes, batchable(req) -> bool, submit_batch(reqs) -> handle,
    submit_single(req) -> handle, poll(handle) -> bool, collect(handle) -> {rid: result}."""

    def __init__(self, backend, budget_bytes=200 * 10**9, max_inflight=4, max_batch=64):
        self.backend = backend
        self.budget = budget_bytes
        self.max_inflight = max_inflight
        self.max_batch = max_batch

    def run(self, requests, clock=time.perf_counter):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        inflight = []                  # (handle, [requests], footprint)
        used = 0
        t0 = clock()
        finished = []
        while pending or queue or inflight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                try:
                    validate(r)
                    r.footprint = self.backend.footprint(r)
                    if r.footprint > self.budget:
                        raise MemoryError(f"request {r.rid} needs {r.footprint} bytes, budget {self.budget}")
                    queue.append(r)
                except (ValueError, MemoryError) as e:
                    r.error = e
                    r.done = now
                    finished.append(r)
            # retire finished work
            still = []
            for h, reqs, fp in inflight:
                if self.backend.poll(h):
                    res = self.backend.collect(h)
                    tdone = clock() - t0
                    for r in reqs:
                        r.done = tdone
                        if isinstance(res.get(r.rid), BaseException):
                            r.error = res[r.rid]
                        else:
                            r.result = res.get(r.rid)
                        finished.append(r)
                    used -= fp
                else:
                    still.append((h, reqs, fp))
            inflight = still
            # admit: FIFO while memory and slots allow
            batch, singles = [], []
            while queue and len(inflight) + (1 if batch else 0) + len(singles) < self.max_inflight:
                r = queue[0]
                if used + r.footprint > self.budget:
                    break
                queue.popleft()
                used += r.footprint
                r.admitted = clock() - t0
                if self.backend.batchable(r) and len(batch) < self.max_batch:
                    batch.append(r)
                else:
                    singles.append(r)
            if batch:
                inflight.append((self.backend.submit_batch(batch), batch, sum(r.footprint for r in batch)))
            for r in singles:
                inflight.append((self.backend.submit_single(r), [r], r.footprint))
            if not batch and not singles:
                if pending and not queue and not inflight:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, 0.002))
                else:
                    time.sleep(0.0002)
        return finished


def latency_stats(finished):
    lat = np.array([r.done - r.arrival for r in finished if r.error is None])
    if lat.size == 0:
        return {"p50_ms": None, "p99_ms": None}
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3)}


# ==================================================================== GPU backend
class GpuBackend:
    """Executes requests on the device over resident tables (configs 0, 1 and 2)."""

    def __init__(self, device, c2_scale=1.0):
        import torch

        from . import workloads
        self.torch = torch
        self.device = device
        t0 = workloads.config0_tables()
        t1 = workloads.config1_tables()
        if c2_scale >= 1.0:
            t2 = workloads.config2_tables()
        else:
            t2 = workloads.config2_tables(n_fact=int(60_000_000 * c2_scale),
                                          n_a=int(15_000_000 * c2_scale),
                                          n_b=max(int(2_000_000 * c2_scale), 1000))
        self.rows = {0: workloads.input_rows(t0, 0), 1: workloads.input_rows(t1, 1),
                     2: workloads.input_rows(t2, 2)}
        self.base = {0: workloads.to_tensor_group(t0, device), 1: workloads.to_tensor_group(t1, device),
                     2: workloads.to_tensor_group(t2, device)}
        self.resident_bytes = sum(v.tensor.numel() * v.tensor.element_size()
                                  for g in self.base.values() for v in g.values())
        self.workloads = workloads
        self._lock = threading.Lock()
        self._single_stream = torch.cuda.Stream()
        self._batch_streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        self._next = 0
        self._bounds = {}

    def plan(self, req):
        w = self.workloads
        if req.shape == 0:
            return w.config0_plan(**req.params)
        if req.shape == 1:
            return w.config1_plan(limit=10, **req.params)[:4]
        return w.config2_plan()

    def footprint(self, req):
        """Upper bound on the request's intermediate device memory (bytes)."""
        if req.shape == 0:
            n = 1_000_000
            return n + 6 * 64 * 8 + 4 * n // 1024            # mask + group table + tile counts
        if req.shape == 1:
            nb, np_ = 1_500_000, 6_000_000
            return nb + np_ + np_ * 24 + nb * 16 + nb * 16   # masks, join output, index, groups
        n = self.base[2][331].tensor.shape[0]
        return 3 * n * 32 + 16 * 100_000                     # three join outputs + group table

    def batchable(self, req):
        return req.shape in (0, 1)

    def submit_batch(self, reqs):
        from . import batched
        torch = self.torch
        stream = self._batch_streams[self._next % 2]
        self._next += 1
        specs = []
        for r in reqs:
            if r.shape == 1:
                specs.append(batched.match_plan(self.plan(r), self.base[1]))
            else:
                specs.append(batched.match_groupby_plan(self.plan(r), self.base[0]))
        with torch.cuda.stream(stream):
            eng = batched.BatchedEngine(specs, key_bounds=self._bounds)
            eng.run()
            ev = torch.cuda.Event()
            ev.record(stream)
        return ("batch", eng, ev, reqs)

    def submit_single(self, req):
        torch = self.torch
        result = {}
        ev_box = {}

        def work():
            from . import executor
            with self._lock:
                with torch.cuda.stream(self._single_stream):
                    tg = dict(self.base[req.shape])
                    try:
                        executor.execute(copy.deepcopy(self.plan(req)), tg)
                        result[req.rid] = _result_columns(req.shape, tg)
                    except Exception as e:     # reported per request
                        result[req.rid] = e
                    ev = torch.cuda.Event()
                    ev.record(self._single_stream)
                    ev_box["ev"] = ev
        th = threading.Thread(target=work, daemon=True)
        th.start()
        return ("single", th, ev_box, result)

    def poll(self, h):
        if h[0] == "batch":
            return h[2].query()
        th, ev_box = h[1], h[2]
        return not th.is_alive() and ev_box["ev"].query()

    def collect(self, h):
        if h[0] == "batch":
            eng, reqs = h[1], h[3]
            out, rerun = eng.results(sync=False)
            res = {}
            for k, r in enumerate(reqs):
                if k not in out:
                    res[r.rid] = RuntimeError(f"request {r.rid} needs the per-operator path")
                elif r.shape == 1:
                    res[r.rid] = {301: out[k][301], 101: out[k][101]}
                else:
                    res[r.rid] = {c: out[k][c] for c in (6, 7, 22, 23, 24, 25, 26, 27, 28)}
            return res
        return h[3]


def _result_columns(shape, tg):
    if shape == 0:
        return {k: tg[k].tensor for k in (6, 7, 22, 23, 24, 25, 26, 27, 28)}
    if shape == 1:
        return {301: tg[301].tensor, 101: tg[101].tensor}
    return {k: tg[k].tensor for k in (322, 313, 401, 402)}
