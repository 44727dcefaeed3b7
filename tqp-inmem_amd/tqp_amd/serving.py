"""Served form: a request loop over HBM-resident base tables with continuous batching.

Requests are parameterised pipelines of three shapes:
- shape 0: filter + 6-group aggregate over the 1M-row fact table;
- shape 1: filter -> join -> aggregate over the 1.5M x 6M tables;
- shape 2: the 3-join star over 60M rows.

The base tables stay resident; a request only carries its constants.  The loop:

* **Admission.** Admits queued requests in arrival order while an execution slot is free and
  the backend can reserve the request's memory.  On the GPU the reservation is a range of one
  pre-sized device slab (`memory.DeviceArena`) of the request's `footprint`, an upper bound on
  its masks, join outputs, group tables and workspaces; the request's operators allocate inside
  it (MemoryError, reported for that request, if they would need more), and it is returned when
  the request's results have been copied out.  Admission therefore counts the slab's real free
  bytes: a request that does not fit waits for earlier ones to finish.  A request larger than
  the whole slab is rejected with MemoryError.
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
                 "footprint", "refused")

    def __init__(self, rid, shape, params, arrival):
        self.rid, self.shape, self.params, self.arrival = rid, shape, params, arrival
        self.admitted = self.done = None
        self.result = self.error = None
        self.footprint = 0
        self.refused = 0          # admission passes that found no room for it


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
    footprint(req) -> bytes, capacity() -> bytes, reserve(req) -> bool (False: no room now),
    release(reqs), batchable(req) -> bool, submit_batch(reqs) -> handle,
    submit_single(req) -> handle, poll(handle) -> bool, collect(handle) -> {rid: result}."""

    def __init__(self, backend, max_inflight=4, max_batch=64):
        self.backend = backend
        self.max_inflight = max_inflight
        self.max_batch = max_batch

    def run(self, requests, clock=time.perf_counter):
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue = collections.deque()
        inflight = []                  # (handle, [requests])
        t0 = clock()
        finished = []
        while pending or queue or inflight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                try:
                    validate(r)
                    r.footprint = self.backend.footprint(r)
                    cap = self.backend.capacity()
                    if r.footprint > cap:
                        raise MemoryError(f"request {r.rid} needs {r.footprint} bytes, "
                                          f"device memory {cap}")
                    queue.append(r)
                except (ValueError, MemoryError) as e:
                    r.error = e
                    r.done = now
                    finished.append(r)
            # retire finished work
            still = []
            for h, reqs in inflight:
                if self.backend.poll(h):
                    res = self.backend.collect(h)
                    self.backend.release(reqs)
                    tdone = clock() - t0
                    for r in reqs:
                        r.done = tdone
                        if isinstance(res.get(r.rid), BaseException):
                            r.error = res[r.rid]
                        else:
                            r.result = res.get(r.rid)
                        finished.append(r)
                else:
                    still.append((h, reqs))
            inflight = still
            # admit: FIFO while memory and slots allow
            batch, singles = [], []
            while queue and len(inflight) + (1 if batch else 0) + len(singles) < self.max_inflight:
                r = queue[0]
                if not self.backend.reserve(r):
                    r.refused += 1
                    break
                queue.popleft()
                r.admitted = clock() - t0
                if self.backend.batchable(r) and len(batch) < self.max_batch:
                    batch.append(r)
                else:
                    singles.append(r)
            if batch:
                inflight.append((self.backend.submit_batch(batch), batch))
            for r in singles:
                inflight.append((self.backend.submit_single(r), [r]))
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
    """Executes requests on the device over resident tables (configs 0, 1 and 2), each in the
    range of the device slab reserved for it at admission."""

    def __init__(self, device, c2_scale=1.0, arena_bytes=32 * 2**30):
        import torch

        from . import memory, workloads
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
        self.memory = memory
        self.arena = memory.DeviceArena(arena_bytes, device)
        self._held = {}            # rid -> the request's child arena
        self.peak = {}             # shape -> most arena bytes a request of it used

    def plan(self, req):
        w = self.workloads
        if req.shape == 0:
            return w.config0_plan(**req.params)
        if req.shape == 1:
            return w.config1_plan(limit=10, **req.params)[:4]
        return w.config2_plan()

    def footprint(self, req):
        """Upper bound on the device memory one request allocates (bytes), on whichever path it
        takes: every output and workspace of its operators, none freed before it ends.  The
        per-operator path needs the most; measured peaks (arena high-water marks) were
        1.30 MB for shape 0, 238 MB for shape 1 (42 MB in a batch), and 321 MB / 5.51 GB for
        shape 2 over 3 M / 60 M fact rows."""
        if req.shape == 0:
            n = self.base[0][1].tensor.shape[0]
            return 2 * n + (1 << 16)                          # 2.07 MB at 1 M rows
        if req.shape == 1:
            nb = self.base[1][101].tensor.shape[0]
            np_ = self.base[1][201].tensor.shape[0]
            return 40 * np_ + 24 * nb                         # 276 MB at 1.5 M x 6 M rows
        n = self.base[2][331].tensor.shape[0]
        na = self.base[2][301].tensor.shape[0]
        nbb = self.base[2][311].tensor.shape[0]
        return 96 * n + 64 * (na + nbb)                       # 6.85 GB at 60 M rows

    def capacity(self):
        return self.arena.nbytes

    def reserve(self, req):
        try:
            self._held[req.rid] = self.arena.reserve(req.footprint)
        except MemoryError:
            return False
        return True

    def release(self, reqs):
        for r in reqs:
            a = self._held.pop(r.rid, None)
            if a is not None:
                a.release()

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
        held = self.memory.ArenaSet(self._held[r.rid] for r in reqs)
        with torch.cuda.stream(stream), self.memory.use(held):
            eng = batched.BatchedEngine(specs, key_bounds=self._bounds)
            eng.run()
            ev = torch.cuda.Event()
            ev.record(stream)
        return ("batch", eng, ev, reqs, stream, held)

    def submit_single(self, req):
        torch = self.torch
        memory = self.memory
        result = {}
        ev_box = {}
        arena = self._held[req.rid]

        def work():
            from . import executor
            with self._lock:
                with torch.cuda.stream(self._single_stream):
                    tg = dict(self.base[req.shape])
                    try:
                        with memory.use(arena):
                            executor.execute(copy.deepcopy(self.plan(req)), tg)
                            cols = _result_columns(req.shape, tg)
                        # copied out of the reservation, which is returned once the copies ran
                        result[req.rid] = {k: v.clone() for k, v in cols.items()}
                    except Exception as e:     # reported per request
                        result[req.rid] = e
                    self.peak[req.shape] = max(self.peak.get(req.shape, 0), arena.peak)
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
            torch = self.torch
            eng, reqs, stream, held = h[1], h[3], h[4], h[5]
            out, rerun = eng.results(sync=False)
            res = {}
            # results are copied out of the reservations on the batch's stream, and the copies
            # finish before the reservations are returned
            with torch.cuda.stream(stream):
                for k, r in enumerate(reqs):
                    if k not in out:
                        res[r.rid] = RuntimeError(f"request {r.rid} needs the per-operator path")
                    elif r.shape == 1:
                        res[r.rid] = {301: out[k][301].clone(), 101: out[k][101].clone()}
                    else:
                        res[r.rid] = {c: out[k][c].clone() for c in (6, 7, 22, 23, 24, 25, 26, 27, 28)}
            stream.synchronize()
            per = sum(a.peak for a in held.arenas) // max(len(held.arenas), 1)
            for r in reqs:
                self.peak[r.shape] = max(self.peak.get(r.shape, 0), per)
            return res
        return h[3]


def _result_columns(shape, tg):
    if shape == 0:
        return {k: tg[k].tensor for k in (6, 7, 22, 23, 24, 25, 26, 27, 28)}
    if shape == 1:
        return {301: tg[301].tensor, 101: tg[101].tensor}
    return {k: tg[k].tensor for k in (322, 313, 401, 402)}
