"""Open-loop sweep of the served form (config 4): Poisson arrivals at several rates over the
resident config-0/1/2 tables, mix 40/40/20, a 200 GB device arena for the requests' memory.
Reports sustained requests/s and input rows/s with p50/p99 latency per rate, and checks every
completed result against the same request run alone through the per-operator path."""
import copy
import time

import numpy as np
import torch

from . import executor, serving


def _alone(be, req):
    tg = dict(be.base[req.shape])
    executor.execute(copy.deepcopy(be.plan(req)), tg)
    torch.cuda.synchronize()
    return serving._result_columns(req.shape, tg)


def _same(a, b):
    for k, x in a.items():
        y = b[k]
        x, y = x.reshape(-1).cpu(), y.reshape(-1).cpu()
        if x.shape != y.shape:
            return False
        if x.dtype.is_floating_point:
            xn, yn = x.numpy(), y.numpy()
            if not np.all(np.abs(xn - yn) <= 1e-12 * np.abs(yn)):
                return False
        elif not torch.equal(x, y):
            return False
    return True


def sweep(dev, rates=(50, 200, 500, 1000, 2000), n_requests=200, budget=200 * 10**9):
    be = serving.GpuBackend(dev, arena_bytes=budget)
    out = {"resident_bytes": be.resident_bytes, "budget_bytes": budget, "mix": [0.4, 0.4, 0.2],
           "rates": []}
    # warm the kernels and the allocator once
    serving.Server(be).run(serving.make_requests(20, 1000, seed=7))
    for rate in rates:
        reqs = serving.make_requests(n_requests, rate, seed=int(rate))
        srv = serving.Server(be)
        t0 = time.perf_counter()
        done = srv.run(reqs)
        wall = time.perf_counter() - t0
        ok = [r for r in done if r.error is None]
        rows = sum(be.rows[r.shape] for r in ok)
        checked = sum(1 for r in ok if _same(r.result, _alone(be, r)))
        out["rates"].append({
            "offered_req_per_s": rate, "requests": len(done), "failed": len(done) - len(ok),
            "sustained_req_per_s": len(ok) / wall, "rows_per_s": rows / wall,
            **serving.latency_stats(done), "identical_to_alone": checked, "checked": len(ok)})
    return out
