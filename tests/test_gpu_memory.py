"""Device arena on the GPU: a query run with a `GPUMemory` slab places its outputs in the slab and
matches the oracle; a slab too small fails loudly; the served form admits requests against the
slab's real free bytes, refusing work that does not fit until earlier requests return their
reservations, and every result equals the same request run alone."""
import copy

import numpy as np
import pytest
import torch

from oracle import ref_ops as R

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-12


def _in_slab(t, arena):
    base = arena.slab.data_ptr()
    return base <= t.data_ptr() < base + arena.slab.numel()


def _assert_col(got, want, name, fp=False):
    g, w = got.detach().cpu(), want.detach().cpu()
    assert tuple(g.shape) == tuple(w.shape) and g.dtype == w.dtype, name
    if fp and w.dtype.is_floating_point:
        np.testing.assert_allclose(g.numpy(), w.numpy(), rtol=FP_RTOL, atol=0, err_msg=name)
    else:
        assert torch.equal(g, w), f"{name}: values differ"


@pytest.mark.parametrize("sync", [True, False])
def test_config1_pipeline_in_a_gpu_memory_slab(device, sync):
    from tqp_amd import executor, workloads
    from tqp_amd.IO.pinned_mem import GPUMemory
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()
    cpu = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), cpu)
    pool = GPUMemory(capacity_gb=2, block_size_mb=40, device=device)
    print(f"config1 reduced in a 2 GB slab, sync={sync}", flush=True)
    for rep in range(2):          # the second query reuses the slab after free_all
        gpu = workloads.to_tensor_group(tables, device)
        executor.execute(copy.deepcopy(plan), gpu, sync=sync, gpu_mem_pool=pool)
        torch.cuda.synchronize()
        assert _in_slab(gpu[301].tensor, pool.arena) and _in_slab(gpu[101].tensor, pool.arena)
        assert pool.get_utilization() > 0
        _assert_col(gpu[101].tensor, cpu[101].tensor, "c1 top keys (slab)")
        _assert_col(gpu[301].tensor, cpu[301].tensor, "c1 top revenue (slab)", fp=True)
        pool.free_all()
        assert pool.get_utilization() == 0


def test_slab_too_small_fails_loudly(device):
    from tqp_amd import executor, workloads
    from tqp_amd.IO.pinned_mem import GPUMemory
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    pool = GPUMemory(capacity_gb=0.0001, block_size_mb=0.04, device=device)   # 100 kB
    gpu = workloads.to_tensor_group(tables, device)
    print("config1 reduced in a 100 kB slab", flush=True)
    with pytest.raises(MemoryError):
        executor.execute(copy.deepcopy(workloads.config1_plan()), gpu, gpu_mem_pool=pool)
    torch.cuda.synchronize()


def test_served_admission_refuses_work_beyond_the_arena(device):
    from tqp_amd import serving, serving_bench
    probe = serving.Request(0, 1, {"segment": 0, "ship_day_min": 0}, 0.0)
    be0 = serving.GpuBackend(device, c2_scale=0.05, arena_bytes=1 << 20)
    fp1 = be0.footprint(probe)
    del be0
    # room for two join requests at a time, and a little more
    cap = int(fp1 * 2.5)
    be = serving.GpuBackend(device, c2_scale=0.05, arena_bytes=cap)
    reqs = serving.make_requests(24, 5000, seed=21, mix=(0.25, 0.75, 0.0))
    print(f"served: 24 requests in a {cap}-byte arena", flush=True)
    done = serving.Server(be, max_inflight=4).run(reqs)
    assert len(done) == 24
    failed = [r for r in done if r.error is not None]
    assert not failed, [(r.rid, r.shape, repr(r.error)) for r in failed]
    refused = [r for r in done if r.refused]
    print("served: refused at least once", len(refused), "peak", be.arena.peak, "of", cap,
          "per-shape peaks", be.peak, flush=True)
    assert refused                                   # admission held work back for memory
    assert be.arena.peak <= be.arena.nbytes and be.arena.used_bytes() == 0
    # at most two join requests were ever admitted together
    spans = sorted((r.admitted, r.done) for r in done if r.shape == 1)
    for a0, _ in spans:
        live = sum(1 for b0, b1 in spans if b0 <= a0 < b1)
        assert live <= 2
    for r in done:
        alone = serving_bench._alone(be, r)
        assert serving_bench._same(r.result, alone), f"request {r.rid} (shape {r.shape})"


def test_served_request_larger_than_the_arena_is_rejected(device):
    from tqp_amd import serving
    be = serving.GpuBackend(device, c2_scale=0.05, arena_bytes=16 << 20)    # 16 MiB
    reqs = serving.make_requests(12, 2000, seed=22, mix=(0.5, 0.5, 0.0))
    print("served: 12 requests in a 16 MiB arena", flush=True)
    done = serving.Server(be).run(reqs)
    for r in done:
        if r.shape == 1:
            assert isinstance(r.error, MemoryError), (r.rid, repr(r.error))
        else:
            assert r.error is None and r.result is not None, (r.rid, repr(r.error))
    assert be.arena.used_bytes() == 0
