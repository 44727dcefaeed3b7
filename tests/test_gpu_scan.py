"""Scan ingress on the GPU: column files into HBM (pinned staging, copies on a side stream,
widening casts on the device), into a GPUMemory slab or torch's allocator, equal to the oracle's
host-side scan; then a plan that scans and filters through the executor."""
import copy

import numpy as np
import pytest
import torch

from oracle import ref_ops as R
from test_scan import ARGS, SCHEMA, write_columns

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("slab", [False, True])
def test_scan_into_hbm_matches_the_oracle(device, tmp_path, slab):
    from tqp_amd import kernels
    from tqp_amd.IO.pinned_mem import GPUMemory
    from tqp_amd.operators.scan import tqp_scan
    write_columns(tmp_path, 300_000, seed=4)
    pool = GPUMemory(capacity_gb=0.1, block_size_mb=1, device=device) if slab else None
    got, want = {}, {}
    print(f"scan 6 columns of 300k rows, slab={slab}", flush=True)
    n = tqp_scan(True, got, ARGS, {'tensors path': str(tmp_path)}, SCHEMA, 1, pool)
    R.tqp_scan(want, ARGS, str(tmp_path), SCHEMA, 1)
    assert n == 300_000
    for k in range(1, 7):
        t = got[k].tensor
        assert t.is_cuda and t.dtype == want[k].tensor.dtype, k
        assert torch.equal(t.cpu(), want[k].tensor), k
        if slab:
            base = pool.arena.slab.data_ptr()
            assert base <= t.data_ptr() < base + pool.arena.nbytes
    assert got[1].is_sorted
    assert kernels.known_bounds(got[1].tensor) == (1, 300_000)
    assert kernels.known_bounds(got[4].tensor) == (int(want[4].tensor.min()), int(want[4].tensor.max()))


def test_scan_then_filter_through_the_executor(device, tmp_path):
    from tqp_amd import executor
    write_columns(tmp_path, 200_000, seed=5)
    plan = [(1, copy.deepcopy(ARGS)),
            (2, {'name': 'Filter', 'filter tree': [['t_day#4', '<', '1200'], 'AND',
                                                   ['t_qty#2', '>', '10']],
                 'materialized in': [4, 2], 'out mask id': 9})]
    gpu, cpu = {}, {}
    print("scan + filter", flush=True)
    executor.execute(copy.deepcopy(plan), gpu, config_json={'tensors path': str(tmp_path)},
                     schema={'T': SCHEMA}, scale_factor=1)
    R.tqp_scan(cpu, ARGS, str(tmp_path), SCHEMA, 1)
    R.execute(copy.deepcopy(plan[1:]), cpu)
    assert torch.equal(gpu[9].tensor.cpu(), cpu[9].tensor)
