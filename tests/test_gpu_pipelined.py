"""Queries queued back to back (executor.execute(..., wait=False)): the host prepares query i+1
while the device runs query i, and each query's counts are collected afterwards from its own
host word slot.  Every result must equal the same query run alone (and the oracle's), including
a query voided in the middle of the queue, which reruns per operator when it is collected."""
import copy

import numpy as np
import pytest
import torch

from oracle import ref_ops as R

pytestmark = pytest.mark.gpu


def _cols_equal(a, b, name):
    ga, gb = a.detach().cpu(), b.detach().cpu()
    assert ga.dtype == gb.dtype and tuple(ga.shape) == tuple(gb.shape), name
    if ga.dtype.is_floating_point:
        np.testing.assert_allclose(ga.numpy(), gb.numpy(), rtol=1e-12, atol=0, err_msg=name)
    else:
        assert torch.equal(ga, gb), name


def test_word_slots_post_and_wait(device):
    """Every slot posted before any is collected; one post past the slots raises; each wait
    returns its own words."""
    from tqp_amd import _lib as L, kernels
    nslots = L.load().tq_word_slots()
    src = torch.arange(nslots * 8, dtype=torch.int64, device=device).view(nslots, 8) * 3 + 1
    print(f"post {nslots} word reads", flush=True)
    tickets = [kernels.post_words([src[i].data_ptr(), src[i, 5:].data_ptr()], [40, 24])
               for i in range(nslots)]
    with pytest.raises(RuntimeError):
        kernels.post_words([src[0].data_ptr()], [8])
    want = src.cpu()
    for i, t in reversed(list(enumerate(tickets))):
        host = kernels.wait_words(t)
        assert [int(host[j]) for j in range(8)] == \
            want[i, :5].tolist() + want[i, 5:].tolist()
    t = kernels.post_words([src[1].data_ptr()], [64])
    assert [int(x) for x in kernels.wait_words(t)[:8]] == want[1].tolist()


def test_queued_queries_match_queries_run_alone(device):
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    base = workloads.to_tensor_group(tables, device)
    params = [(2, 1200), (0, 300), (4, 2400), (1, 0), (3, 1800), (2, 2556)]
    plans = [workloads.config1_plan(segment=s, ship_day_min=d) for s, d in params]
    alone = []
    for p in plans:
        tg = dict(base)
        n = executor.execute(copy.deepcopy(p), tg, sync=False)
        alone.append((tg, n))
    print("queue 6 queries, 3 in flight", flush=True)
    queued, pend = [], []
    for p in plans:
        tg = dict(base)
        q = executor.execute(copy.deepcopy(p), tg, sync=False, wait=False)
        assert isinstance(q, executor.PendingQuery)
        queued.append((tg, q))
        pend.append(q)
        if len(pend) == 3:
            pend.pop(0).result()
    for (tg, q), (tga, na) in zip(queued, alone):
        assert q.result() == na
        assert q.done()
        assert list(tg) == list(tga)
        for k in tga:
            _cols_equal(tg[k].tensor, tga[k].tensor, f"col {k}")
    # and the oracle for two of them
    for i in (0, 4):
        cpu = R.to_cols(tables)
        R.execute(copy.deepcopy(plans[i]), cpu)
        tg = queued[i][0]
        _cols_equal(tg[101].tensor, cpu[101].tensor, "top keys")
        _cols_equal(tg[301].tensor, cpu[301].tensor, "top revenue")


def test_voided_query_in_the_queue(device):
    """A query whose build keys repeat is voided by the fused kernels; collected between two
    queued queries over other tables, it reruns per operator and matches the oracle."""
    from tqp_amd import executor, workloads
    good = workloads.config1_tables(n_build=100_000, n_probe=400_000)
    bad = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    keys = bad[101][1].copy()
    keys[1::2] = keys[0::2][: keys[1::2].shape[0]]
    bad[101] = ('int', keys)
    plan = workloads.config1_plan()
    print("queue good, voided, good", flush=True)
    runs = []
    for tables in (good, bad, good):
        tg = workloads.to_tensor_group(tables, device)
        runs.append((tables, tg, executor.execute(copy.deepcopy(plan), tg, sync=False,
                                                  wait=False)))
    for tables, tg, q in runs:
        n = q.result()
        cpu = R.to_cols(tables)
        R.execute(copy.deepcopy(plan), cpu)
        _cols_equal(tg[101].tensor, cpu[101].tensor, "top keys")
        _cols_equal(tg[301].tensor, cpu[301].tensor, "top revenue")
        full = R.to_cols(tables)
        R.execute(copy.deepcopy(plan[:4]), full)
        assert n == full[101].tensor.shape[0]


def test_plans_without_a_fused_tail_complete_in_the_call(device):
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    tg = workloads.to_tensor_group(tables, device)
    plan = workloads.config1_plan()[:2]          # the two filters alone
    q = executor.execute(copy.deepcopy(plan), tg, sync=False, wait=False)
    assert q.done()
    cpu = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), cpu)
    _cols_equal(tg[10001].tensor, cpu[10001].tensor, "mask")
