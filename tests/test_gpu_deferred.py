"""A Join whose pairs feed the next Aggregate, run with one synchronisation (executor.execute
with sync=False): the results must equal the oracle's and the per-operator path's, whether the
deferral holds (unique build keys inside their bounds), is settled early (a group-by mode that
needs the row count on the host) or is voided and rerun (repeated build keys, keys outside the
bounds the build was given, no pairs at all).  These run the join and the group-by as separate
operators (fuse=False); tests/test_gpu_groupjoin.py covers them fused."""
import copy

import numpy as np
import pytest
import torch

from oracle import ref_ops as R

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-12


def _assert_col(got, want, name, fp=False):
    g = got.detach().cpu()
    w = want.detach().cpu()
    assert tuple(g.shape) == tuple(w.shape), f"{name}: shape {tuple(g.shape)} vs {tuple(w.shape)}"
    assert g.dtype == w.dtype, f"{name}: dtype {g.dtype} vs {w.dtype}"
    if fp and w.dtype.is_floating_point:
        np.testing.assert_allclose(g.numpy(), w.numpy(), rtol=FP_RTOL, atol=0, err_msg=name)
    else:
        assert torch.equal(g, w), f"{name}: values differ"


class _Outcomes:
    """Counts how deferrals end: 'held' (read back once, kept), 'early' (settled before the
    group-by), 'voided' (rerun per operator)."""

    def __init__(self, monkeypatch):
        from tqp_amd import kernels
        self.seen = []
        orig_read = kernels.Deferred._read
        outer = self

        def read(df, with_group):
            try:
                r = orig_read(df, with_group)
            except kernels.DeferredJoinFailed:
                outer.seen.append("voided")
                raise
            outer.seen.append("held" if with_group else "early")
            return r

        monkeypatch.setattr(kernels.Deferred, "_read", read)


def _run(tables, plan, device, stats=True, bounds=None):
    from tqp_amd import executor, kernels, workloads
    gpu = workloads.to_tensor_group(tables, device, stats=stats)
    for cid, (lo, hi) in (bounds or {}).items():
        kernels.note_bounds(gpu[cid].tensor, lo, hi)
    print("deferred run:", [a['name'] for _, a in plan], flush=True)
    n = executor.execute(copy.deepcopy(plan), gpu, sync=False, fuse=False)
    torch.cuda.synchronize()
    return gpu, n


def _check(gpu, cpu, n, plan, keys=(101,), fp=(301,)):
    for k in keys:
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"col {k}")
    for k in fp:
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"col {k}", fp=True)
    if plan[-1][1]['name'] == 'Aggregate':
        # the aggregate returns the row count of the group's first column
        assert n == next(iter(cpu.values())).tensor.shape[0]


def _oracle(tables, plan):
    cpu = R.to_cols(tables)
    n = R.execute(copy.deepcopy(plan), cpu)
    return cpu, n


@pytest.mark.parametrize("upto", [4, 5])
def test_deferred_join_aggregate_matches_oracle(device, monkeypatch, upto):
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()[:upto]
    seen = _Outcomes(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["held"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)
    if upto == 5:
        want = R.to_cols(tables)
        R.execute(copy.deepcopy(plan[:4]), want)
        assert n == want[101].tensor.shape[0]     # the sort's rows before its limit
    if upto == 4:
        # the pair columns the aggregate did not replace are cut to the pair count
        for k in (201, 202, 203):
            _assert_col(gpu[k].tensor, cpu[k].tensor, f"pair col {k}", fp=True)


def test_deferred_top_rows_fewer_than_limit(device, monkeypatch):
    """Five selected build rows: fewer groups than the top-k limit, so the deferred top-k's
    permutation is cut to the group count."""
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    seg = np.zeros(200_000, dtype=np.int8)
    seg[[7, 1000, 50_000, 123_456, 199_999]] = 2
    tables[102] = ('string', seg)
    plan = workloads.config1_plan()
    seen = _Outcomes(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["held"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)
    assert gpu[101].tensor.shape[0] < 10
    want = R.to_cols(tables)
    R.execute(copy.deepcopy(plan[:4]), want)
    assert n == want[101].tensor.shape[0]     # the sort's rows before its limit


def test_deferred_same_as_per_operator_path(device):
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=100_000, n_probe=400_000)
    plan = workloads.config1_plan()
    a = workloads.to_tensor_group(tables, device)
    b = workloads.to_tensor_group(tables, device)
    na = executor.execute(copy.deepcopy(plan), a, sync=False, fuse=False)
    nb = executor.execute(copy.deepcopy(plan), b, sync=True)
    assert na == nb
    assert sorted(a) == sorted(b)
    for k in a:
        _assert_col(a[k].tensor, b[k].tensor, f"col {k}", fp=True)
        assert a[k].tensor_type == b[k].tensor_type


def test_deferred_voided_by_repeated_build_keys(device, monkeypatch):
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    keys = tables[101][1].copy()
    keys[1::2] = keys[0::2][: keys[1::2].shape[0]]      # every key twice
    tables[101] = ('int', keys)
    plan = workloads.config1_plan()[:4]
    seen = _Outcomes(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["voided"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)


def test_deferred_voided_by_keys_outside_bounds(device, monkeypatch):
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    plan = workloads.config1_plan()[:4]
    seen = _Outcomes(monkeypatch)
    # bounds that leave out the upper half of the build keys
    gpu, n = _run(tables, plan, device, bounds={101: (1, 25_000)})
    assert seen.seen == ["voided"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)


def test_deferred_without_column_statistics(device, monkeypatch):
    """The build measures its keys (one synchronisation), the probe is still deferred."""
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=60_000, n_probe=240_000)
    plan = workloads.config1_plan()
    seen = _Outcomes(monkeypatch)
    gpu, n = _run(tables, plan, device, stats=False)
    assert seen.seen == ["held"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)


def test_deferred_settled_early_for_few_groups(device, monkeypatch):
    """40 build keys: the group-by keeps its table in LDS, which needs the row count on the
    host, so the pairs are settled before it runs."""
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=40, n_probe=300_000)
    plan = workloads.config1_plan()[:4]
    seen = _Outcomes(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["early"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _check(gpu, cpu, n, plan)


def test_deferred_without_pairs_matches_per_operator_path(device, monkeypatch):
    """No build row selected: the deferral is voided and the per-operator path gives what the
    reference gives for a join without pairs (empty columns, which the aggregate's pipeline
    passes through as empty outputs)."""
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=20_000, n_probe=80_000)
    plan = workloads.config1_plan(segment=9)[:4]
    a = workloads.to_tensor_group(tables, device)
    na = executor.execute(copy.deepcopy(plan), a, sync=True)
    seen = _Outcomes(monkeypatch)
    b = workloads.to_tensor_group(tables, device)
    nb = executor.execute(copy.deepcopy(plan), b, sync=False, fuse=False)
    assert seen.seen == ["voided"], seen.seen
    assert na == nb
    assert sorted(a) == sorted(b)
    for k in a:
        _assert_col(b[k].tensor, a[k].tensor, f"col {k}")
