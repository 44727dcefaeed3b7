"""The Join + Aggregate of a plan fused into one groupjoin (executor.execute with sync=False)
when the aggregate groups by the join key alone: results must equal the oracle's (keys, counts
and integer sums bit-exact, float sums within 1e-12), the join's output columns must hold the
oracle's pairs when read, and every case the fused kernels cannot settle (repeated or
out-of-bounds build keys, no pairs) must rerun per operator with the same results."""
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


class _Seen:
    """Records how each fused attempt ended: 'fused' (settled), 'voided' (rerun per operator),
    'declined' (not fusable, nothing launched)."""

    def __init__(self, monkeypatch):
        from tqp_amd import executor, groupjoin, kernels
        self.seen = []
        orig_run = groupjoin.run
        orig_resolve = groupjoin.GroupJoinDeferred.resolve
        outer = self

        def run(*a, **k):
            try:
                return orig_run(*a, **k)
            except kernels.DeferredJoinFailed:
                outer.seen.append("declined")
                raise

        def resolve(df):
            if df.G is not None:
                return df.m, df.G
            try:
                r = orig_resolve(df)
            except kernels.DeferredJoinFailed:
                outer.seen.append("voided")
                raise
            outer.seen.append("fused")
            return r

        monkeypatch.setattr(groupjoin, "run", run)
        monkeypatch.setattr(groupjoin.GroupJoinDeferred, "resolve", resolve)


def _run(tables, plan, device, stats=True, bounds=None, fuse=True):
    from tqp_amd import executor, kernels, workloads
    gpu = workloads.to_tensor_group(tables, device, stats=stats)
    for cid, (lo, hi) in (bounds or {}).items():
        kernels.note_bounds(gpu[cid].tensor, lo, hi)
    print("groupjoin run:", [a['name'] for _, a in plan], flush=True)
    n = executor.execute(copy.deepcopy(plan), gpu, sync=False, fuse=fuse)
    torch.cuda.synchronize()
    return gpu, n


def _oracle(tables, plan):
    """The oracle's columns and the row count the plan's last operator returns (an aggregate's:
    its group's first column)."""
    cpu = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), cpu)
    return cpu, next(iter(cpu.values())).tensor.shape[0]


def test_config1_all_groups_fused_full_size(device, monkeypatch):
    """Build 1.5M / probe 6M rows: every one of the ~264k groups, and the row count."""
    from tqp_amd import workloads
    tables = workloads.config1_tables()
    plan = workloads.config1_plan()[:4]
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    cpu, ncpu = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "group keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "revenue", fp=True)
    assert n == ncpu


def test_config1_top10_fused_full_size(device, monkeypatch):
    from tqp_amd import workloads
    tables = workloads.config1_tables()
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue", fp=True)
    want = R.to_cols(tables)
    R.execute(copy.deepcopy(plan[:4]), want)
    assert n == want[101].tensor.shape[0]      # the sort's rows before its limit


def test_join_outputs_computed_when_read(device, monkeypatch):
    """The join's output columns the aggregate did not replace are the oracle's join outputs
    when a later reader asks for them; the fused call itself never wrote them."""
    from tqp_amd import groupjoin, workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()[:4]
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    for k in (201, 202, 203):
        assert isinstance(gpu[k], groupjoin.LazyJoinColumn) and not gpu[k].materialized
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "group keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "revenue", fp=True)
    pairs = R.to_cols(tables)
    R.execute(copy.deepcopy(plan[:3]), pairs)
    for k in (201, 202, 203):
        _assert_col(gpu[k].tensor, pairs[k].tensor, f"join output {k}", fp=True)
        assert gpu[k].tensor_type == pairs[k].tensor_type
    assert gpu[201].rows == pairs[201].tensor.shape[0]


def test_fused_same_as_per_operator_path(device):
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=100_000, n_probe=400_000)
    plan = workloads.config1_plan()
    a = workloads.to_tensor_group(tables, device)
    b = workloads.to_tensor_group(tables, device)
    na = executor.execute(copy.deepcopy(plan), a, sync=False)
    nb = executor.execute(copy.deepcopy(plan), b, sync=True)
    assert na == nb
    assert list(a) == list(b)                     # same columns, same insertion order
    for k in a:
        _assert_col(a[k].tensor, b[k].tensor, f"col {k}", fp=True)
        assert a[k].tensor_type == b[k].tensor_type


def _multi_agg_tables(g, nb, np_, key_dtype=np.int32, lo=1):
    keys = (g.permutation(nb) + lo).astype(key_dtype)
    return {
        101: ('int', keys),
        102: ('string', g.integers(0, 5, nb, dtype=np.int8)),
        201: ('int', g.integers(lo, lo + nb, np_).astype(key_dtype)),
        202: ('float', g.uniform(900.0, 105000.0, np_)),
        203: ('float', g.integers(0, 11, np_) / 100.0),
        204: ('date', g.integers(0, 2557, np_, dtype=np.int32)),
        205: ('int', g.integers(-50, 50, np_, dtype=np.int32)),
    }


def _multi_agg_plan(group_key='key#101', key_id=101, extra=None):
    results = [group_key,
               [[[['sum', [['price#202', '*', ['1.0', '-', 'discount#203']]]], '#', '40'], 'AS', 'rev#301']],
               [[[['count', ['1']], '#', '41'], 'AS', 'cnt#302']],
               [[[['avg', ['discount#203']], '#', '42'], 'AS', 'avgd#303']],
               [[[['min', ['price#202']], '#', '43'], 'AS', 'minp#304']],
               [[[['max', ['q#205']], '#', '44'], 'AS', 'maxq#305']],
               [[[['sum', ['q#205']], '#', '45'], 'AS', 'sumq#306']]]
    if extra:
        results += extra
    return [
        (1, {'name': 'Filter', 'filter tree': [['segment#102', '=', '2']],
             'materialized in': [102], 'out mask id': 10001}),
        (2, {'name': 'Filter', 'filter tree': [['ship_day#204', '>', '1200']],
             'materialized in': [204], 'out mask id': 10002}),
        (3, {'name': 'Join', 'build': 'left', 'left group': [101, 102],
             'right group': [201, 202, 203, 204, 205], 'left keys': [101], 'right keys': [201],
             'mask map': {101: 10001, 201: 10002, 202: 10002, 203: 10002, 205: 10002},
             'materialized in': [101, 201, 202, 203, 205], 'join type': 'Inner'}),
        (4, {'name': 'Aggregate', 'input ids': [key_id, 202, 203, 205], 'keys': [group_key],
             'results': results,
             'out agg ids': [301, 302, 303, 304, 305, 306] + ([307] if extra else []),
             'is partial': False}),
    ]


@pytest.mark.parametrize("group_key,key_id", [("key#101", 101), ("fk#201", 201)])
@pytest.mark.parametrize("key_dtype", [np.int32, np.int64])
def test_several_aggregates_either_join_key(device, monkeypatch, group_key, key_id, key_dtype):
    """sum of an expression, count(*), avg, min, max and an integer sum, grouped by the build or
    the probe key, int32 and int64 keys (the int64 keys start past 2^40)."""
    g = np.random.default_rng(71)
    lo = 1 if key_dtype == np.int32 else (1 << 40) + 5
    tables = _multi_agg_tables(g, 150_000, 700_000, key_dtype, lo)
    plan = _multi_agg_plan(group_key, key_id)
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    cpu, ncpu = _oracle(tables, plan)
    _assert_col(gpu[key_id].tensor, cpu[key_id].tensor, "group keys")
    for k in (302, 305, 306):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"agg {k}")
    for k in (301, 303, 304):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"agg {k}", fp=True)
    assert n == ncpu


def test_counts_of_constants_come_from_pair_counts(device, monkeypatch):
    """count(1) (twice) next to a sum of an expression and a count of a column: the counts of a
    constant have no record word (they are the slots' pair counts), the others keep theirs, and
    every column equals the oracle's."""
    from tqp_amd import groupjoin
    g = np.random.default_rng(77)
    tables = _multi_agg_tables(g, 120_000, 500_000)
    results = ['key#101',
               [[[['sum', [['price#202', '*', ['1.0', '-', 'discount#203']]]], '#', '40'], 'AS', 'rev#301']],
               [[[['count', ['1']], '#', '41'], 'AS', 'cnt#302']],
               [[[['count', ['discount#203']], '#', '42'], 'AS', 'cntd#303']],
               [[[['count', ['1']], '#', '43'], 'AS', 'cnt2#304']]]
    plan = _multi_agg_plan()
    plan[3][1]['input ids'] = [101, 202, 203]
    plan[3][1]['results'] = results
    plan[3][1]['out agg ids'] = [301, 302, 303, 304]
    groupjoin._REC_WORDS.clear()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    assert list(groupjoin._REC_WORDS.values()) == [3]   # slot, revenue, count(discount)
    cpu, ncpu = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "group keys")
    for k in (302, 303, 304):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"agg {k}")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "agg 301", fp=True)
    assert n == ncpu


def test_case_when_argument(device, monkeypatch):
    """A CASE WHEN argument (aggregate.py:74-85) over probe columns, fused."""
    g = np.random.default_rng(72)
    tables = _multi_agg_tables(g, 60_000, 300_000)
    extra = [[[[['sum', [['CASE', 'WHEN', ['q#205', '>', '0'], 'THEN', 'price#202', 'ELSE',
                          '0.0', 'END']]], '#', '46'], 'AS', 'pos#307']]]
    plan = _multi_agg_plan(extra=extra)
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "group keys")
    _assert_col(gpu[307].tensor, cpu[307].tensor, "case sum", fp=True)


def test_voided_by_repeated_build_keys(device, monkeypatch):
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    keys = tables[101][1].copy()
    keys[1::2] = keys[0::2][: keys[1::2].shape[0]]      # every key twice
    tables[101] = ('int', keys)
    plan = workloads.config1_plan()[:4]
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["voided"], seen.seen
    cpu, ncpu = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "group keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "revenue", fp=True)
    assert n == ncpu


def test_voided_by_keys_outside_bounds(device, monkeypatch):
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=50_000, n_probe=200_000)
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device, bounds={101: (1, 25_000)})
    assert seen.seen == ["voided"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue", fp=True)


def test_without_column_statistics(device, monkeypatch):
    """No recorded bounds: the fused call measures the build key column first."""
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=60_000, n_probe=240_000)
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device, stats=False)
    assert seen.seen == ["fused"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue", fp=True)


def test_without_pairs_matches_per_operator_path(device, monkeypatch):
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=20_000, n_probe=80_000)
    plan = workloads.config1_plan(segment=9)[:4]
    a = workloads.to_tensor_group(tables, device)
    na = executor.execute(copy.deepcopy(plan), a, sync=True)
    seen = _Seen(monkeypatch)
    b = workloads.to_tensor_group(tables, device)
    nb = executor.execute(copy.deepcopy(plan), b, sync=False)
    assert seen.seen == ["voided"], seen.seen
    assert na == nb
    assert sorted(a) == sorted(b)
    for k in a:
        _assert_col(b[k].tensor, a[k].tensor, f"col {k}")


def test_few_groups_and_fewer_than_limit(device, monkeypatch):
    """Five selected build rows: fewer groups than the top-k limit."""
    from tqp_amd import workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    seg = np.zeros(200_000, dtype=np.int8)
    seg[[7, 1000, 50_000, 123_456, 199_999]] = 2
    tables[102] = ('string', seg)
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue", fp=True)
    assert gpu[101].tensor.shape[0] < 10


def test_key_range_too_wide_declines(device, monkeypatch):
    """Build keys spread over 2^27 values need more reduce slices than the kernels take: the
    pair is not fused and runs as before."""
    g = np.random.default_rng(73)
    nb, np_ = 40_000, 160_000
    keys = g.choice(1 << 27, nb, replace=False).astype(np.int32)
    tables = {101: ('int', keys), 102: ('string', g.integers(0, 5, nb, dtype=np.int8)),
              103: ('date', g.integers(0, 2557, nb, dtype=np.int32)),
              201: ('int', keys[g.integers(0, nb, np_)]),
              202: ('float', g.uniform(900.0, 105000.0, np_)),
              203: ('float', g.integers(0, 11, np_) / 100.0),
              204: ('date', g.integers(0, 2557, np_, dtype=np.int32))}
    from tqp_amd import workloads
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["declined"], seen.seen
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue", fp=True)


def test_epochs_wrap_around(device):
    """More calls than a presence byte has epochs (255) on one stream: the array is zeroed and
    reused; every call stays exact.  The build filter changes every call, so a stale stamp
    would add groups."""
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=4_000, n_probe=16_000)
    base = workloads.to_tensor_group(tables, device)
    wants = {}
    for seg in range(5):
        cpu, _ = _oracle(tables, workloads.config1_plan(segment=seg)[:4])
        wants[seg] = (cpu[101].tensor, cpu[301].tensor)
    print("epoch wrap: 270 fused calls", flush=True)
    for i in range(270):
        seg = i % 5
        tg = dict(base)
        executor.execute(workloads.config1_plan(segment=seg)[:4], tg, sync=False)
        if i % 27 == 0 or i >= 250:
            _assert_col(tg[101].tensor, wants[seg][0], f"call {i} keys")
            _assert_col(tg[301].tensor, wants[seg][1], f"call {i} revenue", fp=True)


def test_filters_fused_into_the_groupjoin(device, monkeypatch):
    """Both filters are conjunctions of comparisons: the groupjoin evaluates them in its own
    kernels and their masks are never computed; read afterwards, a mask is the oracle's."""
    from tqp_amd import workloads
    from tqp_amd.operators.filter import LazyMask
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    for mid in (10001, 10002):
        assert isinstance(gpu[mid], LazyMask) and not gpu[mid].materialized, mid
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys (fused filters)")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue (fused filters)", fp=True)
    masks = R.to_cols(tables)
    R.execute(copy.deepcopy(plan[:2]), masks)
    for mid in (10001, 10002):
        _assert_col(gpu[mid].tensor, masks[mid].tensor, f"mask {mid} read later")
        assert gpu[mid].tile_counts is not None


def test_filter_that_is_not_a_conjunction_is_computed(device, monkeypatch):
    """An OR on the probe side cannot be fused: its mask is computed and the groupjoin reads it;
    the build side's conjunction is still fused."""
    from tqp_amd import workloads
    from tqp_amd.operators.filter import LazyMask
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()
    plan[1][1]['filter tree'] = [['ship_day#204', '>', '1900'], 'OR', ['ship_day#204', '<', '300']]
    seen = _Seen(monkeypatch)
    gpu, n = _run(tables, plan, device)
    assert seen.seen == ["fused"], seen.seen
    assert isinstance(gpu[10001], LazyMask) and not gpu[10001].materialized
    assert gpu[10002].materialized if isinstance(gpu[10002], LazyMask) else True
    cpu, _ = _oracle(tables, plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "top keys (OR filter)")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "top revenue (OR filter)", fp=True)
