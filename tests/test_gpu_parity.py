"""Parity of the HIP operators with the CPU oracle on the same seeded inputs (GPU).

Bar: bit-exact masks, compacted columns, join outputs (canonical pair order), group keys,
counts, integer sums and sort orders; float64 sums and averages within 1e-12 relative of the
oracle (whose own sums are pinned to sequential sums in tests/test_oracle.py).
"""
import copy

import numpy as np
import pytest
import torch

from oracle import closed_form as CF
from oracle import ref_ops as R

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-12


def _gpu_group(tables, dev, stats=True):
    from tqp_amd import workloads
    return workloads.to_tensor_group(tables, dev, stats=stats)


def _assert_col(got, want, name, fp=False):
    g = got.detach().cpu()
    w = want.detach().cpu()
    assert tuple(g.shape) == tuple(w.shape), f"{name}: shape {tuple(g.shape)} vs {tuple(w.shape)}"
    if fp and w.dtype.is_floating_point:
        assert g.dtype == w.dtype, f"{name}: dtype {g.dtype} vs {w.dtype}"
        np.testing.assert_allclose(g.numpy(), w.numpy(), rtol=FP_RTOL, atol=0, err_msg=name)
    else:
        assert g.dtype == w.dtype, f"{name}: dtype {g.dtype} vs {w.dtype}"
        assert torch.equal(g, w), f"{name}: values differ"


# ------------------------------------------------------------------ filter + compaction
PREDICATES = [
    [[['isnotnull', ['ship_day#1']], 'AND', ['ship_day#1', '<=', '2400']]],
    ['ship_day#1', '>', '1200'],
    ['ship_day#1', '<', '1200.5'],
    [['price#3', '>=', '50000.0'], 'AND', ['discount#4', '=', '0.05']],
    [['flag_a#6', '=', '2'], 'OR', ['NOT', ['tax#5', '<=', '0.04']]],
    ['flag_a#6', 'IN', ['0', '2']],
    [['quantity#2', '<', '25'], 'AND', ['ship_day#1', '>=', '1995-01-01']],
    [['ship_day#1', '>=', '100'], 'AND', ['ship_day#1', '<', '2000']],
    [['flag_b#7', '=', '1'], 'AND', ['price#3', '<', '60000.0']],
    ['tax#5', '>=', '0.04'],
    [[['isnotnull', ['flag_a#6']], 'AND', ['flag_a#6', '>', '0']], 'AND', ['quantity#2', '>=', '10']],
]


def test_filter_f32_narrow_int_columns(device):
    """Comparison conjunctions over float32, int16 and bool columns, with a ragged tail."""
    from tqp_amd import kernels
    g = np.random.default_rng(11)
    n = 70_003
    tables = {1: ('float', g.uniform(-5, 5, n).astype(np.float32)),
              2: ('int', g.integers(-300, 300, n, dtype=np.int16)),
              3: ('bool', g.random(n) < 0.5)}
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    for tree in ([['f#1', '<', '1.5'], 'AND', ['s#2', '>=', '-20']],
                 [['f#1', '>', '-2'], 'AND', [['s#2', '>', '7'], 'AND', ['s#2', '<=', '250']]],
                 ['s#2', '=', '3']):
        mask, counts = kernels.filter_mask(tree, gpu, None)
        want = R.filter_evaluate(cpu, tree, None, None)
        _assert_col(mask, want, f"mask {tree}")
        assert int(counts.sum().item()) == int(want.sum().item())


def test_filter_like_bit_exact(device):
    """'%w1%w2%' over a 2-D string column (ragged row count), alone, negated and in a
    conjunction, with words planted in both orders and overlapping."""
    from tqp_amd import kernels
    g = np.random.default_rng(23)
    n, w = 200_003, 40
    s = np.array([ord(c) for c in "ab q"], dtype=np.int8)[g.integers(0, 4, (n, w))]
    s[g.random((n, w)) < 0.05] = 0
    for word, frac in (("special", 0.3), ("requests", 0.3), ("specialrequests", 0.05)):
        rows = np.flatnonzero(g.random(n) < frac)
        at = g.integers(0, w - len(word) + 1, rows.shape[0])
        for r, p in zip(rows, at):
            s[r, p:p + len(word)] = [ord(c) for c in word]
    tables = {4: ('string', s), 5: ('int', g.integers(0, 100, n, dtype=np.int32))}
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    sums = []
    for tree in (['c#4', 'LIKE', '%special%requests%'],
                 ['NOT', ['c#4', 'LIKE', '%special%requests%']],
                 [['c#4', 'LIKE', '%q%ab%'], 'AND', ['k#5', '<', '50']],
                 ['c#4', 'LIKE', '%%aa%'],
                 ['c#4', 'LIKE', '%a%' + 'b' * 40 + '%']):
        print("like:", tree, flush=True)
        mask, counts = kernels.filter_mask(tree, gpu, None)
        want = R.filter_evaluate(cpu, tree, None, None)
        _assert_col(mask, want, f"mask {tree}")
        assert int(counts.sum().item()) == int(want.sum().item())
        sums.append(int(want.sum()))
    assert 0 < sums[0] < n and 0 < sums[3] < n and sums[4] == 0


@pytest.mark.parametrize("pi", range(len(PREDICATES)))
def test_filter_mask_bit_exact(device, pi):
    from tqp_amd import kernels, workloads
    tables = workloads.config0_tables(n=300_001)
    tree = PREDICATES[pi]
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    mask, counts = kernels.filter_mask(tree, gpu, None)
    want = R.filter_evaluate(cpu, tree, None, None)
    _assert_col(mask, want, f"mask {pi}")
    assert int(counts.sum().item()) == int(want.sum().item())


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 1_000_003])
def test_compaction_bit_exact_and_ordered(device, n):
    from tqp_amd import kernels
    g = np.random.default_rng(n)
    m = g.random(n) < 0.37
    a = g.integers(-100, 100, n, dtype=np.int8)
    b = g.random(n)
    c = g.integers(0, 1 << 40, n, dtype=np.int64)
    s = g.integers(0, 127, (n, 5), dtype=np.int8)
    tm = torch.from_numpy(m).to(device)
    outs = kernels.compact(tm, [torch.from_numpy(x).to(device) for x in (a, b, c, s)])
    for x, o in zip((a, b, c, s), outs):
        assert np.array_equal(o.cpu().numpy(), CF.compact(m, x))
    rows = kernels.compact_rows(tm)
    assert np.array_equal(rows.cpu().numpy(), np.flatnonzero(m))


@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 300_001, 5_000_000])
def test_tile_scan_exact(device, n):
    # one chunk, two-level and ragged scans of per-tile counts into int64 offsets
    from tqp_amd import kernels
    g = np.random.default_rng(n)
    c = g.integers(0, 4097, n, dtype=np.int64).astype(np.int32)
    offs, total = kernels.scan_tiles(torch.from_numpy(c).to(device), n)
    want = np.concatenate([[0], np.cumsum(c.astype(np.int64))])
    assert np.array_equal(offs.cpu().numpy(), want[:-1])
    assert int(total.item()) == int(want[-1])


def test_compaction_empty_and_full(device):
    from tqp_amd import kernels
    for m in (np.zeros(10000, bool), np.ones(10000, bool)):
        v = np.arange(10000, dtype=np.int32)
        o = kernels.compact(torch.from_numpy(m).to(device), [torch.from_numpy(v).to(device)])[0]
        assert np.array_equal(o.cpu().numpy(), v[m])


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("n", [5, 4097, 300_000])
def test_sort_multikey_bit_exact(device, n):
    from tqp_amd import kernels
    g = np.random.default_rng(n)
    a = g.integers(0, 4, n, dtype=np.int8)
    f = np.round(g.normal(size=n), 2)
    f[g.random(n) < 0.05] = 0.0
    f[g.random(n) < 0.05] = -0.0
    f[g.random(n) < 0.02] = np.nan
    f[g.random(n) < 0.01] = np.inf
    i = g.integers(-(1 << 40), 1 << 40, n, dtype=np.int64)
    for descs in ([False, True, False], [True, False, True], [False, False, False]):
        keys = [(torch.from_numpy(x).to(device), d) for x, d in zip((a, f, i), descs)]
        perm = kernels.sort_perm(keys, n)
        assert np.array_equal(perm.cpu().numpy(), CF.sort_perm([a, f, i], descs)), descs
    perm = kernels.sort_perm([(torch.from_numpy(f).to(device), True)], n, limit=10)
    assert np.array_equal(perm.cpu().numpy(), CF.sort_perm([f], [True])[:10])


@pytest.mark.parametrize("n,k", [(4097, 1), (70_000, 10), (70_000, 64), (1_000_000, 64),
                                 (1_000_000, 10), (300_000, 65), (300_000, 256)])
def test_sort_limit_bit_exact(device, n, k):
    # limits up to 64 take the wave-list rounds, 65..256 the radix-select rounds; heavy ties,
    # NaN/-0.0 and two packed keys check the stable (key, row) order of the first k
    from tqp_amd import kernels
    g = np.random.default_rng(n + k)
    f = np.round(g.normal(size=n), 1)
    f[g.random(n) < 0.05] = -0.0
    f[g.random(n) < 0.02] = np.nan
    a = g.integers(-3, 3, n, dtype=np.int32)
    b = g.integers(0, 2, n, dtype=np.int8)
    for keys, descs in (([f], [True]), ([f], [False]), ([a, b], [True, False]),
                        ([b, a], [False, False])):
        dk = [(torch.from_numpy(x).to(device), d) for x, d in zip(keys, descs)]
        perm = kernels.sort_perm(dk, n, limit=k)
        assert np.array_equal(perm.cpu().numpy(), CF.sort_perm(keys, descs)[:k]), (descs, k)


def test_sort_2d_string_and_int32_keys(device):
    from tqp_amd import kernels
    g = np.random.default_rng(7)
    n = 20000
    s = g.integers(65, 70, (n, 3), dtype=np.int8)
    k = g.integers(-5, 5, n, dtype=np.int32)
    perm = kernels.sort_perm([(torch.from_numpy(s).to(device), False),
                              (torch.from_numpy(k).to(device), True)], n)
    assert np.array_equal(perm.cpu().numpy(), CF.sort_perm([s, k], [False, True]))


# ------------------------------------------------------------------ join
def _run_join(device, bk, bm, bv, pk, pm, pv, kind='inner', conds=(), stats=True, bounds=None):
    # stats: load-time column bounds (the build skips its statistics pass); bounds: override
    # the build key's recorded bounds
    from tqp_amd import kernels
    from tqp_amd.operators import hashjoin
    tables = {1: ('int', bk), 2: ('int', bv), 3: ('int', pk), 4: ('float', pv),
              901: ('bool', bm), 902: ('bool', pm)}
    args = {'left group': [1, 2], 'right group': [3, 4], 'materialized in': [1, 2, 3, 4],
            'mask map': {1: 901, 2: 901, 3: 902, 4: 902}}
    gpu = _gpu_group(tables, device, stats=stats)
    if bounds is not None:
        kernels.note_bounds(gpu[1].tensor, *bounds)
    cpu = R.to_cols(tables)
    hashjoin.join_vortex(True, gpu, copy.deepcopy(args), 1, 3, list(conds), kind, None, None, "j")
    R.join_vortex(cpu, copy.deepcopy(args), 1, 3, list(conds), kind)
    return gpu, cpu


@pytest.mark.parametrize("stats", [True, False], ids=["bounded", "stats_pass"])
@pytest.mark.parametrize("case", ["unique", "dups", "skew", "empty_build", "empty_probe", "int64",
                                  "hash_dups", "int64_max", "dense_int64_dups", "narrow_build",
                                  "wide_probe"])
def test_join_inner_bit_exact(device, case, stats):
    # dense directory: unique, dups, skew, dense_int64_dups, narrow_build, wide_probe;
    # hash slots: int64 (unique keys), hash_dups, int64_max (largest int64 key present)
    g = np.random.default_rng(sum(map(ord, case)))
    nb, np_ = 50_000, 200_000
    if case == "hash_dups":
        pool = g.integers(-(1 << 50), 1 << 50, 5000, dtype=np.int64)
        bk = pool[g.integers(0, 5000, nb)]
        pk = np.where(g.random(np_) < 0.7, pool[g.integers(0, 5000, np_)], g.integers(0, 1 << 50, np_))
    elif case == "int64_max":
        bk = g.integers(-(1 << 60), 1 << 60, nb, dtype=np.int64)
        bk[:100] = np.iinfo(np.int64).max
        pk = np.where(g.random(np_) < 0.5, bk[g.integers(0, nb, np_)], g.integers(0, 1 << 60, np_))
    elif case == "dense_int64_dups":
        bk = g.integers(-3000, 3000, nb, dtype=np.int64) + (1 << 40)
        pk = g.integers(-3500, 3500, np_, dtype=np.int64) + (1 << 40)
    elif case == "narrow_build":
        bk = g.integers(-30000, 30000, nb, dtype=np.int16)
        pk = g.integers(-40000, 40000, np_, dtype=np.int32)
    elif case == "wide_probe":
        bk = g.permutation(np.arange(-nb // 2, nb // 2, dtype=np.int32))
        pk = np.where(g.random(np_) < 0.5, g.integers(-nb // 2, nb // 2, np_),
                      g.integers(-(1 << 40), 1 << 40, np_)).astype(np.int64)
    elif case == "unique":
        bk = g.permutation(np.arange(1, nb + 1, dtype=np.int32))
        pk = g.integers(1, nb + 1, np_, dtype=np.int32)
    elif case == "dups":
        bk = g.integers(0, 5000, nb, dtype=np.int32)
        pk = g.integers(0, 6000, np_, dtype=np.int32)
    elif case == "skew":
        bk = np.where(g.random(nb) < 0.3, 7, g.integers(0, 100000, nb)).astype(np.int32)
        pk = np.where(g.random(np_) < 0.05, 7, g.integers(0, 100000, np_)).astype(np.int32)
    elif case == "int64":
        bk = g.integers(-(1 << 50), 1 << 50, nb, dtype=np.int64)
        pk = np.where(g.random(np_) < 0.5, bk[g.integers(0, nb, np_)], g.integers(0, 1 << 50, np_))
    else:
        bk = g.integers(0, 1000, nb, dtype=np.int32)
        pk = g.integers(0, 1000, np_, dtype=np.int32)
    bm = g.random(nb) < (0.0 if case == "empty_build" else 0.2)
    pm = g.random(np_) < (0.0 if case == "empty_probe" else 0.53)
    bv = g.integers(0, 1 << 40, nb, dtype=np.int64)
    pv = g.random(np_)
    gpu, cpu = _run_join(device, bk, bm, bv, pk, pm, pv, stats=stats)
    for k in (1, 2, 3, 4):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"{case} col {k}")


@pytest.mark.parametrize("case", ["narrow_unique", "narrow_dups", "narrow_int64", "wide_hash"])
def test_join_bounds_that_do_not_hold(device, case):
    # recorded bounds that miss some build keys: the first count finds the keys outside them,
    # and the index is built again from the keys' own statistics
    g = np.random.default_rng(sum(map(ord, case)) + 5)
    nb, np_ = 40_000, 120_000
    if case == "narrow_unique":
        bk = g.permutation(np.arange(1, nb + 1, dtype=np.int32))
        bounds = (100, nb - 100)
    elif case == "narrow_dups":
        bk = g.integers(0, 3000, nb, dtype=np.int32)
        bounds = (10, 2000)
    elif case == "narrow_int64":
        bk = g.integers(-(1 << 50), 1 << 50, nb, dtype=np.int64)
        bounds = (-(1 << 49), 1 << 49)
    else:
        bk = g.permutation(np.arange(1, nb + 1, dtype=np.int32)) * 1000
        bounds = (0, 500 * nb)
    pk = np.where(g.random(np_) < 0.6, bk[g.integers(0, nb, np_)], g.integers(0, int(bk.max()), np_))
    pk = pk.astype(bk.dtype)
    bm = g.random(nb) < 0.5
    pm = g.random(np_) < 0.5
    gpu, cpu = _run_join(device, bk, bm, g.integers(0, 99, nb, dtype=np.int64), pk, pm,
                         g.random(np_), bounds=bounds)
    for k in (1, 2, 3, 4):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"{case} col {k}")


@pytest.mark.parametrize("kind", ["right-semi", "right-anti", "right-outer"])
def test_join_other_kinds_bit_exact(device, kind):
    g = np.random.default_rng(3)
    nb, np_ = 3000, 9000
    bk = g.integers(0, 2000, nb, dtype=np.int32)
    pk = g.integers(0, 3000, np_, dtype=np.int32)
    gpu, cpu = _run_join(device, bk, g.random(nb) < 0.5, g.integers(0, 99, nb, dtype=np.int64),
                         pk, g.random(np_) < 0.5, g.random(np_), kind)
    for k in (1, 2, 3, 4):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"{kind} col {k}")


def test_join_with_condition(device):
    g = np.random.default_rng(4)
    nb, np_ = 3000, 9000
    bk = g.integers(0, 500, nb, dtype=np.int32)
    pk = g.integers(0, 500, np_, dtype=np.int32)
    bv = g.integers(0, 4, nb, dtype=np.int64)
    pv = g.integers(0, 4, np_).astype(np.float64)
    gpu, cpu = _run_join(device, bk, g.random(nb) < 0.8, bv, pk, g.random(np_) < 0.8, pv,
                         conds=[['#2', '<=', '#4']])
    for k in (1, 2, 3, 4):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"cond col {k}")


# ------------------------------------------------------------------ aggregate
def _agg_case(device, keys, vals, mask, results, key_ids, agg_ids):
    from tqp_amd.operators import aggregate
    tables = dict(keys)
    tables.update(vals)
    input_ids = list(tables)
    if mask is not None:
        tables[900] = ('bool', mask)
    args = {'keys': [f"k#{k}" for k in key_ids], 'results': results, 'out agg ids': agg_ids,
            'is partial': False, 'input ids': list(input_ids)}
    if mask is not None:
        args['mask map'] = {i: 900 for i in input_ids}
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    errs = []
    for run, grp in ((lambda g: aggregate.tqp_hash_aggregate(True, g, copy.deepcopy(args), None, None), gpu),
                     (lambda g: R.tqp_hash_aggregate(g, copy.deepcopy(args)), cpu)):
        try:
            run(grp)
            errs.append(None)
        except Exception as e:   # the reference's own failure on degenerate inputs
            errs.append(type(e))
    return gpu, cpu, errs


AGG_RESULTS = lambda: [  # noqa: E731
    [[['sum', ['v#20']], '#', '1'], 'AS', 's#50'],
    [[['sum', [['v#20', '*', ['1.0', '-', 'd#21']]]], '#', '2'], 'AS', 'sd#51'],
    [[['avg', ['d#21']], '#', '3'], 'AS', 'a#52'],
    [[['count', ['1']], '#', '4'], 'AS', 'c#53'],
    [[['sum', ['i#22']], '#', '5'], 'AS', 'si#54'],
    [[['min', ['i#22']], '#', '6'], 'AS', 'mn#55'],
    [[['max', ['v#20']], '#', '7'], 'AS', 'mx#56'],
]


@pytest.mark.parametrize("case", ["few_groups", "dense_global", "hash_sparse", "skew_front",
                                  "no_keys", "all_masked", "all_masked_keyed", "all_masked_no_avg"])
def test_aggregate_parity(device, case):
    g = np.random.default_rng(sum(map(ord, case)) + 1)
    n = 400_000
    if case == "few_groups":
        k = {10: ('string', g.integers(0, 3, n, dtype=np.int8)), 11: ('string', g.integers(0, 2, n, dtype=np.int8))}
    elif case == "dense_global":
        k = {10: ('int', g.integers(1, 300_000, n, dtype=np.int32))}
    elif case == "hash_sparse":
        k = {10: ('int', g.integers(0, 1 << 45, 50_000)[g.integers(0, 50_000, n)].astype(np.int64)),
             11: ('int', g.integers(-3, 3, n, dtype=np.int32))}
    elif case in ("skew_front", "all_masked_keyed"):
        hot = np.where(g.random(n) < 0.6, 5, g.integers(0, 2000, n))
        k = {10: ('int', hot.astype(np.int32)), 11: ('int', g.integers(0, 100, n, dtype=np.int32))}
    else:
        k = {}
    vals = {20: ('float', g.uniform(900, 105000, n)), 21: ('float', g.integers(0, 11, n) / 100.0),
            22: ('int', g.integers(-1000, 1000, n, dtype=np.int64))}
    masked = case.startswith("all_masked")
    mask = None if case in ("dense_global",) else (g.random(n) < (0.0 if masked else 0.9))
    key_ids = sorted(k)
    aggs = AGG_RESULTS()
    agg_ids = [50, 51, 52, 53, 54, 55, 56]
    if case == "all_masked_no_avg":
        aggs = [a for a in aggs if a[0][0][0] != 'avg']
        agg_ids = [50, 51, 53, 54, 55, 56]
    results = [f"k#{i}" for i in key_ids] + aggs
    gpu, cpu, errs = _agg_case(device, k, vals, mask, results, key_ids, agg_ids)
    # zero selected rows: the reference divides by the group count (keyed) or divides an integer
    # bincount in place (avg); both must fail the same way, and what it does compute must match
    expect = {"all_masked": RuntimeError, "all_masked_keyed": ZeroDivisionError}.get(case)
    assert errs == [expect, expect], (case, errs)
    if expect is not None:
        return
    for i in key_ids:
        _assert_col(gpu[i].tensor, cpu[i].tensor, f"{case} key {i}")
    for a in agg_ids:
        _assert_col(gpu[a].tensor, cpu[a].tensor, f"{case} agg {a}", fp=True)


@pytest.mark.parametrize("masked", [False, True])
def test_aggregate_key_dictionaries(device, monkeypatch, masked):
    """Keys whose values are spaced apart (10 codes 100 apart, 20 codes 7 apart): the raw key
    space (901 x 134 values) needs a global table, the dictionaries shrink it to at most 200
    slots in LDS.  Groups, counts and sums must match the oracle's, and the dictionaries must
    have been taken (with a mask, only the selected rows' values count)."""
    from tqp_amd import kernels
    taken = []
    orig = kernels.GroupBy._adopt_dictionaries

    def adopt(self, d, *a):
        orig(self, d, *a)
        taken.append(d.key_dict[0] is not None and d.key_dict[1] is not None)

    monkeypatch.setattr(kernels.GroupBy, "_adopt_dictionaries", adopt)
    g = np.random.default_rng(4242 + masked)
    n = 1_600_000   # 1.12M selected rows with the mask: above the dictionaries' row threshold
    k = {10: ('int', (g.integers(0, 10, n) * 100).astype(np.int32)),
         11: ('int', (g.integers(0, 20, n) * 7 + 3).astype(np.int32))}
    vals = {20: ('float', g.uniform(900, 105000, n)), 21: ('float', g.integers(0, 11, n) / 100.0),
            22: ('int', g.integers(-1000, 1000, n, dtype=np.int64))}
    mask = (g.random(n) < 0.7) if masked else None
    key_ids = sorted(k)
    agg_ids = [50, 51, 52, 53, 54, 55, 56]
    results = [f"k#{i}" for i in key_ids] + AGG_RESULTS()
    print("dictionary case, masked", masked, flush=True)
    gpu, cpu, errs = _agg_case(device, k, vals, mask, results, key_ids, agg_ids)
    assert errs == [None, None], errs
    assert taken == [True], taken
    for i in key_ids:
        _assert_col(gpu[i].tensor, cpu[i].tensor, f"dict key {i}")
    for a in agg_ids:
        _assert_col(gpu[a].tensor, cpu[a].tensor, f"dict agg {a}", fp=True)


def test_aggregate_many_columns_global_table(device):
    """Eight loaded columns fill the LDS stage, so the dense key space takes the global table
    with atomics rather than LDS partitions; 600k rows over a 250k-key range."""
    g = np.random.default_rng(77)
    n = 600_000
    k = {10: ('int', g.integers(0, 250_000, n, dtype=np.int32))}
    vals = {20 + i: ('float', g.uniform(1, 100, n)) for i in range(6)}
    vals[26] = ('int', g.integers(-9, 9, n, dtype=np.int64))
    results = ["k#10"] + [[[['sum', [f'v#{20 + i}']], '#', str(i)], 'AS', f's#{50 + i}']
                          for i in range(6)] + [[[['max', ['v#26']], '#', '6'], 'AS', 'm#56']]
    agg_ids = [50, 51, 52, 53, 54, 55, 56]
    gpu, cpu, errs = _agg_case(device, k, vals, None, results, [10], agg_ids)
    assert errs == [None, None]
    _assert_col(gpu[10].tensor, cpu[10].tensor, "wide key")
    for a in agg_ids:
        _assert_col(gpu[a].tensor, cpu[a].tensor, f"wide agg {a}", fp=True)


# ------------------------------------------------------------------ full pipelines
def _run_pipeline(device, tables, plan):
    from tqp_amd import executor
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    executor.execute(copy.deepcopy(plan), gpu)
    R.execute(copy.deepcopy(plan), cpu)
    return gpu, cpu


def test_config0_pipeline(device):
    from tqp_amd import workloads
    print("config0 full size", flush=True)
    gpu, cpu = _run_pipeline(device, workloads.config0_tables(), workloads.config0_plan())
    for k in (6, 7, 28):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"c0 {k}")
    for k in (22, 23, 24, 25, 26, 27):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"c0 {k}", fp=True)
    # sum(quantity) is integer-valued and below 2^53: exact in any order
    _assert_col(gpu[22].tensor, cpu[22].tensor, "c0 sum_qty exact")


def test_config1_pipeline(device):
    from tqp_amd import workloads
    print("config1 full size", flush=True)
    gpu, cpu = _run_pipeline(device, workloads.config1_tables(), workloads.config1_plan())
    _assert_col(gpu[101].tensor, cpu[101].tensor, "c1 top keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "c1 top revenue", fp=True)


@pytest.mark.parametrize("sync", [True, False])
def test_config1_pipeline_under_inference_mode(device, sync):
    """main.py:341-342 runs the plan inside torch.inference_mode(): the operators' outputs are
    then inference tensors (no version counter), including the tables a scan would load there."""
    from tqp_amd import executor, workloads
    tables = workloads.config1_tables(n_build=200_000, n_probe=800_000)
    plan = workloads.config1_plan()
    cpu = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), cpu)
    print(f"config1 under inference_mode, sync={sync}", flush=True)
    with torch.inference_mode():
        gpu = _gpu_group(tables, device)
        executor.execute(copy.deepcopy(plan), gpu, sync=sync)
        torch.cuda.synchronize()
    assert gpu[301].tensor.is_inference()
    _assert_col(gpu[101].tensor, cpu[101].tensor, "c1 top keys (inference)")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "c1 top revenue (inference)", fp=True)


def test_config1_join_aggregate_all_groups(device):
    """Every group of the 264k-group aggregate (not only the top 10)."""
    from tqp_amd import workloads
    plan = workloads.config1_plan()[:4]
    gpu, cpu = _run_pipeline(device, workloads.config1_tables(), plan)
    _assert_col(gpu[101].tensor, cpu[101].tensor, "c1 group keys")
    _assert_col(gpu[301].tensor, cpu[301].tensor, "c1 group revenue", fp=True)


def test_config2_pipeline_reduced(device):
    from tqp_amd import workloads
    tables = workloads.config2_tables(n_fact=2_000_000, n_a=500_000, n_b=100_000)
    gpu, cpu = _run_pipeline(device, tables, workloads.config2_plan())
    for k in (322, 313, 402):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"c2 {k}")
    _assert_col(gpu[401].tensor, cpu[401].tensor, "c2 sum", fp=True)


def test_config2_full_size_properties(device):
    """60M-row star at full size: every fact row joins all three dimensions (keys are drawn from
    them), so the group counts sum to the fact row count and the group sums add up to the total
    amount; groups are distinct and ascending; the count of each (c_attr, b_mod) group equals a
    direct histogram of the fact rows' attribute pairs."""
    from tqp_amd import executor, workloads
    tables = workloads.config2_tables()
    gpu = _gpu_group(tables, device)
    print("config2 full size: tables resident", flush=True)
    executor.execute(copy.deepcopy(workloads.config2_plan()), gpu)
    c = gpu[322].tensor.reshape(-1).cpu().numpy().astype(np.int64)
    b = gpu[313].tensor.reshape(-1).cpu().numpy().astype(np.int64)
    cnt = gpu[402].tensor.cpu().numpy()
    tot = gpu[401].tensor.cpu().numpy()
    n_fact = tables[331][1].shape[0]
    assert cnt.sum() == n_fact
    amount = tables[334][1]
    assert abs(tot.sum() - amount.sum()) <= 1e-10 * amount.sum()
    key = c * 1000 + b
    assert np.all(np.diff(key) > 0)
    # independent histogram of (c_attr, b_attr % 100) over fact rows via the dimension maps
    c_attr = np.zeros(26, np.int64)
    c_attr[tables[321][1]] = tables[322][1]
    b_mod = np.zeros(tables[311][1].max() + 1, np.int64)
    b_mod[tables[311][1]] = tables[313][1]
    fk = c_attr[tables[333][1]] * 1000 + b_mod[tables[332][1]]
    u, n = np.unique(fk, return_counts=True)
    assert np.array_equal(u, key) and np.array_equal(n, cnt)


def test_aggregate_year_case_division_not_equal(device):
    """year, CASE WHEN, '/' (true division in float32 for integer operands) and '<>' inside
    aggregates, computed by the expression kernel, against the oracle (whose own results are
    pinned to the calendar and numpy in tests/test_oracle.py).  Day counts stay below 3652: the
    reference's year table ends with 1999 and fails beyond it."""
    g = np.random.default_rng(77)
    n = 300_000
    k = {10: ('string', g.integers(0, 3, n, dtype=np.int8)),
         11: ('int', g.integers(0, 40, n, dtype=np.int32))}
    vals = {20: ('date', g.integers(-5, 3652, n, dtype=np.int32)),
            21: ('float', g.integers(1, 51, n).astype(np.float64)),
            22: ('int', g.integers(-10, 10, n, dtype=np.int64))}
    results = ['k#10', 'k#11',
               [[['sum', [['year', ['d#20']]]], '#', '1'], 'AS', 'y#50'],
               [[['sum', [['CASE', 'WHEN', ['x#22', '>', '0'], 'THEN', 'q#21', 'ELSE', '0.0',
                           'END']]], '#', '2'], 'AS', 'cw#51'],
               [[['sum', [['x#22', '/', '4']]], '#', '3'], 'AS', 'dv#52'],
               [[['sum', [['x#22', '<>', '0']]], '#', '4'], 'AS', 'ne#53'],
               [[['max', [['year', ['d#20']]]], '#', '5'], 'AS', 'my#54']]
    agg_ids = [50, 51, 52, 53, 54]
    gpu, cpu, errs = _agg_case(device, k, vals, g.random(n) < 0.85, results, [10, 11], agg_ids)
    assert errs == [None, None], errs
    for i in (10, 11):
        _assert_col(gpu[i].tensor, cpu[i].tensor, f"key {i}")
    for a in (50, 53, 54):
        _assert_col(gpu[a].tensor, cpu[a].tensor, f"agg {a}")
    _assert_col(gpu[51].tensor, cpu[51].tensor, "case sum", fp=True)
    # float32 arguments: the oracle sums them with bincount in float64 in row order, the
    # kernels in varying order (the relative error of a float32-valued f64 sum stays ~1e-13)
    g51, c51 = gpu[52].tensor.cpu(), cpu[52].tensor
    assert g51.dtype == c51.dtype and g51.shape == c51.shape
    np.testing.assert_allclose(g51.numpy(), c51.numpy(), rtol=1e-12, atol=0)


def test_filter_contains_endswith_substring_bit_exact(device):
    """Contains / EndsWith (like_contains over the whole row) and substring IN / '=' over a 2-D
    string column, alone, negated and in conjunctions, against the oracle's masks (pinned to
    bytes.find and byte slices in tests/test_oracle.py)."""
    from tqp_amd import kernels
    g = np.random.default_rng(29)
    n, w = 150_001, 25
    s = np.array([ord(c) for c in "abcd "], dtype=np.int8)[g.integers(0, 5, (n, w))]
    s[g.random((n, w)) < 0.08] = 0
    rows = np.flatnonzero(g.random(n) < 0.2)
    for r, p in zip(rows, g.integers(0, w - 4, rows.shape[0])):
        s[r, p:p + 5] = [ord(c) for c in "green"]
    tables = {4: ('string', s), 5: ('int', g.integers(0, 100, n, dtype=np.int32))}
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    for tree in (['Contains', ['c#4', 'green']],
                 ['EndsWith', ['c#4', 'dab']],
                 ['NOT', ['Contains', ['c#4', 'a b']]],
                 [['Contains', ['c#4', 'cc']], 'AND', ['k#5', '<', '30']],
                 [['substring', ['c#4', '3', '2']], 'IN', ['ab', 'gr', 'ee']],
                 [['substring', ['c#4', '1', '3']], '=', 'abc'],
                 [['substring', ['c#4', '20', '10']], 'IN', ['green', 'ab']],
                 ['Contains', ['c#4', 'x' * 30]]):
        print("string function:", tree, flush=True)
        mask, counts = kernels.filter_mask(tree, gpu, None)
        want = R.filter_evaluate(cpu, tree, None, None)
        _assert_col(mask, want, f"mask {tree}")
        assert int(counts.sum().item()) == int(want.sum().item())


@pytest.mark.parametrize("case", ["string", "float", "string_and_int", "float32_masked"])
def test_aggregate_string_and_float_keys(device, case):
    """Group keys that do not pack into an integer range (2-D string rows of width 12, float64
    and float32 values with -0.0 and +0.0 in one group) are grouped by their dense rank in sort
    order; groups, key values (2-D [G, 12] for strings), counts and sums equal the oracle's
    (aggregate.py:302-346, pinned to exact grouping in tests/test_oracle.py)."""
    g = np.random.default_rng(sum(map(ord, case)))
    n = 250_000
    words = [b"ALGERIA", b"ARGENTINA", b"BRAZIL", b"CANADA", b"EGYPT", b"ETHIOPIA", b"FRANCE",
             b"GERMANY", b"INDIA", b"INDONESIA", b"IRAN", b"IRAQ", b"JAPAN", b"JORDAN"]
    names = np.zeros((len(words), 12), dtype=np.int8)
    for i, wd in enumerate(words):
        names[i, :len(wd)] = np.frombuffer(wd, dtype=np.int8)
    fvals = np.array([-2.5, -0.0, 0.0, 0.125, 1.0, 3.75, 1e12, -1e-12], dtype=np.float64)
    keys = {}
    if case in ("string", "string_and_int"):
        keys[10] = ('string', names[g.integers(0, len(words), n)])
    if case == "string_and_int":
        keys[11] = ('int', g.integers(0, 7, n, dtype=np.int32))
    if case == "float":
        keys[10] = ('float', fvals[g.integers(0, len(fvals), n)])
    if case == "float32_masked":
        keys[10] = ('float', fvals[g.integers(0, len(fvals), n)].astype(np.float32))
    vals = {20: ('float', g.uniform(900, 105000, n)), 22: ('int', g.integers(-1000, 1000, n))}
    mask = g.random(n) < 0.8 if case == "float32_masked" else None
    key_ids = sorted(keys)
    results = [f"k#{i}" for i in key_ids] + [
        [[['sum', ['v#20']], '#', '1'], 'AS', 's#50'],
        [[['count', ['1']], '#', '2'], 'AS', 'c#51'],
        [[['max', ['i#22']], '#', '3'], 'AS', 'mx#52']]
    gpu, cpu, errs = _agg_case(device, keys, vals, mask, results, key_ids, [50, 51, 52])
    assert errs == [None, None], errs
    for i in key_ids:
        _assert_col(gpu[i].tensor, cpu[i].tensor, f"{case} key {i}")
    _assert_col(gpu[51].tensor, cpu[51].tensor, f"{case} count")
    _assert_col(gpu[52].tensor, cpu[52].tensor, f"{case} max")
    _assert_col(gpu[50].tensor, cpu[50].tensor, f"{case} sum", fp=True)
    if case.startswith("string"):
        assert gpu[10].tensor.shape[1] == 12


def test_aggregate_evaluate_over_a_group_index(device):
    """aggregate.evaluate (aggregate.py:58-233) over rows numbered into groups by the caller:
    an aggregate call, an expression over aggregates, and a row expression, against the
    oracle's agg_evaluate on the same index."""
    from tqp_amd.operators import aggregate
    g = np.random.default_rng(91)
    n, G = 200_000, 37
    inv = g.integers(0, G, n)
    inv[:G] = np.arange(G)                   # every group holds rows
    v = g.uniform(900, 105000, n)
    d = g.integers(0, 11, n) / 100.0
    q = g.integers(1, 51, n).astype(np.float64)
    tables = {20: ('float', v), 21: ('float', d), 22: ('float', q)}
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    ginv = torch.from_numpy(inv).to(device)
    cinv = torch.from_numpy(inv)
    for tree in ([['sum', [['v#20', '*', ['1.0', '-', 'd#21']]]], '#', '1'],
                 [['0.2', '*', ['avg', ['q#22']]], '#', '2'],
                 [['sum', ['v#20']], '/', ['count', ['1']]],
                 [['max', ['q#22']]],
                 ['v#20', '*', ['1.0', '-', 'd#21']]):
        got = aggregate.evaluate('cuda', gpu, tree, n, G, ginv)
        want = R.agg_evaluate(cpu, tree, n, G, cinv)
        print("evaluate:", tree, flush=True)
        _assert_col(got, want, f"evaluate {tree}", fp=True)


def test_filter_evaluate_with_leaf_function(device):
    """filter.evaluate with leaf_function (filter.py:27-35, the join condition path of
    hashjoin.py:15-22): leaves come from the function, here columns gathered by a row index."""
    from tqp_amd import kernels
    from tqp_amd.operators import filter as F
    g = np.random.default_rng(93)
    n, m = 100_000, 60_000
    a = g.integers(0, 100, n, dtype=np.int32)
    b = g.uniform(0, 10, n)
    idx = g.integers(0, n, m)
    ga = torch.from_numpy(a).to(device)
    gbv = torch.from_numpy(b).to(device)
    gi = torch.from_numpy(idx).to(device)
    gcols = {1: ga, 2: gbv}
    ccols = {1: torch.from_numpy(a), 2: torch.from_numpy(b)}
    tree = [['#1', '<', '40'], 'AND', ['#2', '>=', '2.5']]
    got = F.evaluate('cuda', {}, tree, m, None,
                     leaf_function=lambda cid: kernels.gather(gi, [gcols[cid]])[0])
    want = R.filter_evaluate({}, tree, m, None,
                             leaf_function=lambda cid: ccols[cid][torch.from_numpy(idx)])
    _assert_col(got, want, "leaf mask")


# ------------------------------------------------------------------ project
def _project_case(n, seed=82):
    g = np.random.default_rng(seed)
    tables = {202: ('float', g.uniform(900, 100000, n)),
              203: ('float', np.round(g.integers(0, 11, n) / 100.0, 2)),
              204: ('date', g.integers(0, 3652, n, dtype=np.int32)),
              205: ('float', np.round(g.integers(0, 9, n) / 100.0, 2)),
              206: ('int', g.integers(1, 51, n).astype(np.int64))}
    args = {'name': 'Project', 'rename map': {202: 602},
            'used expr': [
                'price#202',
                [['discount#203', 'AS', 'disc#603']],
                [[['price#202', '*', ['1.0', '-', 'discount#203']], 'AS', 'rev#604']],
                [[[['price#202', '*', ['1.0', '-', 'discount#203']], '*', ['1.0', '+', 'tax#205']],
                  'AS', 'charge#605']],
                [[['year', 'ship_day#204'], 'AS', 'y#606']],
                [[['CASE', 'WHEN', ['q#206', '>', '25'], 'THEN', 'q#206', 'ELSE', '0', 'END'],
                  'AS', 'bigq#607']],
                [[['ship_day#204', '<', '1200'], 'AS', 'early#608']]],
            'materialized in': [202, 203, 204, 205, 206],
            'materialized out': [604, 605, 606, 607, 608]}
    return tables, args


def test_project_bit_exact(device):
    """project.py:10-69 in one expression launch: products, year, CASE, a comparison; renames."""
    from tqp_amd import executor
    tables, args = _project_case(2_000_000)
    gpu = _gpu_group(tables, device)
    cpu = R.to_cols(tables)
    print("project: 5 expressions over 2M rows", flush=True)
    executor.execute([(1, copy.deepcopy(args))], gpu)
    R.execute([(1, copy.deepcopy(args))], cpu)
    assert gpu[602] is gpu[202] and 603 not in gpu
    for k in (604, 605):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"projected {k}", fp=True)
    for k in (606, 607, 608):
        _assert_col(gpu[k].tensor, cpu[k].tensor, f"projected {k}")
        assert gpu[k].tensor_type == cpu[k].tensor_type


def test_project_many_expressions_then_aggregate(device):
    """Fourteen expressions (more than one program's outputs), then a group-by over projected
    columns, with and without per-operator synchronisation."""
    from tqp_amd import executor
    tables, args = _project_case(300_000, seed=83)
    extra = [[[['q#206', '*', str(i)], 'AS', f'm{i}#{620 + i}']] for i in range(9)]
    args['used expr'] = args['used expr'] + extra
    args['materialized out'] = args['materialized out'] + [620 + i for i in range(9)]
    agg = {'name': 'Aggregate', 'input ids': [607, 604, 628], 'keys': ['bigq#607'],
           'results': ['bigq#607',
                       [[[['sum', ['rev#604']], '#', '50'], 'AS', 'srev#701']],
                       [[[['max', ['m8#628']], '#', '51'], 'AS', 'mx#702']]],
           'out agg ids': [701, 702], 'is partial': False}
    plan = [(1, args), (2, agg)]
    cpu = R.to_cols(tables)
    R.execute(copy.deepcopy(plan), cpu)
    for sync in (True, False):
        gpu = _gpu_group(tables, device)
        print(f"project + aggregate, sync={sync}", flush=True)
        executor.execute(copy.deepcopy(plan), gpu, sync=sync)
        torch.cuda.synchronize()
        _assert_col(gpu[607].tensor, cpu[607].tensor, "group keys")
        _assert_col(gpu[701].tensor, cpu[701].tensor, "sum of a projected column", fp=True)
        _assert_col(gpu[702].tensor, cpu[702].tensor, "max of a projected column")
        for i in range(9):
            _assert_col(gpu[620 + i].tensor.cpu(), cpu[620 + i].tensor, f"m{i}")
