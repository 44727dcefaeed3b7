"""Pins the CPU oracle (oracle/ref_ops.py) against closed forms and brute force (CPU only).

The reference holds no usable golden vectors for this path (its only known-answer file,
tests/verification/SF1_reference.out, needs the TPC-H SF1 tensor files, which are not
available), so each oracle function is checked against an independent statement of its result
(oracle/closed_form.py) on random small inputs, plus hand-worked cases.  The planted-mistake
tests show the checks have teeth: a plausible bug injected into the oracle makes them fail.
"""
import re

import numpy as np
import pytest
import torch

from oracle import closed_form as CF
from oracle import ref_ops as R


def cols(**arrays):
    return {int(k[1:]): R.Col(torch.from_numpy(v), 'int') for k, v in arrays.items()}


# ------------------------------------------------------------------ filter
@pytest.mark.parametrize("op", ['<', '<=', '>', '>=', '='])
@pytest.mark.parametrize("lit", [1200, 1200.5, 0.0, 2556])
def test_filter_int_column_matches_closed_form(op, lit):
    g = np.random.default_rng(1)
    x = g.integers(0, 2557, 5000, dtype=np.int32)
    t = {1: R.Col(torch.from_numpy(x), 'date')}
    tree = ['ship_day#1', op, str(lit)]
    got = R.filter_evaluate(t, tree, None, None).numpy()
    want = CF.mask_compare(x, op, lit)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("op", ['<', '<=', '>', '>=', '='])
def test_filter_float_column_eps_and_isclose(op):
    x = np.array([0.05, 0.05 + 5e-6, 0.05 - 5e-6, 0.05 + 2e-5, 0.06, -0.0, np.nan], dtype=np.float64)
    t = {3: R.Col(torch.from_numpy(x), 'float')}
    got = R.filter_evaluate(t, ['d#3', op, '0.05'], None, None).numpy()
    assert np.array_equal(got, CF.mask_compare(x, op, 0.05))


def test_filter_eps_known_answers():
    # hand-worked: <= 0.05 widens by 1e-5, = is isclose(atol=1e-5, rtol=1e-5)
    x = torch.tensor([0.05, 0.050009, 0.050011, 0.049991], dtype=torch.float64)
    t = {3: R.Col(x, 'float')}
    assert R.filter_evaluate(t, ['d#3', '<=', '0.05'], None, None).tolist() == [True, True, False, True]
    assert R.filter_evaluate(t, ['d#3', '=', '0.05'], None, None).tolist() == [True, True, False, True]


def test_filter_and_or_not_isnotnull_in():
    g = np.random.default_rng(2)
    a = g.integers(0, 10, 1000, dtype=np.int32)
    b = g.integers(0, 5, 1000, dtype=np.int8)
    t = {1: R.Col(torch.from_numpy(a), 'date'), 2: R.Col(torch.from_numpy(b), 'string')}
    tree = [[['isnotnull', ['a#1']], 'AND', [['a#1', '<', '4'], 'OR', ['NOT', ['b#2', '=', '2']]]]]
    got = R.filter_evaluate(t, tree, None, None).numpy()
    assert np.array_equal(got, (a < 4) | ~(b == 2))
    got = R.filter_evaluate(t, ['a#1', 'IN', ['1', '3', '7']], None, None).numpy()
    assert np.array_equal(got, np.isin(a, [1, 3, 7]))


def test_filter_string_equality_rows():
    s = np.zeros((4, 10), dtype=np.int8)
    for i, w in enumerate(["BUILDING", "BUILD", "MACHINERY", "BUILDING"]):
        s[i, :len(w)] = [ord(c) for c in w]
    t = {5: R.Col(torch.from_numpy(s), 'string')}
    assert R.filter_evaluate(t, ['seg#5', '=', 'BUILDING'], None, None).tolist() == [True, False, False, True]
    assert R.filter_evaluate(t, ['StartsWith', ['seg#5', 'BUILD']], None, None).tolist() == [True, True, False, True]


def _like_regex(s, w1, w2):
    # '%w1%w2%' over the full row bytes, padding included: a regular-expression search
    text = bytes(int(c) & 0xFF for c in s).decode('latin-1')
    return re.search(re.escape(w1) + '.*' + re.escape(w2), text, re.DOTALL) is not None


def test_like_matches_regex_search():
    g = np.random.default_rng(5)
    rows = ["special requests", "requests special", "specialrequests", "specialrequest",
            "xspecial yrequestsz", "special", "", "requests", "spec requests special requests",
            "special%requests", "sspecialrequestss"]
    s = np.zeros((len(rows) + 400, 32), dtype=np.int8)
    for i, w in enumerate(rows):
        s[i, :len(w)] = [ord(c) for c in w]
    # random rows over a small alphabet, so both words occur by chance in many orders
    rnd = g.integers(0, 4, (400, 32))
    s[len(rows):] = np.array([ord(c) for c in "abc "], dtype=np.int8)[rnd]
    s[len(rows):][g.random((400, 32)) < 0.1] = 0
    t = {5: R.Col(torch.from_numpy(s), 'string')}
    for pat in ('%special%requests%', '%ab%ca%', '%a%a%', '%abc%%', '%%cab%', '%%%',
                '%special requests%s%', '%a%' + 'b' * 31 + '%'):
        _, w1, w2, _ = pat.split('%')
        got = R.filter_evaluate(t, ['c#5', 'LIKE', pat], None, None).tolist()
        want = [_like_regex(r, w1, w2) for r in s]
        assert got == want, pat
    assert R.filter_evaluate(t, ['c#5', 'LIKE', '%special%requests%'], None, None).tolist()[:11] == \
        [True, False, True, False, True, False, False, False, True, True, True]
    with pytest.raises(AssertionError):
        R.filter_evaluate(t, ['c#5', 'LIKE', '%special%'], None, None)


def test_compaction_order_preserving():
    g = np.random.default_rng(3)
    m = g.random(3000) < 0.3
    v = g.random(3000)
    got = R.apply_masks({1: R.Col(torch.from_numpy(v), 'float'), 9: R.Col(torch.from_numpy(m), 'bool')},
                        {1: 9})
    assert list(got) == [1]
    assert np.array_equal(got[1].tensor.numpy(), CF.compact(m, v))


# ------------------------------------------------------------------ join
def _join_case(g, nb, np_, dom, bsel=0.7, psel=0.6, dup=True):
    bk = g.integers(0, dom, nb, dtype=np.int32) if dup else g.permutation(2 * nb)[:nb].astype(np.int32)
    pk = g.integers(0, dom + 3, np_, dtype=np.int32)
    bm = g.random(nb) < bsel
    pm = g.random(np_) < psel
    return bk, bm, pk, pm


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("dup", [True, False])
def test_join_inner_matches_brute_force(seed, dup):
    g = np.random.default_rng(seed)
    bk, bm, pk, pm = _join_case(g, 40, 60, 30, dup=dup)
    bv = np.arange(100, 140, dtype=np.int64)
    pv = g.random(60)
    tg = {1: R.Col(torch.from_numpy(bk), 'int'), 2: R.Col(torch.from_numpy(bv), 'int'),
          3: R.Col(torch.from_numpy(pk), 'int'), 4: R.Col(torch.from_numpy(pv), 'float'),
          901: R.Col(torch.from_numpy(bm), 'bool'), 902: R.Col(torch.from_numpy(pm), 'bool')}
    args = {'left group': [1, 2], 'right group': [3, 4], 'materialized in': [1, 2, 3, 4],
            'mask map': {1: 901, 2: 901, 3: 902, 4: 902}}
    R.join_vortex(tg, args, 1, 3, [], 'inner')
    pairs = CF.join_pairs_brute(bk, bm, pk, pm)
    assert tg[2].tensor.tolist() == [int(bv[b]) for b, _ in pairs]
    assert tg[4].tensor.tolist() == [float(pv[p]) for _, p in pairs]
    assert tg[1].tensor.tolist() == [int(bk[b]) for b, _ in pairs]
    assert args['mask map'] == {}


@pytest.mark.parametrize("kind", ['right-semi', 'right-anti', 'right-outer'])
def test_join_other_kinds_match_brute_force(kind):
    g = np.random.default_rng(11)
    bk, bm, pk, pm = _join_case(g, 30, 50, 20)
    bkc, pkc = bk[bm], pk[pm]
    tl = {1: R.Col(torch.from_numpy(bkc), 'int')}
    tr = {3: R.Col(torch.from_numpy(pkc), 'int')}
    out = R.join_kernel(tl, tr, 1, 3, [], [1, 3], kind)
    pairs = CF.join_pairs_brute(bkc, None, pkc, None)
    matched = sorted({p for _, p in pairs})
    if kind == 'right-semi':
        assert out[1].tolist() == [int(pkc[p]) for p in matched]
    elif kind == 'right-anti':
        assert out[1].tolist() == [int(pkc[p]) for p in range(len(pkc)) if p not in matched]
    else:
        un = [p for p in range(len(pkc)) if p not in matched]
        assert out[1].tolist() == [int(pkc[p]) for _, p in pairs] + [int(pkc[p]) for p in un]
        assert out[0].tolist() == [int(bkc[b]) for b, _ in pairs] + [-64] * len(un)


def test_join_condition_list():
    bk = np.array([1, 1, 2, 3], dtype=np.int32)
    bx = np.array([5, 6, 7, 8], dtype=np.int32)
    pk = np.array([1, 2, 3, 1], dtype=np.int32)
    px = np.array([6, 7, 0, 5], dtype=np.int32)
    tl = {1: R.Col(torch.from_numpy(bk), 'int'), 2: R.Col(torch.from_numpy(bx), 'int')}
    tr = {3: R.Col(torch.from_numpy(pk), 'int'), 4: R.Col(torch.from_numpy(px), 'int')}
    out = R.join_kernel(tl, tr, 1, 3, [['#2', '=', '#4']], [1, 2, 3, 4], 'inner')
    assert [t.tolist() for t in out] == [[1, 2, 1], [6, 7, 5], [1, 2, 1], [6, 7, 5]]


# ------------------------------------------------------------------ aggregate
def _agg_args(keys, results, agg_ids, input_ids, mask_map=None):
    a = {'keys': keys, 'results': results, 'out agg ids': agg_ids, 'is partial': False,
         'input ids': list(input_ids)}
    if mask_map:
        a['mask map'] = dict(mask_map)
    return a


@pytest.mark.parametrize("seed", range(4))
def test_aggregate_matches_sequential_group_reduce(seed):
    g = np.random.default_rng(seed)
    n = 3000
    k1 = g.integers(-3, 4, n, dtype=np.int8)
    k2 = g.integers(0, 50, n, dtype=np.int32)
    q = g.integers(1, 51, n).astype(np.float64)
    p = g.uniform(900, 105000, n)
    d = g.integers(0, 11, n) / 100.0
    iv = g.integers(-1000, 1000, n, dtype=np.int64)
    m = g.random(n) < 0.8
    tg = {1: R.Col(torch.from_numpy(k1), 'string'), 2: R.Col(torch.from_numpy(k2), 'date'),
          3: R.Col(torch.from_numpy(q), 'float'), 4: R.Col(torch.from_numpy(p), 'float'),
          5: R.Col(torch.from_numpy(d), 'float'), 6: R.Col(torch.from_numpy(iv), 'int'),
          900: R.Col(torch.from_numpy(m), 'bool')}
    disc = ['p#4', '*', ['1.0', '-', 'd#5']]
    res = ['a#1', 'b#2',
           [[['sum', ['q#3']], '#', '1'], 'AS', 's#20'],
           [[['sum', [disc]], '#', '2'], 'AS', 'sd#21'],
           [[['avg', ['d#5']], '#', '3'], 'AS', 'ad#22'],
           [[['count', ['1']], '#', '4'], 'AS', 'c#23'],
           [[['sum', ['i#6']], '#', '5'], 'AS', 'si#24'],
           [[['max', ['i#6']], '#', '6'], 'AS', 'mx#25'],
           [[['min', ['i#6']], '#', '7'], 'AS', 'mn#26']]
    args = _agg_args(['a#1', 'b#2'], res, [20, 21, 22, 23, 24, 25, 26], [1, 2, 3, 4, 5, 6],
                     {i: 900 for i in range(1, 7)})
    R.tqp_hash_aggregate(tg, args)
    assert 900 in args['input ids']
    dv = p * (1.0 - d)
    keys, vals = CF.group_reduce([k1, k2], m, {'s': ('sum', q), 'sd': ('sum', dv), 'ad': ('avg', d),
                                               'c': ('count', np.ones(n)), 'si': ('sum', iv),
                                               'mx': ('max', iv), 'mn': ('min', iv)})
    assert tg[1].tensor.reshape(-1).tolist() == keys[0].tolist()
    assert tg[2].tensor.reshape(-1).tolist() == keys[1].tolist()
    assert tg[1].tensor.dtype == torch.int8 and tg[1].tensor.shape[1] == 1
    assert tg[23].tensor.tolist() == vals['c']
    assert tg[24].tensor.tolist() == vals['si'] and tg[24].tensor.dtype == torch.int64
    assert tg[25].tensor.tolist() == vals['mx'] and tg[26].tensor.tolist() == vals['mn']
    np.testing.assert_allclose(tg[20].tensor.numpy(), vals['s'], rtol=0, atol=0)
    np.testing.assert_allclose(tg[21].tensor.numpy(), vals['sd'], rtol=1e-12)
    np.testing.assert_allclose(tg[22].tensor.numpy(), vals['ad'], rtol=1e-12)


def test_aggregate_two_step_path_and_no_keys():
    # rows/groups >= 65536 takes the privatised path (aggregate.py:127-147); no keys -> 1 group
    n = 200_000
    g = np.random.default_rng(5)
    v = g.random(n)
    k = g.integers(0, 2, n, dtype=np.int8)
    tg = {1: R.Col(torch.from_numpy(k), 'string'), 2: R.Col(torch.from_numpy(v), 'float')}
    args = _agg_args(['k#1'], ['k#1', [[['sum', ['v#2']], '#', '1'], 'AS', 's#5'],
                                [[['max', ['v#2']], '#', '2'], 'AS', 'm#6']], [5, 6], [1, 2])
    R.tqp_hash_aggregate(tg, args)
    _, vals = CF.group_reduce([k], None, {'s': ('sum', v), 'm': ('max', v)})
    np.testing.assert_allclose(tg[5].tensor.numpy(), vals['s'], rtol=1e-12)
    assert tg[6].tensor.tolist() == vals['m']
    tg = {2: R.Col(torch.from_numpy(v[:10]), 'float')}
    args = _agg_args([], [[[['sum', ['v#2']], '#', '1'], 'AS', 's#5']], [5], [2])
    R.tqp_hash_aggregate(tg, args)
    assert tg[5].tensor.shape == (1,)
    assert abs(tg[5].tensor.item() - CF.seq_sum(v[:10])) <= 1e-12 * abs(CF.seq_sum(v[:10]))


def test_aggregate_max_initialised_to_zero():
    # reference quirk (aggregate.py:161-167): max starts at 0, so an all-negative group gives 0
    tg = {1: R.Col(torch.tensor([0, 0, 1], dtype=torch.int8), 'string'),
          2: R.Col(torch.tensor([-5, -3, 4], dtype=torch.int64), 'int')}
    args = _agg_args(['k#1'], ['k#1', [[['max', ['x#2']], '#', '1'], 'AS', 'm#5']], [5], [1, 2])
    R.tqp_hash_aggregate(tg, args)
    assert tg[5].tensor.tolist() == [0, 4]


def test_aggregate_zero_selected_rows():
    # every row masked out: keyed -> 0 groups and cnt_rows / cnt_groups fails (aggregate.py:128);
    # keyless -> one group over an empty index: torch.bincount of an empty index is int64
    # whatever the weights, so sum/count are int64 [0], avg's in-place divide fails
    # (aggregate.py:175-177), min is the dtype's maximum and max is 0 (aggregate.py:154-166)
    n = 50
    k = torch.arange(n, dtype=torch.int32) % 3
    v = torch.linspace(1.0, 2.0, n, dtype=torch.float64)
    x = torch.arange(n, dtype=torch.int64) - 7

    def group():
        return {1: R.Col(k.clone(), 'int'), 2: R.Col(v.clone(), 'float'),
                3: R.Col(x.clone(), 'int'), 9: R.Col(torch.zeros(n, dtype=torch.bool), 'bool')}

    def args(keys, results, out):
        a = _agg_args(keys, results, out, [1, 2, 3])
        a['mask map'] = {1: 9, 2: 9, 3: 9}
        return a

    s = [[['sum', ['v#2']], '#', '1'], 'AS', 's#5']
    with pytest.raises(ZeroDivisionError):
        R.tqp_hash_aggregate(group(), args(['k#1'], ['k#1', s], [5]))
    with pytest.raises(RuntimeError):
        R.tqp_hash_aggregate(group(), args([], [[[['avg', ['v#2']], '#', '1'], 'AS', 'a#5']], [5]))
    tg = group()
    R.tqp_hash_aggregate(tg, args([], [s, [[['count', ['1']], '#', '2'], 'AS', 'c#6'],
                                       [[['min', ['x#3']], '#', '3'], 'AS', 'n#7'],
                                       [[['max', ['v#2']], '#', '4'], 'AS', 'm#8']], [5, 6, 7, 8]))
    assert tg[5].tensor.dtype == torch.int64 and tg[5].tensor.tolist() == [0]
    assert tg[6].tensor.tolist() == [0]
    assert tg[7].tensor.tolist() == [torch.iinfo(torch.int64).max]
    assert tg[8].tensor.dtype == torch.float64 and tg[8].tensor.tolist() == [0.0]


# ------------------------------------------------------------------ sort
@pytest.mark.parametrize("seed", range(5))
def test_sort_matches_lexicographic_stable(seed):
    g = np.random.default_rng(seed)
    n = 500
    a = g.integers(0, 4, n, dtype=np.int8)
    f = np.round(g.normal(size=n), 1)
    f[g.random(n) < 0.05] = 0.0
    f[g.random(n) < 0.05] = -0.0
    f[g.random(n) < 0.03] = np.nan
    i = g.integers(-5, 5, n, dtype=np.int64)
    payload = np.arange(n, dtype=np.int64)
    for descs in ([False, True, False], [True, False, True], [True, True, True]):
        tg = {1: R.Col(torch.from_numpy(a.copy()), 'string'), 2: R.Col(torch.from_numpy(f.copy()), 'float'),
              3: R.Col(torch.from_numpy(i.copy()), 'int'), 4: R.Col(torch.from_numpy(payload.copy()), 'int')}
        R.tqp_sort(tg, {'input ids': [1, 2, 3, 4], 'key ids': [1, 2, 3], 'key orders': descs}, False)
        want = CF.sort_perm([a, f, i], descs)
        assert tg[4].tensor.tolist() == want.tolist()


def test_sort_2d_string_keys_and_limit():
    s = np.array([[66, 0], [65, 66], [65, 0], [66, 0], [65, 66]], dtype=np.int8)
    tg = {1: R.Col(torch.from_numpy(s), 'string'), 2: R.Col(torch.arange(5), 'int')}
    R.tqp_sort(tg, {'input ids': [1, 2], 'key ids': [1], 'key orders': [False], 'limit': 3}, True)
    assert tg[2].tensor.tolist() == [2, 1, 4]


# ------------------------------------------------------------------ planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """A dropped eps, a flipped sort direction and an unstable build sort each break a pin."""
    x = np.array([0.05, 0.050009], dtype=np.float64)
    t = {3: R.Col(torch.from_numpy(x), 'float')}
    monkeypatch.setattr(R, "EPS", 0.0)
    got = R.filter_evaluate(t, ['d#3', '<=', '0.05'], None, None).numpy()
    assert not np.array_equal(got, CF.mask_compare(x, '<=', 0.05))
    monkeypatch.undo()

    f = np.array([3.0, 1.0, 2.0, 1.0])
    pay = np.arange(4)
    tg = {1: R.Col(torch.from_numpy(f), 'float'), 2: R.Col(torch.from_numpy(pay), 'int')}
    orig = torch.sort

    def flipped(x, stable=True, descending=False):
        return orig(x, stable=stable, descending=not descending)
    monkeypatch.setattr(torch, "sort", flipped)
    R.tqp_sort(tg, {'input ids': [1, 2], 'key ids': [1], 'key orders': [False]}, False)
    monkeypatch.undo()
    assert tg[2].tensor.tolist() != CF.sort_perm([f], [False]).tolist()

    # a group-by that sums in float32 misses the 1e-12 bound
    g = np.random.default_rng(9)
    v = g.uniform(900, 105000, 10000)
    seq = CF.seq_sum(v)
    assert abs(float(np.sum(v.astype(np.float32), dtype=np.float32)) - seq) > 1e-12 * seq


def test_aggregate_float_and_string_keys_group_by_exact_value():
    """Float keys group by value (-0.0 with +0.0), 2-D string keys by their bytes; the key
    outputs are the groups' values in their own dtype and shape (aggregate.py:302-346)."""
    g = np.random.default_rng(41)
    n = 4000
    names = np.array([list(b"alpha\0"), list(b"beta\0\0"), list(b"alp\0\0\0"), list(b"zeta\0\0")],
                     dtype=np.int8)
    s = names[g.integers(0, 4, n)]
    f = g.choice(np.array([-1.5, -0.0, 0.0, 0.25, 0.3, 1e300, -7.0]), n)
    v = g.uniform(0, 10, n)
    m = g.random(n) < 0.9
    tg = {1: R.Col(torch.from_numpy(s), 'string'), 2: R.Col(torch.from_numpy(f), 'float'),
          3: R.Col(torch.from_numpy(v), 'float'), 900: R.Col(torch.from_numpy(m), 'bool')}
    res = ['a#1', 'b#2', [[['sum', ['v#3']], '#', '1'], 'AS', 's#20'],
           [[['count', ['1']], '#', '2'], 'AS', 'c#21']]
    args = _agg_args(['a#1', 'b#2'], res, [20, 21], [1, 2, 3], {1: 900, 2: 900, 3: 900})
    R.tqp_hash_aggregate(tg, args)
    keys, vals = CF.group_reduce([s, f], m, {'s': ('sum', v), 'c': ('count', np.ones(n))})
    assert tg[1].tensor.dtype == torch.int8 and tg[1].tensor.shape[1] == 6
    assert tg[1].tensor.numpy().tolist() == keys[0].tolist()
    assert tg[2].tensor.dtype == torch.float64 and tg[2].tensor.shape[1] == 1
    got_f = tg[2].tensor.reshape(-1).numpy()
    assert np.array_equal(np.where(got_f == 0, 0.0, got_f), np.where(keys[1] == 0, 0.0, keys[1]))
    assert len({(a, b) for a, b in zip(map(bytes, keys[0].astype(np.uint8)), keys[1])}) == len(got_f)
    assert tg[21].tensor.tolist() == vals['c']
    np.testing.assert_allclose(tg[20].tensor.numpy(), vals['s'], rtol=1e-12)


def test_year_case_division_and_not_equal_match_closed_forms():
    """year (aggregate.py:97-101) against the calendar, CASE WHEN (aggregate.py:74-85) against
    numpy.where, '/' as true division and '<>' as numpy's !=, each inside a sum."""
    g = np.random.default_rng(43)
    n = 2500
    day = g.integers(-5, 3652, n, dtype=np.int32)
    q = g.integers(1, 51, n).astype(np.float64)
    x = g.integers(-10, 10, n, dtype=np.int64)
    k = g.integers(0, 3, n, dtype=np.int8)
    tg = {1: R.Col(torch.from_numpy(k), 'string'), 2: R.Col(torch.from_numpy(day), 'date'),
          3: R.Col(torch.from_numpy(q), 'float'), 4: R.Col(torch.from_numpy(x), 'int')}
    res = ['k#1',
           [[['sum', [['year', ['d#2']]]], '#', '1'], 'AS', 'y#20'],
           [[['sum', [['CASE', 'WHEN', ['x#4', '>', '0'], 'THEN', 'q#3', 'ELSE', '0.0', 'END']]],
             '#', '2'], 'AS', 'cw#21'],
           [[['sum', [['x#4', '/', '4']]], '#', '3'], 'AS', 'dv#22'],
           [[['sum', [['x#4', '<>', '0']]], '#', '4'], 'AS', 'ne#23']]
    args = _agg_args(['k#1'], res, [20, 21, 22, 23], [1, 2, 3, 4])
    R.tqp_hash_aggregate(tg, args)
    years = np.array([CF.year_of_day(d) if d >= 0 else 1990 for d in day], dtype=np.int64)
    _, vals = CF.group_reduce([k], None, {
        'y': ('sum', years), 'cw': ('sum', np.where(x > 0, q, 0.0)),
        'dv': ('sum', (x / 4).astype(np.float32)), 'ne': ('sum', (x != 0).astype(np.int64))})
    assert tg[20].tensor.tolist() == vals['y'] and tg[20].tensor.dtype == torch.int64
    np.testing.assert_allclose(tg[21].tensor.numpy(), vals['cw'], rtol=1e-12)
    np.testing.assert_allclose(tg[22].tensor.numpy(), vals['dv'], rtol=1e-6)
    assert tg[23].tensor.tolist() == vals['ne']


def test_year_fails_past_1999_as_the_reference_does():
    """A day count of 3652 (2000-01-01) indexes past the 10-entry year table."""
    tg = {1: R.Col(torch.tensor([0, 0], dtype=torch.int8), 'string'),
          2: R.Col(torch.tensor([100, 3652], dtype=torch.int32), 'date')}
    args = _agg_args(['k#1'], ['k#1', [[['sum', [['year', ['d#2']]]], '#', '1'], 'AS', 'y#20']],
                     [20], [1, 2])
    with pytest.raises(IndexError):
        R.tqp_hash_aggregate(tg, args)
    assert CF.year_of_day(3651) == 1999 and CF.year_of_day(3652) == 2000


def test_contains_endswith_substring_match_closed_forms():
    """Contains and EndsWith (filter.py:56-62 via like.py:3-11: the word at any position of the
    row, padding included) against bytes.find over each row; substring (filter.py:71-76) IN /
    '=' against the row's bytes [l - 1, l - 1 + w) compared with the zero-padded literals."""
    g = np.random.default_rng(47)
    n, w = 3000, 15
    s = np.array([ord(c) for c in "abcd "], dtype=np.int8)[g.integers(0, 5, (n, w))]
    s[g.random((n, w)) < 0.1] = 0
    rows = np.flatnonzero(g.random(n) < 0.2)
    for r, p in zip(rows, g.integers(0, w - 4, rows.shape[0])):
        s[r, p:p + 5] = [ord(c) for c in "green"]
    tg = {1: R.Col(torch.from_numpy(s), 'string')}
    raw = [bytes(row.astype(np.uint8)) for row in s]
    for head in ('Contains', 'EndsWith'):
        for word in ("green", "ab", "d a", "x"):
            got = R.filter_evaluate(tg, [head, ['c#1', word]], None, None).numpy()
            want = np.array([r.find(word.encode()) >= 0 for r in raw])
            assert np.array_equal(got, want), (head, word)
    got = R.filter_evaluate(tg, [['substring', ['c#1', '2', '3']], 'IN', ['abc', 'gre', 'ree']],
                            None, None).numpy()
    want = np.array([r[1:4] in (b"abc", b"gre", b"ree") for r in raw])
    assert np.array_equal(got, want)
    got = R.filter_evaluate(tg, [['substring', ['c#1', '1', '2']], '=', 'ab'], None, None).numpy()
    want = np.array([r[0:2] == b"ab" for r in raw])
    assert np.array_equal(got, want) and want.any()


# ------------------------------------------------------------------ project
PROJECT_ARGS = {
    'name': 'Project',
    'rename map': {202: 602},
    'used expr': [
        'price#202',                                                   # a bare column: skipped
        [['discount#203', 'AS', 'disc#603']],                          # a renamed column: skipped
        [[['price#202', '*', ['1.0', '-', 'discount#203']], 'AS', 'rev#604']],
        [[[['price#202', '*', ['1.0', '-', 'discount#203']], '*', ['1.0', '+', 'tax#205']],
          'AS', 'charge#605']],
        [[['year', 'ship_day#204'], 'AS', 'y#606']],
        [[['CASE', 'WHEN', ['q#206', '>', '25'], 'THEN', 'q#206', 'ELSE', '0', 'END'],
          'AS', 'bigq#607']],
        [[['ship_day#204', '<', '1200'], 'AS', 'early#608']],
    ],
    'materialized in': [202, 203, 204, 205, 206],
    'materialized out': [604, 605, 606, 607, 608],
}


def project_tables(g, n):
    return {202: ('float', g.uniform(900, 100000, n)),
            203: ('float', np.round(g.integers(0, 11, n) / 100.0, 2)),
            204: ('date', g.integers(0, 3652, n, dtype=np.int32)),
            205: ('float', np.round(g.integers(0, 9, n) / 100.0, 2)),
            206: ('int', g.integers(1, 51, n).astype(np.int64))}


def test_project_matches_closed_forms():
    """project.py:10-69: bare and renamed columns give no output, the rename map aliases, and
    each evaluated expression is the per-row value (through float64 and back to its dtype)."""
    import copy
    g = np.random.default_rng(81)
    t = project_tables(g, 4000)
    tg = R.to_cols(t)
    assert R.tqp_project(tg, copy.deepcopy(PROJECT_ARGS)) == -1
    p, d, day, tax, q = (t[k][1] for k in (202, 203, 204, 205, 206))
    assert tg[602] is tg[202]
    np.testing.assert_array_equal(tg[604].tensor.numpy(), p * (1.0 - d))
    np.testing.assert_array_equal(tg[605].tensor.numpy(), p * (1.0 - d) * (1.0 + tax))
    assert tg[606].tensor.dtype == torch.int64
    np.testing.assert_array_equal(tg[606].tensor.numpy(), [CF.year_of_day(x) for x in day])
    np.testing.assert_array_equal(tg[607].tensor.numpy(), np.where(q > 25, q, 0))
    assert tg[608].tensor.dtype == torch.bool
    np.testing.assert_array_equal(tg[608].tensor.numpy(), day < 1200)
    assert [tg[k].tensor_type for k in (604, 606, 608)] == ['float', 'int', 'string']
    assert 603 not in tg
