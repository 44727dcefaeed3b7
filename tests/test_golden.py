"""Golden fixtures (tests/golden/*.json, written by make_golden.py from the closed forms): the CPU
oracle must reproduce them (CPU tests) and so must the HIP path (GPU tests)."""
import json
import os

import numpy as np
import pytest
import torch

from oracle import ref_ops as R

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def _f64(xs):
    return np.array([np.nan if x is None else x for x in xs], dtype=np.float64)


# ------------------------------------------------------------------ oracle vs fixtures (CPU)
def test_oracle_join_fixture():
    d = load("join_small.json")
    tg = {1: R.Col(torch.tensor(d["bkeys"], dtype=torch.int32), 'int'),
          2: R.Col(torch.arange(len(d["bkeys"])), 'int'),
          3: R.Col(torch.tensor(d["pkeys"], dtype=torch.int32), 'int'),
          4: R.Col(torch.arange(len(d["pkeys"])), 'int'),
          901: R.Col(torch.tensor(d["bmask"]), 'bool'), 902: R.Col(torch.tensor(d["pmask"]), 'bool')}
    # row ids of the compacted sides map back to input rows through these position columns
    args = {'left group': [1, 2], 'right group': [3, 4], 'materialized in': [1, 2, 3, 4],
            'mask map': {1: 901, 2: 901, 3: 902, 4: 902}}
    R.join_vortex(tg, args, 1, 3, [], 'inner')
    assert list(zip(tg[2].tensor.tolist(), tg[4].tensor.tolist())) == [tuple(p) for p in d["pairs"]]


def test_oracle_groupby_fixture():
    d = load("groupby_small.json")
    tg = {1: R.Col(torch.tensor(d["k1"], dtype=torch.int8), 'string'),
          2: R.Col(torch.tensor(d["k2"], dtype=torch.int32), 'date'),
          3: R.Col(torch.tensor(d["v"], dtype=torch.float64), 'float'),
          4: R.Col(torch.tensor(d["d"], dtype=torch.float64), 'float'),
          900: R.Col(torch.tensor(d["mask"]), 'bool')}
    args = {'keys': ['a#1', 'b#2'], 'is partial': False, 'input ids': [1, 2, 3, 4],
            'mask map': {i: 900 for i in (1, 2, 3, 4)}, 'out agg ids': [10, 11, 12],
            'results': ['a#1', 'b#2', [[['sum', ['v#3']], '#', '1'], 'AS', 's#10'],
                        [[['avg', ['d#4']], '#', '2'], 'AS', 'a#11'],
                        [[['count', ['1']], '#', '3'], 'AS', 'c#12']]}
    R.tqp_hash_aggregate(tg, args)
    assert tg[1].tensor.reshape(-1).tolist() == d["keys"][0]
    assert tg[2].tensor.reshape(-1).tolist() == d["keys"][1]
    assert tg[12].tensor.tolist() == d["cnt"]
    np.testing.assert_allclose(tg[10].tensor.numpy(), d["sum_v"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(tg[11].tensor.numpy(), d["avg_d"], rtol=1e-12, atol=0)


def test_oracle_sort_fixture():
    d = load("sort_small.json")
    f = torch.from_numpy(_f64(d["f"]))
    i = torch.tensor(d["i"], dtype=torch.int64)
    for descs, key in (([True, False], "perm_desc_asc"), ([False, True], "perm_asc_desc")):
        tg = {1: R.Col(f.clone(), 'float'), 2: R.Col(i.clone(), 'int'),
              3: R.Col(torch.arange(len(d["i"])), 'int')}
        R.tqp_sort(tg, {'input ids': [1, 2, 3], 'key ids': [1, 2], 'key orders': descs}, False)
        assert tg[3].tensor.tolist() == d[key]


def test_oracle_filter_fixture():
    d = load("filter_small.json")
    t = {1: R.Col(torch.tensor(d["x"], dtype=torch.float64), 'float'),
         2: R.Col(torch.tensor(d["y"], dtype=torch.int32), 'date')}
    for case in d["cases"]:
        assert R.filter_evaluate(t, case["tree"], None, None).tolist() == case["mask"], case["tree"]


# ------------------------------------------------------------------ HIP path vs fixtures (GPU)
@pytest.mark.gpu
def test_gpu_join_fixture(device):
    from tqp_amd import kernels
    d = load("join_small.json")
    bk = torch.tensor(d["bkeys"], dtype=torch.int32, device=device)
    pk = torch.tensor(d["pkeys"], dtype=torch.int32, device=device)
    idx = kernels.JoinIndex(bk, torch.tensor(d["bmask"], device=device))
    _, _, b, p = idx.probe(pk, torch.tensor(d["pmask"], device=device), [], [], want_index=True)
    assert list(zip(b.tolist(), p.tolist())) == [tuple(x) for x in d["pairs"]]


@pytest.mark.gpu
def test_gpu_groupby_fixture(device):
    from tqp_amd.operators import aggregate
    from tqp_amd.variable import Variable
    d = load("groupby_small.json")
    tg = {1: Variable(torch.tensor(d["k1"], dtype=torch.int8, device=device), 'string'),
          2: Variable(torch.tensor(d["k2"], dtype=torch.int32, device=device), 'date'),
          3: Variable(torch.tensor(d["v"], dtype=torch.float64, device=device), 'float'),
          4: Variable(torch.tensor(d["d"], dtype=torch.float64, device=device), 'float'),
          900: Variable(torch.tensor(d["mask"], device=device), 'bool')}
    args = {'keys': ['a#1', 'b#2'], 'is partial': False, 'input ids': [1, 2, 3, 4],
            'mask map': {i: 900 for i in (1, 2, 3, 4)}, 'out agg ids': [10, 11, 12],
            'results': ['a#1', 'b#2', [[['sum', ['v#3']], '#', '1'], 'AS', 's#10'],
                        [[['avg', ['d#4']], '#', '2'], 'AS', 'a#11'],
                        [[['count', ['1']], '#', '3'], 'AS', 'c#12']]}
    aggregate.tqp_hash_aggregate(True, tg, args, None, None)
    assert tg[1].tensor.reshape(-1).tolist() == d["keys"][0]
    assert tg[2].tensor.reshape(-1).tolist() == d["keys"][1]
    assert tg[12].tensor.tolist() == d["cnt"]
    np.testing.assert_allclose(tg[10].tensor.cpu().numpy(), d["sum_v"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(tg[11].tensor.cpu().numpy(), d["avg_d"], rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_gpu_sort_fixture(device):
    from tqp_amd import kernels
    d = load("sort_small.json")
    f = torch.from_numpy(_f64(d["f"])).to(device)
    i = torch.tensor(d["i"], dtype=torch.int64, device=device)
    for descs, key in (([True, False], "perm_desc_asc"), ([False, True], "perm_asc_desc")):
        perm = kernels.sort_perm([(f, descs[0]), (i, descs[1])], len(d["i"]))
        assert perm.tolist() == d[key]


@pytest.mark.gpu
def test_gpu_filter_fixture(device):
    from tqp_amd import kernels
    from tqp_amd.variable import Variable
    d = load("filter_small.json")
    t = {1: Variable(torch.tensor(d["x"], dtype=torch.float64, device=device), 'float'),
         2: Variable(torch.tensor(d["y"], dtype=torch.int32, device=device), 'date')}
    for case in d["cases"]:
        mask, _ = kernels.filter_mask(case["tree"], t, None)
        assert mask.tolist() == case["mask"], case["tree"]
