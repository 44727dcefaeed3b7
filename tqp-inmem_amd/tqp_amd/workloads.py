"""Seeded synthetic tables and plan arguments for the benchmark configurations.

Tables are generated on the host with numpy's PCG64 (seed = 20261203 + 1000*config + table) so
the GPU path and the CPU restatement see identical bytes.  Plan arguments follow the schema the
reference's planner emits (parsing.py:304-507): Q1's filter -> aggregate -> sort shape for the
1M-row fact table and Q3's filter -> join -> aggregate -> take-ordered shape for the 6M-row
join; ids 10000+ are filter masks as the planner numbers them (parsing.py:134).
"""
import numpy as np
import torch

from .variable import Variable, VariableState

SEED = 20261203


def rng(config: int, table: int):
    return np.random.Generator(np.random.PCG64(SEED + 1000 * config + table))


# ---------------------------------------------------------------- config 0: filter + 6-group aggregate
def config0_tables(n=1_000_000):
    g = rng(0, 0)
    return {
        1: ('date', g.integers(0, 2557, n, dtype=np.int32)),                     # ship_day
        2: ('float', g.integers(1, 51, n).astype(np.float64)),                   # quantity
        3: ('float', g.uniform(900.0, 105000.0, n)),                             # price
        4: ('float', g.integers(0, 11, n) / 100.0),                              # discount
        5: ('float', g.integers(0, 9, n) / 100.0),                               # tax
        6: ('string', g.integers(0, 3, n, dtype=np.int8)),                       # flag_a
        7: ('string', g.integers(0, 2, n, dtype=np.int8)),                       # flag_b
    }


def config0_plan(ship_day_max=2400):
    agg_ids = [22, 23, 24, 25, 26, 27, 28]
    disc = ['price#3', '*', ['1.0', '-', 'discount#4']]
    return [
        (1, {'name': 'Filter',
             'filter tree': [[['isnotnull', ['ship_day#1']], 'AND', ['ship_day#1', '<=', str(ship_day_max)]]],
             'materialized in': [1], 'out mask id': 10000}),
        (2, {'name': 'Aggregate', 'input ids': [2, 3, 4, 5, 6, 7],
             'keys': ['flag_a#6', 'flag_b#7'],
             'mask map': {2: 10000, 3: 10000, 4: 10000, 5: 10000, 6: 10000, 7: 10000},
             'results': [
                 'flag_a#6', 'flag_b#7',
                 [[[['sum', ['quantity#2']], '#', '30'], 'AS', 'sum_qty#22']],
                 [[[['sum', ['price#3']], '#', '31'], 'AS', 'sum_base_price#23']],
                 [[[['sum', [disc]], '#', '32'], 'AS', 'sum_disc_price#24']],
                 [[[['sum', [[disc, '*', ['1.0', '+', 'tax#5']]]], '#', '33'], 'AS', 'sum_charge#25']],
                 [[[['avg', ['quantity#2']], '#', '34'], 'AS', 'avg_qty#26']],
                 [[[['avg', ['discount#4']], '#', '35'], 'AS', 'avg_disc#27']],
                 [[[['count', ['1']], '#', '36'], 'AS', 'count_order#28']],
             ],
             'out agg ids': agg_ids, 'is partial': False}),
        (3, {'name': 'Sort', 'input ids': [6, 7] + agg_ids, 'key ids': [6, 7],
             'key orders': [False, False]}),
    ]


# ---------------------------------------------------------------- config 1: filter -> join -> aggregate -> top-k
def config1_tables(n_build=1_500_000, n_probe=6_000_000):
    gb = rng(1, 0)
    gp = rng(1, 1)
    keys = gb.permutation(np.arange(1, n_build + 1, dtype=np.int32))
    return {
        101: ('int', keys),                                                       # key
        102: ('string', gb.integers(0, 5, n_build, dtype=np.int8)),               # segment
        103: ('date', gb.integers(0, 2557, n_build, dtype=np.int32)),             # order_day
        201: ('int', gp.integers(1, n_build + 1, n_probe, dtype=np.int32)),       # fk
        202: ('float', gp.uniform(900.0, 105000.0, n_probe)),                     # price
        203: ('float', gp.integers(0, 11, n_probe) / 100.0),                      # discount
        204: ('date', gp.integers(0, 2557, n_probe, dtype=np.int32)),             # ship_day
    }


def config1_plan(segment=2, ship_day_min=1200, limit=10):
    return [
        (1, {'name': 'Filter',
             'filter tree': [[['isnotnull', ['segment#102']], 'AND', ['segment#102', '=', str(segment)]]],
             'materialized in': [102], 'out mask id': 10001}),
        (2, {'name': 'Filter',
             'filter tree': [[['isnotnull', ['ship_day#204']], 'AND', ['ship_day#204', '>', str(ship_day_min)]]],
             'materialized in': [204], 'out mask id': 10002}),
        (3, {'name': 'Join', 'build': 'left', 'left group': [101, 102, 103],
             'right group': [201, 202, 203, 204], 'left keys': [101], 'right keys': [201],
             'mask map': {101: 10001, 201: 10002, 202: 10002, 203: 10002},
             'materialized in': [101, 201, 202, 203], 'join type': 'Inner'}),
        (4, {'name': 'Aggregate', 'input ids': [101, 202, 203], 'keys': ['key#101'],
             'results': ['key#101',
                         [[[['sum', [['price#202', '*', ['1.0', '-', 'discount#203']]]], '#', '40'],
                           'AS', 'revenue#301']]],
             'out agg ids': [301], 'is partial': False}),
        (5, {'name': 'take ordered', 'input ids': [101, 301], 'key ids': [301], 'key orders': [True],
             'limit': limit}),
    ]


# ---------------------------------------------------------------- config 2: 3-join star with Zipf skew
def _zipf_ranks(g, n, nkeys, s=1.1):
    w = 1.0 / np.power(np.arange(1, nkeys + 1, dtype=np.float64), s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, g.random(n)), nkeys - 1)


def config2_tables(n_fact=60_000_000, n_a=15_000_000, n_b=2_000_000, n_c=25):
    ga, gbt, gc, gf = rng(2, 0), rng(2, 1), rng(2, 2), rng(2, 3)
    a_key = ga.permutation(np.arange(1, n_a + 1, dtype=np.int32))
    b_key = gbt.permutation(np.arange(1, n_b + 1, dtype=np.int32))
    b_attr = gbt.integers(0, 1000, n_b, dtype=np.int32)
    c_key = np.arange(1, n_c + 1, dtype=np.int32)
    c_attr = (gc.permutation(n_c) * 40).astype(np.int32)   # 25 distinct values
    hot = _zipf_ranks(gf, n_fact, n_b)
    return {
        301: ('int', a_key), 302: ('int', ga.integers(0, 10, n_a, dtype=np.int32)),
        311: ('int', b_key), 312: ('int', b_attr), 313: ('int', (b_attr % 100).astype(np.int32)),
        321: ('int', c_key), 322: ('int', c_attr),
        331: ('int', gf.integers(1, n_a + 1, n_fact, dtype=np.int32)),            # fk_a
        332: ('int', b_key[hot]),                                                 # fk_b (Zipf)
        333: ('int', gf.integers(1, n_c + 1, n_fact, dtype=np.int32)),            # fk_c
        334: ('float', gf.lognormal(8.0, 1.0, n_fact)),                           # amount
    }


def config2_plan():
    return [
        (1, {'name': 'Join', 'build': 'left', 'left group': [301, 302],
             'right group': [331, 332, 333, 334], 'left keys': [301], 'right keys': [331],
             'materialized in': [301, 331, 332, 333, 334], 'join type': 'Inner'}),
        (2, {'name': 'Join', 'build': 'left', 'left group': [311, 312, 313],
             'right group': [301, 331, 332, 333, 334], 'left keys': [311], 'right keys': [332],
             'materialized in': [311, 313, 332, 333, 334], 'join type': 'Inner'}),
        (3, {'name': 'Join', 'build': 'left', 'left group': [321, 322],
             'right group': [311, 313, 332, 333, 334], 'left keys': [321], 'right keys': [333],
             'materialized in': [321, 322, 313, 333, 334], 'join type': 'Inner'}),
        (4, {'name': 'Aggregate', 'input ids': [322, 313, 334], 'keys': ['c_attr#322', 'b_mod#313'],
             'results': ['c_attr#322', 'b_mod#313',
                         [[[['sum', ['amount#334']], '#', '50'], 'AS', 'total#401']],
                         [[[['count', ['1']], '#', '51'], 'AS', 'cnt#402']]],
             'out agg ids': [401, 402], 'is partial': False}),
        (5, {'name': 'Sort', 'input ids': [322, 313, 401, 402], 'key ids': [322, 313],
             'key orders': [False, False]}),
    ]


# ---------------------------------------------------------------- config 3: batched pipelines
def config3_sizes(n_pipes=256, lo=10_000, hi=6_000_000):
    """Probe row counts log-uniform in [lo, hi]; build side a quarter of each."""
    g = rng(3, 0)
    probe = np.exp(g.uniform(np.log(lo), np.log(hi), n_pipes)).astype(np.int64)
    build = np.maximum(probe // 4, 1)
    consts = g.integers(0, 2557, (n_pipes, 2))
    return probe, build, consts


def config3_pipeline(i, n_probe, n_build, c_build, c_probe, device):
    """Tables (generated on `device` from a per-pipeline seed) and config-1-shaped plan of
    pipeline i: build filter order_day <= c_build, probe filter ship_day > c_probe, join on
    key = fk, group by key, sum(price * (1 - discount)) and count(*)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(SEED + 3000 + int(i))
    nb, np_ = int(n_build), int(n_probe)
    kw = dict(device=device, generator=gen)
    group = {
        101: Variable((torch.randperm(nb, **kw) + 1).to(torch.int32), 'int', None, VariableState.GPU),
        103: Variable(torch.randint(0, 2557, (nb,), dtype=torch.int32, **kw), 'date', None, VariableState.GPU),
        201: Variable(torch.randint(1, nb + 1, (np_,), dtype=torch.int32, **kw), 'int', None, VariableState.GPU),
        202: Variable(torch.rand(np_, dtype=torch.float64, **kw) * 104100.0 + 900.0, 'float', None, VariableState.GPU),
        203: Variable(torch.randint(0, 11, (np_,), dtype=torch.int64, **kw).to(torch.float64) / 100.0, 'float', None, VariableState.GPU),
        204: Variable(torch.randint(0, 2557, (np_,), dtype=torch.int32, **kw), 'date', None, VariableState.GPU),
    }
    plan = [
        (1, {'name': 'Filter', 'filter tree': [['order_day#103', '<=', str(int(c_build))]],
             'materialized in': [103], 'out mask id': 10001}),
        (2, {'name': 'Filter', 'filter tree': [['ship_day#204', '>', str(int(c_probe))]],
             'materialized in': [204], 'out mask id': 10002}),
        (3, {'name': 'Join', 'build': 'left', 'left group': [101, 103],
             'right group': [201, 202, 203, 204], 'left keys': [101], 'right keys': [201],
             'mask map': {101: 10001, 201: 10002, 202: 10002, 203: 10002},
             'materialized in': [101, 201, 202, 203], 'join type': 'Inner'}),
        (4, {'name': 'Aggregate', 'input ids': [101, 202, 203], 'keys': ['key#101'],
             'results': ['key#101',
                         [[[['sum', [['price#202', '*', ['1.0', '-', 'discount#203']]]], '#', '40'],
                           'AS', 'revenue#301']],
                         [[[['count', ['1']], '#', '41'], 'AS', 'cnt#302']]],
             'out agg ids': [301, 302], 'is partial': False}),
    ]
    return group, plan


# ---------------------------------------------------------------- helpers
def to_tensor_group(tables, device, stats=True):
    """{id: (type, ndarray)} -> {id: Variable} on `device`.  With stats, the minimum and maximum
    of every 1-D integer column are taken on the host while it is loaded and recorded as the
    tensor's value bounds (column statistics, as a columnar file's footer carries them), which
    lets a join build over that column pick its layout without a device round trip."""
    from . import kernels
    group = {}
    for cid, (ty, arr) in tables.items():
        a = np.ascontiguousarray(arr)
        t = torch.from_numpy(a).to(device)
        if stats and a.ndim == 1 and a.size and a.dtype.kind in "iu" and t.is_cuda:
            kernels.note_bounds(t, int(a.min()), int(a.max()))
        group[cid] = Variable(t, ty, None, VariableState.GPU if t.is_cuda else VariableState.CPU)
    return group


def input_rows(tables, config: int) -> int:
    """Base-table rows a pipeline consumes (the rows/s numerator)."""
    if config == 0:
        return len(tables[1][1])
    if config == 1:
        return len(tables[101][1]) + len(tables[201][1])
    if config == 2:
        return sum(len(tables[i][1]) for i in (301, 311, 321, 331))
    raise ValueError(config)
