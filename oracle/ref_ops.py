"""CPU restatement of the reference engine's hot-path operators (test infrastructure only).

This module is the parity oracle and the CPU baseline.  It restates, on torch CPU tensors, the
algorithms of the reference's Python operators, each function citing the file:line it follows:

  filter_evaluate / tqp_filter    operators/filter.py:14-162 / 164-205
  like_mask                       operators/like.py:13-25
  apply_masks                     IO/vortex_pipeline.py:51-70
  join_kernel / join_vortex       operators/hashjoin.py:24-177 / 179-248
  agg_evaluate / tqp_hash_aggregate  operators/aggregate.py:58-233 / 236-393
  tqp_sort                        operators/sort.py:6-58
  execute                         main.py:130-274 (operator dispatch)

Two places are made canonical, because the reference's own result there is unspecified:
  * the build-side sort of the join is stable (the reference's torch.sort(stable=False),
    hashjoin.py:43, leaves the order inside a duplicate-key bucket unspecified);
  * groups are numbered by an exact sorted unique of the key rows (the reference ranks a linear
    hash with random coefficients, aggregate.py:320-327, whose order is arbitrary and which can
    merge distinct keys).
Everything else -- dtype promotion, float literal eps, isclose, integer sums through float64,
the 2-step privatised sum path, max initialised to 0, count skipping -64 -- follows the
reference.  Nothing in the product imports this module; only tests, `smoke()` and the CPU
baseline of bench.py do.
"""
import time

import numpy as np
import torch

EPS = 1e-5
NULL = -64
STRING = torch.int8
INT = torch.int64
FLOAT = torch.float64
DATE = torch.int32
_EPOCH_ORD = np.datetime64('1990-01-01', 'D')


class Col:
    """Minimal column: tensor + SQL type name (variable.py:9-43)."""

    def __init__(self, tensor, tensor_type):
        self.tensor = tensor
        self.tensor_type = tensor_type

    def normalize(self):
        if self.tensor.dim() == 2:
            self.tensor = self.tensor.squeeze()
        if self.tensor.dim() == 0:
            self.tensor = self.tensor.unsqueeze(0)


def _type_of(dtype):
    if dtype == FLOAT:
        return 'float'
    if dtype == DATE:
        return 'date'
    if dtype == INT:
        return 'int'
    return 'string'


def get_id(name):
    return int(name.split('#')[-1].replace('L', ''))


def right_most(tree):
    return right_most(tree[-1]) if isinstance(tree, list) else tree


def is_date(s):
    return len(s) == 10 and s.count('-') == 2 and s.replace('-', '').isdigit()


def is_float(s):
    return s.replace('.', '').replace('-', '').replace('E', '').isdigit()


def date_days(s):
    return int((np.datetime64(s, 'D') - _EPOCH_ORD).astype(int))


def str_codes(s, n):
    return torch.tensor([ord(c) for c in s] + [0] * (n - len(s)), dtype=STRING)


# ====================================================================== filter
def filter_evaluate(tensors, tree, rows, subquery_result, leaf_function=None):
    """operators/filter.py:14-162."""
    if rows is None:
        rows = list(tensors.values())[0].tensor.size(0)
    if isinstance(tree, str):
        if tree == 'Subquery':
            return subquery_result
        if is_date(tree):
            return date_days(tree)
        if is_float(tree):
            return float(tree) if '.' in tree else int(tree)
        if '#' in tree and not tree.startswith('Brand#'):
            if leaf_function:
                return leaf_function(get_id(tree))
            return tensors[get_id(tree)].tensor.squeeze()
        return tree
    if len(tree) == 1:
        return filter_evaluate(tensors, tree[0], rows, subquery_result, leaf_function)
    if len(tree) == 3 and tree[1] == '#':
        if leaf_function:
            return leaf_function(int(tree[-1]))
        return tensors[int(tree[-1])].tensor.squeeze()
    if len(tree) == 2:
        head = tree[0]
        if head == 'StartsWith':
            t = tensors[get_id(tree[1][0])].tensor
            name = tree[1][1]
            return torch.eq(t[:, :len(name)], str_codes(name, len(name))).all(dim=1)
        if head in ('EndsWith', 'Contains'):
            # filter.py:56-62: both through like_contains
            t = tensors[get_id(tree[1][0])].tensor
            return like_contains(t, tree[1][1])
        if head == 'substring':
            # filter.py:71-76
            assert len(tree[1]) == 3
            sub, l, w = tree[1]
            l, w = int(l) - 1, int(w)
            return filter_evaluate(tensors, sub, rows, subquery_result, leaf_function)[:, l:l + w]
        if head == 'NOT':
            return ~filter_evaluate(tensors, tree[1], rows, subquery_result, leaf_function)
        if head == 'isnotnull':
            return torch.ones(rows, dtype=torch.bool)
        if head == 'cast':
            literal, _, cast_type = tree[1][0].split(' ')
            if cast_type == 'double':
                return filter_evaluate(tensors, literal, rows, subquery_result, leaf_function)
        raise ValueError(f"Unsupported Function: {head}")
    op = tree[-2]
    if op == 'IN':
        lhs = filter_evaluate(tensors, tree[0], rows, subquery_result, leaf_function)
        if lhs.dtype == STRING and lhs.dim() == 2:
            mask = torch.zeros(rows, dtype=torch.bool)
            for r in tree[-1]:
                mask = torch.logical_or(mask, (lhs == str_codes(r, lhs.shape[-1])).all(dim=1))
            return mask
        vals = [int(r) for r in tree[-1]] if not lhs.is_floating_point() else [float(r) for r in tree[-1]]
        mask = torch.zeros(rows, dtype=torch.bool)
        for r in vals:
            mask = torch.logical_or(mask, torch.eq(lhs, r))
        return mask
    if op == 'LIKE':
        lhs = filter_evaluate(tensors, tree[0], rows, subquery_result, leaf_function)
        assert tree[-1].count('%') == 3
        _, w1, w2, _ = tree[-1].split('%')
        return like_mask(lhs, w1, w2)
    ops = {'<=': torch.le, '>=': torch.ge, '<': torch.lt, '>': torch.gt, '=': torch.eq,
           'AND': torch.logical_and, 'OR': torch.logical_or}
    if op not in ops:
        raise ValueError(f"Unsupported Operation: {op}")
    lhs = filter_evaluate(tensors, tree[0], rows, subquery_result, leaf_function)
    rhs = filter_evaluate(tensors, tree[-1], rows, subquery_result, leaf_function)
    if isinstance(rhs, float):
        if op == '<=':
            return torch.le(lhs, rhs + EPS)
        if op == '>=':
            return torch.ge(lhs, rhs - EPS)
    if isinstance(rhs, str):
        if len(rhs) == 1:
            return ops[op](lhs, ord(rhs))
        return (lhs == str_codes(rhs, lhs.shape[-1])).all(dim=1)
    if op == '=' and lhs.dtype == FLOAT:
        other = torch.full((lhs.shape[0],), rhs, dtype=lhs.dtype) if isinstance(rhs, float) else rhs
        return torch.isclose(lhs, other.to(lhs.dtype) if torch.is_tensor(other) else other, atol=EPS)
    return ops[op](lhs, rhs)


def like_contains(a, word):
    """operators/like.py:3-11: rows of the [n, m] string tensor `a` that hold `word` at some
    position pos in [0, m - len(word)], padding bytes included."""
    p1 = str_codes(word, len(word))
    n1 = p1.shape[0]
    n, m = a.shape
    appear = torch.zeros(n, dtype=torch.bool)
    for pos in range(m - n1 + 1):
        appear.logical_or_((a[:, pos:pos + n1] == p1).all(dim=-1))
    return appear


def like_mask(a, w1, w2):
    """operators/like.py:13-25 ('%w1%w2%', filter.py:121-127): rows of the [n, m] string tensor
    with w1 at some position q and w2 at a position >= q + len(w1), both windows inside the m
    bytes (padding included).  Restated with window matches over every start position and a
    running OR of the w1 hits."""
    x = a.numpy()
    n, m = x.shape
    n1, n2 = len(w1), len(w2)
    if m - n1 - n2 + 1 <= 0:
        return torch.zeros(n, dtype=torch.bool)

    def hits(w):
        if not w:
            return np.ones((n, m + 1), dtype=bool)
        win = np.lib.stride_tricks.sliding_window_view(x, len(w), axis=1)
        return (win == str_codes(w, len(w)).numpy()).all(axis=-1)

    npos = m - n1 - n2 + 1
    seen1 = np.logical_or.accumulate(hits(w1)[:, :npos], axis=1)
    at2 = hits(w2)[:, n1:n1 + npos]
    return torch.from_numpy((seen1 & at2).any(axis=1))


def tqp_filter(tensor_group, args, subquery_result=None):
    """operators/filter.py:164-205: stores the bool mask under args['out mask id']."""
    assert 'mask map' not in args
    cols = {k: tensor_group[k] for k in args['materialized in']}
    if not cols:
        return -1
    mask = filter_evaluate(cols, args['filter tree'], None, subquery_result)
    if mask.dim() == 2:
        mask = mask.squeeze(0)
    tensor_group[args['out mask id']] = Col(mask, 'bool')
    return -1


# ====================================================================== pipeline mask application
def apply_masks(input_columns, mask_map):
    """IO/vortex_pipeline.py:51-70: masked columns become tensor[mask]; masks are dropped."""
    mask_ids = set(mask_map.values())
    out = {}
    for k, v in input_columns.items():
        if k in mask_ids:
            continue
        if k in mask_map:
            out[k] = Col(v.tensor[input_columns[mask_map[k]].tensor], v.tensor_type)
        else:
            out[k] = Col(v.tensor, v.tensor_type)
    return out


def _all_empty(input_columns):
    return all(v.tensor.numel() == 0 for v in input_columns.values())


# ====================================================================== join
def join_kernel(tensors_left, tensors_right, left_key, right_key, condition_list,
                join_output_keys, type):
    """operators/hashjoin.py:24-177 with a stable build sort."""
    tensors_left[left_key].normalize()
    tensors_right[right_key].normalize()
    left = tensors_left[left_key].tensor
    right = tensors_right[right_key].tensor
    nL, nR = left.size(0), right.size(0)
    sorted_keys, perm = torch.sort(left, stable=True)
    sorted_ids = torch.arange(nL, dtype=torch.int64)[perm]
    uniq, counts = torch.unique_consecutive(sorted_keys, return_counts=True)
    starts = torch.cumsum(counts, 0) - counts
    pos = torch.searchsorted(uniq, right)
    in_range = pos < uniq.size(0)
    hit = torch.zeros(nR, dtype=torch.bool)
    hit[in_range] = uniq[pos[in_range]] == right[in_range]
    m_idx = torch.nonzero(hit).view(-1)
    if m_idx.numel() > 0:
        b = pos[m_idx]
        lens = counts[b]
        # expand each matched probe row over its bucket
        reps = torch.repeat_interleave(m_idx, lens)
        base = torch.repeat_interleave(starts[b], lens)
        within = torch.arange(reps.numel()) - torch.repeat_interleave(torch.cumsum(lens, 0) - lens, lens)
        cs = sorted_ids[base + within]
        cr = reps
        keep = left[cs] == right[cr]
        for cond in condition_list:
            def leaf(cid, cs=cs, cr=cr):
                if cid in tensors_left:
                    return tensors_left[cid].tensor[cs].squeeze()
                return tensors_right[cid].tensor[cr].squeeze()
            keep &= filter_evaluate(None, cond, cs.size(0), None, leaf)
        sel = torch.nonzero(keep).view(-1)
        lm, rm = cs[sel], cr[sel]
    else:
        lm = torch.tensor([], dtype=torch.int64)
        rm = torch.tensor([], dtype=torch.int64)
    nulls = 0
    if type == 'inner':
        lo, ro = lm, rm
    elif type == 'right-semi':
        ro = torch.unique(rm) if rm.numel() else torch.tensor([], dtype=torch.int64)
        lo = torch.tensor([], dtype=torch.int64)
    elif type == 'right-anti':
        allr = torch.ones(nR, dtype=torch.bool)
        if rm.numel():
            allr[torch.unique(rm)] = False
        ro = torch.nonzero(allr).view(-1)
        lo = torch.tensor([], dtype=torch.int64)
    elif type == 'right-outer':
        allr = torch.ones(nR, dtype=torch.bool)
        if rm.numel():
            allr[torch.unique(rm)] = False
        un = torch.nonzero(allr).view(-1)
        lo, ro = lm, torch.cat([rm, un])
        nulls = un.size(0)
    else:
        raise NotImplementedError("unhandled join type")

    def sel_left(x, idx):
        if type == 'right-outer' and idx.numel() > 0:
            full = torch.cat([idx, torch.full((nulls,), NULL, dtype=torch.int64)])
            msk = full == NULL
            out = x[full.masked_fill(msk, 0)]
            return out.masked_fill_(msk, NULL)
        return x[idx]

    res = []
    for k in join_output_keys:
        if k in tensors_left:
            t = tensors_left[k].tensor
            res.append(sel_left(t, lo) if t.shape[0] > 0 else torch.tensor([]))
        else:
            t = tensors_right[k].tensor
            res.append(t[ro] if t.shape[0] > 0 else torch.tensor([]))
    return res


def join_vortex(tensor_group, args, leftcol, rightcol, condition_list, type):
    """operators/hashjoin.py:179-248."""
    groups = [args['left group'], args['right group']]
    keys = [leftcol, rightcol]
    jin, jout, jmask = {}, [], {}
    for side in (0, 1):
        tensor_group[keys[side]].normalize()
        for c in args['materialized in']:
            if c in groups[side]:
                jin[c] = tensor_group[c]
                jout.append(c)
                if 'mask map' in args and c in args['mask map']:
                    mid = args['mask map'].pop(c)
                    jmask[c] = mid
                    jin[mid] = tensor_group[mid]
    if _all_empty(jin):
        res = [torch.tensor([]) for _ in jout]
    else:
        cols = apply_masks(jin, jmask)
        sides = [{k: cols[k] for k in groups[s] if k in cols} for s in (0, 1)]
        res = join_kernel(sides[0], sides[1], keys[0], keys[1], condition_list, jout, type)
    for i, k in enumerate(jout):
        tensor_group[k] = Col(res[i], tensor_group[k].tensor_type)
    return next(iter(tensor_group.values())).tensor.size(0)


# ====================================================================== aggregate
def agg_evaluate(tensors, tree, cnt_rows, cnt_groups, inverted_index):
    """operators/aggregate.py:58-233 (the 2-step privatised path included)."""
    if isinstance(tree, str):
        if tree.isdigit():
            return int(tree)
        if is_float(tree):
            return float(tree)
        if '#' in tree and tree.split('#')[-1].replace('L', '').isdigit():
            return tensors[get_id(tree)].tensor
        return tree
    if len(tree) == 1:
        return agg_evaluate(tensors, tree[0], cnt_rows, cnt_groups, inverted_index)
    if len(tree) == 3 and tree[1] == '#':
        return agg_evaluate(tensors, tree[0], cnt_rows, cnt_groups, inverted_index)
    if len(tree) == 8:
        m = agg_evaluate(tensors, tree[2], cnt_rows, cnt_groups, inverted_index)
        a = agg_evaluate(tensors, tree[4], cnt_rows, cnt_groups, inverted_index)
        b = agg_evaluate(tensors, tree[6], cnt_rows, cnt_groups, inverted_index)
        return torch.where(m, a, b)
    if len(tree) == 2:
        head = tree[0]
        if head == 'year':
            # aggregate.py:97-101 with conversion.py:10-12: bucketize on the cumulative year
            # lengths 1990..1999 indexes the 10-entry year_range, so a day count of 3652
            # (2000-01-01) or more indexes past its end and fails
            r = agg_evaluate(tensors, tree[1], cnt_rows, cnt_groups, inverted_index)
            edges = torch.tensor([365, 730, 1096, 1461, 1826, 2191, 2557, 2922, 3287, 3652])
            return torch.tensor(list(range(1990, 2000)))[torch.bucketize(r, edges, right=True)]
        if head == 'NOT':
            return ~agg_evaluate(tensors, tree[1], cnt_rows, cnt_groups, inverted_index)
        if head == 'substring':
            # aggregate.py:115-120
            assert len(tree[1]) == 3
            sub, l, w = tree[1]
            l, w = int(l) - 1, int(w)
            return agg_evaluate(tensors, sub, cnt_rows, cnt_groups, inverted_index)[:, l:l + w]
        if head == 'cast':
            literal, _, cast_type = tree[1][0].split(' ')
            if cast_type == 'double':
                return agg_evaluate(tensors, literal, cnt_rows, cnt_groups, inverted_index)
            raise ValueError(cast_type)
        if head not in ('sum', 'avg', 'count', 'min', 'max'):
            raise ValueError(f"Unsupported aggregation function: {head}")
        x = agg_evaluate(tensors, tree[1], cnt_rows, cnt_groups, inverted_index)
        two_step = cnt_rows / cnt_groups >= (1 << 16)
        sub = max(1, cnt_rows // (cnt_groups << 6)) if two_step else 1
        if isinstance(x, int):
            x = torch.full((cnt_rows,), x, dtype=INT)
        elif isinstance(x, float):
            x = torch.full((cnt_rows,), x, dtype=FLOAT)

        def two(x, reduce, init):
            n = inverted_index.size(0)
            flat = inverted_index * sub + torch.arange(n) % sub
            acc = torch.full((cnt_groups * sub,), init, dtype=x.dtype)
            acc = torch.scatter_reduce(acc, 0, flat, x, reduce=reduce).view(cnt_groups, sub)
            return {'amax': acc.max(dim=1).values, 'amin': acc.min(dim=1).values,
                    'sum': acc.sum(dim=1)}[reduce]

        if head == 'min':
            mx = torch.finfo(x.dtype).max if x.is_floating_point() else torch.iinfo(x.dtype).max
            if two_step:
                return two(x, 'amin', mx)
            out = torch.full((cnt_groups,), mx, dtype=x.dtype)
            return torch.scatter_reduce(out, 0, inverted_index, x, reduce='amin')
        if head == 'max':
            if two_step:
                return two(x, 'amax', 0)
            out = torch.zeros(cnt_groups, dtype=x.dtype)
            return torch.scatter_reduce(out, 0, inverted_index, x, reduce='amax')
        if head == 'avg':
            if two_step:
                s = two(x, 'sum', 0).to(torch.float64)
                c = two(torch.ones(inverted_index.size(0), dtype=INT), 'sum', 0)
                return s.div_(c)
            s = torch.bincount(inverted_index, weights=x, minlength=cnt_groups)
            return s.div_(torch.bincount(inverted_index, minlength=cnt_groups))
        if head == 'count':
            x = (x != NULL).to(INT)
        r = two(x, 'sum', 0) if two_step else torch.bincount(inverted_index, weights=x, minlength=cnt_groups)
        if not torch.is_floating_point(x):
            r = torch.round(r).to(INT)
        return r
    op = tree[-2]
    if op == 'AS':
        return agg_evaluate(tensors, tree[0], cnt_rows, cnt_groups, inverted_index)
    ops = {'+': torch.add, '-': torch.sub, '*': torch.mul, '/': torch.div, '<=': torch.le,
           '>=': torch.ge, '<': torch.lt, '>': torch.gt, '=': torch.eq, '<>': torch.ne,
           'AND': torch.logical_and, 'OR': torch.logical_or}
    if op not in ops:
        raise ValueError(f"Unsupported Operation: {op}")
    a = agg_evaluate(tensors, tree[0], cnt_rows, cnt_groups, inverted_index)
    b = agg_evaluate(tensors, tree[-1], cnt_rows, cnt_groups, inverted_index)
    if isinstance(b, str):
        if len(b) == 1:
            return ops[op](a, ord(b))
        return (a == str_codes(b, a.shape[-1])).all(dim=1)
    return ops[op](a, b)


def _key_ordinals(t):
    """int64 columns whose lexicographic order and equality are those of the key values:
    integers as they are, 2-D string rows byte by byte, floats by an order-preserving map of
    their bits with -0.0 taken as +0.0 and every NaN alike, above +inf (torch.sort's order)."""
    t = t if t.dim() == 2 else t.unsqueeze(1)
    if not t.is_floating_point():
        return t.to(torch.int64)
    x = t.to(torch.float64)
    x = torch.where(x == 0, torch.zeros_like(x), x)
    x = torch.where(torch.isnan(x), torch.full_like(x, float('nan')), x)
    b = x.view(torch.int64)
    return torch.where(b < 0, b ^ 0x7FFFFFFFFFFFFFFF, b)


def _unwrap(tree):
    while isinstance(tree, list) and len(tree) == 1:
        tree = tree[0]
    return tree


def _is_key(tree):
    return isinstance(tree, str) or (len(tree) == 3 and tree[1] == 'AS' and isinstance(tree[0], str))


def tqp_hash_aggregate(tensor_group, args):
    """operators/aggregate.py:236-393 with exact, key-ordered group numbering."""
    assert not args['is partial']
    transfer = args['input ids']
    amask = {}
    if 'mask map' in args:
        for k, v in args['mask map'].items():
            if k in transfer:
                amask[k] = v
                if v not in transfer:
                    transfer.append(v)
    key_new, key_names, agg_ids = [], [], args['out agg ids']
    for tree in args['results']:
        tree = _unwrap(tree)
        if _is_key(tree):
            kid = get_id(tree if isinstance(tree, str) else tree[0])
            key_new.append(get_id(tree if isinstance(tree, str) else tree[-1]))
            key_names.append(tensor_group[kid].tensor_type)
    jin = {i: tensor_group[i] for i in transfer}
    if _all_empty(jin):
        res = [torch.tensor([]) for _ in key_new + agg_ids]
    else:
        cols = apply_masks(jin, amask)
        n = next(iter(cols.values())).tensor.size(0)
        out = {}
        if not args['keys']:
            uniq, inv = torch.tensor([[1]]), torch.zeros(n, dtype=torch.int64)
            kids, ksz = [], []
        else:
            kids = [get_id(k) for k in args['keys']]
            # groups: distinct key tuples (aggregate.py:302-327 hashes the stacked key columns;
            # the exact tuples are the groups a collision-free hash gives), numbered in
            # ascending lexicographic order
            stacked = torch.cat([_key_ordinals(cols[k].tensor) for k in kids], dim=1)
            uniq, inv = torch.unique(stacked, dim=0, return_inverse=True)
            # each group's key values, read from its first row (exact for every dtype)
            first = torch.full((uniq.shape[0],), n, dtype=torch.int64).scatter_reduce(
                0, inv, torch.arange(n), reduce='amin')
        groups = uniq.shape[0]
        for tree in args['results']:
            tree = _unwrap(tree)
            oid = get_id(right_most(tree))
            if _is_key(tree):
                kid = get_id(tree if isinstance(tree, str) else tree[0])
                kt = cols[kid].tensor
                out[oid] = (kt if kt.dim() == 2 else kt.unsqueeze(1))[first]
            else:
                out[oid] = agg_evaluate(cols, tree[0], n, groups, inv)
        res = [out[i].contiguous() for i in key_new + agg_ids]
    nk = len(key_new)
    for i, aid in enumerate(agg_ids):
        tensor_group[aid] = Col(res[nk + i], _type_of(res[nk + i].dtype))
    for i, kid in enumerate(key_new):
        tensor_group[kid] = Col(res[i], key_names[i])
    return list(tensor_group.values())[0].tensor.size(0)


# ====================================================================== sort
def tqp_sort(tensor_group, args, with_limit):
    """operators/sort.py:6-58: one stable pass per key, least significant first."""
    ids = args['input ids']
    rows = tensor_group[ids[0]].tensor.shape[0]
    if tensor_group[args['key ids'][0]].tensor.numel() * tensor_group[args['key ids'][0]].tensor.element_size() >= 8_000_000_000:
        raise RuntimeError("sort input exceeds in-memory limit")
    perm = torch.arange(rows)
    for kid, desc in zip(reversed(args['key ids']), reversed(args['key orders'])):
        key = tensor_group[kid].tensor
        if key.dim() == 2:
            _, inv = torch.unique(key[perm], sorted=True, return_inverse=True, dim=0)
            _, idx = torch.sort(inv, stable=True, descending=desc)
        else:
            _, idx = torch.sort(key[perm], stable=True, descending=desc)
        perm = perm[idx]
    if with_limit:
        perm = perm[:args['limit']]
    for c in ids:
        tensor_group[c] = Col(tensor_group[c].tensor[perm], tensor_group[c].tensor_type)
    return rows


# ====================================================================== project
def tqp_project(tensor_group, args):
    """operators/project.py:10-69: renames, then every used expression that is not a bare or
    renamed column is evaluated per row (aggregate.py's evaluate), cast through float64 and back
    to its own dtype, and stored under `materialized out` in order."""
    for orig, new in args['rename map'].items():
        tensor_group[new] = tensor_group[orig]
    if len(args['used expr']) == 0:
        return -1
    cols = {k: tensor_group[k] for k in args['materialized in']}
    rows = next(iter(cols.values())).tensor.shape[0]
    results, types = [], []
    for tree in args['used expr']:
        while isinstance(tree, list) and len(tree) == 1:
            tree = tree[0]
        if isinstance(tree, str):
            continue
        if tree[1] == 'AS' and isinstance(tree[0], str):
            continue
        r = agg_evaluate(cols, tree[0] if len(tree) > 1 and tree[1] == 'AS' else tree, rows,
                         None, None)
        types.append(r.dtype)
        results.append(r.to(torch.float64))
    for i, oid in enumerate(args['materialized out']):
        t = results[i].to(types[i])
        tensor_group[oid] = Col(t, _type_name(types[i]))
    return -1


def _type_name(dt):
    """variable.py's torch_to_type."""
    if dt == torch.float64:
        return 'float'
    if dt == torch.int32:
        return 'date'
    if dt == torch.int64:
        return 'int'
    return 'string'


# ====================================================================== scan
SCAN_TYPES = {"identifier": "int", "integer": "int", "decimal": "float", "date": "date",
              "variable text": "string", "fixed text": "string"}


def tqp_scan(tensor_group, args, tensors_path, schema_table, SF):
    """operators/scan.py:8-54 on the host: the first parameter of each column's TorchScript
    archive, typed by the schema, cast to the engine's dtypes (dates int32, floating float64,
    text int8, else int64); the primary key marked sorted.  Returns the row count."""
    import os
    rows = 0
    for name, cid in zip(args['output names'], args['output ids']):
        name = name.upper()
        path = os.path.join(tensors_path, f'SF{SF}-tensor-{name}.pth')
        t = list(torch.jit.load(path, map_location='cpu').parameters())[0].detach()
        ty = SCAN_TYPES[schema_table[name].split(',')[0]]
        rows = t.shape[0]
        if ty == 'date':
            t = t.to(torch.int32)
        elif torch.is_floating_point(t):
            t = t.to(torch.float64)
        elif ty == 'string':
            t = t.to(torch.int8)
        else:
            assert t.dim() == 1
            t = t.to(torch.int64)
        col = Col(t, ty)
        col.is_sorted = name == schema_table['PRIMARY_KEY']
        tensor_group[cid] = col
    return rows


# ====================================================================== executor
JOIN_TYPE = {'Inner': 'inner', 'LeftSemi': 'right-semi', 'LeftOuter': 'right-outer',
             'LeftAnti': 'right-anti'}


def execute(plan, tensor_group, subquery_result=None, op_times=None):
    """main.py:130-274 dispatch on CPU tensors.  `plan` is a list of (u, args)."""
    for u, args in plan:
        t0 = time.perf_counter()
        name = args['name']
        if name == 'Filter':
            tqp_filter(tensor_group, args, subquery_result)
        elif name == 'Aggregate':
            tqp_hash_aggregate(tensor_group, args)
        elif name == 'Project':
            tqp_project(tensor_group, args)
        elif name == 'Sort':
            tqp_sort(tensor_group, args, False)
        elif name == 'take ordered':
            tqp_sort(tensor_group, args, True)
        elif name == 'Join':
            lk, rk = args['left keys'], args['right keys']
            conds = [args['join condition']] if 'join condition' in args else []
            jl, jr = lk[0], rk[0]
            if len(lk) == 2:
                if tensor_group[lk[0]].tensor.is_floating_point():
                    jl, jr = lk[1], rk[1]
                    conds.append([f'#{lk[0]}', '=', f'#{rk[0]}'])
                else:
                    conds.append([f'#{lk[1]}', '=', f'#{rk[1]}'])
            join_vortex(tensor_group, args, jl, jr, conds, JOIN_TYPE[args['join type']])
        else:
            raise Exception(f"Unsupported Operator: {name}")
        if op_times is not None:
            op_times.setdefault(f"{name} {u}", []).append(time.perf_counter() - t0)


def to_cols(tables):
    """{id: (type, ndarray)} -> {id: Col} on the CPU."""
    return {cid: Col(torch.from_numpy(np.ascontiguousarray(arr)), ty) for cid, (ty, arr) in tables.items()}
