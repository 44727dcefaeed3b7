"""Closed-form and brute-force statements of the hot path's results (test infrastructure only).

These pin the CPU restatement in `ref_ops` independently of how it computes: each function
states what the result must be, in plain numpy or Python loops, without the reference's
algorithm (no sort + searchsorted join, no hash + unique grouping, no per-key sort passes).

  mask_compare      filter.py:142-162: comparison in torch's promoted dtype, float literal eps
  compact           vortex_pipeline.py:58: selected rows in input order
  join_pairs_brute  hashjoin.py:24-177: every (build, probe) pair with equal keys, nested loops
  group_reduce      aggregate.py:282-364: exact grouping, sums accumulated in row order
  sort_perm         sort.py:20-42: lexicographic stable order, -0.0 == +0.0, NaN largest
"""
import math

import numpy as np

EPS = 1e-5


def mask_compare(col: np.ndarray, op: str, lit):
    """Bool mask of `col op lit` as the reference's filter computes it."""
    if isinstance(lit, float):
        if op == '<=':
            lit = lit + EPS
        elif op == '>=':
            lit = lit - EPS
        if col.dtype.kind in 'iub':
            # int tensor vs python float: torch compares in the default float dtype (float32)
            a, b = col.astype(np.float32), np.float32(lit)
        else:
            a, b = col, np.float64(lit) if col.dtype == np.float64 else np.float32(lit)
        if op == '=' and col.dtype == np.float64:
            return (a == b) | (np.isfinite(np.abs(a - b)) & (np.abs(a - b) <= EPS + np.abs(1e-5 * b)))
    else:
        a, b = col.astype(np.int64) if col.dtype.kind in 'iub' else col, lit
        if op == '=' and col.dtype == np.float64:
            b = np.float64(lit)
            return (a == b) | (np.isfinite(np.abs(a - b)) & (np.abs(a - b) <= EPS + np.abs(1e-5 * b)))
    return {'<': a < b, '<=': a <= b, '>': a > b, '>=': a >= b, '=': a == b}[op]


def compact(mask: np.ndarray, col: np.ndarray):
    return col[np.flatnonzero(mask)]


def join_pairs_brute(bkeys, bmask, pkeys, pmask):
    """[(build_row, probe_row)] for equal keys among selected rows, probe-major, build ascending.
    Row numbers are positions in the uncompacted inputs."""
    pairs = []
    for p in range(len(pkeys)):
        if pmask is not None and not pmask[p]:
            continue
        for b in range(len(bkeys)):
            if bmask is not None and not bmask[b]:
                continue
            if int(bkeys[b]) == int(pkeys[p]):
                pairs.append((b, p))
    return pairs


def _key_part(col, r):
    """A hashable, ordered stand-in for row r of a key column: a 2-D (string) row is its tuple
    of bytes; a float is its value with -0.0 as +0.0 and every NaN one value above +inf."""
    if col.ndim == 2:
        return tuple(int(b) for b in col[r])
    v = col[r]
    if col.dtype.kind == 'f':
        v = float(v)
        return (1, 0.0) if math.isnan(v) else (0, 0.0 if v == 0 else v)
    return int(v)


def group_reduce(keys, mask, values):
    """Exact grouping of selected rows by the key tuple.

    keys: list of arrays (integer, float, or 2-D int8 strings); values: {name: (op, array)}
    with op in sum/avg/count/min/max.  Returns (one array per key holding each group's key
    value from its first row, groups in ascending key order, {name: result}); float sums are
    accumulated sequentially in row order, integer sums exactly."""
    n = len(keys[0]) if keys else len(next(iter(values.values()))[1])
    sel = np.arange(n) if mask is None else np.flatnonzero(mask)
    groups = {}
    order = []
    for r in sel:
        k = tuple(_key_part(kk, r) for kk in keys)
        if k not in groups:
            groups[k] = []
            order.append(k)
        groups[k].append(r)
    if not keys:
        groups = {(): list(sel)}
        order = [()]
    ks = sorted(order)
    out = {}
    for name, (op, arr) in values.items():
        res = []
        for k in ks:
            rows = groups[k]
            vals = arr[rows]
            if op in ('sum', 'avg'):
                if arr.dtype.kind == 'f':
                    acc = 0.0
                    for v in vals:
                        acc += float(v)
                else:
                    acc = int(sum(int(v) for v in vals))
                res.append(acc / len(rows) if op == 'avg' else acc)
            elif op == 'count':
                res.append(int(sum(1 for v in vals if v != -64)))
            elif op == 'min':
                res.append(min(vals.tolist()) if len(vals) else None)
            elif op == 'max':
                res.append(max([0] + vals.tolist()))
        out[name] = res
    key_cols = [np.array([kk[groups[k][0]] for k in ks]) for kk in keys]
    return key_cols, out


def year_of_day(d: int) -> int:
    """Calendar year of day d counted from 1990-01-01 (the date epoch of conversion.py:9)."""
    import datetime
    return (datetime.date(1990, 1, 1) + datetime.timedelta(days=int(d))).year


def _ranks(x: np.ndarray):
    """Dense ranks under torch.sort's order: -0.0 ties +0.0, NaN above everything."""
    if x.dtype.kind == 'f':
        xx = np.where(x == 0, 0.0, x).astype(np.float64)
        nan = np.isnan(xx)
        u = np.unique(xx[~nan])
        r = np.searchsorted(u, xx)
        r[nan] = len(u)
        return r
    u = np.unique(x)
    return np.searchsorted(u, x)


def sort_perm(keys, descending):
    """Stable lexicographic order of rows; keys most significant first; 2-D keys row-wise."""
    cols = []
    for k, d in zip(keys, descending):
        if k.ndim == 2:
            _, inv = np.unique(k, axis=0, return_inverse=True)
            r = inv.reshape(-1)
        else:
            r = _ranks(k)
        cols.append(-r if d else r)
    # np.lexsort sorts by the last key first and is stable
    return np.lexsort(tuple(reversed(cols))) if cols else np.arange(0)


def seq_sum(vals):
    """Plain left-to-right float64 sum."""
    acc = 0.0
    for v in vals:
        acc += float(v)
    return acc


def exact_sum(vals):
    return math.fsum(float(v) for v in vals)
