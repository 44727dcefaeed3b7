"""Writes the small golden fixtures of tests/golden/ (run from the repo root).

Each fixture holds inputs and the expected outputs computed by the closed forms of
oracle/closed_form.py (brute-force join pairs, exact grouping with sequential sums,
lexicographic stable order), not by the oracle restatement they are used to check.  The cases
cover the edge cases the reference's operators have: duplicate and missing join keys, masks
selecting nothing, -0.0 / NaN sort keys, float literal eps and isclose.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import closed_form as CF  # noqa: E402


def dump(name, obj):
    with open(os.path.join(HERE, name), "w") as f:
        json.dump(obj, f, indent=1)


def main():
    g = np.random.default_rng(20261203)

    # join: duplicates on both sides, masks, a probe key above every build key
    bk = np.array([5, 3, 9, 3, 7, 5, 5, 1, 8, 3], dtype=np.int32)
    bm = np.array([1, 1, 1, 0, 1, 1, 1, 1, 0, 1], dtype=bool)
    pk = np.array([3, 5, 2, 9, 5, 100, 3, 8, 1, 7, 5, 0], dtype=np.int32)
    pm = np.array([1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 1, 1], dtype=bool)
    dump("join_small.json", {"bkeys": bk.tolist(), "bmask": bm.tolist(), "pkeys": pk.tolist(),
                             "pmask": pm.tolist(),
                             "pairs": CF.join_pairs_brute(bk, bm, pk, pm)})

    # group-by over two int keys with a mask; sums accumulated in row order
    n = 64
    k1 = g.integers(-2, 2, n).astype(np.int8)
    k2 = g.integers(0, 3, n).astype(np.int32)
    v = np.round(g.uniform(900, 105000, n), 2)
    d = g.integers(0, 11, n) / 100.0
    m = g.random(n) < 0.75
    keys, vals = CF.group_reduce([k1, k2], m, {"sum_v": ("sum", v), "avg_d": ("avg", d),
                                               "cnt": ("count", np.ones(n))})
    dump("groupby_small.json", {"k1": k1.tolist(), "k2": k2.tolist(), "v": v.tolist(),
                                "d": d.tolist(), "mask": m.tolist(),
                                "keys": [k.tolist() for k in keys], "sum_v": vals["sum_v"],
                                "avg_d": vals["avg_d"], "cnt": vals["cnt"]})

    # sort: float key with -0.0, NaN, inf ties, descending; int tie-breaker ascending
    f = np.array([1.5, -0.0, np.nan, 0.0, 1.5, -np.inf, np.inf, np.nan, 2.0, 0.0, -1.0], dtype=np.float64)
    i = np.array([3, 1, 2, 2, 1, 0, 5, 1, 4, 1, 9], dtype=np.int64)
    dump("sort_small.json", {"f": [None if np.isnan(x) else x for x in f.tolist()], "i": i.tolist(),
                             "perm_desc_asc": CF.sort_perm([f, i], [True, False]).tolist(),
                             "perm_asc_desc": CF.sort_perm([f, i], [False, True]).tolist()})

    # filter: float literal eps / isclose edges and an int column against a float literal
    x = np.array([0.05, 0.050009, 0.050011, 0.049991, 0.04999, 0.06], dtype=np.float64)
    y = np.array([1199, 1200, 1201, 2400, 2401, 0], dtype=np.int32)
    dump("filter_small.json", {"x": x.tolist(), "y": y.tolist(), "cases": [
        {"tree": ["x#1", op, "0.05"], "mask": CF.mask_compare(x, op, 0.05).tolist()}
        for op in ("<=", ">=", "=", "<", ">")] + [
        {"tree": ["y#2", op, lit], "mask": CF.mask_compare(y, op, float(lit) if "." in lit else int(lit)).tolist()}
        for op, lit in (("<=", "2400"), (">", "1200"), ("<", "1200.5"), ("=", "2400.0"))]})


if __name__ == "__main__":
    main()
