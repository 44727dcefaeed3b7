"""Literal decoding shared by the predicate and aggregate compilers.

Follows conversion.py:13-31 of the reference: dates are whole days since 1990-01-01, a literal is
"float-like" when stripping '.', '-' and 'E' leaves only digits, and `year` buckets a day count by
the cumulative year lengths of 1990..1999.
"""
import datetime
from itertools import accumulate

YEAR_START = 1990
YEAR_END = 2000

_epoch = datetime.date(YEAR_START, 1, 1)
year_range = list(range(YEAR_START, YEAR_END))
year_edges = list(accumulate([365 if year % 4 else 366 for year in year_range]))


def is_date(s: str) -> bool:
    return len(s) == 10 and s.count('-') == 2 and s.replace('-', '').isdigit()


def is_float(s: str) -> bool:
    return s.replace('.', '').replace('-', '').replace('E', '').isdigit()


def date_to_float(date_str: str) -> int:
    """Days between 1990-01-01 and an ISO date (conversion.py:25-28)."""
    y, m, d = (int(p) for p in date_str.split('-'))
    return (datetime.date(y, m, d) - _epoch).days


def float_to_date(diff) -> str:
    return (_epoch + datetime.timedelta(days=int(diff))).strftime('%Y-%m-%d')


def str_to_codes(s: str, n: int):
    """Byte codes of `s` right-padded with zeros to width n (conversion.py:36-37)."""
    codes = [ord(c) for c in s]
    return codes + [0] * (n - len(codes))


def get_id(name: str) -> int:
    """'name#id' -> id, with Spark's trailing 'L' of long ids dropped (parsing.py:15-16)."""
    return int(name.split('#')[-1].replace('L', ''))


def right_most(tree):
    """Right-most leaf of an expression tree (parsing.py:42-45)."""
    if isinstance(tree, list):
        return right_most(tree[-1])
    return tree
