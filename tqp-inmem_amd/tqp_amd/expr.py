"""Compiler from the plan's expression trees to the register programs the HIP kernels run.

Expression trees are the nested lists of strings the planner emits (`'name#id'` leaves refer to
columns).  Two dialects exist because the reference evaluates them with two different
functions:

* predicates, as `operators/filter.py:14-162` (`evaluate`): float literals widen `<=`/`>=` by
  `eps`, float `=` is `torch.isclose(atol=eps)`, string comparisons are equality of the padded
  byte rows whatever the operator, `isnotnull` is all-true, `IN` is an OR of equalities;
* row expressions inside aggregates and projections, as `operators/aggregate.py:58-233`:
  `+ - * /`, comparisons, AND/OR/NOT, CASE WHEN (torch.where), `year`, `cast`.

Every node gets the dtype torch would give it (`torch.result_type` with the same tensor/scalar
operands, true division of integers in the default float dtype, integer results wrapped to
their width), so the kernels compute in exactly the dtype the reference's ATen ops would.
"""
import ctypes
import struct

import torch

from . import _lib
from . import constants
from .conversion import date_to_float, get_id, is_date, is_float, str_to_codes

L = _lib


class Value:
    """A compiled sub-expression: a register of a dtype, a Python literal, a string literal,
    or a 2-D string column (which only string predicates can consume).  A string column made by
    `substring` is narrow: its column-set entry starts at the substring's first byte, keeps the
    full row width as its stride, and `width` is the substring's length."""

    __slots__ = ("kind", "reg", "dtype", "value", "col", "width", "narrow")

    def __init__(self, kind, reg=None, dtype=None, value=None, col=None, width=None,
                 narrow=False):
        self.kind, self.reg, self.dtype, self.value = kind, reg, dtype, value
        self.col, self.width, self.narrow = col, width, narrow

    @property
    def is_lit(self):
        return self.kind == "lit"


def _cls(dt):
    if dt == torch.float64:
        return "F64"
    if dt == torch.float32:
        return "F32"
    if dt.is_floating_point:
        raise NotImplementedError(f"unsupported float dtype {dt}")
    return "I64"


def _proto(v: Value):
    return torch.empty(1, dtype=v.dtype) if v.kind == "reg" else v.value


_RESULT_TYPES = {}


def _result_type(a: Value, b: Value):
    """torch.result_type of two operands, memoised: it depends only on the tensors' dtypes and
    the Python types of scalar literals."""
    key = (a.dtype if a.kind == "reg" else type(a.value), a.kind == "reg",
           b.dtype if b.kind == "reg" else type(b.value), b.kind == "reg")
    dt = _RESULT_TYPES.get(key)
    if dt is None:
        dt = torch.result_type(_proto(a), _proto(b))
        _RESULT_TYPES[key] = dt
    return dt


class ProgramBuilder:
    """Accumulates instructions, load slots and the column set of one program."""

    def __init__(self):
        self.prog = L.Program()
        self.cols = L.ColSet()
        self.col_tensors = []       # keep tensors alive while the program is in use
        self.col_offsets = []       # byte offset of each colset entry into its tensor
        self.col_ids = {}           # (address, dtype, shape, offset) -> colset index
        self.load_of = {}           # colset index -> load slot
        self.nload = 0
        self.ninstr = 0
        self.nseg = 0
        self.strpool = bytearray()
        self.free = list(range(L.TQ_NREG))
        self.n = None
        self.n_col = None

    # ---------------------------------------------------------------- columns and registers
    def column(self, t: torch.Tensor, offset: int = 0) -> int:
        key = (t.data_ptr(), t.dtype, tuple(t.shape), offset)
        if key in self.col_ids:
            return self.col_ids[key]
        idx = len(self.col_tensors)
        if idx >= L.TQ_MAX_COLS:
            raise NotImplementedError(f"expression uses more than {L.TQ_MAX_COLS} columns")
        L.require_gpu(t)
        if not t.is_contiguous():
            raise ValueError("expression columns must be contiguous")
        self.col_tensors.append(t)
        self.col_offsets.append(offset)
        self.col_ids[key] = idx
        self.cols.ptr[idx] = t.data_ptr() + offset
        self.cols.width[idx] = t.shape[1] if t.dim() == 2 else 1
        return idx

    def load_slot(self, t: torch.Tensor) -> int:
        c = self.column(t)
        if c in self.load_of:
            return self.load_of[c]
        if self.nload >= L.TQ_MAX_LOADS:
            raise NotImplementedError(f"expression reads more than {L.TQ_MAX_LOADS} columns")
        s = self.nload
        self.prog.load_col[s] = c
        self.prog.load_dtype[s] = L.dtype_code(t.dtype)
        self.load_of[c] = s
        self.nload += 1
        self.prog.nload = self.nload
        return s

    def alloc(self) -> int:
        if not self.free:
            raise NotImplementedError("expression too deep for the 8-register evaluator")
        return self.free.pop(0)

    def release(self, v: Value):
        if v is not None and v.kind == "reg" and v.reg not in self.free:
            self.free.append(v.reg)
            self.free.sort()

    def emit(self, op, dst=0, a=0, b=0, c=0, imm=0):
        if self.ninstr >= L.TQ_MAX_INSTR:
            raise NotImplementedError(f"expression longer than {L.TQ_MAX_INSTR} instructions")
        ins = self.prog.code[self.ninstr]
        ins.op, ins.dst, ins.a, ins.b, ins.c, ins.imm = L.OP[op], dst, a, b, c, imm
        self.ninstr += 1
        self.prog.n_instr = self.ninstr

    def end_segment(self, result: Value):
        if self.nseg >= L.TQ_MAX_SEGS:
            raise NotImplementedError(f"more than {L.TQ_MAX_SEGS} program outputs")
        self.prog.seg_end[self.nseg] = self.ninstr
        self.prog.seg_result[self.nseg] = result.reg
        self.nseg += 1
        self.prog.nseg = self.nseg
        self.free = list(range(L.TQ_NREG))

    def pool_string(self, s: str) -> int:
        codes = bytes(c & 0xFF for c in (ord(ch) for ch in s))
        off = self.strpool.find(codes) if codes else 0
        if off < 0:
            off = len(self.strpool)
            if off + len(codes) > L.TQ_STRPOOL:
                raise NotImplementedError("string literals exceed the program's string pool")
            self.strpool.extend(codes)
            for i, b in enumerate(codes):
                self.prog.strpool[off + i] = b
        return off | (len(codes) << 32)

    # ---------------------------------------------------------------- value plumbing
    def load_column(self, t: torch.Tensor) -> Value:
        if t.dim() == 2 and t.shape[1] == 1:
            t = t.reshape(-1)
        if t.dim() == 2:
            return Value("col2d", col=self.column(t), width=t.shape[1], dtype=t.dtype)
        if self.n is None:
            self.n = t.shape[0]
            self.n_col = self.column(t)
        slot = self.load_slot(t)
        r = self.alloc()
        self.emit("LOAD", dst=r, a=slot)
        return Value("reg", reg=r, dtype=t.dtype)

    def const(self, value, dt) -> Value:
        r = self.alloc()
        if dt == torch.float64:
            imm = struct.unpack("<q", struct.pack("<d", float(value)))[0]
        elif dt == torch.float32:
            imm = struct.unpack("<I", struct.pack("<f", float(value)))[0]
        elif dt == torch.bool:
            imm = 1 if value else 0
        else:
            imm = int(value)
            if not -(1 << 63) <= imm < (1 << 63):
                raise ValueError(f"integer literal {value} outside int64")
        self.emit("CONST", dst=r, imm=imm)
        return Value("reg", reg=r, dtype=dt)

    def to_dtype(self, v: Value, dt) -> Value:
        """Materialise v in a register of dtype dt (torch's cast semantics)."""
        if v.kind == "lit":
            return self.const(v.value, dt)
        if v.kind != "reg":
            raise NotImplementedError("string operand in arithmetic")
        src = v.dtype
        if src == dt:
            return v
        sc, dc = _cls(src), _cls(dt)
        if sc == "I64" and dc == "I64":
            if dt == torch.bool:
                raise NotImplementedError("cast to bool")
            width = {torch.int8: "WRAP_I8", torch.int16: "WRAP_I16", torch.int32: "WRAP_I32"}.get(dt)
            if width and src not in (torch.bool, torch.int8) and torch.iinfo(src).bits > torch.iinfo(dt).bits:
                self.emit(width, dst=v.reg, a=v.reg)
            return Value("reg", reg=v.reg, dtype=dt)
        op = {("I64", "F64"): "CVT_I64_F64", ("I64", "F32"): "CVT_I64_F32",
              ("F32", "F64"): "CVT_F32_F64", ("F64", "F32"): "CVT_F64_F32"}.get((sc, dc))
        if op is None:
            raise NotImplementedError(f"cast {src} -> {dt}")
        self.emit(op, dst=v.reg, a=v.reg)
        return Value("reg", reg=v.reg, dtype=dt)

    def binary(self, opname, a: Value, b: Value, dt, out_dt):
        ra = self.to_dtype(a, dt)
        rb = self.to_dtype(b, dt)
        self.emit(f"{opname}_{_cls(dt)}", dst=ra.reg, a=ra.reg, b=rb.reg)
        if rb.reg != ra.reg:
            self.release(rb)
        return Value("reg", reg=ra.reg, dtype=out_dt)

    def compare(self, op: str, a: Value, b: Value) -> Value:
        name = {"<": "LT", "<=": "LE", ">": "GT", ">=": "GE", "=": "EQ", "<>": "NE", "!=": "NE"}[op]
        if a.is_lit and b.is_lit:
            raise NotImplementedError("comparison of two literals")
        dt = _result_type(a, b)
        if dt == torch.bool:
            dt = torch.int64
        return self.binary(name, a, b, dt, torch.bool)

    def arith(self, op: str, a: Value, b: Value) -> Value:
        if a.is_lit and b.is_lit:
            raise NotImplementedError("arithmetic on two literals")
        dt = _result_type(a, b)
        if op == "/" and not dt.is_floating_point:
            dt = torch.get_default_dtype()
        if dt == torch.bool:
            raise NotImplementedError("arithmetic on booleans")
        name = {"+": "ADD", "-": "SUB", "*": "MUL", "/": "DIV"}[op]
        if name == "DIV" and _cls(dt) == "I64":
            raise NotImplementedError("integer division")
        v = self.binary(name, a, b, dt, dt)
        if _cls(dt) == "I64" and dt != torch.int64:
            self.emit({torch.int8: "WRAP_I8", torch.int16: "WRAP_I16", torch.int32: "WRAP_I32",
                       torch.uint8: "WRAP_I8"}.get(dt, "WRAP_I32"), dst=v.reg, a=v.reg)
        return v

    def logic(self, op: str, a: Value, b: Value) -> Value:
        ra = self.truth(a)
        rb = self.truth(b)
        self.emit(op, dst=ra.reg, a=ra.reg, b=rb.reg)
        self.release(rb)
        return Value("reg", reg=ra.reg, dtype=torch.bool)

    def copy(self, v: Value) -> Value:
        """A second register holding v's bits (WHERE with the value as its own condition)."""
        r = self.alloc()
        self.emit("WHERE", dst=r, a=v.reg, b=v.reg, c=v.reg)
        return Value("reg", reg=r, dtype=v.dtype)

    def truth(self, v: Value) -> Value:
        if v.kind == "lit":
            return self.const(bool(v.value), torch.bool)
        if v.kind != "reg":
            raise NotImplementedError("string operand in a boolean expression")
        return v

    def substring(self, v: Value, start: int, length: int) -> Value:
        """Columns [start - 1, start - 1 + length) of a 2-D string column, as the reference's
        `[:, l:l + w]` slice (filter.py:71-76, aggregate.py:115-120), clamped to the row."""
        if v.kind != "col2d":
            raise NotImplementedError("substring of a non-string value")
        lo = start - 1
        if lo < 0 or length < 0:
            raise NotImplementedError("substring with a start before the first character")
        t = self.col_tensors[v.col]
        full = int(self.cols.width[v.col])
        base = self.col_offsets[v.col]
        lo = min(base + lo, full)
        hi = min(lo + length, base + v.width, full)
        if hi - lo > 255:
            raise NotImplementedError("substring longer than 255 bytes")
        return Value("col2d", col=self.column(t, lo), width=hi - lo, dtype=v.dtype, narrow=True)

    @staticmethod
    def _narrow(col: Value) -> int:
        """The string instructions' c field: the compared width of a narrow column, else 0."""
        return col.width if col.narrow else 0

    def str_eq(self, col: Value, s: str, prefix=False) -> Value:
        if col.kind != "col2d":
            raise NotImplementedError("string comparison needs a 2-D string column")
        if not prefix and len(s) > col.width:
            raise ValueError(f"string literal '{s}' longer than column width {col.width}")
        r = self.alloc()
        self.emit("STR_PREFIX" if prefix else "STR_EQ", dst=r, a=col.col, c=self._narrow(col),
                  imm=self.pool_string(s))
        return Value("reg", reg=r, dtype=torch.bool)

    def str_contains(self, col: Value, s: str) -> Value:
        """Rows of a 2-D string column holding s at some position (like.py:3-11)."""
        if col.kind != "col2d":
            raise NotImplementedError("Contains needs a 2-D string column")
        r = self.alloc()
        self.emit("STR_CONTAINS", dst=r, a=col.col, c=self._narrow(col), imm=self.pool_string(s))
        return Value("reg", reg=r, dtype=torch.bool)

    def str_like2(self, col: Value, w1: str, w2: str) -> Value:
        """Rows of a 2-D string column that hold w1 and, after it, w2 (like.py:13-25)."""
        if col.kind != "col2d":
            raise NotImplementedError("LIKE needs a 2-D string column")
        p1, p2 = self.pool_string(w1), self.pool_string(w2)
        imm = ((p1 & 0xFFFF) | ((p1 >> 32) << 16) | ((p2 & 0xFFFF) << 32) | ((p2 >> 32) << 48))
        r = self.alloc()
        self.emit("STR_LIKE2", dst=r, a=col.col, c=self._narrow(col), imm=imm)
        return Value("reg", reg=r, dtype=torch.bool)

    def load_bytes(self) -> int:
        """Bytes per row the load table reads."""
        return sum(self.col_tensors[self.prog.load_col[s]].element_size() for s in range(self.nload))

    def finish(self):
        self.prog.n_instr = self.ninstr
        self.prog.nseg = self.nseg
        self.prog.nload = self.nload
        return self


# ==================================================================== predicates (filter.py)
def _filter_leaf(tree: str, tensors, subquery_result, leaf_function, b: ProgramBuilder):
    """filter.py:18-35."""
    if tree == 'Subquery':
        return Value("lit", value=subquery_result)
    if is_date(tree):
        return Value("lit", value=date_to_float(tree))
    if is_float(tree):
        return Value("lit", value=float(tree) if '.' in tree else int(tree))
    if '#' in tree and not tree.startswith('Brand#'):
        t = leaf_function(get_id(tree)) if leaf_function else tensors[get_id(tree)].tensor
        return b.load_column(t)
    return Value("str", value=tree)


def compile_predicate(b: ProgramBuilder, tree, tensors, subquery_result=None, leaf_function=None):
    """Compile a predicate tree; follows operators/filter.py:14-162 node by node."""
    def ev(tree):
        if isinstance(tree, str):
            return _filter_leaf(tree, tensors, subquery_result, leaf_function, b)
        if len(tree) == 1:
            return ev(tree[0])
        if len(tree) == 3 and tree[1] == '#':
            cid = int(tree[-1])
            t = leaf_function(cid) if leaf_function else tensors[cid].tensor
            return b.load_column(t)
        if len(tree) == 2:
            head = tree[0]
            if head == 'StartsWith':
                col = b.load_column(tensors[get_id(tree[1][0])].tensor)
                return b.str_eq(col, tree[1][1], prefix=True)
            if head in ('Contains', 'EndsWith'):
                # filter.py:56-62: both run like_contains over the whole row
                col = b.load_column(tensors[get_id(tree[1][0])].tensor)
                return b.str_contains(col, tree[1][1])
            if head == 'substring':
                assert len(tree[1]) == 3
                sub, l, w = tree[1]
                return b.substring(ev(sub), int(l), int(w))
            if head == 'NOT':
                v = b.truth(ev(tree[1]))
                b.emit("NOT", dst=v.reg, a=v.reg)
                return Value("reg", reg=v.reg, dtype=torch.bool)
            if head == 'isnotnull':
                return b.const(True, torch.bool)
            if head == 'cast':
                literal, _, cast_type = tree[1][0].split(' ')
                if cast_type == 'double':
                    return ev(literal)
            raise NotImplementedError(f"Unsupported Function: {head}")
        op = tree[-2]
        if op == 'IN':
            lhs = ev(tree[0])
            names = tree[-1]
            acc = None
            for name in names:
                if lhs.kind == "col2d":
                    if len(name) == 1:
                        raise NotImplementedError("IN over single-character strings")
                    term = b.str_eq(lhs, name)
                else:
                    if lhs.dtype in (constants.int_dtype, constants.date_dtype) or not lhs.dtype.is_floating_point:
                        lit = int(name)
                    else:
                        lit = float(name)
                    term = b.compare('=', b.copy(lhs), Value("lit", value=lit))
                acc = term if acc is None else b.logic("OR", acc, term)
            b.release(lhs)
            return acc if acc is not None else b.const(False, torch.bool)
        if op == 'LIKE':
            # '%w1%w2%' only, as the reference asserts (filter.py:121-127)
            lhs = ev(tree[0])
            assert tree[-1].count('%') == 3
            _, w1, w2, _ = tree[-1].split('%')
            return b.str_like2(lhs, w1, w2)
        if op in ('AND', 'OR'):
            return b.logic(op, ev(tree[0]), ev(tree[-1]))
        if op not in ('<=', '>=', '<', '>', '='):
            raise ValueError(f"Unsupported Operation: {op}")
        lhs = ev(tree[0])
        rhs = ev(tree[-1])
        if rhs.kind == "lit" and isinstance(rhs.value, float):
            if op == '<=':
                return b.compare('<=', lhs, Value("lit", value=rhs.value + constants.eps))
            if op == '>=':
                return b.compare('>=', lhs, Value("lit", value=rhs.value - constants.eps))
        if rhs.kind == "str":
            if len(rhs.value) == 1:
                return b.compare(op, lhs, Value("lit", value=ord(rhs.value)))
            return b.str_eq(lhs, rhs.value)
        if op == '=' and lhs.kind == "reg" and lhs.dtype == constants.float_dtype:
            ra = lhs
            rb = b.to_dtype(rhs, constants.float_dtype)
            atol = struct.unpack("<q", struct.pack("<d", constants.eps))[0]
            b.emit("ISCLOSE_F64", dst=ra.reg, a=ra.reg, b=rb.reg, imm=atol)
            b.release(rb)
            return Value("reg", reg=ra.reg, dtype=torch.bool)
        return b.compare(op, lhs, rhs)

    result = b.truth(ev(tree))
    b.end_segment(result)
    return b.finish()


# ==================================================================== row expressions (aggregate.py)
AGG_FUNCS = ('sum', 'avg', 'count', 'min', 'max')


def _agg_leaf(tree: str, tensors, b: ProgramBuilder):
    """aggregate.py:59-67."""
    if tree.isdigit():
        return Value("lit", value=int(tree))
    if is_float(tree):
        return Value("lit", value=float(tree))
    if '#' in tree and tree.split('#')[-1].replace('L', '').isdigit():
        return b.load_column(tensors[get_id(tree)].tensor)
    return Value("str", value=tree)


def compile_row_expr(b: ProgramBuilder, tree, tensors, agg_ref=None) -> Value:
    """Compile a row expression as operators/aggregate.py:58-233 evaluates it.  `agg_ref`
    maps an aggregate call node (by id) to a column holding its per-group results, which is
    how expressions over aggregates (e.g. 0.2 * avg(x)) are compiled."""
    def ev(tree):
        if agg_ref is not None and id(tree) in agg_ref:
            return b.load_column(agg_ref[id(tree)])
        if isinstance(tree, str):
            return _agg_leaf(tree, tensors, b)
        if len(tree) == 1:
            return ev(tree[0])
        if len(tree) == 3 and tree[1] == '#':
            return ev(tree[0])
        if len(tree) == 8:
            assert tree[0] == 'CASE' and tree[1] == 'WHEN' and tree[3] == 'THEN' \
                and tree[5] == 'ELSE' and tree[7] == 'END'
            cond = b.truth(ev(tree[2]))
            then = ev(tree[4])
            other = ev(tree[6])
            dt = _result_type(then, other)
            rt = b.to_dtype(then, dt)
            ro = b.to_dtype(other, dt)
            b.emit("WHERE", dst=rt.reg, a=rt.reg, b=ro.reg, c=cond.reg)
            b.release(ro)
            b.release(cond)
            return Value("reg", reg=rt.reg, dtype=dt)
        if len(tree) == 2:
            head = tree[0]
            if head == 'StartsWith':
                col = b.load_column(tensors[get_id(tree[1][0])].tensor)
                return b.str_eq(col, tree[1][1], prefix=True)
            if head == 'substring':
                assert len(tree[1]) == 3
                sub, l, w = tree[1]
                return b.substring(ev(sub), int(l), int(w))
            if head == 'year':
                v = b.to_dtype(ev(tree[1]), torch.int64)
                b.emit("YEAR", dst=v.reg, a=v.reg)
                return Value("reg", reg=v.reg, dtype=torch.int64)
            if head == 'NOT':
                v = b.truth(ev(tree[1]))
                b.emit("NOT", dst=v.reg, a=v.reg)
                return Value("reg", reg=v.reg, dtype=torch.bool)
            if head == 'cast':
                literal, _, cast_type = tree[1][0].split(' ')
                if cast_type == 'double':
                    return ev(literal)
                raise ValueError(f"Unsupported cast type: {cast_type}")
            if head in AGG_FUNCS:
                raise NotImplementedError("nested aggregate inside a row expression")
            raise ValueError(f"Unsupported aggregation function: {head}")
        assert len(tree) == 3
        op = tree[-2]
        if op == 'AS':
            return ev(tree[0])
        if op not in ('+', '-', '*', '/', '<=', '>=', '<', '>', '=', '<>', 'AND', 'OR'):
            raise ValueError(f"Unsupported Operation: {op}")
        lhs = ev(tree[0])
        rhs = ev(tree[-1])
        if rhs.kind == "str":
            if lhs.kind == "lit":
                raise NotImplementedError("string-valued expression")
            if len(rhs.value) == 1:
                rhs = Value("lit", value=ord(rhs.value))
            else:
                return b.str_eq(lhs, rhs.value)
        if op in ('+', '-', '*', '/'):
            return b.arith(op, lhs, rhs)
        if op in ('AND', 'OR'):
            return b.logic(op, lhs, rhs)
        return b.compare(op, lhs, rhs)

    return ev(tree)


def find_aggregates(tree, out):
    """Collect aggregate call nodes ['sum'|'avg'|..., arg] in evaluation order."""
    if isinstance(tree, list):
        if len(tree) == 2 and isinstance(tree[0], str) and tree[0] in AGG_FUNCS:
            out.append(tree)
            return out
        for e in tree:
            find_aggregates(e, out)
    return out


def literal_dtype(value):
    return torch.int64 if isinstance(value, int) and not isinstance(value, bool) else torch.float64


# ---------------------------------------------------------------- compiled-program cache
class BoundProgram:
    """A cached program bound to the columns of this call: the program is shared (read-only once
    finished), the column set is this call's."""

    def __init__(self, prog, cols, col_tensors, n, load_bytes, extra):
        self.prog = prog
        self.cols = cols
        self.col_tensors = col_tensors
        self.n = n
        self.nload = prog.nload
        self.extra = extra
        self._load_bytes = load_bytes

    def load_bytes(self) -> int:
        return self._load_bytes


_PROGRAMS = {}
_PROGRAMS_MAX = 512


def cached_program(key, tensors, compile_fn):
    """Compile with compile_fn(builder) -> extra once per key and bind later calls' columns to
    the same program.  The key must fix everything the program text depends on except the
    referenced columns themselves; those are found again by column id in `tensors` ({id: object
    with .tensor}) and must keep their dtype and row shape, else the program is compiled anew.
    Programs reading a column that is not one of `tensors` (a derived or reshaped tensor) are
    not cached."""
    ent = _PROGRAMS.get(key)
    if ent is not None:
        prog, refs, n_col, lb, extra = ent
        cols = L.ColSet()
        col_tensors = []
        for i, (cid, dt, tail, width, off) in enumerate(refs):
            v = tensors.get(cid)
            t = v.tensor if v is not None else None
            if t is None or t.dtype != dt or tuple(t.shape[1:]) != tail or not t.is_contiguous():
                break
            cols.ptr[i] = t.data_ptr() + off
            cols.width[i] = width
            col_tensors.append(t)
        else:
            n = col_tensors[n_col].shape[0] if n_col is not None else None
            return BoundProgram(prog, cols, col_tensors, n, lb, extra)
    b = ProgramBuilder()
    extra = compile_fn(b)
    b.finish()
    b.extra = extra
    ids = {id(v.tensor): cid for cid, v in tensors.items()}
    refs = []
    for i, t in enumerate(b.col_tensors):
        cid = ids.get(id(t))
        if cid is None:
            return b
        refs.append((cid, t.dtype, tuple(t.shape[1:]), int(b.cols.width[i]), b.col_offsets[i]))
    if len(_PROGRAMS) >= _PROGRAMS_MAX:
        _PROGRAMS.clear()
    _PROGRAMS[key] = (b.prog, refs, b.n_col, b.load_bytes(), extra)
    return b


def colset_ptr(b: ProgramBuilder):
    return ctypes.byref(b.cols)


def program_ptr(b: ProgramBuilder):
    return ctypes.byref(b.prog)
