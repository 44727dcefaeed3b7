"""Project operator (operators/project.py:10-69 of the reference).

`tqp_project` keeps the reference's signature, argument schema and side effects: the `rename map`
entries alias existing columns, then every `used expr` that is neither a bare column nor a
column renamed with AS is evaluated per row and stored, in order, under `materialized out`, with
the type name of its result dtype; it returns -1.  All the expressions run as one HIP launch
(`tq_eval_columns`): the input columns are staged once and every expression is evaluated by the
same expression program, where the reference runs one ATen kernel per tree node and expression
and casts every result through float64 and back.  The results are the same values (an int64
result above 2^53 is exact here; the reference's float64 round trip would round it).
"""
import torch

from .. import expr as E
from .. import kernels
from ..utility.logger import perf_logger
from ..variable import Variable, VariableState, torch_to_type


def _projected(used_expr):
    """The trees the reference evaluates (project.py:26-40), in order."""
    out = []
    for tree in used_expr:
        while isinstance(tree, list) and len(tree) == 1:
            tree = tree[0]
        if isinstance(tree, str):
            continue
        if tree[1] == 'AS' and isinstance(tree[0], str):
            continue
        out.append(tree[0] if len(tree) > 1 and tree[1] == 'AS' else tree)
    return out


def _compile(trees, cols):
    """One program with a segment per tree (a new program every TQ_MAX_SEGS trees, or when a
    program runs out of registers); returns [(builder, [result dtypes])]."""
    from .. import _lib as L
    progs = []
    b, dts = E.ProgramBuilder(), []
    for tree in trees:
        for attempt in range(2):
            try:
                v = E.compile_row_expr(b, tree, cols)
                if v.kind == "lit":
                    v = b.const(v.value, E.literal_dtype(v.value))
                if v.kind != "reg":
                    raise NotImplementedError("string-valued projection")
                b.end_segment(v)
                dts.append(v.dtype)
                break
            except (NotImplementedError, ValueError) as e:
                # a full program (columns, load slots, instructions): finish it and compile
                # the tree into a fresh one; an error of the tree itself raises again there
                if attempt == 1 or not dts:
                    raise e
                progs.append((b, dts))
                b, dts = E.ProgramBuilder(), []
        if len(dts) == L.TQ_MAX_SEGS:
            progs.append((b, dts))
            b, dts = E.ProgramBuilder(), []
    if dts:
        progs.append((b, dts))
    for b, _ in progs:
        b.finish()
    return progs


def tqp_project(gpu_enable, tensor_group, args, cpu_mem_pool, gpu_mem_pool, name="project"):
    perf_logger().start(name)
    if not gpu_enable:
        raise NotImplementedError("tqp_amd executes operators on the GPU only (gpu_enable=False has no CPU path)")
    for orig, new in args['rename map'].items():
        tensor_group[new] = tensor_group[orig]
    trees = _projected(args['used expr'])
    if trees:
        assert 'mask map' not in args
        cols = {k: tensor_group[k] for k in args['materialized in']}
        if all(v.tensor.numel() == 0 for v in cols.values()):
            # the reference's pipeline returns no results then, and reading the result types
            # fails (project.py:59-60)
            raise IndexError("list index out of range")
        first = next(iter(cols.values())).tensor
        rows, dev = first.shape[0], first.device
        outs, dtypes = [], []
        for b, dts in _compile(trees, cols):
            n = b.n if b.n is not None else rows
            outs += kernels.eval_programs(b, n, dts, dev)
            dtypes += dts
        for oid, t, dt in zip(args['materialized out'], outs, dtypes):
            tensor_group[oid] = Variable(t, torch_to_type(dt), None, VariableState.GPU)
    perf_logger().stop(name)
    return -1
