"""Filter operator: predicate tree -> bool mask (operators/filter.py:164-205 of the reference).

`tqp_filter` keeps the reference's signature, argument schema and side effects: it evaluates
`args['filter tree']` over the `materialized in` columns, stores the mask as
`tensor_group[args['out mask id']] = Variable(mask, 'bool', ...)` and returns -1; consumers
compact by the mask later (late materialisation).  The whole tree runs as one HIP kernel (the
reference launches one ATen kernel per tree node) that also writes per-tile match counts, which
ride along on the mask Variable for the consumer's compaction.

With `lazy=True` (the executor's unsynchronised mode) the mask is a `LazyMask`: its kernel runs
when a consumer first reads `.tensor`, and a consumer that can evaluate the predicate inside its
own kernels (the fused groupjoin, for a conjunction of comparisons) never reads it, so the mask
is never written.  The mask's value is the same either way.
"""
import torch

from .. import _lib as L
from .. import expr as E
from .. import kernels
from ..IO.vortex_pipeline import InMemoryPipeline
from ..utility.logger import message_logger, perf_logger
from ..variable import Variable, VariableState


def evaluate(device_name, tensors, tree, rows, subquery_result, leaf_function=None):
    """Bool mask of `tree` over `tensors` (operators/filter.py:14-162), computed on the GPU.

    With `leaf_function`, column leaves are `leaf_function(id)` instead of `tensors[id]` (the
    join's condition evaluation over gathered pair columns, hashjoin.py:15-22); the tree then
    runs as one expression-kernel launch over those columns."""
    if leaf_function is None:
        mask, _ = kernels.filter_mask(tree, tensors, subquery_result, n=rows)
        return mask
    b = E.ProgramBuilder()
    E.compile_predicate(b, tree, tensors, subquery_result, leaf_function=leaf_function)
    n = b.n if b.n is not None else rows
    if n is None:
        n = next(iter(tensors.values())).tensor.shape[0]
    dev = b.col_tensors[0].device if b.col_tensors else next(iter(tensors.values())).tensor.device
    return kernels.eval_program(b, n, torch.bool, dev)


_CONJ = {}   # id(program) -> (program, is a conjunction): the answer for a cached program


def _is_conjunction(b):
    """Whether the kernels' conjunction form covers the bound predicate program b (the answer
    depends on the program and its columns' widths, fixed for a cached program)."""
    ent = _CONJ.get(id(b.prog))
    if ent is not None and ent[0] is b.prog:
        return ent[1]
    yes = bool(L.load().tq_filter_is_conjunction(E.program_ptr(b), E.colset_ptr(b)))
    if len(_CONJ) > 512:
        _CONJ.clear()
    _CONJ[id(b.prog)] = (b.prog, yes)
    return yes


class LazyMask(Variable):
    """A filter's bool mask whose kernel runs when `.tensor` or `.tile_counts` is first read."""

    def __init__(self, tree, columns, subquery_result):
        self._tree = tree
        self._cols = columns          # {id: Variable} snapshot of the filter's inputs
        self._sub = subquery_result
        self._t = None
        self._tc = None
        self._conj = False            # not yet asked
        self.rows = next(iter(columns.values())).tensor.shape[0]
        super().__init__(None, 'bool', None, VariableState.GPU)

    def _run(self):
        self._t, self._tc = kernels.filter_mask(self._tree, self._cols, self._sub)
        self._cols = None

    @property
    def tensor(self):
        if self._t is None:
            self._run()
        return self._t

    @tensor.setter
    def tensor(self, t):
        if t is not None:
            self._t = t

    @property
    def tile_counts(self):
        if self._t is None:
            self._run()
        return self._tc

    @tile_counts.setter
    def tile_counts(self, v):
        if v is not None:
            self._tc = v

    @property
    def materialized(self):
        return self._t is not None

    def conjunction(self):
        """The predicate bound to its columns when it is a conjunction of column-versus-constant
        comparisons over columns of `rows` rows (what a consumer's kernel can evaluate itself),
        else None.  Only for a mask not yet computed."""
        if self._conj is False:
            self._conj = None
            if self._t is None:
                b = kernels.filter_program(self._tree, self._cols, self._sub)
                if b.n == self.rows and b.col_tensors and _is_conjunction(b):
                    self._conj = b
        return self._conj

    def __repr__(self):
        if self._t is None:
            return "Variable (filter mask, not yet computed, tensor type = 'bool')"
        return super().__repr__()


def tqp_filter(gpu_enable, tensor_group, args, subquery_result, cpu_mem_pool, gpu_mem_pool, name,
               lazy=False):
    perf_logger().start(name)
    assert 'mask map' not in args, "Filter not accepting masks"
    if not gpu_enable:
        raise NotImplementedError("tqp_amd executes operators on the GPU only (gpu_enable=False has no CPU path)")
    tree = args['filter tree']
    input_columns = {k: tensor_group[k] for k in args['materialized in']}

    if len(input_columns) == 0:
        message_logger().debug("Filter not dispatched")
        perf_logger().stop(name)
        return -1

    if lazy and all(type(v) is Variable and v.tensor.is_cuda and v.tensor.numel() > 0
                    and v.tensor.is_contiguous() for v in input_columns.values()):
        cols = {k: Variable(v.tensor, v.tensor_type, None, VariableState.GPU)
                for k, v in input_columns.items()}
        tensor_group[args['out mask id']] = LazyMask(tree, cols, subquery_result)
        perf_logger().stop(name)
        return -1

    counts_box = []

    def filter_wrapper(cols):
        mask, counts = kernels.filter_mask(tree, cols, subquery_result)
        counts_box.append(counts)
        return (mask,)

    pipe = InMemoryPipeline(
        input_columns=input_columns,
        output_columns_type=[torch.bool],
        operator=filter_wrapper,
        chunk_size=2_000_000_000,
        cpu_mem_pool=cpu_mem_pool,
        gpu_mem_pool=gpu_mem_pool,
        name=f"{name} filter pipe",
    )
    pipe.do_exchange(20_000_000)
    mask = pipe.get_result()[0]

    t = mask.tensor
    if t.dim() == 2:
        assert t.shape[0] == 1, "squeeze failed."
        t = t.squeeze(0)
    out = Variable(t, 'bool', mask.memory_ptr, VariableState.GPU)
    if counts_box:
        out.tile_counts = counts_box[0]
    tensor_group[args['out mask id']] = out
    perf_logger().stop(name)
    return -1
