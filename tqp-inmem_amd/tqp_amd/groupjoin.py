"""A Join whose pairs feed only the next Aggregate, grouped by the join key, run as one fused
operation ("groupjoin"): the join of operators/hashjoin.py:179-248 followed by the aggregate of
operators/aggregate.py:236-393 when the aggregate's single grouping key is the join key.

With unique selected build keys every group is one build key and every pair one probe row, so
the aggregate reduces probe-row values into the build keys' slots (tq_groupjoin_run: three
launches, no pair columns, no synchronisation).  The plan's side effects are the two
operators': the join's `mask map` entries are popped, every `materialized in` column of the
join is replaced (by a column that runs the join when first read: nothing in a plain
filter -> join -> aggregate -> top-k plan reads them, so their pairs are never written), and
the aggregate's key and value columns are written as tqp_hash_aggregate writes them.

The result words are read back once, when a consumer needs a count (`GroupJoinDeferred`).  A
repeated selected build key, a build key outside the column's recorded bounds, a join without
pairs or a look-back that gave up voids the fused result: the executor then runs the two
operators as usual from the state saved before the join.
"""
import ctypes
import functools

import torch

from . import _lib as L
from . import expr as E
from . import kernels, memory
from .conversion import get_id
from .operators.aggregate import _is_key, _parse_results, _strip_alias, _unwrap
from .operators.filter import LazyMask
from .variable import Variable, VariableState, torch_to_type

OPS = kernels.GroupBy.OPS


# ---------------------------------------------------------------- eligibility
class _Plan:
    """What the fused launch needs, worked out on the host from the two nodes' arguments."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def _one_mask(ids, mask_map):
    """The single mask id shared by all of `ids` (None: unmasked), or False when they differ."""
    masks = {mask_map.get(i) for i in ids}
    return masks.pop() if len(masks) == 1 else False


_STATIC = {}           # argument signature -> _Plan (or None: never fusable)
_STATIC_MAX = 256


def _signature(ja, aa):
    """Everything in the two nodes' arguments that decides fusability and the fused program."""
    return (ja.get('name'), ja.get('join type'), 'join condition' in ja,
            tuple(ja.get('left keys', ())), tuple(ja.get('right keys', ())),
            tuple(ja.get('left group', ())), tuple(ja.get('right group', ())),
            tuple(ja.get('materialized in', ())), tuple(sorted(ja.get('mask map', {}).items())),
            aa.get('name'), bool(aa.get('is partial')), tuple(aa.get('input ids', ())),
            repr(aa.get('keys')), repr(aa.get('results')),
            tuple(sorted(aa.get('mask map', {}).items())), tuple(aa.get('out agg ids', ())))


def _static_plan(ja, aa):
    """The checks and parsing that depend on the arguments alone (cached by their signature)."""
    if ja.get('name') != 'Join' or aa.get('name') != 'Aggregate' or aa.get('is partial'):
        return None
    if ja.get('join type') != 'Inner' or 'join condition' in ja:
        return None
    if len(ja['left keys']) != 1 or len(ja['right keys']) != 1:
        return None
    lk, rk = ja['left keys'][0], ja['right keys'][0]
    left, right = set(ja['left group']), set(ja['right group'])
    mat = list(ja['materialized in'])
    mask_map = ja.get('mask map', {})
    if lk not in left or rk not in right:
        return None
    # the aggregate reads only the join's outputs, unmasked, and groups by the join key alone
    ids = list(aa['input ids'])
    if not set(ids) <= set(mat) or any(i in aa.get('mask map', {}) for i in ids):
        return None
    keys = [get_id(k) for k in aa['keys']] if aa['keys'] else []
    if len(keys) != 1 or keys[0] not in (lk, rk):
        return None
    gkey = keys[0]
    try:
        calls, plans = _parse_results(aa['results'])
    except Exception:
        return None
    call_ids = {id(c) for c in calls}
    for _, kind, what in plans:
        if kind == 'key':
            if what != gkey:
                return None
        elif id(_strip_alias(what)) not in call_ids:
            return None       # an expression over aggregates: evaluated over the groups
    if not calls or len(calls) > L.TQ_MAX_SEGS:
        return None
    # aggregate arguments over probe columns only (read at the matched probe row)
    arg_ids = set()
    for c in calls:
        _collect_ids(c[1], arg_ids)
    if not arg_ids <= (right & set(mat)):
        return None
    bmask_id = mask_map.get(lk)
    pmask_id = _one_mask([rk] + sorted(arg_ids), mask_map)
    if pmask_id is False:
        return None
    return _Plan(lk=lk, rk=rk, gkey=gkey, calls=calls, plans=plans, arg_ids=sorted(arg_ids),
                 bmask_id=bmask_id, pmask_id=pmask_id,
                 prog_key=("groupjoin", repr([c for c in calls])))


def plan(tensor_group, ja, aa):
    """A _Plan when (Join ja, Aggregate aa) can run fused, else None.  Nothing is modified."""
    sig = _signature(ja, aa)
    if sig in _STATIC:
        p = _STATIC[sig]
    else:
        p = _static_plan(ja, aa)
        if len(_STATIC) >= _STATIC_MAX:
            _STATIC.clear()
        _STATIC[sig] = p
    if p is None:
        return None
    # the columns themselves
    if p.lk not in tensor_group or p.rk not in tensor_group:
        return None
    bk, pk = tensor_group[p.lk].tensor, tensor_group[p.rk].tensor
    for t in (bk, pk):
        if t.dim() == 2 and t.shape[1] == 1:
            continue
        if t.dim() != 1:
            return None
    if bk.dtype.is_floating_point or pk.dtype.is_floating_point or bk.dtype == torch.bool \
            or pk.dtype == torch.bool or not bk.is_cuda or not pk.is_cuda:
        return None
    for mid, n in ((p.bmask_id, bk.shape[0]), (p.pmask_id, pk.shape[0])):
        if mid is None:
            continue
        if mid not in tensor_group:
            return None
        mv = tensor_group[mid]
        if isinstance(mv, LazyMask) and not mv.materialized:
            if mv.rows != n:          # (a mask not yet computed is a 1-D bool of `rows` rows)
                return None
            continue
        m = mv.tensor
        if m.dtype != torch.bool or m.dim() != 1 or m.shape[0] != n:
            return None
    np_ = pk.shape[0]
    for i in p.arg_ids:
        if tensor_group[i].tensor.shape[0] != np_:
            return None
    return p


def _collect_ids(tree, out):
    if isinstance(tree, str):
        if '#' in tree and tree.split('#')[-1].replace('L', '').isdigit() \
                and not tree.startswith('Brand#'):
            out.add(get_id(tree))
        return
    for e in tree:
        _collect_ids(e, out)


# ---------------------------------------------------------------- presence arrays
_PRES = {}   # (device, stream) -> [byte tensor, last epoch]


def _presence(nbytes, device):
    """The byte array of this stream (zeroed when allocated or grown, and when its 255 epochs
    are used up) and the epoch for the next call.  It outlives queries and requests (its epochs
    are what spare the clearing), so it comes from torch, never from a device arena."""
    key = (str(device), torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))
    ent = _PRES.get(key)
    if ent is None or ent[0].numel() < nbytes:
        ent = [torch.zeros(max(nbytes, 256) + max(nbytes, 256) // 4, dtype=torch.uint8,
                           device=device), 0]
        _PRES[key] = ent
    if ent[1] >= 255:
        ent[0].zero_()     # a fill kernel on this stream, ordered before the next build
        ent[1] = 0
    ent[1] += 1
    return ent[0], ent[1]


# ---------------------------------------------------------------- the deferred result
class GroupJoinDeferred(kernels.Deferred):
    """The fused call's counts, read back once.  While it is the thread's active deferral, a
    top-k (limit <= 64) over the aggregate's outputs reads the group count on the device, as it
    does after a deferred group-by (kernels.sort_perm)."""

    def __init__(self, words, group, trace_at, bytes_fn):
        super().__init__()
        self.words = words
        self.group = group
        self.trace_at = trace_at
        self.bytes_fn = bytes_fn
        self.ticket = None

    def resolve_join(self):
        return self.resolve()[0]

    def post(self):
        """Queue the read of the counts behind the work on the stream (the top-k's included);
        `resolve` collects them."""
        if self.ticket is None and self.G is None:
            self.ticket = kernels.post_words([self.words], [40], 80)

    def resolve(self):
        if self.G is not None:
            return self.m, self.G
        self.post()
        ticket, self.ticket = self.ticket, None
        host = kernels.wait_words(ticket)
        G, m, nsel, nset, err = (int(host[i]) for i in range(5))
        if err or nsel != nset or m <= 0:
            raise kernels.DeferredJoinFailed()
        self.m, self.G = m, G
        for kern, nbytes in self.bytes_fn(m, G).items():
            L.trace_set_bytes(self.trace_at, kern, nbytes)
        self.sorted_bytes(G)
        return m, G


def _flat(t):
    """A key column as a contiguous 1-D tensor (a view when it already is one)."""
    if t.dim() != 1:
        t = t.reshape(-1)
    return t if t.is_contiguous() else t.contiguous()


def _clone(v):
    """A shallow copy of a Variable (of any subclass): a new object with the same fields."""
    c = object.__new__(type(v))
    c.__dict__.update(v.__dict__)
    return c


def _copy_args(x):
    """A copy of plan arguments (nested dicts, lists and tuples of immutable leaves)."""
    if isinstance(x, dict):
        return {k: _copy_args(v) for k, v in x.items()}
    if isinstance(x, list):
        return [_copy_args(v) for v in x]
    if isinstance(x, tuple):
        return tuple(_copy_args(v) for v in x)
    return x


def _snapshot_args(ja):
    """The join's arguments as they are now, safe from the in-place edits a join makes to them
    (join_vortex pops 'mask map' entries, hashjoin.py:196-197): a new dict whose dict and list
    values are copied one level deep.  The fused path only takes joins without conditions, whose
    other values are flat lists of ids and names; the replay deep-copies the snapshot before
    running the join."""
    return {k: dict(v) if type(v) is dict else list(v) if type(v) is list else v
            for k, v in ja.items()}


# ---------------------------------------------------------------- lazily materialised join outputs
class _JoinReplay:
    """Runs the join itself (operators/hashjoin.py:179-248) on the state saved before the fused
    call, the first time one of its output columns is read."""

    def __init__(self, run_join, saved_group, saved_args):
        self.run_join = run_join
        self.saved_group = saved_group
        self.saved_args = saved_args
        self.cols = None
        self.arena = memory.active()      # the replay allocates where the query did

    def column(self, cid):
        if self.cols is None:
            tg = {k: _clone(v) for k, v in self.saved_group.items()}
            active = kernels._deferred()
            kernels.set_deferred(None)     # the replay is an ordinary, settled join
            try:
                with memory.use(self.arena):
                    self.run_join(tg, _copy_args(self.saved_args))
            finally:
                kernels.set_deferred(active)
            self.cols = {k: tg[k].tensor for k in self.saved_args['materialized in']}
            self.saved_group = None
        return self.cols[cid]


class LazyJoinColumn(Variable):
    """A join output column whose rows are computed when `.tensor` is first read."""

    def __init__(self, replay, cid, tensor_type, rows):
        self._replay = replay
        self._cid = cid
        self._t = None
        self.rows = rows            # the pair count, once the fused result is settled
        super().__init__(None, tensor_type, None, VariableState.GPU)

    @property
    def tensor(self):
        if self._t is None:
            self._t = self._replay.column(self._cid)
        return self._t

    @tensor.setter
    def tensor(self, t):
        if t is not None:
            self._t = t

    @property
    def materialized(self):
        return self._t is not None

    def __repr__(self):
        if self._t is None:
            return f"Variable (join output {self._cid}, not yet computed, tensor type = " \
                   f"'{self.tensor_type}')"
        return super().__repr__()


# ---------------------------------------------------------------- the fused call
def _out_dtype(op, dt):
    if op == L.TQ_AGG_COUNT:
        return torch.int64
    if op == L.TQ_AGG_AVG:
        return torch.float32 if dt == torch.float32 else torch.float64
    if op == L.TQ_AGG_SUM:
        return dt if dt.is_floating_point else torch.int64
    return dt


def run(tensor_group, ja, aa, p, run_join):
    """Run Join ja and Aggregate aa fused (p = plan(...)), applying both operators' effects to
    tensor_group and the args.  Returns the GroupJoinDeferred holding the counts on the device.
    Raises kernels.DeferredJoinFailed before changing anything when the launch cannot be made
    (no key bounds, a key range the kernels do not cover)."""
    lk, rk = p.lk, p.rk
    bk = _flat(tensor_group[lk].tensor)
    pk = _flat(tensor_group[rk].tensor)
    dev = pk.device
    kb = kernels.known_bounds(tensor_group[lk].tensor)
    if kb is None:
        kb = kernels.known_bounds(bk)
    if kb is None:
        kb = _key_bounds(bk)
        if kb is None:
            raise kernels.DeferredJoinFailed()
    lo, hi = kb
    rng = hi - lo + 1
    nb, np_ = bk.shape[0], pk.shape[0]
    naggs = len(p.calls)
    ws_bytes, pres_bytes = _sizes(nb, np_, rng, naggs, 1)
    if ws_bytes == 0:
        raise kernels.DeferredJoinFailed()
    bmask, bpred = _mask_or_predicate(tensor_group, p.bmask_id)
    pres, epoch = _presence(pres_bytes, dev)
    ws = L.workspace(ws_bytes, dev, "groupjoin")
    key_dtype = tensor_group[p.gkey].tensor.dtype

    # the descriptor: the build side first, then the probe side and the outputs
    d = L.GroupJoinDesc()
    d.bkey, d.bkey_dtype = bk.data_ptr(), L.dtype_code(bk.dtype)
    d.bmask, d.nb = (bmask.data_ptr() if bmask is not None else None), nb
    d.pkey, d.pkey_dtype = pk.data_ptr(), L.dtype_code(pk.dtype)
    d.np = np_
    d.key_lo, d.key_hi, d.epoch, d.naggs = lo, hi, epoch, naggs
    d.out_key_dtype = L.dtype_code(key_dtype)
    if bpred is not None:
        d.bpred = ctypes.addressof(bpred.prog)
        d.bpred_cols = ctypes.addressof(bpred.cols)
    # a fused filter reads its columns in place of the mask byte
    bsel = (1 if bmask is not None else 0) + (bpred.load_bytes() if bpred is not None else 0)
    words = ctypes.c_void_p()
    build_bytes = nb * (bk.element_size() + bsel)

    pmask, ppred = _mask_or_predicate(tensor_group, p.pmask_id)
    probe_cols = {i: tensor_group[i] for i in p.arg_ids}

    def compile_fn(b):
        dtypes = []
        for c in p.calls:
            v = E.compile_row_expr(b, c[1], probe_cols)
            if v.kind == "lit":
                v = b.const(v.value, E.literal_dtype(v.value))
            if v.kind != "reg":
                raise NotImplementedError("string-valued aggregate argument")
            dtypes.append(v.dtype)
            b.end_segment(v)
        return dtypes

    try:
        b = E.cached_program(p.prog_key, probe_cols, compile_fn)
    except NotImplementedError:
        raise kernels.DeferredJoinFailed()   # nothing has been launched
    arg_dtypes = list(b.extra)
    if not b.col_tensors:
        # no column read (count(*) only): point the column set at the probe key
        b.col_tensors = [pk]
        b.cols.ptr[0] = pk.data_ptr()
        b.cols.width[0] = 1
    ops = [OPS[c[0]] for c in p.calls]
    out_dtypes = [_out_dtype(op, dt) for op, dt in zip(ops, arg_dtypes)]
    cap = max(1, min(rng, np_))
    key_full = memory.empty(cap, dtype=key_dtype, device=dev)
    agg_full = [memory.empty(cap, dtype=dt, device=dev) for dt in out_dtypes]
    d.pmask = pmask.data_ptr() if pmask is not None else None
    for j, (op, dt) in enumerate(zip(ops, arg_dtypes)):
        d.agg_op[j] = op
        d.agg_dtype[j] = L.dtype_code(dt)
        d.out_agg[j] = agg_full[j].data_ptr()
    d.out_key, d.out_cap = key_full.data_ptr(), cap
    if ppred is not None:
        d.ppred = ctypes.addressof(ppred.prog)
        d.ppred_cols = ctypes.addressof(ppred.cols)

    at = L.trace_len()
    pkbytes = pk.element_size()
    load_bytes = b.load_bytes()
    rec_bytes = 8 * _record_words(p.prog_key, b, d, naggs)
    out_bytes = key_dtype.itemsize + sum(dt.itemsize for dt in out_dtypes)
    psel = (1 if pmask is not None else 0) + (ppred.load_bytes() if ppred is not None else 0)
    # one call launches the build, the probe and the reduce (the step is host-bound: a
    # separate earlier build launch would only add a call)
    d.phases = 0
    L.call("tq_groupjoin_run", ctypes.byref(d), E.program_ptr(b), E.colset_ptr(b), L.ptr(pres),
           pres.numel(), L.ptr(ws), ws_bytes, ctypes.byref(words), L.stream_handle(),
           kernel={"k_gj_build": build_bytes, "k_gj_probe": 0, "k_gj_reduce": 0})
    key_type = tensor_group[p.gkey].tensor_type

    def bytes_fn(m, G):
        # probe: key and mask of every probe row, every gathered column read once over its
        # table (the matches touch nearly every line), one record per pair; reduce: the
        # records and the presence bytes, then the groups' key and aggregate columns
        return {"k_gj_probe": np_ * (pkbytes + psel + load_bytes) + m * rec_bytes,
                "k_gj_reduce": m * rec_bytes + rng + G * out_bytes}

    df = GroupJoinDeferred(words.value, {"ngroups": words.value, "outs": [key_full] + agg_full,
                                         "cap": cap, "ws": ws, "pres": pres},
                           at, bytes_fn)

    # the join's inputs and arguments as they were before the join (nothing has changed them
    # yet), for the replay of its outputs; Variables are copied, as a later operator may
    # reassign their fields
    join_ids = set(ja['left group']) | set(ja['right group']) | set(ja.get('mask map', {}).values())
    replay = _JoinReplay(run_join, {k: _clone(v) for k, v in tensor_group.items() if k in join_ids},
                         _snapshot_args(ja))

    # ---- side effects of join_vortex (hashjoin.py:198-238): mask map pops, outputs replaced
    lgroup, rgroup = ja['left group'], ja['right group']
    join_output = []
    for side in (lgroup, rgroup):
        for used in ja['materialized in']:
            if used in side:
                join_output.append(used)
                if 'mask map' in ja and used in ja['mask map']:
                    ja['mask map'].pop(used)
    for k in join_output:
        tensor_group[k] = LazyJoinColumn(replay, k, tensor_group[k].tensor_type, None)
    df.lazy = [tensor_group[k] for k in join_output]

    # ---- side effects of tqp_hash_aggregate (aggregate.py:380-390): values, then keys
    by_call = {id(c): t for c, t in zip(p.calls, agg_full)}
    key_ids_new, agg_ids = [], aa['out agg ids']
    for tree in aa['results']:
        tree = _unwrap(tree)
        if _is_key(tree):
            key_ids_new.append(get_id(tree if isinstance(tree, str) else tree[-1]))
    vals = []
    for out_id, kind, what in p.plans:
        if kind == 'agg':
            vals.append(by_call[id(_strip_alias(what))])
    for i, t in enumerate(vals):
        tensor_group[agg_ids[i]] = Variable(t, torch_to_type(t.dtype))
    for kid in key_ids_new:
        tensor_group[kid] = Variable(key_full.reshape(-1, 1), key_type)
    df.group["outs"] = [key_full] + agg_full
    df.written = {**{agg_ids[i]: "agg" for i in range(len(vals))},
                  **{kid: "key" for kid in key_ids_new}}
    return df


_REC_WORDS = {}


def _record_words(prog_key, b, d, naggs):
    """Words per probe record (tq_groupjoin_record_words), per program and aggregate dtypes."""
    key = (prog_key, tuple(d.agg_op[:naggs]), tuple(d.agg_dtype[:naggs]))
    w = _REC_WORDS.get(key)
    if w is None:
        w = int(L.load().tq_groupjoin_record_words(E.program_ptr(b), naggs,
                                                   ctypes.addressof(d.agg_op),
                                                   ctypes.addressof(d.agg_dtype)))
        if w < 1:
            w = 1 + naggs
        if len(_REC_WORDS) > 256:
            _REC_WORDS.clear()
        _REC_WORDS[key] = w
    return w


@functools.lru_cache(maxsize=256)
def _sizes(nb, np_, rng, naggs, nload):
    """(workspace bytes, presence bytes) of a call of this shape."""
    return (L.size_call("tq_groupjoin_ws_bytes", nb, np_, rng, naggs, nload),
            L.size_call("tq_groupjoin_pres_bytes", rng))


def _mask_or_predicate(tensor_group, mid):
    """(mask tensor, None), or (None, bound conjunction) for a filter mask not yet computed
    whose predicate the kernels evaluate themselves, or (None, None) without a mask."""
    if mid is None:
        return None, None
    mv = tensor_group[mid]
    if isinstance(mv, LazyMask) and not mv.materialized:
        b = mv.conjunction()
        if b is not None:
            return None, b
    return mv.tensor, None


def _key_bounds(bk):
    """Min and max of a whole key column (one read-back), or None when it is empty."""
    n = bk.shape[0]
    if n == 0:
        return None
    d = L.GroupByDesc()
    d.nkeys = 1
    d.key_dtype[0] = L.dtype_code(bk.dtype)
    d.key_ptr[0] = bk.data_ptr()
    out = memory.empty(3, dtype=torch.int64, device=bk.device)
    ws = L.workspace(L.size_call("tq_groupby_bounds_ws_bytes", n), bk.device, "gb_bounds")
    L.call("tq_groupby_bounds", ctypes.byref(d), None, n, L.ptr(out), L.ptr(ws),
           L.stream_handle(), kernel={"k_gb_bounds": n * bk.element_size()})
    lo, hi, _ = out.cpu().tolist()
    kernels.note_bounds(bk, lo, hi)
    return lo, hi
