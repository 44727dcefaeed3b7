"""The fused join + group-by entry point that allocates nothing (tq_groupjoin_run_noalloc: its
workspace and presence bytes live in the library's own per-GPU scratch array) against the
entry point that takes caller workspaces (tq_groupjoin_run), through the C ABI: the same
groups, keys and counts bit-exact, sums within 1e-12, on consecutive calls of different
sizes, and a size beyond the scratch refused before anything runs."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

FP_RTOL = 1e-12


class _Col:
    def __init__(self, t):
        self.tensor = t


def _inputs(device, nb, np_, seed, key_lo=1):
    g = np.random.default_rng(seed)
    bkey = (g.permutation(nb) + key_lo).astype(np.int32)
    bmask = g.random(nb) < 0.2
    pkey = g.integers(key_lo, key_lo + nb, np_).astype(np.int32)
    price = g.uniform(900.0, 105000.0, np_)
    disc = g.integers(0, 11, np_) / 100.0
    pmask = g.random(np_) < 0.53
    t = lambda a: torch.from_numpy(a).to(device)   # noqa: E731
    host = dict(bkey=bkey, bmask=bmask, pkey=pkey, price=price, disc=disc, pmask=pmask)
    return {k: t(v) for k, v in host.items()}, host


def _call(dev_cols, noalloc, epoch=1):
    from tqp_amd import _lib as L
    from tqp_amd import expr as E
    c = dev_cols
    nb, np_ = c["bkey"].shape[0], c["pkey"].shape[0]
    lo, hi = int(c["bkey"].min().item()), int(c["bkey"].max().item())
    b = E.ProgramBuilder()
    cols = {202: _Col(c["price"]), 203: _Col(c["disc"])}
    v = E.compile_row_expr(b, ['price#202', '*', ['1.0', '-', 'disc#203']], cols)
    b.end_segment(v)
    one = b.const(1, torch.int64)
    b.end_segment(one)
    b.finish()
    cap = min(hi - lo + 1, np_)
    dev = c["pkey"].device
    key_out = torch.empty(cap, dtype=torch.int32, device=dev)
    sums = torch.empty(cap, dtype=torch.float64, device=dev)
    counts = torch.empty(cap, dtype=torch.int64, device=dev)
    d = L.GroupJoinDesc()
    d.bkey, d.bkey_dtype, d.bmask, d.nb = c["bkey"].data_ptr(), L.TQ_I32, c["bmask"].data_ptr(), nb
    d.pkey, d.pkey_dtype, d.pmask, d.np = c["pkey"].data_ptr(), L.TQ_I32, c["pmask"].data_ptr(), np_
    d.key_lo, d.key_hi, d.epoch, d.naggs = lo, hi, epoch, 2
    d.agg_op[0], d.agg_dtype[0], d.out_agg[0] = L.TQ_AGG_SUM, L.TQ_F64, sums.data_ptr()
    d.agg_op[1], d.agg_dtype[1], d.out_agg[1] = L.TQ_AGG_COUNT, L.TQ_I64, counts.data_ptr()
    d.out_key, d.out_key_dtype, d.out_cap = key_out.data_ptr(), L.TQ_I32, cap
    s = L.stream_handle()
    if noalloc:
        words = torch.zeros(8, dtype=torch.int64, device=dev)
        L.call("tq_groupjoin_run_noalloc", ctypes.byref(d), E.program_ptr(b), E.colset_ptr(b),
               L.ptr(words), s)
        torch.cuda.synchronize()
        w = words.cpu().tolist()
    else:
        pres = torch.zeros(L.size_call("tq_groupjoin_pres_bytes", hi - lo + 1), dtype=torch.uint8,
                           device=dev)
        wsb = L.size_call("tq_groupjoin_ws_bytes", nb, np_, hi - lo + 1, 2, b.nload)
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        wp = ctypes.c_void_p()
        L.call("tq_groupjoin_run", ctypes.byref(d), E.program_ptr(b), E.colset_ptr(b), L.ptr(pres),
               pres.numel(), L.ptr(ws), wsb, ctypes.byref(wp), s)
        torch.cuda.synchronize()
        host = (ctypes.c_int64 * 8)()
        L.call("tq_read_words", 1, (ctypes.c_void_p * 1)(wp.value), (ctypes.c_size_t * 1)(64),
               host, s)
        w = list(host)
    G = w[0]
    return w[:5], key_out[:G].cpu(), sums[:G].cpu(), counts[:G].cpu()


def _closed_form(h):
    """Groups of the selected probe rows whose key is a selected build key: key, sum, count."""
    sel_b = set(h["bkey"][h["bmask"]].tolist())
    rows = np.flatnonzero(h["pmask"] & np.isin(h["pkey"], list(sel_b)))
    keys, inv = np.unique(h["pkey"][rows], return_inverse=True)
    vals = h["price"][rows] * (1.0 - h["disc"][rows])
    return keys, np.bincount(inv, weights=vals), np.bincount(inv)


def _check(a, b, what):
    wa, ka, sa, ca = a
    wb, kb, sb, cb = b
    assert wa == wb, (what, wa, wb)
    assert wa[4] == 0 and wa[2] == wa[3] and wa[1] > 0, (what, wa)
    assert torch.equal(ka, kb), what
    assert torch.equal(ca, cb), what
    np.testing.assert_allclose(sa.numpy(), sb.numpy(), rtol=FP_RTOL, atol=0, err_msg=what)


def test_noalloc_matches_caller_workspace_entry(device):
    from tqp_amd import _lib as L
    L.load()
    print("no-alloc groupjoin: 300k x 1.2M, then 40k x 150k, then 500k x 900k", flush=True)
    first, host = _inputs(device, 300_000, 1_200_000, seed=1)
    a = _call(first, noalloc=True)
    b = _call(first, noalloc=False)
    _check(a, b, "first call")
    keys, sums, counts = _closed_form(host)
    assert np.array_equal(a[1].numpy(), keys) and np.array_equal(a[3].numpy(), counts)
    np.testing.assert_allclose(a[2].numpy(), sums, rtol=FP_RTOL, atol=0)
    # a smaller call right after: the scratch keeps the last call's bytes
    second, _ = _inputs(device, 40_000, 150_000, seed=2, key_lo=7)
    _check(_call(second, noalloc=True), _call(second, noalloc=False), "second, smaller call")
    # a wider key range than before: the presence bytes grow over old workspace bytes
    third, _ = _inputs(device, 500_000, 900_000, seed=3)
    _check(_call(third, noalloc=True), _call(third, noalloc=False), "third, wider call")


def test_noalloc_refuses_sizes_beyond_its_scratch(device):
    from tqp_amd import _lib as L
    big, _ = _inputs(device, 1000, 5_000_000, seed=4)
    with pytest.raises(MemoryError, match="no allocation"):
        _call(big, noalloc=True)
