"""Host-side bookkeeping of the operator layer, on the CPU (no kernel launches)."""
import gc

import torch

from tqp_amd import kernels


def test_value_bounds_follow_the_tensor_and_its_version():
    t = torch.arange(10)
    kernels.note_bounds(t, 0, 9)
    assert kernels.known_bounds(t) == (0, 9)
    assert kernels.known_bounds(t.clone()) is None        # another tensor
    t.add_(1)                                             # written in place: bounds are stale
    assert kernels.known_bounds(t) is None


def test_value_bounds_skip_inference_tensors():
    """main.py runs every query under torch.inference_mode(); tensors made there carry no
    version counter, so no bounds are recorded or trusted for them (and no call raises)."""
    with torch.inference_mode():
        t = torch.arange(10)
        kernels.note_bounds(t, 0, 9)
        assert kernels.known_bounds(t) is None
        t.add_(1)
        assert kernels.known_bounds(t) is None
    assert kernels.known_bounds(t) is None
    u = torch.arange(4)                       # an ordinary tensor read inside inference mode
    kernels.note_bounds(u, 0, 3)
    with torch.inference_mode():
        assert kernels.known_bounds(u) == (0, 3)
        v = u[:2]                             # a view of an ordinary tensor stays versioned
        kernels.note_bounds(v, 0, 1)
        assert kernels.known_bounds(v) == (0, 1)


def test_value_bounds_entry_dies_with_the_tensor():
    u = torch.arange(5)
    kernels.note_bounds(u, 0, 4)
    k = id(u)
    del u
    gc.collect()
    assert k not in kernels._BOUNDS


def test_untransform_inverts_the_order_transform():
    for v in (0, 1, -1, 2**31 - 1, -2**31):
        assert kernels._untransform((v & 0xFFFFFFFF) ^ 0x80000000, 32) == v
    for v in (0, 5, -5, 2**63 - 1, -2**63):
        assert kernels._untransform((v & 0xFFFFFFFFFFFFFFFF) ^ 0x8000000000000000, 64) == v


def test_program_cache_rebinds_columns_and_recompiles_on_dtype_change(monkeypatch):
    from tqp_amd import _lib as L
    from tqp_amd import expr as E
    monkeypatch.setattr(L, "require_gpu", lambda *a, **k: None)
    E._PROGRAMS.clear()

    class Col:
        def __init__(self, t):
            self.tensor = t

    tree = [['d#5', '>', '10'], 'AND', ['d#5', '<', '20']]
    key = ("t", repr(tree))
    a = {5: Col(torch.arange(100, dtype=torch.int32))}
    b1 = E.cached_program(key, a, lambda pb: E.compile_predicate(pb, tree, a))
    assert isinstance(b1, E.ProgramBuilder) and key in E._PROGRAMS
    c = {5: Col(torch.arange(50, dtype=torch.int32))}
    b2 = E.cached_program(key, c, lambda pb: E.compile_predicate(pb, tree, c))
    assert isinstance(b2, E.BoundProgram)
    assert b2.prog is b1.prog and b2.n == 50
    assert b2.cols.ptr[0] == c[5].tensor.data_ptr()
    # same tree over an int64 column compiles again (the compare dtype may change)
    d = {5: Col(torch.arange(30, dtype=torch.int64))}
    b3 = E.cached_program(key, d, lambda pb: E.compile_predicate(pb, tree, d))
    assert isinstance(b3, E.ProgramBuilder) and b3.cols.ptr[0] == d[5].tensor.data_ptr()


def test_like_compiles_to_one_string_instruction(monkeypatch):
    """'%w1%w2%' becomes one STR_LIKE2 over the 2-D column, both words in the string pool and
    their offsets and lengths packed into the immediate; other % counts are refused."""
    import pytest

    from tqp_amd import _lib as L
    from tqp_amd import expr as E
    monkeypatch.setattr(L, "require_gpu", lambda *a, **k: None)

    class Col:
        def __init__(self, t):
            self.tensor = t

    cols = {7: Col(torch.zeros((10, 16), dtype=torch.int8))}
    pb = E.ProgramBuilder()
    E.compile_predicate(pb, ['c#7', 'LIKE', '%special%requests%'], cols)
    ins = [pb.prog.code[i] for i in range(pb.ninstr)]
    like = [i for i in ins if i.op == L.OP["STR_LIKE2"]]
    assert len(like) == 1
    imm = like[0].imm
    off1, len1, off2, len2 = imm & 0xFFFF, (imm >> 16) & 0xFFFF, (imm >> 32) & 0xFFFF, imm >> 48
    pool = bytes(pb.prog.strpool[i] for i in range(L.TQ_STRPOOL))
    assert pool[off1:off1 + len1] == b"special" and pool[off2:off2 + len2] == b"requests"
    with pytest.raises(AssertionError):
        E.compile_predicate(E.ProgramBuilder(), ['c#7', 'LIKE', '%special%'], cols)
