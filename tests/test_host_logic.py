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
