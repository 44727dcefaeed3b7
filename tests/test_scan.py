"""Scan ingress on the host (gpu_enable=False): TorchScript column files as the reference's scan
reads them (operators/scan.py:8-54), schema types, casts and the primary-key flag, against the
oracle's restatement; the file format written here follows the reference's reader (the first
parameter of a TorchScript module)."""
import numpy as np
import pytest
import torch

from oracle import ref_ops as R

SCHEMA = {"T_KEY": "identifier", "T_QTY": "integer", "T_PRICE": "decimal",
          "T_DAY": "date", "T_FLAG": "fixed text, size 1", "T_NAME": "variable text, size 8",
          "PRIMARY_KEY": "T_KEY"}


class _Holder(torch.nn.Module):
    def __init__(self, t):
        super().__init__()
        self.col = torch.nn.Parameter(t, requires_grad=False)

    def forward(self):
        return self.col


def write_columns(path, n, seed=0, sf=1):
    """Column files like the reference's data: keys as int32, quantities as int64, prices as
    float32, days as int64, text as 2-D uint8.  Returns the arrays written."""
    g = np.random.default_rng(seed)
    cols = {
        "T_KEY": torch.from_numpy(np.arange(1, n + 1, dtype=np.int32)),
        "T_QTY": torch.from_numpy(g.integers(1, 51, n).astype(np.int64)),
        "T_PRICE": torch.from_numpy(g.uniform(900, 1e5, n).astype(np.float32)),
        "T_DAY": torch.from_numpy(g.integers(0, 2557, n).astype(np.int64)),
        "T_FLAG": torch.from_numpy(g.integers(65, 91, (n, 1)).astype(np.uint8)),
        "T_NAME": torch.from_numpy(g.integers(97, 123, (n, 8)).astype(np.uint8)),
    }
    for name, t in cols.items():
        torch.jit.save(torch.jit.script(_Holder(t)), str(path / f"SF{sf}-tensor-{name}.pth"))
    return cols


ARGS = {'name': 'Scan', 'table name': 'T',
        'output names': ['t_key', 't_qty', 't_price', 't_day', 't_flag', 't_name'],
        'output ids': [1, 2, 3, 4, 5, 6]}


def test_scan_on_the_host_matches_the_oracle(tmp_path):
    from tqp_amd.operators.scan import tqp_scan
    write_columns(tmp_path, 1000)
    got, want = {}, {}
    n = tqp_scan(False, got, ARGS, {'tensors path': str(tmp_path)}, SCHEMA, 1, None)
    m = R.tqp_scan(want, ARGS, str(tmp_path), SCHEMA, 1)
    assert n == m == 1000
    for k in range(1, 7):
        assert got[k].tensor.dtype == want[k].tensor.dtype, k
        assert torch.equal(got[k].tensor, want[k].tensor), k
        assert got[k].tensor_type == want[k].tensor_type, k
    assert [got[k].tensor.dtype for k in range(1, 7)] == \
        [torch.int64, torch.int64, torch.float64, torch.int32, torch.int8, torch.int8]
    assert got[1].is_sorted and not got[2].is_sorted
    assert [got[k].tensor_type for k in range(1, 7)] == \
        ['int', 'int', 'float', 'date', 'string', 'string']


def test_scan_reads_plain_tensor_files(tmp_path):
    """A file holding a bare tensor (torch.save) is read with weights_only=True."""
    from tqp_amd.operators.scan import tqp_scan
    torch.save(torch.arange(5, dtype=torch.int32), str(tmp_path / "SF1-tensor-T_KEY.pth"))
    got = {}
    args = {'output names': ['t_key'], 'output ids': [1]}
    assert tqp_scan(False, got, args, {'tensors path': str(tmp_path)}, SCHEMA, 1, None) == 5
    assert got[1].tensor.tolist() == [0, 1, 2, 3, 4] and got[1].tensor.dtype == torch.int64


def test_scan_missing_file_fails(tmp_path):
    from tqp_amd.operators.scan import tqp_scan
    with pytest.raises((RuntimeError, FileNotFoundError, ValueError)):
        tqp_scan(False, {}, ARGS, {'tensors path': str(tmp_path)}, SCHEMA, 1, None)
