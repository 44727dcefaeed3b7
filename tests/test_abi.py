"""The C-ABI library loads and exports every entry point include/tqhip.h declares (CPU only:
no compute calls), and the Python mirror of the header's constants and structs matches it."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tqhip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tq_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from tqp_amd import _lib
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.tq_version() >= 10000
    assert lib.tq_tile_rows() == 4096


def test_python_signatures_cover_header():
    from tqp_amd import _lib
    assert set(declared_functions()) == set(_lib._SIGNATURES)


def test_opcode_table_matches_header():
    from tqp_amd import _lib
    text = open(HEADER).read()
    body = text[text.index("enum tq_opcode {"):text.index("TQ_OP_COUNT_")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = re.findall(r"TQ_OP_(\w+)", body)
    assert names == _lib.OPCODES


def test_struct_sizes_match_c_layout():
    from tqp_amd import _lib
    # tq_program_t: 4+4 + 12*2 + 12 + 4 + 8 + 8 (=64, 8-aligned) + 64*16 + 128
    assert ctypes.sizeof(_lib.Instr) == 16
    assert ctypes.sizeof(_lib.Program) == 64 + 64 * 16 + 128
    assert ctypes.sizeof(_lib.ColSet) == 16 * 8 + 16 * 4
    assert ctypes.sizeof(_lib.SortKey) == 24
    assert ctypes.sizeof(_lib.JoinState) == 128
    assert ctypes.sizeof(_lib.KTime) == 16


def test_error_codes_map_to_reference_exceptions():
    from tqp_amd import _lib
    # an invalid argument is rejected on the host before any device work
    with pytest.raises(ValueError):
        _lib.call("tq_filter_mask", None, None, 10, None, None, None)
    assert "null" in _lib.load().tq_last_error().decode()


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof/offsetof from gcc on include/tqhip.h equal the ctypes mirrors."""
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    from tqp_amd import _lib, batched
    checks = {
        "tq_program_t": (_lib.Program, ["nload", "code", "strpool"]),
        "tq_colset_t": (_lib.ColSet, ["width"]),
        "tq_sortkey_t": (_lib.SortKey, ["descending"]),
        "tq_join_state_t": (_lib.JoinState, ["layout", "base", "range", "bits", "rows", "pending", "nb", "key_and"]),
        "tq_ktime_t": (_lib.KTime, ["ms"]),
        "tq_groupby_desc_t": (_lib.GroupByDesc, ["key_ptr", "key_stride", "agg_dtype", "nsel", "nrows"]),
        "tq_batch_pipe_t": (batched.BatchPipe, ["pprog", "acols", "pres", "agg_out", "ntiles",
                                                "gkey_min", "gkey_out", "records", "nslices",
                                                "rec", "loff", "keycount", "fast", "prep"]),
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){']
    for cname, (_, fields) in checks.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for f in fields:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append('return 0;}')
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                   text=True).stdout.splitlines())
    for cname, (cls, fields) in checks.items():
        assert int(got[cname]) == ctypes.sizeof(cls), cname
        for f in fields:
            assert int(got[f"{cname}.{f}"]) == getattr(cls, f).offset, f"{cname}.{f}"
