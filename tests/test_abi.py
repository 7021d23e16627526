"""The C-ABI library loads, exports every symbol include/dba/dba.h declares,
and the ctypes struct mirrors match the C layout (no GPU calls)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from dba_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dba", "dba.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:dba_status|const char\*)\s+(dba_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "dba_step" in names and "dba_init" in names
    assert len(names) >= 18


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(abi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    bound = {s[0] for s in abi.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_struct_layout_matches_c(tmp_path):
    fields = [f for f, _ in abi.Params._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){',
           'printf("%zu\\n", sizeof(dba_params));']
    src += [f'printf("%zu\\n", offsetof(dba_params, {f}));' for f in fields]
    src += ['printf("%zu\\n", sizeof(dba_info));']
    src += [f'printf("%zu\\n", offsetof(dba_info, {f}));' for f, _ in abi.Info._fields_]
    src += ['return 0;}']
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(c)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == C.sizeof(abi.Params)
    for k, f in enumerate(fields):
        assert vals[1 + k] == getattr(abi.Params, f).offset, f
    base = 1 + len(fields)
    assert vals[base] == C.sizeof(abi.Info)
    for k, (f, _) in enumerate(abi.Info._fields_):
        assert vals[base + 1 + k] == getattr(abi.Info, f).offset, f


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(OSError):
        abi.load(str(tmp_path / "nope.so"))


def test_partial_tile_one_cell_wide_is_refused():
    """A grid that leaves a partial last tile one cell wide is refused at creation
    (validation happens before any device work)."""
    from dba_amd import configs
    lib = abi.load()
    p = configs.ch_params(n=65, ny=64, tile=32)
    # no context is needed to see the refusal: dba_footprint validates too
    b = C.c_int64()
    assert lib.dba_footprint(C.byref(p), C.byref(b)) != abi.OK
    assert "one cell wide" in lib.dba_last_error().decode()
    p = configs.ch_params(n=66, ny=64, tile=32)
    assert lib.dba_footprint(C.byref(p), C.byref(b)) == abi.OK
