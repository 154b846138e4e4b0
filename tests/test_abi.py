"""The C-ABI library loads and exports every entry point include/aero_hip.h
declares (no device calls: these run without a GPU)."""

import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "aero_hip.h"


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(aero_\w+)\s*\(", text, re.M)))


def test_header_declares_the_binding():
    from aero_amd import _abi
    assert set(declared()) == set(_abi.EXPORTED)


def test_library_exports_every_symbol():
    from aero_amd import _abi
    if not _abi.LIB_PATH.exists():
        pytest.fail("libaero_hip.so is not built (python -m aero_amd.build)")
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    missing = [name for name in declared() if not hasattr(lib, name)]
    assert not missing, missing
    assert lib.aero_version() == 1


def test_desc_layout_matches_header():
    """sizeof/offsets of aero_desc agree between C and ctypes (gcc on the header)."""
    import subprocess
    import tempfile
    from aero_amd import _abi
    src = ('#include <stdio.h>\n#include <stddef.h>\n#include "aero_hip.h"\n'
           'int main(){printf("%zu %zu %zu %zu\\n", sizeof(aero_desc), offsetof(aero_desc, omega),'
           ' offsetof(aero_desc, center), offsetof(aero_desc, mrt_minv));return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        c = pathlib.Path(d) / "t.c"
        c.write_text(src)
        exe = pathlib.Path(d) / "t"
        subprocess.check_call(["gcc", f"-I{ROOT / 'include'}", str(c), "-o", str(exe)])
        got = subprocess.check_output([str(exe)]).decode().split()
    expect = [ctypes.sizeof(_abi.AeroDesc), _abi.AeroDesc.omega.offset, _abi.AeroDesc.center.offset,
              _abi.AeroDesc.mrt_minv.offset]
    assert [int(v) for v in got] == expect


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from aero_amd import _abi
    monkeypatch.setattr(_abi, "LIB_PATH", tmp_path / "absent.so")
    monkeypatch.setattr(_abi, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _abi.lib()
