"""The C-ABI library loads and exports every entry point include/pfc_hip.h declares (CPU only)."""
import ctypes
import os
import re

from pfc_hip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "pfc_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pfc_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    bound = _lib.load()
    assert bound.pfc_version() == 2


def test_params_struct_layout_matches_c_compiler(tmp_path):
    """ctypes mirror of pfc_params has the C compiler's size and field offsets."""
    import shutil
    import subprocess
    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    names = [f[0] for f in _lib.PfcParams._fields_]
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "pfc_hip.h"', "int main(void){",
           'printf("%zu\\n", sizeof(pfc_params));']
    src += [f'printf("%zu\\n", offsetof(pfc_params, {n}));' for n in names]
    src += ["return 0;}"]
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(_lib.PfcParams)
    assert out[1:] == [getattr(_lib.PfcParams, n).offset for n in names]
