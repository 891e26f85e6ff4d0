"""The C-ABI library loads and exports every symbol include/xpr_hip.h declares
(no GPU work: only the pure version call is made)."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xpr_hip.h")


def declared_functions(path=HEADER):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(xpr_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_hot_path():
    names = declared_functions()
    for n in ("xpr_forward", "xpr_error", "xpr_backward", "xpr_preprocess"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from xpulse_recon_amd import _lib
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(xpr_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing
    for n in declared_functions():
        assert hasattr(lib, n)
    # every binding in the Python front end names a declared symbol
    assert set(_lib.exported_symbols()) <= set(declared_functions())
    assert lib.xpr_abi_version() == 2


def test_proj_args_layout_matches_the_header():
    """xpr_proj_args as the kernels take it by value: interp fills the
    padding before view_index (80 bytes, interp at 68, view_index at 72)."""
    import ctypes as C
    from xpulse_recon_amd import _lib
    assert C.sizeof(_lib.ProjArgs) == 80
    assert _lib.ProjArgs.interp.offset == 68
    assert _lib.ProjArgs.view_index.offset == 72


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from xpulse_recon_amd import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load()
