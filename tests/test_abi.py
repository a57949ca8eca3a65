"""The C-ABI library loads and exports every entry point include/ds2d.h
declares (no device work is done here)."""
import os
import re

import ds2d_amd.lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "ds2d.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ds2d_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_surface():
    syms = declared_symbols()
    assert len(syms) >= 19
    assert set(syms) == set(L.SIGNATURES), set(syms) ^ set(L.SIGNATURES)


def test_library_exports_all_symbols():
    lib = L.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert lib.ds2d_last_error() is not None


def test_invalid_arguments_fail_loudly_without_device_work():
    lib = L.load()
    assert lib.ds2d_create(None, None, None) == -1
    assert b"null" in lib.ds2d_last_error()
    assert lib.ds2d_laplace_filter(0, 0, 2, 1.0, None) == -1
    assert lib.ds2d_ricker(0, 15.0, 0.001, None) == -1


def test_missing_library_raises(tmp_path):
    old = L._lib
    L._lib = None
    try:
        try:
            L.load(str(tmp_path / "nope.so"))
        except L.DS2DError as e:
            assert "not built" in str(e)
        else:
            raise AssertionError("expected DS2DError")
    finally:
        L._lib = old
