"""The C-ABI library loads and exports every entry point include/ds2d.h
declares (no device work is done here)."""
import os
import re

import ctypes as C

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
    # U-Net / learned entries validate before touching a device
    h = L._H()
    assert lib.ds2d_unet_create(C.byref(L.UnetCfg(4, 32, 8, 100, 64, 0)), C.byref(h)) == -1  # 100 % 8 != 0
    assert b"multiple" in lib.ds2d_last_error()
    assert lib.ds2d_unet_create(C.byref(L.UnetCfg(4, 48, 8, 64, 64, 0)), C.byref(h)) == -1
    assert lib.ds2d_unet_create(C.byref(L.UnetCfg(5, 32, 8, 64, 64, 0)), C.byref(h)) == -1  # 512 channels
    assert lib.ds2d_conv3x3_bf16(3, 1, 8, 8, 32, 0, 0, 32, None, None, None, None, None, 0, 1, None, None) == -1
    assert lib.ds2d_conv3x3_bf16(0, 1, 8, 8, 24, 0, 0, 32, None, None, None, None, None, 0, 1, None, None) == -1
    assert lib.ds2d_set_batch(None, 1) == -1
    assert lib.ds2d_lsrtm_learned(None, None, 0, None, None, 0, None, None, None, None) == -1
    assert lib.ds2d_unet_train(None, 1, 1, 1e-3, 0, None, None) == -1


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
