"""The C-ABI library loads and exports every symbol include/shtk.h declares
(no compute calls: these run without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "shtk.h")).read()
    return sorted(set(re.findall(r"\b(shtk_[a-z_0-9]+)\s*\(", src)))


def test_header_matches_binding_list(shtk):
    assert _declared() == sorted(shtk.EXPORTS)


def test_library_exports_all_symbols(shtk):
    lib = shtk.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.shtk_version().decode().count(".") == 2


def test_plan_validation_without_gpu(shtk):
    """Invalid descriptors are rejected before any device call."""
    lib = shtk.load_library()
    from shtnskit_jl_amd._lib import ShtkDesc, SHTK_EINVAL

    bad = ShtkDesc(10, 11, 1, 11, 32, 0, 1, 0, 0, 0)   # mmax > lmax
    h = ctypes.c_void_p()
    assert lib.shtk_plan_create(ctypes.byref(bad), ctypes.byref(h)) == SHTK_EINVAL
    buf = ctypes.create_string_buffer(256)
    assert lib.shtk_last_error(buf, 256) == SHTK_EINVAL
    assert b"mmax" in buf.value
    bad = ShtkDesc(10, 10, 1, 10, 32, 0, 1, 0, 0, 0)   # nlat < lmax+1
    assert lib.shtk_plan_create(ctypes.byref(bad), ctypes.byref(h)) == SHTK_EINVAL
    bad = ShtkDesc(10, 10, 1, 11, 20, 0, 1, 0, 0, 0)   # nlon < 2*mmax+1
    assert lib.shtk_plan_create(ctypes.byref(bad), ctypes.byref(h)) == SHTK_EINVAL


def test_config_mirror_validation(shtk):
    import pytest

    with pytest.raises(ValueError):
        shtk.create_gauss_config(10, 10)
    with pytest.raises(ValueError):
        shtk.create_gauss_config(10, 11, nlon=20)
    cfg = shtk.create_gauss_config(10, 11)
    assert cfg.nlon == 22 and cfg.nlm == 66
