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


def test_grid_descriptor_validation_without_gpu(shtk):
    """Grid minimums of _min_nlat (src/config.jl:452-458) in the C ABI and the mirror."""
    import pytest

    lib = shtk.load_library()
    from shtnskit_jl_amd._lib import ShtkDesc, SHTK_EINVAL

    h = ctypes.c_void_p()
    for lmax, nlat, grid in ((10, 11, 1),     # regular needs lmax + 2
                             (10, 10, 2),     # pole grid needs max(2, lmax + 1)
                             (0, 1, 2),
                             (10, 20, 3),     # Driscoll-Healy needs 2 (lmax + 1)
                             (10, 22, 4)):    # unknown grid
        bad = ShtkDesc(lmax, lmax, 1, nlat, 2 * lmax + 2, 0, 1, 0, 0, 0, grid)
        assert lib.shtk_plan_create(ctypes.byref(bad), ctypes.byref(h)) == SHTK_EINVAL, (lmax, nlat, grid)
    with pytest.raises(ValueError):
        shtk.create_regular_config(10, 11)
    with pytest.raises(ValueError):
        shtk.create_regular_config(10, 21, include_poles=True, use_dh_weights=True)
    with pytest.raises(ValueError):
        shtk.create_regular_config(10, 22, use_dh_weights=True)  # DH needs include_poles
    assert shtk.create_regular_config(10, 12).grid_type == "regular"
    assert shtk.create_regular_config(10, 11, include_poles=True).grid_type == "regular_poles"
    assert shtk.create_regular_config(10, 22, include_poles=True, use_dh_weights=True).grid_type == "driscoll_healy"


def test_create_config_defaults(shtk):
    """create_config (src/config.jl:1376-1415): the default nlat = lmax + 2 becomes
    max(2 lmax + 1, 2) on equiangular grids and 2 (lmax + 1) for Driscoll-Healy."""
    assert shtk.create_config(10).nlat == 12 and shtk.create_config(10).grid_type == "gauss"
    assert shtk.create_config(10, grid_type="regular").nlat == 21
    c = shtk.create_config(10, grid_type="regular_poles")
    assert c.nlat == 21 and c.grid_type == "regular_poles"
    c = shtk.create_config(10, grid_type="driscoll_healy")
    assert c.nlat == 22 and c.grid_type == "driscoll_healy"
    assert shtk.create_config(10, grid_type="regular", nlat=5).nlat == 12  # raised to the minimum
    assert shtk.create_config(10, nlon=4).nlon == 21


def test_uneven_batch_rejects_mixed_arrays(shtk):
    """The batch kernels read every group with one dtype: mismatched precisions,
    real coefficients or complex grids are rejected before any device call."""
    import pytest
    import torch

    from shtnskit_jl_amd.batch import check_group_arrays

    cfgs = [shtk.create_gauss_config(7, 8, nlon=16), shtk.create_gauss_config(3, 4, nlon=8)]
    nfs = [2, 1]
    good = [torch.zeros((8, 8, 2), dtype=torch.complex128), torch.zeros((4, 4, 1), dtype=torch.complex128)]
    mixed = [good[0], good[1].to(torch.complex64)]
    real = [torch.zeros((8, 8, 2), dtype=torch.float64), good[1]]
    with pytest.raises(ValueError, match="same precision"):
        check_group_arrays(cfgs, nfs, mixed, "spec", "synthesis")
    with pytest.raises(ValueError, match="complex64 or complex128"):
        check_group_arrays(cfgs, nfs, real, "spec", "synthesis")
    grids = [torch.zeros((8, 16, 2), dtype=torch.complex128), torch.zeros((4, 8, 1), dtype=torch.float64)]
    with pytest.raises(ValueError, match="float32 or float64"):
        check_group_arrays(cfgs, nfs, grids, "grid", "analysis")
    with pytest.raises(ValueError, match="expected 2 arrays"):
        check_group_arrays(cfgs, nfs, good[:1], "spec", "synthesis")
    with pytest.raises(ValueError, match="device tensors"):
        check_group_arrays(cfgs, nfs, good, "spec", "synthesis")
