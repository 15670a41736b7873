"""SHTConfig mirror: the grid configuration of SHTnsKit.jl backed by a device
plan of libshtk_mi355x.so.

Reference: ``SHTConfig`` (src/config.jl:170-206), ``create_gauss_config``
(src/config.jl:633-655), ``create_gauss_fly_config`` (:689-707, identical here:
this backend always computes Legendre rows on the fly), ``create_regular_config``
(:1235-1330), ``create_config`` (:1376-1415), ``_min_nlat`` (:452-458), ``_default_nlon``
(:76-80), ``_validate_spectral_layout`` (:465-481), ``nlm_calc``
(src/layout.jl:88-102), and the phi-scale conventions of src/SHTnsKit.jl:141-177.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import _lib

_NORMS = {"orthonormal": 0, "fourpi": 1, "schmidt": 2}
_GRIDS = {"gauss": 0, "regular": 1, "regular_poles": 2, "driscoll_healy": 3}


def _min_nlat(grid_type: str, lmax: int) -> int:
    if grid_type == "gauss":
        return lmax + 1
    if grid_type == "regular_poles":
        return lmax + 1
    if grid_type == "regular":
        return lmax + 2
    if grid_type == "driscoll_healy":
        return 2 * (lmax + 1)
    return 1


def _default_nlon(lmax: int) -> int:
    n = 2 * lmax + 1
    return n + (n & 1)


def nlm_calc(lmax: int, mmax: int, mres: int) -> int:
    if lmax < 0 or mmax < 0 or mres < 1:
        raise ValueError("invalid sizes")
    if mmax > lmax:
        return 0
    return sum(lmax - m + 1 for m in range(0, mmax + 1, mres))


class SHTConfig:
    """Grid descriptor (Gauss-Legendre, regular Fejer, regular with poles) with a device plan.

    Attributes follow the reference: ``lmax, mmax, mres, nlat, nlon, x, w, theta,
    phi, cphi, nlm, norm, cs_phase, real_norm, robert_form, phi_scale, grid_type``.
    ``x``/``w`` are host arrays, north first.
    """

    def __init__(self, lmax: int, nlat: int, *, mmax: int | None = None, mres: int = 1,
                 nlon: int | None = None, norm: str = "orthonormal", cs_phase: bool = True,
                 real_norm: bool = False, robert_form: bool = False, phi_scale: str | None = None,
                 grid_type: str = "gauss"):
        mmax = lmax if mmax is None else int(mmax)
        nlon = _default_nlon(lmax) if nlon is None else int(nlon)
        if lmax < 0:
            raise ValueError(f"lmax must be ≥ 0, got {lmax}")
        if mmax < 0:
            raise ValueError(f"mmax must be ≥ 0, got {mmax}")
        if mmax > lmax:
            raise ValueError(f"mmax must be ≤ lmax, got mmax={mmax}, lmax={lmax}")
        if mres < 1:
            raise ValueError(f"mres must be ≥ 1, got {mres}")
        if grid_type not in _GRIDS:
            raise ValueError(f"unsupported grid_type={grid_type}; choose gauss, regular, regular_poles "
                             "or driscoll_healy")
        if grid_type == "gauss" and nlat < lmax + 1:
            raise ValueError(f"Gauss–Legendre quadrature requires nlat ≥ lmax+1, got nlat={nlat} and lmax={lmax}")
        if grid_type != "gauss" and nlat < _min_nlat(grid_type, lmax):
            raise ValueError(f"nlat must be ≥ {_min_nlat(grid_type, lmax)} for regular grids")
        if grid_type == "driscoll_healy" and (nlat % 2 or nlat != 2 * (lmax + 1)):
            raise ValueError(f"DH weights require nlat=2*(lmax+1)={2 * (lmax + 1)}, got nlat={nlat}")
        if grid_type == "regular_poles" and nlat < 2:
            raise ValueError(f"pole-inclusive grids need nlat ≥ 2 (got nlat={nlat}); "
                             "use grid_type='regular' for a single-latitude grid")
        if nlon < 2 * mmax + 1:
            raise ValueError(f"nlon must be ≥ 2*mmax+1 = {2 * mmax + 1} to resolve every azimuthal order, got nlon={nlon}")
        if norm not in _NORMS:
            raise ValueError(f"Unsupported normalization: {norm}")
        env = os.environ.get("SHTNSKIT_PHI_SCALE", "")
        if env in ("dft", "quad"):
            phi_scale = env
        phi_scale = phi_scale or "dft"
        if phi_scale not in ("dft", "quad"):
            raise ValueError(f"phi_scale must be 'dft' or 'quad', got {phi_scale}")
        self.lmax, self.mmax, self.mres, self.nlat, self.nlon = int(lmax), mmax, int(mres), int(nlat), nlon
        self.norm, self.cs_phase, self.real_norm = norm, bool(cs_phase), bool(real_norm)
        self.robert_form, self.phi_scale = bool(robert_form), phi_scale
        self.grid_type = grid_type
        self.cphi = 2 * math.pi / nlon
        self.nlm = nlm_calc(lmax, mmax, mres)
        self.nspat = nlat * nlon
        self._plan = None
        self._desc = _lib.ShtkDesc(self.lmax, mmax, self.mres, self.nlat, nlon, _NORMS[norm],
                                   int(cs_phase), int(real_norm), int(robert_form),
                                   1 if phi_scale == "quad" else 0, _GRIDS[grid_type])
        self._x = None
        self._w = None

    # --- plan ------------------------------------------------------------------
    @property
    def plan(self) -> int:
        """Opaque ``shtk_plan*`` (created on first use, on the current HIP device)."""
        if self._plan is None:
            lib = _lib.load()
            handle = ctypes.c_void_p()
            _lib.check(lib.shtk_plan_create(ctypes.byref(self._desc), ctypes.byref(handle)),
                       "create_gauss_config")
            self._plan = handle.value
        return self._plan

    def close(self) -> None:
        """Frees the device plan (``gpu_clear_cache!`` for this configuration)."""
        if self._plan is not None:
            _lib.load().shtk_plan_destroy(ctypes.c_void_p(self._plan))
            self._plan = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass

    # --- grid ------------------------------------------------------------------
    def _nodes(self):
        if self._x is None:
            x = np.zeros(self.nlat)
            w = np.zeros(self.nlat)
            _lib.check(_lib.load().shtk_plan_nodes(ctypes.c_void_p(self.plan), x.ctypes.data, w.ctypes.data),
                       "nodes")
            self._x, self._w = x, w
        return self._x, self._w

    @property
    def x(self) -> np.ndarray:
        return self._nodes()[0]

    @property
    def w(self) -> np.ndarray:
        return self._nodes()[1]

    @property
    def theta(self) -> np.ndarray:
        return np.arccos(self.x)

    @property
    def phi(self) -> np.ndarray:
        return (2 * math.pi / self.nlon) * np.arange(self.nlon)

    def phi_inv_scale(self) -> float:
        return self.nlon / (2 * math.pi) if self.phi_scale == "quad" else float(self.nlon)

    def __repr__(self) -> str:
        return (f"SHTConfig({self.grid_type}, lmax={self.lmax}, mmax={self.mmax}, mres={self.mres}, "
                f"nlat={self.nlat}, nlon={self.nlon}, norm={self.norm}, cs_phase={self.cs_phase}, "
                f"real_norm={self.real_norm}, robert_form={self.robert_form})")


def create_gauss_config(lmax: int, nlat: int, **kwargs) -> SHTConfig:
    """``create_gauss_config(lmax, nlat; mmax, mres, nlon, norm, cs_phase, real_norm, robert_form)``."""
    return SHTConfig(lmax, nlat, **kwargs)


def create_gauss_fly_config(lmax: int, nlat: int, **kwargs) -> SHTConfig:
    """On-the-fly Gauss configuration; every plan of this backend is on-the-fly."""
    return SHTConfig(lmax, nlat, **kwargs)


def create_regular_config(lmax: int, nlat: int, *, include_poles: bool = False, precompute_plm: bool = True,
                          use_dh_weights: bool = False, **kwargs) -> SHTConfig:
    """``create_regular_config(lmax, nlat; ..., include_poles, precompute_plm, use_dh_weights)``.

    ``include_poles = False``: Fejer-1 midpoints ``theta_i = (i + 1/2) pi / nlat``;
    ``True``: Clenshaw-Curtis nodes ``theta_i = i pi / (nlat - 1)``; with
    ``use_dh_weights`` the Driscoll-Healy nodes ``theta_i = i pi / nlat`` (north pole
    included, south pole not) and weights, ``nlat = 2 (lmax + 1)``.  Legendre rows
    are always computed on the fly, so ``precompute_plm`` has no effect.
    """
    if use_dh_weights and not include_poles:
        raise ValueError("Driscoll-Healy weights require include_poles=true")
    grid = "driscoll_healy" if use_dh_weights else ("regular_poles" if include_poles else "regular")
    return SHTConfig(lmax, nlat, grid_type=grid, **kwargs)


def create_config(lmax: int, *, mmax: int | None = None, mres: int = 1, nlat: int | None = None,
                  nlon: int | None = None, norm: str = "orthonormal", cs_phase: bool = True,
                  real_norm: bool = False, robert_form: bool = False, grid_type: str = "gauss",
                  include_poles: bool = False, precompute_plm: bool | None = None) -> SHTConfig:
    """``create_config(lmax; ..., grid_type)``: the default ``nlat`` (lmax + 2) becomes
    ``max(2 lmax + 1, 2)`` on the equiangular grids; other counts are raised to
    the grid's minimum (src/config.jl:1376-1415)."""
    mmax = lmax if mmax is None else int(mmax)
    nlat = lmax + 2 if nlat is None else int(nlat)
    nlon = _default_nlon(lmax) if nlon is None else int(nlon)
    if grid_type not in ("gauss", "regular", "regular_poles", "driscoll_healy"):
        raise ValueError(f"unsupported grid_type={grid_type}; choose gauss, regular, regular_poles, "
                         "or driscoll_healy")
    pole_grid = grid_type in ("regular_poles", "driscoll_healy")
    min_lat = lmax + 1 if grid_type == "gauss" else (lmax + 1 if pole_grid else lmax + 2)
    if grid_type != "gauss" and nlat == lmax + 2:
        nlat_eff = 2 * (lmax + 1) if grid_type == "driscoll_healy" else max(2 * lmax + 1, 2)
    else:
        nlat_eff = max(nlat, min_lat)
    nlon_eff = max(nlon, 2 * mmax + 1)
    kw = dict(mmax=mmax, mres=mres, nlon=nlon_eff, norm=norm, cs_phase=cs_phase, real_norm=real_norm,
              robert_form=robert_form)
    if grid_type == "gauss":
        return create_gauss_config(lmax, nlat_eff, **kw)
    return create_regular_config(lmax, nlat_eff, include_poles=include_poles or pole_grid,
                                 use_dh_weights=grid_type == "driscoll_healy", **kw)
