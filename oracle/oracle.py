"""ctypes front end of the CPU oracle (``shtk_oracle.c``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, the smoke check and the
``cpu_baseline`` leg of ``bench.py``, never by the product package.

Arrays use the reference's index order and Julia's column-major memory layout,
i.e. numpy arrays with ``order='F'``:

* grid      ``f[i, j(, k)]``      shape ``(nlat, nlon[, nf])``, latitude fastest, north first
* spectrum  ``alm[l, m(, k)]``    shape ``(lmax+1, mmax+1[, nf])`` complex128

Reference behaviour restated: see the header of ``shtk_oracle.c``.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

NORMS = {"orthonormal": 0, "fourpi": 1, "schmidt": 2}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        _lib.ora_gauss_legendre.argtypes = [i, p, p]
        _lib.ora_plm_row.argtypes = [p, d, i, i]
        _lib.ora_plm_dtheta_over_sinth_row.argtypes = [p, p, p, p, d, i, i]
        _lib.ora_fft.argtypes = [p, i, i]
        _lib.ora_set_threads.argtypes = [i]
        _lib.ora_max_threads.restype = i
        _lib.ora_coef_scale.argtypes = [p, i, i]
        _lib.ora_coef_scale.restype = d
        _lib.ora_synthesis.argtypes = [p, p, p, i, i, i]
        _lib.ora_analysis.argtypes = [p, p, p, i, i]
        _lib.ora_analysis_cplx.argtypes = [p, p, p, i, i]
        _lib.ora_synthesis_sphtor.argtypes = [p, p, p, p, p, i, i, i]
        _lib.ora_analysis_sphtor.argtypes = [p, p, p, p, p, i, i]
        _lib.ora_adjoint_synthesis.argtypes = [p, p, p, i, i]
        _lib.ora_adjoint_analysis.argtypes = [p, p, p, i]
        _lib.ora_adjoint_synthesis_sphtor.argtypes = [p, p, p, p, p, i, i]
        _lib.ora_adjoint_analysis_sphtor.argtypes = [p, p, p, p, p, i]
    return _lib


class _CCfg(ctypes.Structure):
    _fields_ = [
        ("lmax", ctypes.c_int), ("mmax", ctypes.c_int), ("mres", ctypes.c_int),
        ("nlat", ctypes.c_int), ("nlon", ctypes.c_int), ("norm", ctypes.c_int),
        ("cs_phase", ctypes.c_int), ("real_norm", ctypes.c_int),
        ("robert_form", ctypes.c_int), ("phi_quad", ctypes.c_int),
        ("x", ctypes.c_void_p), ("w", ctypes.c_void_p),
    ]


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def set_threads(n: int) -> None:
    lib().ora_set_threads(int(n))


def max_threads() -> int:
    return int(lib().ora_max_threads())


def gausslegendre(n: int):
    """Ascending nodes/weights (legendre.jl:956-1003)."""
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().ora_gauss_legendre(n, _ptr(x), _ptr(w)) != 0:
        raise ValueError("n must be positive")
    return x, w


def plm_row(x: float, lmax: int, m: int) -> np.ndarray:
    P = np.zeros(lmax + 1)
    lib().ora_plm_row(_ptr(P), float(x), lmax, m)
    return P


def plm_dtheta_over_sinth_row(x: float, lmax: int, m: int):
    P = np.zeros(lmax + 1)
    dP = np.zeros(lmax + 1)
    Ps = np.zeros(lmax + 1)
    Pb = np.zeros(lmax + 2)
    lib().ora_plm_dtheta_over_sinth_row(_ptr(P), _ptr(dP), _ptr(Ps), _ptr(Pb), float(x), lmax, m)
    return P, dP, Ps


def fft(v: np.ndarray, inverse: bool = False) -> np.ndarray:
    a = np.ascontiguousarray(v, dtype=np.complex128).copy()
    lib().ora_fft(_ptr(a), a.size, 1 if inverse else 0)
    return a


def fejer1(n):
    """Fejer's first rule, north first: theta_j = (j + 1/2) pi / n and
    w_j = (2/n) (1 - 2 sum_{k=1}^{n//2} cos(2 k theta_j) / (4 k^2 - 1))
    (_fejer1_weights, src/legendre.jl:1017-1032; create_regular_config
    src/config.jl:1294-1300).  Returns (x, w)."""
    th = (np.arange(n) + 0.5) * (math.pi / n)
    k = np.arange(1, n // 2 + 1)
    series = (np.cos(2.0 * np.outer(th, k)) / (4.0 * k * k - 1.0)).sum(axis=1) if k.size else np.zeros(n)
    return np.cos(th), (2.0 / n) * (1.0 - 2.0 * series)


def clenshaw_curtis(n):
    """Clenshaw-Curtis rule with both poles, north first: theta_j = j pi / N,
    N = n - 1, endpoint weight 1/(N^2 - 1) (N even) or 1/N^2 (N odd), interior
    (2/N) (1 - 2 sum_k cos(2 k theta)/(4k^2 - 1) [- cos(N theta)/(N^2 - 1), N even])
    (_clenshaw_curtis_weights, src/legendre.jl:1044-1075; create_regular_config
    src/config.jl:1285-1292).  Returns (x, w) with x = +-1 exactly at the poles."""
    N = n - 1
    th = np.arange(n) * (math.pi / N)
    x = np.cos(th)
    x[0], x[-1] = 1.0, -1.0
    kmax = N // 2 - 1 if N % 2 == 0 else (N - 1) // 2
    k = np.arange(1, kmax + 1)
    v = 1.0 - 2.0 * (np.cos(2.0 * np.outer(th, k)) / (4.0 * k * k - 1.0)).sum(axis=1)
    if N % 2 == 0:
        v -= np.cos(N * th) / (N * N - 1.0)
    w = 2.0 * v / N
    w[0] = w[-1] = 1.0 / (N * N - 1.0) if N % 2 == 0 else 1.0 / (N * N)
    return x, w


def driscoll_healy(n):
    """Driscoll-Healy nodes theta_j = pi j / n, j < n (north pole included, south
    pole not), north first, and weights (sqrt(8)/n sqrt(2)) sin(pi j/n)
    sum_{l=0}^{n/2-1} sin((2l+1) pi j / n) / (2l+1) (driscoll_healy_weights,
    src/mathutils.jl:149-183; create_regular_config src/config.jl:1258-1273)."""
    j = np.arange(n)
    th = np.pi * j / n
    l = np.arange(n // 2)
    s = (np.sin(np.outer(j, 2 * l + 1) * np.pi / n) / (2 * l + 1)).sum(axis=1)
    return np.cos(th), (math.sqrt(8.0) / n * math.sqrt(2.0)) * np.sin(th) * s


class Config:
    """Grid configuration mirroring ``create_gauss_config`` (src/config.jl:633-655)
    and ``create_regular_config`` (src/config.jl:1235-1330, grids "regular" and
    "regular_poles"): nodes stored north first."""

    def __init__(self, lmax, nlat, mmax=None, mres=1, nlon=None, norm="orthonormal",
                 cs_phase=True, real_norm=False, robert_form=False, phi_scale="dft", grid="gauss"):
        mmax = lmax if mmax is None else mmax
        if nlon is None:
            nlon = 2 * lmax + 1
            nlon += nlon & 1
        min_nlat = {"gauss": lmax + 1, "regular": lmax + 2, "regular_poles": max(2, lmax + 1),
                    "driscoll_healy": 2 * (lmax + 1)}[grid]
        if grid == "driscoll_healy" and nlat != 2 * (lmax + 1):
            raise ValueError("DH weights require nlat = 2 (lmax + 1)")
        if not (0 <= mmax <= lmax) or nlat < min_nlat or nlon < 2 * mmax + 1 or mres < 1:
            raise ValueError("invalid spectral layout")
        self.lmax, self.mmax, self.mres, self.nlat, self.nlon = lmax, mmax, mres, nlat, nlon
        self.norm, self.cs_phase, self.real_norm = norm, cs_phase, real_norm
        self.robert_form, self.phi_scale, self.grid = robert_form, phi_scale, grid
        if grid == "gauss":
            xa, wa = gausslegendre(nlat)
            self.x = np.ascontiguousarray(xa[::-1])
            self.w = np.ascontiguousarray(wa[::-1])
        else:
            xa, wa = {"regular": fejer1, "regular_poles": clenshaw_curtis, "driscoll_healy": driscoll_healy}[grid](nlat)
            self.x = np.ascontiguousarray(xa)
            self.w = np.ascontiguousarray(wa)
        self._c = _CCfg(lmax, mmax, mres, nlat, nlon, NORMS[norm], int(cs_phase), int(real_norm),
                        int(robert_form), 1 if phi_scale == "quad" else 0,
                        _ptr(self.x), _ptr(self.w))

    @property
    def cref(self):
        return ctypes.byref(self._c)

    def coef_scale(self, l, m):
        return lib().ora_coef_scale(self.cref, l, m)

    # --- single order (pure-Python loops over the oracle's Legendre rows) -----
    def _phi_inv_scale(self):
        return self.nlon / (2 * math.pi) if self.phi_scale == "quad" else float(self.nlon)

    def synthesis_ml(self, im, Ql, ltr):
        """synthesis_packed_ml (src/transforms.jl:394-437): Vr_m[i] = phi_inv_scale *
        sum_{l=m}^{ltr} conv(l, m) Q_l P_l^m(x_i)."""
        m = im * self.mres
        Q = np.array([Ql[l - m] * self.coef_scale(l, m) for l in range(m, ltr + 1)])
        out = np.zeros(self.nlat, dtype=np.complex128)
        for i in range(self.nlat):
            P = plm_row(self.x[i], ltr, m)
            out[i] = np.dot(Q, P[m:]) * self._phi_inv_scale()
        return out

    def analysis_ml(self, im, Vm, ltr):
        """analysis_packed_ml (src/transforms.jl:351-385): Q_l = phi_scale sum_i w_i V_i
        P_l^m(x_i), then the convention map."""
        m = im * self.mres
        sphi = (2 * math.pi / self.nlon) * self.nlon / self._phi_inv_scale()
        Q = np.zeros(ltr - m + 1, dtype=np.complex128)
        for i in range(self.nlat):
            P = plm_row(self.x[i], ltr, m)
            Q += self.w[i] * Vm[i] * P[m:]
        return np.array([Q[l - m] * sphi / self.coef_scale(l, m) for l in range(m, ltr + 1)])

    def synthesis_sphtor_ml(self, im, Sl, Tl, ltr):
        """synthesis_sphtor_ml (src/sphtor_transforms.jl:1264-1322): Vt = dS/dtheta -
        i m (P/s) T, Vp = i m (P/s) S + dT/dtheta, times sin(theta) in Robert form."""
        m = im * self.mres
        S = [Sl[l - m] * self.coef_scale(l, m) for l in range(m, ltr + 1)]
        T = [Tl[l - m] * self.coef_scale(l, m) for l in range(m, ltr + 1)]
        Vt = np.zeros(self.nlat, dtype=np.complex128)
        Vp = np.zeros(self.nlat, dtype=np.complex128)
        for i in range(self.nlat):
            x = self.x[i]
            st = math.sqrt(max(0.0, 1 - x * x))
            _, dP, Ps = plm_dtheta_over_sinth_row(x, ltr, m)
            gt = gp = 0j
            for l in range(max(1, m), ltr + 1):
                term = 1j * m * Ps[l]
                gt += dP[l] * S[l - m] - term * T[l - m]
                gp += term * S[l - m] + dP[l] * T[l - m]
            if self.robert_form:
                gt *= st
                gp *= st
            Vt[i] = gt * self._phi_inv_scale()
            Vp[i] = gp * self._phi_inv_scale()
        return Vt, Vp

    def analysis_sphtor_ml(self, im, Vt, Vp, ltr):
        """analysis_sphtor_ml (src/sphtor_transforms.jl:1188-1243)."""
        m = im * self.mres
        sphi = (2 * math.pi / self.nlon) * self.nlon / self._phi_inv_scale()
        S = np.zeros(ltr - m + 1, dtype=np.complex128)
        T = np.zeros(ltr - m + 1, dtype=np.complex128)
        for i in range(self.nlat):
            x = self.x[i]
            st = math.sqrt(max(0.0, 1 - x * x))
            ft, fp = Vt[i], Vp[i]
            if self.robert_form and st > 0:
                ft, fp = ft / st, fp / st
            _, dP, Ps = plm_dtheta_over_sinth_row(x, ltr, m)
            for l in range(max(1, m), ltr + 1):
                c = self.w[i] * sphi / (l * (l + 1))
                term = 1j * m * Ps[l]
                S[l - m] += c * (ft * dP[l] + np.conj(term) * fp)
                T[l - m] += c * (-np.conj(term) * ft + dP[l] * fp)
        cs = np.array([self.coef_scale(l, m) for l in range(m, ltr + 1)])
        return S / cs, T / cs

    # --- shapes ---------------------------------------------------------
    def _nf(self, a, lead):
        return 1 if a.ndim == lead else a.shape[lead]

    def _spec(self, nf):
        return np.zeros((self.lmax + 1, self.mmax + 1, nf), dtype=np.complex128, order="F")

    # --- transforms -------------------------------------------------------
    def synthesis(self, alm, real_output=True, ltr=-1):
        alm = np.asfortranarray(alm, dtype=np.complex128)
        nf = self._nf(alm, 2)
        dt = np.float64 if real_output else np.complex128
        out = np.zeros((self.nlat, self.nlon, nf), dtype=dt, order="F")
        lib().ora_synthesis(self.cref, _ptr(alm), _ptr(out), nf, int(real_output), ltr)
        return out if alm.ndim == 3 else out[:, :, 0]

    def analysis(self, f, ltr=-1):
        f = np.asfortranarray(f, dtype=np.float64)
        nf = self._nf(f, 2)
        alm = self._spec(nf)
        lib().ora_analysis(self.cref, _ptr(f), _ptr(alm), nf, ltr)
        return alm if f.ndim == 3 else alm[:, :, 0]

    def analysis_cplx(self, f, ltr=-1):
        """Analysis of complex fields (bins m >= 0 of the complex FFT)."""
        f = np.asfortranarray(f, dtype=np.complex128)
        nf = self._nf(f, 2)
        alm = self._spec(nf)
        lib().ora_analysis_cplx(self.cref, _ptr(f), _ptr(alm), nf, ltr)
        return alm if f.ndim == 3 else alm[:, :, 0]

    def synthesis_sphtor(self, S, T, real_output=True, ltr=-1):
        S = np.asfortranarray(S, dtype=np.complex128)
        T = np.asfortranarray(T, dtype=np.complex128)
        nf = self._nf(S, 2)
        dt = np.float64 if real_output else np.complex128
        Vt = np.zeros((self.nlat, self.nlon, nf), dtype=dt, order="F")
        Vp = np.zeros((self.nlat, self.nlon, nf), dtype=dt, order="F")
        lib().ora_synthesis_sphtor(self.cref, _ptr(S), _ptr(T), _ptr(Vt), _ptr(Vp), nf,
                                   int(real_output), ltr)
        if S.ndim == 2:
            return Vt[:, :, 0], Vp[:, :, 0]
        return Vt, Vp

    def analysis_sphtor(self, Vt, Vp, ltr=-1):
        Vt = np.asfortranarray(Vt, dtype=np.float64)
        Vp = np.asfortranarray(Vp, dtype=np.float64)
        nf = self._nf(Vt, 2)
        S, T = self._spec(nf), self._spec(nf)
        lib().ora_analysis_sphtor(self.cref, _ptr(Vt), _ptr(Vp), _ptr(S), _ptr(T), nf, ltr)
        if Vt.ndim == 2:
            return S[:, :, 0], T[:, :, 0]
        return S, T

    def adjoint_synthesis(self, fbar, real_output=True):
        fbar = np.asfortranarray(fbar, dtype=np.float64)
        nf = self._nf(fbar, 2)
        out = self._spec(nf)
        lib().ora_adjoint_synthesis(self.cref, _ptr(fbar), _ptr(out), nf, int(real_output))
        return out if fbar.ndim == 3 else out[:, :, 0]

    def adjoint_analysis(self, almbar):
        almbar = np.asfortranarray(almbar, dtype=np.complex128)
        nf = self._nf(almbar, 2)
        out = np.zeros((self.nlat, self.nlon, nf), dtype=np.complex128, order="F")
        lib().ora_adjoint_analysis(self.cref, _ptr(almbar), _ptr(out), nf)
        return out if almbar.ndim == 3 else out[:, :, 0]

    def adjoint_synthesis_sphtor(self, Vtb, Vpb, real_output=True):
        Vtb = np.asfortranarray(Vtb, dtype=np.float64)
        Vpb = np.asfortranarray(Vpb, dtype=np.float64)
        nf = self._nf(Vtb, 2)
        S, T = self._spec(nf), self._spec(nf)
        lib().ora_adjoint_synthesis_sphtor(self.cref, _ptr(Vtb), _ptr(Vpb), _ptr(S), _ptr(T), nf,
                                           int(real_output))
        if Vtb.ndim == 2:
            return S[:, :, 0], T[:, :, 0]
        return S, T

    def adjoint_analysis_sphtor(self, Sb, Tb):
        Sb = np.asfortranarray(Sb, dtype=np.complex128)
        Tb = np.asfortranarray(Tb, dtype=np.complex128)
        nf = self._nf(Sb, 2)
        Vt = np.zeros((self.nlat, self.nlon, nf), dtype=np.complex128, order="F")
        Vp = np.zeros((self.nlat, self.nlon, nf), dtype=np.complex128, order="F")
        lib().ora_adjoint_analysis_sphtor(self.cref, _ptr(Sb), _ptr(Tb), _ptr(Vt), _ptr(Vp), nf)
        if Sb.ndim == 2:
            return Vt[:, :, 0], Vp[:, :, 0]
        return Vt, Vp


# --- packed SHTns LM order (src/layout.jl:143-162, 282-339) -----------------
def lm_index(lmax: int, mres: int, l: int, m: int) -> int:
    im = m // mres
    return (im * (2 * lmax + 2 - (im + 1) * mres)) // 2 + l


def unpack_lm(cfg, q):
    A = np.zeros((cfg.lmax + 1, cfg.mmax + 1), dtype=np.complex128, order="F")
    for m in range(0, cfg.mmax + 1, cfg.mres):
        for l in range(m, cfg.lmax + 1):
            A[l, m] = q[lm_index(cfg.lmax, cfg.mres, l, m)]
    return A


def pack_lm(cfg, A):
    n = sum(cfg.lmax - m + 1 for m in range(0, cfg.mmax + 1, cfg.mres))
    q = np.zeros(n, dtype=np.complex128)
    for m in range(0, cfg.mmax + 1, cfg.mres):
        for l in range(m, cfg.lmax + 1):
            q[lm_index(cfg.lmax, cfg.mres, l, m)] = A[l, m]
    return q


def random_alm(lmax, mmax, seed, power=1.0, nf=None, lmin=0):
    """Complex Gaussian spectrum scaled by (l+1)^-power; l<m and m=0 imaginary
    parts zero (the input recipe of the benchmark configurations)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (lmax + 1, mmax + 1) + (() if nf is None else (nf,))
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    l = np.arange(lmax + 1).reshape((-1, 1) + (() if nf is None else (1,)))
    m = np.arange(mmax + 1).reshape((1, -1) + (() if nf is None else (1,)))
    a = a * (l + 1.0) ** (-power)
    a = np.where(l >= m, a, 0)
    a = np.where(l >= lmin, a, 0)
    a[:, 0, ...] = a[:, 0, ...].real
    return np.asfortranarray(a)


def rel_l2(x, ref) -> float:
    x = np.asarray(x)
    ref = np.asarray(ref)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((x - ref).ravel()) / (den if den > 0 else 1.0))


__all__ = ["Config", "build", "lib", "gausslegendre", "plm_row", "plm_dtheta_over_sinth_row", "fft",
           "lm_index", "unpack_lm", "pack_lm", "random_alm", "rel_l2", "set_threads", "max_threads",
           ]
