#!/usr/bin/env python3
"""Generates the high-precision golden vectors in tests/golden/ (run here, once;
the outputs are committed and the GPU box never runs this script).

* plm_mpmath.npz    orthonormal + Condon-Shortley rows P_l^m(x) computed with the
                    plain three-term recurrence in 256-bit mpmath arithmetic --
                    the BigFloat reference construction of the reference's own
                    test (test/serial/test_legendre.jl:462-478), at the points that
                    test uses (:505-526) plus lmax=1023 rows at Gauss nodes nearest
                    the pole and the equator, and zonal theta-derivatives from the
                    identity dP_l^0/dtheta = sqrt(l(l+1)) P_l^1 (:480-503).
* gauss_mpmath.npz  Gauss-Legendre nodes/weights for n in (6, 8, 128, 1024),
                    Newton-refined in 60-digit arithmetic.
"""
from __future__ import annotations

import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def plm_row_mp(lmax: int, m: int, x) -> list:
    x = mp.mpf(x)
    s = mp.sqrt(1 - x * x)
    pmm = 1 / mp.sqrt(4 * mp.pi)
    for k in range(1, m + 1):
        pmm = -mp.sqrt(mp.mpf(2 * k + 1) / (2 * k)) * s * pmm
    P = [mp.mpf(0)] * (lmax + 1)
    P[m] = pmm
    if m < lmax:
        P[m + 1] = mp.sqrt(2 * m + 3) * x * pmm
    for l in range(m + 2, lmax + 1):
        a = mp.sqrt(mp.mpf((2 * l - 1) * (2 * l + 1)) / ((l - m) * (l + m)))
        b = mp.sqrt(mp.mpf((2 * l + 1) * (l - 1 - m) * (l - 1 + m)) / ((2 * l - 3) * (l - m) * (l + m)))
        P[l] = a * x * P[l - 1] - b * P[l - 2]
    return P


def gauss_mp(n: int):
    x0, _ = np.polynomial.legendre.leggauss(n)
    xs, ws = [], []
    for z0 in x0:
        z = mp.mpf(z0)
        for _ in range(3):
            p0, p1 = mp.mpf(1), z
            for l in range(2, n + 1):
                p0, p1 = p1, ((2 * l - 1) * z * p1 - (l - 1) * p0) / l
            dp = n * (z * p1 - p0) / (z * z - 1)
            z = z - p1 / dp
        p0, p1 = mp.mpf(1), z
        for l in range(2, n + 1):
            p0, p1 = p1, ((2 * l - 1) * z * p1 - (l - 1) * p0) / l
        dp = n * (z * p1 - p0) / (z * z - 1)
        xs.append(float(z))
        ws.append(float(2 / ((1 - z * z) * dp * dp)))
    return np.array(xs), np.array(ws)


def main():
    mp.mp.prec = 256
    out = {}
    pts = [(2047, 1190, 0.5), (2500, 1500, 0.6), (3000, 1743, 0.5), (4000, 3000, 0.05)]
    for k, (lmax, m, s) in enumerate(pts):
        x = mp.sqrt(1 - mp.mpf(s) ** 2)
        out[f"row{k}"] = np.array([float(v) for v in plm_row_mp(lmax, m, x)])
        out[f"row{k}_meta"] = np.array([lmax, m, s, float(x)])
    # lmax=1023 rows at Gauss nodes of nlat=1024 nearest the pole and the equator
    mp.mp.prec = 128
    xg, wg = gauss_mp(1024)
    xg = xg[::-1]  # north first
    k = len(pts)
    for m in (0, 1, 300, 1023):
        for i in (0, 3, 511):
            out[f"row{k}"] = np.array([float(v) for v in plm_row_mp(1023, m, xg[i])])
            out[f"row{k}_meta"] = np.array([1023, m, float(np.sqrt((1 - xg[i]) * (1 + xg[i]))), xg[i]])
            k += 1
    out["nrows"] = np.array([k])
    # zonal derivative: dP_l^0/dtheta = sqrt(l(l+1)) P_l^1
    mp.mp.prec = 256
    for j, x in enumerate((float(np.nextafter(1.0, 0.0)), -float(np.nextafter(1.0, 0.0)), 0.25, xg[0])):
        row1 = plm_row_mp(32, 1, x)
        out[f"zonal{j}"] = np.array([float(mp.sqrt(l * (l + 1)) * row1[l]) for l in range(33)])
        out[f"zonal{j}_x"] = np.array([x])
    np.savez(os.path.join(HERE, "plm_mpmath.npz"), **out)

    mp.mp.prec = 200
    g = {}
    for n in (6, 8, 128, 1024):
        xs, ws = gauss_mp(n)
        g[f"x{n}"] = xs
        g[f"w{n}"] = ws
    np.savez(os.path.join(HERE, "gauss_mpmath.npz"), **g)
    print("wrote plm_mpmath.npz, gauss_mpmath.npz")


if __name__ == "__main__":
    main()
