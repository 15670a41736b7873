"""Reader of the SHTns 3.7 golden fixtures in tests/golden/shtns37 (payload format
of the reference's test/parity/shtns37_fixtures.jl:10-30) and the comparison rule
of Julia's isapprox on arrays: norm(x - y) <= max(atol, rtol * max(norm(x), norm(y)))."""
import json
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "shtns37")


def manifest():
    with open(os.path.join(DIR, "manifest.json")) as fh:
        return json.load(fh)


def fixture(fid):
    for f in manifest()["fixtures"]:
        if f["id"] == fid:
            return f
    raise KeyError(fid)


def payloads(fx):
    out = {}
    for p in fx["payloads"]:
        raw = open(os.path.join(DIR, p["file"]), "rb").read()
        et = p["eltype"]
        if et == "float64":
            v = np.frombuffer(raw, dtype="<f8")
        elif et == "float32":
            v = np.frombuffer(raw, dtype="<f4")
        elif et == "complex64":
            v = np.frombuffer(raw, dtype="<f8").view(np.complex128)
        elif et == "complex32":
            v = np.frombuffer(raw, dtype="<f4").view(np.complex64)
        else:
            raise ValueError(et)
        out[p["name"]] = v.reshape(tuple(p["shape"]), order="F")
    return out


def isapprox(x, y, atol, rtol):
    x = np.asarray(x, dtype=np.complex128).ravel()
    y = np.asarray(y, dtype=np.complex128).ravel()
    d = np.linalg.norm(x - y)
    return d <= max(atol, rtol * max(np.linalg.norm(x), np.linalg.norm(y))), d


def grid_of(fx):
    """Grid name of a fixture for the configurations here ("gauss_fly" is a Gauss grid)."""
    return {"gauss": "gauss", "gauss_fly": "gauss", "regular": "regular", "regular_poles": "regular_poles"}[fx["grid"]]


def config_kwargs(fx):
    return dict(mmax=fx["mmax"], mres=fx["mres"], nlon=fx["nphi"], norm=fx["norm"],
                cs_phase=fx["cs_phase"], real_norm=fx["real_norm"])


def lm_pairs(lmax, mmax, mres):
    return [(l, m) for m in range(0, mmax + 1, mres) for l in range(m, lmax + 1)]


def unpack(fx, q):
    """Packed LM vector (or [nlm, nf]) -> dense (lmax+1, mmax+1[, nf])."""
    lmax, mmax, mres = fx["lmax"], fx["mmax"], fx["mres"]
    q = np.asarray(q)
    A = np.zeros((lmax + 1, mmax + 1) + q.shape[1:], dtype=np.complex128, order="F")
    for k, (l, m) in enumerate(lm_pairs(lmax, mmax, mres)):
        A[l, m] = q[k]
    return A


def pack(fx, A):
    lmax, mmax, mres = fx["lmax"], fx["mmax"], fx["mres"]
    A = np.asarray(A)
    return np.array([A[l, m] for (l, m) in lm_pairs(lmax, mmax, mres)])


def truncated(fx, q, ltr):
    lmax, mmax, mres = fx["lmax"], fx["mmax"], fx["mres"]
    q = np.array(q, copy=True)
    for k, (l, m) in enumerate(lm_pairs(lmax, mmax, mres)):
        if l > ltr:
            q[k] = 0
    return q
