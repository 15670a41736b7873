"""Uneven batches: many fields of several Gauss configurations (different lmax)
transformed together through one batch object of libshtk_mi355x.so.

The reference handles such a set one configuration (or one field) at a time,
``analysis_batch``/``synthesis_batch`` per configuration
(src/batch_transforms.jl:312-655).  Here every stage is one launch for all
configurations; with ``precision="split"`` the Legendre contractions run as
persistent work-queue kernels on the int8 matrix cores, and with
``graph=True`` each call's launch sequence is captured once in a hipGraph and
replayed with a single launch (the stream must not be the default stream).
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import torch

from . import _lib
from .config import SHTConfig
from .transforms import _PRECISIONS, _code, _colmajor, _empty_colmajor, _ptr, _stream

_SPEC_DTYPES = (torch.complex64, torch.complex128)
_GRID_DTYPES = (torch.float32, torch.float64)


def check_group_arrays(cfgs, nfields, xs, kind, op):
    """Validates the per-group arrays of an uneven batch call and returns them
    column-major: one array per group of shape (lmax+1, mmax+1, nf) (kind
    "spec", complex64/complex128) or (nlat, nlon, nf) (kind "grid",
    float32/float64), every group of the same precision and on one device.
    The batch kernels read every group with the one dtype they are given, so a
    mismatch is an error, never a conversion."""
    xs = list(xs)
    if len(xs) != len(cfgs):
        raise ValueError(f"{op}: expected {len(cfgs)} arrays, got {len(xs)}")
    allowed = _SPEC_DTYPES if kind == "spec" else _GRID_DTYPES
    out = []
    for g, (c, nf, x) in enumerate(zip(cfgs, nfields, xs)):
        want = ((c.lmax + 1, c.mmax + 1) if kind == "spec" else (c.nlat, c.nlon)) + (nf,)
        if tuple(x.shape) != want:
            raise ValueError(f"{op}: array of shape {tuple(x.shape)} where {want} is expected")
        if x.dtype not in allowed:
            what = "complex64 or complex128 coefficients" if kind == "spec" else "float32 or float64 grids"
            raise ValueError(f"{op}: group {g} has dtype {x.dtype}; the batch takes {what}")
        if x.dtype != xs[0].dtype:
            raise ValueError(f"{op}: group {g} has dtype {x.dtype}, group 0 {xs[0].dtype}; "
                             "every group must have the same precision")
    for x in xs:
        if not x.is_cuda:
            raise ValueError(f"{op}: arrays must be device tensors")
        if x.device != xs[0].device:
            raise ValueError(f"{op}: every array must be on one device")
        out.append(_colmajor(x))
    return out


class UnevenBatch:
    """``UnevenBatch(cfgs, nfields, precision="split", graph=False)``: group g holds
    ``nfields[g]`` fields of configuration ``cfgs[g]`` (at most 8 groups)."""

    def __init__(self, cfgs: Sequence[SHTConfig], nfields: Sequence[int], precision: str = "split",
                 graph: bool = False):
        if len(cfgs) != len(nfields):
            raise ValueError("cfgs and nfields must have the same length")
        self.cfgs = list(cfgs)
        self.nfields = [int(n) for n in nfields]
        prec = _PRECISIONS.get(str(precision).lstrip(":").lower())
        if prec is None:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}, got {precision!r}")
        n = len(self.cfgs)
        descs = (_lib.ShtkDesc * n)(*[c._desc for c in self.cfgs])
        nfs = (ctypes.c_int * n)(*self.nfields)
        handle = ctypes.c_void_p()
        lib = _lib.load()
        _lib.check(lib.shtk_batch_create(descs, nfs, n, prec, ctypes.byref(handle)), "UnevenBatch")
        self._b = handle.value
        if graph:
            _lib.check(lib.shtk_batch_set_graph(ctypes.c_void_p(self._b), 1), "UnevenBatch")

    def _arrays(self, xs, kind, op):
        return check_group_arrays(self.cfgs, self.nfields, xs, kind, op)

    def synthesis(self, alms: Sequence[torch.Tensor]) -> list:
        """Coefficients (lmax+1, mmax+1, nf) per group -> grids (nlat, nlon, nf)."""
        alms = self._arrays(alms, "spec", "synthesis")
        rdt = torch.float32 if alms[0].dtype == torch.complex64 else torch.float64
        outs = [_empty_colmajor((c.nlat, c.nlon, nf), rdt, alms[0].device) for c, nf in zip(self.cfgs, self.nfields)]
        ins = (ctypes.c_void_p * len(alms))(*[_ptr(a).value for a in alms])
        outp = (ctypes.c_void_p * len(outs))(*[_ptr(o).value for o in outs])
        _lib.check(_lib.load().shtk_batch_synthesis(ctypes.c_void_p(self._b), ins, outp, _code(rdt),
                                                    _stream(alms[0].device)), "batch synthesis")
        return outs

    def analysis(self, fields: Sequence[torch.Tensor]) -> list:
        """Grids (nlat, nlon, nf) per group -> coefficients (lmax+1, mmax+1, nf)."""
        fields = self._arrays(fields, "grid", "analysis")
        rdt = fields[0].dtype
        cdt = torch.complex64 if rdt == torch.float32 else torch.complex128
        outs = [_empty_colmajor((c.lmax + 1, c.mmax + 1, nf), cdt, fields[0].device)
                for c, nf in zip(self.cfgs, self.nfields)]
        ins = (ctypes.c_void_p * len(fields))(*[_ptr(a).value for a in fields])
        outp = (ctypes.c_void_p * len(outs))(*[_ptr(o).value for o in outs])
        _lib.check(_lib.load().shtk_batch_analysis(ctypes.c_void_p(self._b), ins, outp, _code(rdt),
                                                   _stream(fields[0].device)), "batch analysis")
        return outs

    def set_tiles(self, synth_blocks: int = 0, anal_blocks: int = 0) -> None:
        """Grid sizes of the persistent synthesis / analysis kernels (0: two blocks per CU)."""
        _lib.check(_lib.load().shtk_batch_set_tiles(ctypes.c_void_p(self._b), int(synth_blocks),
                                                    int(anal_blocks)), "set_tiles")

    def autotune(self, alms: Sequence[torch.Tensor], fields: Sequence[torch.Tensor], reps: int = 3) -> dict:
        """Times every candidate launch shape on these arrays (round trips that
        overwrite them) and keeps the fastest; returns the choice."""
        alms = self._arrays(alms, "spec", "autotune")
        fields = self._arrays(fields, "grid", "autotune")
        if (alms[0].dtype == torch.complex64) != (fields[0].dtype == torch.float32):
            raise ValueError("autotune: coefficients and grids must have the same precision")
        rdt = fields[0].dtype
        ins = (ctypes.c_void_p * len(alms))(*[_ptr(a).value for a in alms])
        outs = (ctypes.c_void_p * len(fields))(*[_ptr(f).value for f in fields])
        nb, fp, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _lib.check(_lib.load().shtk_batch_autotune(ctypes.c_void_p(self._b), ins, outs, _code(rdt), int(reps),
                                                   _stream(alms[0].device), ctypes.byref(nb), ctypes.byref(fp),
                                                   ctypes.byref(ms)), "autotune")
        return {"synth_blocks": nb.value, "anal_blocks": fp.value, "round_trip_ms": ms.value}

    def close(self) -> None:
        if getattr(self, "_b", None):
            _lib.load().shtk_batch_destroy(ctypes.c_void_p(self._b))
            self._b = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order varies
        try:
            self.close()
        except Exception:
            pass
