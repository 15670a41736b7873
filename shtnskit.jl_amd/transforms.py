"""Transforms of SHTnsKit.jl's hot path, executed by libshtk_mi355x.so.

Signatures and semantics mirror the reference:

* ``synthesis(cfg, alm; real_output=True, use_rfft=False)``   src/core_transforms.jl:271-344
* ``analysis(cfg, f; use_rfft=False)``                         src/core_transforms.jl:169-207
* ``synthesis_sphtor(cfg, S, T; real_output=True)``            src/sphtor_transforms.jl:162-241
* ``analysis_sphtor(cfg, Vt, Vp)``                             src/sphtor_transforms.jl:252-289
* ``synthesis_batch / analysis_batch``                         src/batch_transforms.jl:312-655
* ``synthesis_sphtor_batch / analysis_sphtor_batch``           src/batch_transforms.jl:689-815
* ``_adjoint_synthesis / _adjoint_analysis`` and the vector adjoints
                                                               src/core_transforms.jl:477-652,
                                                               src/sphtor_transforms.jl:385-452

Arrays are torch tensors on a ROCm device in the reference's index order and
Julia's column-major memory layout: a grid ``f[i, j(, k)]`` of shape
``(nlat, nlon[, nf])`` is latitude-fastest, a spectrum ``alm[l, m(, k)]`` of shape
``(lmax+1, mmax+1[, nf])`` is degree-fastest.  Inputs in another layout are
copied once into that layout.  Host inputs are rejected unless the call names
the device (``synthesis(GPU(), cfg, alm)``), which moves them to the current
device as SHTnsKit's ``_gpu_adapter_adapt`` does.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import BackendUnavailableError
from .config import SHTConfig


class CPU:
    """Typed device marker (src/devices.jl:7-13)."""


class GPU:
    """Typed device marker (src/devices.jl:15-17); host operands are adapted."""


def on_device(x) -> type:
    return GPU if isinstance(x, torch.Tensor) and x.is_cuda else CPU


# --------------------------------------------------------------------------- helpers
def _split_device(args):
    if args and isinstance(args[0], (GPU, CPU)) or (args and args[0] in (GPU, CPU)):
        dev = args[0]
        if dev is CPU or isinstance(dev, CPU):
            raise BackendUnavailableError("cpu", "this package is the GPU backend; the CPU path is SHTnsKit's own")
        return True, args[1:]
    return False, args


def _adapt(x, adapt: bool, op: str):
    if isinstance(x, np.ndarray):
        x = torch.from_numpy(np.asarray(x))
    if not isinstance(x, torch.Tensor):
        raise TypeError(f"{op}: expected a torch tensor, got {type(x).__name__}")
    if not x.is_cuda:
        if not adapt:
            raise BackendUnavailableError(
                op, "operand is host memory; pass a device tensor or call with GPU() as first argument")
        if not torch.cuda.is_available():
            raise BackendUnavailableError(op, "no ROCm device is available")
        x = x.to("cuda")
    return x


def _colmajor(t: torch.Tensor) -> torch.Tensor:
    """Same logical tensor with Julia column-major memory (first index fastest)."""
    perm = list(range(t.dim() - 1, -1, -1))
    return t.permute(*perm).contiguous().permute(*perm)


def _empty_colmajor(shape, dtype, device) -> torch.Tensor:
    perm = list(range(len(shape) - 1, -1, -1))
    return torch.empty([shape[p] for p in perm], dtype=dtype, device=device).permute(*perm)


def _real_dtype(t: torch.Tensor) -> torch.dtype:
    if t.dtype in (torch.float32, torch.complex64):
        return torch.float32
    if t.dtype in (torch.float64, torch.complex128):
        return torch.float64
    raise TypeError(f"unsupported element type {t.dtype}; use float32/64 or complex64/128")


def _code(rdt: torch.dtype) -> int:
    return _lib.SHTK_F32 if rdt == torch.float32 else _lib.SHTK_F64


def _cplx(rdt: torch.dtype) -> torch.dtype:
    return torch.complex64 if rdt == torch.float32 else torch.complex128


def _stream(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _spec_in(cfg: SHTConfig, alm, adapt: bool, op: str, batch: bool):
    alm = _adapt(alm, adapt, op)
    want = 3 if batch else 2
    if alm.dim() != want:
        raise ValueError(f"{op}: coefficients must have {want} dimensions, got shape {tuple(alm.shape)}")
    if alm.shape[0] != cfg.lmax + 1:
        raise ValueError(f"{op}: first dim must be lmax+1={cfg.lmax + 1}")
    if alm.shape[1] != cfg.mmax + 1:
        raise ValueError(f"{op}: second dim must be mmax+1={cfg.mmax + 1}")
    if batch and alm.shape[2] < 1:
        raise ValueError(f"{op} requires at least one field")
    rdt = _real_dtype(alm)
    alm = _colmajor(alm.to(_cplx(rdt)))
    return alm, rdt


def _grid_in(cfg: SHTConfig, f, adapt: bool, op: str, batch: bool, allow_complex: bool = False):
    f = _adapt(f, adapt, op)
    want = 3 if batch else 2
    if f.dim() != want:
        raise ValueError(f"{op}: field must have {want} dimensions, got shape {tuple(f.shape)}")
    if f.shape[0] != cfg.nlat:
        raise ValueError(f"{op}: first dim must be nlat={cfg.nlat}")
    if f.shape[1] != cfg.nlon:
        raise ValueError(f"{op}: second dim must be nlon={cfg.nlon}")
    if batch and f.shape[2] < 1:
        raise ValueError(f"{op} requires at least one field")
    rdt = _real_dtype(f)
    if f.is_complex():
        if not allow_complex:
            raise ValueError(f"{op}: the fields must be real-valued")
        return _colmajor(f.to(_cplx(rdt))), rdt
    f = _colmajor(f.to(rdt))
    return f, rdt


def _nf(t: torch.Tensor, batch: bool) -> int:
    return int(t.shape[2]) if batch else 1


_PRECISIONS = {"fp64": _lib.SHTK_PREC_FP64, "split": _lib.SHTK_PREC_SPLIT, "fp32": _lib.SHTK_PREC_FP32}


def _precision(op, prec) -> int:
    """Contraction precision keyword of the batch transforms: "fp64" (the plain
    FP64 kernels), "split" (Ozaki split on int8 matrix cores, FP64-grade) or
    "fp32" (FP32 matrix cores, FP32-grade; FP64 Legendre values)."""
    try:
        return _PRECISIONS[str(prec).lstrip(":").lower()]
    except KeyError:
        raise ValueError(f"{op}: precision must be one of {sorted(_PRECISIONS)}, got {prec!r}") from None


def _check_kw(op, real_output=True, fft_scratch=None):
    if fft_scratch is not None and not (isinstance(fft_scratch, torch.Tensor) and fft_scratch.is_cuda):
        raise ValueError(f"{op}: fft_scratch must be device storage")


# --------------------------------------------------------------------------- scalar
def _synthesis(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, alm = args
    real_output = kw.pop("real_output", True)
    kw.pop("use_rfft", None)
    ltr = kw.pop("ltr", -1)
    prec = _precision(op, kw.pop("precision", "fp64"))
    _check_kw(op, real_output, kw.pop("fft_scratch", None))
    if kw:
        raise TypeError(f"{op}: unexpected keywords {sorted(kw)}")
    alm, rdt = _spec_in(cfg, alm, adapt, op, batch)
    nf = _nf(alm, batch)
    shape = (cfg.nlat, cfg.nlon) + ((nf,) if batch else ())
    out = _empty_colmajor(shape, rdt if real_output else _cplx(rdt), alm.device)
    lib = _lib.load()
    _lib.check(lib.shtk_synthesis_ex(ctypes.c_void_p(cfg.plan), _ptr(alm), _ptr(out), nf, _code(rdt),
                                     1 if real_output else 0, int(ltr), prec, _stream(alm.device)), op)
    return out


def _analysis(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, f = args
    use_rfft = kw.pop("use_rfft", False)
    ltr = kw.pop("ltr", -1)
    prec = _precision(op, kw.pop("precision", "fp64"))
    _check_kw(op, True, kw.pop("fft_scratch", None))
    if kw:
        raise TypeError(f"{op}: unexpected keywords {sorted(kw)}")
    f, rdt = _grid_in(cfg, f, adapt, op, batch, allow_complex=True)
    nf = _nf(f, batch)
    shape = (cfg.lmax + 1, cfg.mmax + 1) + ((nf,) if batch else ())
    out = _empty_colmajor(shape, _cplx(rdt), f.device)
    lib = _lib.load()
    if f.is_complex():
        # complex fields: the m-loop over the bins m >= 0 of the complex FFT
        # (src/core_transforms.jl:169-207); use_rfft needs a real input (:181)
        if use_rfft:
            raise ValueError(f"{op}: use_rfft=true requires a real-valued input")
        _lib.check(lib.shtk_analysis_cplx(ctypes.c_void_p(cfg.plan), _ptr(f), _ptr(out), nf, _code(rdt), int(ltr),
                                          prec, _stream(f.device)), op)
        return out
    _lib.check(lib.shtk_analysis_ex(ctypes.c_void_p(cfg.plan), _ptr(f), _ptr(out), nf, _code(rdt), int(ltr), prec,
                                    _stream(f.device)), op)
    return out


def synthesis(*args, **kw):
    """``synthesis(cfg, alm; real_output=True, use_rfft=False)`` -> grid (nlat, nlon)."""
    return _synthesis(args, dict(kw), False, "synthesis")


def analysis(*args, **kw):
    """``analysis(cfg, f; use_rfft=False)`` -> coefficients (lmax+1, mmax+1)."""
    return _analysis(args, dict(kw), False, "analysis")


def synthesis_cplx(cfg, alm):
    """``synthesis_cplx(cfg, alm)`` -> complex grid (nlat, nlon): the bins m >= 0 alone,
    as ``synthesis(cfg, alm; real_output=false)`` (src/core_transforms.jl:289-292)."""
    return _synthesis((cfg, alm), {"real_output": False}, alm.dim() == 3, "synthesis_cplx")


def synthesis_batch(*args, **kw):
    """``synthesis_batch(cfg, alm (lmax+1, mmax+1, nf); real_output=True, precision="fp64")``
    -> (nlat, nlon, nf).  ``precision="split"`` runs the shared-recurrence Ozaki
    contraction on the int8 matrix cores (FP64-grade results)."""
    return _synthesis(args, dict(kw), True, "synthesis_batch")


def analysis_batch(*args, **kw):
    """``analysis_batch(cfg, fields (nlat, nlon, nf); precision="fp64")`` -> (lmax+1, mmax+1, nf).
    ``precision="split"`` contracts on the int8 matrix cores (FP64-grade results)."""
    return _analysis(args, dict(kw), True, "analysis_batch")


def synthesis_l(cfg, alm, ltr: int, **kw):
    """Degree-truncated synthesis (``_synthesis_l``, src/core_transforms.jl:346-368)."""
    return _synthesis((cfg, alm), dict(kw, ltr=ltr), False, "synthesis_l")


def analysis_l(cfg, f, ltr: int, **kw):
    """Degree-truncated analysis: coefficients above ``ltr`` are zero."""
    return _analysis((cfg, f), dict(kw, ltr=ltr), False, "analysis_l")


# --------------------------------------------------------------------------- vector
def _synthesis_sphtor(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, S, T = args
    real_output = kw.pop("real_output", True)
    kw.pop("use_rfft", None)
    ltr = kw.pop("ltr", -1)
    prec = _precision(op, kw.pop("precision", "fp64"))
    if kw:
        raise TypeError(f"{op}: unexpected keywords {sorted(kw)}")
    S, rdt = _spec_in(cfg, S, adapt, op, batch)
    T, rdt2 = _spec_in(cfg, T, adapt, op, batch)
    if rdt2 != rdt:
        rdt = torch.float64
        S, T = S.to(torch.complex128), T.to(torch.complex128)
        S, T = _colmajor(S), _colmajor(T)
    if S.shape != T.shape:
        raise ValueError(f"{op}: Slm and Tlm dims differ")
    nf = _nf(S, batch)
    shape = (cfg.nlat, cfg.nlon) + ((nf,) if batch else ())
    odt = rdt if real_output else _cplx(rdt)  # complex grids of the bins m >= 0 for real_output=False
    Vt = _empty_colmajor(shape, odt, S.device)
    Vp = _empty_colmajor(shape, odt, S.device)
    lib = _lib.load()
    _lib.check(lib.shtk_synthesis_sphtor_ex(ctypes.c_void_p(cfg.plan), _ptr(S), _ptr(T), _ptr(Vt), _ptr(Vp), nf,
                                            _code(rdt), 1 if real_output else 0, int(ltr), prec,
                                            _stream(S.device)), op)
    return Vt, Vp


def _analysis_sphtor(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, Vt, Vp = args
    kw.pop("use_rfft", None)
    ltr = kw.pop("ltr", -1)
    prec = _precision(op, kw.pop("precision", "fp64"))
    if kw:
        raise TypeError(f"{op}: unexpected keywords {sorted(kw)}")
    Vt, rdt = _grid_in(cfg, Vt, adapt, op, batch)
    Vp, rdt2 = _grid_in(cfg, Vp, adapt, op, batch)
    if rdt2 != rdt:
        rdt = torch.float64
        Vt, Vp = _colmajor(Vt.to(rdt)), _colmajor(Vp.to(rdt))
    if Vt.shape != Vp.shape:
        raise ValueError(f"{op}: Vt and Vp dims differ")
    nf = _nf(Vt, batch)
    shape = (cfg.lmax + 1, cfg.mmax + 1) + ((nf,) if batch else ())
    S = _empty_colmajor(shape, _cplx(rdt), Vt.device)
    T = _empty_colmajor(shape, _cplx(rdt), Vt.device)
    lib = _lib.load()
    _lib.check(lib.shtk_analysis_sphtor_ex(ctypes.c_void_p(cfg.plan), _ptr(Vt), _ptr(Vp), _ptr(S), _ptr(T), nf,
                                           _code(rdt), int(ltr), prec, _stream(Vt.device)), op)
    return S, T


def synthesis_sphtor(*args, **kw):
    """``synthesis_sphtor(cfg, Slm, Tlm; real_output=True)`` -> (Vt, Vp)."""
    return _synthesis_sphtor(args, dict(kw), False, "synthesis_sphtor")


def analysis_sphtor(*args, **kw):
    """``analysis_sphtor(cfg, Vt, Vp)`` -> (Slm, Tlm)."""
    return _analysis_sphtor(args, dict(kw), False, "analysis_sphtor")


def synthesis_sphtor_batch(*args, **kw):
    """``synthesis_sphtor_batch(cfg, Slm, Tlm (lmax+1, mmax+1, nf); real_output=True, precision="fp64")``
    -> (Vt, Vp) each (nlat, nlon, nf).  ``precision="split"`` runs the contraction
    on the int8 matrix cores (FP64-grade results)."""
    return _synthesis_sphtor(args, dict(kw), True, "synthesis_sphtor_batch")


def analysis_sphtor_batch(*args, **kw):
    """``analysis_sphtor_batch(cfg, Vt, Vp (nlat, nlon, nf); precision="fp64")`` -> (Slm, Tlm)."""
    return _analysis_sphtor(args, dict(kw), True, "analysis_sphtor_batch")


def synthesis_sphtor_l(cfg, S, T, ltr: int, **kw):
    """Degree-truncated vector synthesis (``synthesis_sphtor_l``, src/sphtor_transforms.jl:956-1005):
    coefficients above ``ltr`` are ignored."""
    return _synthesis_sphtor((cfg, S, T), dict(kw, ltr=ltr), False, "synthesis_sphtor_l")


def analysis_sphtor_l(cfg, Vt, Vp, ltr: int, **kw):
    """Degree-truncated vector analysis (``analysis_sphtor_l``, src/sphtor_transforms.jl:1007-1035):
    coefficients above ``ltr`` are zero."""
    return _analysis_sphtor((cfg, Vt, Vp), dict(kw, ltr=ltr), False, "analysis_sphtor_l")


def synthesis_sph_l(cfg, S, ltr: int, real_output: bool = True):
    """Spheroidal part only (``synthesis_sph_l``, src/sphtor_transforms.jl:1049-1085): T = 0."""
    return synthesis_sphtor_l(cfg, S, torch.zeros_like(S), ltr, real_output=real_output)


def synthesis_tor_l(cfg, T, ltr: int, real_output: bool = True):
    """Toroidal part only (``synthesis_tor_l``, src/sphtor_transforms.jl:1087-1123): S = 0."""
    return synthesis_sphtor_l(cfg, torch.zeros_like(T), T, ltr, real_output=real_output)


# --------------------------------------------------------------------------- single order
def _ml_ltr(cfg: SHTConfig, im: int, ltr, op: str) -> tuple[int, int]:
    """(m, ltr) after the checks of _validate_stored_order (src/sphtor_transforms.jl:1156-1169)."""
    im = int(im)
    if not 0 <= im <= cfg.mmax // cfg.mres:
        raise ValueError(f"{op}: im must satisfy 0 <= im <= {cfg.mmax // cfg.mres}")
    ltr = cfg.lmax if ltr is None or int(ltr) < 0 else int(ltr)
    if ltr > cfg.lmax:
        raise ValueError(f"{op}: ltr must be <= lmax={cfg.lmax}")
    m = im * cfg.mres
    if ltr < m:
        raise ValueError(f"{op}: ltr must satisfy ltr >= im*mres={m}")
    return m, ltr


def _col_in(x, n: int, op: str, what: str):
    """A complex column (n,) or a batch of columns (n, nf) on the device, column-major."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise BackendUnavailableError(op, "inputs must be ROCm device tensors")
    if x.dim() not in (1, 2) or x.shape[0] != n:
        raise ValueError(f"{op}: {what} must have length {n}, got shape {tuple(x.shape)}")
    rdt = _real_dtype(x)
    return _colmajor(x.to(_cplx(rdt))), rdt


def _ml_call(fn, cfg, im, ins, n_in, n_out, ltr, op):
    m, ltr = _ml_ltr(cfg, im, ltr, op)
    cols = [_col_in(x, n_in(m, ltr), op, "column") for x in ins]
    rdt = torch.float32 if all(r == torch.float32 for _, r in cols) else torch.float64
    ts = [_colmajor(t.to(_cplx(rdt))) for t, _ in cols]
    if any(t.shape != ts[0].shape for t in ts):
        raise ValueError(f"{op}: inputs must have equal shapes")
    nf = 1 if ts[0].dim() == 1 else int(ts[0].shape[1])
    shape = (n_out(m, ltr),) + (() if ts[0].dim() == 1 else (nf,))
    outs = [_empty_colmajor(shape, _cplx(rdt), ts[0].device) for _ in ts]
    args = [_ptr(t) for t in ts] + [_ptr(o) for o in outs]
    _lib.check(fn(ctypes.c_void_p(cfg.plan), int(im), *args, nf, _code(rdt), ltr, _stream(ts[0].device)), op)
    return outs


def analysis_packed(*args):
    """``analysis_packed(cfg, Vr_flat)`` (src/transforms.jl:125-136): a flattened real
    grid (length nlat*nlon, latitude fastest as Julia's ``vec``) -> packed Qlm in SHTns
    LM order (length ``cfg.nlm``; ``pack_lm`` skips the orders that are not multiples
    of mres)."""
    from .layout import pack_lm

    adapt, args = _split_device(args)
    cfg, Vr = args
    Vr = _adapt(Vr, adapt, "analysis_packed")
    if Vr.dim() != 1 or Vr.shape[0] != cfg.nspat:
        raise ValueError(f"analysis_packed: Vr must have length {cfg.nspat}")
    if Vr.is_complex():
        raise TypeError("analysis_packed: Vr must be real-valued")
    f = Vr.reshape(cfg.nlon, cfg.nlat).T  # reshape(Vr, nlat, nlon) of column-major storage
    return pack_lm(cfg, analysis(cfg, f))


def synthesis_packed(*args):
    """``synthesis_packed(cfg, Qlm)`` (src/transforms.jl:145-157): packed Qlm (SHTns LM
    order, length ``cfg.nlm``) -> the flattened real grid (length nlat*nlon,
    latitude fastest)."""
    from .layout import unpack_lm

    adapt, args = _split_device(args)
    cfg, Qlm = args
    Qlm = _adapt(Qlm, adapt, "synthesis_packed")
    if Qlm.dim() != 1 or Qlm.shape[0] != cfg.nlm:
        raise ValueError(f"synthesis_packed: Qlm must have length {cfg.nlm}")
    if not Qlm.is_complex():
        raise TypeError("synthesis_packed: Qlm must be complex-valued")
    f = synthesis(cfg, unpack_lm(cfg, Qlm), real_output=True)
    return f.T.reshape(-1)  # vec(f): latitude fastest


def synthesis_packed_ml(cfg: SHTConfig, im: int, Ql, ltr: int):
    """``synthesis_packed_ml(cfg, mval, Ql, ltr)`` (src/transforms.jl:394-437): one stored
    order's coefficients (degrees m..ltr) -> its complex Fourier column (nlat,)."""
    lib = _lib.load()
    return _ml_call(lib.shtk_synthesis_ml, cfg, im, (Ql,), lambda m, l: l - m + 1, lambda m, l: cfg.nlat, ltr,
                    "synthesis_packed_ml")[0]


def analysis_packed_ml(cfg: SHTConfig, im: int, Vr_m, ltr: int):
    """``analysis_packed_ml(cfg, mval, Vr_m, ltr)`` (src/transforms.jl:351-385)."""
    lib = _lib.load()
    return _ml_call(lib.shtk_analysis_ml, cfg, im, (Vr_m,), lambda m, l: cfg.nlat, lambda m, l: l - m + 1, ltr,
                    "analysis_packed_ml")[0]


def synthesis_sphtor_ml(cfg: SHTConfig, im: int, Sl, Tl, ltr: int):
    """``synthesis_sphtor_ml(cfg, im, Sl, Tl, ltr)`` (src/sphtor_transforms.jl:1264-1322)."""
    lib = _lib.load()
    Vt, Vp = _ml_call(lib.shtk_synthesis_sphtor_ml, cfg, im, (Sl, Tl), lambda m, l: l - m + 1,
                      lambda m, l: cfg.nlat, ltr, "synthesis_sphtor_ml")
    return Vt, Vp


def analysis_sphtor_ml(cfg: SHTConfig, im: int, Vt_m, Vp_m, ltr: int):
    """``analysis_sphtor_ml(cfg, im, Vt_m, Vp_m, ltr)`` (src/sphtor_transforms.jl:1188-1243)."""
    lib = _lib.load()
    Sl, Tl = _ml_call(lib.shtk_analysis_sphtor_ml, cfg, im, (Vt_m, Vp_m), lambda m, l: cfg.nlat,
                      lambda m, l: l - m + 1, ltr, "analysis_sphtor_ml")
    return Sl, Tl


def synthesis_sph_ml(cfg: SHTConfig, im: int, Sl, ltr: int):
    """Spheroidal part of one order (``synthesis_sph_ml``, src/sphtor_transforms.jl:1125-1130): T = 0."""
    return synthesis_sphtor_ml(cfg, im, Sl, torch.zeros_like(Sl), ltr)


def synthesis_tor_ml(cfg: SHTConfig, im: int, Tl, ltr: int):
    """Toroidal part of one order (``synthesis_tor_ml``, src/sphtor_transforms.jl:1143-1148): S = 0."""
    return synthesis_sphtor_ml(cfg, im, torch.zeros_like(Tl), Tl, ltr)


def synthesis_grad_ml(cfg: SHTConfig, im: int, Sl, ltr: int):
    """``synthesis_grad_ml`` (src/sphtor_transforms.jl:1362-1364): the spheroidal synthesis."""
    return synthesis_sph_ml(cfg, im, Sl, ltr)


# --------------------------------------------------------------------------- QST
def _qst_promoted(xs):
    """The three components in one precision, the widest of the inputs (``_qst_promoted``
    of src/qst_transforms.jl): FP32 only when every input is single precision."""
    wide = any(isinstance(x, torch.Tensor) and x.dtype in (torch.float64, torch.complex128) or
               not isinstance(x, torch.Tensor) and np.asarray(x).dtype in (np.float64, np.complex128)
               for x in xs)
    if wide:
        return [x.to(torch.complex128 if x.is_complex() else torch.float64) if isinstance(x, torch.Tensor)
                else np.asarray(x).astype(np.complex128 if np.iscomplexobj(x) else np.float64) for x in xs]
    return list(xs)


def _synthesis_qst(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, Q, S, T = args
    Q, S, T = _qst_promoted((Q, S, T))
    dev = (GPU,) if adapt else ()
    Vr = _synthesis(dev + (cfg, Q), dict(kw), batch, op)
    Vt, Vp = _synthesis_sphtor(dev + (cfg, S, T), dict(kw), batch, op)
    return Vr, Vt, Vp


def _analysis_qst(args, kw, batch, op):
    adapt, args = _split_device(args)
    cfg, Vr, Vt, Vp = args
    Vr, Vt, Vp = _qst_promoted((Vr, Vt, Vp))
    dev = (GPU,) if adapt else ()
    Q = _analysis(dev + (cfg, Vr), dict(kw), batch, op)
    S, T = _analysis_sphtor(dev + (cfg, Vt, Vp), dict(kw), batch, op)
    return Q, S, T


def synthesis_qst(*args, **kw):
    """``synthesis_qst(cfg, Qlm, Slm, Tlm; real_output=True)`` -> (Vr, Vt, Vp): the scalar
    synthesis of Q and the sphtor synthesis of (S, T) on one grid, as
    src/qst_transforms.jl:102-128 and its device form :206-217."""
    return _synthesis_qst(args, dict(kw), False, "synthesis_qst")


def analysis_qst(*args, **kw):
    """``analysis_qst(cfg, Vr, Vt, Vp)`` -> (Qlm, Slm, Tlm) (src/qst_transforms.jl:137-160, :219-230)."""
    return _analysis_qst(args, dict(kw), False, "analysis_qst")


def synthesis_qst_batch(*args, **kw):
    """Batched QST synthesis (src/batch_transforms.jl:830-985); ``precision="split"`` applies to
    all three components."""
    return _synthesis_qst(args, dict(kw), True, "synthesis_qst_batch")


def analysis_qst_batch(*args, **kw):
    """Batched QST analysis (src/batch_transforms.jl:830-985)."""
    return _analysis_qst(args, dict(kw), True, "analysis_qst_batch")


def synthesis_qst_l(cfg, Q, S, T, ltr: int, **kw):
    """Degree-truncated QST synthesis (``synthesis_qst_l``, src/qst_transforms.jl:342-384)."""
    return _synthesis_qst((cfg, Q, S, T), dict(kw, ltr=ltr), False, "synthesis_qst_l")


def analysis_qst_l(cfg, Vr, Vt, Vp, ltr: int, **kw):
    """Degree-truncated QST analysis (``analysis_qst_l``, src/qst_transforms.jl:308-340)."""
    return _analysis_qst((cfg, Vr, Vt, Vp), dict(kw, ltr=ltr), False, "analysis_qst_l")


# --------------------------------------------------------------------------- adjoints
def _batchdim(t: torch.Tensor) -> bool:
    return t.dim() == 3


def _adjoint_synthesis(cfg: SHTConfig, fbar, real_output: bool = True):
    """Adjoint of ``synthesis`` (src/core_transforms.jl:477-519; Zygote/torch convention
    ``dL/dRe + i dL/dIm``).  A complex cotangent is taken as its real and imaginary
    parts: the adjoint is complex-linear in ``fbar``."""
    if isinstance(fbar, torch.Tensor) and fbar.is_complex():
        re = _adjoint_synthesis(cfg, fbar.real.contiguous(), real_output)
        im = _adjoint_synthesis(cfg, fbar.imag.contiguous(), real_output)
        return re + 1j * im
    batch = _batchdim(fbar)
    fbar, rdt = _grid_in(cfg, fbar, False, "_adjoint_synthesis", batch)
    nf = _nf(fbar, batch)
    shape = (cfg.lmax + 1, cfg.mmax + 1) + ((nf,) if batch else ())
    out = _empty_colmajor(shape, _cplx(rdt), fbar.device)
    _lib.check(_lib.load().shtk_adjoint_synthesis(ctypes.c_void_p(cfg.plan), _ptr(fbar), _ptr(out), nf, _code(rdt),
                                                  1 if real_output else 0, _stream(fbar.device)),
               "_adjoint_synthesis")
    return out


def _adjoint_analysis(cfg: SHTConfig, almbar, project_real: bool = False):
    """Adjoint of ``analysis`` (src/core_transforms.jl:611-652): the complex grid of the
    inverse FFT over the bins m >= 0.  ``project_real=True`` returns its real part, what
    ``ProjectTo`` keeps for a real primal (ext/SHTnsKitAdvancedADExt.jl:55-67)."""
    batch = _batchdim(almbar)
    almbar, rdt = _spec_in(cfg, almbar, False, "_adjoint_analysis", batch)
    nf = _nf(almbar, batch)
    shape = (cfg.nlat, cfg.nlon) + ((nf,) if batch else ())
    out = _empty_colmajor(shape, rdt if project_real else _cplx(rdt), almbar.device)
    fn = _lib.load().shtk_adjoint_analysis if project_real else _lib.load().shtk_adjoint_analysis_cplx
    _lib.check(fn(ctypes.c_void_p(cfg.plan), _ptr(almbar), _ptr(out), nf, _code(rdt), _stream(almbar.device)),
               "_adjoint_analysis")
    return out


def _adjoint_synthesis_sphtor(cfg: SHTConfig, Vtb, Vpb, real_output: bool = True):
    """Adjoint of ``synthesis_sphtor`` (src/core_transforms.jl:537-597); complex-linear
    in the cotangents, so complex ones are taken as real and imaginary parts."""
    if any(isinstance(v, torch.Tensor) and v.is_complex() for v in (Vtb, Vpb)):
        Vtb, Vpb = torch.as_tensor(Vtb), torch.as_tensor(Vpb)
        sr, tr = _adjoint_synthesis_sphtor(cfg, Vtb.real.contiguous(), Vpb.real.contiguous(), real_output)
        zt = Vtb.imag if Vtb.is_complex() else torch.zeros_like(Vtb)
        zp = Vpb.imag if Vpb.is_complex() else torch.zeros_like(Vpb)
        si, ti = _adjoint_synthesis_sphtor(cfg, zt.contiguous(), zp.contiguous(), real_output)
        return sr + 1j * si, tr + 1j * ti
    batch = _batchdim(Vtb)
    Vtb, rdt = _grid_in(cfg, Vtb, False, "_adjoint_synthesis_sphtor", batch)
    Vpb, _ = _grid_in(cfg, Vpb.to(Vtb.dtype), False, "_adjoint_synthesis_sphtor", batch)
    if Vtb.shape != Vpb.shape:
        raise ValueError("_adjoint_synthesis_sphtor: Vt and Vp cotangents must have equal shapes")
    nf = _nf(Vtb, batch)
    shape = (cfg.lmax + 1, cfg.mmax + 1) + ((nf,) if batch else ())
    S = _empty_colmajor(shape, _cplx(rdt), Vtb.device)
    T = _empty_colmajor(shape, _cplx(rdt), Vtb.device)
    _lib.check(_lib.load().shtk_adjoint_synthesis_sphtor(ctypes.c_void_p(cfg.plan), _ptr(Vtb), _ptr(Vpb), _ptr(S),
                                                         _ptr(T), nf, _code(rdt), 1 if real_output else 0,
                                                         _stream(Vtb.device)), "_adjoint_synthesis_sphtor")
    return S, T


def _adjoint_analysis_sphtor(cfg: SHTConfig, Sb, Tb, project_real: bool = False):
    """Adjoint of ``analysis_sphtor`` (src/sphtor_transforms.jl:385-452): complex grids;
    ``project_real=True`` returns their real parts (a real primal's projection)."""
    batch = _batchdim(Sb)
    Sb, rdt = _spec_in(cfg, Sb, False, "_adjoint_analysis_sphtor", batch)
    Tb, _ = _spec_in(cfg, Tb.to(Sb.dtype), False, "_adjoint_analysis_sphtor", batch)
    if Sb.shape != Tb.shape:
        raise ValueError("_adjoint_analysis_sphtor: S and T cotangents must have equal shapes")
    nf = _nf(Sb, batch)
    shape = (cfg.nlat, cfg.nlon) + ((nf,) if batch else ())
    odt = rdt if project_real else _cplx(rdt)
    Vt = _empty_colmajor(shape, odt, Sb.device)
    Vp = _empty_colmajor(shape, odt, Sb.device)
    lib = _lib.load()
    fn = lib.shtk_adjoint_analysis_sphtor if project_real else lib.shtk_adjoint_analysis_sphtor_cplx
    _lib.check(fn(ctypes.c_void_p(cfg.plan), _ptr(Sb), _ptr(Tb), _ptr(Vt), _ptr(Vp), nf, _code(rdt),
                  _stream(Sb.device)), "_adjoint_analysis_sphtor")
    return Vt, Vp


def synthesis_qst_ml(cfg: SHTConfig, im: int, Ql, Sl, Tl, ltr: int):
    """``synthesis_qst_ml(cfg, mval, Ql, Sl, Tl, ltr)`` (src/qst_transforms.jl:414-424):
    the scalar and vector single-order syntheses in one promoted precision."""
    Ql, Sl, Tl = _qst_promoted((Ql, Sl, Tl))
    return (synthesis_packed_ml(cfg, im, Ql, ltr),) + tuple(synthesis_sphtor_ml(cfg, im, Sl, Tl, ltr))


def analysis_qst_ml(cfg: SHTConfig, im: int, Vr_m, Vt_m, Vp_m, ltr: int):
    """``analysis_qst_ml(cfg, mval, Vr_m, Vt_m, Vp_m, ltr)`` (src/qst_transforms.jl:386-397)."""
    Vr_m, Vt_m, Vp_m = _qst_promoted((Vr_m, Vt_m, Vp_m))
    return (analysis_packed_ml(cfg, im, Vr_m, ltr),) + tuple(analysis_sphtor_ml(cfg, im, Vt_m, Vp_m, ltr))
