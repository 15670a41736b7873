"""ctypes binding of libshtk_mi355x.so (the C ABI declared in include/shtk.h).

The library is loaded from the package directory.  There is no fallback: if the
library is missing or cannot be loaded, every transform raises
BackendUnavailableError, as SHTnsKit's strict routing does when a GPU backend
cannot run (src/devices.jl:19-26, src/device_utils.jl:106-132).
"""
from __future__ import annotations

import ctypes
import os

from ._build import LIB_PATH

SHTK_OK, SHTK_EINVAL, SHTK_EUNSUPPORTED, SHTK_ENOMEM, SHTK_EHIP = 0, 1, 2, 3, 4
SHTK_F64, SHTK_F32 = 0, 1
SHTK_PREC_FP64, SHTK_PREC_SPLIT, SHTK_PREC_FP32 = 0, 1, 2

# Every symbol include/shtk.h declares.
EXPORTS = (
    "shtk_plan_create", "shtk_plan_destroy", "shtk_plan_nodes", "shtk_workspace_size",
    "shtk_synthesis", "shtk_analysis", "shtk_synthesis_ex", "shtk_analysis_ex", "shtk_analysis_cplx",
    "shtk_synthesis_static", "shtk_synthesis_sphtor", "shtk_analysis_sphtor",
    "shtk_synthesis_sphtor_ex", "shtk_analysis_sphtor_ex",
    "shtk_synthesis_ml", "shtk_analysis_ml", "shtk_synthesis_sphtor_ml", "shtk_analysis_sphtor_ml",
    "shtk_adjoint_synthesis", "shtk_adjoint_analysis", "shtk_adjoint_synthesis_sphtor",
    "shtk_adjoint_analysis_sphtor", "shtk_adjoint_analysis_cplx", "shtk_adjoint_analysis_sphtor_cplx",
    "shtk_batch_create", "shtk_batch_destroy", "shtk_batch_plan",
    "shtk_batch_set_graph", "shtk_batch_set_tiles", "shtk_batch_synthesis", "shtk_batch_analysis",
    "shtk_batch_autotune", "shtk_profile_enable", "shtk_profile_read", "shtk_stream_sync",
    "shtk_last_error", "shtk_version",
)


class BackendUnavailableError(RuntimeError):
    """Mirror of SHTnsKit.BackendUnavailableError(operation, detail)."""

    def __init__(self, operation: str, detail: str):
        super().__init__(f"GPU backend unavailable for `{operation}`: {detail}")
        self.operation = operation
        self.detail = detail


class ShtkDesc(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "lmax", "mmax", "mres", "nlat", "nlon", "norm", "cs_phase", "real_norm", "robert_form",
        "phi_scale", "grid")]


_lib = None


def load():
    """Loads the library once and declares the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BackendUnavailableError(
            "load", f"{os.path.basename(LIB_PATH)} is not built; run __graft_entry__.build()")
    try:
        lib = ctypes.CDLL(LIB_PATH)
    except OSError as err:  # pragma: no cover - depends on the machine
        raise BackendUnavailableError("load", str(err)) from err
    i, p, sz = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
    lib.shtk_plan_create.argtypes = [ctypes.POINTER(ShtkDesc), ctypes.POINTER(p)]
    lib.shtk_plan_destroy.argtypes = [p]
    lib.shtk_plan_nodes.argtypes = [p, p, p]
    lib.shtk_workspace_size.argtypes = [p, i, i]
    lib.shtk_workspace_size.restype = sz
    lib.shtk_synthesis.argtypes = [p, p, p, i, i, i, i, p]
    lib.shtk_analysis.argtypes = [p, p, p, i, i, i, p]
    lib.shtk_synthesis_ex.argtypes = [p, p, p, i, i, i, i, i, p]
    lib.shtk_analysis_ex.argtypes = [p, p, p, i, i, i, i, p]
    lib.shtk_analysis_cplx.argtypes = [p, p, p, i, i, i, i, p]
    lib.shtk_synthesis_static.argtypes = [p, p, p, i, i, i, i, p]
    lib.shtk_synthesis_sphtor.argtypes = [p, p, p, p, p, i, i, i, i, p]
    lib.shtk_analysis_sphtor.argtypes = [p, p, p, p, p, i, i, i, p]
    lib.shtk_synthesis_sphtor_ex.argtypes = [p, p, p, p, p, i, i, i, i, i, p]
    lib.shtk_analysis_sphtor_ex.argtypes = [p, p, p, p, p, i, i, i, i, p]
    lib.shtk_synthesis_ml.argtypes = [p, i, p, p, i, i, i, p]
    lib.shtk_analysis_ml.argtypes = [p, i, p, p, i, i, i, p]
    lib.shtk_synthesis_sphtor_ml.argtypes = [p, i, p, p, p, p, i, i, i, p]
    lib.shtk_analysis_sphtor_ml.argtypes = [p, i, p, p, p, p, i, i, i, p]
    lib.shtk_adjoint_synthesis.argtypes = [p, p, p, i, i, i, p]
    lib.shtk_adjoint_analysis.argtypes = [p, p, p, i, i, p]
    lib.shtk_adjoint_synthesis_sphtor.argtypes = [p, p, p, p, p, i, i, i, p]
    lib.shtk_adjoint_analysis_sphtor.argtypes = [p, p, p, p, p, i, i, p]
    lib.shtk_adjoint_analysis_cplx.argtypes = [p, p, p, i, i, p]
    lib.shtk_adjoint_analysis_sphtor_cplx.argtypes = [p, p, p, p, p, i, i, p]
    lib.shtk_batch_create.argtypes = [ctypes.POINTER(ShtkDesc), p, i, i, ctypes.POINTER(p)]
    lib.shtk_batch_destroy.argtypes = [p]
    lib.shtk_batch_plan.argtypes = [p, i, ctypes.POINTER(p)]
    lib.shtk_batch_set_graph.argtypes = [p, i]
    lib.shtk_batch_set_tiles.argtypes = [p, i, i]
    lib.shtk_batch_synthesis.argtypes = [p, p, p, i, p]
    lib.shtk_batch_analysis.argtypes = [p, p, p, i, p]
    lib.shtk_batch_autotune.argtypes = [p, p, p, i, i, p, p, p, p]
    lib.shtk_profile_enable.argtypes = [p, i]
    lib.shtk_profile_read.argtypes = [p, i, p, p, p, p]
    lib.shtk_stream_sync.argtypes = [p]
    lib.shtk_last_error.argtypes = [ctypes.c_char_p, sz]
    lib.shtk_version.restype = ctypes.c_char_p
    _lib = lib
    return lib


def last_error() -> tuple[int, str]:
    buf = ctypes.create_string_buffer(512)
    code = load().shtk_last_error(buf, 512)
    return code, buf.value.decode(errors="replace")


def check(rc: int, operation: str) -> None:
    """Maps a C status to the reference's exception types."""
    if rc == SHTK_OK:
        return
    _, msg = last_error()
    if rc == SHTK_EINVAL:
        raise ValueError(f"{operation}: {msg}")
    if rc == SHTK_EUNSUPPORTED:
        raise BackendUnavailableError(operation, msg)
    if rc == SHTK_ENOMEM:
        raise MemoryError(f"{operation}: {msg}")
    raise RuntimeError(f"{operation}: HIP error: {msg}")


def profile_enable(plan: int, on: bool) -> None:
    check(load().shtk_profile_enable(ctypes.c_void_p(plan), 1 if on else 0), "profile_enable")


def profile_read(plan: int) -> dict:
    """{kernel name: (total_ms, launches)} since the last read (waits for the events)."""
    n = 16
    names = ctypes.create_string_buffer(48 * n)
    ms = (ctypes.c_double * n)()
    cnt = (ctypes.c_int * n)()
    nout = ctypes.c_int()
    check(load().shtk_profile_read(ctypes.c_void_p(plan), n, names, ms, cnt, ctypes.byref(nout)), "profile_read")
    out = {}
    for k in range(nout.value):
        nm = names.raw[48 * k:48 * (k + 1)].split(b"\0", 1)[0].decode()
        out[nm] = (ms[k], cnt[k])
    return out
