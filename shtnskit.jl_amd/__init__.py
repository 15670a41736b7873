"""shtnskit.jl_amd -- MI355X (gfx950) backend for SHTnsKit.jl's spherical
harmonic transforms.

The compute path is libshtk_mi355x.so (hand-written HIP kernels behind the C ABI
of include/shtk.h).  This Python layer mirrors the reference's operator API so
host code (and the parity tests) read like SHTnsKit's; the Julia adapter that
binds the same C ABI lives in ``julia/``.
"""
from ._lib import BackendUnavailableError, EXPORTS
from ._lib import load as load_library
from .config import (SHTConfig, create_config, create_gauss_config, create_gauss_fly_config,
                     create_regular_config, nlm_calc)
from .layout import LM_index, pack_lm, unpack_lm
from .transforms import (CPU, GPU, on_device, synthesis, analysis, synthesis_batch, analysis_batch, synthesis_cplx,
                         synthesis_l, analysis_l, synthesis_sphtor, analysis_sphtor,
                         synthesis_sphtor_batch, analysis_sphtor_batch, synthesis_sphtor_l, analysis_sphtor_l,
                         synthesis_sph_l, synthesis_tor_l, synthesis_qst, analysis_qst, synthesis_qst_batch,
                         analysis_qst_batch, synthesis_qst_l, analysis_qst_l, synthesis_packed, analysis_packed,
                         synthesis_packed_ml,
                         analysis_packed_ml, synthesis_sphtor_ml, analysis_sphtor_ml, synthesis_sph_ml,
                         synthesis_tor_ml, synthesis_grad_ml, synthesis_qst_ml, analysis_qst_ml, _adjoint_synthesis,
                         _adjoint_analysis, _adjoint_synthesis_sphtor, _adjoint_analysis_sphtor)
from . import ad
from .batch import UnevenBatch

__version__ = "0.1.0"
