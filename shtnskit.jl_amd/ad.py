"""Reverse-mode rules for the transforms, the torch counterpart of SHTnsKit's
ChainRulesCore rrules (ext/SHTnsKitAdvancedADExt.jl:55-241).

Each pullback calls the hand-written adjoint entry point of libshtk_mi355x.so.
The complex-gradient convention is the reference's Zygote one,
``g = dL/dRe + i dL/dIm``, which is also torch's, so gradients agree one to one.
"""
from __future__ import annotations

import torch

from . import transforms as T


class _Synthesis(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, alm, real_output):
        ctx.cfg, ctx.real_output = cfg, real_output
        fn = T.synthesis_batch if alm.dim() == 3 else T.synthesis
        return fn(cfg, alm, real_output=real_output)

    @staticmethod
    def backward(ctx, ybar):
        g = T._adjoint_synthesis(ctx.cfg, ybar, real_output=ctx.real_output)
        return None, g, None


class _Analysis(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, f):
        ctx.cfg = cfg
        fn = T.analysis_batch if f.dim() == 3 else T.analysis
        return fn(cfg, f)

    @staticmethod
    def backward(ctx, abar):
        return None, T._adjoint_analysis(ctx.cfg, abar, project_real=True)


class _SynthesisSphtor(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, S, Tlm, real_output):
        ctx.cfg, ctx.real_output = cfg, real_output
        fn = T.synthesis_sphtor_batch if S.dim() == 3 else T.synthesis_sphtor
        return fn(cfg, S, Tlm, real_output=real_output)

    @staticmethod
    def backward(ctx, vtb, vpb):
        if vtb is None:
            vtb = torch.zeros_like(vpb)
        if vpb is None:
            vpb = torch.zeros_like(vtb)
        sb, tb = T._adjoint_synthesis_sphtor(ctx.cfg, vtb, vpb, real_output=ctx.real_output)
        return None, sb, tb, None


class _AnalysisSphtor(torch.autograd.Function):
    @staticmethod
    def forward(ctx, cfg, Vt, Vp):
        ctx.cfg = cfg
        fn = T.analysis_sphtor_batch if Vt.dim() == 3 else T.analysis_sphtor
        return fn(cfg, Vt, Vp)

    @staticmethod
    def backward(ctx, sb, tb):
        if sb is None:
            sb = torch.zeros_like(tb)
        if tb is None:
            tb = torch.zeros_like(sb)
        vt, vp = T._adjoint_analysis_sphtor(ctx.cfg, sb, tb, project_real=True)
        return None, vt, vp


def synthesis(cfg, alm, real_output: bool = True):
    """Differentiable ``synthesis`` / ``synthesis_batch`` (rrule :76-89, :113-133)."""
    return _Synthesis.apply(cfg, alm, real_output)


def analysis(cfg, f):
    """Differentiable ``analysis`` / ``analysis_batch`` (rrule :55-67, :95-111)."""
    return _Analysis.apply(cfg, f)


def synthesis_sphtor(cfg, S, Tlm, real_output: bool = True):
    """Differentiable ``synthesis_sphtor`` (rrule :225-241)."""
    return _SynthesisSphtor.apply(cfg, S, Tlm, real_output)


def analysis_sphtor(cfg, Vt, Vp):
    """Differentiable ``analysis_sphtor`` (rrule :198-215)."""
    return _AnalysisSphtor.apply(cfg, Vt, Vp)
