# SHTnsKitMI355XExt -- the Julia side of the MI355X backend: registers an
# adapter with SHTnsKit's vendor-neutral GPU routing and reaches the HIP
# kernels through `ccall` into libshtk_mi355x.so (C ABI in include/shtk.h).
#
# Drop-in point: src/device_utils.jl:11-16 (`_register_gpu_adapter!`) and the
# adapter protocol of src/device_utils.jl:33-99 and
# src/sphtor_transforms.jl:298-311.  The extension owns its adapter in a
# `const` because core keeps only a WeakRef (as ext/SHTnsKitAMDGPUExt.jl:87-95).
#
# Coexistence with SHTnsKitAMDGPUExt (both load with AMDGPU.jl): core routing
# needs exactly one adapter to recognize an array (src/device_utils.jl:140-147).
# This adapter recognizes ROCm arrays on gfx950 (MI355X) devices only, and the
# AMDGPU adapter's `_gpu_adapter_matches` is narrowed by one line to the other
# devices (the patch is in INTEGRATION.md).  The direct overloads below (batch,
# in-place, planned) take dense `ROCArray`s, which is more specific than the
# AMDGPU extension's `AMDGPU.AnyROCArray` methods with the same element-type
# bounds, so dispatch is unambiguous; on other devices they hand the call to
# those methods with `invoke`.  There is no second kernel path on an MI355X: a
# failing call raises, it never falls back.
module SHTnsKitMI355XExt

using AMDGPU
using SHTnsKit
using SHTnsKit: SHTConfig, SHTPlan, BackendUnavailableError
import SHTnsKit: _register_gpu_adapter!, _gpu_adapter_functional, _gpu_adapter_matches,
                 _gpu_adapter_adapt, _gpu_adapter_dense, _gpu_adapter_analysis,
                 _gpu_adapter_synthesis, _gpu_adapter_analysis_sphtor,
                 _gpu_adapter_synthesis_sphtor, _gpu_adapter_clear_cache!

const LIB = get(ENV, "SHTK_MI355X_LIB", joinpath(@__DIR__, "..", "libshtk_mi355x.so"))

# status codes and enums of include/shtk.h
const SHTK_OK, SHTK_EINVAL, SHTK_EUNSUPPORTED, SHTK_ENOMEM, SHTK_EHIP = 0, 1, 2, 3, 4
const SHTK_F64, SHTK_F32 = 0, 1
const SHTK_PREC_FP64, SHTK_PREC_SPLIT, SHTK_PREC_FP32 = 0, 1, 2

struct ShtkDesc          # layout of `shtk_desc`
    lmax::Cint; mmax::Cint; mres::Cint; nlat::Cint; nlon::Cint
    norm::Cint; cs_phase::Cint; real_norm::Cint; robert_form::Cint; phi_scale::Cint
    grid::Cint
end

function _last_error()
    buf = Vector{UInt8}(undef, 512)
    ccall((:shtk_last_error, LIB), Cint, (Ptr{UInt8}, Csize_t), buf, length(buf))
    return unsafe_string(pointer(buf))
end

# status -> the reference's exception kinds (src/devices.jl:19-26)
function _check(op::Symbol, status::Cint)
    status == SHTK_OK && return nothing
    msg = _last_error()
    status == SHTK_EINVAL && throw(ArgumentError(msg))
    status == SHTK_EUNSUPPORTED && throw(BackendUnavailableError(op, msg))
    status == SHTK_ENOMEM && throw(OutOfMemoryError())
    throw(BackendUnavailableError(op, "HIP error: " * msg))
end

_norm_code(cfg) = cfg.norm === :orthonormal ? 0 : cfg.norm === :fourpi ? 1 : 2
# shtk_grid: Gauss-Legendre, Fejer midpoints, Clenshaw-Curtis with poles, Driscoll-Healy
function _grid_code(cfg)
    g = cfg.grid_type
    (g === :gauss || g === :gauss_fly) && return 0
    g === :regular && return 1
    g === :regular_poles && return 2
    g === :driscoll_healy && return 3
    throw(BackendUnavailableError(:plan, "grid_type $(g) is not provided by the MI355X backend"))
end
_desc(cfg::SHTConfig) = ShtkDesc(cfg.lmax, cfg.mmax, cfg.mres, cfg.nlat, cfg.nlon, _norm_code(cfg),
                                 cfg.cs_phase, cfg.real_norm, cfg.robert_form,
                                 cfg.phi_scale === :quad ? 1 : 0, _grid_code(cfg))

# Plan cache keyed by (device, config signature).  Plans are read-only after
# creation; the library orders calls that share a plan (its workspace) across
# streams, so tasks with their own AMDGPU streams may share one plan.
mutable struct MI355XAdapter
    plans::Dict{Any,Ptr{Cvoid}}
    lock::ReentrantLock
end
const MI355X_ADAPTER = MI355XAdapter(Dict{Any,Ptr{Cvoid}}(), ReentrantLock())

function _plan(cfg::SHTConfig)
    key = (AMDGPU.device_id(AMDGPU.device()), _desc(cfg))
    lock(MI355X_ADAPTER.lock) do
        get!(MI355X_ADAPTER.plans, key) do
            d = Ref(_desc(cfg)); p = Ref{Ptr{Cvoid}}(C_NULL)
            _check(:plan, ccall((:shtk_plan_create, LIB), Cint, (Ref{ShtkDesc}, Ref{Ptr{Cvoid}}), d, p))
            p[]
        end
    end
end

_dtype(::Type{Float64}) = SHTK_F64
_dtype(::Type{Float32}) = SHTK_F32
_dtype(::Type{T}) where {T} = throw(ArgumentError("MI355X transforms take Float32 or Float64 data, got $T"))
_stream() = AMDGPU.stream().stream       # the task's HIP stream (hipStream_t)
# contraction precision of the batch transforms: :fp64 (plain), :split (Ozaki
# split on the int8 matrix cores, FP64-grade), :fp32 (FP32 matrix cores)
_precision(p::Symbol) = p === :fp64 ? SHTK_PREC_FP64 : p === :split ? SHTK_PREC_SPLIT :
    p === :fp32 ? SHTK_PREC_FP32 : throw(ArgumentError("precision must be :fp64, :split or :fp32, got :$p"))

# The library is built for gfx950; the adapter claims arrays on such devices.
const _ARCH_PREFIX = get(ENV, "SHTK_MI355X_ARCH", "gfx950")
_is_mi355x(dev) = startswith(AMDGPU.HIP.gcn_arch(dev), _ARCH_PREFIX)
_on_mi355x(x::AMDGPU.AnyROCArray) = _is_mi355x(AMDGPU.device(x))

function __init__()
    _register_gpu_adapter!(:mi355x, MI355X_ADAPTER)
    return nothing
end

# --- adapter protocol (src/device_utils.jl:33-99) -------------------------------
_gpu_adapter_functional(::MI355XAdapter) =
    AMDGPU.functional() && ccall((:shtk_version, LIB), Cstring, ()) != C_NULL
_gpu_adapter_matches(::MI355XAdapter, x::AMDGPU.AnyROCArray) = _on_mi355x(x)
_gpu_adapter_matches(a::MI355XAdapter, x::AbstractArray) = (p = parent(x)) !== x && _gpu_adapter_matches(a, p)
_gpu_adapter_adapt(::MI355XAdapter, x::AbstractArray) = ROCArray(x)
_gpu_adapter_dense(::MI355XAdapter, x::AbstractArray) = x isa ROCArray ? x : ROCArray(x)

# dense device storage of an operand (wrappers are copied once on the device)
_dense(x::ROCArray) = x
_dense(x::AbstractArray) = similar(x, eltype(x), size(x)) .= x

_nfields(x) = ndims(x) == 3 ? size(x, 3) : 1
_check_spec(op, cfg, x) = size(x, 1) == cfg.lmax + 1 && size(x, 2) == cfg.mmax + 1 ||
    throw(DimensionMismatch("$op: coefficients must be (lmax+1, mmax+1[, nf]) = ($(cfg.lmax + 1), $(cfg.mmax + 1)), got $(size(x))"))
_check_grid(op, cfg, x) = size(x, 1) == cfg.nlat && size(x, 2) == cfg.nlon ||
    throw(DimensionMismatch("$op: fields must be (nlat, nlon[, nf]) = ($(cfg.nlat), $(cfg.nlon)), got $(size(x))"))
function _check_pair(op, a, b)
    size(a) == size(b) || throw(DimensionMismatch("$op: the two components differ in size, $(size(a)) vs $(size(b))"))
    ndims(a) in (2, 3) || throw(DimensionMismatch("$op: expected 2 or 3 dimensions, got $(ndims(a))"))
    _nfields(a) > 0 || throw(ArgumentError("$op requires at least one field"))
end

# ---- the C calls, writing into caller-provided outputs -------------------------
function _synthesis_into!(f, cfg, alm; real_output::Bool=true, precision::Symbol=:fp64)
    _check_spec(:synthesis, cfg, alm); _check_grid(:synthesis, cfg, f)
    nf = _nfields(alm)
    _nfields(f) == nf || throw(DimensionMismatch("synthesis: $(nf) coefficient fields, $(_nfields(f)) output fields"))
    RT = float(real(eltype(f)))
    (real_output ? eltype(f) <: Real : eltype(f) <: Complex) ||
        throw(ArgumentError("real_output=$(real_output) does not match output eltype $(eltype(f))"))
    a = ROCArray{Complex{RT}}(alm)
    _check(:synthesis, ccall((:shtk_synthesis_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(a), pointer(f), nf, _dtype(RT), real_output ? 1 : 0, cfg.lmax, _precision(precision),
        _stream()))
    AMDGPU.synchronize()                 # results ready on return, as the AMDGPU adapter
    return f
end

function _analysis_into!(alm, cfg, f; use_rfft::Bool=false, precision::Symbol=:fp64)
    _check_grid(:analysis, cfg, f); _check_spec(:analysis, cfg, alm)
    nf = _nfields(f)
    _nfields(alm) == nf || throw(DimensionMismatch("analysis: $(nf) fields, $(_nfields(alm)) output fields"))
    RT = float(real(eltype(alm)))
    if eltype(f) <: Complex
        # complex fields: the bins m >= 0 of the complex FFT (src/core_transforms.jl:169-207)
        use_rfft && throw(ArgumentError("use_rfft=true requires a real-valued input"))
        x = ROCArray{Complex{RT}}(f)
        _check(:analysis, ccall((:shtk_analysis_cplx, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
            _plan(cfg), pointer(x), pointer(alm), nf, _dtype(RT), cfg.lmax, _precision(precision), _stream()))
    else
        x = ROCArray{RT}(f)
        _check(:analysis, ccall((:shtk_analysis_ex, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
            _plan(cfg), pointer(x), pointer(alm), nf, _dtype(RT), cfg.lmax, _precision(precision), _stream()))
    end
    AMDGPU.synchronize()
    return alm
end

function _synthesis_sphtor_into!(Vt, Vp, cfg, S, T; real_output::Bool=true, precision::Symbol=:fp64)
    _check_pair(:synthesis_sphtor, S, T); _check_pair(:synthesis_sphtor, Vt, Vp)
    _check_spec(:synthesis_sphtor, cfg, S); _check_grid(:synthesis_sphtor, cfg, Vt)
    nf = _nfields(S)
    _nfields(Vt) == nf || throw(DimensionMismatch("synthesis_sphtor: $(nf) coefficient fields, $(_nfields(Vt)) outputs"))
    RT = float(real(eltype(Vt)))
    s = ROCArray{Complex{RT}}(S); t = ROCArray{Complex{RT}}(T)
    _check(:synthesis_sphtor, ccall((:shtk_synthesis_sphtor_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(s), pointer(t), pointer(Vt), pointer(Vp), nf, _dtype(RT), real_output ? 1 : 0, cfg.lmax,
        _precision(precision), _stream()))
    AMDGPU.synchronize()
    return Vt, Vp
end

function _analysis_sphtor_into!(S, T, cfg, Vt, Vp; precision::Symbol=:fp64)
    _check_pair(:analysis_sphtor, Vt, Vp); _check_pair(:analysis_sphtor, S, T)
    _check_grid(:analysis_sphtor, cfg, Vt); _check_spec(:analysis_sphtor, cfg, S)
    nf = _nfields(Vt)
    _nfields(S) == nf || throw(DimensionMismatch("analysis_sphtor: $(nf) fields, $(_nfields(S)) outputs"))
    RT = float(real(eltype(S)))
    vt = ROCArray{RT}(Vt); vp = ROCArray{RT}(Vp)
    _check(:analysis_sphtor, ccall((:shtk_analysis_sphtor_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(vt), pointer(vp), pointer(S), pointer(T), nf, _dtype(RT), cfg.lmax,
        _precision(precision), _stream()))
    AMDGPU.synchronize()
    return S, T
end

_grid_dims(cfg, x) = ndims(x) == 3 ? (cfg.nlat, cfg.nlon, size(x, 3)) : (cfg.nlat, cfg.nlon)
_spec_dims(cfg, x) = ndims(x) == 3 ? (cfg.lmax + 1, cfg.mmax + 1, size(x, 3)) : (cfg.lmax + 1, cfg.mmax + 1)

function _synthesis_new(cfg, alm; real_output::Bool=true, precision::Symbol=:fp64)
    RT = float(real(eltype(alm)))
    f = AMDGPU.zeros(real_output ? RT : Complex{RT}, _grid_dims(cfg, alm)...)
    return _synthesis_into!(f, cfg, _dense(alm); real_output, precision)
end
function _analysis_new(cfg, f; use_rfft::Bool=false, precision::Symbol=:fp64)
    alm = AMDGPU.zeros(Complex{float(real(eltype(f)))}, _spec_dims(cfg, f)...)
    return _analysis_into!(alm, cfg, _dense(f); use_rfft, precision)
end
function _synthesis_sphtor_new(cfg, S, T; real_output::Bool=true, precision::Symbol=:fp64)
    RT = float(real(promote_type(eltype(S), eltype(T))))
    OT = real_output ? RT : Complex{RT}   # complex grids of the bins m >= 0
    Vt = AMDGPU.zeros(OT, _grid_dims(cfg, S)...); Vp = AMDGPU.zeros(OT, _grid_dims(cfg, S)...)
    return _synthesis_sphtor_into!(Vt, Vp, cfg, _dense(S), _dense(T); real_output, precision)
end
function _analysis_sphtor_new(cfg, Vt, Vp; precision::Symbol=:fp64)
    CT = Complex{float(real(promote_type(eltype(Vt), eltype(Vp))))}
    S = AMDGPU.zeros(CT, _spec_dims(cfg, Vt)...); T = AMDGPU.zeros(CT, _spec_dims(cfg, Vt)...)
    return _analysis_sphtor_into!(S, T, cfg, _dense(Vt), _dense(Vp); precision)
end

_no_scratch(fft_scratch) = fft_scratch === nothing ||
    throw(ArgumentError("MI355X transforms do not accept a host fft_scratch"))

# ---- protocol transforms --------------------------------------------------------
function _gpu_adapter_synthesis(::MI355XAdapter, cfg::SHTConfig, alm::AbstractArray;
                                real_output::Bool=true, use_rfft::Bool=false, fft_scratch=nothing,
                                precision::Symbol=:fp64)
    _no_scratch(fft_scratch)
    return _synthesis_new(cfg, alm; real_output, precision)
end

function _gpu_adapter_analysis(::MI355XAdapter, cfg::SHTConfig, f::AbstractArray;
                               use_rfft::Bool=false, fft_scratch=nothing, precision::Symbol=:fp64)
    _no_scratch(fft_scratch)
    return _analysis_new(cfg, f; use_rfft, precision)
end

function _gpu_adapter_synthesis_sphtor(::MI355XAdapter, cfg::SHTConfig, S, T;
                                       real_output::Bool=true, use_rfft::Bool=false, precision::Symbol=:fp64)
    return _synthesis_sphtor_new(cfg, S, T; real_output, precision)
end

function _gpu_adapter_analysis_sphtor(::MI355XAdapter, cfg::SHTConfig, Vt, Vp; use_rfft::Bool=false,
                                      precision::Symbol=:fp64)
    return _analysis_sphtor_new(cfg, Vt, Vp; precision)
end

function _gpu_adapter_clear_cache!(a::MI355XAdapter)
    lock(a.lock) do
        for p in values(a.plans)
            ccall((:shtk_plan_destroy, LIB), Cint, (Ptr{Cvoid},), p)
        end
        empty!(a.plans)
    end
    return nothing
end

# ---- direct overloads (no protocol hook in core; ext/SHTnsKitAMDGPUExt.jl:1283-1301,
# 1865-2030).  Dense ROCArrays on an MI355X run here; anything else goes to the
# AMDGPU extension's AnyROCArray method of the same function via `invoke`. ------------
const _G = SHTnsKit.GPU
const _Any2{T} = AMDGPU.AnyROCArray{T,2}
const _Any3{T} = AMDGPU.AnyROCArray{T,3}

# batch transforms (:1865-1866, 1963-1980, 1283-1301)
function SHTnsKit.analysis_batch(g::_G, cfg::SHTConfig, f::ROCArray{T,3}; use_rfft::Bool=false,
                                 precision::Symbol=:fp64) where {T<:Real}
    _on_mi355x(f) || return invoke(SHTnsKit.analysis_batch, Tuple{_G,SHTConfig,_Any3{T}}, g, cfg, f; use_rfft)
    return SHTnsKit._with_gpu_device(f) do
        _analysis_new(cfg, f; use_rfft, precision)
    end
end
SHTnsKit.analysis_batch(cfg::SHTConfig, f::ROCArray{T,3}; kwargs...) where {T<:Real} =
    SHTnsKit.analysis_batch(_G(), cfg, f; kwargs...)

function SHTnsKit.synthesis_batch(g::_G, cfg::SHTConfig, alm::ROCArray{T,3}; real_output::Bool=true,
                                  use_rfft::Bool=false, precision::Symbol=:fp64) where {T<:Complex}
    _on_mi355x(alm) ||
        return invoke(SHTnsKit.synthesis_batch, Tuple{_G,SHTConfig,_Any3{T}}, g, cfg, alm; real_output, use_rfft)
    return SHTnsKit._with_gpu_device(alm) do
        _synthesis_new(cfg, alm; real_output, precision)
    end
end
SHTnsKit.synthesis_batch(cfg::SHTConfig, alm::ROCArray{T,3}; kwargs...) where {T<:Complex} =
    SHTnsKit.synthesis_batch(_G(), cfg, alm; kwargs...)
SHTnsKit.synthesis_batch_cplx(g::_G, cfg::SHTConfig, alm::ROCArray{T,3}) where {T<:Complex} =
    SHTnsKit.synthesis_batch(g, cfg, alm; real_output=false)
SHTnsKit.synthesis_batch_cplx(cfg::SHTConfig, alm::ROCArray{T,3}) where {T<:Complex} =
    SHTnsKit.synthesis_batch(_G(), cfg, alm; real_output=false)

function SHTnsKit.analysis_batch!(g::_G, cfg::SHTConfig, out::ROCArray{T,3}, f::ROCArray{R,3};
                                  use_rfft::Bool=false, precision::Symbol=:fp64) where {T<:Complex,R<:Real}
    _on_mi355x(f) ||
        return invoke(SHTnsKit.analysis_batch!, Tuple{_G,SHTConfig,_Any3{T},_Any3{R}}, g, cfg, out, f; use_rfft)
    return SHTnsKit._with_gpu_device(f, out) do
        _analysis_into!(out, cfg, f; use_rfft, precision)
    end
end
SHTnsKit.analysis_batch!(cfg::SHTConfig, out::ROCArray{T,3}, f::ROCArray{R,3}; kwargs...) where {T<:Complex,R<:Real} =
    SHTnsKit.analysis_batch!(_G(), cfg, out, f; kwargs...)

function _synthesis_batch_bang(g, cfg, out::ROCArray{T,3}, alm::ROCArray{R,3}; real_output::Bool=true,
                               use_rfft::Bool=false, precision::Symbol=:fp64) where {T,R}
    # invoke on the AnyROCArray signature picks the AMDGPU extension's method for
    # this output eltype (its two methods, :1981-1993)
    _on_mi355x(alm) || return invoke(SHTnsKit.synthesis_batch!, Tuple{_G,SHTConfig,_Any3{T},_Any3{R}},
                                     g, cfg, out, alm; real_output, use_rfft)
    return SHTnsKit._with_gpu_device(alm, out) do
        _synthesis_into!(out, cfg, alm; real_output, precision)
    end
end
# two methods, as the reference's (:1981-1993), each more specific than its counterpart
SHTnsKit.synthesis_batch!(g::_G, cfg::SHTConfig, out::ROCArray{T,3}, alm::ROCArray{R,3};
                          kwargs...) where {T,R<:Complex} = _synthesis_batch_bang(g, cfg, out, alm; kwargs...)
SHTnsKit.synthesis_batch!(g::_G, cfg::SHTConfig, out::ROCArray{T,3}, alm::ROCArray{R,3};
                          kwargs...) where {T<:Real,R<:Complex} = _synthesis_batch_bang(g, cfg, out, alm; kwargs...)
SHTnsKit.synthesis_batch!(cfg::SHTConfig, out::ROCArray{T,3}, alm::ROCArray{R,3}; kwargs...) where {T,R<:Complex} =
    SHTnsKit.synthesis_batch!(_G(), cfg, out, alm; kwargs...)

function SHTnsKit.analysis_sphtor_batch(g::_G, cfg::SHTConfig, Vt::ROCArray{T,3}, Vp::ROCArray{R,3};
                                        precision::Symbol=:fp64) where {T<:Real,R<:Real}
    _on_mi355x(Vt) ||
        return invoke(SHTnsKit.analysis_sphtor_batch, Tuple{_G,SHTConfig,_Any3{T},_Any3{R}}, g, cfg, Vt, Vp)
    return SHTnsKit._with_gpu_device(Vt, Vp) do
        _analysis_sphtor_new(cfg, Vt, Vp; precision)
    end
end
SHTnsKit.analysis_sphtor_batch(cfg::SHTConfig, Vt::ROCArray{T,3}, Vp::ROCArray{R,3}; kwargs...) where {T<:Real,R<:Real} =
    SHTnsKit.analysis_sphtor_batch(_G(), cfg, Vt, Vp; kwargs...)

function SHTnsKit.synthesis_sphtor_batch(g::_G, cfg::SHTConfig, S::ROCArray{T,3}, Tl::ROCArray{R,3};
                                         real_output::Bool=true, precision::Symbol=:fp64) where {T<:Complex,R<:Complex}
    _on_mi355x(S) || return invoke(SHTnsKit.synthesis_sphtor_batch, Tuple{_G,SHTConfig,_Any3{T},_Any3{R}},
                                   g, cfg, S, Tl; real_output)
    return SHTnsKit._with_gpu_device(S, Tl) do
        _synthesis_sphtor_new(cfg, S, Tl; real_output, precision)
    end
end
SHTnsKit.synthesis_sphtor_batch(cfg::SHTConfig, S::ROCArray{T,3}, Tl::ROCArray{R,3};
                                kwargs...) where {T<:Complex,R<:Complex} =
    SHTnsKit.synthesis_sphtor_batch(_G(), cfg, S, Tl; kwargs...)
SHTnsKit.synthesis_sphtor_batch_cplx(g::_G, cfg::SHTConfig, S::ROCArray{T,3}, Tl::ROCArray{R,3}) where {T<:Complex,R<:Complex} =
    SHTnsKit.synthesis_sphtor_batch(g, cfg, S, Tl; real_output=false)
SHTnsKit.synthesis_sphtor_batch_cplx(cfg::SHTConfig, S::ROCArray{T,3}, Tl::ROCArray{R,3}) where {T<:Complex,R<:Complex} =
    SHTnsKit.synthesis_sphtor_batch(_G(), cfg, S, Tl; real_output=false)

# in-place transforms without a plan (:2018-2030): the output is written in place
function SHTnsKit.analysis!(cfg::SHTConfig, out::ROCArray{T,2}, f::ROCArray{R,2};
                            fft_scratch=nothing, use_rfft::Bool=false) where {T,R}
    _on_mi355x(f) || return invoke(SHTnsKit.analysis!, Tuple{SHTConfig,_Any2{T},_Any2{R}}, cfg, out, f;
                                   fft_scratch, use_rfft)
    _no_scratch(fft_scratch)
    eltype(out) <: Complex || throw(ArgumentError("analysis! output must be complex, got $(eltype(out))"))
    return SHTnsKit._with_gpu_device(f, out) do
        _analysis_into!(out, cfg, f; use_rfft)
    end
end
function SHTnsKit.synthesis!(cfg::SHTConfig, out::ROCArray{T,2}, alm::ROCArray{R,2};
                             real_output::Bool=true, fft_scratch=nothing, use_rfft::Bool=false) where {T,R}
    _on_mi355x(alm) || return invoke(SHTnsKit.synthesis!, Tuple{SHTConfig,_Any2{T},_Any2{R}}, cfg, out, alm;
                                     real_output, fft_scratch, use_rfft)
    _no_scratch(fft_scratch)
    return SHTnsKit._with_gpu_device(alm, out) do
        _synthesis_into!(out, cfg, alm; real_output)
    end
end

# planned transforms (:1995-2016, 1337-1345): the plan's configuration, device workspace
function SHTnsKit.analysis!(g::_G, plan::SHTPlan, out::ROCArray{T,2}, f::ROCArray{R,2}) where {T<:Complex,R<:Number}
    _on_mi355x(f) || return invoke(SHTnsKit.analysis!, Tuple{_G,SHTPlan,_Any2{T},_Any2{R}}, g, plan, out, f)
    return SHTnsKit._with_gpu_device(f, out) do
        _analysis_into!(out, plan.cfg, f; use_rfft=plan.use_rfft)
    end
end
SHTnsKit.analysis!(plan::SHTPlan, out::ROCArray{T,2}, f::ROCArray{R,2}) where {T<:Complex,R<:Number} =
    SHTnsKit.analysis!(_G(), plan, out, f)

function SHTnsKit.synthesis!(g::_G, plan::SHTPlan, out::ROCArray{T,2}, alm::ROCArray{R,2};
                             real_output::Bool=true) where {T<:Number,R<:Complex}
    _on_mi355x(alm) || return invoke(SHTnsKit.synthesis!, Tuple{_G,SHTPlan,_Any2{T},_Any2{R}}, g, plan, out, alm;
                                     real_output)
    return SHTnsKit._with_gpu_device(alm, out) do
        _synthesis_into!(out, plan.cfg, alm; real_output)
    end
end
SHTnsKit.synthesis!(plan::SHTPlan, out::ROCArray{T,2}, alm::ROCArray{R,2}; real_output::Bool=true) where {T<:Number,R<:Complex} =
    SHTnsKit.synthesis!(_G(), plan, out, alm; real_output)

function SHTnsKit.analysis_sphtor!(plan::SHTPlan, S::ROCArray, T::ROCArray, Vt::ROCArray, Vp::ROCArray)
    _on_mi355x(Vt) || return invoke(SHTnsKit.analysis_sphtor!,
                                    Tuple{SHTPlan,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray},
                                    plan, S, T, Vt, Vp)
    return SHTnsKit._with_gpu_device(Vt, Vp, S, T) do
        _analysis_sphtor_into!(S, T, plan.cfg, Vt, Vp)
    end
end
function SHTnsKit.synthesis_sphtor!(plan::SHTPlan, Vt::ROCArray, Vp::ROCArray, S::ROCArray, T::ROCArray;
                                    real_output::Bool=true)
    _on_mi355x(S) || return invoke(SHTnsKit.synthesis_sphtor!,
                                   Tuple{SHTPlan,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray,AMDGPU.AnyROCArray},
                                   plan, Vt, Vp, S, T; real_output)
    return SHTnsKit._with_gpu_device(S, T, Vt, Vp) do
        _synthesis_sphtor_into!(Vt, Vp, plan.cfg, S, T; real_output)
    end
end

# ---- adjoints behind the reference's rrules ------------------------------------
# ext/SHTnsKitAdvancedADExt.jl:55-283 call SHTnsKit._adjoint_synthesis,
# _adjoint_analysis, _adjoint_synthesis_sphtor and _adjoint_analysis_sphtor
# (src/core_transforms.jl:477-652, src/sphtor_transforms.jl:385-452).  These device
# methods run the HIP adjoint kernels; the rrules stay unchanged.  Arrays on other
# devices keep the generic methods.
_full_theta(cfg, θ) = θ == 1:cfg.nlat || throw(ArgumentError(
    "the MI355X adjoints cover the whole latitude range (θ_globals = 1:nlat)"))

function _adjoint_synthesis_dev(cfg, fbar, real_output)
    RT = float(real(eltype(fbar)))
    CT = Complex{RT}
    if eltype(fbar) <: Complex  # complex-linear: real and imaginary parts separately
        return _adjoint_synthesis_dev(cfg, real.(fbar), real_output) .+
               im .* _adjoint_synthesis_dev(cfg, imag.(fbar), real_output)
    end
    _check_grid(:_adjoint_synthesis, cfg, fbar)
    y = ROCArray{RT}(fbar)
    out = AMDGPU.zeros(CT, _spec_dims(cfg, fbar)...)
    _check(:_adjoint_synthesis, ccall((:shtk_adjoint_synthesis, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(y), pointer(out), _nfields(fbar), _dtype(RT), real_output ? 1 : 0, _stream()))
    AMDGPU.synchronize()
    return out
end

function _adjoint_analysis_dev(cfg, abar)
    _check_spec(:_adjoint_analysis, cfg, abar)
    RT = float(real(eltype(abar)))
    a = ROCArray{Complex{RT}}(abar)
    out = AMDGPU.zeros(Complex{RT}, _grid_dims(cfg, abar)...)
    _check(:_adjoint_analysis, ccall((:shtk_adjoint_analysis_cplx, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(a), pointer(out), _nfields(abar), _dtype(RT), _stream()))
    AMDGPU.synchronize()
    return out
end

function _adjoint_synthesis_sphtor_dev(cfg, Vtb, Vpb, real_output)
    _check_pair(:_adjoint_synthesis_sphtor, Vtb, Vpb)
    RT = float(real(promote_type(eltype(Vtb), eltype(Vpb))))
    if eltype(Vtb) <: Complex || eltype(Vpb) <: Complex  # complex-linear
        Sr, Tr = _adjoint_synthesis_sphtor_dev(cfg, real.(Vtb), real.(Vpb), real_output)
        Si, Ti = _adjoint_synthesis_sphtor_dev(cfg, imag.(Vtb), imag.(Vpb), real_output)
        return Sr .+ im .* Si, Tr .+ im .* Ti
    end
    _check_grid(:_adjoint_synthesis_sphtor, cfg, Vtb)
    vt = ROCArray{RT}(Vtb); vp = ROCArray{RT}(Vpb)
    S = AMDGPU.zeros(Complex{RT}, _spec_dims(cfg, Vtb)...); T = AMDGPU.zeros(Complex{RT}, _spec_dims(cfg, Vtb)...)
    _check(:_adjoint_synthesis_sphtor, ccall((:shtk_adjoint_synthesis_sphtor, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(vt), pointer(vp), pointer(S), pointer(T), _nfields(Vtb), _dtype(RT),
        real_output ? 1 : 0, _stream()))
    AMDGPU.synchronize()
    return S, T
end

function _adjoint_analysis_sphtor_dev(cfg, Sb, Tb)
    _check_pair(:_adjoint_analysis_sphtor, Sb, Tb)
    _check_spec(:_adjoint_analysis_sphtor, cfg, Sb)
    RT = float(real(promote_type(eltype(Sb), eltype(Tb))))
    s = ROCArray{Complex{RT}}(Sb); t = ROCArray{Complex{RT}}(Tb)
    Vt = AMDGPU.zeros(Complex{RT}, _grid_dims(cfg, Sb)...); Vp = AMDGPU.zeros(Complex{RT}, _grid_dims(cfg, Sb)...)
    _check(:_adjoint_analysis_sphtor, ccall((:shtk_adjoint_analysis_sphtor_cplx, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(s), pointer(t), pointer(Vt), pointer(Vp), _nfields(Sb), _dtype(RT), _stream()))
    AMDGPU.synchronize()
    return Vt, Vp
end

function SHTnsKit._adjoint_synthesis(cfg::SHTConfig, fbar::AMDGPU.AnyROCArray{T,2};
                                     θ_globals::AbstractVector{<:Integer}=1:cfg.nlat,
                                     real_output::Bool=true) where {T}
    _on_mi355x(fbar) || return invoke(SHTnsKit._adjoint_synthesis, Tuple{SHTConfig,AbstractMatrix}, cfg, fbar;
                                      θ_globals, real_output)
    _full_theta(cfg, θ_globals)
    return SHTnsKit._with_gpu_device(fbar) do
        _adjoint_synthesis_dev(cfg, _dense(fbar), real_output)
    end
end

function SHTnsKit._adjoint_analysis(cfg::SHTConfig, abar::AMDGPU.AnyROCArray{T,2};
                                    θ_globals::AbstractVector{<:Integer}=1:cfg.nlat,
                                    φ_window::Union{Nothing,UnitRange{Int}}=nothing) where {T}
    _on_mi355x(abar) || return invoke(SHTnsKit._adjoint_analysis, Tuple{SHTConfig,AbstractMatrix}, cfg, abar;
                                      θ_globals, φ_window)
    _full_theta(cfg, θ_globals)
    out = SHTnsKit._with_gpu_device(abar) do
        _adjoint_analysis_dev(cfg, _dense(abar))
    end
    return φ_window === nothing ? out : out[:, φ_window]
end

function SHTnsKit._adjoint_synthesis_sphtor(cfg::SHTConfig, Vtb::AMDGPU.AnyROCArray{T,2},
                                            Vpb::AMDGPU.AnyROCArray{R,2};
                                            θ_globals::AbstractVector{<:Integer}=1:cfg.nlat,
                                            real_output::Bool=true) where {T,R}
    _on_mi355x(Vtb) || return invoke(SHTnsKit._adjoint_synthesis_sphtor,
                                     Tuple{SHTConfig,AbstractMatrix,AbstractMatrix}, cfg, Vtb, Vpb;
                                     θ_globals, real_output)
    _full_theta(cfg, θ_globals)
    return SHTnsKit._with_gpu_device(Vtb, Vpb) do
        _adjoint_synthesis_sphtor_dev(cfg, _dense(Vtb), _dense(Vpb), real_output)
    end
end

function SHTnsKit._adjoint_analysis_sphtor(cfg::SHTConfig, Sb::AMDGPU.AnyROCArray{T,2},
                                           Tb::AMDGPU.AnyROCArray{R,2}; kwargs...) where {T,R}
    _on_mi355x(Sb) || return invoke(SHTnsKit._adjoint_analysis_sphtor,
                                    Tuple{SHTConfig,AbstractMatrix,AbstractMatrix}, cfg, Sb, Tb; kwargs...)
    θ = get(kwargs, :θ_globals, 1:cfg.nlat)
    _full_theta(cfg, θ)
    return SHTnsKit._with_gpu_device(Sb, Tb) do
        _adjoint_analysis_sphtor_dev(cfg, _dense(Sb), _dense(Tb))
    end
end

end # module
