# SHTnsKitMI355XExt -- the Julia side of the MI355X backend: registers an
# adapter with SHTnsKit's vendor-neutral GPU routing and reaches the HIP
# kernels through `ccall` into libshtk_mi355x.so (C ABI in include/shtk.h).
#
# Drop-in point: src/device_utils.jl:11-16 (`_register_gpu_adapter!`) and the
# adapter protocol of src/device_utils.jl:33-99 and
# src/sphtor_transforms.jl:298-311.  The extension owns its adapter in a
# `const` because core keeps only a WeakRef (as ext/SHTnsKitAMDGPUExt.jl:87-95).
# There is no second kernel path: a failing call raises, it never falls back.
module SHTnsKitMI355XExt

using AMDGPU
using SHTnsKit
using SHTnsKit: SHTConfig, BackendUnavailableError
import SHTnsKit: _register_gpu_adapter!, _gpu_adapter_functional, _gpu_adapter_matches,
                 _gpu_adapter_adapt, _gpu_adapter_dense, _gpu_adapter_analysis,
                 _gpu_adapter_synthesis, _gpu_adapter_analysis_sphtor,
                 _gpu_adapter_synthesis_sphtor, _gpu_adapter_clear_cache!, on_device

const LIB = get(ENV, "SHTK_MI355X_LIB", joinpath(@__DIR__, "..", "libshtk_mi355x.so"))

# status codes of include/shtk.h
const SHTK_OK, SHTK_EINVAL, SHTK_EUNSUPPORTED, SHTK_ENOMEM, SHTK_EHIP = 0, 1, 2, 3, 4
const SHTK_F64, SHTK_F32 = 0, 1
const SHTK_PREC_FP64, SHTK_PREC_SPLIT = 0, 1

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

# plan cache keyed by (device, config signature); plans are read-only after creation
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
_stream() = AMDGPU.stream().stream       # the task's HIP stream (hipStream_t)
# contraction precision of the batch transforms: :fp64 (plain) or :split (Ozaki
# split on the int8 matrix cores, FP64-grade)
_precision(p::Symbol) = p === :fp64 ? SHTK_PREC_FP64 : p === :split ? SHTK_PREC_SPLIT :
    throw(ArgumentError("precision must be :fp64 or :split, got :$p"))

function __init__()
    _register_gpu_adapter!(:mi355x, MI355X_ADAPTER)
    return nothing
end

_gpu_adapter_functional(::MI355XAdapter) =
    AMDGPU.functional() && ccall((:shtk_version, LIB), Cstring, ()) != C_NULL
_gpu_adapter_matches(::MI355XAdapter, ::AMDGPU.AnyROCArray) = true
_gpu_adapter_matches(a::MI355XAdapter, x::AbstractArray) = (p = parent(x)) !== x && _gpu_adapter_matches(a, p)
_gpu_adapter_adapt(::MI355XAdapter, x::AbstractArray) = ROCArray(x)
_gpu_adapter_dense(::MI355XAdapter, x::AbstractArray) = x isa ROCArray ? x : ROCArray(x)

function _gpu_adapter_synthesis(::MI355XAdapter, cfg::SHTConfig, alm::AbstractArray;
                                real_output::Bool=true, use_rfft::Bool=false, fft_scratch=nothing,
                                precision::Symbol=:fp64)
    fft_scratch === nothing || throw(ArgumentError("MI355X transforms do not accept a host fft_scratch"))
    size(alm)[1:2] == (cfg.lmax + 1, cfg.mmax + 1) || throw(DimensionMismatch("alm size $(size(alm))"))
    RT = float(real(eltype(alm)))
    OT = real_output ? RT : Complex{RT}   # synthesis_cplx: complex grid of the bins m >= 0
    nf = ndims(alm) == 3 ? size(alm, 3) : 1
    a = ROCArray{Complex{RT}}(alm)
    f = ndims(alm) == 3 ? AMDGPU.zeros(OT, cfg.nlat, cfg.nlon, nf) : AMDGPU.zeros(OT, cfg.nlat, cfg.nlon)
    _check(:synthesis, ccall((:shtk_synthesis_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(a), pointer(f), nf, _dtype(RT), real_output ? 1 : 0, cfg.lmax, _precision(precision),
        _stream()))
    AMDGPU.synchronize()                 # results ready on return, as the AMDGPU adapter
    return f
end

function _gpu_adapter_analysis(::MI355XAdapter, cfg::SHTConfig, f::AbstractArray;
                               use_rfft::Bool=false, fft_scratch=nothing, precision::Symbol=:fp64)
    fft_scratch === nothing || throw(ArgumentError("MI355X transforms do not accept a host fft_scratch"))
    size(f)[1:2] == (cfg.nlat, cfg.nlon) || throw(DimensionMismatch("field size $(size(f))"))
    RT = float(real(eltype(f)))
    nf = ndims(f) == 3 ? size(f, 3) : 1
    x = ROCArray{RT}(f)
    alm = ndims(f) == 3 ? AMDGPU.zeros(Complex{RT}, cfg.lmax + 1, cfg.mmax + 1, nf) :
                          AMDGPU.zeros(Complex{RT}, cfg.lmax + 1, cfg.mmax + 1)
    _check(:analysis, ccall((:shtk_analysis_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(x), pointer(alm), nf, _dtype(RT), cfg.lmax, _precision(precision), _stream()))
    AMDGPU.synchronize()
    return alm
end

function _gpu_adapter_synthesis_sphtor(::MI355XAdapter, cfg::SHTConfig, S, T;
                                       real_output::Bool=true, use_rfft::Bool=false, precision::Symbol=:fp64)
    RT = float(real(eltype(S)))
    OT = real_output ? RT : Complex{RT}   # complex grids of the bins m >= 0
    nf = ndims(S) == 3 ? size(S, 3) : 1
    s = ROCArray{Complex{RT}}(S); t = ROCArray{Complex{RT}}(T)
    dims = ndims(S) == 3 ? (cfg.nlat, cfg.nlon, nf) : (cfg.nlat, cfg.nlon)
    Vt = AMDGPU.zeros(OT, dims...); Vp = AMDGPU.zeros(OT, dims...)
    _check(:synthesis_sphtor, ccall((:shtk_synthesis_sphtor_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(s), pointer(t), pointer(Vt), pointer(Vp), nf, _dtype(RT), real_output ? 1 : 0, cfg.lmax,
        _precision(precision), _stream()))
    AMDGPU.synchronize()
    return Vt, Vp
end

function _gpu_adapter_analysis_sphtor(::MI355XAdapter, cfg::SHTConfig, Vt, Vp; use_rfft::Bool=false,
                                      precision::Symbol=:fp64)
    RT = float(real(eltype(Vt)))
    nf = ndims(Vt) == 3 ? size(Vt, 3) : 1
    vt = ROCArray{RT}(Vt); vp = ROCArray{RT}(Vp)
    dims = ndims(Vt) == 3 ? (cfg.lmax + 1, cfg.mmax + 1, nf) : (cfg.lmax + 1, cfg.mmax + 1)
    S = AMDGPU.zeros(Complex{RT}, dims...); T = AMDGPU.zeros(Complex{RT}, dims...)
    _check(:analysis_sphtor, ccall((:shtk_analysis_sphtor_ex, LIB), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Cvoid}),
        _plan(cfg), pointer(vt), pointer(vp), pointer(S), pointer(T), nf, _dtype(RT), cfg.lmax,
        _precision(precision), _stream()))
    AMDGPU.synchronize()
    return S, T
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

# Batch entry points have no protocol hook in core; the vendor extension adds
# direct overloads (as ext/SHTnsKitAMDGPUExt.jl:1865-1994).  The C calls are the
# same with nfields = size(x, 3); `precision = :split` selects the int8-matrix-core
# contraction for many fields of one configuration.
SHTnsKit.analysis_batch(cfg::SHTConfig, f::ROCArray{T,3}; kwargs...) where {T} =
    _gpu_adapter_analysis(MI355X_ADAPTER, cfg, f; kwargs...)
SHTnsKit.synthesis_batch(cfg::SHTConfig, alm::ROCArray{T,3}; kwargs...) where {T} =
    _gpu_adapter_synthesis(MI355X_ADAPTER, cfg, alm; kwargs...)
SHTnsKit.analysis_sphtor_batch(cfg::SHTConfig, Vt::ROCArray{T,3}, Vp::ROCArray{T,3}; kwargs...) where {T} =
    _gpu_adapter_analysis_sphtor(MI355X_ADAPTER, cfg, Vt, Vp; kwargs...)
SHTnsKit.synthesis_sphtor_batch(cfg::SHTConfig, S::ROCArray{T,3}, Tl::ROCArray{T,3}; kwargs...) where {T} =
    _gpu_adapter_synthesis_sphtor(MI355X_ADAPTER, cfg, S, Tl; kwargs...)

end # module
