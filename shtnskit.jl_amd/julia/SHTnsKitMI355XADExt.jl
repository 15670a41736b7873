# SHTnsKitMI355XADExt -- reverse-mode rules of the batch transforms for dense
# ROCArrays on an MI355X (loaded with AMDGPU.jl and ChainRulesCore.jl).
#
# The single-field rrules of ext/SHTnsKitAdvancedADExt.jl:55-89, 190-283 need
# nothing from here: they call SHTnsKit._adjoint_synthesis / _adjoint_analysis /
# _adjoint_*_sphtor, whose device methods (SHTnsKitMI355XExt) run the HIP adjoint
# kernels.  Its batch rrules (:95-133) loop the scalar adjoint per field into host
# arrays; these methods, more specific in the array type, pull a whole batch back
# through one call of the batched HIP adjoint instead, with the same result (the
# batch adjoint is the per-field adjoint of every slice).
module SHTnsKitMI355XADExt

using AMDGPU
using ChainRulesCore
using SHTnsKit
using SHTnsKit: SHTConfig

# the transform extension (loaded with AMDGPU.jl), looked up when a rule runs
_mi() = Base.get_extension(SHTnsKit, :SHTnsKitMI355XExt)

_zero_tangents() = (NoTangent(), NoTangent(), ZeroTangent())

function ChainRulesCore.rrule(::typeof(SHTnsKit.synthesis_batch), cfg::SHTConfig,
                              alm::ROCArray{<:Complex,3}; real_output::Bool=true, use_rfft::Bool=false,
                              precision::Symbol=:fp64)
    y = SHTnsKit.synthesis_batch(cfg, alm; real_output, use_rfft, precision)
    project = ProjectTo(alm)
    function synthesis_batch_pullback(ȳ)
        ȳ = unthunk(ȳ)
        ȳ isa ChainRulesCore.AbstractZero && return _zero_tangents()
        MI = _mi()
        MI._on_mi355x(alm) || throw(SHTnsKit.BackendUnavailableError(:synthesis_batch,
            "the MI355X batch adjoint needs arrays on an MI355X"))
        ā = SHTnsKit._with_gpu_device(alm) do
            MI._adjoint_synthesis_dev(cfg, MI._dense(ȳ), real_output)
        end
        return NoTangent(), NoTangent(), project(ā)
    end
    return y, synthesis_batch_pullback
end

function ChainRulesCore.rrule(::typeof(SHTnsKit.analysis_batch), cfg::SHTConfig,
                              fields::ROCArray{<:Real,3}; use_rfft::Bool=false, precision::Symbol=:fp64)
    y = SHTnsKit.analysis_batch(cfg, fields; use_rfft, precision)
    project = ProjectTo(fields)
    function analysis_batch_pullback(ȳ)
        ȳ = unthunk(ȳ)
        ȳ isa ChainRulesCore.AbstractZero && return _zero_tangents()
        MI = _mi()
        MI._on_mi355x(fields) || throw(SHTnsKit.BackendUnavailableError(:analysis_batch,
            "the MI355X batch adjoint needs arrays on an MI355X"))
        f̄ = SHTnsKit._with_gpu_device(fields) do
            MI._adjoint_analysis_dev(cfg, MI._dense(ȳ))
        end
        return NoTangent(), NoTangent(), project(f̄)   # a real primal keeps the real part
    end
    return y, analysis_batch_pullback
end

end # module
