#
# ChainRulesCore rules of OctopusHIP: build_edges_diff (the reference's
# ext/OctopusChainRulesCoreExt.jl:103-192, ext/OctopusCUDAChainRulesCoreExt.jl:
# 118-167) with the HIP adjoint kernel as its pullback, and apply_zsort
# (ext/OctopusChainRulesCoreExt.jl:65-94) with the library's scatter.
#
module OctopusHIPChainRulesCoreExt

using ChainRulesCore
using OctopusHIP
using OctopusHIP: TNS, build_edges, build_edges_diff_pullback!, apply_zsort, apply_zsort_adjoint

# Primal: fresh copies of the pair's edges, so the pullback closes over values
# that a later build on the pair cannot overwrite (the reference snapshots the
# same four arrays).
function OctopusHIP.build_edges_diff(coords::AbstractMatrix{<:Union{Float32,Float64}}, t::TNS, id::Integer,
                                     radius::Real;
                                     mode::Symbol=:exact)
    e = build_edges(t, id, id; mode=mode)
    return (senders = copy(e.senders), receivers = copy(e.receivers),
            rel_displacement = copy(e.rel_displacement), rel_dist_norm = copy(e.rel_dist_norm))
end

# One tangent field from whatever the AD system passes (a Tangent, a
# NamedTuple, a zero or nothing); `nothing` stands for a zero tangent, which
# the library takes as a NULL pointer.
_field(::Nothing, ::Symbol) = nothing
_field(::AbstractZero, ::Symbol) = nothing
function _field(d, name::Symbol)
    v = hasproperty(d, name) ? getproperty(d, name) : nothing
    (v === nothing || v isa AbstractZero) && return nothing
    return unthunk(v)
end

# A tangent in the layout of the primal output (a dense device array of the
# same type), converting lazy or broadcasted tangents.
_dense(::Nothing, like) = nothing
_dense(x, like) = x isa typeof(like) ? x : copyto!(similar(like), x)

function ChainRulesCore.rrule(::typeof(OctopusHIP.build_edges_diff), coords::AbstractMatrix{<:Union{Float32,Float64}},
                              t::TNS, id::Integer, radius::Real; mode::Symbol=:exact)
    e = OctopusHIP.build_edges_diff(coords, t, id, radius; mode=mode)
    function build_edges_diff_pullback(d)
        dd = _dense(_field(d, :rel_displacement), e.rel_displacement)
        dn = _dense(_field(d, :rel_dist_norm), e.rel_dist_norm)
        grad = similar(coords)  # overwritten (zeros when both tangents are zero)
        build_edges_diff_pullback!(t, size(coords, 2), e, dd, dn, radius, grad)
        return (NoTangent(), grad, NoTangent(), NoTangent(), NoTangent())
    end
    return e, build_edges_diff_pullback
end

function ChainRulesCore.rrule(::typeof(apply_zsort), t::TNS, id::Integer, A::AbstractArray)
    y = apply_zsort(t, id, A)
    function apply_zsort_pullback(d)
        dy = unthunk(d)
        dy isa AbstractZero && return (NoTangent(), NoTangent(), NoTangent(), ZeroTangent())
        return (NoTangent(), NoTangent(), NoTangent(), apply_zsort_adjoint(t, id, _dense(dy, y)))
    end
    return y, apply_zsort_pullback
end

end # module
