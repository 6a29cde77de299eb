#
# OctopusHIP: the public TNS API of Octopus.jl (src/Octopus.jl:31-44) over the
# C ABI of liboctohip.so (include/octohip.h).  Same names, argument meaning and
# exception types as the reference; the backend symbol is :hip and every
# computation runs in the HIP library on one MI355X.
#
# Coordinates are device arrays of Float32 with 3 rows (column-major (3, N),
# i.e. xyz interleaved), referenced and never copied (src/types.jl:13).  The
# binding only needs a raw device pointer from them, so any device array type
# that converts to Ptr{Float32} works (AMDGPU.jl's ROCArray does via
# `pointer`).  Results come back as raw device pointers plus sizes; wrap them
# with `unsafe_wrap(ROCArray, ptr, dims)` (the library keeps owning them).
#
module OctopusHIP

export TNS, set_search_radius!, set_refit_mode!, add_point_set!, update_point_set!,
       resize_point_set!, set_active_search!, run!, build_edges!, build_edges, device_view,
       for_each_neighbor, build_edges_diff

const libocto = joinpath(@__DIR__, "..", "octohip", "liboctohip.so")

const OCT_OK = Cint(0)
const OCT_ERR_ARGUMENT = Cint(-1)
const OCT_ERR_DIMENSION = Cint(-2)
const OCT_ERR_STATE = Cint(-3)

struct OctEdges
    senders::Ptr{Int32}
    receivers::Ptr{Int32}
    rel_displacement::Ptr{Float32}
    rel_dist_norm::Ptr{Float32}
    n_edges::Int64
end

struct OctView            # field order and sizes of oct_view
    coords_q::Ptr{Float32}
    coords_t::Ptr{Float32}
    bounds_min::Ptr{Float32}
    bounds_max::Ptr{Float32}
    node_first::Ptr{Int32}
    node_last::Ptr{Int32}
    node_children::Ptr{Int32}
    perm::Ptr{Int32}
    nodes::Ptr{Cvoid}
    sorted_t::Ptr{Float32}
    n_nodes::Int32
    root::Int32
    radius::Float32
    radius_sq::Float32
    n_q::Int64
    index_base::Int32
    pad::Int32
end

mutable struct TNS
    ctx::Ptr{Cvoid}
    radius::Float32
    coords::Vector{Any}          # keeps user arrays alive (referenced, not copied)
    active_pairs::Vector{Tuple{Int32,Int32}}
    function TNS(::Type{Float32}=Float32; ndims::Int=3, device::Int=0)
        ndims == 3 || throw(ArgumentError("the HIP engine builds octrees (ndims = 3)"))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:oct_create, libocto), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint), h, device, 3, 1)
        rc == OCT_OK || error("oct_create failed ($rc)")
        t = new(h[], 0f0, Any[], Tuple{Int32,Int32}[])
        finalizer(x -> ccall((:oct_destroy, libocto), Cint, (Ptr{Cvoid},), x.ctx), t)
        return t
    end
end

function _check(t::TNS, rc::Cint)
    rc == OCT_OK && return nothing
    msg = unsafe_string(ccall((:oct_last_error, libocto), Cstring, (Ptr{Cvoid},), t.ctx))
    rc == OCT_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == OCT_ERR_DIMENSION && throw(DimensionMismatch(msg))
    error(msg)
end

function set_search_radius!(t::TNS, r::Real)
    r > 0 || throw(ArgumentError("search radius must be positive"))
    t.radius = Float32(r)
    _check(t, ccall((:oct_set_search_radius, libocto), Cint, (Ptr{Cvoid}, Cfloat), t.ctx, t.radius))
    return t
end

function set_refit_mode!(t::TNS, on::Bool)
    _check(t, ccall((:oct_set_refit_mode, libocto), Cint, (Ptr{Cvoid}, Cint), t.ctx, on))
    return t
end

function add_point_set!(t::TNS, coords::AbstractMatrix{Float32})
    size(coords, 1) == 3 || throw(DimensionMismatch("coords must be (3, N)"))
    id = Ref{Cint}(0)
    _check(t, ccall((:oct_add_point_set, libocto), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Cint}),
                    t.ctx, pointer(coords), size(coords, 2), id))
    push!(t.coords, coords)
    return Int(id[])
end

function update_point_set!(t::TNS, id::Integer, coords::AbstractMatrix{Float32})
    size(coords, 1) == 3 || throw(DimensionMismatch("coords must be (3, N)"))
    _check(t, ccall((:oct_update_point_set, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, Int64),
                    t.ctx, id, pointer(coords), size(coords, 2)))
    t.coords[id] = coords
    return t
end

function resize_point_set!(t::TNS, id::Integer, coords::AbstractMatrix{Float32})
    size(coords, 1) == 3 || throw(DimensionMismatch("coords must be (3, N)"))
    _check(t, ccall((:oct_resize_point_set, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, Int64),
                    t.ctx, id, pointer(coords), size(coords, 2)))
    t.coords[id] = coords
    return t
end

function set_active_search!(t::TNS, q::Integer, s::Integer)
    _check(t, ccall((:oct_set_active_search, libocto), Cint, (Ptr{Cvoid}, Cint, Cint), t.ctx, q, s))
    (Int32(q), Int32(s)) in t.active_pairs || push!(t.active_pairs, (Int32(q), Int32(s)))
    return t
end

function run!(t::TNS)
    _check(t, ccall((:oct_run, libocto), Cint, (Ptr{Cvoid},), t.ctx))
    return t
end

"""
    build_edges!(tns, q, t; mode=:exact, order=:leaf) -> OctEdges

Raw device pointers + count of the pair's flat-COO buffers (owned by the
context, valid until the next build on the pair).
"""
function build_edges!(t::TNS, q::Integer, s::Integer; mode::Symbol=:exact, order::Symbol=:leaf)
    out = Ref{OctEdges}()
    m = mode === :learned ? 1 : 0
    o = order === :receiver ? 1 : 0
    _check(t, ccall((:oct_build_edges, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{OctEdges}),
                    t.ctx, q, s, m, o, out))
    return out[]
end

# Needs a device array type to wrap the pointers; AMDGPU.jl users pass ROCArray.
function build_edges(t::TNS, q::Integer, s::Integer; array_type=nothing, kw...)
    e = build_edges!(t, q, s; kw...)
    array_type === nothing && return e
    n = Int(e.n_edges)
    return (senders = unsafe_wrap(array_type, e.senders, n),
            receivers = unsafe_wrap(array_type, e.receivers, n),
            rel_displacement = unsafe_wrap(array_type, e.rel_displacement, (3, n)),
            rel_dist_norm = unsafe_wrap(array_type, e.rel_dist_norm, (1, n)))
end

function device_view(t::TNS, q::Integer, s::Integer)
    v = Ref{OctView}()
    _check(t, ccall((:oct_device_view, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{OctView}), t.ctx, q, s, v))
    return v[]
end

for_each_neighbor(f, t::TNS, q, s, i) =
    error("for_each_neighbor (host) requires a CPU TNS; use for_each_neighbor_device on GPU")

"""
    build_edges_diff_pullback!(tns, n_points, edges::OctEdges, d_disp, d_dist, radius, grad)

The HIP adjoint behind the `build_edges_diff` rrule
(ext/OctopusChainRulesCoreExt.jl:137-192): device pointers in, grad (3, N)
overwritten.  A ChainRulesCore `rrule` wraps this exactly as the reference's
extension does: forward through `build_edges`, pullback through this call.
"""
function build_edges_diff_pullback!(t::TNS, n_points::Integer, e::OctEdges, d_disp, d_dist, radius::Real,
                                    grad)
    _check(t, ccall((:oct_build_edges_backward, libocto), Cint,
                    (Ptr{Cvoid}, Int64, Ref{OctEdges}, Ptr{Float32}, Ptr{Float32}, Cfloat, Ptr{Float32}),
                    t.ctx, n_points, e, d_disp === nothing ? C_NULL : pointer(d_disp),
                    d_dist === nothing ? C_NULL : pointer(d_dist), Float32(radius), pointer(grad)))
    return grad
end

function build_edges_diff end  # the rrule lives in a ChainRulesCore package extension, not here

end # module
