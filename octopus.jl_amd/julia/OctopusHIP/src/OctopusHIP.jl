#
# OctopusHIP: the public TNS API of Octopus.jl (src/Octopus.jl:31-44) over the
# C ABI of liboctohip.so (include/octohip.h).  Same names, argument meaning and
# exception types as the reference; the backend symbol is :hip and every
# computation runs in the HIP library on one MI355X.
#
# Coordinates are device arrays of the TNS's element type (Float32 or
# Float64) with D = 3 or 2 rows (column-major (D, N), i.e. xyz interleaved),
# referenced and never copied (src/types.jl:13).  The binding needs only a raw
# device pointer from them (`pointer(A)`); the C ABI types such pointers as
# float*, and a Float64 context reads them as double (oct_create_typed).
#
# Device arrays and streams come from a GPU array package through two hooks
# that the AMDGPU extension (ext/OctopusHIPAMDGPUExt.jl) fills:
#   _device_wrap(ptr, dims, owner): a device array over library memory that
#       keeps `owner` (the TNS) alive as long as the array lives;
#   _current_stream(): the task-local HIP stream every call is ordered on.
# Without them results are raw pointers (build_edges!, get_tree!) and each
# call that returns them waits for its kernels (oct_synchronize).
#
module OctopusHIP

export TNS, set_search_radius!, set_refit_mode!, add_point_set!, update_point_set!,
       resize_point_set!, set_active_search!, run!, build_edges!, build_edges, device_view,
       for_each_neighbor, build_edges_diff, materialize_all_neighbors!, get_neighborlist,
       prepare_zsort, apply_zsort, get_tree

const libocto = get(ENV, "OCTOHIP_LIB", joinpath(@__DIR__, "..", "..", "..", "octohip", "liboctohip.so"))

const OCT_OK = Cint(0)
const OCT_ERR_ARGUMENT = Cint(-1)
const OCT_ERR_DIMENSION = Cint(-2)
const OCT_ERR_STATE = Cint(-3)

const OCT_EDGES_EXACT = Cint(0)
const OCT_EDGES_LEARNED = Cint(1)
const OCT_ORDER_LEAF = Cint(0)
const OCT_ORDER_RECEIVER = Cint(1)

# ---- C structs (field order and sizes of include/octohip.h) ----------------

struct OctEdges
    senders::Ptr{Int32}
    receivers::Ptr{Int32}
    rel_displacement::Ptr{Float32}
    rel_dist_norm::Ptr{Float32}
    n_edges::Int64
end

struct OctTree
    node_bounds_min::Ptr{Float32}
    node_bounds_max::Ptr{Float32}
    node_first::Ptr{Int32}
    node_last::Ptr{Int32}
    node_children::Ptr{Int32}
    permutation::Ptr{Int32}
    n_nodes::Int32
    n_leaves::Int32
    n_points::Int64
    origin::NTuple{3,Float32}
end

struct OctView
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

struct OctNeighbors
    offsets::Ptr{Int32}
    flat::Ptr{Int32}
    n_queries::Int64
    total::Int64
    materialized::Int32
    pad::Int32
end

# ---- hooks filled by a GPU array package extension ---------------------------

"""Device array over library-owned memory, kept valid (and `owner` alive) as
long as the array lives.  Defined by the AMDGPU extension."""
function _device_wrap end
_device_wrap(ptr::Ptr, dims, owner) =
    error("OctopusHIP: device arrays need a GPU array package (`using AMDGPU`); use build_edges! for raw pointers")

"""The HIP stream (hipStream_t) the caller's task works on, or `nothing`."""
_current_stream() = nothing

# ---- the context ------------------------------------------------------------

mutable struct TNS
    ctx::Ptr{Cvoid}
    eltype::DataType             # Float32 or Float64 (TNS(T), src/types.jl:167-183)
    ndims::Int                   # 3 (octrees) or 2 (quadtrees)
    radius::Float64
    coords::Vector{Any}          # keeps user arrays alive (referenced, not copied)
    dirty::Vector{Bool}
    active_pairs::Vector{Tuple{Int32,Int32}}
    function TNS(::Type{T}=Float32; ndims::Int=3, device::Int=0) where {T<:Union{Float32,Float64}}
        ndims in (2, 3) || throw(ArgumentError("ndims must be 2 or 3"))
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:oct_create_typed, libocto), Cint, (Ptr{Ptr{Cvoid}}, Cint, Cint, Cint, Cint), h, device, ndims, 1,
                   sizeof(T))
        rc == OCT_OK || error("oct_create failed ($rc)")
        t = new(h[], T, ndims, 0.0, Any[], Bool[], Tuple{Int32,Int32}[])
        # every device array handed out keeps `t` reachable (see _device_wrap),
        # so the context is destroyed only after the last of them is gone
        finalizer(x -> ccall((:oct_destroy, libocto), Cint, (Ptr{Cvoid},), x.ctx), t)
        return t
    end
end
TNS(::Type{T}; kw...) where {T} = throw(MethodError(TNS, (T,)))  # Float32 and Float64 only

function _check(t::TNS, rc::Cint)
    rc == OCT_OK && return nothing
    msg = unsafe_string(ccall((:oct_last_error, libocto), Cstring, (Ptr{Cvoid},), t.ctx))
    rc == OCT_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == OCT_ERR_DIMENSION && throw(DimensionMismatch(msg))
    error(msg)
end

"""Orders the library's work on `stream` (a hipStream_t) from now on."""
function set_stream!(t::TNS, stream::Ptr{Cvoid})
    _check(t, ccall((:oct_set_stream, libocto), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), t.ctx, stream))
    return t
end

# Before a call: follow the caller's current stream (free while unchanged).
function _enter(t::TNS)
    s = _current_stream()
    s === nothing || set_stream!(t, Ptr{Cvoid}(s))
    return t.ctx
end

# After a call that hands out device memory: without a stream to order on,
# wait for the kernels so raw pointers are complete when they are used.
function _leave(t::TNS)
    _current_stream() === nothing &&
        _check(t, ccall((:oct_synchronize, libocto), Cint, (Ptr{Cvoid},), t.ctx))
    return nothing
end

synchronize(t::TNS) = _check(t, ccall((:oct_synchronize, libocto), Cint, (Ptr{Cvoid},), t.ctx))

"""Frees allocations that larger results replaced (arrays taken from them
must no longer be read)."""
trim!(t::TNS) = (_check(t, ccall((:oct_trim, libocto), Cint, (Ptr{Cvoid},), t.ctx)); t)

# ---- setters (src/api.jl:24-165) ----------------------------------------------

function set_search_radius!(t::TNS, r::Real)
    r > 0 || throw(ArgumentError("search radius must be positive"))
    t.radius = Float64(t.eltype(r))  # T(r)
    _check(t, ccall((:oct_set_search_radius64, libocto), Cint, (Ptr{Cvoid}, Cdouble), t.ctx, t.radius))
    fill!(t.dirty, true)
    return t
end

function set_refit_mode!(t::TNS, on::Bool)
    _check(t, ccall((:oct_set_refit_mode, libocto), Cint, (Ptr{Cvoid}, Cint), t.ctx, on))
    fill!(t.dirty, true)
    return t
end

function set_leaf_size!(t::TNS, leaf_size::Integer)
    _check(t, ccall((:oct_set_leaf_size, libocto), Cint, (Ptr{Cvoid}, Cint), t.ctx, leaf_size))
    fill!(t.dirty, true)
    return t
end

function _npoints(t::TNS, coords)
    # an eltype other than the TNS's is a MethodError, as in the reference (test/float64.jl:138-148)
    eltype(coords) === t.eltype || throw(MethodError(add_point_set!, (t, coords)))
    size(coords, 1) == t.ndims || throw(DimensionMismatch("coords must be ($(t.ndims), N)"))
    return size(coords, 2)
end

function add_point_set!(t::TNS, coords::AbstractMatrix{<:Union{Float32,Float64}})
    n = _npoints(t, coords)
    id = Ref{Cint}(0)
    _check(t, ccall((:oct_add_point_set, libocto), Cint, (Ptr{Cvoid}, Ptr{Float32}, Int64, Ptr{Cint}),
                    t.ctx, Ptr{Float32}(pointer(coords)), n, id))
    push!(t.coords, coords)
    push!(t.dirty, true)
    return Int(id[])
end

function update_point_set!(t::TNS, id::Integer, coords::AbstractMatrix{<:Union{Float32,Float64}})
    n = _npoints(t, coords)
    _check(t, ccall((:oct_update_point_set, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, Int64),
                    t.ctx, id, Ptr{Float32}(pointer(coords)), n))
    t.coords[id] = coords
    t.dirty[id] = true
    return t
end

function resize_point_set!(t::TNS, id::Integer, coords::AbstractMatrix{<:Union{Float32,Float64}})
    n = _npoints(t, coords)
    _check(t, ccall((:oct_resize_point_set, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Float32}, Int64),
                    t.ctx, id, Ptr{Float32}(pointer(coords)), n))
    t.coords[id] = coords
    t.dirty[id] = true
    return t
end

function set_active_search!(t::TNS, q::Integer, s::Integer)
    _check(t, ccall((:oct_set_active_search, libocto), Cint, (Ptr{Cvoid}, Cint, Cint), t.ctx, q, s))
    (Int32(q), Int32(s)) in t.active_pairs || push!(t.active_pairs, (Int32(q), Int32(s)))
    return t
end

# ---- run! (src/api.jl:183-250) ----------------------------------------------------

function run!(t::TNS)
    _check(t, ccall((:oct_run, libocto), Cint, (Ptr{Cvoid},), _enter(t)))
    fill!(t.dirty, false)
    return t
end

"""Raw device pointers of set `id`'s tree (tns.trees[id], tns.permutation[id])."""
function get_tree!(t::TNS, id::Integer)
    out = Ref{OctTree}()
    _check(t, ccall((:oct_get_tree, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{OctTree}), _enter(t), id, out))
    _leave(t)
    return out[]
end

"""The tree of set `id` in the reference layout (src/types.jl:23-42) as device
arrays: node_bounds_min/max (D, M), node_first/last (M), node_children
(2^D, M), permutation (N); valid until the next run!."""
function get_tree(t::TNS, id::Integer)
    x = get_tree!(t, id)
    m, n = Int(x.n_nodes), Int(x.n_points)
    d, T = t.ndims, t.eltype
    return (node_bounds_min = _device_wrap(Ptr{T}(x.node_bounds_min), (d, m), t),
            node_bounds_max = _device_wrap(Ptr{T}(x.node_bounds_max), (d, m), t),
            node_first = _device_wrap(x.node_first, (m,), t),
            node_last = _device_wrap(x.node_last, (m,), t),
            node_children = _device_wrap(x.node_children, (1 << d, m), t),
            permutation = _device_wrap(x.permutation, (n,), t),
            origin = x.origin)
end

# ---- build_edges (src/api.jl:304-337) ------------------------------------------

_mode(m::Symbol) = m === :exact ? OCT_EDGES_EXACT : m === :learned ? OCT_EDGES_LEARNED :
                   throw(ArgumentError("mode must be :exact or :learned"))
_order(o::Symbol) = o === :receiver ? OCT_ORDER_RECEIVER : o === :leaf ? OCT_ORDER_LEAF :
                    throw(ArgumentError("order must be :receiver or :leaf"))

"""
    build_edges!(tns, q, t; mode=:exact, order=:receiver) -> OctEdges

Raw device pointers + count of the pair's flat-COO buffers (owned by the
context, valid until the next build on the pair).  order=:receiver writes the
reference's layout (receivers ascending, src/cpu/edges_cpu.jl:209-247) and
needs mode=:exact; mode=:learned (the single pass) writes each receiver's
edges as one run in tree-leaf order (order=:leaf).
"""
function build_edges!(t::TNS, q::Integer, s::Integer; mode::Symbol=:exact,
                      order::Symbol=(mode === :exact ? :receiver : :leaf))
    (Int32(q), Int32(s)) in t.active_pairs ||
        error("build_edges!: pair ($q, $s) not registered; call set_active_search! first")
    out = Ref{OctEdges}()
    _check(t, ccall((:oct_build_edges, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{OctEdges}),
                    _enter(t), q, s, _mode(mode), _order(order), out))
    _leave(t)
    return out[]
end

"""
    build_edges(tns, q, t; mode=:exact, order=...) -> NamedTuple

`(senders, receivers, rel_displacement (D, E), rel_dist_norm (1, E))` as
device arrays aliasing the pair's buffers, as the reference's NamedTuple
aliases its EdgeBuffer (src/api.jl:329-337, test/contracts.jl:223-240): the
next build on the pair overwrites them.
"""
function build_edges(t::TNS, q::Integer, s::Integer; kw...)
    e = build_edges!(t, q, s; kw...)
    n = Int(e.n_edges)
    return (senders = _device_wrap(e.senders, (n,), t),
            receivers = _device_wrap(e.receivers, (n,), t),
            rel_displacement = _device_wrap(Ptr{t.eltype}(e.rel_displacement), (t.ndims, n), t),
            rel_dist_norm = _device_wrap(Ptr{t.eltype}(e.rel_dist_norm), (1, n), t))
end

# ---- neighbour lists (src/api.jl:282-292, 348-362) --------------------------------

"""materialize_all_neighbors!(tns): CSR lists of every active pair, built on
the GPU (the reference builds them on the CPU only, src/api.jl:350)."""
function materialize_all_neighbors!(t::TNS)
    ctx = _enter(t)
    for (q, s) in t.active_pairs
        out = Ref{OctNeighbors}()
        _check(t, ccall((:oct_materialize_neighbors, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{OctNeighbors}),
                        ctx, q, s, out))
    end
    _leave(t)
    return t
end

"""The pair's CSR lists (tns.neighbor_buffers[pair], src/types.jl:44-48) as
device arrays (offsets (n_q + 1), flat), or `nothing` before a materialize."""
function neighbor_buffer(t::TNS, q::Integer, s::Integer)
    out = Ref{OctNeighbors}()
    _check(t, ccall((:oct_get_neighbors, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{OctNeighbors}),
                    t.ctx, q, s, out))
    nb = out[]
    nb.materialized == 0 && return nothing
    return (offsets = _device_wrap(nb.offsets, (Int(nb.n_queries) + 1,), t),
            flat = _device_wrap(nb.flat, (Int(nb.total),), t))
end

"""get_neighborlist(tns, q, t, i) -> Vector{Int32}: query i's neighbours
(self excluded for the self pair), from the pair's materialized lists."""
function get_neighborlist(t::TNS, q::Integer, s::Integer, i::Integer)
    nb = neighbor_buffer(t, q, s)
    if nb === nothing
        materialize_all_neighbors!(t)
        nb = neighbor_buffer(t, q, s)
    end
    lo, hi = Array(nb.offsets[i:i+1])
    return Array(nb.flat[lo+1:hi])
end

for_each_neighbor(f, t::TNS, q, s, i) =
    error("for_each_neighbor (host) requires a CPU TNS; use for_each_neighbor_device on GPU")

# ---- device view (src/api.jl:451-454) ----------------------------------------------

function device_view(t::TNS, q::Integer, s::Integer)
    v = Ref{OctView}()
    _check(t, ccall((:oct_device_view, libocto), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{OctView}), _enter(t), q, s, v))
    return v[]
end

# ---- z-sort (src/api.jl:381-410) ------------------------------------------------------

function prepare_zsort(t::TNS)
    for id in eachindex(t.dirty)
        t.dirty[id] && error("prepare_zsort: call run! first so the permutation for set $id is available")
    end
    return t
end

function _zsort(t::TNS, id::Integer, A::AbstractArray, adjoint::Bool)
    ndims(A) <= 2 || throw(ArgumentError("apply_zsort supports 1-D or 2-D arrays only; got $(ndims(A))-D"))
    n = size(t.coords[id], 2)
    size(A, ndims(A)) == n ||
        throw(DimensionMismatch("perm length $n does not match last axis $(size(A, ndims(A)))"))
    out = similar(A)
    width = ndims(A) == 1 ? 1 : size(A, 1)
    rc = adjoint ?
        ccall((:oct_apply_zsort_adjoint, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint),
              _enter(t), id, pointer(A), pointer(out), width, sizeof(eltype(A))) :
        ccall((:oct_apply_zsort, libocto), Cint, (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint),
              _enter(t), id, pointer(A), pointer(out), width, sizeof(eltype(A)))
    _check(t, rc)
    return out
end

"""apply_zsort(tns, id, A): a new device array, A's columns (or entries)
in the z-order permutation of set `id` (src/api.jl:405-410)."""
apply_zsort(t::TNS, id::Integer, A::AbstractArray) = _zsort(t, id, A, false)

"""The adjoint of apply_zsort: grad[:, perm[k]] = dy[:, k] (the rrule's
scatter, ext/OctopusChainRulesCoreExt.jl:85-94)."""
apply_zsort_adjoint(t::TNS, id::Integer, dy::AbstractArray) = _zsort(t, id, dy, true)

# ---- build_edges_diff (src/Octopus.jl:56-73) ------------------------------------------

"""
    build_edges_diff_pullback!(tns, n_points, e, d_disp, d_dist, radius, grad)

The HIP adjoint behind the `build_edges_diff` rrule
(ext/OctopusChainRulesCoreExt.jl:137-192): `e` is the NamedTuple the primal
returned, d_disp (3, E) / d_dist (1, E) device arrays or `nothing` (a zero
tangent), grad (3, N) overwritten with dL/dcoords.
"""
function build_edges_diff_pullback!(t::TNS, n_points::Integer, e, d_disp, d_dist, radius::Real, grad)
    edges = OctEdges(pointer(e.senders), pointer(e.receivers), Ptr{Float32}(pointer(e.rel_displacement)),
                     Ptr{Float32}(pointer(e.rel_dist_norm)), Int64(length(e.senders)))
    dd = d_disp === nothing ? C_NULL : pointer(d_disp)
    dn = d_dist === nothing ? C_NULL : pointer(d_dist)
    if t.eltype === Float64
        _check(t, ccall((:oct_build_edges_backward64, libocto), Cint,
                        (Ptr{Cvoid}, Int64, Ref{OctEdges}, Ptr{Float64}, Ptr{Float64}, Cdouble, Ptr{Float64}),
                        _enter(t), n_points, edges, dd, dn, Float64(radius), pointer(grad)))
    else
        _check(t, ccall((:oct_build_edges_backward, libocto), Cint,
                        (Ptr{Cvoid}, Int64, Ref{OctEdges}, Ptr{Float32}, Ptr{Float32}, Cfloat, Ptr{Float32}),
                        _enter(t), n_points, edges, dd, dn, Float32(radius), pointer(grad)))
    end
    return grad
end

"""
    build_edges_diff(coords, tns, id, radius; mode=:exact) -> NamedTuple

Differentiable copies of `build_edges(tns, id, id)`; the primal and its rrule
live in the ChainRulesCore extension (ext/OctopusHIPChainRulesCoreExt.jl), as
in the reference (ext/OctopusChainRulesCoreExt.jl:103-192).
"""
function build_edges_diff end

end # module
