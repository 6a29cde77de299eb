"""octohip: Python front-end of the MI355X fixed-radius neighbour search engine.

It mirrors the public API of una-auxme/Octopus.jl (src/Octopus.jl:31-44) —
TNS, set_search_radius, set_refit_mode, add_point_set, update_point_set,
resize_point_set, set_active_search, run, build_edges, build_edges_diff,
device_view, for_each_neighbor — with the same argument meaning and error
types, over the C ABI of liboctohip.so (hand-written HIP kernels for gfx950).
Julia's ``!`` suffix is dropped; the TNS methods carry the same names.

Coordinates are (N, 3) float32 CUDA tensors (one point per row, the bytes of
the reference's column-major (3, N) matrix); they are referenced, never copied
(src/types.jl:13).  torch supplies device memory and the stream; every
computation runs in the library.  Indices are 0-based by default
(index_base=1 reproduces the reference's 1-based values).
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

from . import _lib as L

__all__ = [
    "TNS", "ArgumentError", "DimensionMismatch", "OctopusError",
    "set_search_radius", "set_refit_mode", "set_leaf_size", "add_point_set", "update_point_set",
    "resize_point_set", "set_active_search", "run", "build_edges", "build_edges_diff",
    "device_view", "for_each_neighbor", "count_neighbors_device", "get_tree", "trim", "retired_count",
    "materialize_all_neighbors", "neighbors", "get_neighborlist", "prepare_zsort", "apply_zsort",
]


class ArgumentError(ValueError):
    """Julia ArgumentError (src/api.jl:25,63-65,184,190)."""


class DimensionMismatch(ValueError):
    """Julia DimensionMismatch (src/api.jl:97,137-138)."""


class OctopusError(RuntimeError):
    """Julia ErrorException (src/api.jl:306,452, src/cpu/iterator_cpu.jl:246)."""


Edges = namedtuple("Edges", ["senders", "receivers", "rel_displacement", "rel_dist_norm"])
Tree = namedtuple("Tree", ["node_bounds_min", "node_bounds_max", "node_first", "node_last",
                           "node_children", "permutation", "n_nodes", "n_leaves", "origin"])

_MODES = {"exact": L.OCT_EDGES_EXACT, "learned": L.OCT_EDGES_LEARNED}
_ORDERS = {"leaf": L.OCT_ORDER_LEAF, "receiver": L.OCT_ORDER_RECEIVER}


def _torch():
    import torch

    return torch


# DLPack (dlpack.h, v0.8 ABI): a capsule of a DLManagedTensor that names its
# device, so torch takes library-owned memory without asking the runtime
# which device a pointer belongs to.  (torch.as_tensor on a
# __cuda_array_interface__ does ask, and that query was seen to wait for the
# kernels still writing the memory: up to 0.4 ms per build_edges call.)
class _DLDevice(ctypes.Structure):
    _fields_ = [("device_type", ctypes.c_int), ("device_id", ctypes.c_int)]


class _DLDataType(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint8), ("bits", ctypes.c_uint8), ("lanes", ctypes.c_uint16)]


class _DLTensor(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("device", _DLDevice), ("ndim", ctypes.c_int),
                ("dtype", _DLDataType), ("shape", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("byte_offset", ctypes.c_uint64)]


class _DLManagedTensor(ctypes.Structure):
    pass


# The deleter takes the struct's address (a pointer argument, so the C ABI
# is dlpack.h's).
_DL_DELETER = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
_DLManagedTensor._fields_ = [("dl_tensor", _DLTensor), ("manager_ctx", ctypes.c_void_p),
                             ("deleter", _DL_DELETER)]
_KDL_ROCM = 10
_DL_TYPES = {"<i4": (0, 32), "<i8": (0, 64), "<f4": (2, 32), "<f8": (2, 64)}  # (kDLInt | kDLFloat, bits)
_dl_live = {}  # address of a handed-out DLManagedTensor -> (struct, shape, owner) until torch deletes it


def _make_dl_deleter(live):
    @_DL_DELETER
    def delete(addr):
        live.pop(addr, None)

    return delete


_dl_delete = _make_dl_deleter(_dl_live)
# A tensor can outlive this module: at interpreter exit a caller's global is
# freed after the module's dict has been cleared.  torch then reads the
# struct (kept by _dl_live) and calls the deleter, so neither may be freed
# before the process ends: both get a reference nothing releases.  (The
# deleter reaches the dict through its closure, not a module global.)
ctypes.pythonapi.Py_IncRef(ctypes.py_object(_dl_live))
ctypes.pythonapi.Py_IncRef(ctypes.py_object(_dl_delete))


_PyCapsule_New = ctypes.pythonapi.PyCapsule_New
_PyCapsule_New.restype = ctypes.py_object
_PyCapsule_New.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p]


def _dl_capsule(ptr, shape, typestr, device_id, owner):
    """Zero-copy DLPack capsule of library-owned device memory (row-major).
    Until torch calls the deleter, _dl_live keeps the struct, its shape array
    and `owner` (the TNS or batch whose context owns the memory) alive, so
    the context outlives every view of its buffers."""
    m = _DLManagedTensor()
    shp = (ctypes.c_int64 * len(shape))(*shape)
    code, bits = _DL_TYPES[typestr]
    m.dl_tensor.data = int(ptr)
    m.dl_tensor.device = _DLDevice(_KDL_ROCM, int(device_id))
    m.dl_tensor.ndim = len(shape)
    m.dl_tensor.dtype = _DLDataType(code, bits, 1)
    m.dl_tensor.shape = shp
    m.dl_tensor.strides = None
    m.dl_tensor.byte_offset = 0
    m.deleter = _dl_delete
    _dl_live[ctypes.addressof(m)] = (m, shp, owner)
    return _PyCapsule_New(ctypes.addressof(m), b"dltensor", None)


def _wrap(ptr, shape, dtype, device, owner):
    torch = _torch()
    n = 1
    for s in shape:
        n *= s
    if n == 0 or not ptr:
        return torch.empty(shape, dtype=dtype, device=device)
    typestr = {torch.int32: "<i4", torch.float32: "<f4", torch.int64: "<i8", torch.float64: "<f8"}[dtype]
    index = device.index if device.index is not None else torch.cuda.current_device()
    return torch.from_dlpack(_dl_capsule(ptr, shape, typestr, index, owner))


class _Core:
    """Owns the C context.  Library memory views hold this object rather
    than their TNS, so the context (and the memory) lives until the TNS and
    every such view are gone, and a TNS may cache views without a cycle."""

    __slots__ = ("lib", "h")

    def __init__(self, lib, h):
        self.lib = lib
        self.h = h

    def __del__(self):
        if self.h is not None:
            try:
                self.lib.oct_destroy(self.h)
            except Exception:
                pass
            self.h = None


class TNS:
    """TNS(T = Float32; ndims = 3) (src/types.jl:167-183) on one HIP device.

    The C context is created on first use, so argument checks that need no
    device behave as in the reference even without a GPU."""

    def __init__(self, dtype=None, ndims: int = 3, device=None, index_base: int = 0,
                 private_stream: bool = False):
        torch = _torch()
        dtype = torch.float32 if dtype is None else dtype
        if dtype not in (torch.float32, torch.float64):
            raise ArgumentError("TNS(T): T must be Float32 or Float64")
        if ndims not in (2, 3):
            raise ArgumentError("ndims must be 2 (quadtrees) or 3 (octrees)")
        if index_base not in (0, 1):
            raise ArgumentError("index_base must be 0 or 1")
        self.dtype = dtype
        self.ndims = ndims
        self.index_base = index_base
        self.device = torch.device("cuda", 0 if device is None else int(device))
        self.radius = 0.0
        self.refit_mode = False
        self.target_leaf_size = 32
        self.point_sets = []        # coords tensors, kept alive (referenced, not copied)
        self.active_pairs = []
        self.dirty = []
        self._core = None
        self._views = {}  # (pair, field) -> (address, view of the buffer): build_edges slices these
        self._stream = None
        self._private_stream = bool(private_stream)
        self._lib = L.load()

    # ------------------------------------------------------------ context
    @property
    def _ctx(self):
        return self._core.h if self._core is not None else None

    def _c(self):
        if self._core is None:
            torch = _torch()
            if not torch.cuda.is_available():
                raise OctopusError("no HIP device: the octohip engine runs on an MI355X only")
            h = ctypes.c_void_p()
            rc = self._lib.oct_create_typed(ctypes.byref(h), self.device.index, self.ndims, self.index_base,
                                            8 if self.dtype == torch.float64 else 4)
            if rc != L.OCT_OK:
                raise OctopusError(f"oct_create failed ({rc})")
            self._core = _Core(self._lib, h)
            # By default the library runs on the caller's current stream, like
            # any torch op (set again on every call: free while it is the same
            # stream).  private_stream: its own stream (default priority),
            # ordered after the caller's current stream and back (_enter/_leave).
            if self._private_stream:
                self._stream = torch.cuda.Stream(device=self.device)
                self._check(self._lib.oct_set_stream(h, ctypes.c_void_p(self._stream.cuda_stream)))
            else:
                self._check(self._lib.oct_set_stream(h, self._caller_stream()))
            if self.radius > 0:
                self._set_radius(h)
            if self.target_leaf_size != 32:
                self._check(self._lib.oct_set_leaf_size(h, self.target_leaf_size))
            if self.refit_mode:
                self._check(self._lib.oct_set_refit_mode(h, 1))
        return self._ctx

    def _enter(self):
        """Library work is ordered after everything queued on the caller's stream."""
        ctx = self._c()
        if self._private_stream:
            self._check(self._lib.oct_stream_join(ctx, self._caller_stream(), 0))
        else:
            self._check(self._lib.oct_set_stream(ctx, self._caller_stream()))
        return ctx

    def _leave(self):
        """The caller's stream waits for the library's work (outputs ready)."""
        if self._private_stream:
            self._check(self._lib.oct_stream_join(self._ctx, self._caller_stream(), 1))

    def _caller_stream(self):
        torch = _torch()
        raw = getattr(torch._C, "_cuda_getCurrentRawStream", None)
        if raw is not None:  # the handle without building a Stream object
            return ctypes.c_void_p(raw(self.device.index))
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    @property
    def f64(self):
        return self.dtype == _torch().float64

    def _set_radius(self, h):
        if self.f64:  # T(r) in double
            self._check(self._lib.oct_set_search_radius64(h, self.radius))
        else:
            self._check(self._lib.oct_set_search_radius(h, self.radius))

    def _check(self, rc):
        if rc == L.OCT_OK:
            return
        msg = self._lib.oct_last_error(self._ctx).decode() if self._ctx is not None else ""
        if rc == L.OCT_ERR_ARGUMENT:
            raise ArgumentError(msg)
        if rc == L.OCT_ERR_DIMENSION:
            raise DimensionMismatch(msg)
        if rc == L.OCT_ERR_STATE:
            raise OctopusError(msg)
        raise OctopusError(f"octohip error {rc}: {msg}")

    def _view(self, key, ptr, n, tail, dtype):
        """The first n rows of the library buffer at address ptr, as a slice
        of a view kept per key: wrapping a buffer (a DLPack capsule, then
        torch.from_dlpack) costs several microseconds per tensor, a slice
        about one.  A buffer keeps its address while it does not grow (a
        grown one gets a new address), and n never exceeds what the buffer
        at ptr holds now, so the slice lies within it."""
        if n == 0 or not ptr:
            return _torch().empty((n,) + tail, dtype=dtype, device=self.device)
        ent = self._views.get(key)
        if ent is None or ent[0] != ptr or ent[1].shape[0] < n:
            ent = (ptr, _wrap(ptr, (n,) + tail, dtype, self.device, self._core))
            self._views[key] = ent
        return ent[1][:n]

    # method spellings of the free functions
    def set_search_radius(self, r):
        return set_search_radius(self, r)

    def add_point_set(self, coords):
        return add_point_set(self, coords)

    def set_active_search(self, q, t):
        return set_active_search(self, q, t)

    def run(self):
        return run(self)

    def build_edges(self, q, t, **kw):
        return build_edges(self, q, t, **kw)


def _coords(tns: TNS, coords):
    torch = _torch()
    if not isinstance(coords, torch.Tensor):
        raise ArgumentError("coords must be a torch tensor on the TNS device")
    if coords.dim() != 2 or coords.shape[1] != tns.ndims:
        raise DimensionMismatch(f"coords must be ({tns.ndims}, N) points, i.e. an (N, {tns.ndims}) tensor")
    if coords.dtype != tns.dtype:
        raise ArgumentError(f"coords must be {tns.dtype}")
    if coords.device.type != "cuda":
        raise ArgumentError("TNS was initialized on :hip but got coords on :cpu; "
                            "all point sets must share a backend")
    if not coords.is_contiguous():
        raise ArgumentError("coords must be contiguous")
    return coords


def _sid(tns, set_id):
    k = int(set_id) - tns.index_base
    if k < 0 or k >= len(tns.point_sets):
        raise ArgumentError(f"unknown point set {set_id}")
    return k


def set_search_radius(tns: TNS, r):
    """set_search_radius!(tns, r) (src/api.jl:24-29)."""
    if not r > 0:
        raise ArgumentError("search radius must be positive")
    import numpy as np

    tns.radius = float(r) if tns.f64 else float(np.float32(r))
    tns.dirty = [True] * len(tns.dirty)
    if tns._ctx is not None:
        tns._set_radius(tns._ctx)
    return tns


def set_refit_mode(tns: TNS, on: bool):
    """set_refit_mode!(tns, on) (src/api.jl:43-47)."""
    tns.refit_mode = bool(on)
    tns.dirty = [True] * len(tns.dirty)
    if tns._ctx is not None:
        tns._check(tns._lib.oct_set_refit_mode(tns._ctx, int(bool(on))))
    return tns


def set_leaf_size(tns: TNS, leaf_size: int):
    """target_leaf_size (fixed to 32 in src/types.jl:177; settable here)."""
    if not 1 <= int(leaf_size) <= 256:
        raise ArgumentError("leaf size must be in 1..256")
    tns.target_leaf_size = int(leaf_size)
    tns.dirty = [True] * len(tns.dirty)
    if tns._ctx is not None:
        tns._check(tns._lib.oct_set_leaf_size(tns._ctx, tns.target_leaf_size))
    return tns


def add_point_set(tns: TNS, coords) -> int:
    """add_point_set!(tns, coords) -> set id (src/api.jl:79-109)."""
    c = _coords(tns, coords)
    sid = ctypes.c_int()
    tns._check(tns._lib.oct_add_point_set(tns._c(), ctypes.c_void_p(c.data_ptr()), c.shape[0],
                                          ctypes.byref(sid)))
    tns.point_sets.append(c)
    tns.dirty.append(True)
    return sid.value


def update_point_set(tns: TNS, set_id: int, coords):
    """update_point_set!(tns, id, coords), same N (src/api.jl:131-140)."""
    c = _coords(tns, coords)
    k = _sid(tns, set_id)
    if c.shape[0] != tns.point_sets[k].shape[0]:
        raise DimensionMismatch("update_point_set! expects same size; use resize_point_set! otherwise")
    tns._check(tns._lib.oct_update_point_set(tns._c(), int(set_id), ctypes.c_void_p(c.data_ptr()), c.shape[0]))
    tns.point_sets[k] = c
    tns.dirty[k] = True
    return tns


def resize_point_set(tns: TNS, set_id: int, coords):
    """resize_point_set!(tns, id, coords), N may change (src/api.jl:111-129)."""
    c = _coords(tns, coords)
    k = _sid(tns, set_id)
    tns._check(tns._lib.oct_resize_point_set(tns._c(), int(set_id), ctypes.c_void_p(c.data_ptr()), c.shape[0]))
    tns.point_sets[k] = c
    tns.dirty[k] = True
    return tns


def set_active_search(tns: TNS, q: int, t: int):
    """set_active_search!(tns, q, t): directed, idempotent (src/api.jl:142-165)."""
    _sid(tns, q)
    _sid(tns, t)
    if (q, t) in tns.active_pairs:
        return tns
    tns._check(tns._lib.oct_set_active_search(tns._c(), int(q), int(t)))
    tns.active_pairs.append((q, t))
    return tns


def run(tns: TNS):
    """run!(tns) (src/api.jl:183-250): rebuild every dirty set."""
    if not tns.radius > 0:
        raise ArgumentError("set_search_radius! before run!")
    if not tns.point_sets:
        raise ArgumentError("add at least one point set before run!")
    tns._check(tns._lib.oct_run(tns._enter()))
    tns._leave()
    tns.dirty = [False] * len(tns.dirty)
    return tns


def get_tree(tns: TNS, set_id: int) -> Tree:
    """tns.trees[id] and tns.permutation[id] as device tensors (reference
    layout, src/types.jl:23-42); valid until the next run."""
    torch = _torch()
    _sid(tns, set_id)
    t = L.OctTree()
    tns._check(tns._lib.oct_get_tree(tns._c(), int(set_id), ctypes.byref(t)))
    m, n, dev, d = int(t.n_nodes), int(t.n_points), tns.device, tns.ndims
    return Tree(
        node_bounds_min=_wrap(t.node_bounds_min, (m, d), tns.dtype, dev, tns),
        node_bounds_max=_wrap(t.node_bounds_max, (m, d), tns.dtype, dev, tns),
        node_first=_wrap(t.node_first, (m,), torch.int32, dev, tns),
        node_last=_wrap(t.node_last, (m,), torch.int32, dev, tns),
        node_children=_wrap(t.node_children, (m, 1 << d), torch.int32, dev, tns),
        permutation=_wrap(t.permutation, (n,), torch.int32, dev, tns),
        n_nodes=m, n_leaves=int(t.n_leaves), origin=tuple(t.origin)[:d],
    )


def build_info(tns: TNS, set_id: int) -> dict:
    """How the last run built a set: {"refit_used": bool, "changed_points": int}
    (changed_points -1 when no previous sort was tried)."""
    _sid(tns, set_id)
    b = L.OctBuildInfo()
    tns._check(tns._lib.oct_build_info_get(tns._c(), int(set_id), ctypes.byref(b)))
    return {"refit_used": bool(b.refit_used), "changed_points": int(b.changed_points)}


def build_edges(tns: TNS, q: int, t: int, mode: str = "exact", order: str = "leaf") -> Edges:
    """build_edges(tns, q, t) (src/api.jl:304-337): flat-COO
    (senders, receivers, rel_displacement (E,3), rel_dist_norm (E,1)).

    The tensors alias library-owned buffers that the next build on this pair
    overwrites (the reference's EdgeBuffer aliasing, test/contracts.jl:223-240).
    order="leaf" groups each receiver's edges in tree-leaf order;
    order="receiver" writes receivers ascending like the reference."""
    torch = _torch()
    if (q, t) not in tns.active_pairs:
        raise OctopusError(f"build_edges!: pair ({q}, {t}) not registered; call set_active_search! first")
    if mode not in _MODES:
        raise ArgumentError(f"mode must be one of {sorted(_MODES)}")
    if order not in _ORDERS:
        raise ArgumentError(f"order must be one of {sorted(_ORDERS)}")
    e = L.OctEdges()
    tns._check(tns._lib.oct_build_edges(tns._enter(), int(q), int(t), _MODES[mode], _ORDERS[order], ctypes.byref(e)))
    tns._leave()
    n = int(e.n_edges)
    return Edges(
        senders=tns._view((q, t, 0), e.senders, n, (), torch.int32),
        receivers=tns._view((q, t, 1), e.receivers, n, (), torch.int32),
        rel_displacement=tns._view((q, t, 2), e.rel_displacement, n, (tns.ndims,), tns.dtype),
        rel_dist_norm=tns._view((q, t, 3), e.rel_dist_norm, n, (1,), tns.dtype),
    )


Neighbors = namedtuple("Neighbors", ["offsets", "flat"])


def _neighbors_of(tns, nb):
    torch = _torch()
    if not nb.materialized:
        return None
    n, total, dev = int(nb.n_queries), int(nb.total), tns.device
    return Neighbors(offsets=_wrap(nb.offsets, (n + 1,), torch.int32, dev, tns),
                     flat=_wrap(nb.flat, (total,), torch.int32, dev, tns))


def materialize_all_neighbors(tns: TNS):
    """materialize_all_neighbors!(tns) (src/api.jl:348-362): CSR neighbour
    lists of every active pair, on the device (the reference builds them on
    the CPU only, src/api.jl:350).  Read them with neighbors(tns, q, t)."""
    ctx = tns._enter()
    for q, t in tns.active_pairs:
        nb = L.OctNeighbors()
        tns._check(tns._lib.oct_materialize_neighbors(ctx, int(q), int(t), ctypes.byref(nb)))
    tns._leave()
    return tns


def neighbors(tns: TNS, q: int, t: int):
    """tns.neighbor_buffers[pair] (src/types.jl:44-48) as device tensors
    (offsets (n_q + 1,), flat (total,)): query i's neighbours are
    flat[offsets[i]:offsets[i + 1]]; None until materialized (run clears it)."""
    nb = L.OctNeighbors()
    tns._check(tns._lib.oct_get_neighbors(tns._c(), int(q), int(t), ctypes.byref(nb)))
    return _neighbors_of(tns, nb)


def get_neighborlist(tns: TNS, q: int, t: int, i: int):
    """get_neighborlist(tns, q, t, i) (src/api.jl:282-292): the neighbours of
    query i (self excluded for the self pair) as an int32 device tensor, from
    the pair's materialized lists (materialized first when they are not)."""
    if (q, t) not in tns.active_pairs:
        raise OctopusError(f"pair ({q}, {t}) not registered; call set_active_search! first")
    nb = neighbors(tns, q, t)
    if nb is None:
        ctx = tns._enter()
        raw = L.OctNeighbors()
        tns._check(tns._lib.oct_materialize_neighbors(ctx, int(q), int(t), ctypes.byref(raw)))
        tns._leave()
        nb = _neighbors_of(tns, raw)
    k = int(i) - tns.index_base
    lo, hi = (int(v) for v in nb.offsets[k:k + 2].cpu())
    return nb.flat[lo:hi]


def device_view(tns: TNS, q: int, t: int) -> L.OctView:
    """device_view(tns, q, t) (src/api.jl:451-454): a by-value handle for HIP
    kernels (include/octohip_device.h); invalid after the next run."""
    v = L.OctView()
    tns._check(tns._lib.oct_device_view(tns._c(), int(q), int(t), ctypes.byref(v)))
    return v


def count_neighbors_device(tns: TNS, view: L.OctView):
    """Per-query neighbour counts from a kernel that uses the device iterator
    (self not filtered), the reference's test/query_cuda.jl:40-56 pattern."""
    torch = _torch()
    out = torch.empty(int(view.n_q), dtype=torch.int32, device=tns.device)
    tns._check(tns._lib.oct_count_neighbors_device(tns._enter(), ctypes.byref(view),
                                                   ctypes.c_void_p(out.data_ptr())))
    tns._leave()
    return out


def set_backward_form(tns: TNS, pull: bool = True):
    """The build_edges_diff adjoint pulls each edge's sender term from its
    reverse edge (pull=True, the default) or pushes every term with an
    atomic (pull=False, the reference GPU kernel's form)."""
    tns._check(tns._lib.oct_set_backward_form(tns._c(), int(bool(pull))))


def set_run_leaves(tns: TNS, run_leaves: int):
    """Leaves per single-pass run without a plan (1..64, default 6)."""
    tns._check(tns._lib.oct_set_run_leaves(tns._c(), int(run_leaves)))


def set_run_chunk(tns: TNS, run_chunk: int):
    """Single-pass runs a block takes per ticket (1..4, default 4)."""
    tns._check(tns._lib.oct_set_run_chunk(tns._c(), int(run_chunk)))


def _tune_table(ls, rs, cs, ms, k):
    """(leaf, run, chunk) of candidate index k and the table of all times."""
    def key(i):
        return (int(ls[i // (rs.shape[0] * cs.shape[0])]), int(rs[(i // cs.shape[0]) % rs.shape[0]]),
                int(cs[i % cs.shape[0]]))

    return key(k), {key(i): float(ms[i]) for i in range(ms.shape[0])}


def autotune(tns: TNS, q: int, t: int, mode: str = "learned", leaf_sizes=(8, 16, 32, 64), runs=(1, 2, 4, 8),
             chunks=(1, 4), reps: int = 3) -> dict:
    """Times run! + build_edges(q, t, mode) for every (leaf size, leaves per
    run, runs per block ticket) candidate and keeps the fastest (the
    reference fixes the leaf size at 32, src/types.jl:177).  Returns
    {"best": (leaf, run, chunk), "ms": {...}}; the edges never depend on the
    choice."""
    import numpy as np

    if mode not in _MODES:
        raise ArgumentError(f"mode must be one of {sorted(_MODES)}")
    ls = np.ascontiguousarray(np.asarray(leaf_sizes, dtype=np.int32))
    rs = np.ascontiguousarray(np.asarray(runs, dtype=np.int32))
    cs = np.ascontiguousarray(np.asarray(chunks, dtype=np.int32))
    ms = np.zeros(ls.shape[0] * rs.shape[0] * cs.shape[0], dtype=np.float64)
    best = ctypes.c_int()
    tns._check(tns._lib.oct_autotune(tns._enter(), int(q), int(t), _MODES[mode], ls.ctypes.data_as(ctypes.c_void_p),
                                     int(ls.shape[0]), rs.ctypes.data_as(ctypes.c_void_p), int(rs.shape[0]),
                                     cs.ctypes.data_as(ctypes.c_void_p), int(cs.shape[0]), int(reps),
                                     ms.ctypes.data_as(ctypes.c_void_p), ctypes.byref(best)))
    tns._leave()
    top, table = _tune_table(ls, rs, cs, ms, best.value)
    tns.target_leaf_size = top[0]
    tns.dirty = [True] * len(tns.dirty)
    return {"best": top, "ms": table}


def profile(tns: TNS, enable: bool = True, phases=None):
    """Turn per-phase kernel timing (HIP events on the TNS stream) on or off;
    phases: names from PROF_PHASES to time (default: all)."""
    if phases is None or not enable:
        tns._check(tns._lib.oct_profile(tns._c(), int(bool(enable))))
        return
    mask = 0
    for name in phases:
        if name not in L.PROF_PHASES:
            raise ArgumentError(f"unknown phase {name!r}; one of {L.PROF_PHASES}")
        mask |= 1 << L.PROF_PHASES.index(name)
    tns._check(tns._lib.oct_profile_phases(tns._c(), mask))


def profile_read(tns: TNS, reset: bool = True) -> dict:
    """{phase: (milliseconds, launches)} since the last reset."""
    n = len(L.PROF_PHASES)
    ms = (ctypes.c_double * n)()
    calls = (ctypes.c_int64 * n)()
    tns._check(tns._lib.oct_profile_read(tns._c(), int(bool(reset)), ms, calls))
    return {name: (ms[k], calls[k]) for k, name in enumerate(L.PROF_PHASES)}


def synchronize(tns: TNS):
    tns._check(tns._lib.oct_synchronize(tns._c()))


def trim(tns: TNS) -> int:
    """Frees the allocations that later, larger results replaced (they stay
    allocated, stale, so that tensors taken from them remain readable
    memory); returns how many were freed.  Tensors taken before their buffer
    was replaced must not be read afterwards."""
    n = retired_count(tns)
    tns._check(tns._lib.oct_trim(tns._c()))
    return n


def retired_count(tns: TNS) -> int:
    """Allocations kept on the context's retire list."""
    n = ctypes.c_int64()
    tns._check(tns._lib.oct_retired_count(tns._c(), ctypes.byref(n)))
    return int(n.value)


def for_each_neighbor(f, tns: TNS, q: int, t: int, i: int):
    """Host for_each_neighbor: the reference allows it on CPU handles only and
    raises on device handles (src/cpu/iterator_cpu.jl:246); this engine's
    handles are device handles, so use the device iterator instead."""
    raise OctopusError("for_each_neighbor (host) requires a CPU TNS; use for_each_neighbor_device on GPU")


# ------------------------------------------------------------------ z-sort


def prepare_zsort(tns: TNS):
    """prepare_zsort(tns) (src/api.jl:381-388): every point set's permutation
    must be current (run! since the last change)."""
    for k, d in enumerate(tns.dirty):
        if d:
            raise OctopusError(f"prepare_zsort: call run! first so the permutation for set "
                               f"{k + tns.index_base} is available")
    return tns


def _zsort_call(tns, set_id, src, adjoint):
    torch = _torch()
    k = _sid(tns, set_id)
    n = tns.point_sets[k].shape[0]
    if src.dim() > 2:
        raise ArgumentError(f"apply_zsort supports 1-D or 2-D arrays only; got {src.dim()}-D")
    if src.shape[0] != n:
        raise DimensionMismatch(f"perm length {n} does not match the point axis {src.shape[0]}")
    if src.device != tns.device:
        raise ArgumentError("apply_zsort: the array must live on the TNS device")
    src = src.contiguous()
    width = 1 if src.dim() == 1 else src.shape[1]
    out = torch.empty_like(src)
    if src.numel() == 0:
        return out
    fn = tns._lib.oct_apply_zsort_adjoint if adjoint else tns._lib.oct_apply_zsort
    tns._check(fn(tns._enter(), int(set_id), ctypes.c_void_p(src.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                  int(width), int(src.element_size())))
    tns._leave()
    return out


def _make_zsort_fn():
    torch = _torch()

    class ApplyZsort(torch.autograd.Function):
        """rrule of apply_zsort (ext/OctopusChainRulesCoreExt.jl:65-94): the
        pullback scatters the tangent back through the permutation."""

        @staticmethod
        def forward(ctx, user_array, tns, set_id):
            ctx.tns, ctx.set_id = tns, set_id
            return _zsort_call(tns, set_id, user_array, False)

        @staticmethod
        def backward(ctx, dy):
            return _zsort_call(ctx.tns, ctx.set_id, dy, True), None, None

    return ApplyZsort


_ApplyZsort = None


def apply_zsort(tns: TNS, set_id: int, user_array):
    """apply_zsort(tns, set_id, user_array) (src/api.jl:405-410): a new array
    whose entries along the point axis follow the z-order permutation of the
    set (a (N,) vector or an (N, F) matrix, the bytes of the reference's
    (F, N) column-major matrix); user_array is not written.  Differentiable:
    the backward is the permutation's scatter (the reference's rrule)."""
    global _ApplyZsort
    if _ApplyZsort is None:
        _ApplyZsort = _make_zsort_fn()
    if user_array.requires_grad:
        return _ApplyZsort.apply(user_array, tns, set_id)
    return _zsort_call(tns, set_id, user_array, False)


# ----------------------------------------------------------------- autograd


def _pullback(tns: TNS, n_points: int, senders, receivers, rdisp, rdist, d_disp, d_dist, radius):
    torch = _torch()
    grad = torch.empty((n_points, tns.ndims), dtype=tns.dtype, device=tns.device)
    e = L.OctEdges()
    e.senders = senders.data_ptr() if senders.numel() else None
    e.receivers = receivers.data_ptr() if receivers.numel() else None
    e.rel_displacement = rdisp.data_ptr() if rdisp.numel() else None
    e.rel_dist_norm = rdist.data_ptr() if rdist.numel() else None
    e.n_edges = senders.numel()
    dd = None if d_disp is None else ctypes.c_void_p(d_disp.contiguous().data_ptr())
    dn = None if d_dist is None else ctypes.c_void_p(d_dist.contiguous().data_ptr())
    fn = tns._lib.oct_build_edges_backward64 if tns.f64 else tns._lib.oct_build_edges_backward
    tns._check(fn(tns._enter(), n_points, ctypes.byref(e), dd, dn, float(radius), ctypes.c_void_p(grad.data_ptr())))
    tns._leave()
    return grad


def _make_fn():
    torch = _torch()

    class BuildEdgesDiff(torch.autograd.Function):
        """rrule of build_edges_diff (ext/OctopusChainRulesCoreExt.jl:137-192):
        primal = build_edges copies; pullback = the HIP adjoint kernel."""

        @staticmethod
        def forward(ctx, coords, tns, set_id, radius):
            e = build_edges(tns, set_id, set_id)
            s, r, d, n = (x.clone() for x in e)
            ctx.save_for_backward(s, r, d, n)
            ctx.tns, ctx.n_points, ctx.radius = tns, coords.shape[0], radius
            ctx.mark_non_differentiable(s, r)
            ctx.set_materialize_grads(False)  # an unused output passes a NULL tangent
            return s, r, d, n

        @staticmethod
        def backward(ctx, g_s, g_r, g_disp, g_dist):
            s, r, d, n = ctx.saved_tensors
            grad = _pullback(ctx.tns, ctx.n_points, s, r, d, n, g_disp, g_dist, ctx.radius)
            return grad, None, None, None

    return BuildEdgesDiff


_BuildEdgesDiff = None


def build_edges_diff(coords, tns: TNS, set_id: int, radius) -> Edges:
    """build_edges_diff(coords, tns, id, radius) (src/Octopus.jl:56-73):
    differentiable copies of build_edges(tns, id, id); gradients flow to coords
    through the HIP adjoint.  Keep coords in sync with the TNS
    (update_point_set + run) before calling, as in the reference."""
    global _BuildEdgesDiff
    if _BuildEdgesDiff is None:
        _BuildEdgesDiff = _make_fn()
    return Edges(*_BuildEdgesDiff.apply(coords, tns, set_id, radius))


from . import predictor  # noqa: E402  (learned per-leaf count predictor)
from .batch import Batch  # noqa: E402  (batched sets)
