"""ctypes binding of the CPU oracle (liboracle.so, built from octopus_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
as the checker, and by bench.py's ``cpu_baseline`` leg; the product package
never imports it.

Indices are 0-based; the reference (Julia) values are these plus one.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


class _Tree(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("radius", ctypes.c_float),
        ("leaf_size", ctypes.c_int32),
        ("n_nodes", ctypes.c_int32),
        ("origin", ctypes.c_float * 3),
        ("keys", _u64p),
        ("perm", _i32p),
        ("node_first", _i32p),
        ("node_last", _i32p),
        ("node_level", _i32p),
        ("node_children", _i32p),
        ("bmin", _f32p),
        ("bmax", _f32p),
    ]


class _Tree2(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("radius", ctypes.c_float),
        ("leaf_size", ctypes.c_int32),
        ("n_nodes", ctypes.c_int32),
        ("origin", ctypes.c_float * 2),
        ("keys", _u64p),
        ("perm", _i32p),
        ("node_first", _i32p),
        ("node_last", _i32p),
        ("node_level", _i32p),
        ("node_children", _i32p),
        ("bmin", _f32p),
        ("bmax", _f32p),
    ]


class _Edges(ctypes.Structure):
    _fields_ = [
        ("nq", ctypes.c_int64),
        ("n_edges", ctypes.c_int64),
        ("offsets", _i64p),
        ("senders", _i32p),
        ("receivers", _i32p),
        ("rdisp", _f32p),
        ("rdist", _f32p),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess

    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_split_by_3.restype = ctypes.c_uint64
        L.orc_split_by_3.argtypes = [ctypes.c_uint64]
        L.orc_compact_by_3.restype = ctypes.c_uint64
        L.orc_compact_by_3.argtypes = [ctypes.c_uint64]
        L.orc_morton_encode3.restype = ctypes.c_uint64
        L.orc_morton_encode3.argtypes = [ctypes.c_uint32] * 3
        L.orc_morton_decode3.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_bin_point3.argtypes = [ctypes.c_float] * 7 + [_i32p]
        L.orc_point_origin3.argtypes = [_f32p, ctypes.c_int64, _f32p]
        L.orc_bin3.argtypes = [_f32p, ctypes.c_int64, _f32p, ctypes.c_float, _u64p]
        L.orc_sort_by_key.argtypes = [_u64p, ctypes.c_int64, _i32p]
        L.orc_run.restype = ctypes.POINTER(_Tree)
        L.orc_run.argtypes = [_f32p, ctypes.c_int64, ctypes.c_float, ctypes.c_int32]
        L.orc_tree_free.argtypes = [ctypes.POINTER(_Tree)]
        L.orc_for_each_neighbor.restype = ctypes.c_int64
        L.orc_for_each_neighbor.argtypes = [ctypes.POINTER(_Tree), _f32p, _f32p, ctypes.c_int64,
                                            _i32p, ctypes.c_int64]
        L.orc_build_edges.restype = ctypes.POINTER(_Edges)
        L.orc_build_edges.argtypes = [ctypes.POINTER(_Tree), _f32p, ctypes.c_int64, _f32p,
                                      ctypes.c_int, ctypes.c_int]
        L.orc_edges_free.argtypes = [ctypes.POINTER(_Edges)]
        L.orc_count_edges.restype = ctypes.c_int64
        L.orc_count_edges.argtypes = [ctypes.POINTER(_Tree), _f32p, ctypes.c_int64, _f32p, ctypes.c_int,
                                      ctypes.c_int, _i32p]
        L.orc_build_edges_rows.restype = ctypes.POINTER(_Edges)
        L.orc_build_edges_rows.argtypes = [ctypes.POINTER(_Tree), _f32p, _i64p, ctypes.c_int64, _f32p,
                                           ctypes.c_int, ctypes.c_int]
        L.orc_pullback.argtypes = [ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f32p, _f32p,
                                   _f32p, _f32p, ctypes.c_float, _f32p]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_morton_encode2.restype = ctypes.c_uint64
        L.orc_morton_encode2.argtypes = [ctypes.c_uint32] * 2
        L.orc_morton_decode2.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32)]
        L.orc_bin_point2.argtypes = [ctypes.c_float] * 5 + [_i32p]
        L.orc_run2.restype = ctypes.POINTER(_Tree2)
        L.orc_run2.argtypes = [_f32p, ctypes.c_int64, ctypes.c_float, ctypes.c_int32]
        L.orc_tree2_free.argtypes = [ctypes.POINTER(_Tree2)]
        L.orc_build_edges2.restype = ctypes.POINTER(_Edges)
        L.orc_build_edges2.argtypes = [ctypes.POINTER(_Tree2), _f32p, ctypes.c_int64, _f32p, ctypes.c_int,
                                       ctypes.c_int]
        L.orc_pullback2.argtypes = [ctypes.c_int64, ctypes.c_int64, _i32p, _i32p, _f32p, _f32p, _f32p, _f32p,
                                    ctypes.c_float, _f32p]
        _lib = L
    return _lib


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def _coords(c) -> np.ndarray:
    c = np.ascontiguousarray(c, dtype=np.float32)
    if c.ndim != 2 or c.shape[1] != 3:
        raise ValueError("coords must be (N, 3) float32, one point per row (xyz interleaved)")
    return c


# ------------------------------------------------------------------ Morton


def morton_encode3(ix: int, iy: int, iz: int) -> int:
    return int(lib().orc_morton_encode3(ix, iy, iz))


def morton_decode3(code: int):
    out = (ctypes.c_uint32 * 3)()
    lib().orc_morton_decode3(code, out)
    return tuple(int(v) for v in out)


def bin_point(p, o, inv_cs):
    out = (ctypes.c_int32 * 3)()
    lib().orc_bin_point3(*(float(v) for v in p), *(float(v) for v in o), float(inv_cs), out)
    return tuple(int(v) for v in out)


def point_origin(coords) -> np.ndarray:
    c = _coords(coords)
    o = np.zeros(3, np.float32)
    lib().orc_point_origin3(_ptr(c, _f32p), c.shape[0], _ptr(o, _f32p))
    return o


def morton_keys(coords, origin, cell_size: float) -> np.ndarray:
    c = _coords(coords)
    o = np.ascontiguousarray(origin, np.float32)
    k = np.zeros(c.shape[0], np.uint64)
    lib().orc_bin3(_ptr(c, _f32p), c.shape[0], _ptr(o, _f32p), float(cell_size), _ptr(k, _u64p))
    return k


def sort_by_key(keys) -> np.ndarray:
    k = np.ascontiguousarray(keys, np.uint64)
    p = np.zeros(k.shape[0], np.int32)
    lib().orc_sort_by_key(_ptr(k, _u64p), k.shape[0], _ptr(p, _i32p))
    return p


# -------------------------------------------------------------------- run!


@dataclass
class Tree:
    n: int
    radius: float
    leaf_size: int
    origin: np.ndarray
    keys: np.ndarray
    perm: np.ndarray
    node_first: np.ndarray
    node_last: np.ndarray
    node_level: np.ndarray
    node_children: np.ndarray  # (M, 8)
    bmin: np.ndarray  # (M, 3)
    bmax: np.ndarray  # (M, 3)
    _handle: object = None
    _coords: object = None

    @property
    def n_nodes(self) -> int:
        return int(self.node_first.shape[0])

    def __del__(self):
        if self._handle is not None and _lib is not None:
            _lib.orc_tree_free(self._handle)
            self._handle = None


def _arr(p, n, dtype, shape=None):
    if n == 0:
        a = np.zeros(0, dtype)
    else:
        a = np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)
    return a.reshape(shape) if shape is not None else a


def run(coords, radius: float, leaf_size: int = 32) -> Tree:
    """run! for one 3D point set (src/api.jl:194-250)."""
    c = _coords(coords)
    h = lib().orc_run(_ptr(c, _f32p), c.shape[0], float(radius), int(leaf_size))
    t = h.contents
    n, m = int(t.n), int(t.n_nodes)
    tree = Tree(
        n=n, radius=float(np.float32(radius)), leaf_size=int(leaf_size),
        origin=np.array(list(t.origin), np.float32),
        keys=_arr(t.keys, n, np.uint64), perm=_arr(t.perm, n, np.int32),
        node_first=_arr(t.node_first, m, np.int32), node_last=_arr(t.node_last, m, np.int32),
        node_level=_arr(t.node_level, m, np.int32),
        node_children=_arr(t.node_children, 8 * m, np.int32, (m, 8)),
        bmin=_arr(t.bmin, 3 * m, np.float32, (m, 3)), bmax=_arr(t.bmax, 3 * m, np.float32, (m, 3)),
        _handle=h, _coords=c,
    )
    return tree


# ------------------------------------------------------------ build_edges


@dataclass
class Edges:
    offsets: np.ndarray
    senders: np.ndarray
    receivers: np.ndarray
    rel_displacement: np.ndarray  # (E, 3)
    rel_dist_norm: np.ndarray  # (E,)

    @property
    def n_edges(self) -> int:
        return int(self.senders.shape[0])


def build_edges(tree: Tree, coords_q, coords_t=None, exclude_self=None, nthreads=None) -> Edges:
    """build_edges for query set coords_q against the target set the tree was
    built on (src/cpu/edges_cpu.jl:172-247)."""
    cq = _coords(coords_q)
    ct = tree._coords if coords_t is None else _coords(coords_t)
    if exclude_self is None:
        exclude_self = coords_t is None
    nt = lib().orc_max_threads() if nthreads is None else int(nthreads)
    h = lib().orc_build_edges(tree._handle, _ptr(cq, _f32p), cq.shape[0], _ptr(ct, _f32p),
                              int(bool(exclude_self)), nt)
    e = h.contents
    ne, nq = int(e.n_edges), int(e.nq)
    out = Edges(
        offsets=_arr(e.offsets, nq + 1, np.int64),
        senders=_arr(e.senders, ne, np.int32), receivers=_arr(e.receivers, ne, np.int32),
        rel_displacement=_arr(e.rdisp, 3 * ne, np.float32, (ne, 3)),
        rel_dist_norm=_arr(e.rdist, ne, np.float32),
    )
    lib().orc_edges_free(h)
    return out


def count_edges(tree: Tree, coords_q, coords_t=None, exclude_self=None, nthreads=None):
    """Pass A of build_edges alone (src/cpu/edges_cpu.jl:36-61): per-query
    counts (int32) and their total, without storing any edge."""
    cq = _coords(coords_q)
    ct = tree._coords if coords_t is None else _coords(coords_t)
    if exclude_self is None:
        exclude_self = coords_t is None
    nt = lib().orc_max_threads() if nthreads is None else int(nthreads)
    counts = np.zeros(cq.shape[0], np.int32)
    total = lib().orc_count_edges(tree._handle, _ptr(cq, _f32p), cq.shape[0], _ptr(ct, _f32p),
                                  int(bool(exclude_self)), nt, _ptr(counts, _i32p))
    return counts, int(total)


def build_edges_rows(tree: Tree, coords_q, rows, coords_t=None, exclude_self=None, nthreads=None) -> Edges:
    """Pass B of build_edges (src/cpu/edges_cpu.jl:92-130) for the queries in
    `rows` only, in that order; receivers are the queries' own indices."""
    cq = _coords(coords_q)
    ct = tree._coords if coords_t is None else _coords(coords_t)
    if exclude_self is None:
        exclude_self = coords_t is None
    rw = np.ascontiguousarray(rows, np.int64)
    if rw.size and (rw.min() < 0 or rw.max() >= cq.shape[0]):
        raise ValueError("rows out of range")
    nt = lib().orc_max_threads() if nthreads is None else int(nthreads)
    h = lib().orc_build_edges_rows(tree._handle, _ptr(cq, _f32p), _ptr(rw, _i64p), rw.shape[0],
                                   _ptr(ct, _f32p), int(bool(exclude_self)), nt)
    e = h.contents
    ne, nq = int(e.n_edges), int(e.nq)
    out = Edges(
        offsets=_arr(e.offsets, nq + 1, np.int64),
        senders=_arr(e.senders, ne, np.int32), receivers=_arr(e.receivers, ne, np.int32),
        rel_displacement=_arr(e.rdisp, 3 * ne, np.float32, (ne, 3)),
        rel_dist_norm=_arr(e.rdist, ne, np.float32),
    )
    lib().orc_edges_free(h)
    return out


def apply_zsort(perm, a):
    """apply_zsort_cpu! (src/cpu/zsort_cpu.jl:11-32): out[..., k] =
    a[..., perm[k]] along the point axis (axis 0 of an (N,) or (N, F) array
    here, the last axis of the reference's column-major (F, N))."""
    a = np.asarray(a)
    if a.ndim > 2:
        raise ValueError("1-D or 2-D arrays only")
    out = np.empty_like(a)
    for k in range(perm.shape[0]):
        out[k] = a[perm[k]]
    return out


def zsort_pullback(perm, dy):
    """The rrule's scatter (ext/OctopusChainRulesCoreExt.jl:85-94): grad =
    zero(a); grad[perm[k]] = dy[k]."""
    dy = np.asarray(dy)
    g = np.zeros_like(dy)
    for k in range(perm.shape[0]):
        g[perm[k]] = dy[k]
    return g


def materialize(tree: Tree, coords_q, coords_t=None, exclude_self=None, nthreads=None):
    """materialize_cpu! (src/cpu/query_cpu.jl:139-192): CSR (offsets int32
    of n_q + 1 starting at 0, flat int32 targets).  Its passes are those of
    build_edges without the features (the same traversal, the same j != i
    rule for the self pair, the same serial exclusive scan), so the lists
    are build_edges' offsets and senders."""
    e = build_edges(tree, coords_q, coords_t, exclude_self, nthreads)
    return e.offsets.astype(np.int32), e.senders.copy()


def for_each_neighbor(tree: Tree, coords_q, i: int, coords_t=None) -> np.ndarray:
    """Host for_each_neighbor (src/cpu/iterator_cpu.jl:245-300): every j within
    r of query i, traversal order, self not filtered."""
    cq = _coords(coords_q)
    ct = tree._coords if coords_t is None else _coords(coords_t)
    cap = 1024
    while True:
        buf = np.zeros(cap, np.int32)
        n = lib().orc_for_each_neighbor(tree._handle, _ptr(cq, _f32p), _ptr(ct, _f32p), int(i),
                                        _ptr(buf, _i32p), cap)
        if n <= cap:
            return buf[:n].copy()
        cap = int(n)


def pullback(n_points: int, senders, receivers, rdisp, rdist, d_disp, d_dist, r: float):
    """build_edges_diff CPU pullback (ext/OctopusChainRulesCoreExt.jl:137-192)."""
    s = np.ascontiguousarray(senders, np.int32)
    rc = np.ascontiguousarray(receivers, np.int32)
    dp = np.ascontiguousarray(rdisp, np.float32).reshape(-1, 3)
    dn = np.ascontiguousarray(rdist, np.float32).reshape(-1)
    g = np.zeros((n_points, 3), np.float32)
    a = None if d_disp is None else np.ascontiguousarray(d_disp, np.float32).reshape(-1, 3)
    b = None if d_dist is None else np.ascontiguousarray(d_dist, np.float32).reshape(-1)
    lib().orc_pullback(n_points, s.shape[0], _ptr(s, _i32p), _ptr(rc, _i32p), _ptr(dp, _f32p),
                       _ptr(dn, _f32p), None if a is None else _ptr(a, _f32p),
                       None if b is None else _ptr(b, _f32p), float(r), _ptr(g, _f32p))
    return g


def max_threads() -> int:
    return int(lib().orc_max_threads())


# ------------------------------------------------------------------------ 2D


def _coords2(c) -> np.ndarray:
    c = np.ascontiguousarray(c, dtype=np.float32)
    if c.ndim != 2 or c.shape[1] != 2:
        raise ValueError("coords must be (N, 2) float32, one point per row (xy interleaved)")
    return c


def morton_encode2(ix: int, iy: int) -> int:
    return int(lib().orc_morton_encode2(ix, iy))


def morton_decode2(code: int):
    out = (ctypes.c_uint32 * 2)()
    lib().orc_morton_decode2(code, out)
    return tuple(int(v) for v in out)


def bin_point2(p, o, inv_cs):
    out = (ctypes.c_int32 * 2)()
    lib().orc_bin_point2(*(float(v) for v in p), *(float(v) for v in o), float(inv_cs), out)
    return tuple(int(v) for v in out)


@dataclass
class Tree2:
    n: int
    radius: float
    leaf_size: int
    origin: np.ndarray
    keys: np.ndarray
    perm: np.ndarray
    node_first: np.ndarray
    node_last: np.ndarray
    node_level: np.ndarray
    node_children: np.ndarray  # (M, 4)
    bmin: np.ndarray  # (M, 2)
    bmax: np.ndarray  # (M, 2)
    _handle: object = None
    _coords: object = None

    @property
    def n_nodes(self) -> int:
        return int(self.node_first.shape[0])

    def __del__(self):
        if self._handle is not None and _lib is not None:
            _lib.orc_tree2_free(self._handle)
            self._handle = None


def run2(coords, radius: float, leaf_size: int = 32) -> Tree2:
    """run! for one 2D point set (quadtree)."""
    c = _coords2(coords)
    h = lib().orc_run2(_ptr(c, _f32p), c.shape[0], float(radius), int(leaf_size))
    t = h.contents
    n, m = int(t.n), int(t.n_nodes)
    return Tree2(
        n=n, radius=float(np.float32(radius)), leaf_size=int(leaf_size),
        origin=np.array(list(t.origin), np.float32),
        keys=_arr(t.keys, n, np.uint64), perm=_arr(t.perm, n, np.int32),
        node_first=_arr(t.node_first, m, np.int32), node_last=_arr(t.node_last, m, np.int32),
        node_level=_arr(t.node_level, m, np.int32),
        node_children=_arr(t.node_children, 4 * m, np.int32, (m, 4)),
        bmin=_arr(t.bmin, 2 * m, np.float32, (m, 2)), bmax=_arr(t.bmax, 2 * m, np.float32, (m, 2)),
        _handle=h, _coords=c,
    )


def build_edges2(tree: Tree2, coords_q, coords_t=None, exclude_self=None, nthreads=None) -> Edges:
    """2D build_edges (src/cpu/edges_cpu.jl:132-170 fill): rel_displacement (E, 2)."""
    cq = _coords2(coords_q)
    ct = tree._coords if coords_t is None else _coords2(coords_t)
    if exclude_self is None:
        exclude_self = coords_t is None
    nt = lib().orc_max_threads() if nthreads is None else int(nthreads)
    h = lib().orc_build_edges2(tree._handle, _ptr(cq, _f32p), cq.shape[0], _ptr(ct, _f32p),
                               int(bool(exclude_self)), nt)
    e = h.contents
    ne, nq = int(e.n_edges), int(e.nq)
    out = Edges(
        offsets=_arr(e.offsets, nq + 1, np.int64),
        senders=_arr(e.senders, ne, np.int32), receivers=_arr(e.receivers, ne, np.int32),
        rel_displacement=_arr(e.rdisp, 2 * ne, np.float32, (ne, 2)),
        rel_dist_norm=_arr(e.rdist, ne, np.float32),
    )
    lib().orc_edges_free(h)
    return out


def pullback2(n_points: int, senders, receivers, rdisp, rdist, d_disp, d_dist, r: float):
    """The CPU pullback with D = 2 (ext/OctopusChainRulesCoreExt.jl:137-192)."""
    s = np.ascontiguousarray(senders, np.int32)
    rc = np.ascontiguousarray(receivers, np.int32)
    dp = np.ascontiguousarray(rdisp, np.float32).reshape(-1, 2)
    dn = np.ascontiguousarray(rdist, np.float32).reshape(-1)
    g = np.zeros((n_points, 2), np.float32)
    a = None if d_disp is None else np.ascontiguousarray(d_disp, np.float32).reshape(-1, 2)
    b = None if d_dist is None else np.ascontiguousarray(d_dist, np.float32).reshape(-1)
    lib().orc_pullback2(n_points, s.shape[0], _ptr(s, _i32p), _ptr(rc, _i32p), _ptr(dp, _f32p), _ptr(dn, _f32p),
                        None if a is None else _ptr(a, _f32p), None if b is None else _ptr(b, _f32p), float(r),
                        _ptr(g, _f32p))
    return g
