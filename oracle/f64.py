"""NumPy restatement of the Float64 path (TNS(Float64), test/float64.jl):
binning, stable sort, BFS build, bounds refit, edges and the pullback, every
value in float64, for 3D octrees and 2D quadtrees.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker of the
engine's Float64 sets; the product package never imports it.

numpy float64 array operations round each operation once and never contract
to FMA, so the expressions below are the reference's IEEE double values:
  origin = min - Float64(1e-6)            src/cpu/bin_cpu.jl:12-35
  cell = clamp(floor((p - o) * (1 / r)))  src/cpu/bin_cpu.jl:52-95, src/morton.jl:83-95
  keys: 21 bits per axis (3D), 31 (2D)    src/morton.jl:13-80
  BFS: leaf iff count <= T or level >= 21 (3D) / 31 (2D), children in
       octant order                       src/cpu/build_cpu.jl:20-52,100-194
  bounds: exact min / max                 src/cpu/refit_cpu.jl:21-70
  edges: d2 = (dx*dx + dy*dy) + dz*dz <= r*r with dx = t - q,
       rdisp = -dx / r as -dx * inv_r, rdist = sqrt(d2) * inv_r
                                          src/cpu/iterator_cpu.jl:96-97,
                                          src/cpu/edges_cpu.jl:106-115
The traversal only prunes with a lower bound of the point distances, so the
edge set is this closed form (as for Float32); edges come back sorted by
(receiver, sender).  Small inputs only (Python loops over nodes, an O(N^2)
pair scan).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_BITS = {3: 21, 2: 31}


def _ordered(a):
    """Order-preserving uint64 image of float64 values (-0.0 below +0.0),
    so min / max follow Julia's Base.min / Base.max."""
    u = np.ascontiguousarray(a, np.float64).view(np.uint64)
    neg = (u >> np.uint64(63)) == np.uint64(1)
    return np.where(neg, ~u, u | np.uint64(1 << 63))


def _jmin(a, axis=0):
    a = np.asarray(a, np.float64)
    return np.take_along_axis(a, np.expand_dims(np.argmin(_ordered(a), axis=axis), axis), axis).squeeze(axis)


def _jmax(a, axis=0):
    a = np.asarray(a, np.float64)
    return np.take_along_axis(a, np.expand_dims(np.argmax(_ordered(a), axis=axis), axis), axis).squeeze(axis)


def spread(v, dims):
    """Bit b of v to bit dims * b (the magic-bits split, src/morton.jl)."""
    v = np.asarray(v, np.uint64)
    out = np.zeros_like(v)
    for b in range(_BITS[dims]):
        out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(dims * b)
    return out


def morton_keys(coords, radius):
    """(keys, origin) of (N, D) float64 points."""
    c = np.asarray(coords, np.float64)
    d = c.shape[1]
    origin = _jmin(c, 0) - 1e-6
    inv = 1.0 / float(radius)
    hi = float((1 << _BITS[d]) - 1)
    cells = np.clip(np.floor((c - origin) * inv), 0.0, hi).astype(np.uint64)
    keys = np.zeros(c.shape[0], np.uint64)
    for k in range(d):
        keys |= spread(cells[:, k], d) << np.uint64(k)
    return keys, origin


@dataclass
class Tree:
    dims: int
    origin: np.ndarray
    keys: np.ndarray
    perm: np.ndarray
    node_first: np.ndarray
    node_last: np.ndarray
    node_level: np.ndarray
    node_children: np.ndarray  # (M, 2^D)
    bmin: np.ndarray  # (M, D)
    bmax: np.ndarray

    @property
    def n_nodes(self):
        return int(self.node_first.shape[0])


def run(coords, radius, leaf_size=32):
    """run! for one Float64 set (3D or 2D by the number of columns)."""
    c = np.asarray(coords, np.float64)
    n, d = c.shape
    nch = 1 << d
    maxl = _BITS[d]
    keys, origin = morton_keys(c, radius)
    perm = np.argsort(keys, kind="stable").astype(np.int32)
    sk = keys[perm]
    first, last, level, kids = [0], [n - 1], [0], []
    nid = 0
    while nid < len(first):
        f, l, lv = first[nid], last[nid], level[nid]
        row = [-1] * nch
        if l - f + 1 > leaf_size and lv < maxl:
            shift = np.uint64(d * maxl - d * (lv + 1))
            oc = ((sk[f:l + 1] >> shift) & np.uint64(nch - 1)).astype(np.int64)
            for o in range(nch):
                idx = np.nonzero(oc == o)[0]
                if idx.size:
                    row[o] = len(first)
                    first.append(f + int(idx[0]))
                    last.append(f + int(idx[-1]))
                    level.append(lv + 1)
        kids.append(row)
        nid += 1
    m = len(first)
    first = np.array(first, np.int32)
    last = np.array(last, np.int32)
    kids = np.array(kids, np.int32).reshape(m, nch)
    bmin = np.zeros((m, d))
    bmax = np.zeros((m, d))
    for k in range(m - 1, -1, -1):
        if np.all(kids[k] == -1):
            p = c[perm[first[k]:last[k] + 1]]
        else:
            live = kids[k][kids[k] != -1]
            p = np.concatenate([bmin[live], bmax[live]])
        bmin[k] = _jmin(p, 0)
        bmax[k] = _jmax(p, 0)
    return Tree(d, origin, keys, perm, first, last, np.array(level, np.int32), kids, bmin, bmax)


@dataclass
class Edges:
    offsets: np.ndarray
    senders: np.ndarray
    receivers: np.ndarray
    rel_displacement: np.ndarray  # (E, D)
    rel_dist_norm: np.ndarray  # (E,)

    @property
    def n_edges(self):
        return int(self.senders.shape[0])


def build_edges(coords_q, coords_t, radius, exclude_self):
    """All (i, j) with d2 <= r*r, sorted by (receiver, sender), with the
    features in double, and the per-query offsets."""
    cq = np.asarray(coords_q, np.float64)
    ct = np.asarray(coords_t, np.float64)
    r = float(radius)
    r2, inv = r * r, 1.0 / r
    S, R, D, N = [], [], [], []
    for i in range(cq.shape[0]):
        dd = ct - cq[i]
        d2 = dd[:, 0] * dd[:, 0] + dd[:, 1] * dd[:, 1]
        if cq.shape[1] == 3:
            d2 = d2 + dd[:, 2] * dd[:, 2]
        m = d2 <= r2
        if exclude_self:
            m[i] = False
        j = np.nonzero(m)[0].astype(np.int32)
        S.append(j)
        R.append(np.full(j.shape[0], i, np.int32))
        D.append(-dd[j] * inv)
        N.append(np.sqrt(d2[j]) * inv)
    s, rc = np.concatenate(S), np.concatenate(R)
    off = np.zeros(cq.shape[0] + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(rc, minlength=cq.shape[0]))
    return Edges(off, s, rc, np.concatenate(D).reshape(-1, cq.shape[1]), np.concatenate(N))


def pullback(n_points, senders, receivers, rdisp, rdist, d_disp, d_dist, r):
    """ext/OctopusChainRulesCoreExt.jl:137-192 in double (summed with
    np.add.at, so the order of the float64 additions is numpy's)."""
    s = np.asarray(senders, np.int64)
    rc = np.asarray(receivers, np.int64)
    rd = np.asarray(rdisp, np.float64)
    dims = rd.shape[1]
    g = np.zeros((n_points, dims))
    v = np.zeros_like(rd)
    if d_disp is not None:
        v = v + np.asarray(d_disp, np.float64).reshape(-1, dims) / r
    if d_dist is not None:
        dn = np.asarray(rdist, np.float64).reshape(-1)
        ok = dn > 1e-8
        inv_dnr = np.zeros_like(dn)
        inv_dnr[ok] = np.asarray(d_dist, np.float64).reshape(-1)[ok] / (dn[ok] * r)
        v = v + rd * inv_dnr[:, None]
    np.add.at(g, rc, v)
    np.add.at(g, s, -v)
    return g
