"""Pins the 2D (quadtree) oracle: the reference tests' 2D Morton known
answers (test/morton.jl:40-75), the 2D tree invariants (test/build_2d.jl),
edges and features against an O(N^2) 2D brute force (test/edges_cpu_2d.jl),
the D = 2 pullback against float64 central differences, and the identity the
engine's 2D sets rest on: the quadtree of (x, y) is a chain of ten
single-child nodes on top of the octree of (x, y, 2^20) - same ranges,
bounds and levels shifted by ten, children in the first four octants - and
the two point sets have the same edges."""
import json
import os

import numpy as np
import pytest

from brute import bits
from octohip import datasets

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))
PLANE_Z = np.float32(2.0 ** 20)


def brute2(cq, ct, r, exclude_self):
    """All (i, j) with fl(dx*dx + dy*dy) <= r*r, dx = t - q, sorted by (i, j);
    rdisp = -dx * inv_r, rdist = sqrt(d2) * inv_r (src/cpu/edges_cpu.jl:150-156)."""
    r = np.float32(r)
    r2, inv = r * r, np.float32(1.0) / r
    S, R, D, N = [], [], [], []
    for i in range(cq.shape[0]):
        d = ct - cq[i]
        d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]
        m = d2 <= r2
        if exclude_self:
            m[i] = False
        j = np.nonzero(m)[0].astype(np.int32)
        S.append(j)
        R.append(np.full(j.shape[0], i, np.int32))
        D.append(-d[j] * inv)
        N.append(np.sqrt(d2[j]) * inv)
    return np.concatenate(S), np.concatenate(R), np.concatenate(D).astype(np.float32), np.concatenate(N)


def sort2(s, rc, d, n):
    o = np.lexsort((s, rc))
    return s[o], rc[o], d.reshape(-1, 2)[o], n.reshape(-1)[o]


def test_morton2_known_answers(orc):
    for x, y in GOLD["morton_roundtrip_2d"]["cases"]:
        assert orc.morton_decode2(orc.morton_encode2(x, y)) == (x, y)
    for x, y in GOLD["morton_top_bits_2d"]["cases"]:
        assert orc.morton_encode2(x, y) >> 62 == 0
    g = GOLD["morton_monotone_x_2d"]
    prev = orc.morton_encode2(0, g["y"])
    for x in range(1, g["x_max"] + 1):
        cur = orc.morton_encode2(x, g["y"])
        assert cur > prev
        prev = cur
    assert orc.morton_encode2(1, 0) == 1 and orc.morton_encode2(0, 1) == 2 and orc.morton_encode2(2, 0) == 4
    for c in GOLD["bin_point_2d"]["cases"]:
        inv = np.float32(1.0) if c["inv_cs"] == "1f0" else np.float32(1.0) / np.float32(0.1)
        assert orc.bin_point2(np.float32(c["p"]), np.float32(c["o"]), inv) == tuple(c["want"])


def assert_tree2_invariants(t, pts):
    """test/build_2d.jl: ranges, quadrant order, 4 children, exact bounds."""
    n, m = t.n, t.n_nodes
    assert t.node_first[0] == 0 and t.node_last[0] == n - 1
    assert np.array_equal(np.sort(t.perm), np.arange(n))
    sk = t.keys[t.perm]
    assert np.all(sk[1:] >= sk[:-1])
    leaf = np.all(t.node_children == -1, axis=1)
    for nid in range(m):
        f, l = t.node_first[nid], t.node_last[nid]
        if leaf[nid]:
            assert l - f + 1 <= t.leaf_size or t.node_level[nid] >= 31
            p = pts[t.perm[f:l + 1]]
            assert np.array_equal(t.bmin[nid], p.min(axis=0)) and np.array_equal(t.bmax[nid], p.max(axis=0))
        else:
            kids = t.node_children[nid]
            prev = f - 1
            for q in range(4):
                c = kids[q]
                if c == -1:
                    continue
                assert t.node_level[c] == t.node_level[nid] + 1 and t.node_first[c] == prev + 1
                prev = t.node_last[c]
                shift = 62 - 2 * (t.node_level[nid] + 1)
                assert np.all((sk[t.node_first[c]:t.node_last[c] + 1] >> np.uint64(shift)) & np.uint64(3) == q)
            assert prev == l
            live = kids[kids != -1]
            assert np.array_equal(t.bmin[nid], t.bmin[live].min(axis=0))


@pytest.mark.parametrize("n,r,seed", [(60, 0.2, 1), (1000, 0.05, 2), (3000, 0.02, 3)])
def test_tree2_and_edges_vs_brute(orc, n, r, seed):
    pts = datasets.uniform_cube(n, seed)[:, :2].copy()
    t = orc.run2(pts, r)
    assert_tree2_invariants(t, pts)
    e = orc.build_edges2(t, pts)
    assert np.all(np.diff(e.receivers) >= 0)
    got = sort2(e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm)
    want = brute2(pts, pts, r, True)
    for x, y in zip(got, want):
        assert np.array_equal(bits(x), bits(y))


def test_edges2_cross_pair_and_coincident(orc):
    a = datasets.uniform_cube(150, 5)[:, :2].copy()
    b = datasets.uniform_cube(250, 6)[:, :2].copy()
    e = orc.build_edges2(orc.run2(b, 0.12), a, b)
    for x, y in zip(sort2(e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm), brute2(a, b, 0.12, False)):
        assert np.array_equal(bits(x), bits(y))
    z = np.zeros((50, 2), np.float32)
    t = orc.run2(z, 0.1)
    assert t.node_level.max() == 31  # the coincident chain runs to the last level
    assert orc.build_edges2(t, z).n_edges == 50 * 49


def test_pullback2_fd64(orc):
    pts = datasets.uniform_cube(300, 8)[:, :2].astype(np.float32).copy()
    r = 0.15
    e = orc.build_edges2(orc.run2(pts, r), pts)
    s, rc = e.senders.astype(np.int64), e.receivers.astype(np.int64)
    g = orc.pullback2(300, e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm,
                      2 * e.rel_displacement, e.rel_dist_norm, r)
    rd = float(np.float32(r))
    xd = pts.astype(np.float64)

    def f(z):
        d = (z[rc] - z[s]) / rd
        q = np.sum(d * d, axis=1)
        return np.sum(q) + 0.5 * np.sum(q)

    for k in (0, 17, 301, 555):
        zp, zm = xd.copy().reshape(-1), xd.copy().reshape(-1)
        zp[k] += 1e-5
        zm[k] -= 1e-5
        fd = (f(zp.reshape(-1, 2)) - f(zm.reshape(-1, 2))) / 2e-5
        assert abs(g.reshape(-1)[k] - fd) <= 1e-3 * abs(fd) + 1e-2, (k, g.reshape(-1)[k], fd)


@pytest.mark.parametrize("n,r,leaf,seed", [(5000, 0.01, 32, 11), (5000, 0.01, 8, 12), (20, 0.3, 32, 13),
                                           (64, 0.1, 32, 0)])
def test_quadtree_is_chain_over_planar_octree(orc, n, r, leaf, seed):
    """The engine builds a 2D set as the octree of (x, y, 2^20): at z = 2^20
    the origin's z is 2^20 itself (2^20 - 1e-6 rounds back in FP32), so every
    z cell is 0, and the octants are the quadrants.  Below 2^21 cells per
    axis the quadtree's levels 0-9 are one-child nodes over everything, and
    levels 10-31 are the octree's levels 0-21, node for node."""
    pts = datasets.uniform_cube(n, seed)[:, :2].copy() if seed else np.zeros((n, 2), np.float32)
    p3 = np.concatenate([pts, np.full((n, 1), PLANE_Z, np.float32)], axis=1)
    q = orc.run2(pts, r, leaf_size=leaf)
    o = orc.run(p3, r, leaf_size=leaf)
    assert np.array_equal(q.perm, o.perm)
    chain = 10 if o.n_nodes > 1 or n > leaf else 0
    assert q.n_nodes == o.n_nodes + chain
    assert np.all(q.node_children[:chain, 1:] == -1)
    assert np.array_equal(q.node_children[:chain, 0], np.arange(1, chain + 1))
    assert np.all(q.node_first[:chain] == 0) and np.all(q.node_last[:chain] == n - 1)
    assert np.array_equal(q.node_first[chain:], o.node_first)
    assert np.array_equal(q.node_last[chain:], o.node_last)
    assert np.array_equal(q.node_level[chain:], o.node_level + chain)
    assert np.all(o.node_children[:, 4:] == -1)
    kids = o.node_children[:, :4].copy()
    kids[kids >= 0] += chain
    assert np.array_equal(q.node_children[chain:], kids)
    assert np.array_equal(q.bmin[chain:], o.bmin[:, :2]) and np.array_equal(q.bmax[chain:], o.bmax[:, :2])
    assert np.all(o.bmin[:, 2] == PLANE_Z)
    e2 = orc.build_edges2(q, pts)
    e3 = orc.build_edges(o, p3)
    a = sort2(e2.senders, e2.receivers, e2.rel_displacement, e2.rel_dist_norm)
    o3 = np.lexsort((e3.senders, e3.receivers))
    assert np.array_equal(a[0], e3.senders[o3]) and np.array_equal(a[1], e3.receivers[o3])
    assert np.array_equal(bits(a[2]), bits(e3.rel_displacement[o3][:, :2]))
    assert np.array_equal(bits(a[3]), bits(e3.rel_dist_norm[o3]))
