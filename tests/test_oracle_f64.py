"""Pins the Float64 oracle (oracle/f64.py): on dyadic inputs, where every
value is exact in both precisions, its keys, tree and edges must equal the
Float32 C oracle's (itself pinned by test_oracle.py / test_oracle_2d.py),
which checks the restated formulas (origin, cells, octant order, the leaf
rule, the feature signs) against an independent implementation; on random
doubles the tree invariants hold and the edges are the brute-force set of
test/float64.jl; the pullback matches central differences."""
import numpy as np
import pytest

import oracle.f64 as f64
from octohip import datasets


def _dyadic(n, dims, seed):
    # multiples of 1/64 in [0, 1): exact in float32 and float64
    g = np.random.default_rng(seed)
    return (g.integers(0, 64, size=(n, dims)) / 64.0).astype(np.float64)


@pytest.mark.parametrize("dims,leaf", [(3, 32), (3, 8), (2, 32), (2, 8)])
def test_f64_matches_the_float32_oracle_on_exact_inputs(orc, dims, leaf):
    pts = _dyadic(3000, dims, 5 + dims)
    r = 0.0625  # 2^-4: 1 / r and every product below are exact
    t = f64.run(pts, r, leaf_size=leaf)
    p32 = pts.astype(np.float32)
    o = orc.run(p32, r, leaf_size=leaf) if dims == 3 else orc.run2(p32, r, leaf_size=leaf)
    assert np.array_equal(t.keys, o.keys)
    assert np.array_equal(t.perm, o.perm)
    assert np.array_equal(t.node_first, o.node_first) and np.array_equal(t.node_last, o.node_last)
    assert np.array_equal(t.node_children, o.node_children)
    assert np.array_equal(t.node_level, o.node_level)
    assert np.array_equal(t.bmin, o.bmin.astype(np.float64)) and np.array_equal(t.bmax, o.bmax.astype(np.float64))
    e = f64.build_edges(pts, pts, r, True)
    oe = orc.build_edges(o, p32) if dims == 3 else orc.build_edges2(o, p32)
    w = np.lexsort((oe.senders, oe.receivers))
    assert np.array_equal(e.senders, oe.senders[w]) and np.array_equal(e.receivers, oe.receivers[w])
    assert np.array_equal(e.rel_displacement, oe.rel_displacement[w].astype(np.float64))
    # sqrt rounds in each precision: the double result rounds to the Float32 one
    assert np.array_equal(e.rel_dist_norm.astype(np.float32), oe.rel_dist_norm[w])
    assert np.array_equal(e.offsets, oe.offsets)


@pytest.mark.parametrize("dims", [3, 2])
def test_f64_tree_invariants_and_brute_force(dims):
    """test/float64.jl:33-110: build + query + edges against brute force."""
    g = np.random.default_rng(40 + dims)
    pts = g.random((400, dims))
    r = 0.1
    t = f64.run(pts, r)
    assert np.array_equal(np.sort(t.perm), np.arange(400))
    assert np.all(np.diff(t.keys[t.perm].astype(np.float64)) >= 0)
    leaf = np.all(t.node_children == -1, axis=1)
    for k in range(t.n_nodes):
        p = pts[t.perm[t.node_first[k]:t.node_last[k] + 1]]
        assert np.array_equal(t.bmin[k], p.min(axis=0)) and np.array_equal(t.bmax[k], p.max(axis=0))
        if not leaf[k]:
            assert t.node_last[k] - t.node_first[k] + 1 > 32
    e = f64.build_edges(pts, pts, r, True)
    for i in (0, 123, 399):
        d = pts - pts[i]
        d2 = np.sum(d * d, axis=1)
        want = np.nonzero(d2 <= r * r)[0]
        want = want[want != i]
        assert np.array_equal(e.senders[e.receivers == i], want)
    assert np.all(e.rel_dist_norm <= 1.0 + 1e-12)
    assert e.rel_displacement.shape == (e.n_edges, dims)


def test_f64_pullback_fd():
    g = np.random.default_rng(9)
    pts = g.random((200, 3))
    r = 0.2
    e = f64.build_edges(pts, pts, r, True)
    s, rc = e.senders.astype(np.int64), e.receivers.astype(np.int64)
    gr = f64.pullback(200, e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm, 2 * e.rel_displacement,
                      e.rel_dist_norm, r)

    def f(z):
        d = (z[rc] - z[s]) / r
        q = np.sum(d * d, axis=1)
        return np.sum(q) + 0.5 * np.sum(q)

    for k in (0, 31, 302, 599):
        zp, zm = pts.copy().reshape(-1), pts.copy().reshape(-1)
        zp[k] += 1e-6
        zm[k] -= 1e-6
        fd = (f(zp.reshape(-1, 3)) - f(zm.reshape(-1, 3))) / 2e-6
        assert abs(gr.reshape(-1)[k] - fd) <= 1e-6 * abs(fd) + 1e-5
