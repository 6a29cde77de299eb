"""Pins the CPU oracle before anything is compared against it.

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a swapped index, a wrong sign, an off-by-one level shift) fails it:
literal known answers from the reference's tests, O(N^2) brute force for the
edge set and features, the reference's tree invariants, and closed forms /
finite differences for the pullback.
"""
import json
import os

import numpy as np
import pytest

from brute import bits, brute_edges, sort_edges
from octohip import datasets

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))


# ------------------------------------------------------------------ Morton


def test_morton_roundtrip(orc):
    for x, y, z in GOLD["morton_roundtrip_3d"]["cases"]:
        assert orc.morton_decode3(orc.morton_encode3(x, y, z)) == (x, y, z)


def test_morton_bit_layout(orc):
    # x -> bit 0, y -> bit 1, z -> bit 2 of every triple (src/morton.jl:35-40)
    assert orc.morton_encode3(1, 0, 0) == 1
    assert orc.morton_encode3(0, 1, 0) == 2
    assert orc.morton_encode3(0, 0, 1) == 4
    assert orc.morton_encode3(2, 0, 0) == 8
    top = orc.morton_encode3(0x1FFFFF, 0x1FFFFF, 0x1FFFFF)
    assert top == (1 << 63) - 1  # 63-bit code, top bit zero


def test_morton_monotone_x(orc):
    g = GOLD["morton_monotone_x"]
    prev = orc.morton_encode3(0, g["y"], g["z"])
    for x in range(1, g["x_max"] + 1):
        cur = orc.morton_encode3(x, g["y"], g["z"])
        assert cur > prev
        prev = cur


def test_bin_point_known_answers(orc):
    for c in GOLD["bin_point"]["cases"]:
        inv = np.float32(1.0) if c["inv_cs"] == "1f0" else np.float32(1.0) / np.float32(0.1)
        assert orc.bin_point(np.float32(c["p"]), np.float32(c["o"]), inv) == tuple(c["want"])


def test_origin_is_min_minus_1e6(orc):
    pts = datasets.uniform_cube(1000, seed=3, lo=-2.0, hi=5.0)
    o = orc.point_origin(pts)
    want = pts.min(axis=0) - np.float32(1e-6)
    assert np.array_equal(o, want.astype(np.float32))


def test_keys_closed_form(orc):
    pts = datasets.uniform_cube(5000, seed=4)
    r = np.float32(0.03)
    o = orc.point_origin(pts)
    k = orc.morton_keys(pts, o, r)
    inv = np.float32(1.0) / r
    cell = np.clip(np.floor((pts - o) * inv), 0, 0x1FFFFF).astype(np.uint64)

    def spread(v):
        out = np.zeros_like(v)
        for b in range(21):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return out

    want = spread(cell[:, 0]) | (spread(cell[:, 1]) << np.uint64(1)) | (spread(cell[:, 2]) << np.uint64(2))
    assert np.array_equal(k, want)


def test_sort_is_stable_permutation(orc):
    keys = (datasets.splitmix64(9, 0, 20000) % np.uint64(300)).astype(np.uint64)
    p = orc.sort_by_key(keys)
    assert np.array_equal(np.sort(p), np.arange(keys.shape[0]))
    assert np.array_equal(p, np.argsort(keys, kind="stable"))


# -------------------------------------------------------------------- tree


def assert_tree_invariants(t, pts):
    """test/build.jl:30-81, plus exact bounds and the leaf rule."""
    n, m = t.n, t.n_nodes
    assert m >= 1
    assert t.node_first[0] == 0 and t.node_last[0] == n - 1
    assert np.array_equal(np.sort(t.perm), np.arange(n))
    sk = t.keys[t.perm]
    assert np.all(sk[1:] >= sk[:-1])
    # A true leaf has all 8 children empty.  The reference tests only
    # children[1] == -1 (src/types.jl:22), which an internal node whose first
    # octant is empty also passes; traversal then scans that node's whole
    # range, which changes the work but not the edge set.
    leaf = np.all(t.node_children == -1, axis=1)
    for nid in range(m):
        f, l = t.node_first[nid], t.node_last[nid]
        cnt = l - f + 1
        assert cnt >= 1
        if leaf[nid]:
            assert np.all(t.node_children[nid] == -1)
            assert cnt <= t.leaf_size or t.node_level[nid] >= 21
            p = pts[t.perm[f:l + 1]]
            assert np.array_equal(t.bmin[nid], p.min(axis=0))
            assert np.array_equal(t.bmax[nid], p.max(axis=0))
        else:
            assert cnt > t.leaf_size and t.node_level[nid] < 21
            kids = t.node_children[nid]
            tot = 0
            prev_last = f - 1
            for o in range(8):
                c = kids[o]
                if c == -1:
                    continue
                assert c > nid
                assert t.node_level[c] == t.node_level[nid] + 1
                assert t.node_first[c] == prev_last + 1  # octant order, contiguous
                prev_last = t.node_last[c]
                shift = 63 - 3 * (t.node_level[nid] + 1)
                oc = (sk[t.node_first[c]:t.node_last[c] + 1] >> np.uint64(shift)) & np.uint64(7)
                assert np.all(oc == o)
                tot += t.node_last[c] - t.node_first[c] + 1
            assert tot == cnt and prev_last == l
            live = kids[kids != -1]
            assert np.array_equal(t.bmin[nid], t.bmin[live].min(axis=0))
            assert np.array_equal(t.bmax[nid], t.bmax[live].max(axis=0))
    # BFS numbering: levels are nondecreasing with node id
    assert np.all(np.diff(t.node_level) >= 0)


@pytest.mark.parametrize("n,r,seed", [(500, 0.08, 1), (3000, 0.05, 2)])
def test_tree_invariants_uniform(orc, n, r, seed):
    pts = datasets.uniform_cube(n, seed)
    assert_tree_invariants(orc.run(pts, r), pts)


def test_tree_invariants_clustered(orc):
    a = datasets.gaussian_blob(1000, 21, 0.1, (0.25, 0.25, 0.25))
    b = datasets.gaussian_blob(1000, 22, 0.1, (-0.25, -0.25, -0.25))
    pts = np.concatenate([a, b])
    assert_tree_invariants(orc.run(pts, 0.05), pts)


def test_tree_coincident_chain(orc):
    pts = np.zeros((64, 3), np.float32)
    t = orc.run(pts, 0.1)
    # All keys equal: a single-child chain down to level 21, one exhausted leaf.
    assert t.n_nodes == 22
    assert t.node_level[-1] == 21
    assert_tree_invariants(t, pts)


def test_single_point(orc):
    g = GOLD["single_point"]
    pts = np.zeros((1, 3), np.float32)
    t = orc.run(pts, g["r"])
    assert t.n_nodes == g["n_nodes"]
    nb = orc.for_each_neighbor(t, pts, 0)
    assert list(nb[nb != 0]) == g["neighbors"]


def test_empty_set(orc):
    t = orc.run(np.zeros((0, 3), np.float32), 0.1)
    assert t.n_nodes == GOLD["empty_set"]["n_nodes"]
    e = orc.build_edges(t, np.zeros((0, 3), np.float32))
    assert e.n_edges == GOLD["empty_set"]["n_edges"]


def test_node_count_per_point(orc):
    g = GOLD["memory_nodes_per_point"]
    pts = datasets.uniform_cube(g["n"], seed=7)
    t = orc.run(pts, g["r"])
    assert t.n_nodes / g["n"] <= g["ratio_max"]


# ------------------------------------------------------------------- edges


def _check_edges(orc, cq, r, ct=None):
    self_pair = ct is None
    t = orc.run(cq if self_pair else ct, r)
    e = orc.build_edges(t, cq, None if self_pair else ct)
    # receivers ascending, as the reference's offsets layout gives
    assert np.all(np.diff(e.receivers) >= 0)
    assert np.array_equal(np.diff(e.offsets), np.bincount(e.receivers, minlength=cq.shape[0]))
    got = sort_edges(e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm)
    want = brute_edges(cq, cq if self_pair else ct, r, exclude_self=self_pair)
    assert np.array_equal(got[0], want[0])
    assert np.array_equal(got[1], want[1])
    # features bit-exact (stronger than the reference's atol=1f-5)
    assert np.array_equal(bits(got[2]), bits(want[2]))
    assert np.array_equal(bits(got[3]), bits(want[3]))
    assert np.all(got[3] <= np.float32(1.0) + np.float32(1e-5))
    return e


def test_edges_n200_r015(orc):  # test/edges_cpu.jl:81-85
    _check_edges(orc, datasets.uniform_cube(200, 11), 0.15)


def test_edges_n1000_r005(orc):  # test/edges_cpu.jl:87-91
    _check_edges(orc, datasets.uniform_cube(1000, 21), 0.05)


def test_edges_clustered_2000(orc):  # test/edges_cpu.jl:93-97
    pts = (np.float32(0.1) * datasets.normal(31, 0, 6000).astype(np.float32)).reshape(2000, 3)
    _check_edges(orc, pts, 0.03)


def test_edges_asymmetric(orc):  # test/edges_cpu.jl:99-124
    a = datasets.uniform_cube(150, 41)
    b = datasets.uniform_cube(250, 42)
    _check_edges(orc, a, 0.12, ct=b)


def test_edges_chunked_matches_serial(orc):
    pts = datasets.uniform_cube(6000, 12)
    t = orc.run(pts, 0.06)
    e1 = orc.build_edges(t, pts, nthreads=1)
    e4 = orc.build_edges(t, pts, nthreads=4)
    assert np.array_equal(e1.senders, e4.senders)
    assert np.array_equal(bits(e1.rel_displacement), bits(e4.rel_displacement))


def test_count_and_rows_match_brute_force(orc):
    """The count-only pass and the selected-rows fill (the checkers of the
    4M-point workload) against the brute force, self pair and cross pair."""
    pts = datasets.uniform_cube(3000, 13)
    r = 0.07
    t = orc.run(pts, r)
    counts, total = orc.count_edges(t, pts, nthreads=3)
    want = brute_edges(pts, pts, r, exclude_self=True)
    assert total == want[0].shape[0]
    assert np.array_equal(counts, np.bincount(want[1], minlength=pts.shape[0]))
    rows = np.array([2999, 0, 17, 17, 1500], np.int64)  # any order, repeats allowed
    e = orc.build_edges_rows(t, pts, rows, nthreads=2)
    assert np.array_equal(np.diff(e.offsets), counts[rows])
    for k, i in enumerate(rows):
        a, b = int(e.offsets[k]), int(e.offsets[k + 1])
        assert np.all(e.receivers[a:b] == i)
        m = want[1] == i
        got = sort_edges(e.senders[a:b], e.receivers[a:b], e.rel_displacement[a:b], e.rel_dist_norm[a:b])
        ref = (want[0][m], want[1][m], want[2][m], want[3][m])
        for x, y in zip(got, ref):
            assert np.array_equal(bits(x), bits(y))
    q = datasets.uniform_cube(400, 14)
    cq, cross_total = orc.count_edges(t, q, pts)
    wq = brute_edges(q, pts, r, exclude_self=False)
    assert cross_total == wq[0].shape[0]
    assert np.array_equal(cq, np.bincount(wq[1], minlength=q.shape[0]))
    with pytest.raises(ValueError):
        orc.build_edges_rows(t, pts, np.array([3000], np.int64))


def test_boundary_exact_r(orc):
    for key in ("boundary_exact_r_included", "boundary_past_r_excluded"):
        g = GOLD[key]
        pts = np.array(g["coords"], np.float32)
        t = orc.run(pts, g["r"])
        n1 = [j + 1 for j in orc.for_each_neighbor(t, pts, 0) if j != 0]
        n2 = [j + 1 for j in orc.for_each_neighbor(t, pts, 1) if j != 1]
        assert n1 == g["neighbors_of_1"] and n2 == g["neighbors_of_2"]


def test_coincident_points(orc):
    g = GOLD["coincident_points"]
    pts = np.zeros((g["n"], 3), np.float32)
    e = _check_edges(orc, pts, g["r"])
    assert e.n_edges == g["n"] * (g["n"] - 1)


def test_negative_coords(orc):  # test/edge_cases.jl:132-155
    pts = datasets.uniform_cube(200, 3001, lo=-1.0, hi=1.0)
    _check_edges(orc, pts, 0.1)


def test_tiny_and_huge_radius(orc):
    g = GOLD["tiny_radius"]
    e = _check_edges(orc, datasets.uniform_cube(g["n"], 3002), g["r"])
    assert e.n_edges == 0
    g = GOLD["huge_radius"]
    e = _check_edges(orc, datasets.uniform_cube(g["n"], 3003), g["r"])
    assert e.n_edges == g["n"] * g["neighbors_each"]


def test_neighborlist_doctest(orc):
    g = GOLD["neighborlist_doc"]
    pts = np.array(g["coords"], np.float32)
    t = orc.run(pts, g["r"])
    got = [j + 1 for j in orc.for_each_neighbor(t, pts, 0) if j != 0]
    assert got == g["neighbors_of_1"]


def test_for_each_neighbor_includes_self(orc):
    pts = datasets.uniform_cube(300, 5)
    t = orc.run(pts, 0.1)
    for i in (0, 17, 299):
        nb = orc.for_each_neighbor(t, pts, i)
        assert i in set(nb.tolist())


# ---------------------------------------------------------------- pullback


def _central_diff(loss, x, eps):
    """4-point central difference in FP64 (test/chainrules.jl:23-34)."""
    g = np.zeros_like(x)
    flat = x.reshape(-1)
    gf = g.reshape(-1)
    for k in range(flat.shape[0]):
        v = flat[k]
        vals = []
        for s in (2, 1, -1, -2):
            flat[k] = v + s * eps
            vals.append(loss(x))
        flat[k] = v
        gf[k] = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * eps)
    return g


def _loss64(x, s, rc, r, wd, wn):
    d = (x[rc] - x[s]) / r
    return wd * np.sum(d * d) + wn * np.sum(np.sum(d * d, axis=1))


@pytest.mark.parametrize("wd,wn", [(1.0, 0.5), (1.0, 0.0), (0.0, 1.0)])
def test_pullback_vs_fd64(orc, wd, wn):
    pts = datasets.uniform_cube(40, 8002)
    r = 0.25
    t = orc.run(pts, r)
    e = orc.build_edges(t, pts)
    dd = None if wd == 0 else 2 * wd * e.rel_displacement
    dn = None if wn == 0 else 2 * wn * e.rel_dist_norm
    g = orc.pullback(40, e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm, dd, dn, r)
    x = pts.astype(np.float64)
    fd = _central_diff(lambda c: _loss64(c, e.senders, e.receivers, r, wd, wn), x, 1e-4)
    np.testing.assert_allclose(g, fd, rtol=1e-4, atol=1e-4)


def test_pullback_closed_form(orc):
    """L = sum|rdisp|^2 + 0.5 sum rdist^2 on a symmetric edge set gives
    g_i = (6/r^2) * sum_j (x_i - x_j)."""
    pts, _ = datasets.differentiable(4096, seed=5)
    r = np.float32(0.1)
    t = orc.run(pts, r)
    e = orc.build_edges(t, pts)
    g = orc.pullback(pts.shape[0], e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm,
                     2 * e.rel_displacement, e.rel_dist_norm, r)
    x = pts.astype(np.float64)
    # receiver side only: each i sums over its own neighbour list
    want = np.zeros_like(x)
    np.add.at(want, e.receivers, 6.0 / float(r) ** 2 * (x[e.receivers] - x[e.senders]))
    np.testing.assert_allclose(g, want, rtol=2e-4, atol=2e-3)


def test_pullback_zero_tangents(orc):
    pts = datasets.uniform_cube(30, 8006)
    t = orc.run(pts, 0.2)
    e = orc.build_edges(t, pts)
    g = orc.pullback(30, e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm, None, None, 0.2)
    assert np.all(g == 0)


def test_zero_displacement_component_is_negative_zero(orc):
    """The CPU fill writes -dx * inv_r with dx = target - query
    (src/cpu/edges_cpu.jl:111-113, src/cpu/iterator_cpu.jl:96): a component
    where the two points agree is -(+0.0) * inv_r = -0.0, sign bit set."""
    pts = np.array([[0.25, 0.0, 0.5], [0.25, 0.5, 0.5]], np.float32)
    e = orc.build_edges(orc.run(pts, 1.0), pts)
    assert e.n_edges == 2
    d = e.rel_displacement.view(np.int32)
    assert np.all(d[:, 0] == np.int32(-2**31)) and np.all(d[:, 2] == np.int32(-2**31))
    assert np.all(e.rel_displacement[:, 1] == np.float32([-0.5, 0.5])[e.receivers])
    b = brute_edges(pts, pts, 1.0, exclude_self=True)
    assert np.array_equal(bits(sort_edges(*b)[2]), bits(sort_edges(e.senders, e.receivers, e.rel_displacement,
                                                                     e.rel_dist_norm)[2]))


@pytest.mark.parametrize("n,r,seed", [(50, 0.15, 10), (500, 0.05, 20), (2000, 0.1, 30), (500, 0.08, 60)])
def test_materialize_matches_brute_force(orc, n, r, seed):
    """CSR lists (test/query_cpu.jl:14-52,83-106): each query's slice is its
    brute-force neighbour set, self excluded."""
    pts = datasets.uniform_cube(n, seed)
    off, flat = orc.materialize(orc.run(pts, r), pts)
    assert off.dtype == np.int32 and off[0] == 0 and off.shape[0] == n + 1
    s, rc, _, _ = brute_edges(pts, pts, r, exclude_self=True)
    assert np.array_equal(np.diff(off), np.bincount(rc, minlength=n))
    for i in range(n):
        assert np.array_equal(np.sort(flat[off[i]:off[i + 1]]), s[rc == i])


def test_materialize_cross_pair_and_clustered(orc):
    a = datasets.uniform_cube(200, 51)
    b = datasets.uniform_cube(300, 52)
    off, flat = orc.materialize(orc.run(b, 0.1), a, b)
    s, rc, _, _ = brute_edges(a, b, 0.1, exclude_self=False)
    for i in range(200):
        assert np.array_equal(np.sort(flat[off[i]:off[i + 1]]), s[rc == i])
    pts = (np.float32(0.1) * datasets.normal(40, 0, 15000).astype(np.float32)).reshape(5000, 3)
    off, flat = orc.materialize(orc.run(pts, 0.03), pts)
    s, rc, _, _ = brute_edges(pts, pts, 0.03, exclude_self=True)
    assert np.array_equal(np.diff(off), np.bincount(rc, minlength=5000))
    assert np.array_equal(np.sort(flat[off[100]:off[101]]), s[rc == 100])


def test_zsort_gather_and_scatter(orc):
    """apply_zsort on the reference's layout: the gather follows perm, the
    scatter is its transpose (<S x, y> = <x, S^T y> for random x, y) and
    undoes it."""
    pts = datasets.uniform_cube(700, 71)
    t = orc.run(pts, 0.05)
    g = np.random.default_rng(3)
    a = g.normal(size=(700, 5)).astype(np.float32)
    z = orc.apply_zsort(t.perm, a)
    assert np.array_equal(z[0], a[t.perm[0]]) and np.array_equal(z[-1], a[t.perm[-1]])
    v = np.arange(700, dtype=np.int32)
    assert np.array_equal(orc.apply_zsort(t.perm, v), t.perm)  # gathering the identity gives perm itself
    y = g.normal(size=(700, 5))
    x = g.normal(size=(700, 5))
    assert np.isclose((orc.apply_zsort(t.perm, x) * y).sum(), (x * orc.zsort_pullback(t.perm, y)).sum())
    assert np.array_equal(orc.zsort_pullback(t.perm, z), a)
