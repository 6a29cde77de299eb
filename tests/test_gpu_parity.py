"""GPU parity: run! and build_edges through the C ABI against the CPU oracle.

Bar: bit-exact for indices, tree arrays and (sorted) edge sets; the edge
features are compared bit-exactly too (the engine builds with
-ffp-contract=off, so no 1-ulp slack is needed).  Inputs are the reference
tests' shapes plus the workload configurations, seeded.
"""
import numpy as np
import pytest

from brute import bits, brute_edges, sort_edges
from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _gpu(pts):
    return _torch().from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()


def _build(pts, r, leaf=32, base=0):
    import octohip

    tns = octohip.TNS(index_base=base)
    octohip.set_search_radius(tns, r)
    if leaf != 32:
        octohip.set_leaf_size(tns, leaf)
    c = _gpu(pts)
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid, c


def _np(t):
    return t.detach().cpu().numpy()


TREE_CASES = {
    "uniform500": (lambda: datasets.uniform_cube(500, 1), 0.08),
    "uniform3000": (lambda: datasets.uniform_cube(3000, 2), 0.05),
    "blobs2000": (lambda: np.concatenate([datasets.gaussian_blob(1000, 21, 0.1, (0.25,) * 3),
                                          datasets.gaussian_blob(1000, 22, 0.1, (-0.25,) * 3)]), 0.05),
    "coincident64": (lambda: np.zeros((64, 3), np.float32), 0.1),
    "single": (lambda: np.zeros((1, 3), np.float32), 1.0),
    "plain100k": (lambda: datasets.plain_cpu()[0], 0.02),
    "negative": (lambda: datasets.uniform_cube(2000, 3001, lo=-1.0, hi=1.0), 0.1),
}


@pytest.mark.parametrize("case", sorted(TREE_CASES))
@pytest.mark.parametrize("leaf", [32, 8])
def test_tree_parity(orc, case, leaf):
    import octohip

    make, r = TREE_CASES[case]
    pts = make()
    print(f"tree parity case={case} n={pts.shape[0]} leaf={leaf}", flush=True)
    tns, sid, _ = _build(pts, r, leaf=leaf)
    t = octohip.get_tree(tns, sid)
    o = orc.run(pts, r, leaf_size=leaf)
    assert t.n_nodes == o.n_nodes
    assert np.array_equal(_np(t.permutation), o.perm)
    assert np.array_equal(_np(t.node_first), o.node_first)
    assert np.array_equal(_np(t.node_last), o.node_last)
    assert np.array_equal(_np(t.node_children), o.node_children)
    assert np.array_equal(_np(t.node_bounds_min), o.bmin)
    assert np.array_equal(_np(t.node_bounds_max), o.bmax)
    assert np.array_equal(np.array(t.origin, np.float32), o.origin)
    assert t.n_leaves == int(np.sum(np.all(o.node_children == -1, axis=1)))


def test_tree_parity_many_leaves(orc):
    # 740k leaves at levels 12-21: the bounds refit strides over more leaf
    # blocks than its grid holds, and the per-level merge of the runs cut at
    # block edges goes through several chunks with the open run carried
    import octohip

    pts = np.concatenate([datasets.gaussian_blob(1_000_000, 31, 0.1, (0.25,) * 3),
                          datasets.gaussian_blob(1_000_000, 32, 0.05, (-0.25,) * 3)])
    r = 0.002
    print(f"tree parity many leaves n={pts.shape[0]} r={r} leaf=8", flush=True)
    tns, sid, _ = _build(pts, r, leaf=8)
    t = octohip.get_tree(tns, sid)
    o = orc.run(pts, r, leaf_size=8)
    assert t.n_leaves > 600_000
    assert t.n_nodes == o.n_nodes
    assert np.array_equal(_np(t.node_first), o.node_first)
    assert np.array_equal(_np(t.node_last), o.node_last)
    assert np.array_equal(_np(t.node_children), o.node_children)
    assert np.array_equal(_np(t.node_bounds_min), o.bmin)
    assert np.array_equal(_np(t.node_bounds_max), o.bmax)


def test_tree_index_base_one(orc):
    import octohip

    pts = datasets.uniform_cube(3000, 2)
    tns, sid, _ = _build(pts, 0.05, base=1)
    assert sid == 1
    t = octohip.get_tree(tns, sid)
    o = orc.run(pts, 0.05)
    assert np.array_equal(_np(t.permutation), o.perm + 1)
    assert np.array_equal(_np(t.node_first), o.node_first + 1)
    kids = o.node_children.copy()
    kids[kids >= 0] += 1
    assert np.array_equal(_np(t.node_children), kids)


def _edges_equal(got, want):
    g = sort_edges(*got)
    w = sort_edges(*want)
    assert np.array_equal(g[0], w[0]), "senders"
    assert np.array_equal(g[1], w[1]), "receivers"
    assert np.array_equal(bits(g[2]), bits(w[2])), "rel_displacement"
    assert np.array_equal(bits(g[3]), bits(w[3])), "rel_dist_norm"


def _gpu_edges(e):
    return (_np(e.senders), _np(e.receivers), _np(e.rel_displacement), _np(e.rel_dist_norm).reshape(-1))


def _assert_receiver_runs(recv):
    # each receiver's edges are one contiguous run
    if recv.shape[0] == 0:
        return
    starts = np.concatenate([[True], recv[1:] != recv[:-1]])
    assert np.unique(recv[starts]).shape[0] == starts.sum()


EDGE_CASES = {
    "n200_r015": (lambda: datasets.uniform_cube(200, 11), 0.15),
    "n1000_r005": (lambda: datasets.uniform_cube(1000, 21), 0.05),
    "clustered2000": (lambda: (np.float32(0.1) * datasets.normal(31, 0, 6000).astype(np.float32)).reshape(2000, 3), 0.03),
    "coincident100": (lambda: np.zeros((100, 3), np.float32), 0.1),
    "negative200": (lambda: datasets.uniform_cube(200, 3001, lo=-1.0, hi=1.0), 0.1),
    "tiny_radius": (lambda: datasets.uniform_cube(200, 3002), 1e-8),
    "huge_radius": (lambda: datasets.uniform_cube(100, 3003), 100.0),
    "plain100k": (lambda: datasets.plain_cpu()[0], 0.02),
    "diff262k": (lambda: datasets.differentiable()[0], float(datasets.differentiable()[1])),
}


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_edges_parity(orc, case):
    import octohip

    make, r = EDGE_CASES[case]
    pts = make()
    print(f"edges parity case={case} n={pts.shape[0]} r={r}", flush=True)
    tns, sid, _ = _build(pts, r)
    e = octohip.build_edges(tns, sid, sid)
    got = _gpu_edges(e)
    o = orc.build_edges(orc.run(pts, r), pts)
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))
    _assert_receiver_runs(got[1])
    if pts.shape[0] <= 2000:
        _edges_equal(got, brute_edges(pts, pts, r, exclude_self=True))


def test_edges_receiver_order(orc):
    import octohip

    pts = datasets.uniform_cube(20000, 5)
    r = 0.04
    tns, sid, _ = _build(pts, r)
    e = octohip.build_edges(tns, sid, sid, order="receiver")
    got = _gpu_edges(e)
    o = orc.build_edges(orc.run(pts, r), pts)
    assert np.all(np.diff(got[1]) >= 0)  # receivers ascending, as the reference writes them
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


def test_edges_asymmetric(orc):
    import octohip

    a = datasets.uniform_cube(150, 41)
    b = datasets.uniform_cube(250, 42)
    r = 0.12
    tns = octohip.TNS()
    octohip.set_search_radius(tns, r)
    ia = octohip.add_point_set(tns, _gpu(a))
    ib = octohip.add_point_set(tns, _gpu(b))
    octohip.set_active_search(tns, ia, ib)
    octohip.run(tns)
    e = octohip.build_edges(tns, ia, ib)
    _edges_equal(_gpu_edges(e), brute_edges(a, b, r, exclude_self=False))
    o = orc.build_edges(orc.run(b, r), a, b)
    _edges_equal(_gpu_edges(e), (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


def test_edges_sph_lattice(orc):
    """Config: 1M-point SPH block (100^3 lattice, h = 0.01, r = 2h)."""
    import octohip

    pts, r = datasets.sph_lattice()
    tns, sid, _ = _build(pts, r)
    e = octohip.build_edges(tns, sid, sid)
    got = _gpu_edges(e)
    o = orc.build_edges(orc.run(pts, r), pts)
    assert got[0].shape[0] == o.n_edges
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


def test_edge_count_beyond_int32_is_an_error():
    """50,000 points with a radius covering the cube give 2.5e9 > 2^31 edges:
    the engine must refuse before allocating (the reference's Int32 limit)."""
    import octohip

    pts = datasets.uniform_cube(50_000, 91)
    tns, sid, _ = _build(pts, 2.0)
    with pytest.raises(octohip.OctopusError, match="Int32"):
        octohip.build_edges(tns, sid, sid)


def test_empty_and_single():
    import octohip

    tns, sid, _ = _build(np.zeros((0, 3), np.float32), 0.1)
    e = octohip.build_edges(tns, sid, sid)
    assert e.senders.numel() == 0 and tuple(e.rel_displacement.shape) == (0, 3)
    tns, sid, _ = _build(np.zeros((1, 3), np.float32), 1.0)
    assert octohip.get_tree(tns, sid).n_nodes == 1
    e = octohip.build_edges(tns, sid, sid)
    assert e.senders.numel() == 0


def test_rebuild_after_update(orc):
    import octohip

    p1 = datasets.uniform_cube(150, 71)
    p2 = datasets.uniform_cube(150, 72)
    tns, sid, c1 = _build(p1, 0.1)
    n1 = octohip.build_edges(tns, sid, sid).senders.numel()
    assert n1 > 0
    octohip.update_point_set(tns, sid, _gpu(p2))
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid)
    _edges_equal(_gpu_edges(e), brute_edges(p2, p2, 0.1, exclude_self=True))


def test_device_iterator_counts(orc):
    import octohip

    pts = datasets.uniform_cube(5000, 9)
    r = 0.06
    tns, sid, _ = _build(pts, r)
    v = octohip.device_view(tns, sid, sid)
    cnt = _np(octohip.count_neighbors_device(tns, v))
    o = orc.run(pts, r)
    want = np.array([orc.for_each_neighbor(o, pts, i).shape[0] for i in range(pts.shape[0])])
    assert np.array_equal(cnt, want)  # self included, as the device iterator does not filter


# ----------------------------------------------------------------- backward


def test_backward_vs_oracle_pullback(orc):
    import octohip

    pts, r = datasets.differentiable()
    tns, sid, c = _build(pts, r)
    e = octohip.build_edges(tns, sid, sid)
    s, rc, d, n = _gpu_edges(e)
    torch = _torch()
    dd = torch.randn(d.shape, generator=torch.Generator().manual_seed(0)).float()
    dn = torch.randn(n.shape, generator=torch.Generator().manual_seed(1)).float()
    g = octohip._pullback(tns, pts.shape[0], e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm,
                          dd.cuda(), dn.cuda(), r)
    want = orc.pullback(pts.shape[0], s, rc, d, n, dd.numpy(), dn.numpy(), r)
    np.testing.assert_allclose(_np(g), want, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("layout", ["leaf", "receiver", "shuffled", "one_way"])
def test_backward_pull_and_push_forms(orc, layout):
    """The pull form (sender terms taken from the reverse edges) and the push
    form (atomics) against the oracle pullback on the differentiable
    configuration: build_edges' two layouts, a shuffled edge list (receivers
    not in runs: the pull form must notice and push), and a list with half of
    the reverse edges dropped (unpaired edges push)."""
    import octohip

    torch = _torch()
    pts, r = datasets.differentiable()
    tns, sid, c = _build(pts, r)
    e = octohip.build_edges(tns, sid, sid, order="receiver" if layout == "receiver" else "leaf")
    s, rc, d, n = (x.clone() for x in e)
    if layout == "shuffled":
        p = torch.randperm(s.shape[0], generator=torch.Generator().manual_seed(3)).cuda()
        s, rc, d, n = s[p], rc[p], d[p], n[p]
    if layout == "one_way":
        keep = (s < rc) | (rc % 2 == 0)
        s, rc, d, n = s[keep], rc[keep], d[keep], n[keep]
    g0 = torch.Generator().manual_seed(0)
    dd = torch.randn(d.shape, generator=g0).float()
    dn = torch.randn(n.shape, generator=g0).float()
    want = orc.pullback(pts.shape[0], _np(s), _np(rc), _np(d), _np(n).reshape(-1), dd.numpy(), dn.numpy(), r)
    for pull in (True, False):
        octohip.set_backward_form(tns, pull)
        g = octohip._pullback(tns, pts.shape[0], s, rc, d, n, dd.cuda(), dn.cuda(), r)
        print(f"backward layout={layout} pull={pull} edges={s.shape[0]}", flush=True)
        np.testing.assert_allclose(_np(g), want, rtol=1e-4, atol=1e-3)


def test_backward_closed_form():
    """loss = sum|rdisp|^2 + 0.5 sum rdist^2 -> g_i = (6/r^2) sum_j (x_i - x_j)."""
    import octohip

    torch = _torch()
    pts, r = datasets.differentiable()
    tns, sid, c = _build(pts, r)
    x = c.clone().requires_grad_(True)
    e = octohip.build_edges_diff(x, tns, sid, r)
    loss = (e.rel_displacement ** 2).sum() + 0.5 * (e.rel_dist_norm ** 2).sum()
    loss.backward()
    s, rc = _np(e.senders), _np(e.receivers)
    xd = pts.astype(np.float64)
    want = np.zeros_like(xd)
    np.add.at(want, rc, 6.0 / float(np.float32(r)) ** 2 * (xd[rc] - xd[s]))
    np.testing.assert_allclose(_np(x.grad), want, rtol=1e-3, atol=5e-2)


def test_backward_fd64_subset():
    """Central differences in FP64 on a 2,048-point subset (radius ~0.15)."""
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(2048, 55)
    r = 0.15
    tns, sid, c = _build(pts, r)
    x = c.clone().requires_grad_(True)
    e = octohip.build_edges_diff(x, tns, sid, r)
    loss = (e.rel_displacement ** 2).sum() + 0.5 * (e.rel_dist_norm ** 2).sum()
    loss.backward()
    s, rc = _np(e.senders).astype(np.int64), _np(e.receivers).astype(np.int64)
    xd = pts.astype(np.float64)
    rd = float(np.float32(r))

    def f(z):
        d = (z[rc] - z[s]) / rd
        q = np.sum(d * d, axis=1)
        return np.sum(q) + 0.5 * np.sum(q)

    rng = np.random.default_rng(0)
    idx = rng.choice(xd.size, 64, replace=False)
    g = _np(x.grad).reshape(-1)
    eps = 1e-4
    for k in idx:
        vals = []
        for sgn in (2, 1, -1, -2):
            z = xd.copy().reshape(-1)
            z[k] += sgn * eps
            vals.append(f(z.reshape(xd.shape)))
        fd = (-vals[0] + 8 * vals[1] - 8 * vals[2] + vals[3]) / (12 * eps)
        assert abs(g[k] - fd) <= 1e-4 * abs(fd) + 2e-3, (k, g[k], fd)


def test_backward_partial_tangents(orc):
    import octohip

    pts = datasets.uniform_cube(3000, 57)
    r = 0.08
    tns, sid, c = _build(pts, r)
    for wd, wn in ((1.0, 0.0), (0.0, 1.0)):
        x = c.clone().requires_grad_(True)
        e = octohip.build_edges_diff(x, tns, sid, r)
        loss = wd * (e.rel_displacement ** 2).sum() + wn * (e.rel_dist_norm ** 2).sum()
        loss.backward()
        s, rc, d, n = (_np(e.senders), _np(e.receivers), _np(e.rel_displacement), _np(e.rel_dist_norm).reshape(-1))
        want = orc.pullback(3000, s, rc, d, n, None if wd == 0 else 2 * d, None if wn == 0 else 2 * n, r)
        np.testing.assert_allclose(_np(x.grad), want, rtol=1e-4, atol=1e-3)


# ------------------------------------------------------------- single pass

SINGLE_CASES = ["clustered2000", "coincident100", "huge_radius", "plain100k", "diff262k", "n200_r015"]


@pytest.mark.parametrize("case", SINGLE_CASES)
def test_single_pass_parity(orc, case):
    """mode="learned": one traversal with in-kernel exact counts and a
    look-back for the offsets; the first call sizes the output buffers, the
    second runs the single pass.  Edge set and features bit-exact."""
    import octohip

    make, r = EDGE_CASES[case]
    pts = make()
    print(f"single pass case={case} n={pts.shape[0]}", flush=True)
    tns, sid, _ = _build(pts, r)
    o = orc.build_edges(orc.run(pts, r), pts)
    want = (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    for call in range(2):
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        got = _gpu_edges(e)
        assert got[0].shape[0] == o.n_edges, (call, got[0].shape[0], o.n_edges)
        _edges_equal(got, want)
        _assert_receiver_runs(got[1])


def test_single_pass_capacity_regrow(orc):
    """A larger point set after a small one: the single pass runs out of room,
    reports its total and is rerun with a larger buffer."""
    import octohip

    small = datasets.uniform_cube(2000, 81)
    big = datasets.uniform_cube(30000, 82)
    r = 0.05
    tns, sid, _ = _build(small, r)
    octohip.build_edges(tns, sid, sid, mode="learned")
    octohip.build_edges(tns, sid, sid, mode="learned")
    octohip.resize_point_set(tns, sid, _gpu(big))
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid, mode="learned")
    o = orc.build_edges(orc.run(big, r), big)
    _edges_equal(_gpu_edges(e), (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


def test_single_pass_heavy_leaves(orc):
    """Uneven leaves (a Gaussian mixture at ~20 neighbours on average, so
    cores hold leaves with thousands of hits): the heavy ones are counted by
    the parallel count pass before the single pass.  The learned edges must be
    exactly the exact path's, in the same order, and the oracle's set; on
    even data the pre-count switches off after the first call."""
    import octohip
    from octohip import predictor

    pts = datasets.gaussian_mixture(30000, 41, 4, 0.02, 0.1)
    r = float(datasets.calibrate_radius(pts, 20.0, seed=1))
    tns, sid, _ = _build(pts, r)
    ex = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="exact"))
    o = orc.build_edges(orc.run(pts, r), pts)
    octohip.build_edges(tns, sid, sid, mode="learned")  # sizes the single pass's buffers
    for call in range(3):
        got = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="learned"))
        st = predictor.edge_stats(tns, sid, sid)
        print(f"mixture call {call}: heavy={st['heavy_leaves']} recounted={st['recounted_leaves']} "
              f"recorded={st['recorded_leaves']} single={st['single_pass']}", flush=True)
        assert st["single_pass"]
        assert st["heavy_leaves"] > 0
        for a, b in zip(got, ex):  # same edges in the same order
            assert np.array_equal(bits(a), bits(b))
        _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))

    even = datasets.uniform_cube(20000, 42)
    tns, sid, _ = _build(even, 0.03)
    o = orc.build_edges(orc.run(even, 0.03), even)
    heavy = []
    for call in range(3):
        got = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="learned"))
        st = predictor.edge_stats(tns, sid, sid)
        heavy.append(st["heavy_leaves"])
        _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))
    print(f"uniform heavy per call {heavy}", flush=True)
    assert heavy[2] == 0


def test_dense_fill_many_queries_and_candidates(orc):
    """A dense core (2,500 points in a cube of side r/4, so every pair is an
    edge) over a uniform background: its leaves hold hundreds to thousands of
    queries (the per-query cursors are scanned in several block-wide
    chunks) and more candidates within r of the leaf box than one staging
    round of the dense fill takes.  The learned call (heavy leaves through
    the dense fill) writes exactly the exact path's edges, in the same order,
    and both are the oracle's set."""
    import octohip
    from octohip import predictor

    r = 0.02
    core = 0.5 + (datasets.uniform(71, 0, 3 * 2500).reshape(2500, 3) * (r / 4))
    back = datasets.uniform(72, 0, 3 * 3000).reshape(3000, 3)
    pts = np.concatenate([core, back]).astype(np.float32)
    tns, sid, _ = _build(pts, r)
    ex = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="exact"))
    octohip.build_edges(tns, sid, sid, mode="learned")  # sizes the single pass's buffers
    got = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="learned"))
    st = predictor.edge_stats(tns, sid, sid)
    print(f"dense core: {st['heavy_leaves']} heavy leaves, {st['n_edges']} edges", flush=True)
    assert st["single_pass"] and st["heavy_leaves"] > 0
    assert st["n_edges"] > 2500 * 2499
    for a, b in zip(got, ex):  # same edges in the same order
        assert np.array_equal(bits(a), bits(b))
    o = orc.build_edges(orc.run(pts, r), pts)
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


def test_planned_first_call_with_heavy_leaves(orc):
    """A pair's first learned call under the predictor's plan where many
    leaves are heavy (1M points of the clustered distribution, the shipped
    clustered4m predictor): heavy leaves cost nothing in the plan's packing,
    their edges come from the dense fill, and the edges equal the exact
    path's position for position and the oracle's set."""
    import octohip
    from octohip import predictor

    pts, r = datasets.clustered(n=1_000_000, seed=9)
    tns, sid, _ = _build(pts, float(r))
    predictor.load(tns, "clustered4m")
    got = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="learned"))
    st = predictor.edge_stats(tns, sid, sid)
    ex = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="exact"))
    print(f"clustered 1M: {st['heavy_leaves']} heavy of {st['n_leaves']} leaves, {st['n_edges']} edges, "
          f"predicted {st['predicted_total']:.0f}", flush=True)
    assert st["single_pass"] and st["used_predictor"] and st["heavy_leaves"] > 0
    for a, b in zip(got, ex):  # same edges in the same order
        assert np.array_equal(bits(a), bits(b))
    o = orc.build_edges(orc.run(pts, float(r)), pts)
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


PLAN_CASES = {
    "plain100k_leaf32": ("plain100k", 32),
    "plain100k_leaf8": ("plain100k", 8),
    "clustered2000": ("clustered2000", 32),
    "coincident100": ("coincident100", 32),
    "huge_radius": ("huge_radius", 32),
}


@pytest.mark.parametrize("case", sorted(PLAN_CASES))
def test_single_pass_plan_every_call(orc, case):
    """The single pass under the predictor's plan on every call (runs of
    uneven length, packed by predicted hits) writes the same edges in the
    same positions as with fixed runs, and both are the oracle's set."""
    import octohip
    from octohip import predictor

    name, leaf = PLAN_CASES[case]
    make, r = EDGE_CASES[name]
    pts = make()
    print(f"single pass plan case={case}", flush=True)
    o = orc.build_edges(orc.run(pts, r, leaf_size=leaf), pts)
    want = (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    outs = []
    for plan in (False, True):
        tns, sid, _ = _build(pts, r, leaf=leaf)
        if plan:
            predictor.load(tns, "sph_lattice")
            predictor.set_policy(tns, True)
        for call in range(2):
            e = octohip.build_edges(tns, sid, sid, mode="learned")
            got = tuple(_np(x).copy() for x in (e.senders, e.receivers, e.rel_displacement, e.rel_dist_norm))
            _edges_equal(_gpu_edges(e), want)
        st = predictor.edge_stats(tns, sid, sid)
        assert st["single_pass"], st
        assert st["used_predictor"] == plan, st
        outs.append(got)
    for a, b in zip(*outs):
        assert np.array_equal(bits(a), bits(b))


def test_single_pass_list_overflow_and_heavy(orc):
    """Leaves whose hits overflow their run's hit list (long runs of dense
    leaves) are re-swept in phase B from the per-query counts of their phase
    A sweep, and heavy leaves are filled after the pass: the edges stay the
    exact path's, in the same order."""
    import octohip
    from octohip import predictor

    pts = datasets.gaussian_mixture(30000, 43, 4, 0.02, 0.1)
    r = float(datasets.calibrate_radius(pts, 40.0, seed=3))
    tns, sid, _ = _build(pts, r)
    ex = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="exact"))
    octohip.set_run_leaves(tns, 64)  # runs far beyond one slot's hits
    for call in range(3):
        got = _gpu_edges(octohip.build_edges(tns, sid, sid, mode="learned"))
        st = predictor.edge_stats(tns, sid, sid)
        print(f"overflow call {call}: {st['recounted_leaves']} recounted, {st['heavy_leaves']} heavy", flush=True)
        for a, b in zip(got, ex):
            assert np.array_equal(bits(a), bits(b))
    assert st["recounted_leaves"] > 0
    o = orc.build_edges(orc.run(pts, r), pts)
    _edges_equal(got, (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


@pytest.mark.parametrize("mode", ["exact", "learned"])
def test_leaf_count_growth_between_builds(orc, mode):
    """run! does not wait for the tree counts; build_edges sizes its per-leaf
    arrays from the last count read back plus a margin.  A rebuild with four
    times as many leaves (leaf size 32 -> 8) overflows that margin: the
    kernels see a clamped leaf count, the call notices at its wait and redoes
    the edges with the real count.  Edges must still match the oracle."""
    import octohip
    from octohip import predictor

    pts = datasets.uniform_cube(120000, 61)
    r = 0.02
    tns, sid, c = _build(pts, r)
    octohip.build_edges(tns, sid, sid, mode=mode)
    octohip.build_edges(tns, sid, sid, mode=mode)
    n32 = octohip.get_tree(tns, sid).node_first.shape[0]
    octohip.set_leaf_size(tns, 8)
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid, mode=mode)
    st = predictor.edge_stats(tns, sid, sid)
    print(f"mode={mode} nodes@32={n32} leaves@8={st['n_leaves']} attempts={st['attempts']}", flush=True)
    assert st["attempts"] >= 2
    o = orc.build_edges(orc.run(pts, r, 8), pts)
    _edges_equal(_gpu_edges(e), (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))
    e = octohip.build_edges(tns, sid, sid, mode=mode)  # the count is known now: one attempt
    assert predictor.edge_stats(tns, sid, sid)["attempts"] == 1
    _edges_equal(_gpu_edges(e), (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm))


@pytest.mark.parametrize("private", [False, True])
def test_caller_streams(orc, private):
    """The library follows torch's current stream (or, private_stream, joins
    its own stream to it): results produced under a side stream and then on
    the default stream are the oracle's, with no explicit synchronization."""
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(30000, 71)
    r = 0.03
    o = orc.build_edges(orc.run(pts, r), pts)
    want = (o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    tns = octohip.TNS(private_stream=private)
    octohip.set_search_radius(tns, r)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        c = torch.from_numpy(pts).cuda()
        sid = octohip.add_point_set(tns, c)
        octohip.set_active_search(tns, sid, sid)
        octohip.run(tns)
        for mode in ("exact", "learned"):
            _edges_equal(_gpu_edges(octohip.build_edges(tns, sid, sid, mode=mode)), want)
    c2 = c.clone()  # on the default stream, after the side stream's work
    torch.cuda.current_stream().wait_stream(side)
    octohip.update_point_set(tns, sid, c2)
    octohip.run(tns)
    for mode in ("exact", "learned"):
        _edges_equal(_gpu_edges(octohip.build_edges(tns, sid, sid, mode=mode)), want)


def test_single_pass_returns_on_its_total(orc):
    """With fixed runs the learned call returns as soon as the single pass
    knows its total, while the pass's emit still runs: work the caller then
    queues on its stream (here: sums of the outputs, then the next step) must
    see complete edges, step after step, and the statistics must be final."""
    import octohip
    from octohip import predictor

    torch = _torch()
    rng = np.random.default_rng(17)
    pts = datasets.uniform_cube(40000, 77)
    r = 0.03
    tns, sid, c = _build(pts, r)
    octohip.set_refit_mode(tns, True)
    octohip.run(tns)
    octohip.build_edges(tns, sid, sid, mode="learned")  # sizes the buffers
    sums = []
    steps = []
    cur = pts
    for k in range(4):
        cur = (cur.astype(np.float64) + rng.normal(0, 2e-4, cur.shape)).astype(np.float32)
        octohip.update_point_set(tns, sid, _gpu(cur))
        octohip.run(tns)
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        # queued on the caller's stream behind the emit: no explicit wait
        sums.append(torch.stack([e.senders.double().sum(), e.receivers.double().sum(),
                                 e.rel_dist_norm.double().sum(), e.rel_displacement.double().abs().sum()]))
        steps.append(cur)
    pk = predictor.edge_stats(tns, sid, sid, tensors=False, wait=False)  # no wait for the stream
    st = predictor.edge_stats(tns, sid, sid)
    assert st["single_pass"] and st["recorded_leaves"] + st["recounted_leaves"] == st["n_leaves"]
    for k in ("n_edges", "n_leaves", "single_pass", "used_predictor", "attempts", "heavy_leaves"):
        assert pk[k] == st[k], k
    for k, cur in enumerate(steps):
        o = orc.build_edges(orc.run(cur, r), cur)
        want = np.array([o.senders.astype(np.float64).sum(), o.receivers.astype(np.float64).sum(),
                         o.rel_dist_norm.astype(np.float64).sum(), np.abs(o.rel_displacement.astype(np.float64)).sum()])
        got = sums[k].cpu().numpy()
        print(f"step {k}: edges {o.n_edges}", flush=True)
        np.testing.assert_array_equal(got[:2], want[:2])
        np.testing.assert_allclose(got[2:], want[2:], rtol=1e-12)


def test_unaligned_coordinates(orc):
    """Coordinates that start 4 bytes past a 16-byte boundary (a view into a
    larger buffer) take the bounds kernel's 4-byte load path; tree and edges
    must be the oracle's, in refit steps too (the step graph reads the caller's
    array through the address the bounds fold leaves on the device)."""
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(20011, 83)  # not a multiple of four: the partial group
    r = 0.03
    buf = torch.zeros(3 * pts.shape[0] + 1, dtype=torch.float32, device="cuda")
    c = buf[1:].view(pts.shape[0], 3)
    c.copy_(torch.from_numpy(pts).cuda())
    assert c.data_ptr() % 16 == 4
    tns = octohip.TNS()
    octohip.set_search_radius(tns, r)
    octohip.set_refit_mode(tns, True)
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    rng = np.random.default_rng(3)
    cur = pts
    for step in range(3):
        octohip.run(tns)
        t = octohip.get_tree(tns, sid)
        o = orc.run(cur, r)
        assert np.array_equal(_np(t.node_bounds_min), o.bmin) and np.array_equal(_np(t.node_bounds_max), o.bmax)
        assert np.array_equal(np.array(t.origin, np.float32), o.origin)
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        ob = orc.build_edges(o, cur)
        _edges_equal(_gpu_edges(e), (ob.senders, ob.receivers, ob.rel_displacement, ob.rel_dist_norm))
        cur = (cur.astype(np.float64) + rng.normal(0, 1e-4, cur.shape)).astype(np.float32)
        c.copy_(torch.from_numpy(cur).cuda())
        octohip.update_point_set(tns, sid, c)


def test_views_survive_growth_and_their_tns():
    """A result tensor taken before a larger result replaced its buffer keeps
    its values (the old allocation is retired, not freed, until trim), and a
    result outlives the Python TNS it came from."""
    import gc

    import octohip

    pts = datasets.uniform_cube(4000, 61)
    tns, sid, c = _build(pts, 0.05)
    e1 = octohip.build_edges(tns, sid, sid)
    v1 = [_np(x).copy() for x in e1]
    p1 = e1.senders.data_ptr()
    octohip.set_search_radius(tns, 0.12)  # ~14x the edges: every output buffer grows
    octohip.run(tns)
    e2 = octohip.build_edges(tns, sid, sid)
    assert e2.senders.shape[0] > 8 * e1.senders.shape[0]
    assert e2.senders.data_ptr() != p1
    assert octohip.retired_count(tns) >= 4
    for a, b in zip(e1, v1):  # the retired allocation still holds the first result
        assert np.array_equal(bits(_np(a)), bits(b))
    del e1
    assert octohip.trim(tns) >= 4
    assert octohip.retired_count(tns) == 0

    def make():
        t, s, cc = _build(pts, 0.05)
        return octohip.build_edges(t, s, s), cc

    e3, _ = make()
    gc.collect()  # the TNS is only reachable through e3's tensors
    for a, b in zip(e3, v1):
        assert np.array_equal(bits(_np(a)), bits(b))
