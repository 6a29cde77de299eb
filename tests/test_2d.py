"""2D point sets (quadtrees) on the GPU against the 2D oracle: the tree in
the reference's quadtree layout (perm, node ranges, 4 children, 2D bounds)
entry for entry, edges and features bit for bit in both build_edges modes and
both orders, cross pairs, CSR lists, the D = 2 adjoint, refit steps and the
API checks of test/api_2d.jl.  Shapes follow test/edges_cpu_2d.jl:75-125,
test/build_2d.jl:78-110 and test/query_cpu_2d.jl:29-100."""
import numpy as np
import pytest

from brute import bits
from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _np(t):
    return t.detach().cpu().numpy()


def _pts(n, seed, scale=None):
    if scale is not None:  # 0.1 * randn, the reference's clustered clouds
        return (np.float32(scale) * datasets.normal(seed, 0, 2 * n).astype(np.float32)).reshape(n, 2)
    return datasets.uniform_cube(n, seed)[:, :2].copy()


def _tns(pts, r, leaf=32, base=0, refit=False):
    import octohip

    tns = octohip.TNS(ndims=2, index_base=base)
    octohip.set_search_radius(tns, r)
    if leaf != 32:
        octohip.set_leaf_size(tns, leaf)
    if refit:
        octohip.set_refit_mode(tns, True)
    c = _torch().from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid, c


def _sorted(s, rc, d, n):
    o = np.lexsort((s, rc))
    return s[o], rc[o], d.reshape(-1, 2)[o], n.reshape(-1)[o]


def _same_edges(e, o):
    got = _sorted(_np(e.senders), _np(e.receivers), _np(e.rel_displacement), _np(e.rel_dist_norm))
    want = _sorted(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    for x, y in zip(got, want):
        assert np.array_equal(bits(x), bits(y))


TREES = {
    "uniform500": (lambda: _pts(500, 1), 0.05),
    "clustered2000": (lambda: _pts(2000, 2, 0.1), 0.03),
    "single": (lambda: np.zeros((1, 2), np.float32), 1.0),
    "small20": (lambda: _pts(20, 3), 0.3),
    "coincident64": (lambda: np.zeros((64, 2), np.float32), 0.1),
    "uniform50k": (lambda: _pts(50_000, 4), 0.005),
}


@pytest.mark.parametrize("case", sorted(TREES))
@pytest.mark.parametrize("leaf", [32, 8])
def test_quadtree_parity(orc, case, leaf):
    import octohip

    make, r = TREES[case]
    pts = make()
    print(f"2D tree case={case} n={pts.shape[0]} leaf={leaf}", flush=True)
    tns, sid, _ = _tns(pts, r, leaf)
    t = octohip.get_tree(tns, sid)
    o = orc.run2(pts, r, leaf_size=leaf)
    assert t.n_nodes == o.n_nodes
    assert tuple(t.node_children.shape) == (o.n_nodes, 4) and tuple(t.node_bounds_min.shape) == (o.n_nodes, 2)
    assert np.array_equal(_np(t.permutation), o.perm)
    assert np.array_equal(_np(t.node_first), o.node_first)
    assert np.array_equal(_np(t.node_last), o.node_last)
    assert np.array_equal(_np(t.node_children), o.node_children)
    assert np.array_equal(bits(_np(t.node_bounds_min)), bits(o.bmin))
    assert np.array_equal(bits(_np(t.node_bounds_max)), bits(o.bmax))
    assert np.array_equal(np.array(t.origin, np.float32), o.origin)


EDGES = {
    "n200_r015": (lambda: _pts(200, 11), 0.15),
    "n1000_r005": (lambda: _pts(1000, 12), 0.05),
    "clustered2000": (lambda: _pts(2000, 13, 0.1), 0.03),
    "n300_r008": (lambda: _pts(300, 14), 0.08),
    "uniform200k": (lambda: _pts(200_000, 15), 0.004),
}


@pytest.mark.parametrize("case", sorted(EDGES))
@pytest.mark.parametrize("mode", ["exact", "learned"])
def test_edges_2d_parity(orc, case, mode):
    import octohip

    make, r = EDGES[case]
    pts = make()
    print(f"2D edges case={case} n={pts.shape[0]} mode={mode}", flush=True)
    tns, sid, _ = _tns(pts, r)
    o = orc.build_edges2(orc.run2(pts, r), pts)
    for _ in range(2):  # the learned mode's second call is the single pass
        e = octohip.build_edges(tns, sid, sid, mode=mode)
        assert tuple(e.rel_displacement.shape) == (o.n_edges, 2)
        _same_edges(e, o)
    if mode == "exact":
        e = octohip.build_edges(tns, sid, sid, order="receiver")
        assert np.all(np.diff(_np(e.receivers)) >= 0)
        _same_edges(e, o)


def test_cross_pair_csr_and_index_base_2d(orc):
    import octohip

    torch = _torch()
    a, b = _pts(150, 21), _pts(250, 22)
    r = 0.12
    tns = octohip.TNS(ndims=2, index_base=1)
    octohip.set_search_radius(tns, r)
    ia = octohip.add_point_set(tns, torch.from_numpy(a).cuda())
    ib = octohip.add_point_set(tns, torch.from_numpy(b).cuda())
    octohip.set_active_search(tns, ia, ib)
    octohip.run(tns)
    e = octohip.build_edges(tns, ia, ib)
    o = orc.build_edges2(orc.run2(b, r), a, b)
    o.senders += 1
    o.receivers += 1
    _same_edges(e, o)
    octohip.materialize_all_neighbors(tns)
    nb = octohip.neighbors(tns, ia, ib)
    assert np.array_equal(_np(nb.offsets), o.offsets.astype(np.int32))
    flat = _np(nb.flat)
    for i in (0, 75, 149):
        lo, hi = o.offsets[i], o.offsets[i + 1]
        assert np.array_equal(np.sort(flat[lo:hi]), np.sort(o.senders[lo:hi]))


def test_backward_2d(orc):
    import octohip

    pts = _pts(3000, 31)
    r = 0.04
    tns, sid, c = _tns(pts, r)
    x = c.clone().requires_grad_(True)
    e = octohip.build_edges_diff(x, tns, sid, r)
    assert tuple(e.rel_displacement.shape)[1] == 2
    loss = (e.rel_displacement ** 2).sum() + 0.5 * (e.rel_dist_norm ** 2).sum()
    loss.backward()
    d, n = _np(e.rel_displacement), _np(e.rel_dist_norm).reshape(-1)
    want = orc.pullback2(3000, _np(e.senders), _np(e.receivers), d, n, 2 * d, n, r)
    assert tuple(x.grad.shape) == (3000, 2)
    np.testing.assert_allclose(_np(x.grad), want, rtol=1e-4, atol=1e-3)


def test_refit_steps_2d(orc):
    import octohip

    torch = _torch()
    pts = _pts(20_000, 41)
    r = 0.01
    tns, sid, _ = _tns(pts, r, refit=True)
    rng = np.random.default_rng(2)
    cur = pts
    for k in range(3):
        cur = (cur.astype(np.float64) + rng.normal(0, 1e-4, cur.shape)).astype(np.float32)
        octohip.update_point_set(tns, sid, torch.from_numpy(cur).cuda())
        octohip.run(tns)
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        info = octohip.build_info(tns, sid)
        print(f"2D refit step {k}: {info}", flush=True)
        assert info["refit_used"]
        _same_edges(e, orc.build_edges2(orc.run2(cur, r), cur))
    t = octohip.get_tree(tns, sid)
    assert np.array_equal(_np(t.node_children), orc.run2(cur, r).node_children)


def test_api_2d_checks():
    """test/api_2d.jl:11-50: 3-column coords rejected by a 2D TNS and 2-column
    by a 3D one; update keeps the size."""
    import octohip

    torch = _torch()
    t2 = octohip.TNS(ndims=2)
    octohip.set_search_radius(t2, 0.1)
    with pytest.raises(octohip.DimensionMismatch):
        octohip.add_point_set(t2, torch.zeros(10, 3, device="cuda"))
    t3 = octohip.TNS()
    with pytest.raises(octohip.DimensionMismatch):
        octohip.add_point_set(t3, torch.zeros(10, 2, device="cuda"))
    sid = octohip.add_point_set(t2, torch.zeros(10, 2, device="cuda"))
    with pytest.raises(octohip.DimensionMismatch):
        octohip.update_point_set(t2, sid, torch.zeros(11, 2, device="cuda"))
    with pytest.raises((octohip.ArgumentError, octohip.DimensionMismatch)):  # batches are 3D
        octohip.Batch(t2, torch.zeros(10, 2, device="cuda"), [0, 10], [0.1])


def test_device_view_2d():
    """device_view of a 2D set (test/query_cuda_2d.jl pattern): a user kernel
    iterating the view counts, per query, the points within r in the 2D
    predicate (dx*dx + dy*dy <= r*r in FP32, self included); the view's
    coordinates are the stored (x, y, constant z) copy, whose z differences
    are exactly 0."""
    import octohip

    pts = _pts(4000, 51)
    r = 0.03
    tns, sid, _ = _tns(pts, r)
    v = octohip.device_view(tns, sid, sid)
    cnt = _np(octohip.count_neighbors_device(tns, v))
    p = np.ascontiguousarray(pts, np.float32)
    r2 = np.float32(r) * np.float32(r)
    want = np.empty(p.shape[0], np.int64)
    for i in range(p.shape[0]):
        dx = p[:, 0] - p[i, 0]
        dy = p[:, 1] - p[i, 1]
        want[i] = np.count_nonzero(dx * dx + dy * dy <= r2)
    assert np.array_equal(cnt, want)


def test_jittered_lattice_2d(orc):
    """The configuration of test/edges_cuda.jl:147-200 (D = 2, N = 40,000,
    r = 0.072, leaf 32: a jittered lattice at spacing r / 1.6, dense enough
    that the reference's first node estimate overflowed): tree and edges
    against the oracle."""
    import octohip

    r = 0.072
    s = np.float32(r / 1.6)
    g = np.stack(np.meshgrid(np.arange(200), np.arange(200), indexing="ij"), -1).reshape(-1, 2)
    jit = datasets.uniform(77, 0, 2 * g.shape[0]).astype(np.float32).reshape(-1, 2) - np.float32(0.5)
    pts = ((g.astype(np.float32) + np.float32(0.25) * jit) * s).astype(np.float32)
    tns, sid, _ = _tns(pts, r)
    o = orc.run2(pts, r)
    t = octohip.get_tree(tns, sid)
    assert t.n_nodes == o.n_nodes
    assert np.array_equal(_np(t.node_children), o.node_children)
    assert np.array_equal(_np(t.permutation), o.perm)
    _same_edges(octohip.build_edges(tns, sid, sid), orc.build_edges2(o, pts))
