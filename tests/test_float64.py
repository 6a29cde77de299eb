"""Float64 point sets on the GPU (TNS(Float64), test/float64.jl) against the
Float64 oracle: the tree (perm, node ranges, children, double bounds) entry
for entry, edges and features bit for bit in 3D and 2D (the reference
tests' shapes plus larger sets), cross pairs with 1-based indices, CSR
lists, the double adjoint, and the modes a Float64 context refuses."""
import numpy as np
import pytest

import oracle.f64 as f64

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _np(t):
    return t.detach().cpu().numpy()


def _tns(pts, r, leaf=32, base=0):
    import octohip

    torch = _torch()
    tns = octohip.TNS(dtype=torch.float64, ndims=pts.shape[1], index_base=base)
    octohip.set_search_radius(tns, r)
    if leaf != 32:
        octohip.set_leaf_size(tns, leaf)
    c = torch.from_numpy(np.ascontiguousarray(pts, np.float64)).cuda()
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid, c


def _same(e, o, base=0):
    s, rc = _np(e.senders) - base, _np(e.receivers) - base
    w = np.lexsort((s, rc))
    assert np.array_equal(s[w], o.senders) and np.array_equal(rc[w], o.receivers)
    d = _np(e.rel_displacement)[w]
    n = _np(e.rel_dist_norm).reshape(-1)[w]
    assert d.dtype == np.float64 and n.dtype == np.float64
    assert np.array_equal(d.view(np.int64), o.rel_displacement.view(np.int64))
    assert np.array_equal(n.view(np.int64), o.rel_dist_norm.view(np.int64))


CASES = {
    "3d_n400_r01": (3, 400, 0.1, 4001),      # test/float64.jl:33-50
    "3d_n300_r008": (3, 300, 0.08, 4002),    # test/float64.jl:52-75
    "2d_n300_r01": (2, 300, 0.1, 4003),      # test/float64.jl:77-95
    "3d_n30000_r003": (3, 30000, 0.03, 4004),
    "2d_n30000_r001": (2, 30000, 0.01, 4005),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("leaf", [32, 8])
def test_float64_tree_and_edges(case, leaf):
    import octohip

    dims, n, r, seed = CASES[case]
    pts = np.random.default_rng(seed).random((n, dims))
    print(f"float64 case={case} leaf={leaf}", flush=True)
    tns, sid, _ = _tns(pts, r, leaf)
    t = octohip.get_tree(tns, sid)
    o = f64.run(pts, r, leaf_size=leaf)
    assert t.n_nodes == o.n_nodes
    assert np.array_equal(_np(t.permutation), o.perm)
    assert np.array_equal(_np(t.node_first), o.node_first) and np.array_equal(_np(t.node_last), o.node_last)
    assert np.array_equal(_np(t.node_children), o.node_children)
    assert t.node_bounds_min.dtype == _torch().float64
    assert np.array_equal(_np(t.node_bounds_min).view(np.int64), o.bmin.view(np.int64))
    assert np.array_equal(_np(t.node_bounds_max).view(np.int64), o.bmax.view(np.int64))
    e = octohip.build_edges(tns, sid, sid)
    assert np.all(np.diff(_np(e.receivers)) >= 0)  # receiver order, the reference's layout
    _same(e, f64.build_edges(pts, pts, r, True))


def test_float64_cross_pair_csr_index_base():
    import octohip

    torch = _torch()
    g = np.random.default_rng(7)
    a, b = g.random((500, 3)), g.random((700, 3))
    r = 0.09
    tns = octohip.TNS(dtype=torch.float64, index_base=1)
    octohip.set_search_radius(tns, r)
    ia = octohip.add_point_set(tns, torch.from_numpy(a).cuda())
    ib = octohip.add_point_set(tns, torch.from_numpy(b).cuda())
    octohip.set_active_search(tns, ia, ib)
    octohip.run(tns)
    e = octohip.build_edges(tns, ia, ib)
    o = f64.build_edges(a, b, r, False)
    _same(e, o, base=1)
    octohip.materialize_all_neighbors(tns)
    nb = octohip.neighbors(tns, ia, ib)
    assert np.array_equal(_np(nb.offsets), o.offsets.astype(np.int32))
    flat = _np(nb.flat) - 1
    for i in (0, 250, 499):
        assert np.array_equal(np.sort(flat[o.offsets[i]:o.offsets[i + 1]]), o.senders[o.offsets[i]:o.offsets[i + 1]])


@pytest.mark.parametrize("dims", [3, 2])
def test_float64_backward(dims):
    import octohip

    pts = np.random.default_rng(11 + dims).random((5000, dims))
    r = 0.06 if dims == 3 else 0.02
    tns, sid, c = _tns(pts, r)
    x = c.clone().requires_grad_(True)
    e = octohip.build_edges_diff(x, tns, sid, r)
    loss = (e.rel_displacement ** 2).sum() + 0.5 * (e.rel_dist_norm ** 2).sum()
    loss.backward()
    d, n = _np(e.rel_displacement), _np(e.rel_dist_norm).reshape(-1)
    want = f64.pullback(5000, _np(e.senders), _np(e.receivers), d, n, 2 * d, n, r)
    assert x.grad.dtype == _torch().float64
    np.testing.assert_allclose(_np(x.grad), want, rtol=1e-12, atol=1e-9)


def test_float64_refuses_float32_only_modes():
    import octohip

    pts = np.random.default_rng(3).random((100, 3))
    tns, sid, _ = _tns(pts, 0.2)
    with pytest.raises(octohip.ArgumentError):
        octohip.build_edges(tns, sid, sid, mode="learned")
    with pytest.raises(octohip.OctopusError):
        octohip.device_view(tns, sid, sid)


def test_float64_apply_zsort_and_cross_rejection():
    """test/float64.jl:115-148: apply_zsort of Float64 arrays (1-D and 2-D)
    follows the set's permutation, and a Float64 TNS takes no Float32
    coordinates nor a Float32 TNS Float64 ones."""
    import octohip

    torch = _torch()
    pts = np.random.default_rng(4005).random((200, 3))
    tns, sid, _ = _tns(pts, 0.1)
    perm = _np(octohip.get_tree(tns, sid).permutation)
    masses = torch.rand(200, dtype=torch.float64, device="cuda")
    out = octohip.apply_zsort(tns, sid, masses)
    assert out.dtype == torch.float64
    assert np.array_equal(_np(out), _np(masses)[perm])
    feats = torch.rand(200, 4, dtype=torch.float64, device="cuda")
    assert np.array_equal(_np(octohip.apply_zsort(tns, sid, feats)), _np(feats)[perm])
    with pytest.raises(octohip.ArgumentError):
        octohip.add_point_set(tns, torch.rand(10, 3, device="cuda"))
    t32 = octohip.TNS()
    octohip.set_search_radius(t32, 0.1)
    with pytest.raises(octohip.ArgumentError):
        octohip.add_point_set(t32, torch.rand(10, 3, dtype=torch.float64, device="cuda"))
