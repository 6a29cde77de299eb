"""The reference's API contracts on the HIP path: test/contracts.jl,
test/api.jl, test/backend_errors_cpu.jl and the GPU-relevant cases of
test/edge_cases.jl, one test per reference testset (cited), with the engine's
edges checked against the oracle where a testset checks results."""
import numpy as np
import pytest

from brute import bits
from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _gpu(pts):
    return _torch().from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()


def _np(t):
    return t.detach().cpu().numpy()


def _tns(pts, r, ndims=3):
    import octohip

    tns = octohip.TNS(ndims=ndims)
    octohip.set_search_radius(tns, r)
    sid = octohip.add_point_set(tns, _gpu(pts))
    octohip.set_active_search(tns, sid, sid)
    return tns, sid


def _sorted(s, r, d, n):
    o = np.lexsort((np.asarray(s, np.int64), np.asarray(r, np.int64)))
    d = np.asarray(d)
    d = d.reshape(len(s), d.size // max(len(s), 1))  # (E, D) whatever the layout
    return s[o], r[o], d[o], np.asarray(n).reshape(-1)[o]


def _same_set(e, o):
    g = _sorted(_np(e.senders), _np(e.receivers), _np(e.rel_displacement), _np(e.rel_dist_norm))
    w = _sorted(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    for a, b in zip(g, w):
        assert np.array_equal(bits(a), bits(b))


def test_set_active_search_idempotent_and_directed():
    """contracts.jl:34-83: registering a pair twice keeps one entry, the self
    pair is one entry, (a, b) and (b, a) are separate directed pairs."""
    import octohip

    tns = octohip.TNS()
    octohip.set_search_radius(tns, 0.1)
    a = octohip.add_point_set(tns, _gpu(datasets.uniform_cube(100, 1)))
    b = octohip.add_point_set(tns, _gpu(datasets.uniform_cube(120, 2)))
    octohip.set_active_search(tns, a, a)
    octohip.set_active_search(tns, a, a)
    assert tns.active_pairs == [(a, a)]
    octohip.set_active_search(tns, a, b)
    assert (a, b) in tns.active_pairs and (b, a) not in tns.active_pairs
    octohip.run(tns)
    with pytest.raises(octohip.OctopusError):
        octohip.build_edges(tns, b, a)
    octohip.set_active_search(tns, b, a)
    assert len(tns.active_pairs) == 3
    octohip.run(tns)
    ab = octohip.build_edges(tns, a, b)
    ba = octohip.build_edges(tns, b, a)
    # contracts.jl:303-320: the direction matters; the edge sets are each other's transpose
    assert _np(ab.receivers).max() < 100 and _np(ba.receivers).max() < 120
    s1 = set(zip(_np(ab.receivers).tolist(), _np(ab.senders).tolist()))
    s2 = set(zip(_np(ba.senders).tolist(), _np(ba.receivers).tolist()))
    assert s1 == s2


def test_run_idempotent_on_unchanged_state(orc):
    """contracts.jl:85-101: a second run! with nothing changed keeps the tree."""
    import octohip

    pts = datasets.uniform_cube(200, 3)
    tns, sid = _tns(pts, 0.1)
    octohip.run(tns)
    t1 = octohip.get_tree(tns, sid)
    perm1, n1 = _np(t1.permutation).copy(), t1.n_nodes
    assert not any(tns.dirty)
    octohip.run(tns)
    t2 = octohip.get_tree(tns, sid)
    assert np.array_equal(_np(t2.permutation), perm1) and t2.n_nodes == n1
    assert np.array_equal(_np(t2.permutation), orc.run(pts, 0.1).perm)


def test_radius_change_rebuilds(orc):
    """contracts.jl:103-125: set_search_radius! after run! marks every set
    dirty, and the next run! and build_edges use the new radius."""
    import octohip

    pts = datasets.uniform_cube(200, 4)
    tns, sid = _tns(pts, 0.1)
    octohip.run(tns)
    octohip.set_search_radius(tns, 0.05)
    assert all(tns.dirty)
    octohip.run(tns)
    assert not any(tns.dirty)
    _same_set(octohip.build_edges(tns, sid, sid), orc.build_edges(orc.run(pts, 0.05), pts))
    n_small = len(octohip.get_neighborlist(tns, sid, sid, 0))
    octohip.set_search_radius(tns, 0.2)
    octohip.run(tns)
    assert len(octohip.get_neighborlist(tns, sid, sid, 0)) >= n_small
    _same_set(octohip.build_edges(tns, sid, sid), orc.build_edges(orc.run(pts, 0.2), pts))


def test_refit_mode_marks_dirty():
    """contracts.jl:127-143."""
    import octohip

    tns, sid = _tns(datasets.uniform_cube(100, 5), 0.1)
    octohip.run(tns)
    assert not tns.dirty[sid]
    octohip.set_refit_mode(tns, True)
    assert tns.dirty[sid] and tns.refit_mode
    octohip.run(tns)
    octohip.set_refit_mode(tns, False)
    assert tns.dirty[sid] and not tns.refit_mode


def test_resize_and_update(orc):
    """contracts.jl:145-169, api.jl:35-42: resize_point_set! changes N (to 0
    too), update_point_set! keeps it; the pairs and results follow."""
    import octohip

    tns, sid = _tns(datasets.uniform_cube(100, 6), 0.1)
    octohip.run(tns)
    pts2 = datasets.uniform_cube(75, 7)
    octohip.resize_point_set(tns, sid, _gpu(pts2))
    assert tns.dirty[sid]
    octohip.run(tns)
    _same_set(octohip.build_edges(tns, sid, sid), orc.build_edges(orc.run(pts2, 0.1), pts2))
    with pytest.raises(octohip.DimensionMismatch):
        octohip.update_point_set(tns, sid, _gpu(datasets.uniform_cube(74, 8)))
    pts3 = datasets.uniform_cube(75, 9)
    octohip.update_point_set(tns, sid, _gpu(pts3))
    assert tns.dirty[sid] and tns.active_pairs == [(sid, sid)]
    octohip.run(tns)
    _same_set(octohip.build_edges(tns, sid, sid), orc.build_edges(orc.run(pts3, 0.1), pts3))
    octohip.resize_point_set(tns, sid, _gpu(np.zeros((0, 3), np.float32)))
    octohip.run(tns)
    assert octohip.build_edges(tns, sid, sid).senders.shape[0] == 0


def test_build_edges_aliases_the_pair_buffer():
    """contracts.jl:223-240: build_edges hands out the pair's buffers, not a
    fresh allocation per call: a second call on unchanged state writes the
    same storage."""
    import octohip

    tns, sid = _tns(datasets.uniform_cube(150, 10), 0.1)
    octohip.run(tns)
    e1 = octohip.build_edges(tns, sid, sid)
    e2 = octohip.build_edges(tns, sid, sid)
    for a, b in zip(e1, e2):
        assert a.data_ptr() == b.data_ptr()


def test_unregistered_pair_errors():
    """contracts.jl:242-250: build_edges! of a pair never registered."""
    import octohip

    tns = octohip.TNS()
    octohip.set_search_radius(tns, 0.1)
    sid = octohip.add_point_set(tns, _gpu(datasets.uniform_cube(50, 11)))
    octohip.run(tns)
    with pytest.raises(octohip.OctopusError):
        octohip.build_edges(tns, sid, sid)


def test_two_instances_independent(orc):
    """contracts.jl:273-294."""
    import octohip

    a_pts, b_pts = datasets.uniform_cube(80, 12), datasets.uniform_cube(120, 13)
    t1, a = _tns(a_pts, 0.1)
    t2, b = _tns(b_pts, 0.2)
    octohip.run(t1)
    octohip.run(t2)
    assert t1.radius == np.float32(0.1) and t2.radius == np.float32(0.2)
    _same_set(octohip.build_edges(t1, a, a), orc.build_edges(orc.run(a_pts, 0.1), a_pts))
    _same_set(octohip.build_edges(t2, b, b), orc.build_edges(orc.run(b_pts, 0.2), b_pts))


def test_tns_parameters_and_defaults():
    """contracts.jl:296-301, api.jl:11-16: element type and dimension, the
    default leaf size 32, no radius until one is set."""
    import octohip

    torch = _torch()
    t3, t2 = octohip.TNS(ndims=3), octohip.TNS(ndims=2)
    assert (t3.dtype, t3.ndims) == (torch.float32, 3) and (t2.dtype, t2.ndims) == (torch.float32, 2)
    assert octohip.TNS(dtype=torch.float64).dtype == torch.float64
    assert t3.target_leaf_size == 32 and not t3.radius > 0 and t3.point_sets == [] and t3.active_pairs == []


def test_api_errors():
    """api.jl:18-33, backend_errors_cpu.jl:25-38,80-97: radius > 0, run!
    needs a radius and a set, row counts per dimension."""
    import octohip

    torch = _torch()
    tns = octohip.TNS()
    for r in (0.0, -0.1):
        with pytest.raises(octohip.ArgumentError):
            octohip.set_search_radius(tns, r)
    octohip.set_search_radius(tns, 0.1)
    with pytest.raises(octohip.ArgumentError):
        octohip.run(tns)  # no point set
    t = octohip.TNS()
    sid = octohip.add_point_set(t, torch.rand(50, 3, device="cuda"))  # allowed without a radius
    octohip.set_active_search(t, sid, sid)
    with pytest.raises(octohip.ArgumentError):
        octohip.run(t)  # no radius
    t3, t2 = octohip.TNS(ndims=3), octohip.TNS(ndims=2)
    for tt, bad in ((t3, 2), (t3, 4), (t2, 3), (t2, 1)):
        octohip.set_search_radius(tt, 0.1)
        with pytest.raises(octohip.DimensionMismatch):
            octohip.add_point_set(tt, torch.rand(10, bad, device="cuda"))


def test_build_edges_before_run_is_empty(orc):
    """backend_errors_cpu.jl:40-57: before the first run! the tree is empty
    and build_edges returns no edges (no error); after run! it is the
    oracle's set."""
    import octohip

    pts = datasets.uniform_cube(100, 14)
    tns, sid = _tns(pts, 0.1)
    assert octohip.build_edges(tns, sid, sid).senders.shape[0] == 0
    octohip.run(tns)
    _same_set(octohip.build_edges(tns, sid, sid), orc.build_edges(orc.run(pts, 0.1), pts))


def test_zsort_errors():
    """backend_errors_cpu.jl:59-78: prepare_zsort before run!, apply_zsort
    of the wrong length."""
    import octohip

    torch = _torch()
    tns, sid = _tns(datasets.uniform_cube(50, 15), 0.1)
    with pytest.raises(octohip.OctopusError):
        octohip.prepare_zsort(tns)
    octohip.run(tns)
    octohip.prepare_zsort(tns)
    with pytest.raises((octohip.ArgumentError, octohip.DimensionMismatch)):
        octohip.apply_zsort(tns, sid, torch.rand(49, device="cuda"))


@pytest.mark.parametrize("ndims", [3, 2])
def test_exact_distance_boundary(ndims):
    """edge_cases.jl:68-101: a point at exactly r is a neighbour, one ulp past
    r is not (the FP32 predicate d2 <= r*r)."""
    import octohip

    r = np.float32(0.25)
    p = np.zeros((3, ndims), np.float32)
    p[1, 0] = r  # at exactly r
    p[2, 0] = np.nextafter(np.float32(2 * r), np.float32(1))  # from point 1: just past r
    tns, sid = _tns(p, float(r), ndims=ndims)
    octohip.run(tns)
    for mode in ("exact", "learned"):
        e = octohip.build_edges(tns, sid, sid, mode=mode)
        got = set(zip(_np(e.receivers).tolist(), _np(e.senders).tolist()))
        assert (0, 1) in got and (1, 0) in got
        assert (1, 2) not in got and (2, 1) not in got
        d2 = np.float32(p[2, 0] - p[1, 0]) ** 2
        assert d2 > r * r  # the pair really is past r in FP32


@pytest.mark.parametrize("ndims", [3, 2])
def test_coincident_negative_tiny_huge(orc, ndims):
    """edge_cases.jl:103-195: all points at one location (everyone is everyone's
    neighbour, tree invariants hold), negative coordinates, a radius below
    every distance, a radius above every distance."""
    import octohip

    n = 40
    same = np.full((n, ndims), -3.5, np.float32)
    tns, sid = _tns(same, 0.01, ndims=ndims)
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid)
    assert e.senders.shape[0] == n * (n - 1)
    t = octohip.get_tree(tns, sid)
    assert np.array_equal(np.sort(_np(t.permutation)), np.arange(n))
    neg = (datasets.uniform_cube(300, 16)[:, :ndims] - np.float32(2.0)).copy()
    build = orc.run if ndims == 3 else orc.run2
    edges = orc.build_edges if ndims == 3 else orc.build_edges2
    for r, expect in ((0.12, None), (1e-6, 0), (10.0, 300 * 299)):
        tns, sid = _tns(neg, r, ndims=ndims)
        octohip.run(tns)
        e = octohip.build_edges(tns, sid, sid)
        if expect is not None:
            assert e.senders.shape[0] == expect
        _same_set(e, edges(build(neg, r), neg))
