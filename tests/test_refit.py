"""Almost-sorted refit (refit mode): across time steps the tree, permutation
and edges must be exactly those of a full rebuild and of the CPU oracle (the
reference's test/refit.jl:63-86 pattern: perm validity + edge-set equality
after coordinate updates).  The fix-up path is checked to have run, and the
fallback to the full sort when too many points change cells."""
import numpy as np
import pytest

from brute import bits, sort_edges
from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _np(t):
    return t.detach().cpu().numpy()


def _pair(pts, r, refit):
    import octohip

    tns = octohip.TNS()
    octohip.set_search_radius(tns, float(r))
    octohip.set_refit_mode(tns, refit)
    c = _torch().from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid, c


def _tree_equal(a, b):
    for f in ("permutation", "node_first", "node_last", "node_children", "node_bounds_min", "node_bounds_max"):
        assert np.array_equal(_np(getattr(a, f)), _np(getattr(b, f))), f


def _steps(pts, r, n_steps, sigma, seed, check_oracle=None):
    import octohip

    torch = _torch()
    t_on, s_on, _ = _pair(pts, r, True)
    t_off, s_off, _ = _pair(pts, r, False)
    rng = np.random.default_rng(seed)
    cur = pts.astype(np.float32)
    infos = []
    for k in range(n_steps):
        cur = (cur.astype(np.float64) + rng.normal(0, sigma, cur.shape)).astype(np.float32)
        c = torch.from_numpy(cur).cuda()
        octohip.update_point_set(t_on, s_on, c)
        octohip.update_point_set(t_off, s_off, c)
        octohip.run(t_on)
        octohip.run(t_off)
        info = octohip.build_info(t_on, s_on)
        infos.append(info)
        print(f"step {k}: {info}", flush=True)
        _tree_equal(octohip.get_tree(t_on, s_on), octohip.get_tree(t_off, s_off))
        e_on = octohip.build_edges(t_on, s_on, s_on)
        e_off = octohip.build_edges(t_off, s_off, s_off)
        a = sort_edges(_np(e_on.senders), _np(e_on.receivers), _np(e_on.rel_displacement),
                       _np(e_on.rel_dist_norm).reshape(-1))
        b = sort_edges(_np(e_off.senders), _np(e_off.receivers), _np(e_off.rel_displacement),
                       _np(e_off.rel_dist_norm).reshape(-1))
        for x, y in zip(a, b):
            assert np.array_equal(bits(x), bits(y))
        if check_oracle is not None and k == n_steps - 1:
            o = check_oracle.run(cur, float(r))
            assert np.array_equal(_np(octohip.get_tree(t_on, s_on).permutation), o.perm)
    return infos


def test_refit_sph_steps_match_full_rebuild(orc):
    pts, r = datasets.sph_lattice(seed=7, side=40)
    infos = _steps(pts, r, 5, 1e-4, 1, check_oracle=orc)  # 0.01h per step, as in the SPH workload
    assert all(i["refit_used"] for i in infos)
    assert all(0 < i["changed_points"] < pts.shape[0] // 8 for i in infos)


def test_refit_many_changed_points(orc):
    # more changed points than the rank kernel's grid holds at 16 lanes per
    # point (64,000 points: 8,192 threads, so more than 512 changed), still
    # within the refit limit of N / 8
    pts, r = datasets.sph_lattice(seed=7, side=40)
    infos = _steps(pts, r, 1, 2e-3, 5, check_oracle=orc)
    assert all(i["refit_used"] for i in infos)
    assert all(512 < i["changed_points"] < pts.shape[0] // 8 for i in infos)


def test_refit_ties_and_coincident_points(orc):
    # many exactly coincident points (equal keys, order by index) with movers
    base = np.repeat(datasets.uniform_cube(200, 3), 20, axis=0)
    infos = _steps(base, 0.05, 4, 2e-3, 2, check_oracle=orc)
    assert any(i["refit_used"] for i in infos)


def test_refit_falls_back_when_many_points_move(orc):
    pts = datasets.uniform_cube(20000, 4)
    infos = _steps(pts, 0.02, 2, 0.05, 3, check_oracle=orc)
    assert not any(i["refit_used"] for i in infos)
    assert all(i["changed_points"] > 20000 // 8 for i in infos)


def test_refit_off_reports_no_previous_sort():
    import octohip

    tns, sid, _ = _pair(datasets.uniform_cube(1000, 5), 0.05, False)
    assert octohip.build_info(tns, sid) == {"refit_used": False, "changed_points": -1}


@pytest.mark.parametrize("mode", ["exact", "learned"])
def test_refit_failure_found_at_build_edges(orc, mode):
    """run! refits without waiting; the device finds too many changed points
    and keeps a valid tree of the previous order, and build_edges (the first
    read-back) rebuilds with the full sort and redoes the edges."""
    import octohip

    pts = datasets.uniform_cube(20000, 14)
    r = 0.03
    tns, sid, _ = _pair(pts, r, True)
    octohip.build_edges(tns, sid, sid, mode=mode)
    rng = np.random.default_rng(5)
    moved = (pts.astype(np.float64) + rng.normal(0, 0.05, pts.shape)).astype(np.float32)
    octohip.update_point_set(tns, sid, _torch().from_numpy(moved).cuda())
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid, mode=mode)
    info = octohip.build_info(tns, sid)
    print(f"{mode}: {info}", flush=True)
    assert not info["refit_used"] and info["changed_points"] > 20000 // 8
    ot = orc.run(moved, r)
    o = orc.build_edges(ot, moved)
    got = sort_edges(_np(e.senders), _np(e.receivers), _np(e.rel_displacement), _np(e.rel_dist_norm).reshape(-1))
    for x, y in zip(got, sort_edges(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)):
        assert np.array_equal(bits(x), bits(y))
    assert np.array_equal(_np(octohip.get_tree(tns, sid).permutation), ot.perm)
