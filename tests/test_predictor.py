"""Per-leaf hit-count predictor: the NumPy oracle against finite differences
(CPU), the HIP features / forward / gradient against the oracle, training,
and what the trained network gains over the method without it (GPU).

Bars: features within 2 ulp-ish of the oracle (the device log1pf/log2f may
differ from NumPy by one ulp; adjacency counts are exact integers inside);
forward within 1e-4 relative of the float64 oracle; gradient within 1e-3
relative (float32 sums over the batch) of the float64 oracle.  The edges the
predictor plans stay bit-exact against the CPU oracle.
"""
import math

import numpy as np
import pytest

from brute import bits, brute_edges, sort_edges
from octohip import datasets

import oracle.mlp as om


def _rand_params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(om.PARAMS)
    P[om.MU:om.SD] = rng.normal(0, 0.5, om.IN)
    P[om.SD:om.W1] = rng.uniform(0.5, 2.0, om.IN)
    P[om.W1:om.B1] = rng.normal(0, scale * math.sqrt(2 / om.IN), om.IN * om.H)
    P[om.B1:om.W2] = rng.normal(0, 0.1, om.H)
    P[om.W2:om.B2] = rng.normal(0, scale * math.sqrt(2 / om.H), om.H * om.H)
    P[om.B2:om.W3] = rng.normal(0, 0.1, om.H)
    P[om.W3:om.B3] = rng.normal(0, math.sqrt(1 / om.H), om.H)
    P[om.B3] = 0.3
    return P


# ------------------------------------------------------------------ CPU: oracle pinned

def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    P = _rand_params(2)
    X = rng.normal(0, 1, (9, om.IN))
    T = rng.normal(1, 1, 9)
    _, g = om.loss_and_grad(P, X, T)
    idx = np.concatenate([rng.choice(om.TRAINED, 150, replace=False), np.arange(om.TRAINED - om.H - 1, om.TRAINED)])
    h = 1e-6
    for k in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[om.W1 + k] += h
        Pm[om.W1 + k] -= h
        lp = np.mean((om.forward(Pp, X) - T) ** 2)
        lm = np.mean((om.forward(Pm, X) - T) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-5 * abs(fd), (k, fd, g[k])


def test_oracle_forward_special_cases():
    # zero weights: y = b3 everywhere, and only b3 (and W3 through relu(b2)) gets gradient
    P = np.zeros(om.PARAMS)
    P[om.SD:om.W1] = 1.0
    P[om.B3] = 0.7
    X = np.random.default_rng(3).normal(0, 1, (5, om.IN))
    assert np.allclose(om.forward(P, X), 0.7)
    T = np.arange(5.0)
    loss, g = om.loss_and_grad(P, X, T)
    assert math.isclose(loss, float(np.mean((0.7 - T) ** 2)))
    assert math.isclose(g[-1], 2 * float(np.mean(0.7 - T)))
    assert np.all(g[:-1] == 0)


def test_dir_codes_27_placements():
    """Every one of the 27 placements of a unit box around the unit query box
    gets its own code, x weighing 1, y 3 and z 9, below = 0 / level = 1 /
    above = 2 per axis; a centre offset of exactly a quarter of the two
    extents is still level, one step more is not."""
    f = np.float32
    qlo, qhi = np.zeros(3, f), np.ones(3, f)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                off = np.array([2 * dx, 2 * dy, 2 * dz], f)
                code = int(om.dir_codes(qlo, qhi, qlo + off, qhi + off))
                assert code == (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1), (dx, dy, dz, code)
    # threshold: tau = 0.25 * (1 + 1) = 0.5 on each axis
    for shift, want in ((0.5, 1), (0.5625, 2), (-0.5, 1), (-0.5625, 0)):
        clo = qlo + np.array([shift, 0, 0], f)
        assert int(om.dir_codes(qlo, qhi, clo, clo + np.ones(3, f))) == want + 3 + 9, shift
        clo = qlo + np.array([0, 0, shift], f)
        assert int(om.dir_codes(qlo, qhi, clo, clo + np.ones(3, f))) == 1 + 3 + 9 * want, shift


class _FakeTree:
    """A hand-made tree: root 0 with four leaves (the attributes
    leaf_features reads)."""

    def __init__(self):
        self.node_children = np.full((5, 8), -1, np.int32)
        self.node_children[0, :4] = [1, 2, 3, 4]
        self.node_first = np.array([0, 0, 2, 5, 6], np.int32)
        self.node_last = np.array([9, 1, 4, 5, 9], np.int32)
        self.node_level = np.array([0, 1, 2, 3, 2], np.int32)
        self.bmin = np.array([[-1.5, -1.5, 0], [0, 0, 0], [2, 0, 0], [0, 0, 3], [-1.5, -1.5, 0]], np.float32)
        self.bmax = np.array([[2.5, 1, 3], [1, 1, 1], [2.5, 0.5, 0.5], [0, 0, 3], [-0.5, -0.5, 1]], np.float32)


def test_leaf_features_hand_worked():
    """The 29 features of each leaf of a hand-made four-leaf tree (r = 1),
    worked out by hand from the definitions: [0] log1p(points), [1] level / 21,
    [2] max(log2(max(extent / r, 1e-30)), -20) with extent the largest box
    side, [3 + k] log1p(points of the candidate leaves in direction k, the 27
    codes without 13 = the leaf's own cell).  Candidates are leaves whose box
    gap is <= r (L1-L2 gap exactly 1: included).
      L1 box [0,1]^3, 2 points, level 1: sees L2 at code 14 (slot 13, 3 points)
         and L4 at code 9 (slot 9, 4 points); L3 is 2 away in z.
      L2 box [2,2.5]x[0,.5]^2, 3 points, level 2, extent .5: sees L1 at code 12
         (slot 12, 2 points).
      L3 a single point (0,0,3), level 3, extent 0: sees nothing.
      L4 box [-1.5,-.5]^2 x [0,1], 4 points, level 2: sees L1 at code 17
         (slot 16, 2 points)."""
    want = np.zeros((4, 29), np.float64)
    want[0, :3] = [math.log(3.0), 1 / 21, 0.0]
    want[0, 3 + 9] = math.log(5.0)
    want[0, 3 + 13] = math.log(4.0)
    want[1, :3] = [math.log(4.0), 2 / 21, -1.0]
    want[1, 3 + 12] = math.log(3.0)
    want[2, :3] = [math.log(2.0), 3 / 21, -20.0]
    want[3, :3] = [math.log(5.0), 2 / 21, 0.0]
    want[3, 3 + 16] = math.log(3.0)
    t = _FakeTree()
    got = om.leaf_features(t, t, 1.0)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=0)


def test_oracle_candidates_cover_every_pair(orc):
    # every pair within r lies in a candidate leaf of the query's leaf, and the
    # direction code of a leaf seen from itself is 13
    pts = (np.float32(0.1) * datasets.normal(41, 0, 3000).astype(np.float32)).reshape(1000, 3)
    r = 0.04
    t = orc.run(pts, r)
    leaves = om.leaves_of(t)
    leaf_of = np.empty(pts.shape[0], np.int64)
    for k, l in enumerate(leaves):
        leaf_of[t.perm[t.node_first[l]:t.node_last[l] + 1]] = k
    lo, hi = t.bmin[leaves].astype(np.float32), t.bmax[leaves].astype(np.float32)
    r_sq = np.float32(np.float32(r) * np.float32(r))
    s, rc, _, _ = brute_edges(pts, pts, r, exclude_self=True)
    for i, j in zip(rc[:5000], s[:5000]):
        a, b = leaf_of[i], leaf_of[j]
        assert om.box_gap_sq(lo[a], hi[a], lo[b], hi[b]) <= r_sq
    assert np.all(om.dir_codes(lo, hi, lo, hi) == 13)
    f = om.leaf_features(t, t, r)
    assert f.shape == (leaves.shape[0], 29)
    assert np.all(np.isfinite(f))


# ------------------------------------------------------------------ GPU

def _torch():
    import torch

    return torch


def _tns(pts, r, leaf=32):
    import octohip

    torch = _torch()
    tns = octohip.TNS()
    octohip.set_search_radius(tns, float(r))
    if leaf != 32:
        octohip.set_leaf_size(tns, leaf)
    c = torch.from_numpy(np.ascontiguousarray(pts, np.float32)).cuda()
    sid = octohip.add_point_set(tns, c)
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid, c


def _clustered(n, seed):
    pts, r = datasets.clustered(n=n, seed=seed)
    # keep about the 4M-point neighbour count at this size
    return pts, np.float32(float(r) * (4_000_000 / n) ** (1 / 3))


@pytest.mark.gpu
def test_features_match_oracle(orc):
    import octohip

    pts, r = _clustered(3000, 7)
    tns, sid, _ = _tns(pts, r)
    octohip.predictor.collect(tns, sid, sid)
    st = octohip.predictor.edge_stats(tns, sid, sid)
    got = st["features"].cpu().numpy()
    want = om.leaf_features(orc.run(pts, float(r)), orc.run(pts, float(r)), float(r))
    print(f"features n={pts.shape[0]} r={r:.4f} leaves={got.shape[0]}", flush=True)
    assert got.shape == want.shape
    np.testing.assert_allclose(got, want, rtol=1e-6, atol=1e-6)
    # exact per-leaf hits sum to the edge count
    assert int(st["exact"].sum().item()) == st["n_edges"]


@pytest.mark.gpu
def test_forward_and_gradient_match_oracle():
    import octohip

    torch = _torch()
    tns, _, _ = _tns(datasets.uniform_cube(100, 1), 0.1)
    rng = np.random.default_rng(5)
    for seed, n in ((11, 1), (12, 300), (13, 1000)):
        P = _rand_params(seed)
        X = rng.normal(0, 1.5, (n, om.IN))
        T = rng.normal(2, 1, n)
        P32, X32, T32 = P.astype(np.float32), X.astype(np.float32), T.astype(np.float32)
        print(f"mlp parity seed={seed} n={n}", flush=True)
        y = octohip.predictor.forward(tns, torch.from_numpy(P32), torch.from_numpy(X32)).cpu().numpy()
        y64 = om.forward(P32.astype(np.float64), X32.astype(np.float64))
        np.testing.assert_allclose(y, y64, rtol=1e-4, atol=1e-4)
        g, loss = octohip.predictor.gradient(tns, torch.from_numpy(P32), torch.from_numpy(X32),
                                             torch.from_numpy(T32))
        l64, g64 = om.loss_and_grad(P32.astype(np.float64), X32.astype(np.float64), T32.astype(np.float64))
        assert math.isclose(loss, l64, rel_tol=1e-4)
        g = g.cpu().numpy()
        scale = np.abs(g64).max()
        np.testing.assert_allclose(g, g64, rtol=1e-3, atol=1e-5 * scale)


def _mae(pred_hits, exact):
    return float(np.mean(np.abs(np.log1p(pred_hits) - np.log1p(exact))))


@pytest.fixture(scope="module")
def trained():
    """A predictor trained by the library's kernels on 4 clustered instances
    (100k points each, the 4M-point neighbour count), plus held-out data."""
    import octohip

    torch = _torch()
    tns = octohip.TNS()
    octohip.set_search_radius(tns, 1.0)
    coords = []
    rows = 0
    train_targets = []
    for seed in (101, 102, 103, 104):
        pts, r = _clustered(100_000, seed)
        octohip.set_search_radius(tns, float(r))
        c = torch.from_numpy(pts).cuda()
        coords.append(c)
        sid = octohip.add_point_set(tns, c)
        octohip.set_active_search(tns, sid, sid)
        octohip.run(tns)
        rows += octohip.predictor.collect(tns, sid, sid)
        train_targets.append(octohip.predictor.edge_stats(tns, sid, sid)["exact"].cpu().numpy().astype(np.float64))
    loss0 = octohip.predictor.train(tns, steps=1, batch=2048, lr=1e-3, seed=7, init=True)
    loss = octohip.predictor.train(tns, steps=600, batch=2048, lr=2e-3, seed=7)
    params = octohip.predictor.save(tns)
    print(f"trained rows={rows} loss {loss0:.4f} -> {loss:.4f}", flush=True)
    return {"params": params, "loss0": loss0, "loss": loss,
            "const": float(np.mean(np.log1p(np.concatenate(train_targets))))}


@pytest.mark.gpu
def test_training_lowers_loss_and_beats_a_constant(trained):
    import octohip

    assert trained["loss"] < 0.25 * trained["loss0"]
    pts, r = _clustered(100_000, 201)  # held out
    tns, sid, _ = _tns(pts, r)
    octohip.predictor.load(tns, trained["params"])
    octohip.build_edges(tns, sid, sid, mode="learned")
    st = octohip.predictor.edge_stats(tns, sid, sid)
    assert st["used_predictor"] and st["single_pass"]
    pred = st["predicted"].cpu().numpy().astype(np.float64)
    exact = st["exact"].cpu().numpy().astype(np.float64)
    mae = _mae(pred, exact)
    mae_const = float(np.mean(np.abs(trained["const"] - np.log1p(exact))))
    rel_total = abs(st["predicted_total"] - exact.sum()) / exact.sum()
    print(f"held-out MAE(log1p hits) mlp={mae:.3f} constant={mae_const:.3f} total error={rel_total:.3%}",
          flush=True)
    assert mae < 0.5 * mae_const
    assert rel_total < 0.1


@pytest.mark.gpu
def test_first_call_is_single_pass_only_with_the_predictor(orc, trained):
    import octohip

    pts, r = _clustered(100_000, 202)  # held out
    o = orc.build_edges(orc.run(pts, float(r)), pts)
    want = sort_edges(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    for use in (False, True):
        tns, sid, _ = _tns(pts, r)
        if use:
            octohip.predictor.load(tns, trained["params"])
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        st = octohip.predictor.edge_stats(tns, sid, sid)
        print(f"first learned call predictor={use}: {st['single_pass']=} {st['attempts']=} "
              f"{st['recounted_leaves']=} {st['recorded_leaves']=}", flush=True)
        assert st["single_pass"] == use and st["used_predictor"] == use
        got = sort_edges(e.senders.cpu().numpy(), e.receivers.cpu().numpy(), e.rel_displacement.cpu().numpy(),
                         e.rel_dist_norm.cpu().numpy().reshape(-1))
        for a, b in zip(got, want):
            assert np.array_equal(bits(a), bits(b))


@pytest.mark.gpu
def test_plan_records_more_leaves_than_fixed_runs(orc):
    # Uniform points with 8 per radius-sized cell: leaves of about 8 points
    # and 270 hits, so fixed runs of 4 leaves often overflow a wave's hit
    # list (1024 entries) and recount their last leaf, while the plan packs
    # the predicted hits of a run into the list.
    import octohip

    torch = _torch()

    def lattice(seed):
        n = 100_000
        return datasets.uniform_cube(n, seed), np.float32((8.0 / n) ** (1 / 3))

    tr = octohip.TNS()
    keep = []
    for seed in (11, 12):
        pts, r = lattice(seed)
        octohip.set_search_radius(tr, float(r))
        keep.append(torch.from_numpy(pts).cuda())
        s = octohip.add_point_set(tr, keep[-1])
        octohip.set_active_search(tr, s, s)
        octohip.run(tr)
        octohip.predictor.collect(tr, s, s)
    octohip.predictor.train(tr, steps=400, batch=2048, lr=2e-3, seed=3, init=True)
    params = octohip.predictor.save(tr)

    pts, r = lattice(13)
    tns, sid, _ = _tns(pts, r)
    octohip.build_edges(tns, sid, sid, mode="exact")  # sizes the buffers
    e = octohip.build_edges(tns, sid, sid, mode="learned")  # fixed runs, no predictor
    fixed = octohip.predictor.edge_stats(tns, sid, sid)
    got_fixed = sort_edges(e.senders.cpu().numpy(), e.receivers.cpu().numpy(), e.rel_displacement.cpu().numpy(),
                           e.rel_dist_norm.cpu().numpy().reshape(-1))
    octohip.predictor.load(tns, params)
    e = octohip.build_edges(tns, sid, sid, mode="learned")
    plan = octohip.predictor.edge_stats(tns, sid, sid)
    got_plan = sort_edges(e.senders.cpu().numpy(), e.receivers.cpu().numpy(), e.rel_displacement.cpu().numpy(),
                          e.rel_dist_norm.cpu().numpy().reshape(-1))
    print(f"recounted leaves: fixed runs {fixed['recounted_leaves']} of {fixed['n_leaves']}, "
          f"plan {plan['recounted_leaves']} of {plan['n_leaves']}", flush=True)
    assert fixed["single_pass"] and not fixed["used_predictor"]
    assert plan["single_pass"] and plan["used_predictor"]
    for a, b in zip(got_fixed, got_plan):
        assert np.array_equal(bits(a), bits(b))
    o = orc.build_edges(orc.run(pts, float(r)), pts)
    want = sort_edges(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
    for a, b in zip(got_plan, want):
        assert np.array_equal(bits(a), bits(b))
    assert fixed["recounted_leaves"] > fixed["n_leaves"] // 20  # fixed runs do overflow their lists here
    # the leaves left are those whose candidates do not fit one register tile
    assert plan["recounted_leaves"] < 0.8 * fixed["recounted_leaves"]


@pytest.mark.gpu
def test_learned_edges_bit_exact_small_cases(orc, trained):
    import octohip

    for name, pts, r in (("uniform2000", datasets.uniform_cube(2000, 9), 0.08),
                         ("clustered5000", *_clustered(5000, 301)),
                         ("coincident50", np.zeros((50, 3), np.float32), 0.1)):
        print(f"learned small case {name}", flush=True)
        tns, sid, _ = _tns(pts, r)
        octohip.predictor.load(tns, trained["params"])
        for _ in range(2):
            e = octohip.build_edges(tns, sid, sid, mode="learned")
            got = sort_edges(e.senders.cpu().numpy(), e.receivers.cpu().numpy(), e.rel_displacement.cpu().numpy(),
                             e.rel_dist_norm.cpu().numpy().reshape(-1))
            want = sort_edges(*brute_edges(pts, pts, float(r), exclude_self=True))
            for a, b in zip(got, want):
                assert np.array_equal(bits(a), bits(b))
