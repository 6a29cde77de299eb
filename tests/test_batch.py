"""Batched sets (one forest, one launch per stage, graph replay): every set's
edges must be exactly its own self-pair edges (bit-exact against the brute
force / CPU oracle per set), replays must equal eager steps, the result must
not depend on leaf size or run length, and a replay that would overflow
must fall back to an eager step."""
import numpy as np
import pytest

from brute import bits, brute_edges, sort_edges
from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _sets(n_sets, seed, lo, hi):
    sets, radii = datasets.batched_sets(n_sets=n_sets, seed=seed, lo=lo, hi=hi, target=20.0)
    # an empty set and a one-point set ride along
    sets = sets[:3] + [np.zeros((0, 3), np.float32), np.full((1, 3), 0.5, np.float32)] + sets[3:]
    radii = radii[:3] + [np.float32(0.1), np.float32(0.1)] + radii[3:]
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    return sets, np.array(radii, np.float32), off


def _batch(sets, radii, off, leaf=32, run=4, graph=True, chunk=4):
    import octohip

    tns = octohip.TNS()
    c = _torch().from_numpy(np.concatenate(sets).astype(np.float32)).cuda()
    b = octohip.Batch(tns, c, off, radii)
    b.config(leaf, run, graph, chunk)
    return tns, b, c


def _per_set(res, n_sets):
    eoff = res.set_offsets.cpu().numpy()
    s_ = res.senders.cpu().numpy()
    r_ = res.receivers.cpu().numpy()
    d_ = res.rel_displacement.cpu().numpy()
    n_ = res.rel_dist_norm.cpu().numpy().reshape(-1)
    out = []
    for k in range(n_sets):
        a, b = int(eoff[k]), int(eoff[k + 1])
        out.append(sort_edges(s_[a:b], r_[a:b], d_[a:b], n_[a:b]))
    return out


def _check_against_oracle(orc, sets, radii, got):
    for k, (pts, r) in enumerate(zip(sets, radii)):
        if pts.shape[0] <= 3000:
            want = sort_edges(*brute_edges(pts, pts, float(r), exclude_self=True))
        else:
            o = orc.build_edges(orc.run(pts, float(r)), pts)
            want = sort_edges(o.senders, o.receivers, o.rel_displacement, o.rel_dist_norm)
        for x, y in zip(got[k], want):
            assert np.array_equal(bits(x), bits(y)), f"set {k}"


def test_batch_matches_each_set(orc):
    sets, radii, off = _sets(20, 7, 300, 8000)
    print(f"batch sets={len(sets)} points={off[-1]}", flush=True)
    tns, b, _ = _batch(sets, radii, off)
    res = b.run()
    assert not res.replayed
    _check_against_oracle(orc, sets, radii, _per_set(res, len(sets)))


def test_batch_graph_replay_equals_eager(orc):
    sets, radii, off = _sets(12, 8, 300, 6000)
    tns, b, c = _batch(sets, radii, off, graph=True)
    b.run()
    rng = np.random.default_rng(3)
    for step in range(3):
        moved = [(s + rng.normal(0, 1e-3, s.shape)).astype(np.float32) for s in sets]
        c.copy_(_torch().from_numpy(np.concatenate(moved)))
        res = b.run()
        print(f"replay step {step}: replayed={res.replayed} edges={res.senders.shape[0]}", flush=True)
        assert res.replayed
        got = _per_set(res, len(sets))
        tns2, b2, _ = _batch(moved, radii, off, graph=False)
        ref = _per_set(b2.run(), len(sets))
        for k in range(len(sets)):
            for x, y in zip(got[k], ref[k]):
                assert np.array_equal(bits(x), bits(y))
        _check_against_oracle(orc, moved, radii, got)


def test_batch_config_does_not_change_edges():
    sets, radii, off = _sets(10, 9, 500, 5000)
    results = []
    for leaf, run, chunk in ((8, 1, 4), (32, 4, 1), (64, 16, 2), (16, 2, 3)):
        print(f"config leaf={leaf} run={run} chunk={chunk}", flush=True)
        tns, b, _ = _batch(sets, radii, off, leaf=leaf, run=run, chunk=chunk)
        results.append(_per_set(b.run(), len(sets)))
    for other in results[1:]:
        for k in range(len(sets)):
            for x, y in zip(results[0][k], other[k]):
                assert np.array_equal(bits(x), bits(y))


def test_batch_replay_overflow_falls_back(orc):
    sets, radii, off = _sets(6, 10, 1000, 4000)
    tns, b, c = _batch(sets, radii, off)
    b.run()
    # pull every set to a tenth of its size: many more neighbours than the buffers hold
    dense = [((s - s.mean(axis=0)) * 0.1 + s.mean(axis=0)).astype(np.float32) if s.shape[0] else s for s in sets]
    c.copy_(_torch().from_numpy(np.concatenate(dense)))
    res = b.run()
    print(f"dense step: replayed={res.replayed} edges={res.senders.shape[0]}", flush=True)
    assert not res.replayed
    _check_against_oracle(orc, dense, radii, _per_set(res, len(sets)))
