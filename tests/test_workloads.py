"""Parity at the benchmarked configurations, at their full sizes.

- The headline path exactly as the benchmark times it: the 1M-point SPH
  block in refit mode, the learned single pass with the shipped sph_lattice
  predictor, the benchmark's own time steps (the same generator and seed).
- The batched configuration: all 256 sets (sizes log-uniform in
  [1000, 65536]), one eager step and one graph replay after an in-place
  update, every set against the oracle.
- The learned predictor's configuration: 4M clustered points with the
  shipped clustered4m predictor; per-query counts and the total against the
  oracle's count pass, and the full rows of a sample of receivers against
  the oracle's fill of those receivers (the whole edge set is ~15 GB).

Rows are compared bit for bit (int32 views of the features) after sorting
both sides by (receiver, sender) on the device; the reference's
within-receiver order is its DFS order, which the edge set does not depend on
(test/edges_cpu.jl:14-79 compares sorted sets the same way).
"""
import numpy as np
import pytest

from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _keyed(senders, receivers, rdisp, rdist, shift=0):
    """Rows sorted by (receiver, sender) on the device: the int64 keys and the
    int32 bits of rel_displacement (E, 3) and rel_dist_norm (E,)."""
    torch = _torch()
    r = receivers.long() + shift
    s = senders.long() + shift
    key, o = torch.sort((r << 32) | s)
    disp = rdisp.reshape(-1, 3)[o].contiguous().view(torch.int32)
    dist = rdist.reshape(-1)[o].contiguous().view(torch.int32)
    return key, disp, dist


def _oracle_keyed(o, dev, shift=0):
    torch = _torch()
    return _keyed(torch.from_numpy(o.senders).to(dev), torch.from_numpy(o.receivers).to(dev),
                  torch.from_numpy(o.rel_displacement).to(dev), torch.from_numpy(o.rel_dist_norm).to(dev), shift)


def _assert_same(got, want, what):
    torch = _torch()
    assert got[0].shape == want[0].shape, f"{what}: {got[0].shape[0]} edges, oracle {want[0].shape[0]}"
    assert torch.equal(got[0], want[0]), f"{what}: (receiver, sender) pairs differ"
    assert torch.equal(got[1], want[1]), f"{what}: rel_displacement bits differ"
    assert torch.equal(got[2], want[2]), f"{what}: rel_dist_norm bits differ"


def test_headline_sph_learned_refit_steps(orc):
    """1M SPH points, refit mode, learned mode with the sph_lattice predictor,
    four of the benchmark's time steps: the first call (full sort, predictor
    plan) and the last (refit, fixed runs, early total) are compared row for
    row with the oracle; every step's per-query counts equal the oracle's."""
    import octohip

    torch = _torch()
    pts, r = datasets.sph_lattice()
    base = torch.from_numpy(pts).cuda()
    snaps = datasets.sph_snapshots(base, 4, h=0.01, seed=2)
    tns = octohip.TNS()
    octohip.set_search_radius(tns, float(r))
    octohip.set_refit_mode(tns, True)
    sid = octohip.add_point_set(tns, snaps[0])
    octohip.set_active_search(tns, sid, sid)
    octohip.predictor.load(tns, "sph_lattice")
    for k, snap in enumerate(snaps):
        octohip.update_point_set(tns, sid, snap)
        octohip.run(tns)
        e = octohip.build_edges(tns, sid, sid, mode="learned")
        info = octohip.build_info(tns, sid)
        st = octohip.predictor.edge_stats(tns, sid, sid, tensors=False)
        print(f"sph step {k}: E={e.senders.shape[0]} refit={info['refit_used']} single={st['single_pass']} "
              f"predictor={st['used_predictor']} attempts={st['attempts']}", flush=True)
        assert st["single_pass"]
        if k == 0:
            assert st["used_predictor"]
        else:
            assert info["refit_used"]
        host = snap.cpu().numpy()
        t = orc.run(host, float(r))
        counts, total = orc.count_edges(t, host)
        assert e.senders.shape[0] == total
        got_counts = torch.bincount(e.receivers.long(), minlength=host.shape[0]).cpu().numpy()
        assert np.array_equal(got_counts, counts)
        if k in (0, len(snaps) - 1):
            o = orc.build_edges(t, host)
            _assert_same(_keyed(*e), _oracle_keyed(o, base.device), f"step {k}")


def test_batched_256_sets_full_config(orc):
    """All 256 sets of the batched configuration (about 4.3M points): an
    eager step, then an in-place update and a graph replay; every set's edges
    equal the oracle's on both."""
    import octohip

    torch = _torch()
    sets, radii = datasets.batched_sets()
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    print(f"batched sets={len(sets)} points={off[-1]}", flush=True)
    host = np.concatenate(sets).astype(np.float32)
    coords = torch.from_numpy(host).cuda()
    tns = octohip.TNS()
    b = octohip.Batch(tns, coords, off, radii)
    b.config(32, 4, True, 1)

    def check(res, pts_all, what):
        eoff = res.set_offsets.cpu().numpy()
        assert int(eoff[-1]) == res.senders.shape[0]
        for k in range(len(sets)):
            a, z = int(eoff[k]), int(eoff[k + 1])
            pts = pts_all[off[k]:off[k + 1]]
            o = orc.build_edges(orc.run(pts, float(radii[k])), pts)
            got = _keyed(res.senders[a:z], res.receivers[a:z], res.rel_displacement[a:z], res.rel_dist_norm[a:z])
            _assert_same(got, _oracle_keyed(o, coords.device), f"{what} set {k} (n={pts.shape[0]})")

    res = b.run()
    assert not res.replayed
    check(res, host, "eager")
    g = np.random.default_rng(5)
    moved = (host.astype(np.float64) + g.normal(0, 1e-4, host.shape)).astype(np.float32)
    coords.copy_(torch.from_numpy(moved))
    res = b.run()
    assert res.replayed
    check(res, moved, "replay")


def test_clustered_4m_learned_counts_and_rows(orc):
    """4M clustered points, radius 0.01, learned mode with the clustered4m
    predictor (~6.3e8 edges): E and every per-query count equal the oracle's
    count pass; the rows of 12,000 sampled receivers (plus the 64 with the
    most neighbours) equal the oracle's fill of those receivers."""
    import octohip

    torch = _torch()
    pts, r = datasets.clustered()
    n = pts.shape[0]
    coords = torch.from_numpy(pts).cuda()
    tns = octohip.TNS()
    octohip.set_search_radius(tns, float(r))
    octohip.set_refit_mode(tns, True)
    sid = octohip.add_point_set(tns, coords)
    octohip.set_active_search(tns, sid, sid)
    octohip.predictor.load(tns, "clustered4m")
    octohip.run(tns)
    e = octohip.build_edges(tns, sid, sid, mode="learned")
    st = octohip.predictor.edge_stats(tns, sid, sid, tensors=False)
    print(f"clustered 4M: E={e.senders.shape[0]} single={st['single_pass']} predictor={st['used_predictor']} "
          f"heavy={st['heavy_leaves']} of {st['n_leaves']} leaves", flush=True)
    assert st["single_pass"] and st["used_predictor"]
    t = orc.run(pts, float(r))
    counts, total = orc.count_edges(t, pts)
    print(f"oracle count pass: E={total}", flush=True)
    assert e.senders.shape[0] == total
    got_counts = torch.bincount(e.receivers.long(), minlength=n).cpu().numpy()
    assert np.array_equal(got_counts, counts)
    g = np.random.default_rng(11)
    rows = np.union1d(g.choice(n, 12_000, replace=False), np.argsort(counts)[-64:]).astype(np.int64)
    sel = torch.zeros(n, dtype=torch.bool, device=coords.device)
    sel[torch.from_numpy(rows).to(coords.device)] = True
    m = sel[e.receivers.long()]
    got = _keyed(e.senders[m], e.receivers[m], e.rel_displacement[m], e.rel_dist_norm[m])
    o = orc.build_edges_rows(t, pts, rows)
    assert int(o.n_edges) == int(counts[rows].sum())
    _assert_same(got, _oracle_keyed(o, coords.device), "sampled receivers")
