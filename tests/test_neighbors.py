"""materialize_all_neighbors! on the GPU: CSR lists of every active pair
against the oracle's materialize_cpu! restatement (offsets identical, each
query's slice the same set; the slice order is the traversal's in both and
is compared sorted, as test/query_cpu.jl:83-106 does), on the reference
tests' shapes and the 100k-point CPU configuration; run! invalidates the
lists (test/contracts.jl:252-270); empty sets (test/edge_cases.jl:42-52)."""
import numpy as np
import pytest

from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _sorted_slices(off, flat):
    """(query, target) keys of a CSR list, sorted, on the device."""
    torch = _torch()
    n = off.shape[0] - 1
    q = torch.repeat_interleave(torch.arange(n, device=flat.device, dtype=torch.int64),
                                (off[1:] - off[:-1]).long())
    return torch.sort((q << 32) | flat.long())[0]


def _check(orc, tns, q, t, cq, ct, r, self_pair, base=0):
    import octohip

    torch = _torch()
    nb = octohip.neighbors(tns, q, t)
    assert nb is not None
    o_off, o_flat = orc.materialize(orc.run(ct, r), cq, None if self_pair else ct)
    assert np.array_equal(nb.offsets.cpu().numpy(), o_off)
    dev = nb.flat.device
    got = _sorted_slices(nb.offsets, nb.flat - base)
    want = _sorted_slices(torch.from_numpy(o_off).to(dev), torch.from_numpy(o_flat).to(dev))
    assert torch.equal(got, want)


CASES = {
    "n50_r015": (lambda: datasets.uniform_cube(50, 10), 0.15),
    "n500_r005": (lambda: datasets.uniform_cube(500, 20), 0.05),
    "n2000_r01": (lambda: datasets.uniform_cube(2000, 30), 0.1),
    "clustered5000": (lambda: (np.float32(0.1) * datasets.normal(40, 0, 15000).astype(np.float32)).reshape(5000, 3),
                      0.03),
    "materialize500": (lambda: datasets.uniform_cube(500, 60), 0.08),
    "plain100k": (lambda: datasets.plain_cpu()[0], 0.02),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_materialize_self_pair(orc, case):
    import octohip

    make, r = CASES[case]
    pts = make()
    print(f"materialize case={case} n={pts.shape[0]} r={r}", flush=True)
    tns = octohip.TNS()
    octohip.set_search_radius(tns, r)
    sid = octohip.add_point_set(tns, _torch().from_numpy(pts).cuda())
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    assert octohip.neighbors(tns, sid, sid) is None
    octohip.materialize_all_neighbors(tns)
    _check(orc, tns, sid, sid, pts, pts, r, True)


def test_materialize_cross_pairs_and_index_base(orc):
    """Two sets a -> b and b -> a at once, 1-based (Julia) indices."""
    import octohip

    torch = _torch()
    a = datasets.uniform_cube(200, 50)
    b = datasets.uniform_cube(300, 51)
    r = 0.1
    tns = octohip.TNS(index_base=1)
    octohip.set_search_radius(tns, r)
    ia = octohip.add_point_set(tns, torch.from_numpy(a).cuda())
    ib = octohip.add_point_set(tns, torch.from_numpy(b).cuda())
    octohip.set_active_search(tns, ia, ib)
    octohip.set_active_search(tns, ib, ia)
    octohip.run(tns)
    octohip.materialize_all_neighbors(tns)
    _check(orc, tns, ia, ib, a, b, r, False, base=1)
    _check(orc, tns, ib, ia, b, a, r, False, base=1)
    got = np.sort(octohip.get_neighborlist(tns, ia, ib, 1).cpu().numpy()) - 1  # query 1 (1-based)
    d = b - a[0]
    d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    assert np.array_equal(got, np.nonzero(d2 <= np.float32(r) * np.float32(r))[0])


def test_run_invalidates_and_empty_sets(orc):
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(3000, 7)
    tns = octohip.TNS()
    octohip.set_search_radius(tns, 0.05)
    c = torch.from_numpy(pts).cuda()
    sid = octohip.add_point_set(tns, c)
    empty = octohip.add_point_set(tns, torch.zeros((0, 3), dtype=torch.float32, device="cuda"))
    octohip.set_active_search(tns, sid, sid)
    octohip.set_active_search(tns, empty, sid)
    octohip.set_active_search(tns, sid, empty)
    octohip.run(tns)
    octohip.materialize_all_neighbors(tns)
    assert octohip.neighbors(tns, empty, sid).offsets.cpu().tolist() == [0]
    off = octohip.neighbors(tns, sid, empty).offsets
    assert off.shape[0] == 3001 and int(off.abs().sum()) == 0
    _check(orc, tns, sid, sid, pts, pts, 0.05, True)
    octohip.update_point_set(tns, sid, c)
    octohip.run(tns)
    assert octohip.neighbors(tns, sid, sid) is None  # run! clears the materialized lists
    # build_edges on the same pair after a materialize still gives the full edges
    octohip.materialize_all_neighbors(tns)
    e = octohip.build_edges(tns, sid, sid)
    o = orc.build_edges(orc.run(pts, 0.05), pts)
    assert e.senders.shape[0] == o.n_edges
    _check(orc, tns, sid, sid, pts, pts, 0.05, True)
