"""apply_zsort on the GPU and its adjoint, against the oracle's
apply_zsort_cpu! / rrule scatter restatements with the oracle's permutation
(which the tree parity tests show equal to the engine's): vectors and
matrices of 1-, 2-, 4- and 8-byte elements bit for bit, the backward of the
autograd function, 1-based indices, and the reference's error cases
(test/zsort.jl:23-78, test/zsort_cuda.jl:14-70: higher rank rejected,
wrong length, run! first)."""
import numpy as np
import pytest

from octohip import datasets

pytestmark = pytest.mark.gpu


def _torch():
    import torch

    return torch


def _tns(pts, r, base=0):
    import octohip

    tns = octohip.TNS(index_base=base)
    octohip.set_search_radius(tns, r)
    sid = octohip.add_point_set(tns, _torch().from_numpy(pts).cuda())
    octohip.set_active_search(tns, sid, sid)
    octohip.run(tns)
    return tns, sid


@pytest.mark.parametrize("base", [0, 1])
def test_apply_zsort_matches_oracle(orc, base):
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(50_000, 81)
    r = 0.03
    tns, sid = _tns(pts, r, base)
    perm = orc.run(pts, r).perm
    g = np.random.default_rng(4)
    arrays = [
        g.normal(size=50_000).astype(np.float32),
        g.normal(size=(50_000, 3)).astype(np.float32),
        g.normal(size=(50_000, 7)).astype(np.float64),
        g.integers(-2**31, 2**31 - 1, size=(50_000, 2)).astype(np.int32),
        g.integers(0, 255, size=(50_000, 5)).astype(np.uint8),
        g.integers(-30000, 30000, size=50_000).astype(np.int16),
        g.integers(-2**62, 2**62, size=(50_000, 1)).astype(np.int64),
    ]
    for a in arrays:
        z = octohip.apply_zsort(tns, sid, torch.from_numpy(a).cuda())
        assert z.shape == a.shape and z.dtype == torch.from_numpy(a).dtype
        assert np.array_equal(z.cpu().numpy(), orc.apply_zsort(perm, a)), a.dtype
    # the coordinates themselves: the z-sorted copy the tree was built on
    zc = octohip.apply_zsort(tns, sid, torch.from_numpy(pts).cuda())
    assert np.array_equal(zc.cpu().numpy(), pts[perm])


def test_apply_zsort_backward(orc):
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(20_000, 82)
    tns, sid = _tns(pts, 0.04)
    perm = orc.run(pts, 0.04).perm
    x = torch.randn(20_000, 4, device="cuda", requires_grad=True)
    w = torch.randn(20_000, 4, device="cuda")
    y = octohip.apply_zsort(tns, sid, x)
    (y * w).sum().backward()
    want = orc.zsort_pullback(perm, w.cpu().numpy())
    assert np.array_equal(x.grad.cpu().numpy(), want)
    v = torch.randn(20_000, device="cuda", dtype=torch.float64, requires_grad=True)
    octohip.apply_zsort(tns, sid, v).sum().backward()
    assert torch.equal(v.grad, torch.ones_like(v))


def test_apply_zsort_errors():
    import octohip

    torch = _torch()
    pts = datasets.uniform_cube(1000, 83)
    tns, sid = _tns(pts, 0.05)
    with pytest.raises(octohip.ArgumentError):
        octohip.apply_zsort(tns, sid, torch.zeros(1000, 2, 2, device="cuda"))
    with pytest.raises(octohip.DimensionMismatch):
        octohip.apply_zsort(tns, sid, torch.zeros(999, device="cuda"))
    octohip.update_point_set(tns, sid, torch.from_numpy(pts).cuda())
    with pytest.raises(octohip.OctopusError):
        octohip.prepare_zsort(tns)
    with pytest.raises(octohip.OctopusError):  # the permutation is stale until run!
        octohip.apply_zsort(tns, sid, torch.zeros(1000, device="cuda"))
    octohip.run(tns)
    octohip.prepare_zsort(tns)
    assert octohip.apply_zsort(tns, sid, torch.zeros(1000, device="cuda")).shape == (1000,)
