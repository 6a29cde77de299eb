"""Deterministic synthetic point clouds for the five workloads.

One counter-based generator (splitmix64) drives everything, so a (seed,
stream) pair always yields the same FP32 values on any machine: uniforms are
the top 24 bits of a splitmix64 output scaled by 2**-24 (exact in FP32, in
[0, 1)); normals are Box-Muller in FP64 rounded to FP32.  Coordinates are
returned as (N, 3) float32 rows, i.e. xyz interleaved, the same bytes as the
reference's column-major (3, N) matrix.
"""
from __future__ import annotations

import math

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, count: int) -> np.ndarray:
    """count outputs of splitmix64 started at state hash(seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1 + 0x12345) & (2**64 - 1))
        z = base + (np.arange(1, count + 1, dtype=np.uint64) * _GAMMA)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, count: int) -> np.ndarray:
    """FP32 uniforms in [0, 1), exactly representable."""
    u = splitmix64(seed, stream, count) >> np.uint64(40)
    return u.astype(np.float32) * np.float32(2.0**-24)


def normal(seed: int, stream: int, count: int) -> np.ndarray:
    """Standard normals (FP64 Box-Muller) as FP64."""
    m = (count + 1) // 2
    u = splitmix64(seed, stream, 2 * m) >> np.uint64(11)
    u = u.astype(np.float64) * 2.0**-53
    u1 = 1.0 - u[:m]  # (0, 1]
    u2 = u[m:]
    rad = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([rad * np.cos(2 * math.pi * u2), rad * np.sin(2 * math.pi * u2)])
    return z[:count]


def uniform_cube(n: int, seed: int = 1, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
    """n points i.i.d. uniform in [lo, hi)^3."""
    u = uniform(seed, 0, 3 * n).reshape(n, 3)
    if lo == 0.0 and hi == 1.0:
        return u
    return (np.float32(lo) + np.float32(hi - lo) * u).astype(np.float32)


def gaussian_blob(n: int, seed: int, sigma: float, center=(0.0, 0.0, 0.0)) -> np.ndarray:
    z = normal(seed, 1, 3 * n).reshape(n, 3)
    return (np.asarray(center, np.float64) + sigma * z).astype(np.float32)


# ---------------------------------------------------------------- workloads


def plain_cpu(seed: int = 1):
    """100,000 uniform points in [0,1)^3, radius 0.02, searched against itself."""
    return uniform_cube(100_000, seed), np.float32(0.02)


def sph_lattice(seed: int = 2, side: int = 100, h: float = 0.01):
    """side^3 lattice with spacing h and U(+-0.1h) jitter per axis, radius 2h.
    Points are numbered x fastest, then y, then z."""
    g = np.arange(side, dtype=np.float64) * h
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    base = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    n = base.shape[0]
    jit = (uniform(seed, 0, 3 * n).astype(np.float64).reshape(n, 3) * 2.0 - 1.0) * (0.1 * h)
    return (base + jit).astype(np.float32), np.float32(2 * h)


def sph_step(coords: np.ndarray, step: int, h: float = 0.01, seed: int = 2) -> np.ndarray:
    """One time step of the SPH block: add N(0, (0.01h)^2) per coordinate."""
    n = coords.shape[0]
    d = normal(seed * 1000 + step, 2, 3 * n).reshape(n, 3) * (0.01 * h)
    return (coords.astype(np.float64) + d).astype(np.float32)


def sph_snapshots(base, steps: int, h: float = 0.01, seed: int = 2):
    """The SPH block's time steps on the device of `base` (an (N, 3) float32
    torch tensor): snapshot k is snapshot k-1 plus N(0, (0.01h)^2) per
    coordinate, drawn in float64 by a torch generator seeded with `seed` on
    that device and rounded to float32.  Returns `steps` contiguous tensors
    (the first is one step past `base`)."""
    import torch

    g = torch.Generator(device=base.device)
    g.manual_seed(seed)
    snaps = []
    cur = base
    for _ in range(steps):
        cur = (cur.double() + torch.randn(base.shape, generator=g, device=base.device, dtype=torch.float64)
               * (0.01 * h)).float().contiguous()
        snaps.append(cur)
    return snaps


def gaussian_mixture(n: int, seed: int, n_comp: int, sigma_lo: float, sigma_hi: float,
                     stream: int = 3) -> np.ndarray:
    centers = uniform(seed, stream, 3 * n_comp).astype(np.float64).reshape(n_comp, 3)
    sig = sigma_lo + (sigma_hi - sigma_lo) * uniform(seed, stream + 1, n_comp).astype(np.float64)
    comp = (splitmix64(seed, stream + 2, n) % np.uint64(n_comp)).astype(np.int64)
    z = normal(seed, stream + 3, 3 * n).reshape(n, 3)
    return (centers[comp] + sig[comp, None] * z).astype(np.float32)


def mean_neighbors_sample(pts: np.ndarray, r: float, sample: int = 512, seed: int = 0) -> float:
    """Mean neighbour count (self excluded) of a fixed sample of query points."""
    n = pts.shape[0]
    idx = (splitmix64(seed, 9, min(sample, n)) % np.uint64(n)).astype(np.int64)
    q = pts[idx].astype(np.float32)
    r2 = np.float32(r) * np.float32(r)
    tot = 0
    for k in range(q.shape[0]):
        d = pts - q[k]
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        tot += int(np.count_nonzero(d2 <= r2)) - 1
    return tot / q.shape[0]


def calibrate_radius(pts: np.ndarray, target: float, seed: int = 0, sample: int = 128,
                     keep: int = 256) -> np.float32:
    """Bisection (in log r) for the radius whose sample mean neighbour count is
    target.  The FP32 squared distances of a fixed sample of query points are
    computed once; each sample keeps its `keep` smallest (counts saturate
    there, far above the targets used here)."""
    n = pts.shape[0]
    if n < 2:
        return np.float32(0.1)
    idx = (splitmix64(seed, 9, min(sample, n)) % np.uint64(n)).astype(np.int64)
    q = pts[idx].astype(np.float32)
    p32 = pts.astype(np.float32)
    kk = min(keep, n)
    rows = []
    for a in range(0, q.shape[0], 64):
        d = p32[None, :, :] - q[a:a + 64, None, :]
        d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        part = np.partition(d2, kk - 1, axis=1)[:, :kk]
        rows.append(np.sort(part, axis=1))
    near = np.concatenate(rows)

    def mean_count(r):
        r2 = np.float32(r) * np.float32(r)
        return float(np.mean(np.sum(near <= r2, axis=1) - 1))

    lo, hi = 1e-4, 0.5
    for _ in range(30):
        mid = math.sqrt(lo * hi)
        if mean_count(mid) < target:
            lo = mid
        else:
            hi = mid
    return np.float32(math.sqrt(lo * hi))


def batched_sets(n_sets: int = 256, seed: int = 3, lo: int = 1000, hi: int = 65536,
                 target: float = 20.0):
    """n_sets point sets, sizes log-uniform in [lo, hi], each a 4-component
    isotropic Gaussian mixture (centers U[0,1)^3, sigma U[0.02, 0.1]); each
    radius is calibrated so the mean neighbour count is about target."""
    u = uniform(seed, 7, n_sets).astype(np.float64)
    sizes = np.floor(np.exp(math.log(lo) + u * (math.log(hi) - math.log(lo)))).astype(np.int64)
    sets, radii = [], []
    for s in range(n_sets):
        pts = gaussian_mixture(int(sizes[s]), seed * 100003 + s, 4, 0.02, 0.1)
        sets.append(pts)
        radii.append(calibrate_radius(pts, target, seed=s))
    return sets, radii


def clustered(n: int = 4_000_000, seed: int = 4, n_clusters: int = 16, sigma: float = 0.05,
              frac_uniform: float = 0.1):
    """90% from a 16-cluster Gaussian mixture (sigma 0.05, centers U[0,1)^3),
    10% uniform background; radius 0.01."""
    n_bg = int(round(n * frac_uniform))
    n_cl = n - n_bg
    centers = uniform(seed, 11, 3 * n_clusters).astype(np.float64).reshape(n_clusters, 3)
    comp = (splitmix64(seed, 12, n_cl) % np.uint64(n_clusters)).astype(np.int64)
    z = normal(seed, 13, 3 * n_cl).reshape(n_cl, 3)
    cl = (centers[comp] + sigma * z).astype(np.float32)
    bg = uniform(seed, 14, 3 * n_bg).reshape(n_bg, 3)
    return np.concatenate([cl, bg], axis=0), np.float32(0.01)


def differentiable(n: int = 262_144, seed: int = 5):
    """n uniform points in [0,1)^3 with the radius for about 24 neighbours."""
    r = (24.0 / (n * 4.0 / 3.0 * math.pi)) ** (1.0 / 3.0)
    return uniform_cube(n, seed), np.float32(r)
