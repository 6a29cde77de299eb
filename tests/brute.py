"""O(N^2) brute-force neighbour search in numpy FP32, the same check the
reference's tests use (test/edges_cpu.jl:14-39, test/build.jl:13-28).

It is an independent statement of the edge predicate: numpy's FP32 array
operations round each operation once and never contract to FMA, so
d2 = (dx*dx + dy*dy) + dz*dz with dx = q - t is the reference's value.
"""
import numpy as np


def brute_edges(cq, ct, r, exclude_self):
    """All (receiver i, sender j) with d2 <= r*r, sorted by (i, j); features
    rdisp = -(t_j - q_i) * inv_r and rdist = sqrt(d2) * inv_r in FP32, the
    expressions of the CPU fill (src/cpu/edges_cpu.jl:111-114; the negated
    difference makes a zero component -0.0)."""
    cq = np.ascontiguousarray(cq, np.float32)
    ct = np.ascontiguousarray(ct, np.float32)
    r = np.float32(r)
    r2 = r * r
    inv_r = np.float32(1.0) / r
    S, R, D, N = [], [], [], []
    for i in range(cq.shape[0]):
        d = ct - cq[i]  # (nt, 3) FP32, target minus query (src/cpu/iterator_cpu.jl:96)
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        m = d2 <= r2
        if exclude_self and i < ct.shape[0]:
            m[i] = False
        j = np.nonzero(m)[0].astype(np.int32)
        S.append(j)
        R.append(np.full(j.shape[0], i, np.int32))
        D.append(-d[j] * inv_r)
        N.append(np.sqrt(d2[j]) * inv_r)
    if not S:
        return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 3), np.float32),
                np.zeros(0, np.float32))
    return (np.concatenate(S), np.concatenate(R), np.concatenate(D).astype(np.float32),
            np.concatenate(N).astype(np.float32))


def bits(a):
    """The bits of a float32 array as int32 (array_equal on these tells -0.0
    from +0.0); other arrays unchanged."""
    a = np.asarray(a)
    return np.ascontiguousarray(a).view(np.int32) if a.dtype == np.float32 else a


def sort_edges(senders, receivers, rdisp, rdist):
    """Sort an edge list by (receiver, sender); rows move together."""
    s = np.asarray(senders, np.int64)
    r = np.asarray(receivers, np.int64)
    o = np.lexsort((s, r))
    return (np.asarray(senders)[o], np.asarray(receivers)[o],
            np.asarray(rdisp).reshape(-1, 3)[o], np.asarray(rdist).reshape(-1)[o])


def neighbor_counts(cq, ct, r, exclude_self):
    s, rc, _, _ = brute_edges(cq, ct, r, exclude_self)
    return np.bincount(rc, minlength=np.asarray(cq).shape[0])
