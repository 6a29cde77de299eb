"""NumPy restatement of the per-leaf hit-count predictor: features, network
forward and the gradient of its training loss.

TEST INFRASTRUCTURE ONLY.  Imported by tests/ as the checker of the HIP
predictor kernels; the product package never imports it.

The reference has no predictor (its GPU build_edges always runs the exact
count pass, ext/OctopusCUDAExt.jl:373-387), so nothing here follows a
reference file.  What pins it:
  * features: the candidate leaves are every target leaf whose FP32 box gap
    (the expression of the candidate test, a monotone lower bound of the
    point distances) is within r; a brute-force point scan confirms that no
    pair within r lies outside the candidates (tests/test_predictor.py);
  * forward / gradient: central finite differences in float64 of the same
    loss (tests/test_predictor.py), plus a closed form for a one-layer case.

Parameter block (floats): feature mean[29], std[29], W1[29*64] (row =
input), b1[64], W2[64*64] (row = input), b2[64], W3[64], b3[1].
"""
from __future__ import annotations

import numpy as np

IN = 29
H = 64
ADJ = 26
MU = 0
SD = MU + IN
W1 = SD + IN
B1 = W1 + IN * H
W2 = B1 + H
B2 = W2 + H * H
W3 = B2 + H
B3 = W3 + H
PARAMS = B3 + 1
TRAINED = PARAMS - W1


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return dict(mu=P[MU:SD], sd=P[SD:W1], W1=P[W1:B1].reshape(IN, H), b1=P[B1:W2],
                W2=P[W2:B2].reshape(H, H), b2=P[B2:W3], W3=P[W3:B3], b3=P[B3])


def forward(P, X):
    """y = network(X) in float64 (log1p of the predicted hits)."""
    p = unpack(P)
    xn = (np.asarray(X, np.float64) - p["mu"]) / p["sd"]
    h1 = np.maximum(xn @ p["W1"] + p["b1"], 0.0)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(a2, 0.0)
    return h2 @ p["W3"] + p["b3"]


def loss_and_grad(P, X, T):
    """mean((y - T)^2) and its gradient with respect to the trained
    parameters (W1, b1, W2, b2, W3, b3 in storage order), float64."""
    p = unpack(P)
    X = np.asarray(X, np.float64)
    T = np.asarray(T, np.float64)
    n = X.shape[0]
    xn = (X - p["mu"]) / p["sd"]
    a1 = xn @ p["W1"] + p["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ p["W2"] + p["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ p["W3"] + p["b3"]
    diff = y - T
    loss = float(np.mean(diff * diff))
    g = 2.0 * diff / n                                 # dL/dy
    gW3 = h2.T @ g
    gb3 = g.sum()
    d2 = np.outer(g, p["W3"]) * (a2 > 0)             # dL/da2
    gW2 = h1.T @ d2
    gb2 = d2.sum(axis=0)
    d1 = (d2 @ p["W2"].T) * (a1 > 0)                  # dL/da1
    gW1 = xn.T @ d1
    gb1 = d1.sum(axis=0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])
    return loss, grad


def box_gap_sq(alo, ahi, blo, bhi):
    """FP32 squared gap of boxes, evaluated like the candidate test:
    per axis max(max(blo - ahi, alo - bhi), 0), then (dx*dx + dy*dy) + dz*dz."""
    f = np.float32
    d = np.maximum(np.maximum(blo - ahi, alo - bhi), f(0))
    d = d.astype(np.float32)
    return ((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(np.float32)


def dir_codes(qlo, qhi, clo, chi):
    """27-way direction of box c from box q: per axis 0 below / 1 level / 2
    above, off-level when the centres differ by more than a quarter of the
    two extents summed; code = sx + 3 sy + 9 sz."""
    f = np.float32
    d = (f(0.5) * (clo + chi)).astype(np.float32) - (f(0.5) * (qlo + qhi)).astype(np.float32)
    d = d.astype(np.float32)
    tau = (f(0.25) * ((qhi - qlo).astype(np.float32) + (chi - clo).astype(np.float32))).astype(np.float32)
    s = np.where(d > tau, 2, np.where(d < -tau, 0, 1))
    return s[..., 0] + 3 * s[..., 1] + 9 * s[..., 2]


def leaves_of(tree):
    """Leaf node ids of an oracle tree in Morton order (by first position)."""
    kids = tree.node_children
    leaf = (kids < 0).all(axis=1) & (tree.node_last >= tree.node_first)
    ids = np.nonzero(leaf)[0]
    return ids[np.argsort(tree.node_first[ids], kind="stable")]


def leaf_features(tq, tt, r):
    """(L, 29) float32 features of the query leaves of tree tq against the
    target tree tt (brute force over leaf pairs; small trees only)."""
    f = np.float32
    lq = leaves_of(tq)
    lt = leaves_of(tt)
    qlo, qhi = tq.bmin[lq].astype(np.float32), tq.bmax[lq].astype(np.float32)
    tlo, thi = tt.bmin[lt].astype(np.float32), tt.bmax[lt].astype(np.float32)
    tcnt = (tt.node_last[lt] - tt.node_first[lt] + 1).astype(np.float64)
    r_sq = f(f(r) * f(r))
    out = np.zeros((lq.shape[0], IN), np.float32)
    for k in range(lq.shape[0]):
        gap = box_gap_sq(qlo[k], qhi[k], tlo, thi)
        cand = np.nonzero(gap <= r_sq)[0]
        codes = dir_codes(qlo[k], qhi[k], tlo[cand], thi[cand])
        acc = np.zeros(27, np.float64)
        np.add.at(acc, codes, tcnt[cand])
        adj = np.delete(acc, 13).astype(np.float32)
        ext = f(max(qhi[k] - qlo[k]))
        out[k, 0] = np.log1p(f(tq.node_last[lq[k]] - tq.node_first[lq[k]] + 1))
        out[k, 1] = f(tq.node_level[lq[k]]) / f(21)
        out[k, 2] = max(np.log2(max(ext / f(r), f(1e-30))), f(-20))
        out[k, 3:] = np.log1p(adj)
    return out
