"""Learned per-leaf hit-count predictor of the single-pass build_edges.

The reference has no counterpart (its GPU build_edges always runs the exact
count pass, ext/OctopusCUDAExt.jl:373-387).  A 29 -> 64 -> 64 -> 1 ReLU MLP
predicts log1p(hits) of every query leaf from its point count, tree level,
extent and the point counts of its adjacent candidate leaves in 26
directions.  build_edges(mode="learned") uses the predictions to size the
edge buffers on a pair's first call and to plan which leaves each wave emits
in one sweep; the edges never depend on it.  Training, inference and the
plan run in the library's own HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _lib as L

PARAMS = L.MLP_PARAMS
TRAINED = L.MLP_TRAINED
FEATURES = L.MLP_FEATURES
SHIPPED = os.path.join(L.HERE, "predictors")


def _ptr(x):
    return ctypes.c_void_p(x.data_ptr())


def load(tns, params):
    """Install a parameter block (array of PARAMS float32), a path to one
    (raw little-endian float32), or the name of a shipped block."""
    if params is None:
        tns._check(tns._lib.oct_predictor_load(tns._c(), None))
        return
    if isinstance(params, str):
        path = params if os.path.exists(params) else os.path.join(SHIPPED, params + ".f32")
        params = np.fromfile(path, dtype="<f4")
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
    if p.shape != (PARAMS,):
        from . import ArgumentError

        raise ArgumentError(f"a predictor has {PARAMS} parameters, got shape {p.shape}")
    tns._check(tns._lib.oct_predictor_load(tns._c(), p.ctypes.data_as(ctypes.c_void_p)))


def set_policy(tns, always: bool):
    """always=False (default): the learned mode runs the predictor on a pair's
    first call and after a call that overflowed, fell back or recounted more
    than 1/64 of its leaves; always=True: on every call."""
    tns._check(tns._lib.oct_set_predictor_policy(tns._c(), int(bool(always))))


def save(tns, path: str | None = None) -> np.ndarray:
    """The installed parameter block (and written to path when given)."""
    p = np.empty(PARAMS, dtype=np.float32)
    tns._check(tns._lib.oct_predictor_save(tns._c(), p.ctypes.data_as(ctypes.c_void_p)))
    if path is not None:
        p.astype("<f4").tofile(path)
    return p


def collect(tns, q: int, t: int) -> int:
    """Exact per-leaf hit counts of the active pair (q, t) (count pass only,
    no edges are written); their features and counts join the training set.
    Returns the rows added."""
    rows = ctypes.c_int64()
    tns._check(tns._lib.oct_predictor_collect(tns._enter(), int(q), int(t), ctypes.byref(rows)))
    tns._leave()
    return int(rows.value)


def clear_data(tns):
    tns._check(tns._lib.oct_predictor_clear_data(tns._c()))


def train(tns, steps: int = 2000, batch: int = 4096, lr: float = 1e-3, seed: int = 1, init: bool = False) -> float:
    """Adam steps on the training set; returns the last batch's mean squared
    error of log1p(hits)."""
    loss = ctypes.c_float()
    tns._check(tns._lib.oct_predictor_train(tns._enter(), int(steps), int(batch), float(lr), int(seed),
                                            int(bool(init)), ctypes.byref(loss)))
    tns._leave()
    return float(loss.value)


def edge_stats(tns, q: int, t: int, tensors: bool = True, wait: bool = True) -> dict:
    """What the last build of (q, t) did, with device tensors of the per-leaf
    features, predicted hits (None unless the predictor ran) and exact hits
    (all None with tensors=False).  wait=False never waits for the stream:
    the recount counters of a call that returned on the single pass's early
    total are then the previous call's until its kernels end."""
    from . import _wrap, _torch

    torch = _torch()
    s = L.OctEdgeStats()
    get = tns._lib.oct_edge_stats_get if wait else tns._lib.oct_edge_stats_peek
    tns._check(get(tns._c(), int(q), int(t), ctypes.byref(s)))
    n, dev = int(s.n_leaves), tns.device
    return {
        "features": _wrap(s.features, (n, FEATURES), torch.float32, dev, tns) if tensors and s.features else None,
        "predicted": _wrap(s.predicted, (n,), torch.float32, dev, tns) if tensors and s.predicted else None,
        "exact": _wrap(s.exact, (n,), torch.int32, dev, tns) if tensors and s.exact and n else None,
        "n_leaves": n,
        "n_edges": int(s.n_edges),
        "recounted_leaves": int(s.recounted_leaves),
        "recorded_leaves": int(s.recorded_leaves),
        "heavy_leaves": int(s.heavy_leaves),
        "predicted_total": float(s.predicted_total),
        "single_pass": bool(s.single_pass),
        "used_predictor": bool(s.used_predictor),
        "attempts": int(s.attempts),
    }


def forward(tns, params, features):
    """Network output (log1p of the predicted hits) on device float32 rows."""
    torch = __import__("torch")
    params = params.to(device=tns.device, dtype=torch.float32).contiguous()
    x = features.to(device=tns.device, dtype=torch.float32).contiguous()
    y = torch.empty(x.shape[0], dtype=torch.float32, device=tns.device)
    tns._check(tns._lib.oct_mlp_forward(tns._enter(), _ptr(params), _ptr(x), int(x.shape[0]), _ptr(y)))
    tns._leave()
    return y


def gradient(tns, params, features, targets):
    """(gradient over the TRAINED parameters, loss) of mean((f(x) - t)^2)."""
    torch = __import__("torch")
    params = params.to(device=tns.device, dtype=torch.float32).contiguous()
    x = features.to(device=tns.device, dtype=torch.float32).contiguous()
    t = targets.to(device=tns.device, dtype=torch.float32).contiguous()
    g = torch.empty(TRAINED, dtype=torch.float32, device=tns.device)
    loss = torch.empty(1, dtype=torch.float32, device=tns.device)
    tns._check(tns._lib.oct_mlp_gradient(tns._enter(), _ptr(params), _ptr(x), _ptr(t), int(x.shape[0]), _ptr(g),
                                         _ptr(loss)))
    tns._leave()
    return g, float(loss.item())
