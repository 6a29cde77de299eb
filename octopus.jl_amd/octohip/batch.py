"""Batched point sets: many sets, each searched against itself with its own
radius, built and emitted as one forest (one launch per stage for all sets),
the step captured once and replayed as a HIP graph.

The reference runs the sets of a TNS one after the other (src/api.jl:196);
this has no counterpart there.  Edge indices are local to their set.
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np

from . import _lib as L

BatchEdges = namedtuple("BatchEdges", ["senders", "receivers", "rel_displacement", "rel_dist_norm",
                                       "set_offsets", "replayed", "recounted_leaves", "recorded_leaves",
                                       "heavy_leaves"])


class Batch:
    """Sets concatenated in one (N, 3) float32 device tensor; offsets: S + 1
    point offsets; radii: S radii.  The tensor is referenced: update its
    contents in place between runs."""

    def __init__(self, tns, coords, offsets, radii):
        from . import ArgumentError, DimensionMismatch, _torch

        torch = _torch()
        if not (isinstance(coords, torch.Tensor) and coords.is_cuda and coords.dtype == torch.float32):
            raise ArgumentError("batch coords must be a float32 device tensor")
        if coords.dim() != 2 or coords.shape[1] != 3:
            raise DimensionMismatch("batch coords must have shape (N, 3)")
        if not coords.is_contiguous():
            # a private copy would freeze the coordinates: in-place updates of
            # the caller's tensor would then never reach the batch
            raise ArgumentError("batch coords must be contiguous (they are referenced, not copied)")
        self.tns = tns
        self.coords = coords
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        self.radii = np.ascontiguousarray(np.asarray(radii, dtype=np.float32))
        if self.offsets.shape[0] != self.radii.shape[0] + 1 or int(self.offsets[-1]) != coords.shape[0]:
            raise DimensionMismatch("offsets must have n_sets + 1 entries ending at N")
        bid = ctypes.c_int()
        tns._check(tns._lib.oct_batch_create(tns._c(), int(self.radii.shape[0]),
                                             ctypes.c_void_p(self.coords.data_ptr()),
                                             self.offsets.ctypes.data_as(ctypes.c_void_p),
                                             self.radii.ctypes.data_as(ctypes.c_void_p), ctypes.byref(bid)))
        self.id = bid.value
        self.n_sets = int(self.radii.shape[0])

    def config(self, leaf_size: int = 32, run_leaves: int = 6, use_graph: bool = True, run_chunk: int = 4):
        self.tns._check(self.tns._lib.oct_batch_config(self.tns._c(), self.id, int(leaf_size), int(run_leaves),
                                                       int(run_chunk), int(bool(use_graph))))

    def run(self) -> BatchEdges:
        """run! + build_edges of every set; tensors alias library buffers
        until the next run of the batch."""
        from . import _torch, _wrap

        torch = _torch()
        out = L.OctBatchEdges()
        self.tns._check(self.tns._lib.oct_batch_run(self.tns._enter(), self.id, ctypes.byref(out)))
        self.tns._leave()
        n, dev = int(out.n_edges), self.tns.device
        return BatchEdges(
            senders=_wrap(out.senders, (n,), torch.int32, dev, self),
            receivers=_wrap(out.receivers, (n,), torch.int32, dev, self),
            rel_displacement=_wrap(out.rel_displacement, (n, 3), torch.float32, dev, self),
            rel_dist_norm=_wrap(out.rel_dist_norm, (n, 1), torch.float32, dev, self),
            set_offsets=_wrap(out.set_offsets, (self.n_sets + 1,), torch.int64, dev, self),
            replayed=bool(out.replayed),
            recounted_leaves=int(out.recounted_leaves),
            recorded_leaves=int(out.recorded_leaves),
            heavy_leaves=int(out.heavy_leaves),
        )

    def autotune(self, leaf_sizes=(8, 12, 16, 24, 32), runs=(1, 2, 4, 8), chunks=(1, 4), reps: int = 3) -> dict:
        """Times every (leaf size, leaves per run, runs per block ticket)
        candidate on this batch's current coordinates and keeps the fastest;
        returns {"best": (leaf, run, chunk), "ms": {(leaf, run, chunk): ms per
        step}}."""
        from . import _tune_table

        ls = np.ascontiguousarray(np.asarray(leaf_sizes, dtype=np.int32))
        rs = np.ascontiguousarray(np.asarray(runs, dtype=np.int32))
        cs = np.ascontiguousarray(np.asarray(chunks, dtype=np.int32))
        ms = np.zeros(ls.shape[0] * rs.shape[0] * cs.shape[0], dtype=np.float64)
        best = ctypes.c_int()
        self.tns._check(self.tns._lib.oct_batch_autotune(
            self.tns._enter(), self.id, ls.ctypes.data_as(ctypes.c_void_p), int(ls.shape[0]),
            rs.ctypes.data_as(ctypes.c_void_p), int(rs.shape[0]), cs.ctypes.data_as(ctypes.c_void_p),
            int(cs.shape[0]), int(reps), ms.ctypes.data_as(ctypes.c_void_p), ctypes.byref(best)))
        self.tns._leave()
        top, table = _tune_table(ls, rs, cs, ms, best.value)
        return {"best": top, "ms": table}

    def close(self):
        if self.id is not None:
            self.tns._check(self.tns._lib.oct_batch_destroy(self.tns._c(), self.id))
            self.id = None
