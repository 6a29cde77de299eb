"""Train a per-leaf hit-count predictor with the library's own HIP kernels
and write its parameter block.

    python -m octohip.train_predictor --workload clustered --out clustered4m.f32
    python -m octohip.train_predictor --workload sph --out sph_lattice.f32

Workloads (seeded instances of one distribution; the held-out seeds follow
the training seeds):
  clustered  4,000,000 points, 90% from a 16-cluster Gaussian mixture
             (sigma 0.05) and 10% uniform, radius 0.01
  sph        the 100^3 SPH lattice (h = 0.01, jitter 0.1h), radius 2h, each
             instance advanced 10 * (seed % 11) steps of N(0, (0.01h)^2), so
             consecutive seeds cover the configuration's 100 time steps
The exact per-leaf hit counts come from the count pass (no edges are
written).  Prints one JSON line with the training loss and the held-out
error against the best constant predictor.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def instance(workload: str, seed: int, n: int):
    from . import datasets

    if workload == "clustered":
        pts, r = datasets.clustered(n=n, seed=seed)
        if n != 4_000_000:
            r = np.float32(float(r) * (4_000_000 / n) ** (1 / 3))
        return pts, float(r)
    if workload == "sph":
        side = round(n ** (1 / 3))
        pts, r = datasets.sph_lattice(seed=seed, side=side)
        for k in range(10 * (seed % 11)):
            pts = datasets.sph_step(pts, k, seed=seed)
        return pts, float(r)
    raise SystemExit(f"unknown workload {workload}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["clustered", "sph"], required=True)
    ap.add_argument("--instances", type=int, default=None, help="training instances (clustered 50, sph 10)")
    ap.add_argument("--held-out", type=int, default=2)
    ap.add_argument("--n", type=int, default=None, help="points per instance (clustered 4M, sph 1M)")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed0", type=int, default=1000)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    n = a.n or (4_000_000 if a.workload == "clustered" else 1_000_000)
    n_train = a.instances or (50 if a.workload == "clustered" else 10)

    import torch

    import octohip

    tns = octohip.TNS()
    octohip.set_search_radius(tns, 1.0)
    sid = None
    rows = 0
    consts = []
    t0 = time.perf_counter()
    for k in range(n_train):
        pts, r = instance(a.workload, a.seed0 + k, n)
        octohip.set_search_radius(tns, r)
        c = torch.from_numpy(pts).cuda()
        if sid is None:
            sid = octohip.add_point_set(tns, c)
            octohip.set_active_search(tns, sid, sid)
        else:
            octohip.resize_point_set(tns, sid, c)
        octohip.run(tns)
        rows += octohip.predictor.collect(tns, sid, sid)
        ex = octohip.predictor.edge_stats(tns, sid, sid)["exact"].cpu().numpy().astype(np.float64)
        consts.append(np.log1p(ex))
        print(f"instance {k + 1}/{n_train}: {ex.shape[0]} leaves, {int(ex.sum())} hits, "
              f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    const = float(np.mean(np.concatenate(consts)))
    loss0 = octohip.predictor.train(tns, steps=1, batch=a.batch, lr=a.lr, seed=a.seed0, init=True)
    done = 1
    loss = loss0
    while done < a.steps:
        k = min(500, a.steps - done)
        # step decay: the last third at a tenth of the rate
        lr = a.lr if done < 2 * a.steps // 3 else a.lr / 10
        loss = octohip.predictor.train(tns, steps=k, batch=a.batch, lr=lr, seed=a.seed0 + done)
        done += k
        print(f"step {done}: batch loss {loss:.4f}", file=sys.stderr, flush=True)
    params = octohip.predictor.save(tns, a.out)

    # held-out instances through the learned single pass
    ev = []
    for k in range(a.held_out):
        pts, r = instance(a.workload, a.seed0 + n_train + k, n)
        t2 = octohip.TNS()
        octohip.set_search_radius(t2, r)
        c = torch.from_numpy(pts).cuda()
        s2 = octohip.add_point_set(t2, c)
        octohip.set_active_search(t2, s2, s2)
        octohip.run(t2)
        octohip.predictor.load(t2, params)
        octohip.predictor.collect(t2, s2, s2)
        st = octohip.predictor.edge_stats(t2, s2, s2)
        X = st["features"].clone()
        ex = st["exact"].cpu().numpy().astype(np.float64)
        y = octohip.predictor.forward(t2, torch.from_numpy(params), X).cpu().numpy().astype(np.float64)
        pred = np.maximum(np.expm1(y), 0)
        ev.append({"leaves": int(ex.shape[0]), "hits": int(ex.sum()),
                   "mae_log1p": float(np.mean(np.abs(y - np.log1p(ex)))),
                   "mae_log1p_constant": float(np.mean(np.abs(const - np.log1p(ex)))),
                   "total_rel_error": float((pred.sum() - ex.sum()) / ex.sum()),
                   "under_predicted_leaves": float(np.mean(pred < ex))})
        del t2
    print(json.dumps({"workload": a.workload, "points": n, "instances": n_train, "rows": rows, "steps": a.steps,
                      "batch": a.batch, "loss_first": loss0, "loss_last": loss, "held_out": ev,
                      "seconds": time.perf_counter() - t0, "out": os.path.abspath(a.out)}))


if __name__ == "__main__":
    main()
