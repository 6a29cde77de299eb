"""Benchmark of the run! + build_edges hot path on one MI355X.

Workload (one step = one pass of the hot path over one batch): the 3D SPH
block, 1,000,000 FP32 points on a 100^3 lattice (h = 0.01, U(+-0.1h) jitter),
radius 2h; every step adds N(0, (0.01h)^2) per coordinate (the snapshots are
generated on the device before the timed region), then update_point_set +
run! (refit mode: the almost-sorted path of time stepping) + build_edges on
the self pair.  The metric is edges/s of build_edges with a bit-exact edge
set, learned single pass (the headline `value`) against the exact two-pass
path, both measured in the same process on the same snapshots.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "octopus.jl_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "edges/s, learned vs exact-count build_edges, 1M pts 3D, bit-exact edges"
# HBM bytes per launch of the dominant kernels, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a 512 MiB copy (the file records the commands)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_sph1m.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (the SPH configuration's 100 time steps)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--mode", default="both", choices=["both", "exact", "learned"],
                    help="both: learned (headline value) and exact, in that order")
    ap.add_argument("--cpu-steps", type=int, default=30, help="steps of the CPU oracle baseline (0 = skip)")
    ap.add_argument("--side", type=int, default=100, help="lattice side (100 -> 1M points)")
    ap.add_argument("--no-profile", action="store_true", help="time without the per-phase HIP events")
    ap.add_argument("--private-stream", action="store_true",
                    help="run the library on its own stream, joined to torch's current stream per call "
                         "(default: on torch's current stream)")
    ap.add_argument("--predictor", default="sph_lattice",
                    help="shipped predictor for the learned mode ('none': plan-free single pass)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before the workload, copy a MIB-MiB device buffer 3 times (a known-size kernel for "
                         "scaling rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-refit", action="store_true",
                    help="full sort every step (default: refit mode, the almost-sorted path of time stepping)")
    ap.add_argument("--predict-always", action="store_true",
                    help="run the predictor on every step (default: first call and after overflows / recounts)")
    ap.add_argument("--workload", default="sph", choices=["sph", "batched", "clustered"],
                    help="sph: the headline 1M-point SPH block; batched: 256 uneven sets in one replayed graph; "
                         "clustered: 4M clustered points (the learned predictor's own distribution)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched: keep --leaf-size / --run-leaves / --run-chunk")
    ap.add_argument("--leaf-size", type=int, default=32)
    ap.add_argument("--run-leaves", type=int, default=6)
    ap.add_argument("--run-chunk", type=int, default=4, help="single-pass runs per block ticket (1..4)")
    return ap.parse_args()


def run_mode(a, mode, snaps, r):
    """Times a.steps steps of one build_edges mode on a fresh TNS."""
    import torch

    import octohip

    tns = octohip.TNS(private_stream=a.private_stream)
    octohip.set_search_radius(tns, float(r))
    octohip.set_refit_mode(tns, not a.no_refit)
    sid = octohip.add_point_set(tns, snaps[0])
    octohip.set_active_search(tns, sid, sid)
    predictor = None
    if mode == "learned" and a.predictor != "none":
        if os.path.exists(os.path.join(octohip.predictor.SHIPPED, a.predictor + ".f32")):
            octohip.predictor.load(tns, a.predictor)
            predictor = a.predictor
            if a.predict_always:
                octohip.predictor.set_policy(tns, True)
    used = [0]
    refits = [0]

    def step(k):
        octohip.update_point_set(tns, sid, snaps[k])
        octohip.run(tns)
        e = octohip.build_edges(tns, sid, sid, mode=mode)
        # (after build_edges: the refit's outcome is back on the host by then)
        refits[0] += int(octohip.build_info(tns, sid)["refit_used"])
        if predictor:
            # (known when the call returns: no wait for the stream)
            used[0] += int(octohip.predictor.edge_stats(tns, sid, sid, tensors=False, wait=False)["used_predictor"])
        return e.senders.shape[0]

    step(0)
    first_call = octohip.predictor.edge_stats(tns, sid, sid)
    for k in range(1, a.warmup):
        step(k)
    used[0] = 0
    refits[0] = 0
    # the timed loop times only the dominant kernel's phase with events (two
    # per step); the full per-phase breakdown comes from a separate pass
    kphase = "single" if mode == "learned" else "fill"
    torch.cuda.synchronize()
    if not a.no_profile:
        octohip.profile(tns, True, phases=[kphase])
        octohip.profile_read(tns, reset=True)
    edges = 0
    t0 = time.perf_counter()
    for k in range(a.warmup, a.warmup + a.steps):
        edges += step(k)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = octohip.predictor.edge_stats(tns, sid, sid)
    kprof = octohip.profile_read(tns, reset=True)[kphase] if not a.no_profile else (0.0, 0)
    used_timed, refits_timed = used[0], refits[0]
    prof = {}
    if not a.no_profile:
        octohip.profile(tns, True)
        n_prof = min(a.steps, 10)
        for k in range(a.warmup, a.warmup + n_prof):
            step(k)
        torch.cuda.synchronize()
        prof = octohip.profile_read(tns, reset=True)
        octohip.profile(tns, False)
    sec = t1 - t0
    res = {
        "mode": mode,
        "value": edges / sec,
        "ms_per_step": 1e3 * sec / a.steps,
        "edges_per_step": edges / a.steps,
        "breakdown_ms_per_step": {k: v[0] / min(a.steps, 10) for k, v in prof.items() if v[1]},
        "timed_kernel": {"phase": kphase, "ms_total": kprof[0], "launches": kprof[1]},
        "predictor": predictor,
        "first_call": {k: first_call[k] for k in ("single_pass", "used_predictor", "attempts")},
        "timed_steps_using_predictor": used_timed,
        "timed_steps_refit": refits_timed,
        "last_step": {k: st[k] for k in ("single_pass", "used_predictor", "attempts", "recounted_leaves",
                                         "recorded_leaves", "predicted_total", "n_edges", "n_leaves")},
    }
    del tns
    return res


def bench_batched(a):
    """256 independent 3D sets, sizes log-uniform in [1000, 65536], each a
    4-component Gaussian mixture with its radius set for ~20 neighbours; one
    step = in-place coordinate update + run! + build_edges of every set as one
    replayed graph.  Beside it: the same sets one at a time through the
    single-set API (exact mode), and the C oracle on the host."""
    import numpy as np
    import torch

    import octohip
    from octohip import datasets

    sets, radii = datasets.batched_sets()
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    n = int(off[-1])
    dev = torch.device("cuda", 0)
    base = torch.from_numpy(np.concatenate(sets)).to(dev)
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    total_steps = a.warmup + a.steps
    snaps = [(base.double() + torch.randn(base.shape, generator=g, device=dev, dtype=torch.float64) * 1e-4)
             .float().contiguous() for _ in range(total_steps)]
    coords = snaps[0].clone()
    tns = octohip.TNS()
    b = octohip.Batch(tns, coords, off, radii)
    b.config(a.leaf_size, a.run_leaves, True, a.run_chunk)
    tuned = None
    if not a.no_autotune:
        tuned = b.autotune(reps=3)

    last = {}

    def step(k):
        coords.copy_(snaps[k])
        res = b.run()
        last["recounted_leaves"] = res.recounted_leaves
        last["recorded_leaves"] = res.recorded_leaves
        return res.senders.shape[0], res.replayed

    for k in range(a.warmup):
        step(k)
    torch.cuda.synchronize()
    edges, replays = 0, 0
    t0 = time.perf_counter()
    for k in range(a.warmup, total_steps):
        e, rp = step(k)
        edges += e
        replays += int(rp)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    # the same sets one at a time through the single-set API (one TNS per
    # set, since a TNS has one radius), after one untimed warm-up step each
    host_sets = [snaps[-1][off[i]:off[i + 1]].contiguous() for i in range(len(sets))]
    seq = []
    for i, s in enumerate(host_sets):
        t = octohip.TNS()
        octohip.set_search_radius(t, float(radii[i]))
        sid = octohip.add_point_set(t, s)
        octohip.set_active_search(t, sid, sid)
        octohip.run(t)
        octohip.build_edges(t, sid, sid)
        seq.append((t, sid))
    seq_edges = 0
    torch.cuda.synchronize()
    s0 = time.perf_counter()
    for i, (t, sid) in enumerate(seq):
        octohip.update_point_set(t, sid, host_sets[i])
        octohip.run(t)
        seq_edges += octohip.build_edges(t, sid, sid).senders.shape[0]
    torch.cuda.synchronize()
    seq_sec = time.perf_counter() - s0
    cpu = None
    if a.cpu_steps > 0:
        import oracle

        oracle.build()
        host = [s.cpu().numpy() for s in host_sets]
        c0 = time.perf_counter()
        ce = 0
        for i, hp in enumerate(host):
            ce += oracle.build_edges(oracle.run(hp, float(radii[i])), hp).n_edges
        c1 = time.perf_counter()
        cpu = {"value": ce / (c1 - c0), "unit": "edges/s", "cores": oracle.max_threads(), "kind": "port",
               "sample": f"one step of all 256 sets (run! + build_edges per set, C oracle), {c1 - c0:.1f} s"}
    print(json.dumps({
        "metric": "edges/s, batched 256 sets, one replayed graph per step",
        "value": edges / sec, "unit": "edges/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * sec / a.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "256 sets, sizes log-uniform in [1000, 65536], 4-component Gaussian mixtures, "
                               "radius per set for ~20 neighbours; in-place update + run! + build_edges per step",
                   "sets": len(sets), "points": n, "edges_per_step": edges / a.steps,
                   "leaf_size": tuned["best"][0] if tuned else a.leaf_size,
                   "run_leaves": tuned["best"][1] if tuned else a.run_leaves,
                   "run_chunk": tuned["best"][2] if tuned else a.run_chunk, "timed_replays": replays},
        "autotune_ms_per_step": ({f"leaf{k[0]}_run{k[1]}_chunk{k[2]}": v for k, v in tuned["ms"].items()}
                                 if tuned else None),
        "sets_per_s": len(sets) * a.steps / sec,
        "last_step": dict(last),
        "sequential_single_set_api": {"ms_per_step": 1e3 * seq_sec, "edges_per_s": seq_edges / seq_sec},
        "cpu_baseline": cpu,
    }))


def bench_clustered(a):
    """4M points, 90% from a 16-cluster Gaussian mixture (sigma 0.05) and 10%
    uniform, radius 0.01 (about 155 neighbours per point, 6e8 edges): the
    learned predictor's own distribution.  A step is run! (refit of unmoved
    points) + build_edges; learned (with the shipped clustered predictor) and
    exact on the same points."""
    import torch

    import octohip
    from octohip import datasets

    pts, r = datasets.clustered(seed=4)
    n = pts.shape[0]
    coords = torch.from_numpy(pts).cuda()
    steps = min(a.steps, 5)
    res = {}
    for mode in ("learned", "exact"):
        tns = octohip.TNS()
        octohip.set_search_radius(tns, float(r))
        octohip.set_refit_mode(tns, True)
        sid = octohip.add_point_set(tns, coords)
        octohip.set_active_search(tns, sid, sid)
        pred = None
        if mode == "learned" and os.path.exists(os.path.join(octohip.predictor.SHIPPED, "clustered4m.f32")):
            octohip.predictor.load(tns, "clustered4m")
            pred = "clustered4m"
        first = None
        for k in range(2):
            octohip.update_point_set(tns, sid, coords)
            octohip.run(tns)
            octohip.build_edges(tns, sid, sid, mode=mode)
            if k == 0:
                first = octohip.predictor.edge_stats(tns, sid, sid, tensors=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        edges = 0
        for k in range(steps):
            octohip.update_point_set(tns, sid, coords)
            octohip.run(tns)
            edges += octohip.build_edges(tns, sid, sid, mode=mode).senders.shape[0]
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        st = octohip.predictor.edge_stats(tns, sid, sid, tensors=False)
        octohip.profile(tns, True)
        octohip.profile_read(tns, reset=True)
        octohip.update_point_set(tns, sid, coords)
        octohip.run(tns)
        octohip.build_edges(tns, sid, sid, mode=mode)
        torch.cuda.synchronize()
        prof = octohip.profile_read(tns, reset=True)
        octohip.profile(tns, False)
        res[mode] = {"value": edges / sec, "ms_per_step": 1e3 * sec / steps, "edges_per_step": edges / steps,
                     "breakdown_ms_per_step": {k: v[0] for k, v in prof.items() if v[1]},
                     "predictor": pred,
                     "first_call": {k: first[k] for k in ("single_pass", "used_predictor", "attempts",
                                                          "predicted_total", "n_edges")},
                     "last_step": {k: st[k] for k in ("single_pass", "used_predictor", "attempts",
                                                      "recounted_leaves", "recorded_leaves", "heavy_leaves",
                                                      "n_leaves")}}
        del tns
        torch.cuda.synchronize()
    head = res["learned"]
    print(json.dumps({
        "metric": "edges/s, learned vs exact build_edges, 4M clustered points",
        "value": head["value"], "unit": "edges/s", "n_gpus": 1, "steps": steps, "warmup": 2,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "4,000,000 points: 90% from a 16-cluster Gaussian mixture (sigma 0.05), 10% uniform, "
                               "radius 0.01; run! + build_edges per step on unmoved points",
                   "points": n, "edges_per_step": head["edges_per_step"]},
        "learned_over_exact": head["value"] / res["exact"]["value"],
        "modes": res,
    }))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark measures one GPU (one process, one device); --gpus above 1 is not supported",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "batched":
        bench_batched(a)
        return
    if a.workload == "clustered":
        bench_clustered(a)
        return
    import torch

    from octohip import datasets

    if a.calibrate_copy > 0:
        src = torch.ones(a.calibrate_copy * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    h = 0.01
    pts, r = datasets.sph_lattice(side=a.side, h=h)
    n = pts.shape[0]
    dev = torch.device("cuda", 0)
    base = torch.from_numpy(pts).to(dev)
    total_steps = a.warmup + a.steps
    snaps = datasets.sph_snapshots(base, total_steps, h=h, seed=2)
    torch.cuda.synchronize()

    modes = ["learned", "exact"] if a.mode == "both" else [a.mode]
    results = {m: run_mode(a, m, snaps, r) for m in modes}
    head = results[modes[0]]

    # dominant kernel: the single pass (learned) or the fill (exact)
    kname = head["timed_kernel"]["phase"]
    kms = head["timed_kernel"]["ms_total"]  # HIP events around its launches in the timed loop
    kcalls = head["timed_kernel"]["launches"]
    k_avg_s = (kms / max(kcalls, 1)) / 1e3
    e_avg = head["edges_per_step"]
    alg_bytes = 24.0 * e_avg + 12.0 * n  # 24 B per emitted edge + 12 B per query point read
    achieved = alg_bytes / k_avg_s / 1e9 if k_avg_s > 0 else 0.0
    kernel_symbol = "k_edges_single" if kname == "single" else "k_edges<EDGE_FILL>"
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        tr = json.load(open(TRAFFIC_FILE))
        if kernel_symbol in tr.get("kernels", {}):
            traffic = tr["kernels"][kernel_symbol]["bytes_per_launch"]
    step_bytes = 16.0 * n + 24.0 * e_avg + 12.0 * n  # coords + perm write, edge output, query reads

    cpu = None
    if a.cpu_steps > 0:
        import oracle

        oracle.build()
        threads = oracle.max_threads()
        host = [snaps[k].cpu().numpy() for k in range(min(a.cpu_steps, len(snaps)))]
        ce = 0
        c0 = time.perf_counter()
        for k in range(a.cpu_steps):
            hp = host[k % len(host)]
            t = oracle.run(hp, float(r))
            ce += oracle.build_edges(t, hp).n_edges
        c1 = time.perf_counter()
        cpu = {"value": ce / (c1 - c0), "unit": "edges/s", "cores": threads, "kind": "port",
               "sample": f"{a.cpu_steps} steps of the same 1M-point SPH workload (run! + build_edges, C oracle, "
                         f"OpenMP {threads} threads), {c1 - c0:.1f} s"}

    out = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"SPH block: {n} FP32 points on a {a.side}^3 lattice, h=0.01, jitter 0.1h, "
                               f"radius 2h, N(0,(0.01h)^2) step displacement; update + run! (refit mode) + "
                               f"build_edges per step",
                   "mode": head["mode"], "points": n, "edges_per_step": e_avg, "radius": float(r)},
        "modes": results,
        "points_per_s": n / (head["ms_per_step"] / 1e3),
        "hbm_fraction_step": (step_bytes / (head["ms_per_step"] / 1e3)) / 1e9 / HBM_PEAK_GBS,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel_symbol, "kernel_avg_ms": k_avg_s * 1e3, "alg_bytes_per_launch": alg_bytes},
        "cpu_baseline": cpu,
    }
    if "learned" in results and "exact" in results:
        out["learned_over_exact"] = results["learned"]["value"] / results["exact"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
