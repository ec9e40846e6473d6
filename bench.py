"""Benchmark: batched Vamana search QPS at recall@10 >= 0.95 on 1M x 128 (one
MI355X), plus build points/s.

A step is one batched search of all queries (10,000 at 1M x 128) in one
persistent launch, inputs resident in HBM. The build and the recall sweep run
first (untimed by the step clock, timed separately), then W warmup and K timed
steps at the smallest beam width whose recall@10 reaches the target.

Prints one JSON line (rank 0). Progress goes to stderr.
"""
import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "QPS at recall@10≥0.95, 1M×128 (BF16+rerank vs FP32); build points/s"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench %.1fs]" % (time.time() - T0), *a, file=sys.stderr, flush=True)


T0 = time.time()


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--mode", default="fresh", choices=["parth", "fresh"])
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--L-build", type=int, default=128)
    ap.add_argument("--alphas", default="1.0,1.2")
    ap.add_argument("--sweep", default="10,20,40,60,80,100,120,150,200")
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sharded "
              "design is described in DESIGN.md and is not run here.", file=sys.stderr)
        sys.exit(2)
    import torch
    vmn = importlib.import_module("cs5013-project_amd")
    dev = torch.device("cuda:0")
    n, nq, d = args.n, args.nq, 128
    alphas = tuple(float(x) for x in args.alphas.split(","))
    log(f"generating {n}x{d} base + {nq} queries (256-component integer mixture)")
    base = vmn.datasets.integer_mixture(n, d, k=256, sigma=20.0, seed=1234, stream=0)
    qs = vmn.datasets.integer_mixture(nq, d, k=256, sigma=20.0, seed=1234, stream=1)
    base_t = torch.from_numpy(base).to(dev)
    q_t = torch.from_numpy(qs).to(dev)
    torch.cuda.synchronize()

    log(f"building ({args.mode}, R={args.R}, L={args.L_build}, alphas={alphas})")
    t = time.time()
    idx = vmn.Vamana(base_t, mode=args.mode, R=args.R, L_build=args.L_build, alphas=alphas)
    torch.cuda.synchronize()
    build_s = time.time() - t
    log(f"build {build_s:.2f} s = {n / build_s:.0f} points/s, medoid {idx.medoid}")

    log("FP64 ground truth")
    t = time.time()
    gt_ids, _ = vmn.groundtruth(base_t, q_t, 10)
    gt = gt_ids.cpu().numpy()
    log(f"ground truth {time.time() - t:.2f} s")

    stream = torch.cuda.ExternalStream(idx.stream)
    ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
    dists = torch.empty((nq, 10), dtype=torch.float32, device=dev)

    def timed_search(L, reps):
        idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=True)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=False)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    sweep = []
    for L in [int(x) for x in args.sweep.split(",")]:
        ms = timed_search(L, 2)
        rec = vmn.recall_at_k(ids.cpu().numpy(), gt, 10)
        hops, nd = idx.search_stats()
        sweep.append({"L": L, "recall@10": round(rec, 4), "qps": round(nq / (ms / 1e3), 1),
                      "ms": round(ms, 3), "hops_per_q": round(hops / nq, 1),
                      "dists_per_q": round(nd / nq, 1)})
        log(f"L={L:4d} recall@10={rec:.4f} {nq / (ms / 1e3):,.0f} QPS ({ms:.2f} ms)")
    ok = [s for s in sweep if s["recall@10"] >= args.target]
    chosen = ok[0] if ok else max(sweep, key=lambda s: s["recall@10"])
    L = chosen["L"]

    log(f"timed: warmup {args.warmup}, steps {args.steps} at L={L}")
    for _ in range(args.warmup):
        idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=False)
    idx.synchronize()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t = time.time()
    e0.record(stream)
    for _ in range(args.steps):
        idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=False)
    e1.record(stream)
    e1.synchronize()
    torch.cuda.synchronize()
    wall = time.time() - t
    kern_ms = e0.elapsed_time(e1) / args.steps
    ms_step = wall * 1e3 / args.steps
    qps = nq * args.steps / wall
    rec = vmn.recall_at_k(ids.cpu().numpy(), gt, 10)
    hops, nd = idx.search_stats()
    R = args.R
    bytes_q = (d * 4) * nq + hops * 4 * R + nd * d * 4 + nq * 10 * 8
    achieved = bytes_q / (kern_ms / 1e3) / 1e9
    log(f"{qps:,.0f} QPS, recall@10 {rec:.4f}, kernel {kern_ms:.3f} ms, {achieved:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(idx, base, qs, gt, L, args.cpu_seconds, args.mode)

    out = {
        "metric": METRIC, "value": round(qps, 1), "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: seeded 256-component Gaussian mixture, integer-valued in [0,255]",
        "config": {
            "workload": f"Vamana search {n}x{d} fp32, {nq} queries per batch, k=10",
            "mode": args.mode, "n": n, "nq": nq, "R": R, "L_build": args.L_build,
            "alphas": list(alphas), "L_search": L, "recall@10": round(rec, 4),
            "recall_target_met": bool(ok), "build_s": round(build_s, 3),
            "build_points_per_s": round(n / build_s, 1), "sweep": sweep},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(kern_ms, 4),
                     "bytes_per_launch": int(bytes_q)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def cpu_baseline(idx, base, qs, gt, L, seconds, mode):
    """The CPU oracle's search (same graph, same L) on a bounded query sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    adj = idx.adjacency
    threads = orc.num_threads()
    med = idx.medoid
    search = orc.parth_search if mode == "parth" else orc.fresh_search
    n0 = min(64, qs.shape[0])
    t = time.time()
    search(base, adj, qs[:n0], med, L=L, out_n=10)
    per_q = (time.time() - t) / n0
    ns = int(min(qs.shape[0], max(n0, seconds / max(per_q, 1e-6))))
    t = time.time()
    ids, _, _, _ = search(base, adj, qs[:ns], med, L=L, out_n=10)
    el = time.time() - t
    log(f"cpu baseline: {ns} queries in {el:.1f} s on {threads} threads")
    return {"value": round(ns / el, 1), "unit": "queries/s", "cores": threads, "kind": "port",
            "sample": f"{ns} of the {qs.shape[0]} queries, same graph and L={L}, OpenMP over "
                      f"queries", "recall@10_sample": round(
                          float(np.mean([len(set(a) & set(b)) / 10 for a, b in zip(ids, gt[:ns])])), 4)}


if __name__ == "__main__":
    main()
