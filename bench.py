"""Benchmark: batched Vamana search QPS at recall@10 >= 0.95 on 1M x 128 (one
MI355X), plus build points/s.

A step is one batched search of 10,000 queries (1M x 128) in one persistent
launch, inputs resident in HBM. The build and the recall sweep run first
(untimed by the step clock, timed separately), then W warmup and K timed steps
at the smallest beam width whose recall@10 reaches the target. The K steps are
K batches issued round-robin over `--contexts` search contexts (streams,
default 2), each replaying its own captured graph over its own 10,000-query
set, so the drain of one batch overlaps the start of the next; `value` is the
queries of the K batches over the wall time of the K steps. Each form's launch
shape (waves per CU, beam-list capacity) is tuned on that same round-robin
stream of batches. The
same form timed one batch at a time on one context (no overlap) is reported as
`qps_serial`, and the roofline uses that per-launch kernel time.

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
    ap.add_argument("--components", type=int, default=256,
                    help="mixture components of the synthetic data (the 10k x 128 config uses 64)")
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--L-build", type=int, default=128)
    ap.add_argument("--alphas", default="1.0,1.2")
    ap.add_argument("--sweep", default="10,20,40,50,60,70,80,100,120,150,200")
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--rerank", type=int, default=40,
                    help="FP32 re-rank depth of the BF16 form (the top 4k = 40 candidates)")
    ap.add_argument("--rerank-fp8", type=int, default=40,
                    help="FP32 re-rank depth of the FP8 form (deeper: its distance error is larger)")
    ap.add_argument("--workload", default="search",
                    choices=["search", "streaming", "serving", "copy", "gather", "engine", "bang"],
                    help="search: the headline metric; streaming: 10 rounds of 5%% deletes + "
                         "5%% inserts with consolidation, recall@10 per round; serving: the "
                         "continuous-batching request loop; copy: a 1 GiB device-to-device copy "
                         "(a calibration kernel for the HBM traffic counters); gather: the "
                         "library's random-row gather probe over a 4 GiB table, 128 / 256 / "
                         "512-byte rows, each row read once (the calibration for random row "
                         "gathers); engine: the "
                         "parth engine's replay of its own SIFT workload file (batches of 50 "
                         "mixed operations), the reference's qps / recall_10 metric; bang: "
                         "BANG's SIFT10K run (400 queries, L 40) on its exact search schedule")
    ap.add_argument("--contexts", type=int, default=2,
                    help="search contexts (streams) the timed batches are issued round-robin over")
    ap.add_argument("--engine-reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=10)
    ap.add_argument("--L-eval", type=int, default=60)
    ap.add_argument("--serve-n", type=int, default=10_000_000)
    ap.add_argument("--serve-rate", type=float, default=50_000.0, help="offered queries/s")
    ap.add_argument("--serve-seconds", type=float, default=8.0)
    ap.add_argument("--serve-L", type=int, default=64)
    ap.add_argument("--serve-consolidate", type=int, default=2_000,
                    help="serving: consolidate the tombstones after this many deletes")
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the sharded "
              "form (queries split over replicated indexes, or per-shard graphs merged by one "
              "all-gather of k pairs per query) is not run here.", file=sys.stderr)
        sys.exit(2)
    import torch
    vmn = importlib.import_module("cs5013-project_amd")
    dev = torch.device("cuda:0")
    if args.workload == "streaming":
        return streaming(args, vmn, torch, dev)
    if args.workload == "bang":
        b = bang_section(vmn, args)
        print(json.dumps({"metric": "BANG exact search QPS, SIFT10K, 400 queries, L 40",
                          "value": b["qps"], "unit": "queries/s", "n_gpus": 1,
                          "steps": max(args.steps, 20), "warmup": args.warmup,
                          "ms_per_step": b["ms"], "higher_is_better": True, "scaling": "weak",
                          "vs_baseline": b["vs_reference"], "dtype": "fp32",
                          "data": "SIFT10K base and queries, DiskANN graph (tests/golden)",
                          "config": b}), flush=True)
        return
    if args.workload == "serving":
        return serving(args, vmn, torch, dev)
    if args.workload == "copy":
        return copy_calibration(args, torch, dev)
    if args.workload == "gather":
        return gather_calibration(args, vmn, torch, dev)
    if args.workload == "engine":
        return engine_replay(args, vmn, torch)
    n, nq, d = args.n, args.nq, 128
    alphas = tuple(float(x) for x in args.alphas.split(","))
    kc = args.components
    log(f"generating {n}x{d} base + {nq} queries ({kc}-component integer mixture)")
    base = vmn.datasets.integer_mixture(n, d, k=kc, sigma=20.0, seed=1234, stream=0)
    qs = vmn.datasets.integer_mixture(nq, d, k=kc, sigma=20.0, seed=1234, stream=1)
    qs2 = vmn.datasets.integer_mixture(nq, d, k=kc, sigma=20.0, seed=1234, stream=2)
    base_t = torch.from_numpy(base).to(dev)
    q_t = torch.from_numpy(qs).to(dev)
    q2_t = torch.from_numpy(qs2).to(dev)
    torch.cuda.synchronize()

    log(f"building ({args.mode}, R={args.R}, L={args.L_build}, alphas={alphas})")
    t = time.time()
    idx = vmn.Vamana(base_t, mode=args.mode, R=args.R, L_build=args.L_build, alphas=alphas,
                     low_precision="both")
    torch.cuda.synchronize()
    build_s = time.time() - t
    bst = idx.build_stats()
    build_gbs = bst["algorithmic_bytes"] / build_s / 1e9
    log(f"build {build_s:.2f} s = {n / build_s:.0f} points/s, medoid {idx.medoid}, "
        f"{bst['algorithmic_bytes'] / 1e9:.1f} GB algorithmic = {build_gbs:,.0f} GB/s")

    log("FP64 ground truth")
    t = time.time()
    gt_ids, _ = vmn.groundtruth(base_t, q_t, 10)
    torch.cuda.synchronize()
    gt_s = time.time() - t
    gt_fb = vmn.groundtruth_fallbacks()
    gt = gt_ids.cpu().numpy()
    gt2 = vmn.groundtruth(base_t, q2_t, 10)[0].cpu().numpy()
    log(f"ground truth (2 x {nq} queries) {time.time() - t:.2f} s")

    stream = torch.cuda.ExternalStream(idx.stream)
    ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
    dists = torch.empty((nq, 10), dtype=torch.float32, device=dev)
    forms = {"fp32": ("fp32", 0), "bf16+rerank": ("bf16", args.rerank),
             "fp8+rerank": ("fp8", args.rerank_fp8)}

    def run(form, L, reps, sync_first=True):
        prec, rr = forms[form]
        rr = min(rr, L)
        if sync_first:
            idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=True, prec=prec,
                             rerank=rr)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            idx.searchPoints(q_t, k=10, L=L, ids_out=ids, dists_out=dists, sync=False, prec=prec,
                             rerank=rr)
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps

    sweeps, chosen = {}, {}
    for form in forms:
        sw = []
        for L in [int(x) for x in args.sweep.split(",")]:
            ms = run(form, L, 2)
            rec = vmn.recall_at_k(ids.cpu().numpy(), gt, 10)
            hops, nd = idx.search_stats()
            sw.append({"L": L, "recall@10": round(rec, 4), "qps": round(nq / (ms / 1e3), 1),
                       "ms": round(ms, 3), "hops_per_q": round(hops / nq, 1),
                       "dists_per_q": round(nd / nq, 1)})
            log(f"{form:12s} L={L:4d} recall@10={rec:.4f} {nq / (ms / 1e3):,.0f} QPS ({ms:.2f} ms)")
        ok = [x for x in sw if x["recall@10"] >= args.target]
        chosen[form] = (ok[0] if ok else max(sw, key=lambda x: x["recall@10"]), bool(ok))
        sweeps[form] = sw

    # timed steps: every form at its chosen L. The launch shape is autotuned;
    # each context captures its batch as a HIP graph and replays it.
    timed = {}
    nctx = max(2, args.contexts)
    ctx = [vmn.SearchContext(idx) for _ in range(nctx)]
    cst = [torch.cuda.ExternalStream(c.stream) for c in ctx]
    ids2 = torch.empty((nq, 10), dtype=torch.int32, device=dev)
    dists2 = torch.empty((nq, 10), dtype=torch.float32, device=dev)
    # contexts past the first two search the two query sets again, into
    # buffers of their own
    more = [(torch.empty((nq, 10), dtype=torch.int32, device=dev),
             torch.empty((nq, 10), dtype=torch.float32, device=dev)) for _ in range(nctx - 2)]
    for form in forms:
        L = chosen[form][0]["L"]
        prec, rr = forms[form]
        rr = min(rr, L)
        tune = vmn.tuning.launch_shape(idx, q_t, k=10, L=L, prec=prec, rerank=rr, reps=12,
                                       overlap=True, contexts=nctx, queries_alt=q2_t)
        ctx[0].capture(q_t, ids, dists, k=10, L=L, prec=prec, rerank=rr)
        ctx[1].capture(q2_t, ids2, dists2, k=10, L=L, prec=prec, rerank=rr)
        for c_, (i_, d_) in enumerate(more):
            ctx[2 + c_].capture(q_t if c_ % 2 == 0 else q2_t, i_, d_, k=10, L=L, prec=prec,
                                rerank=rr)
        # one batch at a time on one context: per-launch kernel time
        for _ in range(args.warmup):
            ctx[0].replay()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t = time.time()
        e0.record(cst[0])
        for _ in range(args.steps):
            ctx[0].replay()
        e1.record(cst[0])
        e1.synchronize()
        wall_serial = time.time() - t
        kern_ms = e0.elapsed_time(e1) / args.steps
        # the timed steps: batches go round-robin over the contexts; HIP
        # events on each context's stream bracket every launch, so each
        # launch's duration is measured where it runs (overlapping the
        # neighbouring launches on the other stream)
        for s_ in range(max(args.warmup, nctx)):
            ctx[s_ % nctx].replay()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        t = time.time()
        for s_ in range(args.steps):
            evs[s_][0].record(cst[s_ % nctx])
            ctx[s_ % nctx].replay()
            evs[s_][1].record(cst[s_ % nctx])
        torch.cuda.synchronize()
        wall = time.time() - t
        launch_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
        rec = vmn.recall_at_k(ids.cpu().numpy(), gt, 10)
        rec2 = vmn.recall_at_k(ids2.cpu().numpy(), gt2, 10)
        kernel = idx.search_kernel_name(k=10, L=L, prec=prec, list_cap=tune["list_cap"],
                                        tiles=tune["tiles"])
        idx.searchPoints(q_t, k=10, L=L, prec=prec, rerank=rr)  # same traversal, for the counters
        hops, nd = idx.search_stats()
        # per query: the FP32 query row, each expansion's adjacency row (FP8:
        # and the row norms stored beside it, and in FRESH mode the bloom
        # cells), each measured candidate's row, the re-ranked FP32 rows, the
        # k results
        row_b = {"fp32": d * 4, "bf16": d * 2, "fp8": d}[prec]
        adj_b = 4 * args.R * ((3 if args.mode == "fresh" else 2) if prec == "fp8" else 1)
        bytes_launch = (d * 4) * nq + hops * adj_b + nd * row_b + nq * rr * d * 4 \
            + nq * 10 * 8
        timed[form] = {"L": L, "recall@10": round(rec, 4), "recall@10_set2": round(rec2, 4),
                       "target_met": chosen[form][1],
                       "qps": round(nq * args.steps / wall, 1), "ms_per_step": wall * 1e3 / args.steps,
                       "qps_serial": round(nq * args.steps / wall_serial, 1),
                       "launch_ms": launch_ms, "bytes_per_launch": int(bytes_launch),
                       "achieved_GBs": bytes_launch / (launch_ms / 1e3) / 1e9,
                       "kernel_ms_serial": kern_ms,
                       "achieved_GBs_serial": bytes_launch / (kern_ms / 1e3) / 1e9,
                       "launch": {"waves_per_cu": tune["waves_per_cu"], "list_cap": tune["list_cap"],
                                  "tiles": tune["tiles"], "query_groups": tune["groups"]},
                       "kernel": kernel, "rerank": rr,
                       "traffic": measured_traffic(form, n, nq, L, kernel, rr),
                       "counters": measured_counters(form, n, nq, L, kernel, rr, launch_ms),
                       "hops_per_query": round(hops / nq, 1), "dists_per_query": round(nd / nq, 1)}
        log(f"timed {form:12s} L={L} {timed[form]['qps']:,.0f} QPS (serial {timed[form]['qps_serial']:,.0f})"
            f" recall {rec:.4f}/{rec2:.4f} launch {launch_ms:.3f} ms overlapped, {kern_ms:.3f} ms"
            f" serial, {timed[form]['achieved_GBs']:.0f} GB/s"
            f" (launch {tune['waves_per_cu']} waves/CU, cap {tune['list_cap']}, {tune['tiles']} tiles,"
            f" {tune['groups']} query groups)")
    for c in ctx:
        c.close()

    # PCIe-inclusive rate of the headline form: host query buffer in, host ids
    # and distances out (vmn_search with host pointers stages both ways)
    Lh = chosen["bf16+rerank"][0]["L"]
    rrh = min(args.rerank, Lh)
    idx.set_launch(timed["bf16+rerank"]["launch"]["waves_per_cu"], timed["bf16+rerank"]["launch"]["list_cap"])
    idx.set_tiles(timed["bf16+rerank"]["launch"]["tiles"])
    if timed["bf16+rerank"]["launch"]["query_groups"] is not None:
        idx.set_query_groups(timed["bf16+rerank"]["launch"]["query_groups"])
    idx.searchPoints(qs, k=10, L=Lh, prec="bf16", rerank=rrh)
    t = time.time()
    for _ in range(args.steps):
        idx.searchPoints(qs, k=10, L=Lh, prec="bf16", rerank=rrh)
    host_ms = (time.time() - t) * 1e3 / args.steps
    log(f"bf16+rerank with host buffers: {host_ms:.3f} ms per batch ({nq / host_ms * 1e3:,.0f} QPS)")

    head = timed["bf16+rerank"]
    # distance error of the low-precision rows against FP64 on a sample of pairs
    err = {}
    rng = np.random.default_rng(0)
    pi = rng.integers(0, n, 4096)
    qi = rng.integers(0, nq, 4096)
    exact = ((base[pi].astype(np.float64) - qs[qi].astype(np.float64)) ** 2).sum(1)
    for prec in ("bf16", "fp8"):
        rows = idx.lowp_rows(prec)
        approx = ((rows[pi].astype(np.float64) - qs[qi].astype(np.float64)) ** 2).sum(1)
        rel = (approx - exact) / np.maximum(exact, 1.0)
        err[prec] = {"rms_rel": float(np.sqrt(np.mean(rel ** 2))), "max_rel": float(np.abs(rel).max())}
        del rows

    bang = bang_section(vmn, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(idx, base, qs, gt, chosen["fp32"][0]["L"], args.cpu_seconds, args.mode)

    out = {
        "metric": METRIC, "value": head["qps"], "unit": "queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16",
        "data": f"synthetic: seeded {kc}-component Gaussian mixture, integer-valued in [0,255]",
        "config": {
            "workload": f"Vamana search {n}x{d}, BF16 rows on MFMA + FP32 re-rank of "
                        f"{args.rerank}, {nq} queries per batch, k=10",
            "mode": args.mode, "n": n, "nq": nq, "R": args.R, "L_build": args.L_build,
            "alphas": list(alphas), "L_search": head["L"], "recall@10": head["recall@10"],
            "recall_target_met": head["target_met"], "build_s": round(build_s, 3),
            "build_points_per_s": round(n / build_s, 1),
            # the exact FP64 10-NN of the first query set: its FP32 MFMA filter
            # does 2 n d multiply-adds per query (the kept rows' FP64 distances
            # and the per-query selection are a few percent more work)
            "groundtruth": {"bound": "mfma", "ms": round(gt_s * 1e3, 2),
                            "achieved": round(2.0 * n * d * nq / gt_s / 1e12, 1),
                            "peak": 157.3, "unit": "TFLOP/s (FP32 MFMA)",
                            "frac": round(2.0 * n * d * nq / gt_s / 1e12 / 157.3, 4),
                            "fallback_queries": gt_fb,
                            "time": "host wall clock around vmn_groundtruth, synchronous"},
            "build_roofline": {"bound": "hbm", "bytes": int(bst["algorithmic_bytes"]),
                               "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": round(build_gbs / HBM_PEAK_GBS, 4),
                               "work": {k: v for k, v in bst.items()
                                        if k != "algorithmic_bytes"},
                               "time": "host wall clock around vmn_build, device synchronised"},
            "fp32": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in timed["fp32"].items()},
            "fp8+rerank": {k: (round(v, 4) if isinstance(v, float) else v)
                           for k, v in timed["fp8+rerank"].items()},
            "bf16_vs_fp32_qps": round(head["qps"] / timed["fp32"]["qps"], 3),
            "bf16_vs_fp32_qps_serial": round(head["qps_serial"] / timed["fp32"]["qps_serial"], 3),
            "qps_serial": head["qps_serial"], "recall@10_set2": head["recall@10_set2"],
            "qps_host_buffers": round(nq / host_ms * 1e3, 1),
            "batches": f"K batches of {nq:,} queries issued round-robin over {nctx} search "
                       "contexts (query sets: mixture streams 1 and 2, alternating)",
            "distance_error_vs_fp64": err, "sweep": sweeps, "bang": bang},
        "roofline": roofline(head),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


BANG_REF_QPS = 65271.94  # bang_exact/output.txt:46, best warm run (A100 80 GB)


def bang_section(vmn, args):
    """BANG's SIFT10K run (bang_exact/output.txt): 400 queries over the DiskANN
    graph tests/golden/diskann_graph.npz from entry 123 at L 40, k 10 (the
    query file's 100 rows and 300 zero rows), on vmn_bang_search. Kernel time
    from HIP events on the index stream around K back-to-back launches; also
    the same schedule on 10,000 queries (the 100 tiled). Bytes: every distance
    computation reads one 512-byte FP32 row, every round one 256-byte adjacency
    row; SIFT10K's 5 MB of rows stay in cache, so this is not an HBM figure."""
    import ctypes
    import torch
    V = importlib.import_module("cs5013-project_amd.vamana")
    gold = os.path.join(ROOT, "tests", "golden")
    s = np.load(os.path.join(gold, "sift10k.npz"))
    t = np.load(os.path.join(gold, "bang_sift10k.npz"))
    adj = np.load(os.path.join(gold, "diskann_graph.npz"))["adj"]
    base = s["base"].astype(np.float32)
    q = np.zeros((400, 128), np.float32)
    q[:100] = s["query"].astype(np.float32)
    dev = torch.device("cuda:0")
    bidx = vmn.Vamana.from_graph(base, adj, medoid=3732, mode="fresh", R=64)
    stream = torch.cuda.ExternalStream(bidx.stream)

    def timed(qh, reps):
        nq = qh.shape[0]
        qd = torch.from_numpy(qh).to(dev)
        ids = torch.empty((nq, 10), dtype=torch.int32, device=dev)
        d = torch.empty((nq, 10), dtype=torch.float32, device=dev)
        st = torch.empty((nq, 2), dtype=torch.int32, device=dev)
        ptr = [ctypes.c_void_p(x.data_ptr()) for x in (qd, ids, d, st)]

        def launch():
            V._check(V.lib().vmn_bang_search(bidx._h, ptr[0], nq, 10, 40, 123, ptr[1], ptr[2],
                                             ptr[3], 1))
        for _ in range(args.warmup):
            launch()
        bidx.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(reps):
            launch()
        e1.record(stream)
        e1.synchronize()
        return e0.elapsed_time(e1) / reps, ids.cpu().numpy(), st.cpu().numpy()

    ms, ids, st = timed(q, max(args.steps, 20))
    comps, rounds = int(st[:, 0].sum()), int(st[:, 1].sum())
    total_recall = int(round(vmn.recall_at_k(ids, t["gt_ids"], 10, gt_dists=t["gt_dists"]) * 4000))
    ms10, _, st10 = timed(np.tile(q[:100], (100, 1)), max(args.steps, 10))
    gbs = (comps * 512 + rounds * 256) / ms / 1e6
    log(f"BANG SIFT10K 400 queries L40: {ms:.4f} ms ({400 / ms * 1e3:,.0f} QPS, "
        f"{400 / ms * 1e3 / BANG_REF_QPS:.1f}x BANG's best A100 run); {comps} computations, "
        f"{int(st[:, 1].max())} rounds, total recall {total_recall}; 10,000 queries "
        f"{ms10:.4f} ms ({1e4 / ms10 * 1e3:,.0f} QPS)")
    bidx.close()
    return {"workload": "BANG exact search, SIFT10K DiskANN graph, 400 queries (100 + 300 zero "
                        "rows), L 40, k 10, entry 123",
            "kernel": "bang_kernel<128, 64, 8>", "ms": round(ms, 4), "qps": round(400 / ms * 1e3, 1),
            "reference_qps": BANG_REF_QPS,
            "vs_reference": round(400 / ms * 1e3 / BANG_REF_QPS, 2),
            "computations": comps, "rounds": int(st[:, 1].max()), "total_recall": total_recall,
            "matches_printed_run": (comps, int(st[:, 1].max()), total_recall) == (641853, 47, 2190),
            "row_bytes_per_s_GB": round(gbs, 1),
            "qps_10k_queries": round(1e4 / ms10 * 1e3, 1),
            "computations_10k": int(st10[:, 0].sum()),
            "time": "HIP events on the index stream around back-to-back launches"}


def roofline(e):
    """Roofline of one form's search kernel over the timed (overlapped) steps.
    achieved = algorithmic bytes per launch / mean launch duration (HIP events
    on the launch's own stream, neighbouring launches overlapping it);
    traffic = counter HBM bytes per launch from the committed rocprofv3 passes
    (when they cover this workload), and with it the HBM rate of the timed
    steps, traffic / ms_per_step. The one-batch-at-a-time figures are under
    "serial". Counter ratios measured for this kernel (L2 hit rate, MFMA busy)
    come from the same committed profile."""
    r = {"bound": "hbm", "achieved": round(e["achieved_GBs"], 1), "peak": HBM_PEAK_GBS,
         "unit": "GB/s", "frac": round(e["achieved_GBs"] / HBM_PEAK_GBS, 4),
         "traffic": e["traffic"], "kernel": e["kernel"], "launch": e["launch"],
         "launch_ms": round(e["launch_ms"], 4), "bytes_per_launch": e["bytes_per_launch"]}
    if e["traffic"]:
        hbm = e["traffic"] / (e["ms_per_step"] / 1e3) / 1e9
        r["hbm_GBs_per_step"] = round(hbm, 1)
        r["hbm_frac_per_step"] = round(hbm / HBM_PEAK_GBS, 4)
    r.update(e.get("counters") or {})
    r["serial"] = {"kernel_ms": round(e["kernel_ms_serial"], 4),
                   "achieved": round(e["achieved_GBs_serial"], 1),
                   "frac": round(e["achieved_GBs_serial"] / HBM_PEAK_GBS, 4)}
    return r


def copy_calibration(args, torch, dev):
    """1 GiB device-to-device copy (torch copy_, one runtime copy kernel per
    step): its FETCH_SIZE / WRITE_SIZE per dispatch against the 2^30 bytes it
    must read and write gives the scale that turns those counters into bytes."""
    nbytes = 1 << 30
    src = torch.full((nbytes // 4,), 1.0, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    print(json.dumps({"metric": "device copy bandwidth (read + write)", "value": round(2 * nbytes / ms / 1e6, 1),
                      "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
                      "vs_baseline": None, "dtype": "f32", "data": "synthetic",
                      "config": {"workload": "copy 1 GiB device to device", "bytes": nbytes}}),
          flush=True)


def gather_calibration(args, vmn, torch, dev):
    """Random-row gathers of known size (vmn_gather_probe): 4 GiB table,
    rows of 128, 256 and 512 bytes, 2^22 distinct rows per dispatch drawn from
    a seeded permutation. FETCH_SIZE per dispatch against n_reads x row_bytes
    + 4 n_reads gives the counter scale for the search kernels' access
    pattern; the line reports the rows/s and bytes/s the probe reaches."""
    import ctypes
    L = vmn.lib()
    table_bytes = 4 << 30
    table = torch.zeros(table_bytes // 4, dtype=torch.int32, device=dev)
    sink = torch.empty(1 << 22, dtype=torch.int32, device=dev)
    g = torch.Generator(device="cpu").manual_seed(77)
    n_reads = 1 << 22
    out = {}
    for rb in (128, 256, 512):
        rows = table_bytes // rb
        ids = torch.randperm(rows, generator=g)[:n_reads].to(torch.int32).to(dev)
        torch.cuda.synchronize()
        for split in (0, 1):
            ms = ctypes.c_double()
            times = []
            for s_ in range(args.warmup + args.steps):
                rc = L.vmn_gather_probe(ctypes.c_void_p(table.data_ptr()), rows, rb, split,
                                        ctypes.c_void_p(ids.data_ptr()), n_reads,
                                        ctypes.c_void_p(sink.data_ptr()), sink.numel(), 0,
                                        ctypes.byref(ms))
                if rc != 0:
                    raise vmn.VamanaError(rc, L.vmn_last_error().decode())
                if s_ >= args.warmup:
                    times.append(ms.value)
            t = float(np.median(times))
            known = n_reads * rb + 4 * n_reads
            key = str(rb) + ("_split64" if split else "")
            out[key] = {"n_reads": n_reads, "known_bytes": known, "ms": round(t, 4),
                        "GB_per_s": round(known / t / 1e6, 1),
                        "G_rows_per_s": round(n_reads / t / 1e6, 2)}
            log(f"gather probe {key} B rows: {t:.3f} ms, {known / t / 1e6:,.0f} GB/s")
    print(json.dumps({"metric": "random-row gather bandwidth (4 GiB table, rows read once)",
                      "value": out["256"]["GB_per_s"], "unit": "GB/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": out["256"]["ms"], "higher_is_better": True,
                      "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
                      "config": {"workload": "vmn_gather_probe: 2^22 distinct random rows of "
                                             "128 / 256 / 512 B from a 4 GiB table, each row in "
                                             "one load instruction, or 64 B of 16 rows per "
                                             "instruction (_split64)",
                                 "rows": out}}), flush=True)


def log_uniform_batches(rng, total, lo=1000, hi=50000):
    """Split `total` into batch sizes drawn log-uniform in [lo, hi]."""
    out = []
    while total > 0:
        b = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        b = min(b, total)
        out.append(b)
        total -= b
    return out


def streaming(args, vmn, torch, dev):
    """FreshVamana streaming: start from the n x 128 FRESH index of the search
    workload (R=64, L=128, alpha 1.0 then 1.2); updates prune with alpha 1.2;
    each round deletes 5% of the live ids (uniform, seeded) and
    inserts 5% new points from the same mixture, both in batches of
    log-uniform size in [1k, 50k] that arrive interleaved in a seeded order;
    tombstones are consolidated at the end of the round; recall@10 of 10k queries (FP32, L=L_eval) against FP64 truth
    over the live points."""
    n, nq, d = args.n, args.nq, 128
    frac = 0.05
    grow = int(n * frac) * args.rounds
    base = vmn.datasets.integer_mixture(n, d, k=256, sigma=20.0, seed=1234, stream=0)
    extra = vmn.datasets.integer_mixture(grow, d, k=256, sigma=20.0, seed=1234, stream=2)
    qs = vmn.datasets.integer_mixture(nq, d, k=256, sigma=20.0, seed=1234, stream=1)
    allv = torch.from_numpy(np.concatenate([base, extra])).to(dev)
    q_t = torch.from_numpy(qs).to(dev)
    t = time.time()
    idx = vmn.Vamana(allv[:n], mode="fresh", R=args.R, L_build=args.L_build, alphas=(1.0, 1.2),
                     capacity=n + grow)
    torch.cuda.synchronize()
    log(f"streaming: built {n} in {time.time() - t:.2f} s")
    rng = np.random.default_rng(4321)
    deleted = np.zeros(n + grow, dtype=bool)
    size = n
    rounds = []

    def recall_now():
        live = np.flatnonzero(~deleted[:size])
        live_t = torch.from_numpy(live).to(dev)
        gt_local, _ = vmn.groundtruth(allv[live_t].contiguous(), q_t, 10)
        gt = live[gt_local.cpu().numpy()]
        out = []
        for L in (args.L_eval, args.L_build):
            ids, _ = idx.searchPoints(q_t, k=10, L=L)
            found = ids.cpu().numpy()
            assert not np.isin(found[found >= 0], np.flatnonzero(deleted)).any()
            out.append(vmn.recall_at_k(found, gt, 10))
        return out

    r0 = recall_now()
    log(f"round 0: recall@10 {r0[0]:.4f} (L={args.L_eval}), {r0[1]:.4f} (L={args.L_build})")
    for rnd in range(1, args.rounds + 1):
        live = np.flatnonzero(~deleted[:size])
        n_del = int(len(live) * frac)
        kill = rng.choice(live, n_del, replace=False).astype(np.int32)
        n_ins = int(n * frac)
        t_del = t_ins = 0.0
        off = 0
        # the round's delete and insert batches arrive interleaved (seeded order)
        ops = [("del", b) for b in log_uniform_batches(rng, n_del)] + \
              [("ins", b) for b in log_uniform_batches(rng, n_ins)]
        for o in rng.permutation(len(ops)):
            kind, b = ops[o]
            torch.cuda.synchronize()
            t = time.time()
            if kind == "del":
                idx.deleteIds(kill[off:off + b])
                off += b
            else:
                idx.insertPoints(allv[size:size + b])
                size += b
            torch.cuda.synchronize()
            if kind == "del":
                t_del += time.time() - t
            else:
                t_ins += time.time() - t
        deleted[kill] = True
        t = time.time()
        idx.consolidate()
        torch.cuda.synchronize()
        t_con = time.time() - t
        rec = recall_now()
        rounds.append({"round": rnd, f"recall@10_L{args.L_eval}": round(rec[0], 4),
                       f"recall@10_L{args.L_build}": round(rec[1], 4), "deleted": n_del,
                       "inserted": n_ins, "live": int((~deleted[:size]).sum()),
                       "delete_ops_per_s": round(n_del / max(t_del, 1e-9), 1),
                       "insert_points_per_s": round(n_ins / t_ins, 1),
                       "consolidate_s": round(t_con, 3)})
        log(f"round {rnd}: recall@10 {rec[0]:.4f} / {rec[1]:.4f}, insert {n_ins / t_ins:,.0f} pts/s, "
            f"consolidate {t_con:.2f} s, live {rounds[-1]['live']}")
    ins_rate = sum(r["inserted"] for r in rounds) / sum(r["inserted"] / r["insert_points_per_s"]
                                                       for r in rounds)
    out = {"metric": "FreshVamana streaming: insert points/s with recall@10 per round",
           "value": round(ins_rate, 1), "unit": "points/s", "n_gpus": 1, "steps": args.rounds,
           "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic: seeded 256-component Gaussian mixture, integer-valued in [0,255]",
           "config": {"workload": f"FreshVamana {args.rounds} rounds of 5% delete + 5% insert "
                                  f"on {n}x{d}, log-uniform batches [1k, 50k]",
                      "R": args.R, "L_build": args.L_build, "alpha": 1.2,
                      "L_eval": [args.L_eval, args.L_build],
                      "recall_round0": [round(r0[0], 4), round(r0[1], 4)], "rounds": rounds}}
    print(json.dumps(out), flush=True)


def engine_replay(args, vmn, torch):
    """The parth engine's automated sweep (parth/src/engine.cu:578-600) on the
    reference's own SIFT workload (1,000 operations: 727 searches, 183 inserts,
    90 deletes, batches of 50) over its initial graph (SIFT10K, R=32), beams
    10..100, each replayed --engine-reps times from the initial graph. Like the
    reference, qps is searches over the search launches' device time (events
    around each batch's search launch); the reference's engine printed 849.68
    QPS at recall@10 0.9105 for beam 100 (parth/vis_outputs/sift/metrics.json),
    its only published figure for this metric. The CPU baseline is the oracle's
    replay of the same operations (search loops timed)."""
    golden = os.path.join(ROOT, "tests", "golden")
    w = np.load(os.path.join(golden, "parth_workloads.npz"))
    ref = json.load(open(os.path.join(golden, "parth_metrics.json")))["sift"]
    sift = np.load(os.path.join(golden, "sift10k.npz"))
    base = sift["base"].astype(np.float32)
    adj = sift["parth_adj"]
    t, i, ov = w["sift_types"], w["sift_ids"], w["sift_vecs"]
    sweep = []
    for L in (10, 20, 32, 64, 100):
        reps = []
        for _ in range(args.engine_reps):
            idx = vmn.Vamana.from_graph(base, adj, medoid=0, mode="parth",
                                        capacity=ref["max_capacity"])
            reps.append(idx.run_workload(t, i, ov, ref["batch_size"], L))
            idx.close()
        reps.sort(key=lambda r: r["qps"])
        r = reps[len(reps) // 2]
        sweep.append({"L": L, "qps": round(r["qps"], 1), "recall@10": round(r["recall"], 4),
                      "search_ms": round(r["search_ms"], 3),
                      "replay_ops_per_s": round(len(t) / r["wall_s"], 1),
                      "hops_per_search": round(r["expansions"] / r["searches"], 1),
                      "dists_per_search": round(r["dist_evals"] / r["searches"], 1),
                      "batches": r["batches"]})
        log(f"engine L={L}: qps {r['qps']:,.0f}, recall@10 {r['recall']:.4f}, "
            f"search {r['search_ms']:.2f} ms over {r['batches']} batches")
    top = sweep[-1]
    # algorithmic bytes of the search launches at beam 100: per search the
    # query, each expansion's adjacency row and each distance's row
    d, R = 128, int(adj.shape[1])
    per_q = 4 * d + top["hops_per_search"] * 4 * R + top["dists_per_search"] * 4 * d + 40
    n_search = int((t == 2).sum())
    achieved = per_q * n_search / (top["search_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import orc
        st = {}
        orc.engine_run(base, adj, len(base), ref["max_capacity"], t, i, ov, ref["batch_size"],
                       100, stats=st)
        cpu = {"value": round(n_search / st["search_s"], 1), "unit": "queries/s",
               "cores": orc.num_threads(), "kind": "port", "cpu_model": cpu_model(),
               "sample": f"the oracle's replay of the same {len(t)} operations at beam 100; "
                         f"its {n_search} searches timed (OpenMP over each batch's searches)"}
        log(f"engine cpu baseline {cpu['value']:,.0f} searches/s")
    out = {"metric": "parth engine qps (searches / search device time) at beam 100",
           "value": top["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.engine_reps,
           "warmup": 0, "ms_per_step": top["search_ms"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": round(top["qps"] / ref["qps"], 1), "dtype": "fp32",
           "data": "the reference's SIFT10K engine workload (parth/data/sift_workload.bin) "
                   "and initial graph (parth/data/sift_initial_graph.bin)",
           "config": {"workload": "parth engine replay, SIFT10K R=32, 1000 ops in batches of "
                                  "50 (727 searches, 183 inserts, 90 deletes), k=10",
                      "recall@10": top["recall@10"], "reference": {"qps": ref["qps"],
                      "recall_10": ref["recall_10"]}, "sweep": sweep},
           "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                        "traffic": None,
                        "note": "about 36 searches per launch: one wavefront each, so the "
                                "launch is a chain of dependent hops (latency), not bandwidth"},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def serving(args, vmn, torch, dev):
    """Served index: n x 96 unit-normalised 1,024-component mixture, FRESH graph
    (R=64, L=100, alpha 1.0 then 1.2), BF16 rows with FP32 re-rank. Requests
    arrive as a Poisson stream at the offered query rate: 90% searches (batch
    log-uniform 1..256, k=10, L=64), 5% inserts, 5% deletes (batch log-uniform
    1..256); the tombstones are consolidated every --serve-consolidate deletes.
    Latency is arrival to completion on the server clock."""
    n, d, nq_gt = args.serve_n, 96, 10_000
    rng = np.random.default_rng(99)
    log(f"serving: generating {n}x{d} unit mixture")
    base = vmn.datasets.unit_mixture(n, d, k=1024, sigma=0.1, seed=1234, stream=0)
    pool = vmn.datasets.unit_mixture(100_000, d, k=1024, sigma=0.1, seed=1234, stream=1)
    extra = vmn.datasets.unit_mixture(200_000, d, k=1024, sigma=0.1, seed=1234, stream=2)
    base_t = torch.from_numpy(base).to(dev)
    t = time.time()
    idx = vmn.Vamana(base_t, mode="fresh", R=64, L_build=100, alphas=(1.0, 1.2),
                     capacity=n + len(extra), low_precision="bf16")
    torch.cuda.synchronize()
    build_s = time.time() - t
    log(f"serving: built in {build_s:.1f} s ({n / build_s:,.0f} points/s)")
    # request trace
    mean_batch = 255.0 / np.log(256.0)
    req_rate = args.serve_rate / (0.9 * mean_batch)
    t_arr, kinds, sizes = [], [], []
    tcur = 0.0
    while tcur < args.serve_seconds:
        tcur += rng.exponential(1.0 / req_rate)
        r = rng.random()
        kinds.append(vmn.REQ_SEARCH if r < 0.9 else (vmn.REQ_INSERT if r < 0.95 else vmn.REQ_DELETE))
        sizes.append(int(np.exp(rng.uniform(0, np.log(256)))))
        t_arr.append(tcur)
    # launch shape (tile shape, waves per CU) tuned on a batch of the trace's
    # typical size: 64 queries, one batch at a time
    tune = vmn.tuning.launch_shape(idx, torch.from_numpy(pool[:64]).to(dev), k=10, L=args.serve_L,
                                   prec="bf16", rerank=40, reps=20)
    log(f"serving: launch tuned on 64-query batches: {tune['waves_per_cu']} waves/CU, "
        f"{tune['tiles']} tiles, {tune['ms']:.3f} ms")
    srv = vmn.Server(idx, k=10, L=args.serve_L, prec="bf16", rerank=40, max_batch=8192, n_slots=4,
                     consolidate_every=args.serve_consolidate)
    # warm-up, untimed: searches, and inserts / deletes at the largest request
    # size, so every buffer the engine and the update path grow is allocated
    # before the trace starts
    warm = [srv.submit(vmn.REQ_SEARCH, pool[:sz]) for sz in (1, 16, 256, 256)]
    warm += [srv.submit(vmn.REQ_INSERT, extra[len(extra) - 256 * (j + 1):len(extra) - 256 * j])
             for j in range(2)]
    dead = rng.choice(n, 512, replace=False).astype(np.int32)
    warm += [srv.submit(vmn.REQ_DELETE, dead[:256]), srv.submit(vmn.REQ_DELETE, dead[256:])]
    alive = np.ones(n, dtype=bool)
    alive[dead] = False
    for t in warm:
        while srv.poll(t) is None:
            pass
    srv.reset_stats()
    eng0 = srv.stats()
    ins_off = 0
    # deletes draw live base ids uniformly without replacement; the order is
    # drawn once here so the request loop keeps its arrival schedule
    del_order = rng.permutation(np.flatnonzero(alive[:n])).astype(np.int32)
    del_off = 0
    outstanding, lat, done_q, results, lags = [], [], 0, 0, []
    lat_upd = {vmn.REQ_INSERT: [], vmn.REQ_DELETE: []}
    t_first = None
    start = srv.now() + 0.05
    log(f"serving: {len(t_arr)} requests over {args.serve_seconds} s, offered "
        f"{args.serve_rate:,.0f} queries/s")
    for ta, kind, sz in zip(t_arr, kinds, sizes):
        target = start + ta
        while srv.now() < target:
            pass
        if kind == vmn.REQ_SEARCH:
            data = pool[rng.integers(0, len(pool), sz)]
        elif kind == vmn.REQ_INSERT:
            data = extra[ins_off:ins_off + sz]
            ins_off += sz
        else:
            data = del_order[del_off:del_off + sz]
            del_off += sz
            alive[data] = False
        outstanding.append((srv.submit(kind, data, t_arrival=target), kind, sz))
        lags.append(srv.now() - target)
        # reap finished requests from the front
        while outstanding:
            r = srv.poll(outstanding[0][0])
            if r is None:
                break
            _, k_, s_ = outstanding.pop(0)
            if k_ == vmn.REQ_SEARCH:
                lat.append(r["t_done"] - r["t_arrival"])
                done_q += s_
            else:
                lat_upd[k_].append(r["t_done"] - r["t_arrival"])
    t_last_arrival = start + t_arr[-1]
    while outstanding:
        r = srv.poll(outstanding[0][0])
        if r is None:
            continue
        _, k_, s_ = outstanding.pop(0)
        if k_ == vmn.REQ_SEARCH:
            lat.append(r["t_done"] - r["t_arrival"])
            done_q += s_
        else:
            lat_upd[k_].append(r["t_done"] - r["t_arrival"])
    t_end = srv.now()
    eng1 = srv.stats()
    eng = {k: eng1[k] - eng0[k] for k in eng1}
    eng["max_batch_queries"], eng["max_update_s"] = eng1["max_batch_queries"], eng1["max_update_s"]
    srv.close()
    lat_ms = np.array(lat) * 1e3
    thr = done_q / (t_end - start)
    log(f"serving: {done_q} queries, {thr:,.0f} q/s, p50 {np.percentile(lat_ms, 50):.2f} ms, "
        f"p99 {np.percentile(lat_ms, 99):.2f} ms")
    upd = {name: {"n": len(v), "p50_ms": round(float(np.percentile(v, 50)) * 1e3, 3),
                  "p99_ms": round(float(np.percentile(v, 99)) * 1e3, 3)}
           for name, v in (("insert", lat_upd[vmn.REQ_INSERT]), ("delete", lat_upd[vmn.REQ_DELETE]))
           if v}
    engine = {"search_batches": eng["batches"],
              "mean_batch_queries": round(eng["queries"] / max(eng["batches"], 1), 1),
              "max_batch_queries": eng["max_batch_queries"],
              "insert_ms_mean": round(eng["insert_s"] / max(eng["inserts"], 1) * 1e3, 3),
              "delete_ms_mean": round(eng["delete_s"] / max(eng["deletes"], 1) * 1e3, 3),
              "max_update_ms": round(eng["max_update_s"] * 1e3, 3),
              "consolidations": eng["consolidations"],
              "consolidate_every_deletes": args.serve_consolidate,
              "failed_batches": eng["failed_batches"],
              "update_busy_frac": round((eng["insert_s"] + eng["delete_s"]) / (t_end - start), 4),
              "drain_frac": round(eng["drain_s"] / (t_end - start), 4)}
    log(f"serving: engine {engine}; update latency {upd}")
    # recall on a 10k-query FP64 subset of the final index (live points)
    q_gt = pool[:nq_gt]
    deleted = idx.deleted
    live = np.flatnonzero(~deleted)
    # node ids: base, the two warm-up inserts in submission order, the trace's inserts
    m = len(extra)
    allv = torch.from_numpy(np.concatenate([base, extra[m - 256:m], extra[m - 512:m - 256],
                                            extra[:ins_off]]))
    gt_local, _ = vmn.groundtruth(allv[torch.from_numpy(live)].to(dev).contiguous(),
                                  torch.from_numpy(q_gt).to(dev), 10)
    gt = live[gt_local.cpu().numpy()]
    ids, _ = idx.searchPoints(torch.from_numpy(q_gt).to(dev), k=10, L=args.serve_L, prec="bf16",
                              rerank=40)
    rec = vmn.recall_at_k(ids.cpu().numpy(), gt, 10)
    ids32, _ = idx.searchPoints(torch.from_numpy(q_gt).to(dev), k=10, L=args.serve_L)
    rec32 = vmn.recall_at_k(ids32.cpu().numpy(), gt, 10)
    log(f"serving: recall@10 {rec:.4f} (BF16+rerank), {rec32:.4f} (FP32)")
    # the same index's recall at wider beams (context for the configured L)
    curve = {}
    for Lw in (128, 256):
        ids_w, _ = idx.searchPoints(torch.from_numpy(q_gt).to(dev), k=10, L=Lw, prec="bf16",
                                    rerank=40)
        curve[Lw] = round(vmn.recall_at_k(ids_w.cpu().numpy(), gt, 10), 4)
    log(f"serving: recall@10 at L=128 / 256: {curve[128]:.4f} / {curve[256]:.4f}")
    out = {"metric": "served index: queries/s with p50/p99 latency", "value": round(thr, 1),
           "unit": "queries/s", "n_gpus": 1, "steps": len(t_arr), "warmup": len(warm),
           "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "bf16", "data": "synthetic: seeded 1,024-component mixture, unit-normalised",
           "config": {"workload": f"served Vamana {n}x{d}, Poisson {args.serve_rate:,.0f} "
                                  f"offered queries/s, 90% search / 10% insert+delete",
                      "build_s": round(build_s, 2), "requests": len(t_arr),
                      "queries_done": done_q, "inserted": ins_off,
                      "deleted": int((~alive).sum()), "L_search": args.serve_L,
                      "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                      "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                      "p90_ms": round(float(np.percentile(lat_ms, 90)), 3),
                      "p999_ms": round(float(np.percentile(lat_ms, 99.9)), 3),
                      "update_latency": upd, "engine": engine,
                      "launch": {k: tune[k] for k in ("waves_per_cu", "list_cap", "tiles")},
                      "client_submit_lag_ms": {
                          "p99": round(float(np.percentile(lags, 99)) * 1e3, 3),
                          "max": round(float(np.max(lags)) * 1e3, 3)},
                      "recall@10_bf16_rerank": round(rec, 4), "recall@10_fp32": round(rec32, 4),
                      "recall@10_bf16_rerank_at_L": curve}}
    print(json.dumps(out), flush=True)


def _profile_entry(form, n, nq, L, kernel, rerank):
    """The committed counter profile's entry for `form` when it covers this
    exact workload (same n, nq, precision, L and re-rank depth; the beam-list
    capacity and the tile shape of the kernel instance do not change the
    traversal), else None."""
    path = os.path.join(ROOT, "profiles", "search_1m_hbm_traffic.json")
    try:
        with open(path) as f:
            prof = json.load(f)
    except (OSError, ValueError):
        return None
    e = prof.get("forms", {}).get(form)
    # template arguments <dim, capacity, mode, precision, tiles>: compare dim, mode, precision
    same = lambda a, b: [a.split(",")[i].strip() for i in (0, 2, 3)] == \
        [b.split(",")[i].strip() for i in (0, 2, 3)]
    if not e or prof.get("n") != n or prof.get("nq") != nq or e.get("L") != L \
            or e.get("rerank") != rerank or not same(e.get("kernel", ""), kernel):
        return None
    return e


def measured_traffic(form, n, nq, L, kernel, rerank):
    """HBM bytes per launch of the search kernel for `form`, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by what the counters read
    for random row gathers of known size), when that profile covers this exact
    workload; otherwise None."""
    e = _profile_entry(form, n, nq, L, kernel, rerank)
    return None if e is None else int(e["traffic_bytes"])


def measured_counters(form, n, nq, L, kernel, rerank, launch_ms):
    """L2 hit rate and MFMA busy fraction of the search kernel for `form`, and
    the MFMA rate its counted MFMA operations give over the measured launch
    time (against the dense peak of its precision), from the same committed
    profile; {} when the profile does not cover this workload."""
    e = _profile_entry(form, n, nq, L, kernel, rerank)
    if e is None or "l2_hit_rate" not in e:
        return {}
    out = {"l2_hit_rate": e["l2_hit_rate"], "mfma_busy": e["mfma_busy"]}
    peak = {"bf16+rerank": 2500.0, "fp8+rerank": 5000.0}.get(form)
    if peak and e.get("mfma_flop"):
        tf = e["mfma_flop"] / (launch_ms / 1e3) / 1e12
        out.update({"mfma_tflops": round(tf, 1), "mfma_peak_tflops": peak,
                    "mfma_frac": round(tf / peak, 4)})
    return out


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(idx, base, qs, gt, L, seconds, mode):
    """The CPU oracle's search (same graph, same L) on a bounded query sample:
    the query set is cycled until the sample holds about `seconds` of work."""
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
    # the 64-query probe runs at a fraction of the thread count, so it
    # over-estimates the per-query time; size from it, then cap at 40x the set
    ns = int(min(40 * qs.shape[0], max(n0, seconds / max(per_q / threads, 1e-7) / 4)))
    sel = np.arange(ns) % qs.shape[0]
    t = time.time()
    ids, _, _, _ = search(base, adj, qs[sel], med, L=L, out_n=10)
    el = time.time() - t
    log(f"cpu baseline: {ns} queries in {el:.1f} s on {threads} threads")
    return {"value": round(ns / el, 1), "unit": "queries/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"{ns} searches ({el:.1f} s) cycling the {qs.shape[0]} queries, FP32 on the "
                      f"same graph at L={L}, OpenMP over queries",
            "recall@10_sample": round(float(np.mean(
                [len(set(a) & set(b)) / 10 for a, b in zip(ids, gt[sel])])), 4)}


if __name__ == "__main__":
    main()
