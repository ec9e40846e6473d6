"""Search autotuner.

beam_width(): the smallest beam width L whose recall@k against exact ground
truth reaches a target (the reference hard-codes L_g = 199,
vamana/include/constants.cuh:10). launch_shape(): times the batched search
under each persistent-grid size (waves per CU), beam-list capacity template,
tile shape (MFMA tiles gathered per round by the BF16 / FP8 forms) and query
grouping (one group per XCD, or one queue), keeps the fastest, and installs
it on the index. It times either one batch at a time on one search
context, or batches issued round-robin over several contexts, so that each
batch's drain overlaps the next batches' start (the throughput of a stream of
batches).
"""
import time

import numpy as np

from .vamana import SearchContext, recall_at_k


def _time(ctx, q, ids, d, k, L, prec, rerank, reps):
    import torch
    st = torch.cuda.ExternalStream(ctx.stream)
    ctx.search(q, ids, d, k, L, prec, rerank)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        ctx.search(q, ids, d, k, L, prec, rerank)
    e1.record(st)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _time_overlapped(ctxs, qs, outs, k, L, prec, rerank, reps):
    """Milliseconds per batch over `reps` batches issued round-robin over the
    contexts, each replaying its captured search of query set qs[i % len(qs)]
    (host wall clock between device-wide synchronisations)."""
    import torch
    for i, (c, (ids, d)) in enumerate(zip(ctxs, outs)):
        c.capture(qs[i % len(qs)], ids, d, k, L, prec, rerank)
    for r in range(len(ctxs)):
        ctxs[r].replay()
    torch.cuda.synchronize()
    t = time.perf_counter()
    for r in range(reps):
        ctxs[r % len(ctxs)].replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t) * 1e3 / reps


def beam_width(index, queries, gt_ids, target, widths, k=10, prec="fp32", rerank=0):
    """(L, recall) of the smallest width reaching `target`, else the best seen."""
    best = None
    for L in widths:
        ids, _ = index.searchPoints(queries, k=k, L=L, prec=prec, rerank=min(rerank, L))
        r = recall_at_k(ids.cpu().numpy() if hasattr(ids, "cpu") else ids, gt_ids, k)
        if r >= target:
            return L, r
        if best is None or r > best[1]:
            best = (L, r)
    return best


def launch_shape(index, queries, k=10, L=64, prec="fp32", rerank=0,
                 waves=(12, 16, 20, 24, 32), caps=(0, 128, 256), reps=3, overlap=False,
                 contexts=2, groups=(0, 8), queries_alt=None, tiles=(2, 4)):
    """Fastest (waves_per_cu, list_cap, tiles, query groups) for this batch;
    installs it and returns {"waves_per_cu", "list_cap", "tiles", "groups",
    "ms", "table"}. Tile shapes are tried for the BF16 / FP8 forms where the
    4-tile kernels exist (FRESH, dimensions 33 to 128, lists up to 256).
    overlap=True times batches issued round-robin over `contexts` search
    contexts (ms per batch of the stream) instead of one batch at a time; with
    queries_alt the contexts alternate between the two query sets (a stream of
    different batches, as served). Query groups are tried only for a batch
    large enough to be grouped (8 queries per CU)."""
    import torch
    nctx = max(2, contexts) if overlap else 1
    ctxs = [SearchContext(index) for _ in range(nctx)]
    nq = int(queries.shape[0])
    outs = [(torch.empty((nq, k), dtype=torch.int32, device=queries.device),
             torch.empty((nq, k), dtype=torch.float32, device=queries.device))
            for _ in range(nctx)]
    qs = [queries] if queries_alt is None else [queries, queries_alt]
    cus = torch.cuda.get_device_properties(queries.device).multi_processor_count
    if nq < 8 * cus:
        groups = (None,)  # not grouped at this size: the setting is left alone
    rr = min(rerank, L)
    table = []
    for grp in groups:
        if grp is not None:
            index.set_query_groups(grp)
        for cap in caps:
            if cap and cap < L:
                continue
            # a tile shape other than 2 is tried only where its kernels exist
            # (elsewhere the search runs 2 tiles whatever is set)
            two = index.search_kernel_name(k, L, prec, cap, 2)
            for tl in [t for t in tiles if t == 2 or index.search_kernel_name(k, L, prec, cap, t) != two]:
                index.set_tiles(tl)
                for w in waves:
                    index.set_launch(w, cap)
                    if overlap:
                        ms = _time_overlapped(ctxs, qs, outs, k, L, prec, rr, reps)
                    else:
                        ms = _time(ctxs[0], queries, outs[0][0], outs[0][1], k, L, prec, rr, reps)
                    table.append({"waves_per_cu": w, "list_cap": cap, "tiles": tl, "groups": grp,
                                  "ms": round(ms, 4)})
    best = min(table, key=lambda t: t["ms"])
    index.set_tiles(best["tiles"])
    index.set_launch(best["waves_per_cu"], best["list_cap"])
    if best["groups"] is not None:
        index.set_query_groups(best["groups"])
    for c in ctxs:
        c.close()
    return dict(best, table=table)
