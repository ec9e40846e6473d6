"""The 1M x 128 search workload at full size: the GPU FRESH build (R = 64,
L_build = 128, alpha 1.0 then 1.2) must equal the oracle's build of the same
seeded data bit for bit, and its searches the oracle's searches. The oracle
takes about 17 minutes on 8 cores for that build, so its result is committed as
digests (tests/golden/c2_fresh_1m.json, written by
tests/golden/make_c2_digest.py): medoid, SHA-256 of the adjacency, and SHA-256
of the FP32 search results (k = 10, L = 60) of the 10,000 stream-1 queries."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_gpu_full_size_build_and_search_equal_oracle_digests(vmn, gpu):
    import torch
    g = json.load(open(os.path.join(GOLDEN, "c2_fresh_1m.json")))
    n, d = g["n"], g["d"]
    print(f"generating {n} x {d}", flush=True)
    base = vmn.datasets.integer_mixture(n, d, k=256, sigma=20.0, seed=1234, stream=0)
    qs = vmn.datasets.integer_mixture(10_000, d, k=256, sigma=20.0, seed=1234, stream=1)
    base_t = torch.from_numpy(base).to(gpu)
    torch.cuda.synchronize()
    print("building on the GPU", flush=True)
    idx = vmn.Vamana(base_t, mode="fresh", R=g["R"], L_build=g["L_build"],
                     alphas=tuple(g["alphas"]), low_precision="bf16")
    assert idx.build_precision == "bf16"
    assert idx.medoid == g["medoid"]
    adj = idx.adjacency
    deg = (adj >= 0).sum(1)
    assert (int(deg.min()), int(deg.max())) == (g["degree_min"], g["degree_max"])
    assert abs(float(deg.mean()) - g["degree_mean"]) < 1e-9
    assert sha(adj.astype("<i4")) == g["adjacency_sha256"]
    s = g["search"]
    for prec, rr in (("fp32", 0), ("bf16", 40)):
        ids, dists = idx.searchPoints(qs, k=10, L=s["L"], prec=prec, rerank=rr)
        assert sha(ids.astype("<i4")) == s["ids_sha256"], prec
        assert sha(dists.astype("<f4")) == s["dists_sha256"], prec
        hops, nd = idx.search_stats()
        assert (hops, nd) == (s["hops_sum"], s["dist_evals_sum"]), prec
    gt, _ = vmn.groundtruth(base_t, torch.from_numpy(qs).to(gpu), 10)
    rec = vmn.recall_at_k(ids, gt.cpu().numpy(), 10)
    print(f"recall@10 at L={s['L']}: {rec:.4f}", flush=True)
    assert rec >= 0.95
