"""Digest of the oracle's FRESH build at the full 1M x 128 search workload.

Runs the CPU oracle (oracle/oracle.cpp, test infrastructure) on the same
seeded data the benchmark uses -- 1,000,000 x 128 integer-valued points of the
256-component Gaussian mixture (seed 1234, stream 0), R = 64, L_build = 128,
two passes with alpha 1.0 then 1.2, P = 600, M = 500 -- and writes what the
GPU build must reproduce: the medoid, the SHA-256 of the adjacency (int32
little-endian, -1 padded, [n x R]), degree statistics, and the SHA-256 of the
oracle's FP32 search results (ids and distances, k = 10, L = 60) for the 10,000
queries of stream 1. The build takes tens of minutes on a few CPU cores, so it
runs once here and the numbers are committed:

    python tests/golden/make_c2_digest.py tests/golden/c2_fresh_1m.json
"""
import hashlib
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import orc  # noqa: E402

datasets = importlib.import_module("cs5013-project_amd.datasets")

N, D, K, R, L_BUILD, ALPHAS, L_SEARCH, NQ = 1_000_000, 128, 256, 64, 128, (1.0, 1.2), 60, 10_000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main(out):
    t0 = time.time()
    base = datasets.integer_mixture(N, D, k=K, sigma=20.0, seed=1234, stream=0)
    qs = datasets.integer_mixture(NQ, D, k=K, sigma=20.0, seed=1234, stream=1)
    print(f"data {time.time() - t0:.1f} s; oracle build on {orc.num_threads()} threads", flush=True)
    t = time.time()
    adj, med = orc.fresh_build(base, R=R, L=L_BUILD, alphas=ALPHAS, P=600, M=500)
    build_s = time.time() - t
    print(f"build {build_s:.1f} s, medoid {med}", flush=True)
    deg = (adj >= 0).sum(1)
    t = time.time()
    ids, dists, hops, nd = orc.fresh_search(base, adj, qs, med, L=L_SEARCH, out_n=10)
    print(f"search {time.time() - t:.1f} s", flush=True)
    res = {
        "workload": f"FRESH build {N}x{D}, {K}-component integer mixture (seed 1234, stream 0), "
                    f"R={R}, L_build={L_BUILD}, alphas={list(ALPHAS)}, P=600, M=500",
        "n": N, "d": D, "R": R, "L_build": L_BUILD, "alphas": list(ALPHAS),
        "medoid": int(med),
        "adjacency_sha256": sha(adj.astype("<i4")),
        "degree_mean": float(deg.mean()), "degree_min": int(deg.min()),
        "degree_max": int(deg.max()),
        "search": {"queries": f"{NQ} points of stream 1", "k": 10, "L": L_SEARCH,
                   "ids_sha256": sha(ids.astype("<i4")),
                   "dists_sha256": sha(dists.astype("<f4")),
                   "hops_sum": int(hops.sum()), "dist_evals_sum": int(nd.sum())},
        "oracle_build_s": round(build_s, 1), "oracle_threads": orc.num_threads(),
    }
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
        f.write("\n")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(ROOT, "tests", "golden",
                                                            "c2_fresh_1m.json"))
