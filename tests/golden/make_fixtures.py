"""Extract the reference's data fixtures into small, self-contained .npz files.

Reads DATA files only (vectors, graphs, ground truth) from the reference
checkout; no reference code is imported or executed. Run once in a container
where the reference is mounted:

    python tests/golden/make_fixtures.py /root/reference

Outputs (committed):
  sift10k.npz  base (10000x128 u8), query (100x128 u8), gt100 (100x100 i32,
               siftsmall_groundtruth.ivecs), diskann_gt_ids/dists (first 100
               rows of data/sift10k/sift10k_groundtruth.bin, 10 per row),
               parth_adj (10000x32 i32, parth/data/sift_initial_graph.bin),
               custom_adj (5000x32 i32, parth/data/custom_sift_initial_graph.bin)
  g1_graph.npz adj (10000x64 i32, -1 padded past each node's degree) and
               degree of vamana/vout/vamana.out, the G1 GPU build of
               data/sift10k/sift10k_randomgraph.bin (R=64, alpha 1.5, medoid 5000)
  toy2d.npz    toy_vecs/toy_adj (20 nodes) and vis_vecs/vis_adj (50 nodes),
               parth/data/{toy,vis}_initial_graph.bin (float 2-D data)
  randgraph.npz adj (10000x64 i32) and degree of data/sift10k/sift10k_randomgraph.bin,
               the random initial graph the reference's FreshVamana build starts from
               (772-byte records; vectors = SIFT10K base)
  file_digests.json  size and SHA-256 of the reference's graph files, so a
               writer of those formats can be checked byte for byte without the
               files themselves
  diskann_graph.npz    adj (10000x64 i32) of data/sift10k/sift10k_index_disk.bin, the
               DiskANN-built SIFT10K graph in BANG's disk layout (medoid 3732 in its
               _metadata.bin; every node has degree 64, neighbours in ascending order)
  parth_workloads.npz  the parth engine's workload files parth/data/
               {sift,custom_sift,toy,vis}_workload.bin as <name>_types, <name>_ids
               (int32) and <name>_vecs (float32 [ops x dim])
  bang_sift10k.npz     gt_ids / gt_dists: the first 400 rows (10 per row) of
               data/sift10k/sift10k_groundtruth.bin, the rows BANG's recall of its
               400-query run reads (bang_exact/src/main.cu:87-137)
  bang_expected.json   what BANG's SIFT10K run printed (bang_exact/output.txt):
               per run the rounds ("iterations = "), total_recall and total
               computations, plus L, k and the entry point of that build
               (parANN.cuh:40-41,122-128) and numQueries from the run's first line
  parth_metrics.json   per workload: the reference engine's own result
               (parth/vis_outputs/<dir>/metrics.json: qps, recall_10), the batch
               size and MAX_N of its config (parth/config_<name>.json) and the
               beam its file was written at (the last of engine.cu:584's sweep)

SIFT10K is integer-valued (0..180); storing it as uint8 is lossless, which is
asserted below.
"""
import hashlib
import json
import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_fvecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    d = raw[0]
    return raw.reshape(-1, d + 1)[:, 1:].view(np.float32).copy()


def read_ivecs(path):
    raw = np.fromfile(path, dtype=np.int32)
    d = raw[0]
    return raw.reshape(-1, d + 1)[:, 1:].copy()


def read_parth_graph(path):
    """parth graph file: int32 N, DIM, MAX_DEGREE, float[N*DIM], int32[N*MAX_DEGREE]
    (format written by parth/src/build_graph.cu:361-369)."""
    with open(path, "rb") as f:
        n, d, md = struct.unpack("iii", f.read(12))
        vecs = np.frombuffer(f.read(n * d * 4), dtype=np.float32).reshape(n, d).copy()
        adj = np.frombuffer(f.read(n * md * 4), dtype=np.int32).reshape(n, md).copy()
    return vecs, adj


def read_workload(path, dim):
    """parth workload file: records { int32 type; int32 id; float vec[dim] }
    (written by parth/tools/gen_sift_workload.py:80-86, read by
    parth/src/engine.cu:359-372 up to the first incomplete record)."""
    raw = np.fromfile(path, dtype=np.uint8)
    rec = 8 + 4 * dim
    m = len(raw) // rec
    arr = raw[: m * rec].reshape(m, rec)
    types = arr[:, :4].copy().view(np.int32).ravel()
    ids = arr[:, 4:8].copy().view(np.int32).ravel()
    vecs = arr[:, 8:].copy().view(np.float32)
    return types, ids, vecs


def to_u8(a):
    assert np.all(a == np.round(a)) and a.min() >= 0 and a.max() <= 255
    return a.astype(np.uint8)


def main(ref):
    base = read_fvecs(os.path.join(ref, "siftsmall/siftsmall_base.fvecs"))
    query = read_fvecs(os.path.join(ref, "siftsmall/siftsmall_query.fvecs"))
    gt100 = read_ivecs(os.path.join(ref, "siftsmall/siftsmall_groundtruth.ivecs"))
    pv, padj = read_parth_graph(os.path.join(ref, "parth/data/sift_initial_graph.bin"))
    cv, cadj = read_parth_graph(os.path.join(ref, "parth/data/custom_sift_initial_graph.bin"))
    assert np.array_equal(pv, base), "parth sift graph vectors differ from SIFT10K base"
    assert np.array_equal(cv, base[:5000]), "custom graph vectors differ from base[:5000]"
    # DiskANN truthset: uint32 N, K, ids[N*K], float dists[N*K]
    tpath = os.path.join(ref, "data/sift10k/sift10k_groundtruth.bin")
    hdr = np.fromfile(tpath, dtype=np.uint32, count=2)
    n_t, k_t = int(hdr[0]), int(hdr[1])
    body = np.fromfile(tpath, dtype=np.uint32, offset=8)
    t_ids = body[: n_t * k_t].reshape(n_t, k_t)[:100].astype(np.int32)
    t_d = body[n_t * k_t: 2 * n_t * k_t].view(np.float32).reshape(n_t, k_t)[:100].copy()
    np.savez_compressed(
        os.path.join(HERE, "sift10k.npz"),
        base=to_u8(base), query=to_u8(query), gt100=gt100.astype(np.int32),
        diskann_gt_ids=t_ids, diskann_gt_dists=t_d,
        parth_adj=padj, custom_adj=cadj)
    # FreshVamana 772-byte records: float vec[128], uint32 degree, uint32 nbr[64]
    rec = np.fromfile(os.path.join(ref, "vamana/vout/vamana.out"), dtype=np.uint8).reshape(-1, 772)
    assert np.array_equal(rec[:, :512].copy().view(np.float32), base)
    deg = rec[:, 512:516].copy().view(np.uint32).ravel().astype(np.int32)
    nbr = rec[:, 516:].copy().view(np.uint32).astype(np.int64)
    g1 = np.where(np.arange(64)[None, :] < deg[:, None], nbr, -1).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "g1_graph.npz"), adj=g1, degree=deg)
    rrec = np.fromfile(os.path.join(ref, "data/sift10k/sift10k_randomgraph.bin"),
                       dtype=np.uint8).reshape(-1, 772)
    assert np.array_equal(rrec[:, :512].copy().view(np.float32), base)
    rdeg = rrec[:, 512:516].copy().view(np.uint32).ravel().astype(np.int32)
    rnbr = rrec[:, 516:].copy().view(np.uint32).astype(np.int64)
    radj = np.where(np.arange(64)[None, :] < rdeg[:, None], rnbr, -1).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "randgraph.npz"), adj=radj, degree=rdeg)
    digests = {}
    for rel in ("parth/data/sift_initial_graph.bin", "parth/data/custom_sift_initial_graph.bin",
                "parth/data/toy_initial_graph.bin", "parth/data/vis_initial_graph.bin",
                "vamana/vout/vamana.out", "data/sift10k/sift10k_randomgraph.bin",
                "data/sift10k/sift10k_index_disk.bin", "data/sift10k/sift10k_index_disk_metadata.bin",
                "data/sift10k/siftsmall_query.bin", "data/sift10k/siftsmall_groundtruth.bin"):
        raw = open(os.path.join(ref, rel), "rb").read()
        digests[rel] = {"bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest()}
    with open(os.path.join(HERE, "file_digests.json"), "w") as f:
        json.dump(digests, f, indent=1)
    tv, tadj = read_parth_graph(os.path.join(ref, "parth/data/toy_initial_graph.bin"))
    vv, vadj = read_parth_graph(os.path.join(ref, "parth/data/vis_initial_graph.bin"))
    np.savez_compressed(os.path.join(HERE, "toy2d.npz"),
                        toy_vecs=tv, toy_adj=tadj, vis_vecs=vv, vis_adj=vadj)
    drec = np.fromfile(os.path.join(ref, "data/sift10k/sift10k_index_disk.bin"),
                       dtype=np.uint8).reshape(-1, 772)
    assert np.array_equal(drec[:, :512].copy().view(np.float32), base)
    ddeg = drec[:, 512:516].copy().view(np.uint32).ravel()
    assert (ddeg == 64).all()
    np.savez_compressed(os.path.join(HERE, "diskann_graph.npz"),
                        adj=drec[:, 516:].copy().view(np.int32))
    t_all = body[: n_t * k_t].reshape(n_t, k_t)
    d_all = body[n_t * k_t: 2 * n_t * k_t].view(np.float32).reshape(n_t, k_t)
    np.savez_compressed(os.path.join(HERE, "bang_sift10k.npz"),
                        gt_ids=t_all[:400].astype(np.int32), gt_dists=d_all[:400].copy())
    text = open(os.path.join(ref, "bang_exact/output.txt")).read()
    runs, cur = [], {}
    for line in text.splitlines():
        if line.startswith("iterations = "):
            cur = {"rounds": int(line.split("=")[1])}
        elif line.startswith("total_recall = "):
            cur["total_recall"] = float(line.split()[2])
            cur["num_queries"] = int(line.split()[5])
        elif "total computations:" in line:
            cur["computations"] = int(line.rsplit(":", 1)[1])
            cur["L"] = int(line.split()[0])
            cur["recall_pct"] = float(line.split()[1])
            runs.append(cur)
    first = text.splitlines()[0]
    assert first.startswith("[medoidID=")
    entry = int(first.split("medoidID=")[1].split(",")[0])
    with open(os.path.join(HERE, "bang_expected.json"), "w") as f:
        json.dump({"source": "bang_exact/output.txt", "entry": entry, "k": 10,
                   "dim": 128, "bloom_cells": 399887, "runs": runs}, f, indent=1)
    wl, metrics = {}, {}
    for name, dim, vis_dir in (("sift", 128, "sift"), ("custom_sift", 128, "custom_sift"),
                               ("toy", 2, "toy"), ("vis", 2, "step_by_step")):
        t, i, v = read_workload(os.path.join(ref, f"parth/data/{name}_workload.bin"), dim)
        wl[f"{name}_types"], wl[f"{name}_ids"], wl[f"{name}_vecs"] = t, i, v
        with open(os.path.join(ref, f"parth/config_{name}.json")) as f:
            cfg = json.load(f)
        with open(os.path.join(ref, f"parth/vis_outputs/{vis_dir}/metrics.json")) as f:
            m = json.load(f)
        metrics[name] = {"qps": m["qps"], "recall_10": m["recall_10"],
                         "batch_size": cfg["execution"]["batch_size"],
                         "max_capacity": cfg["max_capacity"], "dim": dim, "L": 100,
                         "metrics_file": f"parth/vis_outputs/{vis_dir}/metrics.json"}
    np.savez_compressed(os.path.join(HERE, "parth_workloads.npz"), **wl)
    with open(os.path.join(HERE, "parth_metrics.json"), "w") as f:
        json.dump(metrics, f, indent=1)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
