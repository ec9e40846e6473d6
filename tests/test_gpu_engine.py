"""GPU parth engine path (vmn_engine_*) against the oracle's replay of
VamanaEngine::process_workload (parth/src/engine.cu:407-568).

* On integer-valued workloads every FP32 distance is exact, so found ids,
  ground truth and the final graph must equal the oracle's bit for bit.
* The reference's own workload files insert SIFT rows plus N(0, 0.1) noise
  (parth/tools/gen_sift_workload.py:127-130), so their distances are not
  exact. PARTH-mode FP32 distances on the GPU are one fused multiply-add chain
  over the dimensions in index order, the reference's distance<D>
  (parth/src/device_funcs.cu:17-49), and the engine's ground truth is the
  reference's FP32 brute force (engine.cu:253-279). So there too every found
  id, every ground-truth id and the final graph equal the oracle's bit for
  bit, and recall@10 is the reference engine's own number (metrics.json) to
  the four decimals it printed."""
import json
import os

import numpy as np
import pytest

from test_oracle_engine import engine_graph

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _recall(found, gt):
    return float(np.mean([len(set(x[x >= 0].tolist()) & set(y.tolist())) / 10
                          for x, y in zip(found, gt)]))


@pytest.mark.parametrize("name", ["sift", "custom_sift", "toy", "vis"])
def test_gpu_engine_reference_workloads(vmn, orc, sift, toy2d, gpu, name):
    w = np.load(os.path.join(GOLDEN, "parth_workloads.npz"))
    m = json.load(open(os.path.join(GOLDEN, "parth_metrics.json")))[name]
    v, a = engine_graph(name, sift, toy2d)
    t, i, ov = w[f"{name}_types"], w[f"{name}_ids"], w[f"{name}_vecs"]
    print(f"engine workload {name}: {len(t)} ops, batch {m['batch_size']}", flush=True)
    idx = vmn.Vamana.from_graph(v, a, medoid=0, mode="parth", capacity=m["max_capacity"])
    r = idx.run_workload(t, i, ov, m["batch_size"], m["L"], results=True)
    dm = 3 if m["dim"] == 2 else 0
    of, ogt, V, A, dl, c0, _ = orc.engine_run(v, a, len(v), m["max_capacity"], t, i, ov,
                                              m["batch_size"], m["L"], dist_mode=dm)
    assert r["searches"] == len(of) and r["inserts"] == int((t == 1).sum())
    assert idx.size == c0
    same_f = (r["found"] == of).all(axis=1).mean()
    same_g = (r["gt"] == ogt).all(axis=1).mean()
    assert same_f == 1.0 and same_g == 1.0, (same_f, same_g)
    assert np.array_equal(idx.adjacency, A[:c0])
    assert np.array_equal(idx.vectors, V[:c0])
    assert np.array_equal(idx.deleted.astype(np.uint8), dl[:c0])
    assert abs(r["recall"] - _recall(r["found"], r["gt"])) < 1e-9
    assert round(r["recall"], 4) == m["recall_10"], (r["recall"], m["recall_10"])
    assert r["qps"] > 0 and r["search_ms"] > 0


def _integer_workload(rng, base, queries, n_ops, id_hi):
    """parth-style op mix (70% search, 20% insert, 10% delete) on integers:
    inserts are base rows plus integer noise, deletes hit ids below id_hi
    (repeats and ids past the node count included)."""
    types = rng.choice([2, 1, 0], size=n_ops, p=[0.7, 0.2, 0.1]).astype(np.int32)
    ids = np.full(n_ops, -1, np.int32)
    vecs = np.zeros((n_ops, base.shape[1]), np.float32)
    for j, ty in enumerate(types):
        if ty == 2:
            vecs[j] = queries[rng.integers(len(queries))]
        elif ty == 1:
            vecs[j] = base[rng.integers(len(base))] + rng.integers(-3, 4, base.shape[1])
        else:
            ids[j] = rng.integers(0, id_hi)
    return types, ids, vecs


@pytest.mark.parametrize("graph,batch,id_hi", [("sift", 50, 10000), ("custom_sift", 50, 10000),
                                               ("sift", 7, 10400)])
def test_gpu_engine_integer_workload_bit_exact(vmn, orc, sift, toy2d, gpu, graph, batch, id_hi):
    v, a = engine_graph(graph, sift, toy2d)
    rng = np.random.default_rng(11 + batch)
    q = sift["query"].astype(np.float32)
    t, i, ov = _integer_workload(rng, sift["base"].astype(np.float32), q, 1500, id_hi)
    print(f"engine integer workload {graph} batch {batch}", flush=True)
    idx = vmn.Vamana.from_graph(v, a, medoid=0, mode="parth", capacity=20000)
    r = idx.run_workload(t, i, ov, batch, 40, results=True)
    of, ogt, V, A, dl, c0, _ = orc.engine_run(v, a, len(v), 20000, t, i, ov, batch, 40)
    assert idx.size == c0
    assert np.array_equal(r["found"], of)
    assert np.array_equal(r["gt"], ogt)
    assert np.array_equal(idx.adjacency, A[:c0])
    assert np.array_equal(idx.vectors, V[:c0])
    assert np.array_equal(idx.deleted.astype(np.uint8), dl[:c0])
    assert abs(r["recall"] - _recall(of, ogt)) < 1e-9
    # the inserted rows are integers too: the searches stayed on exact BF16 rows
    assert idx.build_precision == "bf16"


def test_gpu_engine_real_valued_queries_on_integer_index(vmn, orc, sift, toy2d, gpu):
    """An integer index (its rows exact in BF16) searched with real-valued
    queries (SIFT queries + N(0, 0.1)): the engine must not traverse on BF16
    rows for those queries (a query rounded to BF16 gives other distances), so
    found ids and ground truth equal the oracle's FP32 replay bit for bit.
    Integer queries in the same run may still use the BF16 rows."""
    v, a = engine_graph("sift", sift, toy2d)
    rng = np.random.default_rng(23)
    q = sift["query"].astype(np.float32)
    n_ops = 600
    t = np.full(n_ops, 2, np.int32)
    i = np.full(n_ops, -1, np.int32)
    ov = q[rng.integers(0, len(q), n_ops)].copy()
    noisy = rng.random(n_ops) < 0.5
    ov[noisy] += rng.normal(0.0, 0.1, (int(noisy.sum()), 128)).astype(np.float32)
    print("engine real-valued queries on the SIFT graph", flush=True)
    idx = vmn.Vamana.from_graph(v, a, medoid=0, mode="parth", capacity=20000)
    assert idx.build_precision == "bf16"
    r = idx.run_workload(t, i, ov, 50, 64, results=True)
    of, ogt, _, _, _, _, _ = orc.engine_run(v, a, len(v), 20000, t, i, ov, 50, 64)
    assert np.array_equal(r["found"], of)
    assert np.array_equal(r["gt"], ogt)


def test_gpu_parth_build_real_valued_rows_bit_exact(vmn, orc, sift, gpu):
    """PARTH build over real-valued 128-d rows (SIFT base + N(0, 0.1), 3,000
    points): the search lists and the prune use the reference's sequential
    FP32 distance, so the graph equals the oracle's row for row."""
    rng = np.random.default_rng(31)
    base = sift["base"][:3000].astype(np.float32)
    base = base + rng.normal(0.0, 0.1, base.shape).astype(np.float32)
    q = sift["query"].astype(np.float32) + rng.normal(0.0, 0.1, (100, 128)).astype(np.float32)
    print("parth build on real-valued rows, 3000 x 128", flush=True)
    idx = vmn.Vamana(base, mode="parth", R=32, L_build=64)
    assert idx.build_precision == "fp32"
    oadj, omed = orc.parth_build(base, R=32, L=64, alphas=(1.2, 1.2))
    assert idx.medoid == omed
    assert np.array_equal(idx.adjacency, oadj)
    ids, d = idx.searchPoints(q, k=10, L=40)
    oi, od, _, _ = orc.parth_search(base, oadj, q, omed, L=40, out_n=10)
    assert np.array_equal(ids, oi) and np.array_equal(d, od)


def test_gpu_engine_direct_calls_match_replay(vmn, orc, sift, toy2d, gpu):
    """vmn_engine_delete / vmn_engine_insert called directly give the ids and
    graph of the replay of the same operations as one batch each."""
    v, a = engine_graph("custom_sift", sift, toy2d)
    rng = np.random.default_rng(5)
    base = sift["base"].astype(np.float32)
    kill = np.array([17, 17, 4999, 7000, 123], np.int32)
    new = base[rng.integers(0, 10000, 9)] + rng.integers(-2, 3, (9, 128)).astype(np.float32)
    idx = vmn.Vamana.from_graph(v, a, medoid=0, mode="parth", capacity=20000)
    idx.engine_delete(kill)
    got = idx.engine_insert(new)
    t = np.array([0] * 5 + [1] * 9, np.int32)
    ids = np.concatenate([kill, np.full(9, -1, np.int32)])
    ov = np.concatenate([np.zeros((5, 128), np.float32), new])
    _, _, V, A, dl, c0, asg = orc.engine_run(v, a, len(v), 20000, t, ids, ov, 14, 40)
    assert got.tolist() == asg.tolist()
    # five pushes, nine inserts: 123, 7000, 4999, 17, 17 reused, then 5000 .. 5003
    assert got.tolist() == [123, 7000, 4999, 17, 17, 5000, 5001, 5002, 5003]
    assert idx.size == c0 == 5004
    assert np.array_equal(idx.adjacency, A[:c0])
    with pytest.raises(vmn.VamanaError):
        vmn.Vamana(base[:500], mode="fresh", R=32, L_build=32).engine_delete(kill)


def test_gpu_engine_python_mirror(vmn, sift, gpu, tmp_path):
    """VamanaEngine with the reference's file formats and outputs: graph file
    (build_graph.cu:361-369), workload file, metrics.json."""
    w = np.load(os.path.join(GOLDEN, "parth_workloads.npz"))
    m = json.load(open(os.path.join(GOLDEN, "parth_metrics.json")))["sift"]
    base = sift["base"].astype(np.float32)
    g = tmp_path / "sift_initial_graph.bin"
    g.write_bytes(np.array([10000, 128, 32], np.int32).tobytes() + base.tobytes()
                  + sift["parth_adj"].astype(np.int32).tobytes())
    rec = np.zeros(len(w["sift_types"]), dtype=[("t", "<i4"), ("i", "<i4"), ("v", "<f4", (128,))])
    rec["t"], rec["i"], rec["v"] = w["sift_types"], w["sift_ids"], w["sift_vecs"]
    wf = tmp_path / "sift_workload.bin"
    wf.write_bytes(rec.tobytes())
    eng = vmn.VamanaEngine(dim=128, max_n=20000)
    rows = eng.run_automated_benchmark(str(g), str(wf), 50, str(tmp_path), L_values=(32, 100))
    eng.close()
    assert [r["L"] for r in rows] == [32, 100]
    out = json.load(open(tmp_path / "metrics.json"))
    assert abs(out["recall_10"] - m["recall_10"]) <= 0.002 and out["qps"] > m["qps"]


def test_gpu_cpp_driver_replays_reference_workload(sift, gpu, tmp_path):
    """The C++ driver (examples/parth_engine.cpp) run like the reference's
    engine binary on the reference's SIFT graph and workload files: five sweep
    rows and metrics.json at beam 100 with the reference's recall."""
    import subprocess
    from test_abi import compile_example
    w = np.load(os.path.join(GOLDEN, "parth_workloads.npz"))
    m = json.load(open(os.path.join(GOLDEN, "parth_metrics.json")))["sift"]
    g = tmp_path / "sift_initial_graph.bin"
    g.write_bytes(np.array([10000, 128, 32], np.int32).tobytes()
                  + sift["base"].astype(np.float32).tobytes()
                  + sift["parth_adj"].astype(np.int32).tobytes())
    rec = np.zeros(len(w["sift_types"]), dtype=[("t", "<i4"), ("i", "<i4"), ("v", "<f4", (128,))])
    rec["t"], rec["i"], rec["v"] = w["sift_types"], w["sift_ids"], w["sift_vecs"]
    wf = tmp_path / "sift_workload.bin"
    wf.write_bytes(rec.tobytes())
    exe = compile_example("parth_engine", str(tmp_path / "parth_engine"))
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([exe, str(g), str(wf), "50", str(out)], capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stderr
    rows = [ln.split() for ln in r.stdout.splitlines() if ln[:1].isdigit()]
    assert [int(x[0]) for x in rows] == [10, 20, 32, 64, 100]
    res = json.load(open(out / "metrics.json"))
    assert abs(res["recall_10"] - m["recall_10"]) <= 0.002 and res["qps"] > m["qps"]
