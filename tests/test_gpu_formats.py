"""Index files and the reference's graph files.

* parth graph file: a PARTH build of SIFT10K (and of its first 5,000 points)
  at the reference's parameters, written with vmn_save_parth_graph, must be
  the reference's own file byte for byte (size and SHA-256 recorded in
  tests/golden/file_digests.json by make_fixtures.py), and must load back.
* FreshVamana records: the reference's G1 graph written as records has the
  size of the reference's vamana.out and parses, by an independent numpy
  reader, into the same vectors, degrees and neighbours; it loads back.
* Build from the reference's random initial graph (the input its FreshVamana
  constructor builds from): equal to the oracle's build from the same graph,
  and through the record-file loader equal to the in-memory call.
* Index file: a FRESH index after inserts and deletes, saved and loaded,
  answers every precision's searches identically and keeps updating
  identically (the FP8 scales of the build survive the round trip).
* BANG / DiskANN disk index: the DiskANN-built SIFT10K graph installed with
  its medoid 3732 and written with vmn_save_bang_index is the reference's
  sift10k_index_disk.bin and _metadata.bin byte for byte; loaded back it
  searches exactly like the oracle on that graph."""
import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def digest(path):
    raw = open(path, "rb").read()
    return {"bytes": len(raw), "sha256": hashlib.sha256(raw).hexdigest()}


@pytest.fixture(scope="module")
def digests():
    with open(os.path.join(GOLD, "file_digests.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("which,n", [("sift", 10000), ("custom_sift", 5000)])
def test_parth_graph_file_is_the_reference_file(vmn, sift, gpu, tmp_path, digests, which, n):
    base = sift["base"][:n].astype(np.float32)
    idx = vmn.Vamana(base, mode="parth")  # R=32, L=64, alpha 1.2 x 2, seed 1234
    path = tmp_path / f"{which}_initial_graph.bin"
    idx.save_parth_graph(str(path))
    assert digest(path) == digests[f"parth/data/{which}_initial_graph.bin"]
    back = vmn.Vamana.load_parth_graph(str(path))
    assert back.size == n and back.medoid == idx.medoid
    assert np.array_equal(back.adjacency, idx.adjacency)
    assert np.array_equal(back.vectors, base)
    q = sift["query"].astype(np.float32)
    a = idx.searchPoints(q, k=10, L=40)
    b = back.searchPoints(q, k=10, L=40)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    # the engine's own entry point is node 0 (engine.cu:476-477)
    assert vmn.Vamana.load_parth_graph(str(path), medoid=0).medoid == 0


def test_fresh_records_round_trip(vmn, sift, gpu, tmp_path, digests):
    g1 = np.load(os.path.join(GOLD, "g1_graph.npz"))
    base = sift["base"].astype(np.float32)
    idx = vmn.Vamana.from_graph(base, g1["adj"], medoid=5000, mode="fresh", R=64)
    path = tmp_path / "g1.out"
    idx.save_fresh_graph(str(path))
    assert os.path.getsize(path) == digests["vamana/vout/vamana.out"]["bytes"]
    rec = np.fromfile(path, dtype=np.uint8).reshape(-1, 772)
    assert np.array_equal(rec[:, :512].copy().view(np.float32), base)
    deg = rec[:, 512:516].copy().view(np.uint32).ravel()
    nbr = rec[:, 516:].copy().view(np.uint32).astype(np.int64)
    assert np.array_equal(deg, g1["degree"])
    got = np.where(np.arange(64)[None, :] < deg[:, None], nbr, -1)
    assert np.array_equal(got, g1["adj"])
    assert np.all(nbr[np.arange(64)[None, :] >= deg[:, None]] == 0xFFFFFFFF)
    back = vmn.Vamana.load_fresh_graph(str(path), dim=128, R=64, medoid=5000)
    assert np.array_equal(back.adjacency, idx.adjacency) and back.medoid == 5000
    q = sift["query"].astype(np.float32)
    a = idx.searchPoints(q, k=10, L=64)
    b = back.searchPoints(q, k=10, L=64)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_build_from_reference_random_graph(vmn, orc, sift, gpu, tmp_path):
    rg = np.load(os.path.join(GOLD, "randgraph.npz"))
    base = sift["base"].astype(np.float32)
    # the G1 build parameters (R=64, L=150, alpha 1.5, one pass)
    idx = vmn.Vamana(base, mode="fresh", R=64, L_build=150, alphas=(1.5,), init_adj=rg["adj"])
    want, med = orc.fresh_build(base, R=64, L=150, alphas=(1.5,), init=rg["adj"])
    assert idx.medoid == med
    assert np.array_equal(idx.adjacency, want)
    # the reference's record file of that graph, built through the file loader
    path = tmp_path / "sift10k_randomgraph.bin"
    recs = np.zeros((base.shape[0], 772), dtype=np.uint8)
    recs[:, :512] = base.view(np.uint8).reshape(-1, 512)
    recs[:, 512:516] = rg["degree"].astype(np.uint32).view(np.uint8).reshape(-1, 4)
    recs[:, 516:] = np.where(rg["adj"] < 0, 0xFFFFFFFF, rg["adj"]).astype(np.uint32).view(
        np.uint8).reshape(-1, 256)
    recs.tofile(path)
    built = vmn.Vamana.load_fresh_graph(str(path), dim=128, R=64, build=True, L_build=150,
                                        alphas=(1.5,))
    assert np.array_equal(built.adjacency, want)
    # the reference's G1 output from the same random graph (medoid 5000, its
    # own tie order) has mean degree 57.87; the canonical build lands close
    assert abs((want >= 0).sum(1).mean() - 57.87) < 1.5


def test_index_file_round_trip(vmn, gpu, tmp_path):
    ds = vmn.datasets
    base = ds.integer_mixture(6000, 128, k=32, seed=5, stream=0)
    more = ds.integer_mixture(1500, 128, k=32, seed=5, stream=2)
    q = ds.integer_mixture(100, 128, k=32, seed=5, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=9000,
                     low_precision="both")
    idx.insertPoints(more[:1000])
    idx.deleteIds(np.arange(0, 6000, 7, dtype=np.int32))
    idx.consolidate()
    idx.deleteIds(np.arange(3, 6000, 11, dtype=np.int32))  # tombstones left unconsolidated
    path = tmp_path / "index.vmn"
    idx.save(str(path))
    back = vmn.Vamana.load(str(path))
    assert back.size == idx.size and back.medoid == idx.medoid
    assert np.array_equal(back.adjacency, idx.adjacency)
    assert np.array_equal(back.deleted, idx.deleted)
    assert np.array_equal(back.vectors, idx.vectors)
    for prec in ("bf16", "fp8"):
        assert np.array_equal(back.lowp_rows(prec), idx.lowp_rows(prec))
    for prec, rr in (("fp32", 0), ("bf16", 20), ("fp8", 0), ("fp8", 20)):
        a = idx.searchPoints(q, k=10, L=40, prec=prec, rerank=rr)
        b = back.searchPoints(q, k=10, L=40, prec=prec, rerank=rr)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), prec
    for ix in (idx, back):
        ix.insertPoints(more[1000:])
        ix.consolidate()
    assert np.array_equal(back.adjacency, idx.adjacency)
    assert np.array_equal(back.lowp_rows("fp8"), idx.lowp_rows("fp8"))


def test_version1_index_file_loads_with_zero_centre(vmn, gpu, tmp_path):
    """A version-1 index file (FP8 scales, no centre: its rows were coded about
    0) loads with a zero centre: the decoded FP8 rows are e4m3(x / s) * s with
    the file's scales. The v1 file is the v2 file with the version field set
    to 1 and the trailing centre (DP floats) removed."""
    v = vmn.datasets.integer_mixture(2000, 128, k=16, seed=5, stream=0)
    idx = vmn.Vamana(v, mode="fresh", R=24, L_build=48, alphas=(1.2,), low_precision="fp8")
    p2 = tmp_path / "v2.vmn"
    idx.save(str(p2))
    raw = bytearray(p2.read_bytes())
    assert raw[:8] == b"VMNIDX01" and int.from_bytes(raw[8:12], "little") == 2
    dp = int.from_bytes(raw[12:16], "little")
    scales = np.frombuffer(bytes(raw[-8 * dp:-4 * dp]), dtype=np.float32)
    centre = np.frombuffer(bytes(raw[-4 * dp:]), dtype=np.float32)
    assert np.allclose(centre[:128], v.astype(np.float64).mean(0), atol=1e-3)
    raw[8:12] = (1).to_bytes(4, "little")
    p1 = tmp_path / "v1.vmn"
    p1.write_bytes(bytes(raw[:-4 * dp]))
    back = vmn.Vamana.load(str(p1))
    from test_gpu_dims import e4m3
    want = (e4m3(np.clip((v / scales[:128]).astype(np.float32), -448, 448)) * scales[:128]).astype(np.float32)
    assert np.allclose(back.lowp_rows("fp8"), want, rtol=0, atol=1e-4 * scales.max())
    ids, _ = back.searchPoints(v[:32], k=10, L=40, prec="fp8", rerank=20)
    assert (ids[:, 0] == np.arange(32)).all()


def test_bang_index_is_the_reference_file(vmn, orc, sift, gpu, tmp_path, digests):
    base = sift["base"].astype(np.float32)
    adj = np.load(os.path.join(GOLD, "diskann_graph.npz"))["adj"]
    idx = vmn.Vamana.from_graph(base, adj, medoid=3732, mode="fresh", R=64)
    ip, mp = str(tmp_path / "sift10k_index_disk.bin"), str(tmp_path / "sift10k_index_disk_metadata.bin")
    idx.save_bang_index(ip, mp)
    assert digest(ip) == digests["data/sift10k/sift10k_index_disk.bin"]
    assert digest(mp) == digests["data/sift10k/sift10k_index_disk_metadata.bin"]
    back = vmn.Vamana.load_bang_index(ip, mp, mode="fresh")
    assert back.medoid == 3732 and np.array_equal(back.adjacency, adj)
    q = sift["query"].astype(np.float32)
    for L in (10, 40):
        ids, d = back.searchPoints(q, k=10, L=L)
        oi, od, _, _ = orc.fresh_search(base, adj, q, 3732, L=L)
        assert np.array_equal(ids, oi) and np.array_equal(d, od)
    assert vmn.recall_at_k(ids, sift["gt100"][:, :10], 10) > 0.95


def test_cpp_fresh_vamana_driver(sift, gpu, tmp_path):
    """examples/fresh_vamana.cpp, the reference's FreshVamana driver over the
    C ABI, on the reference's random-graph file: builds, finds node 0, finds
    the inserted copy of it (new id 10000, distance 1), writes 10,001 records."""
    import subprocess
    from test_abi import compile_example
    base = sift["base"].astype(np.float32)
    rg = np.load(os.path.join(GOLD, "randgraph.npz"))
    rec = np.zeros(10000, dtype=[("v", "<f4", (128,)), ("deg", "<u4"), ("nbr", "<u4", (64,))])
    rec["v"], rec["deg"] = base, rg["degree"]
    rec["nbr"] = np.where(rg["adj"] >= 0, rg["adj"], 0).astype(np.uint32)
    g = tmp_path / "sift10k_randomgraph.bin"
    g.write_bytes(rec.tobytes())
    assert digest(str(g))["bytes"] == 7720000
    exe = compile_example("fresh_vamana", str(tmp_path / "fresh_vamana"))
    out = tmp_path / "out.bin"
    r = subprocess.run([exe, str(g), "unused", str(out)], capture_output=True, text=True,
                       timeout=100)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    first = [ln for ln in lines if ln.strip().startswith("id ")]
    assert first[0].split()[1] == "0" and first[0].split()[3] == "0.0"
    assert "inserted id 10000" in r.stdout
    # the inserted vector is node 0 plus 1 in its first coordinate
    assert first[5].split()[1] == "10000" and first[5].split()[3] == "0.0"
    assert first[6].split()[1] == "0" and first[6].split()[3] == "1.0"
    assert os.path.getsize(out) == 10001 * 772
