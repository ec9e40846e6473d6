"""BANG's exact search schedule on the GPU (vmn_bang_search) against the oracle.

* SIFT10K over the DiskANN graph from entry 123 at L 40, 400 queries (100
  from the query file, 300 zero rows): ids, distances and per-query
  computations / rounds equal the oracle's, and their totals are BANG's
  printed ones (641,853 computations, 47 rounds, tie-aware recall 2,190).
* Other list sizes (both LDS list capacities), k, entry points, real-valued
  rows at widths that need padding, tombstones, a parent whose neighbours
  share a bloom cell, the round cap, and device-tensor inputs: bit-exact
  against the oracle.
"""
import json
import os

import numpy as np
import pytest

from test_oracle_bang import tie_aware_total

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def sift_bang(vmn, sift, gpu):
    exp = json.load(open(os.path.join(GOLD, "bang_expected.json")))
    t = np.load(os.path.join(GOLD, "bang_sift10k.npz"))
    base = sift["base"].astype(np.float32)
    adj = np.load(os.path.join(GOLD, "diskann_graph.npz"))["adj"]
    q = np.zeros((400, 128), np.float32)
    q[:100] = sift["query"].astype(np.float32)
    idx = vmn.Vamana.from_graph(base, adj, medoid=3732, mode="fresh", R=64)
    return {"exp": exp, "base": base, "adj": adj, "q": q, "idx": idx,
            "gt_ids": t["gt_ids"], "gt_dists": t["gt_dists"]}


def same(got, want):
    for g, w in zip(got, want):
        g = g.cpu().numpy() if hasattr(g, "cpu") else g
        assert np.array_equal(g, w)


def test_bang_gpu_reproduces_printed_sift10k_run(orc, sift_bang):
    b = sift_bang
    print("bang sift10k L=40 entry=123 nq=400", flush=True)
    ids, d, st = b["idx"].bang_search(b["q"], k=10, L=40, entry=123)
    same((ids, d, st), orc.bang_search(b["base"], b["adj"], b["q"], 123, 40, k=10))
    assert int(st[:, 0].sum()) == 641853 and int(st[:, 1].max()) == 47
    assert tie_aware_total(ids, b["gt_ids"], b["gt_dists"], 10) == 2190


@pytest.mark.parametrize("L,k,entry", [(10, 10, 123), (64, 5, 3732), (65, 20, -1),
                                       (200, 100, 0), (1, 1, 77)])
def test_bang_gpu_other_list_sizes(orc, sift_bang, L, k, entry):
    b = sift_bang
    print(f"bang sift10k L={L} k={k} entry={entry}", flush=True)
    got = b["idx"].bang_search(b["q"][:160], k=k, L=L, entry=entry)
    e = 3732 if entry < 0 else entry
    same(got, orc.bang_search(b["base"], b["adj"], b["q"][:160], e, L, k=k))


@pytest.mark.parametrize("dim,n", [(32, 3000), (100, 3000), (128, 3000), (200, 3000),
                                   (128, 40000)])
def test_bang_gpu_real_valued_rows(vmn, orc, gpu, dim, n):
    """Gaussian rows (FP32 rounding everywhere): the GPU sums each distance in
    the oracle's order, so results are bit-identical, also where the width is
    padded (100 -> 128, 200 -> 256). 40,000 nodes: past the ids whose bloom
    cells are all distinct, the filter is the 399,887-cell one."""
    rng = np.random.default_rng(dim)
    v = rng.standard_normal((n, dim)).astype(np.float32)
    adj = orc.random_init(n, 24, 32, seed=dim)
    q = rng.standard_normal((96, dim)).astype(np.float32)
    print(f"bang gaussian dim={dim} n={n}", flush=True)
    idx = vmn.Vamana.from_graph(v, adj, medoid=5, mode="fresh", R=32)
    got = idx.bang_search(q, k=10, L=48, entry=5)
    same(got, orc.bang_search(v, adj, q, 5, 48, k=10))


@pytest.mark.parametrize("dim,n", [(128, 10000), (100, 3000), (128, 40000), (200, 3000)])
def test_bang_gpu_large_batches(vmn, orc, sift_bang, dim, n):
    """Batches of 2,048 queries (all of a round's row loads in flight) and
    4,608 (past 16 queries per CU: 4 candidates' loads per lane group): the
    rows are read 8 lanes per candidate, chain s in lane s, the same sums.
    Bit-identical to the oracle on SIFT10K and on Gaussian rows (padded
    widths 100 -> 128 and 200 -> 256; 40,000 nodes: the 399,887-cell
    filter)."""
    if n == 10000:
        b = sift_bang
        v, adj, entry, idx = b["base"], b["adj"], 123, b["idx"]
        q = np.concatenate([np.tile(b["q"][:100], (20, 1)), np.zeros((48, 128), np.float32)])
        L = 40
    else:
        rng = np.random.default_rng(n + dim)
        v = rng.standard_normal((n, dim)).astype(np.float32)
        adj = orc.random_init(n, 24, 32, seed=dim)
        q = rng.standard_normal((4608 if dim == 200 or n == 40000 else 2048, dim)).astype(np.float32)
        entry, L = 5, 48
        idx = vmn.Vamana.from_graph(v, adj, medoid=entry, mode="fresh", R=32)
    print(f"bang large batch dim={dim} n={n} nq={q.shape[0]}", flush=True)
    got = idx.bang_search(q, k=10, L=L, entry=entry)
    same(got, orc.bang_search(v, adj, q, entry, L, k=10))
    if n == 10000:
        assert int(got[2][:2000, 0].sum()) == 20 * int(got[2][:100, 0].sum())


def test_bang_gpu_on_a_built_graph(vmn, orc, gpu):
    """BANG's schedule over a FRESH graph the library built (100,000 integer
    mixture rows, R 64; past 32,792 nodes, so the 399,887-cell filter with its
    collisions), entry at the index's medoid, L 40 and 80: bit-identical to
    the oracle run on the same adjacency."""
    ds = vmn.datasets
    base = ds.integer_mixture(100000, 128, k=64, sigma=20.0, seed=77, stream=0)
    q = ds.integer_mixture(3000, 128, k=64, sigma=20.0, seed=77, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=64, L_build=64, alphas=(1.0, 1.2))
    adj = idx.adjacency
    for L in (40, 80):
        print(f"bang built graph L={L}", flush=True)
        got = idx.bang_search(q, k=10, L=L)
        same(got, orc.bang_search(base, adj, q, idx.medoid, L, k=10))


def test_bang_gpu_tombstones(vmn, orc, sift_bang):
    b = sift_bang
    idx = vmn.Vamana.from_graph(b["base"], b["adj"], medoid=3732, mode="fresh", R=64)
    dead = np.unique(np.concatenate([np.arange(0, 10000, 37), [3732]])).astype(np.int32)
    idx.deleteIds(dead)
    dm = np.zeros(10000, np.uint8)
    dm[dead] = 1
    print("bang sift10k tombstones", flush=True)
    got = idx.bang_search(b["q"][:100], k=10, L=40, entry=123)
    want = orc.bang_search(b["base"], b["adj"], b["q"][:100], 123, 40, k=10, deleted=dm)
    same(got, want)
    assert not dm[got[0][got[0] >= 0]].any()


def test_bang_gpu_neighbours_sharing_a_cell(vmn, orc, gpu):
    """Two neighbours of the entry point with one bloom cell: only the first in
    adjacency order is a candidate, on the GPU as in the oracle."""
    # the pair with the smallest larger id: ids below 32,792 have cells of their own
    pair = (16573, 32792)
    assert orc.bang_hash(pair[0]) == orc.bang_hash(pair[1])
    n = max(pair) + 1
    rng = np.random.default_rng(3)
    v = rng.integers(0, 100, size=(n, 32)).astype(np.float32)
    adj = np.full((n, 64), -1, np.int32)
    adj[0, :3] = [pair[1], 7, pair[0]]
    adj[7, :2] = [pair[1], 9]
    q = v[[pair[0], pair[1], 9]]
    print(f"bang shared cell {pair}", flush=True)
    idx = vmn.Vamana.from_graph(v, adj, medoid=0, mode="fresh", R=64)
    got = idx.bang_search(q, k=4, L=8, entry=0)
    want = orc.bang_search(v, adj, q, 0, 8, k=4)
    same(got, want)
    assert pair[0] not in got[0][0] and pair[1] in got[0][0]


def test_bang_gpu_entry_without_neighbours(vmn, orc, gpu):
    n = 100
    v = np.arange(n, dtype=np.float32)[:, None].repeat(32, 1)
    adj = np.full((n, 64), -1, np.int32)
    adj[1:-1, 0] = np.arange(2, n)
    q = np.full((3, 32), 7.0, np.float32)
    print("bang entry without neighbours", flush=True)
    idx = vmn.Vamana.from_graph(v, adj, medoid=0, mode="fresh", R=64)
    got = idx.bang_search(q, k=5, L=8, entry=0)
    same(got, orc.bang_search(v, adj, q, 0, 8, k=5))
    assert (got[2] == 1).all() and (got[0][:, 0] == 0).all() and (got[0][:, 1:] == -1).all()


def test_bang_gpu_round_cap(vmn, orc, gpu):
    n = 300
    v = np.arange(n, dtype=np.float32)[:, None].repeat(32, 1)
    adj = np.full((n, 64), -1, np.int32)
    adj[:-1, 0] = np.arange(1, n)
    q = np.full((2, 32), n - 1, np.float32)
    print("bang path graph", flush=True)
    idx = vmn.Vamana.from_graph(v, adj, medoid=0, mode="fresh", R=64)
    for L in (1, 3):
        ids, d, st = idx.bang_search(q, k=L, L=L, entry=0)
        same((ids, d, st), orc.bang_search(v, adj, q, 0, L, k=L))
        assert (st[:, 1] == 4 * L + 19).all()


def test_bang_gpu_device_tensors(orc, sift_bang):
    import torch
    b = sift_bang
    qd = torch.from_numpy(b["q"][:64]).cuda()
    print("bang device tensors", flush=True)
    ids, d, st = b["idx"].bang_search(qd, k=10, L=40, entry=123)
    assert ids.is_cuda and st.is_cuda
    same((ids, d, st), orc.bang_search(b["base"], b["adj"], b["q"][:64], 123, 40, k=10))


def test_cpp_bang_driver(vmn, sift_bang, tmp_path):
    """examples/bang_search.cpp, BANG's command line over the C ABI, on the
    reference's files rebuilt from the fixtures (the DiskANN disk index with
    its metadata, the 100-row query file, the truthset's first 400 rows):
    prints BANG's totals for its 400-query run."""
    import subprocess
    from test_abi import compile_example
    b = sift_bang
    ip, mp = str(tmp_path / "sift10k_index_disk.bin"), str(tmp_path / "sift10k_index_disk_metadata.bin")
    b["idx"].save_bang_index(ip, mp)
    qf, tf = str(tmp_path / "siftsmall_query.bin"), str(tmp_path / "sift10k_groundtruth.bin")
    vmn.write_vectors(qf, b["q"][:100], "bin")
    vmn.write_truthset(tf, b["gt_ids"], b["gt_dists"])
    exe = compile_example("bang_search", str(tmp_path / "bang_search"))
    print("bang_search driver", flush=True)
    r = subprocess.run([exe, ".", ".", ip, qf, ".", ".", tf, "400", "1", "256", "512", "256",
                        "10", "8", "1", "40", "123"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "[medoidID=123, numQueries=400]" in out
    assert "iterations = 47" in out
    assert "total_recall = 2190.00 num_queries = 400 recall_at 10" in out
    assert "40\t54.75\t total parents 0\t total computations: 641853" in out


def test_bang_gpu_argument_checks(vmn, sift_bang):
    idx = sift_bang["idx"]
    for k, L in ((11, 10), (0, 10), (10, 257)):
        with pytest.raises(vmn.VamanaError):
            idx.bang_search(sift_bang["q"][:4], k=k, L=L)
    with pytest.raises(vmn.VamanaError):
        idx.bang_search(sift_bang["q"][:4], k=10, L=40, entry=10000)
    ids, d, st = idx.bang_search(sift_bang["q"][:0], k=10, L=40)
    assert ids.shape == (0, 10) and st.shape == (0, 2)
