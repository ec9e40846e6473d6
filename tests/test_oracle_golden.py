"""The CPU oracle against the reference's own fixtures and closed forms.

Pins: parth-mode build (parth/data/*_initial_graph.bin, produced by
parth/src/build_graph.cu), the medoid (DiskANN metadata medoid 3732 of
data/sift10k/sift10k_index_disk_metadata.bin), FP64 ground truth
(siftsmall_groundtruth.ivecs and data/sift10k/sift10k_groundtruth.bin), and
search (exact k-NN when the beam covers a small connected graph).
"""
import numpy as np
import pytest


def test_sift_is_integer_valued(sift):
    b = sift["base"].astype(np.int64)
    # every squared-L2 partial sum stays below 2^24: FP32 distances are exact
    assert 128 * int(b.max()) ** 2 < 2 ** 24


def test_oracle_parth_build_matches_reference_sift10k(orc, sift):
    base = sift["base"].astype(np.float32)
    adj, med = orc.parth_build(base, R=32, L=64, alphas=(1.2, 1.2), row_width=32)
    assert med == 3732
    assert np.array_equal(adj, sift["parth_adj"])


def test_oracle_parth_build_matches_reference_5k_subset(orc, sift):
    base = sift["base"][:5000].astype(np.float32)
    adj, _ = orc.parth_build(base, R=32, L=64, alphas=(1.2, 1.2), row_width=32)
    assert np.array_equal(adj, sift["custom_adj"])


@pytest.mark.parametrize("name", ["toy", "vis"])
def test_oracle_parth_build_matches_reference_2d(orc, toy2d, name):
    v = toy2d[name + "_vecs"]
    # gen_toy_data.py: R = k = 6, L = min(n, 2k) = 12, alpha 1.2, MAX_DEGREE 32
    adj, _ = orc.parth_build(v, R=6, L=12, alphas=(1.2, 1.2), row_width=32)
    assert np.array_equal(adj, toy2d[name + "_adj"])


def test_random_init_properties(orc):
    adj = orc.random_init(1000, 32, 40, seed=1234)
    assert (adj[:, 32:] == -1).all()
    for i, row in enumerate(adj[:, :32]):
        assert len(set(row.tolist())) == 32 and i not in row and row.min() >= 0


def test_mutation_in_prune_is_caught(orc, sift):
    """A prune with <= instead of < (or another alpha) no longer reproduces the
    fixture: the fixture discriminates the prune rule."""
    base = sift["base"][:5000].astype(np.float32)
    adj, _ = orc.parth_build(base, R=32, L=64, alphas=(1.25, 1.25), row_width=32)
    assert not np.array_equal(adj, sift["custom_adj"])


def test_groundtruth_matches_siftsmall_ivecs(orc, sift):
    base = sift["base"].astype(np.float32)
    q = sift["query"].astype(np.float32)
    ids, d = orc.bruteforce_fp64(base, q, 100)
    gt = sift["gt100"]
    assert all(set(ids[i, :10]) == set(gt[i, :10]) for i in range(100))
    assert (ids[:, 0] == gt[:, 0]).all()
    # the ivecs order agrees up to equal distances
    dd = ((base[gt].astype(np.float64) - q[:, None, :]) ** 2).sum(-1)
    assert np.array_equal(np.sort(dd, axis=1), d)
    # DiskANN truthset: same ids and the same distances (as float32)
    assert np.array_equal(ids[:, :10], sift["diskann_gt_ids"])
    assert np.array_equal(d[:, :10].astype(np.float32), sift["diskann_gt_dists"])


def test_parth_search_exact_when_beam_covers_graph(orc):
    rng = np.random.default_rng(5)
    v = rng.integers(0, 50, size=(60, 16)).astype(np.float32)
    adj, med = orc.parth_build(v, R=8, L=16, alphas=(1.2, 1.2))
    q = rng.integers(0, 50, size=(20, 16)).astype(np.float32)
    # symmetric closure so every node is reachable from the medoid
    full = [set() for _ in range(60)]
    for i, row in enumerate(adj):
        for j in row[row >= 0]:
            full[i].add(int(j)); full[int(j)].add(i)
    W = max(len(s) for s in full)
    sym = -np.ones((60, W), dtype=np.int32)
    for i, s in enumerate(full):
        sym[i, :len(s)] = sorted(s)
    ids, dists, hops, _ = orc.parth_search(v, sym, q, med, L=64, out_n=5)
    gt, gd = orc.bruteforce_fp64(v, q, 5)
    assert np.array_equal(dists, gd.astype(np.float32))
    # ids equal up to ties in distance
    for a, b, dd in zip(ids, gt, gd):
        for t in range(5):
            if not (dd == dd[t]).sum() > 1:
                assert a[t] == b[t]


def test_parth_search_recall_band_on_reference_graph(orc, sift):
    """Recall@10 of the reference's own graph (SIFT10K, R=32) with 100 SIFT
    queries; parth logged 0.9105 at L=100 under a mixed workload
    (parth/vis_outputs/sift/metrics.json)."""
    base = sift["base"].astype(np.float32)
    q = sift["query"].astype(np.float32)
    ids, _, _, _ = orc.parth_search(base, sift["parth_adj"], q, 3732, L=100)
    gt = sift["gt100"][:, :10]
    rec = np.mean([len(set(ids[i]) & set(gt[i])) / 10 for i in range(100)])
    assert 0.88 <= rec <= 0.95
