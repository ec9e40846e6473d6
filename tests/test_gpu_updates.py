"""GPU FreshVamana updates (insert, delete, consolidate) bit-exact against the
oracle, with searches checked after every step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _data(vmn, n, seed=3, stream=0):
    return vmn.datasets.integer_mixture(n, 128, k=32, seed=seed, stream=stream)


def test_gpu_insert_matches_oracle(vmn, orc, gpu):
    v = _data(vmn, 6000)
    idx = vmn.Vamana(v[:5000], mode="fresh", R=32, L_build=64, alphas=(1.2,), capacity=6000)
    oadj, med = orc.fresh_build(v[:5000], R=32, L=64, alphas=(1.2,))
    full = np.full((6000, 32), -1, np.int32)
    full[:5000] = oadj
    assert np.array_equal(idx.adjacency, oadj)
    ids = idx.insertPoints(v[5000:5400])
    assert np.array_equal(ids, np.arange(5000, 5400))
    orc.fresh_insert(v[:5400], 5000, full[:5400], med, L=64, alpha=1.2)
    assert np.array_equal(idx.adjacency, full[:5400])
    idx.insertPoints(v[5400:6000])
    orc.fresh_insert(v, 5400, full, med, L=64, alpha=1.2)
    assert np.array_equal(idx.adjacency, full)
    q = _data(vmn, 50, stream=1)
    gi, gd = idx.searchPoints(q, k=10, L=48)
    oi, od, _, _ = orc.fresh_search(v, full, q, med, L=48)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    with pytest.raises(vmn.VamanaError):
        idx.insertPoints(v[:1])  # capacity exhausted


def test_gpu_small_insert_batches_match_oracle(vmn, orc, gpu):
    """Batches whose reverse pairs are far fewer than the nodes are grouped by a
    device radix sort instead of a count/scan over every node (batch x R x 8 <
    nodes); the graph must not depend on which grouping ran. Batches of 1, 3,
    16 and 31 points go through the sort; 400 points through the node scan;
    a tombstone makes the searches skip a node in between."""
    v = _data(vmn, 9000, seed=6)
    idx = vmn.Vamana(v[:8000], mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=9000)
    oadj, med = orc.fresh_build(v[:8000], R=32, L=64, alphas=(1.0, 1.2))
    full = np.full((9000, 32), -1, np.int32)
    full[:8000] = oadj
    dele = np.zeros(9000, np.uint8)
    size = 8000
    for b in (1, 3, 16, 31, 400, 16):
        if b == 31:
            idx.deleteIds(np.array([17, 4242], dtype=np.int32))
            dele[[17, 4242]] = 1
        idx.insertPoints(v[size:size + b])
        orc.fresh_insert(v[:size + b], size, full[:size + b], med, L=64, alpha=1.2,
                         deleted=dele[:size + b])
        size += b
        assert np.array_equal(idx.adjacency, full[:size]), b


def test_gpu_delete_consolidate_matches_oracle(vmn, orc, gpu):
    v = _data(vmn, 6000, seed=4)
    q = _data(vmn, 80, seed=4, stream=1)
    idx = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2))
    adj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    rng = np.random.default_rng(5)
    kill = rng.choice(6000, 900, replace=False).astype(np.int32)
    kill[0] = med  # the entry point goes too
    dele = np.zeros(6000, np.uint8)
    dele[kill] = 1
    idx.deleteIds(kill)
    assert np.array_equal(idx.deleted, dele.astype(bool))
    gi, gd = idx.searchPoints(q, k=10, L=64)
    oi, od, _, _ = orc.fresh_search(v, adj, q, med, L=64, deleted=dele)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    idx.consolidate()
    orc.consolidate(v, adj, dele, alpha=1.2)
    assert np.array_equal(idx.adjacency, adj)
    m2 = orc.medoid_live(v, dele)
    assert idx.medoid == m2
    gi, gd = idx.searchPoints(q, k=10, L=64)
    oi, od, _, _ = orc.fresh_search(v, adj, q, m2, L=64, deleted=dele)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    assert not np.isin(gi[gi >= 0], kill).any()


def test_gpu_stream_of_updates_matches_oracle(vmn, orc, gpu):
    """delete -> insert (searches skip tombstones) -> consolidate, twice."""
    v = _data(vmn, 7000, seed=6)
    idx = vmn.Vamana(v[:5000], mode="fresh", R=32, L_build=64, alphas=(1.2,), capacity=7000)
    adj5, med = orc.fresh_build(v[:5000], R=32, L=64, alphas=(1.2,))
    full = np.full((7000, 32), -1, np.int32)
    full[:5000] = adj5
    dele = np.zeros(7000, np.uint8)
    rng = np.random.default_rng(7)
    n = 5000
    for rnd in range(2):
        live = np.flatnonzero(dele[:n] == 0)
        kill = rng.choice(live, 250, replace=False).astype(np.int32)
        dele[kill] = 1
        idx.deleteIds(kill)
        idx.insertPoints(v[n:n + 1000])
        orc.fresh_insert(v[:n + 1000], n, full[:n + 1000], med, L=64, alpha=1.2,
                         deleted=dele[:n + 1000])
        n += 1000
        assert np.array_equal(idx.adjacency, full[:n]), f"round {rnd} insert"
        idx.consolidate()
        orc.consolidate(v[:n], full[:n], dele[:n], alpha=1.2)
        if dele[med]:
            med = orc.medoid_live(v[:n], dele[:n])
        assert idx.medoid == med
        assert np.array_equal(idx.adjacency, full[:n]), f"round {rnd} consolidate"


def test_gpu_delete_by_vector(vmn, gpu):
    """vmn_delete_vectors (deletePoints, findPointsKernel kernels.cuh:905-933):
    each vector tombstones the lowest live id whose row equals it exactly, or
    reports -1. Rows are duplicated here so that the lowest-id rule decides,
    and each answer is checked against an exact-match scan over the live rows;
    deleting the same vectors again takes the next copy, then finds none."""
    v = _data(vmn, 3000, seed=11)
    for a, b in ((10, 2900), (10, 1500), (20, 600), (30, 31)):
        v[b] = v[a]
    idx = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.2,))
    probe = np.concatenate([v[[2900, 20, 31, 777]], np.full((1, 128), 999.0, np.float32)])
    live = np.ones(3000, bool)
    for rnd in range(4):
        ids = idx.deletePoints(probe)
        for j, p_ in enumerate(probe):
            hit = np.flatnonzero(live & (v == p_).all(1))
            want = int(hit[0]) if len(hit) else -1
            assert ids[j] == want, (rnd, j, ids[j], want)
            if want >= 0:
                live[want] = False
        assert np.array_equal(idx.deleted, ~live), rnd
    # 10's row has three copies, 20's two, 30's two, 777's one
    assert (~live).sum() == 3 + 2 + 2 + 1
    # deleting the same vectors again finds no live match unless a duplicate exists
    again = idx.deletePoints(probe[:3])
    for a, b in zip(again, ids[:3]):
        assert a == -1 or (a != b and np.array_equal(v[a], v[b]))


def test_gpu_parth_delete_search_matches_oracle(vmn, orc, sift, gpu):
    base = sift["base"].astype(np.float32)
    q = sift["query"].astype(np.float32)
    idx = vmn.Vamana.from_graph(base, sift["parth_adj"], medoid=3732)
    kill = np.arange(0, 10000, 7, dtype=np.int32)
    dele = np.zeros(10000, np.uint8)
    dele[kill] = 1
    idx.deleteIds(kill)
    gi, gd = idx.searchPoints(q, k=10, L=64)
    oi, od, _, _ = orc.parth_search(base, sift["parth_adj"], q, 3732, L=64, deleted=dele)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)
    with pytest.raises(vmn.VamanaError):
        idx.deleteIds(np.array([10000], np.int32))


def test_gpu_staged_inserts_merge_then_patch(vmn, orc, gpu, tmp_path):
    """insertPoints / searchPoints2 / patchGraph semantics (vamana_impl.cuh:22-33,
    :99-180, :182-265): staged rows take the next ids, every search merges their
    exact distances into the graph's results by (distance, id), the savers
    refuse while rows are staged, and patchGraph adds them as one insert batch,
    which must equal the oracle's insert of that batch."""
    v = _data(vmn, 3300, seed=8)
    idx = vmn.Vamana(v[:3000], mode="fresh", R=32, L_build=64, alphas=(1.2,), capacity=3300)
    oadj, med = orc.fresh_build(v[:3000], R=32, L=64, alphas=(1.2,))
    ids = idx.insert_staged(v[3000:3150])
    assert ids.tolist() == list(range(3000, 3150))
    assert idx.staged_count == 150 and idx.size == 3000
    q = _data(vmn, 40, seed=8, stream=1)
    gi, gd = idx.searchPoints(q, k=10, L=48)
    oi, od, _, _ = orc.fresh_search(v[:3000], oadj, q, med, L=48)
    pdist = ((v[3000:3150][None].astype(np.float64) - q[:, None].astype(np.float64)) ** 2).sum(-1)
    for i in range(len(q)):
        cand = [(float(od[i, j]), int(oi[i, j])) for j in range(10) if oi[i, j] >= 0]
        cand += [(float(np.float32(pdist[i, j])), 3000 + j) for j in range(150)]
        cand.sort()
        assert [c[1] for c in cand[:10]] == gi[i].tolist(), i
        assert np.array_equal(np.float32([c[0] for c in cand[:10]]), gd[i]), i
    rows = idx.searchPoints2(q, k=10, L=48)
    assert np.array_equal(rows, v[:3150][gi])
    with pytest.raises(vmn.VamanaError):
        idx.save(str(tmp_path / "pending.vmn"))
    idx.patchGraph()
    assert idx.staged_count == 0 and idx.size == 3150
    full = np.full((3150, 32), -1, np.int32)
    full[:3000] = oadj
    orc.fresh_insert(v[:3150], 3000, full, med, L=64, alpha=1.2)
    assert np.array_equal(idx.adjacency, full)
    # immediate inserts after staged ones: the staged rows go in first
    idx.insert_staged(v[3150:3160])
    assert idx.insertPoints(v[3160:3170]).tolist() == list(range(3160, 3170))
    assert idx.size == 3170 and idx.staged_count == 0


def test_gpu_heavy_delete_consolidation_tiers(vmn, orc, gpu):
    """Consolidation runs pools of at most 256 candidates first and lists rows
    whose pool could be larger (tombstoned neighbours x R > 256) for a
    full-pool launch. With 60% of the nodes deleted most rows at R = 32 have
    more than 8 tombstoned neighbours, so both tiers run; the graph must be
    the oracle's."""
    v = _data(vmn, 4000, seed=9)
    idx = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2))
    adj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    kill = np.random.default_rng(9).choice(4000, 2400, replace=False).astype(np.int32)
    dele = np.zeros(4000, np.uint8)
    dele[kill] = 1
    dead_nbrs = (np.where(adj >= 0, dele[np.maximum(adj, 0)], 0)).sum(1)
    live = dele == 0
    assert (dead_nbrs[live] * 32 > 256).any() and (dead_nbrs[live] * 32 <= 256).any()
    idx.deleteIds(kill)
    idx.consolidate()
    orc.consolidate(v, adj, dele, alpha=1.2)
    assert np.array_equal(idx.adjacency, adj)


@pytest.mark.parametrize("frac", [0.15, 0.6])
def test_gpu_consolidation_on_fp32_rows_matches_oracle(vmn, orc, gpu, frac):
    """Rows that are not exact in BF16 (half-integers up to 255.5) take the
    FP32 occlusion checks and fill, several members or pivots per pass over
    the candidates' rows. Their differences are integers, so every FP32 sum
    is exact in any order and the oracle fixes the graph. At 60% deleted most
    rows have more than 8 tombstoned neighbours, so the full-pool tier runs
    too."""
    v = _data(vmn, 4000, seed=11) + np.float32(0.5)
    idx = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2))
    assert idx.build_precision == "fp32"
    adj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    assert np.array_equal(idx.adjacency, adj)
    kill = np.random.default_rng(12).choice(4000, int(4000 * frac), replace=False).astype(np.int32)
    dele = np.zeros(4000, np.uint8)
    dele[kill] = 1
    idx.deleteIds(kill)
    idx.consolidate()
    orc.consolidate(v, adj, dele, alpha=1.2)
    assert np.array_equal(idx.adjacency, adj)
    assert idx.medoid == orc.medoid_live(v, dele)
