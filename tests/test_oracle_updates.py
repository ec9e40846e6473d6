"""Oracle FreshVamana updates: insert, tombstones, consolidation. The
reference has no consolidation and no reference fixture for updates, so these
are pinned by invariants that a plausible mistake breaks: no live edge into a
tombstone after consolidation, tombstoned rows empty, rows of untouched nodes
unchanged, every re-pruned row alpha-consistent, deleted ids never returned,
inserted points found at distance 0, recall against FP64 truth over live
points."""
import numpy as np
import pytest


def _data(vmn, n, seed=3):
    return vmn.datasets.integer_mixture(n, 128, k=32, seed=seed, stream=0)


def _alpha_consistent(vecs, row, node, alpha):
    """For kept p_i before p_j: alpha * d(p_i, p_j) > d(node, p_j) (else p_j
    would have been pruned by the <= rule)."""
    r = row[row >= 0]
    for j in range(len(r)):
        dj = np.float32(((vecs[r[j]] - vecs[node]) ** 2).sum())
        for i in range(j):
            dij = np.float32(((vecs[r[i]] - vecs[r[j]]) ** 2).sum())
            if not np.float32(alpha) * dij > dj:
                return False
    return True


def test_insert_points_are_found(vmn, orc):
    v = _data(vmn, 3000)
    adj, med = orc.fresh_build(v[:2500], R=32, L=64, alphas=(1.2,))
    full = np.full((3000, 32), -1, np.int32)
    full[:2500] = adj
    orc.fresh_insert(v, 2500, full, med, L=64, alpha=1.2)
    new = full[2500:]
    assert ((new >= 0).sum(1) > 0).all()
    for i, row in enumerate(full):
        r = row[row >= 0]
        assert i not in r and len(set(r.tolist())) == len(r)
    ids, d, _, _ = orc.fresh_search(v, full, v[2500:], med, L=64, out_n=1)
    assert (d[:, 0] == 0).all()
    # an exact duplicate vector may exist; the hit must be a node with that vector
    assert (np.abs(v[ids[:, 0]] - v[2500:]).sum(1) == 0).all()


def test_consolidation_invariants(vmn, orc):
    v = _data(vmn, 4000, seed=8)
    adj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.2,))
    rng = np.random.default_rng(0)
    dele = np.zeros(4000, np.uint8)
    dele[rng.choice(4000, 400, replace=False)] = 1
    before = adj.copy()
    touched = np.array([(not dele[p]) and any(dele[x] for x in row if x >= 0)
                        for p, row in enumerate(before)])
    n = orc.consolidate(v, adj, dele, alpha=1.2)
    assert n == touched.sum()
    live_rows = adj[dele == 0]
    assert not np.isin(live_rows[live_rows >= 0], np.flatnonzero(dele)).any()
    assert (adj[dele == 1] == -1).all()
    assert np.array_equal(adj[~touched & (dele == 0)], before[~touched & (dele == 0)])
    for p in np.flatnonzero(touched)[:50]:
        assert _alpha_consistent(v, adj[p], p, 1.2)


def test_search_skips_tombstones_and_recall_after_consolidation(vmn, orc):
    v = _data(vmn, 5000, seed=9)
    q = vmn.datasets.integer_mixture(100, 128, k=32, seed=9, stream=1)
    adj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    dele = np.zeros(5000, np.uint8)
    dele[np.random.default_rng(1).choice(5000, 1000, replace=False)] = 1
    ids, _, _, _ = orc.fresh_search(v, adj, q, med, L=64, deleted=dele)
    assert not np.isin(ids[ids >= 0], np.flatnonzero(dele)).any()
    orc.consolidate(v, adj, dele, alpha=1.2)
    med2 = orc.medoid_live(v, dele) if dele[med] else med
    ids2, _, _, _ = orc.fresh_search(v, adj, q, med2, L=64, deleted=dele)
    live = np.flatnonzero(dele == 0)
    gt, _ = orc.bruteforce_fp64(v[live], q, 10)
    rec = np.mean([len(set(ids2[i]) & set(live[gt[i]])) / 10 for i in range(100)])
    assert rec >= 0.95


def test_medoid_live_excludes_tombstones(vmn, orc):
    v = _data(vmn, 1000, seed=2)
    m = orc.medoid(v)
    dele = np.zeros(1000, np.uint8)
    assert orc.medoid_live(v, dele) == m
    dele[m] = 1
    m2 = orc.medoid_live(v, dele)
    assert m2 != m and m2 >= 0


def test_consolidation_keeps_live_edges_and_fills_maximally(vmn, orc):
    """Properties that fix the repair rule: each repaired row starts with the
    node's live out-neighbours in their old order; every other entry is a live
    out-neighbour of a tombstoned out-neighbour, not occluded by an entry before
    it (alpha * d(r, c) <= d(p, c)); and a row left short of R has every unused
    candidate occluded by some entry of the row."""
    v = _data(vmn, 3000, seed=12)
    R, alpha = 24, 1.2
    adj, med = orc.fresh_build(v, R=R, L=48, alphas=(1.0, 1.2))
    dele = np.zeros(3000, np.uint8)
    dele[np.random.default_rng(5).choice(3000, 450, replace=False)] = 1
    before = adj.copy()
    orc.consolidate(v, adj, dele, alpha=alpha)
    # integer data: exact squared distances; the occlusion product is one
    # float32 rounding, as in the oracle and the kernel
    d2 = lambda a, b: np.float32(((v[a].astype(np.float64) - v[b]) ** 2).sum())
    alpha = np.float32(alpha)
    checked = 0
    for p in range(3000):
        row0 = before[p][before[p] >= 0]
        if dele[p] or not dele[row0].any():
            continue
        live0 = [x for x in row0 if not dele[x]]
        row = list(adj[p][adj[p] >= 0])
        assert row[:len(live0)] == live0
        cands = {w for x in row0 if dele[x] for w in before[x][before[x] >= 0]
                 if not dele[w] and w != p and w not in live0}
        for j in range(len(live0), len(row)):
            c = row[j]
            assert c in cands
            assert all(alpha * d2(r, c) > d2(p, c) for r in row[:j])
        if len(row) < R:
            for c in cands - set(row):
                assert any(alpha * d2(r, c) <= d2(p, c) for r in row)
        checked += 1
    assert checked > 500
