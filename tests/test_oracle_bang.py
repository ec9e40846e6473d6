"""The oracle's BANG exact search schedule against what BANG printed.

BANG's SIFT10K run (bang_exact/output.txt, built with L 40, entry point 123,
parANN.cuh:40-41,122-128) searched 400 queries over the DiskANN graph
data/sift10k/sift10k_index_disk.bin. Its query file holds 100 rows; the other
300 rows are read past the file's end (parANN.cu:225-226) from a fresh,
zero-filled allocation, so they are zero vectors here. Four of its six runs
print 47 rounds, 641,853 distance computations and a tie-aware total recall of
2,190 (54.75 %); the other two stop the batch one round early (46 rounds,
641,807: the reference's per-round `d_nextIter` flag is cleared by each block
of the next kernel while others have set it, parANN.cu:1010-1012,1486).
"""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def bang(sift):
    exp = json.load(open(os.path.join(GOLD, "bang_expected.json")))
    t = np.load(os.path.join(GOLD, "bang_sift10k.npz"))
    adj = np.load(os.path.join(GOLD, "diskann_graph.npz"))["adj"]
    nq = exp["runs"][0]["num_queries"]
    q = np.zeros((nq, 128), np.float32)
    q[:100] = sift["query"].astype(np.float32)
    return {"exp": exp, "base": sift["base"].astype(np.float32), "adj": adj, "q": q,
            "gt_ids": t["gt_ids"], "gt_dists": t["gt_dists"]}


def tie_aware_total(found, gt_ids, gt_dists, k):
    """Sum over queries of |GT set ∩ first k results|, the GT set taking every
    id whose distance equals the k-th (bang_exact/src/main.cu:87-137)."""
    tot = 0
    for i in range(found.shape[0]):
        tb = k - 1
        while tb < gt_ids.shape[1] and gt_dists[i, tb] == gt_dists[i, k - 1]:
            tb += 1
        tot += len(set(gt_ids[i, :tb].tolist()) & set(found[i, :k].tolist()))
    return tot


def run(orc, b, L=None, **kw):
    e = b["exp"]
    L = L or e["runs"][0]["L"]
    return orc.bang_search(b["base"], b["adj"], b["q"], e["entry"], L, k=e["k"], **kw)


def test_bang_hash_cells(orc):
    # FNV-1a over the four bytes from offset 0xcbf29ce4, prime 0x01000193, in
    # 64-bit arithmetic, modulo the 399,887 cells (parANN.cu:36,1055-1064);
    # id 0 hashes four zero bytes, so its cell is offset * prime^4 mod 2^64.
    h0 = (0xCBF29CE4 * pow(0x01000193, 4, 1 << 64)) % (1 << 64) % 399887
    assert orc.bang_hash(0) == h0
    cells = {orc.bang_hash(i) for i in range(10000)}
    assert max(cells) < 399887 and len(cells) > 9800


def test_bang_cells_are_distinct_below_32792(orc):
    """The GPU filter indexes by id on graphs of at most 32,792 nodes: the
    cells of ids 0 .. 32,791 are distinct, and 32,792 shares one with 16,573."""
    cells = [orc.bang_hash(i) for i in range(32793)]
    assert len(set(cells[:32792])) == 32792
    assert cells[32792] == cells[16573]


def test_bang_distance_is_exact_on_integer_rows(orc, sift):
    b = sift["base"][:50].astype(np.int64)
    q = sift["query"][0].astype(np.int64)
    want = ((b - q) ** 2).sum(1)
    got = [orc.bang_l2sq(q.astype(np.float32), r.astype(np.float32)) for r in b]
    assert np.array_equal(np.array(got, np.float64), want.astype(np.float64))


def test_bang_matches_printed_sift10k_run(orc, bang):
    r0 = bang["exp"]["runs"][0]
    ids, d, st = run(orc, bang)
    assert int(st[:, 0].sum()) == r0["computations"] == 641853
    assert int(st[:, 1].max()) == r0["rounds"] == 47
    tot = tie_aware_total(ids, bang["gt_ids"], bang["gt_dists"], bang["exp"]["k"])
    assert tot == int(r0["total_recall"]) == 2190
    # the majority of the printed runs agree with these totals
    same = [r for r in bang["exp"]["runs"]
            if (r["computations"], r["rounds"]) == (641853, 47)]
    assert len(same) == 4
    # distances are the exact squared L2 of the returned rows (integer data)
    for i in (0, 57, 99, 250):
        x = bang["base"][ids[i]].astype(np.int64)
        want = ((x - bang["q"][i].astype(np.int64)) ** 2).sum(1)
        assert np.array_equal(d[i].astype(np.int64), want)
        assert (np.diff(d[i]) >= 0).all()


def test_bang_planted_mistake_in_the_stop_rule_changes_the_totals(orc, bang):
    """Scanning the whole list for a parent after a round without candidates
    (instead of ending the query) gives 642,920 computations, not the printed
    641,853."""
    try:
        orc.set_mutation(8)
        _, _, st = run(orc, bang)
    finally:
        orc.set_mutation(0)
    assert int(st[:, 0].sum()) == 642920


def test_bang_list_grows_to_L_with_farther_candidates(orc):
    """1-D rows, query at 0, L = 4: 0 (x = 1, entry) -> 1 (x = 2) -> {2 (x = 10),
    3 (x = 20)}. Round 2 finds 2 and 3 farther than the list's last entry, but
    nb = max(0, min(L - B, count)) = 2 (parANN.cu:1596) still adds both: the
    list is [0, 1, 2, 3]; round 3 expands 2 and ends. Without the min(L - B,
    count) floor (the planted mistake) round 2 adds nothing and the query ends
    with two results."""
    v = np.array([[1.0], [2.0], [10.0], [20.0]], np.float32)
    adj = np.full((4, 64), -1, np.int32)
    adj[0, 0], adj[1, 0], adj[1, 1] = 1, 2, 3
    q = np.zeros((1, 1), np.float32)
    ids, dist, st = orc.bang_search(v, adj, q, 0, L=4, k=4)
    assert ids[0].tolist() == [0, 1, 2, 3] and dist[0].tolist() == [1, 4, 100, 400]
    assert st[0].tolist() == [4, 3]
    try:
        orc.set_mutation(10)
        ids10, _, st10 = orc.bang_search(v, adj, q, 0, L=4, k=4)
    finally:
        orc.set_mutation(0)
    assert ids10[0].tolist() == [0, 1, -1, -1] and st10[0].tolist() == [4, 2]


def test_bang_merge_puts_new_entries_first_on_ties(orc):
    """1-D rows, query at 0, L = 2: node 0 (x = 10, entry) -> 1 (x = 5) -> 2
    (x = -5). Round 1 lists [1 (25), 0 (100)] and expands 1; round 2 finds 2 at
    the same distance 25, and the merge's lower_bound / upper_bound positions
    (parANN.cu:1606-1614) put it before 1: the list is [2, 1]. Round 3 expands
    2, finds nothing and ends the query. With list entries first on ties (the
    planted mistake) the list would be [1, 2]."""
    v = np.array([[10.0], [5.0], [-5.0]], np.float32)
    adj = np.full((3, 64), -1, np.int32)
    adj[0, 0], adj[1, 0] = 1, 2
    q = np.zeros((1, 1), np.float32)
    ids, dist, st = orc.bang_search(v, adj, q, 0, L=2, k=2)
    assert ids[0].tolist() == [2, 1] and dist[0].tolist() == [25.0, 25.0]
    assert st[0].tolist() == [3, 3]
    try:
        orc.set_mutation(9)
        ids9, _, _ = orc.bang_search(v, adj, q, 0, L=2, k=2)
    finally:
        orc.set_mutation(0)
    assert ids9[0].tolist() == [1, 2]


def test_bang_entry_without_neighbours(orc):
    """An entry point with an empty row: round 1 computes the entry alone, the
    list holds it visited, and the query ends after that round."""
    v = np.arange(8, dtype=np.float32)[:, None].repeat(4, 1)
    adj = np.full((8, 64), -1, np.int32)
    adj[1, :3] = [2, 3, 4]
    q = np.full((2, 4), 5.0, np.float32)
    ids, dist, st = orc.bang_search(v, adj, q, 0, L=4, k=3)
    assert ids[:, 0].tolist() == [0, 0] and (ids[:, 1:] == -1).all()
    assert dist[:, 0].tolist() == [100.0, 100.0] and st.tolist() == [[1, 1], [1, 1]]


def test_bang_full_beam_on_complete_graph_is_exact_knn(orc):
    """With every node adjacent to every other and L = n, the list ends up
    holding all n nodes in distance order: the results are the exact k-NN."""
    rng = np.random.default_rng(7)
    n, d = 60, 16
    v = rng.integers(0, 50, size=(n, d)).astype(np.float32)
    adj = np.full((n, 64), -1, np.int32)
    for i in range(n):
        adj[i, :n - 1] = [j for j in range(n) if j != i]
    q = rng.integers(0, 50, size=(20, d)).astype(np.float32)
    ids, dist, st = orc.bang_search(v, adj, q, 0, L=n, k=n)
    exact = ((q[:, None, :].astype(np.int64) - v[None].astype(np.int64)) ** 2).sum(-1)
    for i in range(q.shape[0]):
        # stable by distance, equal distances in arrival order
        assert np.array_equal(np.sort(exact[i]), dist[i].astype(np.int64))
        assert sorted(ids[i].tolist()) == list(range(n))
        assert (exact[i][ids[i]] == dist[i]).all()
    # round 1 computes the entry and its n - 1 neighbours; nothing is new later
    assert (st[:, 0] == n).all() and (st[:, 1] == 2).all()


def test_bang_round_cap_and_empty_list_padding(orc):
    """A path graph with L = 1 walks one node a round; the host loop's cap of
    4L + 19 rounds (MAX_PARENTS_PERQUERY - 1) ends it. k beyond the list pads
    with -1 / inf."""
    n = 200
    v = np.arange(n, dtype=np.float32)[:, None].repeat(4, 1)
    adj = np.full((n, 64), -1, np.int32)
    for i in range(n - 1):
        adj[i, 0] = i + 1
    q = np.full((1, 4), n - 1, np.float32)
    ids, dist, st = orc.bang_search(v, adj, q, 0, L=1, k=3)
    assert st[0, 1] == 4 * 1 + 19 and st[0, 0] == 4 * 1 + 19 + 1
    assert ids[0, 0] == 4 + 19 and (ids[0, 1:] == -1).all() and np.isinf(dist[0, 1:]).all()


def test_bang_tombstones_are_never_candidates(orc, bang):
    dm = np.zeros(10000, np.uint8)
    ids0, _, _ = run(orc, bang)
    dm[np.unique(ids0[:100, :3])] = 1
    ids, _, _ = run(orc, bang, deleted=dm)
    hit = ids[ids >= 0]
    assert not dm[hit[hit != bang["exp"]["entry"]]].any()
