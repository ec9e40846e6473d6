"""Pins of the oracle's fresh-mode restatement (vamana/include/search.cuh,
kernels.cuh, bloom_filter.cuh), the graph the benchmark builds.

What pins each part:
* bloom hashes: the closed form of bloom_filter.cuh:11-30 over 14k ids;
* search: exact k-NN when the beam covers a small connected graph whose ids
  have no bloom false positive (then every node is expanded once and every
  distance computed once), plus hand-built cases for the tie rule of
  mergeIntoWorklist (search.cuh:140-223: worklist entries before new ones on
  equal distance) and the self skip (search.cuh:126);
* out-prune and reverse prune: the properties a robust prune with `<=`
  (kernels.cuh:329-353, 737) must have on every row of a build -- sorted,
  no self loop or duplicate, alpha-consistent (no kept entry occludes a later
  kept one), maximal (each dropped candidate is occluded by a nearer kept
  one, or the row is full), the reverse pairs are the out-rows, every
  reverse pair is kept unless occluded or the row is full, candidates only
  from the first M -- and hand-built cases for the tie rules of
  mergeIntoVisitedSet (:171-229) and mergeIntoReverseEdges (:566-624).

Each check is shown to fail when one mistake is planted in the oracle
(orc.set_mutation): `<` for `<=`, new entries first on ties, no self skip,
the M cap one too high or too low, the wrong side first in either candidate
merge. The data are small integers, so every distance and alpha product is
exact and ties are frequent.
"""
import numpy as np
import pytest

M64 = (1 << 64) - 1


def bloom_closed(x, seed, mul):
    """bf_hashFn1/2: FNV-style over the 4 id bytes, 64-bit wrap-around, mod 39,988."""
    h = seed
    for s in (0, 8, 16, 24):
        h = ((h ^ ((x >> s) & 0xFF)) * mul) & M64
    return h % 39988


@pytest.fixture
def mutant(orc):
    yield orc.set_mutation
    orc.set_mutation(0)


def dist(v, a, b):
    # integer rows: the float64 sum is the exact FP32 distance
    return np.float32(((v[a].astype(np.float64) - v[b]) ** 2).sum())


def occludes(v, alpha, r, c, dqc):
    # the reference's `d(p*, p') * alpha <= d(q, p')`, one float32 rounding
    return np.float32(dist(v, r, c)) * np.float32(alpha) <= dqc


def no_bloom_false_positive(orc, ids):
    h = {i: orc.bf_hash(int(i)) for i in ids}
    for i in ids:
        others = {c for j in ids if j != i for c in h[j]}
        if h[i][0] in others and h[i][1] in others:
            return False
    return True


def connected_graph(rng, n, W):
    rows = [set() for _ in range(n)]
    for i in range(n):
        rows[i].add((i + 1) % n)
        rows[(i + 1) % n].add(i)
    for _ in range(3 * n):
        a, b = (int(x) for x in rng.integers(0, n, 2))
        if a != b and len(rows[a]) < W and len(rows[b]) < W:
            rows[a].add(b)
            rows[b].add(a)
    adj = np.full((n, W), -1, np.int32)
    for i, r in enumerate(rows):
        r = sorted(r)
        rng.shuffle(r)
        adj[i, :len(r)] = r
    return adj


# ---------------------------------------------------------------------------
# bloom filter and search
# ---------------------------------------------------------------------------

def test_bloom_hashes_equal_closed_form(orc):
    rng = np.random.default_rng(3)
    ids = list(range(4096)) + [int(x) for x in rng.integers(0, 1 << 32, 10000, dtype=np.uint64)]
    ids += [0xFFFFFFFF, 1 << 31, 255, 256, 65535, 65536, 3732, 5000, 999_999, 10_000_000]
    for x in ids:
        assert orc.bf_hash(x) == (bloom_closed(x, 0xcbf29ce4, 0x01000193),
                                  bloom_closed(x, 0x84222325, 0x1B3)), x


def test_fresh_search_is_exact_knn_when_beam_covers_graph(orc):
    rng = np.random.default_rng(7)
    n, d = 48, 8
    v = rng.integers(0, 16, (n, d)).astype(np.float32)
    adj = connected_graph(rng, n, 8)
    assert no_bloom_false_positive(orc, range(n))
    q = rng.integers(0, 16, (30, d)).astype(np.float32)
    ids, dists, hops, nd = orc.fresh_search(v, adj, q, start=5, L=n, out_n=n)
    gi, gd = orc.bruteforce_fp64(v, q, n)
    assert np.array_equal(dists, gd.astype(np.float32))
    for a, b, dd in zip(ids, gi, gd):
        for val in np.unique(dd):
            assert set(a[dd == val].tolist()) == set(b[dd == val].tolist())
    # every node expanded once and measured once (the start node included)
    assert (hops == n).all() and (nd == n).all()


def tie_graph():
    """Query at the origin. s = (10, 0) at squared distance 100 links to a =
    (3, 0) and c = (0, 3); a links to b = (-3, 0); a, b and c are all at 9."""
    v = np.array([[10, 0], [3, 0], [0, 3], [-3, 0]], np.float32)  # s a c b
    adj = np.array([[1, 2], [3, -1], [-1, -1], [-1, -1]], np.int32)
    return v, adj, np.zeros((1, 2), np.float32)


def test_fresh_search_keeps_worklist_entries_first_on_ties(orc, mutant):
    """L = 2 from s: expanding s lists a, c (9, 9) and drops s; expanding a
    offers b at 9, which goes after the worklist's a and c and falls off the
    list, so the result is [a, c]. New entries first would give [b, a]."""
    v, adj, q = tie_graph()
    assert no_bloom_false_positive(orc, range(4))
    ids, d, _, _ = orc.fresh_search(v, adj, q, start=0, L=2, out_n=2)
    assert ids[0].tolist() == [1, 2] and d[0].tolist() == [9.0, 9.0]
    mutant(2)
    ids, _, _, _ = orc.fresh_search(v, adj, q, start=0, L=2, out_n=2)
    assert ids[0].tolist() == [3, 1]


def _self_visits(orc, v, adj, med):
    _, _, _, _, vi, _, vc = orc.fresh_search(v, adj, v, med, L=16, out_n=1,
                                             queries_are_nodes=True, want_visited=True)
    return sum(int(i in vi[i, :vc[i]]) for i in range(len(v)) if i != med)


def test_fresh_search_of_a_node_skips_itself(orc, mutant):
    """A node searched as a query never enters its own worklist, so it is
    never in its own visited set (search.cuh:126, with the node id). The
    start node is the one exception: every search begins there."""
    rng = np.random.default_rng(9)
    v = rng.integers(0, 8, (80, 6)).astype(np.float32)
    adj = connected_graph(rng, 80, 6)
    assert _self_visits(orc, v, adj, 0) == 0
    mutant(3)
    assert _self_visits(orc, v, adj, 0) > 0


# ---------------------------------------------------------------------------
# prunes: properties of every row of a build
# ---------------------------------------------------------------------------

def row_violations(v, q, row, cands, alpha, R):
    """Violations of a robust prune with `<=` of candidate set `cands` (node q
    excluded) into `row` (kept ids in output order)."""
    bad = []
    if len(row) > R:
        bad.append("longer than R")
    if len(set(row)) != len(row) or q in row:
        bad.append("duplicate or self loop")
    if not set(row) <= set(cands):
        bad.append("entry not a candidate")
        return bad
    dq = {c: dist(v, q, c) for c in cands}
    if any(dq[row[i]] > dq[row[i + 1]] for i in range(len(row) - 1)):
        bad.append("not sorted")
    for j in range(len(row)):
        for i in range(j):
            if occludes(v, alpha, row[i], row[j], dq[row[j]]):
                bad.append(f"kept {row[j]} occluded by kept {row[i]}")
    kept = set(row)
    for c in cands:
        if c in kept:
            continue
        if any(dq[r] <= dq[c] and occludes(v, alpha, r, c, dq[c]) for r in row):
            continue
        if len(row) == R and dq[c] >= dq[row[-1]]:
            continue
        bad.append(f"candidate {c} dropped without cause")
    return bad


def kept(row):
    return [int(x) for x in row if x >= 0]


def fresh_pass_checks(orc, v, adj0, med, L, alpha, R, M):
    """One FRESH pass through the oracle's parts (search all nodes, out-prune,
    reverse prune), checked row by row. Returns (graph after the pass,
    violations)."""
    n = len(v)
    _, _, _, _, vi, vd, vc = orc.fresh_search(v, adj0, v, med, L, out_n=1,
                                              queries_are_nodes=True, want_visited=True)
    out = adj0.copy()
    pdst, pd = orc.fresh_out_prune(v, out, R, alpha, vi, vd, vc)
    rev = out.copy()
    orc.fresh_reverse(v, rev, R, alpha, pdst, pd, M=M)
    bad = []
    # out-prune: candidates = visited set and current row, the node left out
    for q in range(n):
        cands = sorted((set(vi[q, :vc[q]].tolist()) | set(kept(adj0[q]))) - {q})
        bad += [f"out {q}: {b}" for b in row_violations(v, q, kept(out[q]), cands, alpha, R)]
        # each kept p* emits (p* <- q) with d(q, p*)
        if kept(pdst[q]) != kept(out[q]):
            bad.append(f"pairs of {q} are not its row")
        for p, dd in zip(kept(pdst[q]), pd[q]):
            if dd != dist(v, q, p):
                bad.append(f"pair distance {q}->{p}")
    # reverse prune: candidates = the first M of the reverse pairs by
    # (distance, source) merged with the current row
    rc = [[] for _ in range(n)]
    for q in range(n):
        for p in kept(pdst[q]):
            rc[p].append(q)
    for p in range(n):
        if not rc[p]:
            if kept(rev[p]) != kept(out[p]):
                bad.append(f"row {p} changed without reverse pairs")
            continue
        a = sorted(rc[p], key=lambda q: (dist(v, p, q), q))[:M]
        b = kept(out[p])
        merged = sorted([(dist(v, p, x), 0, i, x) for i, x in enumerate(a)]
                        + [(dist(v, p, x), 1, i, x) for i, x in enumerate(b)])
        window = [x for _, _, _, x in merged][:M]
        row = kept(rev[p])
        bad += [f"rev {p}: {b_}" for b_ in row_violations(v, p, row, sorted(set(window)), alpha, R)]
        # reverse completeness, spelled out: a pair inside the window is kept
        # unless a nearer member occludes it or the row is full
        for q in set(rc[p]) & set(window):
            dqp = dist(v, p, q)
            if q in row or len(row) == R or any(
                    dist(v, p, r) <= dqp and occludes(v, alpha, r, q, dqp) for r in row):
                continue
            bad.append(f"reverse pair {p} <- {q} lost")
    return rev, bad


def small_build(orc, seed, n=300, d=6, R=8, L=16, M=10):
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 8, (n, d)).astype(np.float32)
    med = orc.medoid(v)
    g0 = orc.random_init(n, R, R, seed=1234)
    return v, med, g0, R, L, M


@pytest.mark.parametrize("seed", [1, 2])
def test_fresh_build_rows_have_the_prune_properties(orc, seed):
    v, med, g0, R, L, M = small_build(orc, seed)
    g1, bad1 = fresh_pass_checks(orc, v, g0, med, L, 1.0, R, M)
    g2, bad2 = fresh_pass_checks(orc, v, g1, med, L, 1.2, R, M)
    assert not bad1 and not bad2, (bad1 + bad2)[:10]
    # the build is those two passes
    adj, m = orc.fresh_build(v, R=R, L=L, alphas=(1.0, 1.2), M=M)
    assert m == med and np.array_equal(adj, g2)


@pytest.mark.parametrize("mut", [1, 4, 5])
def test_fresh_build_checks_catch_planted_prune_mistakes(orc, mutant, mut):
    """`<` for `<=` (1) breaks alpha-consistency on exact ties at alpha 1.0;
    an M cap one too high (4) keeps a candidate outside the first M, one too
    low (5) drops the M-th without cause."""
    caught = False
    mutant(mut)
    for seed in (1, 2):
        v, med, g0, R, L, M = small_build(orc, seed)
        g1, bad1 = fresh_pass_checks(orc, v, g0, med, L, 1.0, R, M)
        _, bad2 = fresh_pass_checks(orc, v, g1, med, L, 1.2, R, M)
        caught |= bool(bad1 or bad2)
    assert caught


def test_out_prune_takes_visited_entries_first_on_ties(orc, mutant):
    """mergeIntoVisitedSet (kernels.cuh:171-229) puts visited-set entries
    before current neighbours at equal distance. Node 0 at the origin has
    visited node 1 = (3, 0) and current neighbour 2 = (0, 3), both at 9; with
    R = 1 the first of them is the row."""
    v = np.array([[0, 0], [3, 0], [0, 3]], np.float32)
    adj = np.array([[2], [-1], [-1]], np.int32)
    vi = np.array([[1]], np.int32)
    vd = np.array([[9.0]], np.float32)
    vc = np.array([1], np.int32)
    a = adj.copy()
    orc.fresh_out_prune(v, a, 1, 1.0, vi, vd, vc)
    assert a[0].tolist() == [1]
    mutant(6)
    a = adj.copy()
    orc.fresh_out_prune(v, a, 1, 1.0, vi, vd, vc)
    assert a[0].tolist() == [2]


def test_reverse_prune_takes_reverse_pairs_first_on_ties(orc, mutant):
    """mergeIntoReverseEdges (kernels.cuh:566-624) puts reverse candidates
    before current neighbours at equal distance. Node 0 at the origin has
    current neighbour 2 = (0, 3) and receives the pair (0 <- 1), 1 = (3, 0),
    both at 9; with R = 1 the first of them is the row."""
    v = np.array([[0, 0], [3, 0], [0, 3]], np.float32)
    adj = np.array([[2], [0], [-1]], np.int32)
    pdst = np.array([[0]], np.int32)
    pd = np.array([[9.0]], np.float32)
    a = adj.copy()
    orc.fresh_reverse(v, a, 1, 1.0, pdst, pd, base=1, M=8)
    assert a[0].tolist() == [1]
    mutant(7)
    a = adj.copy()
    orc.fresh_reverse(v, a, 1, 1.0, pdst, pd, base=1, M=8)
    assert a[0].tolist() == [2]
