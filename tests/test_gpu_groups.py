"""Query groups (one per XCD, by nearest k-means centroid): a batch large
enough to be grouped gives exactly the results of the same batch searched
from one queue, in both modes and every precision, eagerly, on contexts and
in replayed graphs; every query of the batch is searched exactly once (every
output row is written, and each query's hop count is its own)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NQ = 4096  # above 8 queries per CU on a 256-CU MI355X, so the batch is grouped


@pytest.fixture(scope="module")
def data(vmn):
    base = vmn.datasets.integer_mixture(30000, 128, k=48, seed=21, stream=0)
    qs = vmn.datasets.integer_mixture(NQ, 128, k=48, seed=21, stream=1)
    return base, qs


def _search(idx, torch, gpu, q, groups, **kw):
    idx.set_query_groups(groups)
    ids = torch.full((q.shape[0], 10), -7, dtype=torch.int32, device=gpu)
    d = torch.full((q.shape[0], 10), -7.0, dtype=torch.float32, device=gpu)
    idx.searchPoints(q, k=10, ids_out=ids, dists_out=d, sync=True, **kw)
    hops, nd = idx.search_stats()
    return ids.cpu().numpy(), d.cpu().numpy(), hops, nd


@pytest.mark.parametrize("mode", ["fresh", "parth"])
def test_grouped_batch_equals_one_queue(vmn, gpu, data, mode):
    import torch
    base, qs = data
    assert NQ >= 8 * torch.cuda.get_device_properties(0).multi_processor_count
    idx = vmn.Vamana(base, mode=mode, R=32, L_build=64, alphas=(1.0, 1.2) if mode == "fresh" else (1.2, 1.2),
                     low_precision="both")
    q = torch.from_numpy(qs).to(gpu)
    for prec, rr in (("fp32", 0), ("bf16", 40), ("fp8", 40)):
        a = _search(idx, torch, gpu, q, 0, L=64, prec=prec, rerank=rr)
        b = _search(idx, torch, gpu, q, 8, L=64, prec=prec, rerank=rr)
        assert (a[0] >= 0).all() and (b[0] >= 0).all(), prec
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (mode, prec)
        # the sums of per-query expansions and distances agree: no query was
        # searched twice or skipped
        assert a[2:] == b[2:], (mode, prec)
    # groups of another count and a ragged batch (not a multiple of 64)
    c = _search(idx, torch, gpu, q[:NQ - 37], 3, L=64, prec="bf16", rerank=40)
    a = _search(idx, torch, gpu, q[:NQ - 37], 0, L=64, prec="bf16", rerank=40)
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])
    idx.set_query_groups(8)


def test_grouped_replay_and_contexts(vmn, gpu, data):
    import torch
    base, qs = data
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), low_precision="bf16")
    q = torch.from_numpy(qs).to(gpu)
    ref = _search(idx, torch, gpu, q, 0, L=60, prec="bf16", rerank=40)
    idx.set_query_groups(8)
    ctxs = [vmn.SearchContext(idx) for _ in range(2)]
    outs = [(torch.empty((NQ, 10), dtype=torch.int32, device=gpu),
             torch.empty((NQ, 10), dtype=torch.float32, device=gpu)) for _ in range(2)]
    for c, (i_, d_) in zip(ctxs, outs):
        c.capture(q, i_, d_, k=10, L=60, prec="bf16", rerank=40)
    for rep in range(3):
        for c, (i_, d_) in zip(ctxs, outs):
            i_.fill_(-7)
        torch.cuda.synchronize()
        for c in ctxs:
            c.replay()
        torch.cuda.synchronize()
        for i_, d_ in outs:
            assert np.array_equal(i_.cpu().numpy(), ref[0]), rep
            assert np.array_equal(d_.cpu().numpy(), ref[1]), rep
        # eager grouped work on the index stream between replays
        x = _search(idx, torch, gpu, q, 8, L=60, prec="bf16", rerank=40)
        assert np.array_equal(x[0], ref[0])
    for c in ctxs:
        c.close()


def test_grouped_build_equals_oracle(vmn, orc, gpu):
    # builds search every node: a batch above the threshold, so grouped
    base = vmn.datasets.integer_mixture(6000, 128, k=16, seed=5, stream=0)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=48, alphas=(1.0, 1.2))
    oadj, omed = orc.fresh_build(base, R=32, L=48, alphas=(1.0, 1.2))
    assert idx.medoid == omed
    assert np.array_equal(idx.adjacency, oadj)


def test_group_ids_follow_the_data(vmn, gpu, data):
    # the lists the search reads form a partition of the batch (checked by the
    # library), and queries of one mixture component share a group: a broken
    # assignment (random groups) would put about 1/8 of each component in its
    # most common group
    base, qs = data
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.2,))
    gid = idx.query_group_ids(qs)
    assert gid.shape == (NQ,) and int(gid.max()) < 8
    labels = vmn.datasets._rng(21, 1).integers(0, 48, size=NQ)
    pur = [np.bincount(gid[labels == c]).max() / (labels == c).sum() for c in range(48)
           if (labels == c).sum() >= 20]
    assert float(np.mean(pur)) >= 0.6, np.mean(pur)
    # the same batch gives the same groups
    assert np.array_equal(gid, idx.query_group_ids(qs))


def test_query_groups_argument_checks(vmn, gpu, data):
    base, _ = data
    idx = vmn.Vamana(base[:3000], mode="fresh", R=16, L_build=32, alphas=(1.2,))
    with pytest.raises(Exception):
        idx.set_query_groups(-1)
    with pytest.raises(Exception):
        idx.set_query_groups(33)
    idx.set_query_groups(1)
    idx.set_query_groups(0)


def test_launch_tuner_chooses_grouping(vmn, gpu, data):
    import torch
    base, qs = data
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), low_precision="bf16")
    q = torch.from_numpy(qs).to(gpu)
    q2 = torch.from_numpy(vmn.datasets.integer_mixture(NQ, 128, k=48, seed=21, stream=2)).to(gpu)
    ref = _search(idx, torch, gpu, q, 0, L=60, prec="bf16", rerank=40)
    t = vmn.tuning.launch_shape(idx, q, k=10, L=60, prec="bf16", rerank=40, waves=(20, 24),
                                caps=(0,), reps=2, overlap=True, queries_alt=q2)
    # waves x groups x tile shapes (the BF16 form at d = 128, L = 60 has 4-tile kernels)
    assert len(t["table"]) == 8 and {r["groups"] for r in t["table"]} == {0, 8}
    assert {r["tiles"] for r in t["table"]} == {2, 4}
    assert t["groups"] in (0, 8) and t["tiles"] in (2, 4)
    # the installed choice gives the same results
    ids = torch.full((NQ, 10), -7, dtype=torch.int32, device=gpu)
    d = torch.empty((NQ, 10), dtype=torch.float32, device=gpu)
    idx.searchPoints(q, k=10, L=60, prec="bf16", rerank=40, ids_out=ids, dists_out=d, sync=True)
    assert np.array_equal(ids.cpu().numpy(), ref[0]) and np.array_equal(d.cpu().numpy(), ref[1])
