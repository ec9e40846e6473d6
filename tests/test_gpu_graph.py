"""HIP graph capture / replay of a batched search, search contexts, launch
shapes: results identical to eager searches and to the oracle's search on the
same graph (integer rows: BF16 distances are exact, so the BF16 walk is the
FP32 walk)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replay_matches_eager(vmn, orc, gpu):
    import torch
    base = vmn.datasets.integer_mixture(8000, 128, k=32, seed=9, stream=0)
    qa = vmn.datasets.integer_mixture(300, 128, k=32, seed=9, stream=1)
    qb = vmn.datasets.integer_mixture(300, 128, k=32, seed=9, stream=3)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), low_precision="bf16")
    ea, da = idx.searchPoints(qa, k=10, L=48, prec="bf16", rerank=32)
    eb, db = idx.searchPoints(qb, k=10, L=48, prec="bf16", rerank=32)
    # the eager results are the oracle's
    for qq, e, d in ((qa, ea, da), (qb, eb, db)):
        oi, od, _, _ = orc.fresh_search(base, idx.adjacency, qq, idx.medoid, L=48, out_n=10)
        assert np.array_equal(e, oi) and np.array_equal(d, od)
    ctx = vmn.SearchContext(idx)
    q = torch.from_numpy(qa).to(gpu)
    ids = torch.empty((300, 10), dtype=torch.int32, device=gpu)
    d = torch.empty((300, 10), dtype=torch.float32, device=gpu)
    ctx.capture(q, ids, d, k=10, L=48, prec="bf16", rerank=32)
    st = torch.cuda.ExternalStream(ctx.stream)
    for rep in range(4):
        ids.fill_(-7)
        torch.cuda.synchronize()
        ctx.replay()
        st.synchronize()
        assert np.array_equal(ids.cpu().numpy(), ea) and np.array_equal(d.cpu().numpy(), da), rep
        # eager work on the index's own stream between replays
        x, _ = idx.searchPoints(qb, k=10, L=48, prec="bf16", rerank=32)
        assert np.array_equal(x, eb)
    q.copy_(torch.from_numpy(qb).to(gpu))
    torch.cuda.synchronize()
    ctx.replay()
    st.synchronize()
    assert np.array_equal(ids.cpu().numpy(), eb) and np.array_equal(d.cpu().numpy(), db)
    ctx.close()


def test_contexts_concurrent_and_launch_shapes(vmn, gpu):
    import torch
    base = vmn.datasets.integer_mixture(8000, 128, k=32, seed=2, stream=0)
    qs = vmn.datasets.integer_mixture(512, 128, k=32, seed=2, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.2,))
    ref, refd = idx.searchPoints(qs, k=10, L=64)
    q = torch.from_numpy(qs).to(gpu)
    ctxs = [vmn.SearchContext(idx) for _ in range(3)]
    outs = [(torch.empty((128, 10), dtype=torch.int32, device=gpu),
             torch.empty((128, 10), dtype=torch.float32, device=gpu)) for _ in range(4)]
    for i in range(4):
        c = ctxs[i % 3]
        c.search(q[128 * i:128 * (i + 1)], outs[i][0], outs[i][1], k=10, L=64)
    torch.cuda.synchronize()
    got = np.concatenate([o[0].cpu().numpy() for o in outs])
    assert np.array_equal(got, ref)
    for w, cap in ((8, 0), (16, 128), (48, 256)):
        idx.set_launch(w, cap)
        x, xd = idx.searchPoints(qs, k=10, L=64)
        assert np.array_equal(x, ref) and np.array_equal(xd, refd)
    t = vmn.tuning.launch_shape(idx, q, k=10, L=64, waves=(16, 32), caps=(0, 128), reps=1)
    assert t["waves_per_cu"] in (16, 32)
    t = vmn.tuning.launch_shape(idx, q, k=10, L=64, waves=(12, 24), caps=(0, 128), reps=4,
                                overlap=True)
    assert (t["waves_per_cu"], t["list_cap"]) in ((12, 0), (24, 0), (12, 128), (24, 128))
    assert len(t["table"]) == 4
    # the installed shape gives the same results
    x, xd = idx.searchPoints(qs, k=10, L=64)
    assert np.array_equal(x, ref) and np.array_equal(xd, refd)
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("dim", [128, 96])
def test_tile_shapes_give_the_same_results(vmn, orc, gpu, dim):
    """The BF16 / FP8 searches gathering 4 MFMA tiles a round return what 2
    tiles return (BF16 on integer rows: the oracle's FP32 search), for small
    batches (row norms from the table) and large ones, with beam lists of 64
    and 256; a mode without 4-tile kernels runs 2 whatever is set."""
    import torch
    base = vmn.datasets.integer_mixture(6000, dim, k=24, seed=31, stream=0)
    qs = vmn.datasets.integer_mixture(2300, dim, k=24, seed=31, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), low_precision="both")
    q = torch.from_numpy(qs).to(gpu)
    for L in (60, 150):
        for prec in ("bf16", "fp8"):
            for nq in (7, 2300):
                out = {}
                for tiles in (2, 4):
                    idx.set_tiles(tiles)
                    i_, d_ = idx.searchPoints(q[:nq], k=10, L=L, prec=prec, rerank=40)
                    out[tiles] = (i_.cpu().numpy(), d_.cpu().numpy())
                assert np.array_equal(out[2][0], out[4][0]) and np.array_equal(out[2][1], out[4][1]), \
                    (L, prec, nq)
                if prec == "bf16" and nq == 7:
                    oi, _, _, _ = orc.fresh_search(base, idx.adjacency, qs[:nq], idx.medoid, L=L,
                                                   out_n=10)
                    assert np.array_equal(out[4][0], oi), L
    assert idx.search_kernel_name(10, 60, "bf16", 0, tiles=4).endswith(", 1, 1, 4>(vmn::SearchArgs)")
    assert idx.search_kernel_name(10, 60, "fp32", 0, tiles=4).endswith(", 1, 0, 2>(vmn::SearchArgs)")
    with pytest.raises(Exception):
        idx.set_tiles(3)
    idx.set_tiles(0)
    parth = vmn.Vamana(base[:3000], mode="parth", R=32, L_build=64, low_precision="bf16")
    ref = parth.searchPoints(qs[:200], k=10, L=60, prec="bf16")
    parth.set_tiles(4)
    got = parth.searchPoints(qs[:200], k=10, L=60, prec="bf16")
    assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1])


def test_replay_refused_after_updates_and_recapture_matches_oracle(vmn, orc, gpu):
    """A captured search holds the index state of its capture (entry point,
    tombstones). After a delete or a consolidation the replay is refused
    (VMN_ERR_STATE); a new capture searches the updated index exactly like an
    eager search and the oracle with the same tombstones."""
    import torch
    base = vmn.datasets.integer_mixture(6000, 128, k=24, seed=21, stream=0)
    qs = vmn.datasets.integer_mixture(256, 128, k=24, seed=21, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2),
                     low_precision="bf16")
    ctx = vmn.SearchContext(idx)
    q = torch.from_numpy(qs).to(gpu)
    ids = torch.empty((256, 10), dtype=torch.int32, device=gpu)
    d = torch.empty((256, 10), dtype=torch.float32, device=gpu)
    ctx.capture(q, ids, d, k=10, L=40, prec="bf16", rerank=20)
    st = torch.cuda.ExternalStream(ctx.stream)
    st.synchronize()
    ctx.replay()
    st.synchronize()
    # delete the entry point and a slice of rows: the old graph is refused
    rng = np.random.default_rng(4)
    kill = np.unique(np.concatenate([[idx.medoid], rng.choice(6000, 600, replace=False)]))
    idx.deleteIds(kill.astype(np.int32))
    with pytest.raises(vmn.VamanaError) as e:
        ctx.replay()
    assert e.value.code == -3
    idx.consolidate()
    with pytest.raises(vmn.VamanaError):
        ctx.replay()
    ctx.capture(q, ids, d, k=10, L=40, prec="bf16", rerank=20)
    for _ in range(3):
        ids.fill_(-7)
        torch.cuda.synchronize()
        ctx.replay()
        st.synchronize()
        e_ids, e_d = idx.searchPoints(qs, k=10, L=40, prec="bf16", rerank=20)
        assert np.array_equal(ids.cpu().numpy(), e_ids) and np.array_equal(d.cpu().numpy(), e_d)
    dead = idx.deleted.astype(np.uint8)
    oi, od, _, _ = orc.fresh_search(base, idx.adjacency, qs, idx.medoid, L=40, out_n=10,
                                    deleted=dead)
    assert np.array_equal(e_ids, oi) and np.array_equal(e_d, od)
    assert not np.isin(e_ids[e_ids >= 0], kill).any()
    ctx.close()


def test_context_rejects_bad_tensors(vmn, gpu):
    """Search-context tensors are checked before any device access: wrong
    dtype, width, device, layout or an output too small raise VamanaError."""
    import torch
    base = vmn.datasets.integer_mixture(2000, 128, k=8, seed=3, stream=0)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=32, alphas=(1.2,))
    ctx = vmn.SearchContext(idx)
    q = torch.zeros((64, 128), dtype=torch.float32, device=gpu)
    ids = torch.empty((64, 10), dtype=torch.int32, device=gpu)
    d = torch.empty((64, 10), dtype=torch.float32, device=gpu)
    bad = [
        (q.double(), ids, d), (q.cpu(), ids, d), (q[:, :96], ids, d), (q.t(), ids, d),
        (q, ids[:32], d), (q, ids, d[:, :5]), (q, ids.float(), d), (q, ids.cpu(), d),
    ]
    for qq, ii, dd in bad:
        with pytest.raises(vmn.VamanaError):
            ctx.search(qq, ii, dd, k=10, L=32)
        with pytest.raises(vmn.VamanaError):
            ctx.capture(qq, ii, dd, k=10, L=32)
    ctx.search(q, ids, d, k=10, L=32)
    torch.cuda.synchronize()
    ctx.close()
