"""HIP graph capture / replay of a batched search, search contexts, launch
shapes: results identical to eager searches."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_graph_replay_matches_eager(vmn, gpu):
    import torch
    base = vmn.datasets.integer_mixture(8000, 128, k=32, seed=9, stream=0)
    qa = vmn.datasets.integer_mixture(300, 128, k=32, seed=9, stream=1)
    qb = vmn.datasets.integer_mixture(300, 128, k=32, seed=9, stream=3)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), low_precision="bf16")
    ea, da = idx.searchPoints(qa, k=10, L=48, prec="bf16", rerank=32)
    eb, db = idx.searchPoints(qb, k=10, L=48, prec="bf16", rerank=32)
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
