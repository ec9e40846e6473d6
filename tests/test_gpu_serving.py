"""Continuous-batching server: every result of a mixed, batched run equals the
same request run alone on an index replayed to the same state (bit-exact:
each query is traversed on its own wavefront and updates are ordered)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _trace(vmn, rng, n_base, dim, nreq):
    pool = vmn.datasets.integer_mixture(4000, dim, k=32, seed=17, stream=1)
    ins = vmn.datasets.integer_mixture(4000, dim, k=32, seed=17, stream=2)
    reqs, ins_off, alive = [], 0, list(range(n_base))
    for _ in range(nreq):
        r = rng.random()
        size = int(np.exp(rng.uniform(0, np.log(64))))
        if r < 0.8:
            reqs.append((vmn.REQ_SEARCH, pool[rng.integers(0, 4000, size)]))
        elif r < 0.9:
            reqs.append((vmn.REQ_INSERT, ins[ins_off:ins_off + size]))
            ins_off += size
        else:
            pick = rng.choice(len(alive), size, replace=False)
            ids = np.array([alive[i] for i in pick], np.int32)
            alive = [a for i, a in enumerate(alive) if i not in set(pick.tolist())]
            reqs.append((vmn.REQ_DELETE, ids))
    return reqs


@pytest.mark.parametrize("prec,rr", [("fp32", 0), ("bf16", 32)])
def test_server_mixed_equals_alone(vmn, gpu, prec, rr):
    dim, n = 128, 4000
    base = vmn.datasets.integer_mixture(n, dim, k=32, seed=17, stream=0)
    kw = dict(mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=n + 4000,
              low_precision="bf16")
    rng = np.random.default_rng(3)
    reqs = _trace(vmn, rng, n, dim, 300)
    idx = vmn.Vamana(base, **kw)
    srv = vmn.Server(idx, k=10, L=48, prec=prec, rerank=rr, max_batch=512, n_slots=3)
    tickets = [srv.submit(kind, data) for kind, data in reqs]
    out = {}
    while len(out) < len(tickets):
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
    st = srv.stats()
    srv.close()
    # the engine's counters account for every request
    n_search = sum(len(d) for k, d in reqs if k == vmn.REQ_SEARCH)
    assert st["queries"] == n_search and 1 <= st["batches"] <= n_search
    assert st["max_batch_queries"] <= 512
    assert st["inserts"] == sum(k == vmn.REQ_INSERT for k, _ in reqs)
    assert st["deletes"] == sum(k == vmn.REQ_DELETE for k, _ in reqs)
    assert st["insert_s"] > 0 and st["max_update_s"] <= st["insert_s"] + st["delete_s"]
    srv2_stats = srv_reset_check(vmn, base, kw)
    assert all(v == 0 for v in srv2_stats.values())
    # replay alone, in arrival order, on an identically built index
    ref = vmn.Vamana(base, **kw)
    for t, (kind, data) in zip(tickets, reqs):
        if kind == vmn.REQ_SEARCH:
            ids, d = ref.searchPoints(data, k=10, L=48, prec=prec, rerank=rr)
            assert np.array_equal(out[t]["ids"], ids) and np.array_equal(out[t]["dists"], d)
        elif kind == vmn.REQ_INSERT:
            assert np.array_equal(out[t]["ids"], ref.insertPoints(data))
        else:
            ref.deleteIds(data)
    assert np.array_equal(idx.adjacency, ref.adjacency)
    assert all(r["t_done"] >= r["t_arrival"] for r in out.values())


def test_server_admission_by_memory(vmn, gpu):
    base = vmn.datasets.integer_mixture(2000, 128, k=16, seed=1, stream=0)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=64, alphas=(1.2,))
    srv = vmn.Server(idx, k=10, L=32, max_batch=64, n_slots=1, max_queued_bytes=64 * 128 * 4)
    q = base[:64]
    t1 = srv.submit(vmn.REQ_SEARCH, q)
    rejected = 0
    for _ in range(50):
        try:
            srv.submit(vmn.REQ_SEARCH, q)
        except vmn.VamanaError:
            rejected += 1
    while srv.pending():
        pass
    assert rejected > 0
    srv.close()


def srv_reset_check(vmn, base, kw):
    """vmn_server_reset_stats zeroes every counter, maxima included."""
    idx = vmn.Vamana(base[:2000], **kw)
    srv = vmn.Server(idx, k=10, L=32)
    t = srv.submit(vmn.REQ_SEARCH, base[:5])
    while srv.poll(t) is None:
        pass
    assert srv.stats()["queries"] == 5
    srv.reset_stats()
    st = srv.stats()
    srv.close()
    return st
