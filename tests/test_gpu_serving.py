"""Continuous-batching server: every result of a mixed, batched run equals the
same request run alone on an index replayed to the same state (bit-exact:
each query is traversed on its own wavefront and updates are ordered), with
periodic consolidation; on integer data the whole request stream equals the
oracle's replay bit for bit, and on the served configuration's unit-normalised
real-valued data the server's BF16 + FP32 re-rank results agree with the
oracle's FP32 search within the tolerance stated in that test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(srv, reqs):
    tickets = [srv.submit(kind, data) for kind, data in reqs]
    out = {}
    while len(out) < len(tickets):
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
    return tickets, out


def _replay_alone(vmn, ref, reqs, tickets, out, L, prec, rr, consolidate_every):
    """Each request alone, in arrival order, on an identically built index;
    deletes are consolidated at the same points as the server does."""
    since = 0
    for t, (kind, data) in zip(tickets, reqs):
        if kind == vmn.REQ_SEARCH:
            ids, d = ref.searchPoints(data, k=10, L=L, prec=prec, rerank=rr)
            assert np.array_equal(out[t]["ids"], ids) and np.array_equal(out[t]["dists"], d)
        elif kind == vmn.REQ_INSERT:
            assert np.array_equal(out[t]["ids"], ref.insertPoints(data))
        else:
            ref.deleteIds(data)
            since += len(data)
            if consolidate_every and since >= consolidate_every:
                ref.consolidate()
                since = 0


def _trace(vmn, rng, n_base, dim, nreq, gen=None):
    gen = gen or (lambda n, s: vmn.datasets.integer_mixture(n, dim, k=32, seed=17, stream=s))
    pool = gen(4000, 1)
    ins = gen(4000, 2)
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


def test_server_with_consolidation_equals_oracle_replay(vmn, orc, gpu):
    """Integer data (every distance exact), FP32 traversal, consolidation every
    60 deletes: every search result, every assigned id and the final graph
    equal the oracle's replay of the request stream in arrival order
    (orc_fresh_insert, tombstones, orc_consolidate and the live medoid)."""
    dim, n, R, Lb = 128, 3000, 32, 64
    base = vmn.datasets.integer_mixture(n, dim, k=32, seed=19, stream=0)
    rng = np.random.default_rng(8)
    reqs = _trace(vmn, rng, n, dim, 160,
                  gen=lambda m, s: vmn.datasets.integer_mixture(m, dim, k=32, seed=19, stream=s))
    idx = vmn.Vamana(base, mode="fresh", R=R, L_build=Lb, alphas=(1.0, 1.2), capacity=n + 4000)
    srv = vmn.Server(idx, k=10, L=40, max_batch=256, n_slots=3, consolidate_every=60)
    tickets, out = _run(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["consolidations"] >= 2 and st["failed_batches"] == 0
    full = np.full((n + 4000, R), -1, np.int32)
    full[:n], med = orc.fresh_build(base, R=R, L=Lb, alphas=(1.0, 1.2))
    vecs = np.zeros((n + 4000, dim), np.float32)
    vecs[:n] = base
    dele = np.zeros(n + 4000, np.uint8)
    size, since, n_cons = n, 0, 0
    print(f"oracle replay of {len(reqs)} requests", flush=True)
    for t, (kind, data) in zip(tickets, reqs):
        if kind == vmn.REQ_SEARCH:
            oi, od, _, _ = orc.fresh_search(vecs[:size], full[:size], data, med, L=40,
                                            deleted=dele[:size])
            assert np.array_equal(out[t]["ids"], oi) and np.array_equal(out[t]["dists"], od)
        elif kind == vmn.REQ_INSERT:
            m = len(data)
            vecs[size:size + m] = data
            orc.fresh_insert(vecs[:size + m], size, full[:size + m], med, L=Lb, alpha=1.2,
                             deleted=dele[:size + m])
            assert out[t]["ids"].tolist() == list(range(size, size + m))
            size += m
        else:
            dele[data] = 1
            since += len(data)
            if since >= 60:
                orc.consolidate(vecs[:size], full[:size], dele[:size], alpha=1.2)
                if dele[med]:
                    med = orc.medoid_live(vecs[:size], dele[:size])
                since, n_cons = 0, n_cons + 1
    assert n_cons == st["consolidations"]
    assert idx.size == size and idx.medoid == med
    assert np.array_equal(idx.adjacency, full[:size])


@pytest.mark.parametrize("prec,rr", [("bf16", 40), ("fp32", 0)])
def test_server_served_config_on_unit_data(vmn, orc, gpu, prec, rr):
    """The served configuration's form on its kind of data: unit-normalised
    96-d vectors of a 1,024-component mixture (sigma 0.1), FRESH graph, L = 64,
    BF16 rows with an FP32 re-rank of 40 (and the FP32 form), 10% updates with
    consolidation every 50 deletes.
    * Mixed and alone: bit-exact, as on integer data.
    * Against the oracle's FP32 search on the final graph, over 500 queries:
      the distances here are not exact, the GPU sums FP32 terms in another
      order than the oracle (and BF16 traverses on rounded rows), so near ties
      may order differently. Stated tolerance: mean top-10 overlap with the
      oracle's ids >= 0.998 (BF16) / 0.999 (FP32), and every id both return
      has the same distance within 5e-6 relative (FP32 rounding over 96
      terms). Measured: 0.9996 / 1.0 and 5.6e-7."""
    dim, n = 96, 6000
    gen = lambda m, s: vmn.datasets.unit_mixture(m, dim, k=1024, sigma=0.1, seed=23, stream=s)
    base = gen(n, 0)
    kw = dict(mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=n + 4000,
              low_precision="bf16")
    rng = np.random.default_rng(12)
    reqs = _trace(vmn, rng, n, dim, 240, gen=gen)
    idx = vmn.Vamana(base, **kw)
    assert idx.build_precision == "fp32"
    srv = vmn.Server(idx, k=10, L=64, prec=prec, rerank=rr, max_batch=512, n_slots=4,
                     consolidate_every=50)
    tickets, out = _run(srv, reqs)
    st = srv.stats()
    srv.close()
    assert st["consolidations"] >= 1 and st["failed_batches"] == 0
    ref = vmn.Vamana(base, **kw)
    _replay_alone(vmn, ref, reqs, tickets, out, 64, prec, rr, 50)
    assert np.array_equal(idx.adjacency, ref.adjacency)
    # against the oracle's FP32 search on the same final graph
    q = gen(500, 3)
    ids, d = idx.searchPoints(q, k=10, L=64, prec=prec, rerank=rr)
    vecs = idx.vectors
    oi, od, _, _ = orc.fresh_search(vecs, idx.adjacency, q, idx.medoid, L=64,
                                    deleted=idx.deleted.astype(np.uint8))
    overlap = np.mean([len(set(a.tolist()) & set(b.tolist())) / 10 for a, b in zip(ids, oi)])
    worst = 0.0
    for a, da, b, db in zip(ids, d, oi, od):
        common = {x: y for x, y in zip(b.tolist(), db.tolist())}
        for x, y in zip(a.tolist(), da.tolist()):
            if x in common:
                worst = max(worst, abs(y - common[x]) / max(common[x], 1e-12))
    print(f"served config {prec}: overlap with the oracle {overlap:.4f}, worst relative "
          f"distance difference {worst:.2e}", flush=True)
    assert overlap >= (0.998 if prec == "bf16" else 0.999)
    assert worst <= 5e-6
