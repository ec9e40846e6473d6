"""Row precision of the build and insert searches.

While every row is an integer vector with |v| <= 256 and squared norm <= 2^22,
the build and insert searches read the BF16 copies of the rows: every BF16
distance then equals the FP32 one bit for bit, so the graph must be the
oracle's either way. One row outside that set switches the index to FP32
traversal for good. Each case checks which traversal ran and that the graph
is the oracle's (vamana_impl.cuh:235, :325 always search FP32 rows)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mix(vmn, n, seed, stream=0):
    return vmn.datasets.integer_mixture(n, 128, k=32, seed=seed, stream=stream)


def _build(vmn, orc, v, capacity=0):
    idx = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=capacity)
    oadj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    assert idx.medoid == med
    assert np.array_equal(idx.adjacency, oadj)
    return idx, oadj, med


def test_gpu_build_traversal_on_both_sides_of_the_bounds(vmn, orc, gpu):
    v = _mix(vmn, 3000, seed=21)
    idx, _, _ = _build(vmn, orc, v)
    assert idx.build_precision == "bf16"
    # squared norm exactly 2^22: 64 values of 256 (still exact)
    w = v.copy()
    w[7] = 0.0
    w[7, :64] = 256.0
    idx, _, _ = _build(vmn, orc, w)
    assert idx.build_precision == "bf16"
    # squared norm 2^22 + 1
    w[7, 64] = 1.0
    idx, _, _ = _build(vmn, orc, w)
    assert idx.build_precision == "fp32"
    # |v| > 256 on one value of one row
    w = v.copy()
    w[2999, 5] = 257.0
    idx, _, _ = _build(vmn, orc, w)
    assert idx.build_precision == "fp32"
    # a non-integer value (differences of half-integers are integers, so the
    # FP32 sums stay exact and the oracle still fixes the graph)
    w = v + np.float32(0.5)
    idx, _, _ = _build(vmn, orc, w)
    assert idx.build_precision == "fp32"


def test_gpu_insert_switches_traversal_once(vmn, orc, gpu):
    v = _mix(vmn, 4000, seed=22)
    v[3300:3500] += np.float32(0.5)
    idx, oadj, med = _build(vmn, orc, v[:3000], capacity=4000)
    full = np.full((4000, 32), -1, np.int32)
    full[:3000] = oadj
    size = 3000
    for end, want in ((3300, "bf16"), (3500, "fp32"), (4000, "fp32")):
        idx.insertPoints(v[size:end])
        orc.fresh_insert(v[:end], size, full[:end], med, L=64, alpha=1.2)
        size = end
        assert idx.build_precision == want, end
        assert np.array_equal(idx.adjacency, full[:end]), end
    q = _mix(vmn, 40, seed=22, stream=1)
    gi, gd = idx.searchPoints(q, k=10, L=48)
    oi, od, _, _ = orc.fresh_search(v, full, q, med, L=48)
    assert np.array_equal(gi, oi) and np.array_equal(gd, od)


def test_gpu_prune_tiers_match_oracle(vmn, orc, gpu):
    """The FRESH out-prune of a batch that fills the GPU (at least 16 items
    per CU) runs in two tiers: a 256-entry candidate list first, then the full
    list for the nodes listed as not fitting. L_build = 300 makes every
    visited set overflow the first tier. A hub point (the centre of the data,
    kept by every other node) gets thousands of reverse candidates, more than
    the reverse prune's cap M = 500 keeps, and goes first in its work order.
    The graph must still be the oracle's, in a build and an insert batch."""
    rng = np.random.default_rng(3)
    c = rng.integers(60, 120, 128)
    v = np.clip(c + rng.integers(-6, 7, (5000, 128)), 0, 255).astype(np.float32)
    v[0] = c
    idx = vmn.Vamana(v, mode="fresh", R=24, L_build=300, alphas=(1.0, 1.2), capacity=9500)
    adj, med = orc.fresh_build(v, R=24, L=300, alphas=(1.0, 1.2))
    assert idx.medoid == med
    assert (adj == 0).sum() > 500  # the hub's reverse candidates exceed M
    assert np.array_equal(idx.adjacency, adj)
    w = np.clip(c + rng.integers(-6, 7, (4500, 128)), 0, 255).astype(np.float32)
    full = np.full((9500, 24), -1, np.int32)
    full[:5000] = adj
    idx.insertPoints(w)
    orc.fresh_insert(np.concatenate([v, w]), 5000, full, med, L=300, alpha=1.2)
    assert np.array_equal(idx.adjacency, full)


def test_gpu_chunked_pass_equals_one_batch(vmn, orc, gpu):
    """A FRESH pass searched and out-pruned in chunks (build_chunk), with the
    chunks' rows held in a shadow adjacency until the pass ends, gives the
    one-batch graph: every search still sees the graph from before the pass.
    Checked for a two-pass build (4 chunks) and an insert batch (3 chunks)."""
    v = _mix(vmn, 4000, seed=31)
    w = _mix(vmn, 2500, seed=31, stream=2)
    one = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=6500)
    chk = vmn.Vamana(v, mode="fresh", R=32, L_build=64, alphas=(1.0, 1.2), capacity=6500,
                     build_chunk=1000)
    oadj, med = orc.fresh_build(v, R=32, L=64, alphas=(1.0, 1.2))
    assert chk.medoid == one.medoid == med
    assert np.array_equal(chk.adjacency, oadj) and np.array_equal(one.adjacency, oadj)
    one.insertPoints(w)
    chk.insertPoints(w)
    assert np.array_equal(chk.adjacency, one.adjacency)
    full = np.full((6500, 32), -1, np.int32)
    full[:4000] = oadj
    orc.fresh_insert(np.concatenate([v, w]), 4000, full, med, L=64, alpha=1.2)
    assert np.array_equal(chk.adjacency, full)
