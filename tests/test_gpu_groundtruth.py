"""The FP64 ground truth's MFMA-filtered path (an FP32 q.x filter with a
per-query error bound, then the direct FP64 distances of the kept rows) gives
the direct kernel's results bit for bit: ids and FP64 distances equal the
oracle's FP64 brute force, on integer and on real-valued rows, at every
padded width it serves, for k up to its limit; queries its bound cannot
settle (bulk ties from duplicated rows) go through the direct kernel."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(vmn, orc, base, q, k):
    ids, d = vmn.groundtruth(base, q, k)
    oi, od = orc.bruteforce_fp64(base, q, k)
    assert np.array_equal(ids, oi) and np.array_equal(d, od)
    return vmn.groundtruth_fallbacks()


@pytest.mark.parametrize("dim", [128, 96, 64, 32])
def test_filtered_groundtruth_real_valued_rows(vmn, orc, gpu, dim):
    base = vmn.datasets.unit_mixture(20000, dim, k=64, seed=41, stream=0)
    q = vmn.datasets.unit_mixture(400, dim, k=64, seed=41, stream=1)
    fb = _check(vmn, orc, base, q, 10)
    assert fb <= 8, fb  # the bound settles almost every query


def test_filtered_groundtruth_integer_rows_and_k_limits(vmn, orc, gpu):
    base = vmn.datasets.integer_mixture(30000, 128, k=32, seed=43, stream=0)
    q = vmn.datasets.integer_mixture(300, 128, k=32, seed=43, stream=1)
    for k in (1, 10, 64):
        _check(vmn, orc, base, q, k)
    # k above the filtered path's limit: the direct kernel, no fallback count
    assert _check(vmn, orc, base, q, 65) == 0


def test_filtered_groundtruth_falls_back_on_bulk_ties(vmn, orc, gpu):
    rng = np.random.default_rng(5)
    uniq = rng.integers(0, 4, size=(40, 128)).astype(np.float32)
    base = np.concatenate([np.repeat(uniq, 150, axis=0),
                           vmn.datasets.integer_mixture(6000, 128, k=16, seed=44, stream=0)])
    q = np.concatenate([uniq[:20], vmn.datasets.integer_mixture(60, 128, k=16, seed=44, stream=1)])
    fb = _check(vmn, orc, base, q, 10)
    # each duplicated query has 150 rows at distance 0: more than the filter recomputes
    assert fb >= 20, fb


def test_filtered_groundtruth_small_and_device_inputs(vmn, orc, gpu):
    import torch
    base = vmn.datasets.unit_mixture(5000, 96, k=16, seed=45, stream=0)
    q = vmn.datasets.unit_mixture(3, 96, k=16, seed=45, stream=1)
    _check(vmn, orc, base, q, 10)
    ids, d = vmn.groundtruth(torch.from_numpy(base).to(gpu), torch.from_numpy(q).to(gpu), 10)
    oi, od = orc.bruteforce_fp64(base, q, 10)
    assert np.array_equal(ids.cpu().numpy(), oi) and np.array_equal(d.cpu().numpy(), od)
