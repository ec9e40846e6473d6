"""vmn_search_noalloc, the search entry point that allocates nothing, called
through the C ABI against vmn_search_ex on the same device buffers: two calls
in a row with the second batch smaller, at a width that needs no padding (128)
and one that does (100), in FP32 and in BF16 with an FP32 re-rank. The results
must be identical. Batches whose padded rows exceed the static scratch, and
calls with staged inserts pending, are refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# include/vamana_mi355x.h
VMN_PREC_FP32, VMN_PREC_BF16, VMN_PREC_FP8, VMN_MEM_DEVICE = 0, 1, 2, 1


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.parametrize("dim", [128, 100])
def test_noalloc_search_equals_search_ex(vmn, gpu, dim):
    import torch
    L = vmn.lib()
    base = vmn.datasets.integer_mixture(6000, dim, k=16, seed=41, stream=0)
    qs = vmn.datasets.integer_mixture(3000, dim, k=16, seed=41, stream=1)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=48, alphas=(1.0, 1.2),
                     low_precision="both", capacity=7000)
    forms = ((VMN_PREC_FP32, 0), (VMN_PREC_BF16, 20), (VMN_PREC_FP8, 40))
    for step, (prec, rr) in enumerate(forms + ((VMN_PREC_FP8, 40),)):
        if step == len(forms):
            # an update: the FP8 norms kept beside the adjacency are laid out
            # again by the next search, here the allocation-free one
            idx.insertPoints(vmn.datasets.integer_mixture(500, dim, k=16, seed=41, stream=2))
        for nq in (3000, 700):
            print(f"noalloc search dim {dim} prec {prec} nq {nq}", flush=True)
            q = torch.from_numpy(qs[:nq]).to(gpu)
            ids_a = torch.full((nq, 10), -7, dtype=torch.int32, device=gpu)
            d_a = torch.full((nq, 10), -7.0, dtype=torch.float32, device=gpu)
            ids_b = torch.full((nq, 10), -9, dtype=torch.int32, device=gpu)
            d_b = torch.full((nq, 10), -9.0, dtype=torch.float32, device=gpu)
            torch.cuda.synchronize()
            rc = L.vmn_search_noalloc(idx._h, _ptr(q), nq, 10, 40, prec, rr, _ptr(ids_a),
                                      _ptr(d_a))
            assert rc == 0, L.vmn_last_error()
            idx.synchronize()
            rc = L.vmn_search_ex(idx._h, _ptr(q), nq, 10, 40, prec, rr, _ptr(ids_b), _ptr(d_b),
                                 VMN_MEM_DEVICE)
            assert rc == 0, L.vmn_last_error()
            idx.synchronize()
            assert torch.equal(ids_a, ids_b) and torch.equal(d_a, d_b)
            assert int(ids_a.min()) >= 0
    if dim % 32 == 0:
        # no padding: no limit on the batch
        q = torch.from_numpy(np.tile(qs, (7, 1))).to(gpu)
        ids = torch.empty((len(q), 10), dtype=torch.int32, device=gpu)
        torch.cuda.synchronize()
        assert L.vmn_search_noalloc(idx._h, _ptr(q), len(q), 10, 40, 0, 0, _ptr(ids), None) == 0
        idx.synchronize()
        ref, _ = idx.searchPoints(qs, k=10, L=40)
        assert np.array_equal(ids.cpu().numpy()[:len(qs)], ref)
    else:
        # padded rows beyond the scratch are refused before any device work
        big = 8192
        q = torch.zeros((big, dim), dtype=torch.float32, device=gpu)
        ids = torch.empty((big, 10), dtype=torch.int32, device=gpu)
        rc = L.vmn_search_noalloc(idx._h, _ptr(q), big, 10, 40, 0, 0, _ptr(ids), None)
        assert rc == -4 and b"scratch" in L.vmn_last_error()


def test_noalloc_search_refused_with_staged_inserts(vmn, gpu):
    import torch
    L = vmn.lib()
    base = vmn.datasets.integer_mixture(3000, 128, k=8, seed=5, stream=0)
    idx = vmn.Vamana(base, mode="fresh", R=32, L_build=32, alphas=(1.2,), capacity=4000)
    idx.insert_staged(base[:10] + 1.0)
    q = torch.from_numpy(base[:16]).to(gpu)
    ids = torch.empty((16, 10), dtype=torch.int32, device=gpu)
    torch.cuda.synchronize()
    assert L.vmn_search_noalloc(idx._h, _ptr(q), 16, 10, 32, 0, 0, _ptr(ids), None) == -3
