"""Vector and ground-truth files through the C ABI (host-only I/O, no GPU):
fvecs / ivecs / bvecs / bin round trips, the reference's own
data/sift10k/siftsmall_query.bin (a header that claims 10,000 vectors for
100) and siftsmall_groundtruth.bin (ids only, written back byte for byte),
and a DiskANN truthset with distances."""
import hashlib
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _digest(rel):
    return json.load(open(os.path.join(GOLDEN, "file_digests.json")))[rel]


def _sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.mark.parametrize("fmt,dt", [("fvecs", np.float32), ("ivecs", np.int32),
                                    ("bvecs", np.uint8), ("bin", np.float32)])
def test_vector_file_round_trip(vmn, sift, tmp_path, fmt, dt):
    a = sift["base"][:300].astype(dt)
    path = str(tmp_path / f"v.{fmt}")
    vmn.write_vectors(path, a, fmt)
    rec = 8 + 4 * 128 * 300 if fmt == "bin" else 300 * (4 + a.itemsize * 128)
    assert os.path.getsize(path) == rec
    b = vmn.read_vectors(path, fmt)
    assert b.dtype == dt and np.array_equal(a, b)


def test_reference_query_bin_with_overstated_header(vmn, sift, tmp_path):
    q = sift["query"].astype(np.float32)
    path = tmp_path / "siftsmall_query.bin"
    path.write_bytes(np.array([10000, 128], np.uint32).tobytes() + q.tobytes())
    d = _digest("data/sift10k/siftsmall_query.bin")
    assert os.path.getsize(path) == d["bytes"] and _sha(path) == d["sha256"]
    got = vmn.read_vectors(str(path), "bin")
    assert got.shape == (100, 128) and np.array_equal(got, q)


def test_reference_truthset_without_distances_written_back(vmn, sift, tmp_path):
    """siftsmall_groundtruth.bin holds the 100 x 100 ids of
    siftsmall_groundtruth.ivecs and no distances."""
    path = str(tmp_path / "siftsmall_groundtruth.bin")
    vmn.write_truthset(path, sift["gt100"])
    d = _digest("data/sift10k/siftsmall_groundtruth.bin")
    assert os.path.getsize(path) == d["bytes"] and _sha(path) == d["sha256"]
    ids, dists = vmn.read_truthset(path)
    assert dists is None and np.array_equal(ids, sift["gt100"])
    ivecs = str(tmp_path / "gt.ivecs")
    vmn.write_vectors(ivecs, sift["gt100"], "ivecs")
    assert np.array_equal(vmn.read_vectors(ivecs, "ivecs"), sift["gt100"])


def test_truthset_with_distances(vmn, sift, tmp_path):
    path = str(tmp_path / "gt.bin")
    vmn.write_truthset(path, sift["diskann_gt_ids"], sift["diskann_gt_dists"])
    assert os.path.getsize(path) == 8 + 2 * 100 * 10 * 4
    ids, dists = vmn.read_truthset(path)
    assert np.array_equal(ids, sift["diskann_gt_ids"])
    assert np.array_equal(dists, sift["diskann_gt_dists"])


def test_bad_vector_files(vmn, tmp_path):
    bad = tmp_path / "bad.fvecs"
    bad.write_bytes(np.array([4, 0, 0, 0, 0, 3], np.int32).tobytes())  # second record cut
    with pytest.raises(vmn.VamanaError):
        vmn.read_vectors(str(bad), "fvecs")
    with pytest.raises(vmn.VamanaError):
        vmn.read_truthset(str(tmp_path / "missing.bin"))
    odd = tmp_path / "odd.bin"
    odd.write_bytes(np.array([2, 2], np.uint32).tobytes() + bytes(12))
    with pytest.raises(vmn.VamanaError):
        vmn.read_vectors(str(odd), "bin")


def test_recall_with_and_without_ties(vmn, sift):
    """vmn_recall: found [1, 2, 9] against ground truth [1, 2, 3, 9, 8] whose
    third and fourth ids are tied: 2/3 plain; 3/3 when the tie extends the set
    (calculate_recall, bang_exact/src/main.cu:87-137)."""
    found = np.array([[1, 2, 9, -1]], np.int32)
    gt = np.array([[1, 2, 3, 9, 8]], np.int32)
    d = np.array([[0.0, 1.0, 2.0, 2.0, 3.0]], np.float32)
    assert vmn.recall_at_k(found, gt, 3) == pytest.approx(2 / 3)
    assert vmn.recall_at_k(found, gt, 3, gt_dists=d) == pytest.approx(1.0)
    # -1 results never count; repeated results count once
    assert vmn.recall_at_k(np.array([[-1, 1, 1]], np.int32), gt, 3) == pytest.approx(1 / 3)
    # the SIFT10K top-10 sets of siftsmall_groundtruth.ivecs and of the DiskANN
    # truthset (with its distances) agree for all 100 queries
    assert vmn.recall_at_k(sift["gt100"][:, :10], sift["diskann_gt_ids"], 10,
                           gt_dists=sift["diskann_gt_dists"]) == 1.0
