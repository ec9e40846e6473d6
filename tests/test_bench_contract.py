"""bench.py's host-side contract (no GPU): the one-GPU refusal and the lookup
that attaches the committed HBM traffic figure only to the workload it was
measured on."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_traffic_lookup_matches_its_workload_only(bench):
    with open(os.path.join(ROOT, "profiles", "search_1m_hbm_traffic.json")) as f:
        prof = json.load(f)
    n, nq = prof["n"], prof["nq"]
    for form, e in prof["forms"].items():
        k, L, rr = e["kernel"], e["L"], e["rerank"]
        assert bench.measured_traffic(form, n, nq, L, k, rr) == e["traffic_bytes"]
        # another beam-list capacity instance runs the same traversal
        k256 = k.replace("<128, 64,", "<128, 256,")
        assert bench.measured_traffic(form, n, nq, L, k256, rr) == e["traffic_bytes"]
        # any other workload parameter: no figure
        assert bench.measured_traffic(form, n, nq, L, k, rr + 20) is None
        assert bench.measured_traffic(form, n, nq, L + 10, k, rr) is None
        assert bench.measured_traffic(form, n * 10, nq, L, k, rr) is None
        assert bench.measured_traffic(form, n, nq * 2, L, k, rr) is None
        assert bench.measured_traffic(form, n, nq, L, k.replace("<128,", "<96,"), rr) is None
    # the re-rank depths recorded are the bench defaults of each form
    assert {f: e["rerank"] for f, e in prof["forms"].items()} == {
        "fp32": 0, "bf16+rerank": 20, "fp8+rerank": 40}
    sys.argv, argv = ["bench.py"], sys.argv
    try:
        a = bench.parse()
    finally:
        sys.argv = argv
    assert (a.rerank, a.rerank_fp8) == (20, 40)
