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
        # and so does the other tile shape
        k4 = k.replace(", 2>(", ", 4>(")
        assert k4 != k and bench.measured_traffic(form, n, nq, L, k4, rr) == e["traffic_bytes"]
        # any other workload parameter: no figure
        assert bench.measured_traffic(form, n, nq, L, k, rr + 20) is None
        assert bench.measured_traffic(form, n, nq, L + 10, k, rr) is None
        assert bench.measured_traffic(form, n * 10, nq, L, k, rr) is None
        assert bench.measured_traffic(form, n, nq * 2, L, k, rr) is None
        assert bench.measured_traffic(form, n, nq, L, k.replace("<128,", "<96,"), rr) is None
    # the re-rank depths recorded are the bench defaults of each form
    assert {f: e["rerank"] for f, e in prof["forms"].items()} == {
        "fp32": 0, "bf16+rerank": 40, "fp8+rerank": 40}
    sys.argv, argv = ["bench.py"], sys.argv
    try:
        a = bench.parse()
    finally:
        sys.argv = argv
    assert (a.rerank, a.rerank_fp8) == (40, 40)


def test_roofline_is_stated_on_the_timed_steps(bench):
    """The roofline's `achieved` divides the algorithmic bytes by the mean
    launch duration measured in the overlapped timed steps; the counter HBM
    rate is traffic over ms_per_step; the one-batch-at-a-time figures sit
    under `serial`; the counter ratios come from the committed profile."""
    with open(os.path.join(ROOT, "profiles", "search_1m_hbm_traffic.json")) as f:
        prof = json.load(f)
    e = prof["forms"]["bf16+rerank"]
    t = {"achieved_GBs": 4000.0, "traffic": e["traffic_bytes"], "kernel": e["kernel"],
         "launch": {"waves_per_cu": 32, "list_cap": 0}, "launch_ms": 1.6,
         "bytes_per_launch": e["algorithmic_bytes"], "ms_per_step": 0.85,
         "kernel_ms_serial": 1.25, "achieved_GBs_serial": 6000.0,
         "counters": bench.measured_counters("bf16+rerank", prof["n"], prof["nq"], e["L"],
                                             e["kernel"], e["rerank"], 1.6)}
    r = bench.roofline(t)
    assert r["frac"] == 0.5 and r["serial"]["frac"] == 0.75
    assert abs(r["hbm_GBs_per_step"] - e["traffic_bytes"] / 0.85e-3 / 1e9) < 0.1
    assert r["l2_hit_rate"] == e["l2_hit_rate"] and r["mfma_busy"] == e["mfma_busy"]
    assert abs(r["mfma_tflops"] - e["mfma_flop"] / 1.6e-3 / 1e12) < 0.1
    assert bench.measured_counters("bf16+rerank", prof["n"], prof["nq"], e["L"] + 10,
                                   e["kernel"], e["rerank"], 1.6) == {}
