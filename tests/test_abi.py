"""The C ABI library: loads, exports every declared entry point, validates
arguments on the host (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vamana_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(vmn_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("vmn_create", "vmn_build", "vmn_search", "vmn_groundtruth", "vmn_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol(vmn):
    L = vmn.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_param_validation_without_gpu(vmn):
    L = vmn.lib()
    p = vmn.default_params(vmn.MODE_PARTH)
    p.R = 65  # rows wider than a wavefront are rejected before touching the device
    h = ctypes.c_void_p()
    rc = L.vmn_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and b"R" in L.vmn_last_error()
    p = vmn.default_params(vmn.MODE_FRESH)
    p.P = 10_000
    assert L.vmn_create(ctypes.byref(p), ctypes.byref(h)) == -1


def test_default_params_match_reference(vmn):
    p = vmn.default_params(vmn.MODE_PARTH)
    assert (p.R, p.L_build, p.passes, p.seed, p.visited_cap, p.max_hops) == (32, 64, 2, 1234, 128, 200)
    assert abs(p.alpha[0] - 1.2) < 1e-7
    f = vmn.default_params(vmn.MODE_FRESH)
    assert (f.R, f.P, f.M) == (64, 600, 500) and abs(f.alpha[0] - 1.5) < 1e-7


def test_null_arguments_are_errors(vmn):
    L = vmn.lib()
    assert L.vmn_build(None, None, 10, 0) == -1
    assert L.vmn_search(None, None, 1, 1, 1, None, None, 0) == -1


def test_file_readers_reject_bad_files_without_gpu(vmn, tmp_path):
    """Every reader checks the file before it touches the device."""
    err = vmn.VamanaError
    with pytest.raises(err) as e:
        vmn.Vamana.load(str(tmp_path / "missing.vmn"))
    assert e.value.code == -5
    junk = tmp_path / "junk.vmn"
    junk.write_bytes(b"not an index file at all")
    with pytest.raises(err) as e:
        vmn.Vamana.load(str(junk))
    assert e.value.code == -5 and "not a version-1 or version-2 index file" in str(e.value)
    # parth graph whose header promises more than the file holds
    bad = tmp_path / "short_graph.bin"
    bad.write_bytes(struct_pack(100, 128, 32) + b"\0" * 1000)
    with pytest.raises(err) as e:
        vmn.Vamana.load_parth_graph(str(bad))
    assert e.value.code == -5 and "header" in str(e.value)
    # FreshVamana records: 772-byte records at dim 128, R 64
    rec = tmp_path / "records.bin"
    rec.write_bytes(b"\0" * (772 * 3 + 5))
    with pytest.raises(err) as e:
        vmn.Vamana.load_fresh_graph(str(rec), dim=128, R=64)
    assert e.value.code == -5 and "772-byte records" in str(e.value)


def struct_pack(*ints):
    import struct
    return struct.pack("<%di" % len(ints), *ints)


def test_read_workload_without_gpu(vmn, tmp_path):
    """vmn_read_workload against the reference's sift workload (records of
    {int32 type; int32 id; float vec[128]}, parth/src/common.h:63-67), a
    trailing partial record (ignored, as engine.cu:364-370 stops there) and a
    missing file."""
    import numpy as np
    w = np.load(os.path.join(ROOT, "tests", "golden", "parth_workloads.npz"))
    t, i, v = w["sift_types"], w["sift_ids"], w["sift_vecs"]
    rec = np.zeros(len(t), dtype=[("t", "<i4"), ("i", "<i4"), ("v", "<f4", (128,))])
    rec["t"], rec["i"], rec["v"] = t, i, v
    path = tmp_path / "sift_workload.bin"
    path.write_bytes(rec.tobytes() + b"\x02\x00\x00\x00\xff")
    assert os.path.getsize(path) == 520000 + 5
    rt, ri, rv = vmn.read_workload(str(path), 128)
    assert np.array_equal(rt, t) and np.array_equal(ri, i) and np.array_equal(rv, v)
    assert np.bincount(rt).tolist() == [90, 183, 727]
    with pytest.raises(vmn.VamanaError):
        vmn.read_workload(str(tmp_path / "missing.bin"), 128)


def compile_example(name, out):
    """A C++ driver in cs5013-project_amd/examples/, built against
    include/vamana_mi355x.h and libvmn.so with the host compiler."""
    import subprocess
    pkg = os.path.join(ROOT, "cs5013-project_amd")
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I",
           os.path.join(ROOT, "include"), os.path.join(pkg, "examples", name + ".cpp"),
           "-L", pkg, "-lvmn", "-Wl,-rpath," + pkg, "-Wl,-rpath-link,/opt/rocm/lib", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


@pytest.mark.parametrize("name", ["parth_engine", "fresh_vamana", "bang_search"])
def test_cpp_drivers_build_against_the_c_abi(tmp_path, name):
    import subprocess
    exe = compile_example(name, str(tmp_path / name))
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage:" in r.stdout


def test_device_tensor_checks_without_gpu(vmn):
    """The Python mirror refuses, before any library call, a tensor it would
    hand over as a raw device pointer unless it is a contiguous CUDA tensor
    of the right dtype, device, width and size (here: CPU tensors and numpy
    arrays, which no device pointer may come from)."""
    import importlib

    import numpy as np
    import torch
    v = importlib.import_module("cs5013-project_amd.vamana")
    for x in (torch.zeros((4, 128)), np.zeros((4, 128), np.float32),
              torch.zeros((4, 128), dtype=torch.float64)):
        with pytest.raises(vmn.VamanaError):
            v._device_tensor(x, "queries", np.float32, 0, width=128)
