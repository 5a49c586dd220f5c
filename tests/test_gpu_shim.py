"""The drop-in C++ headers (include/hashinator, include/splitvector) on the GPU:
runs tests/cpp/_bin/test_shim, built by __graft_entry__.build()."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "_bin", "test_shim")


def test_cpp_dropin_headers(gpu):
    assert os.path.exists(BIN), "tests/cpp/_bin/test_shim missing: run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 failure(s)" in r.stdout


def test_bucket_residency_results_agree(gpu):
    """Device-resident and unified-memory bucket arrays give the same results
    (the timings it prints are the reason the table lives in device memory)."""
    exe = os.path.join(os.path.dirname(BIN), "residency")
    r = subprocess.run([exe, "18"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "residency ok" in r.stdout, r.stdout + r.stderr


def test_dropin_compaction_on_device_vectors(gpu):
    """split::tools::copy_if on device-resident SplitVectors with a user rule
    gives the raw C-ABI compaction's elements, order and count (2^22 pairs,
    Bernoulli 0.3, all kept and none kept)."""
    exe = os.path.join(os.path.dirname(BIN), "dropin_compaction")
    for args in (["22", "0.3"], ["20", "1.0"], ["20", "0.0"]):
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr[-2000:])
        assert r.returncode == 0 and '"same_as_abi": true' in r.stdout, args
