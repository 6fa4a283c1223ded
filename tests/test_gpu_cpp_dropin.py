"""Runs the C++ drop-in tests (tests/cpp/test_fem_hip.cpp): the fem::hip layer
over the C ABI, exercised the way the reference's own CUDA tests exercise fem::cuda."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "test_fem_hip")


def test_cpp_dropin(gpu):
    assert os.path.exists(BIN), "tests/cpp/test_fem_hip is not built (run __graft_entry__.build())"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr


REF_BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "test_fem_hip_ref")


def test_cpp_dropin_on_reference_headers(gpu):
    """The same drop-in tests compiled against the reference's own headers and
    sources (built where the reference is mounted; the binary travels with the tree)."""
    if not os.path.exists(REF_BIN):
        pytest.skip("oracle/_ref/test_fem_hip_ref is built only where the reference is mounted")
    r = subprocess.run([REF_BIN], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
