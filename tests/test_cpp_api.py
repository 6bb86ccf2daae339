"""Runs the C++ tests of the drop-in API (tests/cpp/test_*.cpp, built by
`python -m qtnh_amd.build`): host-only cases here, all cases on the GPU."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "bin")
TESTS = ["test_indexing", "test_tensor", "test_network", "test_print"]


def _run(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail("%s is not built (run `python -m qtnh_amd.build`)" % exe)
    r = subprocess.run([exe] + list(args), stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    return r.stdout


@pytest.mark.parametrize("name", TESTS)
def test_cpp_host_cases(name):
    _run(name, "--cpu-only")


@pytest.mark.gpu
@pytest.mark.parametrize("name", TESTS)
def test_cpp_all_cases(name):
    out = _run(name)
    assert " 0 failed" in out
