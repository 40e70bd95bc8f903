"""The C++ host API (include/ipc_amd.hpp) on a device: runs the compiled
test program tests/cpp/_build/test_api (built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from conftest import ROOT

EXE = os.path.join(ROOT, "tests", "cpp", "_build", "test_api")


@pytest.mark.gpu
def test_cpp_api_known_answers():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout[-4000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "0 failure(s)" in r.stdout


def test_cpp_api_compiles():
    """the header compiles standalone (no device needed)"""
    src = os.path.join(ROOT, "tests", "cpp", "test_api.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
