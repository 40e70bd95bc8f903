"""The reference-side shim (ipc-toolkit_amd/integration/ipc_gpu.hpp) is C++
that a maintainer compiles into IPC Toolkit.  Here it is compiled against a
stand-in of the reference interface it overrides and calls (the abstract
BroadPhase, the collision vectors, the CCD parameters; Eigen's dense and
sparse types as far as used) and linked with libipcg.so.  Without a device
the program must refuse to run (no CPU fallback); on a device it drives every
BroadPhase override, the collision build, the barrier and the step size
through the name dispatch."""
import os
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "integration", "shim_check.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "shim_check")


def compile_shim(out=EXE):
    lib = os.path.join(ROOT, "ipc-toolkit_amd", "lib")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Werror",
           "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "ipc-toolkit_amd", "integration"),
           "-I" + os.path.join(ROOT, "tests", "cpp", "integration", "stand_in"),
           SRC, "-L" + lib, "-lipcg", "-Wl,-rpath," + lib, "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_shim_compiles_and_refuses_without_device(tmp_path):
    """compiles, links and runs: without a device it must refuse (exit 3),
    with one it must pass"""
    exe = str(tmp_path / "shim_check")
    compile_shim(exe)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    if "no device" in r.stdout:
        assert r.returncode == 3, (r.returncode, r.stdout, r.stderr)
    else:
        assert r.returncode == 0 and "0 failure(s)" in r.stdout, (r.stdout, r.stderr)


@pytest.mark.gpu
def test_shim_on_device():
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is not built (run __graft_entry__.build())")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout


def test_eigen_conversions_of_the_cpp_mirror(tmp_path):
    """include/ipc_amd.hpp converts to and from Eigen where it is available
    (here: the stand-in, which follows Eigen's storage conventions)"""
    exe = str(tmp_path / "eigen_check")
    src = os.path.join(ROOT, "tests", "cpp", "integration", "eigen_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror",
                        "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "tests", "cpp", "integration", "stand_in"),
                        src, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "0 failure(s)" in r.stdout, r.stdout
