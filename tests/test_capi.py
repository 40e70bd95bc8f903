"""The C ABI library loads and exports every entry point include/ipcg.h
declares (no device needed), and its wrapper fails loudly without a GPU."""
import os
import re
import ctypes

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "ipcg.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ipcg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("ipcg_ctx_create", "ipcg_mesh_upload", "ipcg_positions_upload",
                 "ipcg_broad_phase", "ipcg_collisions_build", "ipcg_barrier_energy",
                 "ipcg_barrier_gradient", "ipcg_barrier_hessian",
                 "ipcg_compute_collision_free_stepsize", "ipcg_is_step_collision_free"):
        assert must in names


def test_library_exports_every_declared_symbol(capi):
    lib = capi.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == declared()


def test_symbols_are_c_linkage(capi):
    out = os.popen(f"nm -D --defined-only {capi.LIB_PATH}").read()
    exported = set(re.findall(r" T (ipcg_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared()) <= exported


def test_no_device_fails_loudly(capi):
    """Without a HIP device the context cannot be created: the product path
    raises instead of falling back to any CPU code."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(capi.IPCGError):
        capi.Context(0)
