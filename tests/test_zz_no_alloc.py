"""ipcg_barrier_energy_noalloc (block partials in the library's per-device
static array, result written to caller-owned device memory) against
ipcg_barrier_energy (the same kernels over a context buffer), through the
C ABI: bit-identical energies on consecutive calls, the second over a
smaller collision set, and an empty set.  The output memory comes from the
HIP runtime the library itself loaded (torch bundles a runtime of its own,
which is not used here)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_MEMCPY_HOST_TO_DEVICE, HIP_MEMCPY_DEVICE_TO_HOST = 1, 2


def _hip(capi):
    """the libamdhip64 instance libipcg.so is linked against"""
    capi.load()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
    assert len(paths) == 1, paths
    h = C.CDLL(paths.pop())
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


class DeviceDoubles:
    def __init__(self, hip, n):
        self.hip, self.n = hip, n
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), 8 * n) == 0
        nan = np.full(n, np.nan)
        assert hip.hipMemcpy(self.p, nan.ctypes.data, 8 * n, HIP_MEMCPY_HOST_TO_DEVICE) == 0

    def at(self, k):
        return C.c_void_p(self.p.value + 8 * k)

    def get(self):
        out = np.zeros(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, 8 * self.n,
                                  HIP_MEMCPY_DEVICE_TO_HOST) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _noalloc(capi):
    L = capi.load()
    f = L.ipcg_barrier_energy_noalloc
    f.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p]
    f.restype = C.c_int
    return L, f


def test_energy_noalloc_matches_energy(capi, scenes):
    L, noalloc = _noalloc(capi)
    ctx = capi.Context(0)  # initialises the device before the allocation
    out = DeviceDoubles(_hip(capi), 4)
    far = scenes.two_sheets(12, seed=2)
    far.V[far.V[:, 2] > 0, 2] += 1.0  # sheets far apart: no collision
    cases = [scenes.stacked_sheets(40, 8, seed=3, h=1.0 / 511),  # larger
             scenes.two_sheets(16, seed=1),                      # smaller
             far]                                                # empty
    try:
        for k, s in enumerate(cases):
            print("case", k, s.name, s.sizes, flush=True)
            ctx.upload_mesh(s.V, s.E, s.F)
            ctx.upload_positions(0, s.V)
            ctx.build_collisions(s.dhat)
            n = int(np.sum(ctx.collision_counts()))
            e_ref = ctx.energy(s.dhat, s.kappa)
            rc = noalloc(ctx._p, s.dhat, s.kappa, out.at(k))
            assert rc == 0, L.ipcg_last_error(ctx._p)
            ctx.synchronize()
            e = float(out.get()[k])
            print("  collisions", n, "energy", e_ref, e, flush=True)
            assert e == e_ref
            if k == 2:
                assert n == 0 and e == 0.0
            else:
                assert n > 0 and e > 0.0
    finally:
        out.free()
        ctx.close()


def test_energy_noalloc_rejects_null_output(capi, scenes):
    L, noalloc = _noalloc(capi)
    s = scenes.two_sheets(8, seed=0)
    ctx = capi.Context(0)
    try:
        ctx.upload_mesh(s.V, s.E, s.F)
        ctx.upload_positions(0, s.V)
        ctx.build_collisions(s.dhat)
        assert noalloc(ctx._p, s.dhat, s.kappa, None) == capi.ERR_ARG
    finally:
        ctx.close()
