"""CCD search tiers against the oracle, with every tier forced to work.

The device runs a Tight-Inclusion search in one of four tiers (LDS lists,
a lane's HBM slab, a whole wave, the level-by-level wavefront) depending on
how large the search grows.  Two environment knobs shrink the first two
(IPCG_CCD_LDS_BOXES: boxes per LDS list, IPCG_CCD_SLAB_BOXES: boxes a slab
search may visit), so that the same scenes run mostly on the wave tier;
per-candidate times of impact and step sizes must stay bit-identical to
the oracle's."""
import os
import sys

import numpy as np
import pytest

from test_gpu_parity import make, upload

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the wave-cooperative search tier these knobs forced did not finish "
                     "its first GPU run within 3 minutes; no cause was found, so the tier was "
                     "removed and the knobs are not read by the library"),
]

KNOBS = [("1", "0"), ("6", "0"), ("2", "8")]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(params=KNOBS, ids=lambda k: f"lds{k[0]}-slab{k[1]}")
def tiers(request):
    old = {k: os.environ.get(k) for k in ("IPCG_CCD_LDS_BOXES", "IPCG_CCD_SLAB_BOXES")}
    os.environ["IPCG_CCD_LDS_BOXES"], os.environ["IPCG_CCD_SLAB_BOXES"] = request.param
    yield request.param
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("name", ["tet", "spheres:4:3:0", "spheres:2:3:5", "stack:24:1"])
def test_per_candidate_toi_tiers(ctx, O, scenes, name, tiers):
    s = make(scenes, name)
    say("case", name, tiers)
    upload(ctx, s)
    ctx.broad_phase(True, 0.0)
    ee, fv = ctx.candidates()
    te, tf = ctx.candidates_toi()
    assert np.array_equal(te, O.candidates_toi(s.V, s.V1, s.E, s.F, ee, True))
    assert np.array_equal(tf, O.candidates_toi(s.V, s.V1, s.E, s.F, fv, False))
    if tiers[1] == "0" and len(ee) + len(fv) > 0 and name != "tet":
        assert ctx.counter("ccd.wave") > 0


@pytest.mark.parametrize("name", ["sheets:64:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"])
def test_stepsize_tiers(ctx, O, scenes, name, tiers):
    s = make(scenes, name)
    say("case", name, tiers)
    upload(ctx, s)
    t = ctx.stepsize()
    assert t == O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F)
    assert ctx.is_step_collision_free() == O.is_step_collision_free(s.V, s.V1, s.E, s.F)


def test_min_distance_tiers(ctx, O, scenes, tiers):
    s = make(scenes, "spheres:4:3:0")
    say("case min_distance", tiers)
    upload(ctx, s)
    for md in (1e-5, 2e-4):
        assert ctx.stepsize(min_distance=md) == O.compute_collision_free_stepsize(
            s.V, s.V1, s.E, s.F, min_distance=md)
