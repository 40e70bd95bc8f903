"""Area-weighted collision sets on the device against the oracle.

Ids and weights must be identical: both sides sum a collision's
contributions in ascending order of value, so the sums agree bit for bit
whatever order the device produced them in; energy, gradient and Hessian
within the usual 1e-12 / 1e-10."""
import sys

import numpy as np
import pytest

from conftest import rel_err, sparse_rel_err
from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["sheets:16:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"])
def test_area_weighted_collisions(ctx, O, scenes, name):
    s = make(scenes, name)
    say("case", name)
    upload(ctx, s)
    ctx.set_area_weighting(True)
    ctx.build_collisions(s.dhat)
    cg = ctx.collisions()
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat, area_weighting=True)
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(cg[k], getattr(co, k)), k
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(cg[k + "_w"], getattr(co, k + "_w")), k
    e_o = O.barrier_energy(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert abs(ctx.energy(s.dhat, s.kappa) - e_o) <= 1e-12 * abs(e_o)
    g_o = O.barrier_gradient(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert rel_err(ctx.gradient(s.dhat, s.kappa), g_o) <= 1e-12
    H_o = O.barrier_hessian(co, s.V, s.E, s.F, s.dhat, s.kappa, "clamp")
    assert sparse_rel_err(ctx.hessian(s.dhat, s.kappa, "clamp"), H_o) <= 1e-10
    # switching back gives the unit weights again
    ctx.set_area_weighting(False)
    ctx.build_collisions(s.dhat)
    cu = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat)
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(ctx.collisions()[k + "_w"], getattr(cu, k + "_w")), k


def test_area_weighted_codim(ctx, O):
    # two point layers (codim vertices) 0.5 dhat apart
    rng = np.random.default_rng(5)
    P = rng.uniform(0, 1, (40, 2)) * 0.01
    V = np.vstack([np.column_stack([P, np.zeros(40)]),
                   np.column_stack([P + rng.normal(0, 1e-4, P.shape), np.full(40, 5e-4)])])
    E = np.zeros((0, 2), np.int32)
    F = np.zeros((0, 3), np.int32)
    say("case codim points")
    ctx.upload_mesh(V, E, F)
    ctx.upload_positions(0, V)
    ctx.set_area_weighting(True)
    ctx.build_collisions(1e-3)
    cg = ctx.collisions()
    vv, ev = O.codim_candidates(V, E, F, 0.5e-3)
    co = O.build_collisions(V, E, F, np.zeros((0, 2), np.int32), np.zeros((0, 2), np.int32),
                            1e-3, vv_cand=vv, ev_cand=ev, area_weighting=True)
    assert np.array_equal(cg["vv"], co.vv) and len(co.vv) > 0
    assert np.array_equal(cg["vv_w"], co.vv_w)
    assert np.all(co.vv_w == 1.0)  # isolated points weigh 1 each: 0.5 (1 + 1)
