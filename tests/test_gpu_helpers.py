"""Device Gauss-Newton and candidate-set helpers against the oracle.

The per-vertex safe distances, the noncandidate step and the CFL step are
order-independent minima / maxima of bit-exact quantities, so they must
match the oracle bit for bit.  The Gauss-Newton diagonal and quadratic form
are FP64 sums over collisions whose order differs: 1e-12 relative."""
import sys

import numpy as np
import pytest

from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu

TOL_SUM = 1e-12
SCENES = ["tet", "sheets:16:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", SCENES)
def test_gauss_newton(ctx, O, scenes, name):
    s = make(scenes, name)
    say("case", name)
    upload(ctx, s)
    ctx.build_collisions(s.dhat)
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat)
    d_g = ctx.gauss_newton_diagonal(s.dhat, s.kappa)
    d_o = O.gauss_newton_diagonal(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert np.linalg.norm(d_g - d_o) <= TOL_SUM * max(np.linalg.norm(d_o), 1e-300)
    p = np.random.default_rng(1).normal(size=3 * len(s.V))
    q_g = ctx.gauss_newton_quadratic_form(s.dhat, s.kappa, p)
    q_o = O.gauss_newton_quadratic_form(co, s.V, s.E, s.F, s.dhat, s.kappa, p)
    assert abs(q_g - q_o) <= TOL_SUM * max(abs(q_o), 1e-300)
    # the records do not disturb the gradient or an assembled Hessian
    ctx.hessian(s.dhat, s.kappa, "clamp")
    g0 = ctx.gradient(s.dhat, s.kappa)
    ctx.gauss_newton_diagonal(s.dhat, s.kappa)
    assert np.array_equal(g0, ctx.gradient(s.dhat, s.kappa))


@pytest.mark.parametrize("name", SCENES)
def test_candidate_helpers(ctx, O, scenes, name):
    s = make(scenes, name)
    say("case", name)
    upload(ctx, s)
    r = 0.5 * s.dhat
    ctx.broad_phase(False, r)
    ee, fv = ctx.candidates()
    for md in (0.0, 0.25 * r):
        got = ctx.per_vertex_safe_distances(r, md)
        want = O.per_vertex_safe_distances(s.V, s.E, s.F, ee, fv, r, md)
        assert np.array_equal(got, want)
    got = ctx.noncandidate_conservative_stepsize(s.dhat)
    assert got == O.noncandidate_conservative_stepsize(s.V, s.V1, s.E, s.F, ee, fv, s.dhat)
    for dhat in (1e-6, s.dhat, 1e-1):
        ctx.broad_phase(False, r)
        got = ctx.cfl_stepsize(dhat)
        want = O.cfl_stepsize(s.V, s.V1, s.E, s.F, ee, fv, dhat)
        assert got == want, (dhat, got, want)


def test_codim_safe_distances(ctx, O):
    # test_per_vertex_safe_distances.cpp through the codim point lists
    V = np.array([[-1.0, 0, 0], [1, 0, 0], [10, 0, 0]])
    E = np.zeros((0, 2), np.int32)
    F = np.zeros((0, 3), np.int32)
    ctx.upload_mesh(V, E, F)
    ctx.upload_positions(0, V)
    ctx.broad_phase(False, 10.0)
    vv, ev = ctx.codim_candidates()
    assert len(vv) == 3 and len(ev) == 0
    assert list(ctx.per_vertex_safe_distances(10.0, 0.5)) == [1.5, 1.5, 8.5]
