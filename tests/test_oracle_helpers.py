"""Oracle Gauss-Newton and candidate-set helpers against closed forms.

Gauss-Newton (normal_potential.cpp:15-83, 236-308): each collision's
distance vector t = sum_i c_i x_i must be the closest-point difference of its
stored distance type, so |t|^2 equals the stencil's squared distance, t is
orthogonal to the edges / triangle it was projected onto, and the
coefficients sum to zero; for point-point collisions the Gauss-Newton
Hessian is the exact barrier Hessian, so the diagonal and the quadratic form
equal those of the oracle's Hessian (computed by automatic
differentiation); the quadratic form at a unit vector is the diagonal
entry.

Candidate helpers (candidates.cpp:294-408): the reference's per-vertex
safe-distance case (test_per_vertex_safe_distances.cpp), two flat sheets
with a dyadic gap where every vertex's safe distance is the gap, a uniform
displacement of norm 13/1024 for the noncandidate step, and the CFL step
bounded by the full CCD step (test_cfl.cpp)."""
import numpy as np
import pytest


def kat_points():
    V = np.array([[-1.0, 0, 0], [1, 0, 0], [10, 0, 0]])
    return V, np.zeros((0, 2), np.int32), np.zeros((0, 3), np.int32)


def test_safe_distances_reference_case(O):
    V, E, F = kat_points()
    vv = np.array([[0, 1], [0, 2], [1, 2]], np.int32)
    d = O.per_vertex_safe_distances(V, E, F, None, None, 10.0, 0.5, vv=vv)
    assert list(d) == [1.5, 1.5, 8.5]


def flat_pair(scenes, n=17, gap=2.0 ** -9):
    Vb, Fb = scenes.grid_sheet(n)
    Vt = Vb.copy()
    Vt[:, 2] = gap
    return scenes.merge([(Vb, Fb), (Vt, Fb)])


def test_safe_distances_flat_sheets(O, scenes):
    g, r, md = 2.0 ** -9, 2.0 ** -8, 2.0 ** -11
    V, E, F = flat_pair(scenes, gap=g)
    ee, fv = O.candidates(V, E, F, r)
    d = O.per_vertex_safe_distances(V, E, F, ee, fv, r, md)
    assert np.all(d == g - md)
    # no candidates: every vertex keeps inflation_radius - min_distance
    d0 = O.per_vertex_safe_distances(V, E, F, ee[:0], fv[:0], r, md)
    assert np.all(d0 == r - md)


def test_noncandidate_stepsize(O, scenes):
    V, E, F = flat_pair(scenes)
    ee, fv = O.candidates(V, E, F, 2.0 ** -8)
    # an isolated far vertex in no candidate, moving far
    V = np.vstack([V, [[5.0, 5.0, 5.0]]])
    V1 = V + np.array([3.0, 4.0, 12.0]) / 1024
    V1[-1] += 100.0
    dhat = 1e-3
    assert O.noncandidate_conservative_stepsize(V, V1, E, F, ee, fv, dhat) == \
        0.5 * dhat / (13.0 / 1024)
    assert O.noncandidate_conservative_stepsize(V, V1, E, F, ee[:0], fv[:0], dhat) == 1.0


@pytest.mark.parametrize("name", ["sheets", "spheres"])
def test_cfl_bounded_by_ccd(O, scenes, name):
    s = scenes.two_sheets(16, seed=1) if name == "sheets" else scenes.sphere_chain(3, 2, seed=2)
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    full = O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F)
    for dhat in (1e-6, s.dhat, 1e-1):
        cfl = O.cfl_stepsize(s.V, s.V1, s.E, s.F, ee, fv, dhat)
        assert cfl <= full


def scene_collisions(O, s, dhat=None):
    dhat = s.dhat if dhat is None else dhat
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * dhat)
    return O.build_collisions(s.V, s.E, s.F, ee, fv, dhat)


@pytest.mark.parametrize("name", ["sheets", "twisted", "spheres"])
def test_distance_vector_is_the_closest_point_difference(O, scenes, name):
    s = {"sheets": lambda: scenes.two_sheets(12, seed=0),
         "twisted": lambda: scenes.two_sheets(12, seed=1, twist=0.3),
         "spheres": lambda: scenes.sphere_chain(3, 2, seed=0)}[name]()
    co = scene_collisions(O, s, 3 * s.dhat)  # the sphere gaps reach 1.5 dhat
    nv, nev, nee = len(co.vv), len(co.ev), len(co.ee)
    assert len(co) > 0
    for i in range(len(co)):
        t, c = O.distance_vector(co, s.V, s.E, s.F, i)
        assert abs(c.sum()) <= 1e-15
        if i < nv:
            kind, ids, dt = 3, co.vv[i], 0
        elif i < nv + nev:
            e, v = co.ev[i - nv]
            kind, ids, dt = 2, [v, *s.E[e]], 2
        elif i < nv + nev + nee:
            j = i - nv - nev
            ea, eb = co.ee[j]
            kind, ids, dt = 1, [*s.E[ea], *s.E[eb]], int(co.ee_dtype[j])
        else:
            f, v = co.fv[i - nv - nev - nee]
            kind, ids, dt = 0, [v, *s.F[f]], 6
        x = s.V[np.asarray(ids)]
        d2 = O.stencil_distance_derivs(kind, x.reshape(-1), dt)[0]
        assert abs(t @ t - d2) <= 1e-12 * d2
        if kind == 0:  # P_T: t is normal to the triangle
            for e in (x[2] - x[1], x[3] - x[1]):
                assert abs(t @ e) <= 1e-12 * np.linalg.norm(t) * np.linalg.norm(e)
        if kind == 1 and dt == 8:  # EA_EB: t is normal to both edges
            for e in (x[1] - x[0], x[3] - x[2]):
                assert abs(t @ e) <= 1e-12 * np.linalg.norm(t) * np.linalg.norm(e)


def vv_scene():
    """two point layers 0.5 dhat apart: point-point collisions only"""
    rng = np.random.default_rng(5)
    P = rng.uniform(0, 1, (40, 2)) * 0.01
    V = np.vstack([np.column_stack([P, np.zeros(40)]),
                   np.column_stack([P + rng.normal(0, 1e-4, P.shape), np.full(40, 5e-4)])])
    return V, np.zeros((0, 2), np.int32), np.zeros((0, 3), np.int32)


def test_gauss_newton_exact_for_point_point(O):
    V, E, F = vv_scene()
    dhat, kappa = 1e-3, 1e3
    vv, ev = O.codim_candidates(V, E, F, 0.5 * dhat)
    co = O.build_collisions(V, E, F, np.zeros((0, 2), np.int32), np.zeros((0, 2), np.int32),
                            dhat, vv_cand=vv, ev_cand=ev)
    assert len(co) == len(co.vv) > 10
    H = O.barrier_hessian(co, V, E, F, dhat, kappa, "none").toarray()
    dg = O.gauss_newton_diagonal(co, V, E, F, dhat, kappa)
    assert np.linalg.norm(dg - np.diag(H)) <= 1e-12 * np.linalg.norm(np.diag(H))
    p = np.random.default_rng(6).normal(size=3 * len(V))
    q = O.gauss_newton_quadratic_form(co, V, E, F, dhat, kappa, p)
    assert abs(q - p @ H @ p) <= 1e-12 * abs(p @ H @ p)


def test_gauss_newton_quadratic_form_unit_vectors(O, scenes):
    s = scenes.two_sheets(10, seed=3)
    co = scene_collisions(O, s)
    dg = O.gauss_newton_diagonal(co, s.V, s.E, s.F, s.dhat, s.kappa)
    for k in np.nonzero(dg)[0][:40]:
        e = np.zeros(3 * len(s.V))
        e[k] = 1.0
        q = O.gauss_newton_quadratic_form(co, s.V, s.E, s.F, s.dhat, s.kappa, e)
        assert abs(q - dg[k]) <= 1e-12 * abs(dg[k])
