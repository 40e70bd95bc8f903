"""Oracle intersection queries against exact rational arithmetic.

has_intersections (ipc.cpp:105-166) = edge-face candidates at 1e-6 x the
world box diagonal (lbvh.cpp:771-784, can_edge_face_collide
lbvh.cpp:836-849) and is_edge_intersecting_triangle (intersection.cpp:
116-138): an exact orient3d rejection, then Eigen's FullPivLU solve of
[t1-t0, t2-t0, e0-e1] (u, v, t) = e0 - t0 with u, v >= 0, u + v <= 1,
0 <= t <= 1.

Pins: orient3d signs equal the sign of the determinant in Python's
Fractions (random, exactly coplanar, one-ulp perturbed and mixed-magnitude
inputs, so that the exact fallback runs); the LU solve matches
numpy.linalg.solve on well-conditioned systems, is exact on permutation and
diagonal systems, and gives Eigen's basic solution (free unknowns zero) on
hand-worked rank-deficient ones; the edge-triangle test and has_intersections
equal an exact rational (Cramer) evaluation over every non-adjacent
edge-face pair of random sphere scenes."""
from fractions import Fraction as Fr

import numpy as np
import pytest

from test_oracle_broadphase import np_boxes, overlap


def det3_sign(pa, pb, pc, pd):
    a = [Fr(pa[k]) - Fr(pd[k]) for k in range(3)]
    b = [Fr(pb[k]) - Fr(pd[k]) for k in range(3)]
    c = [Fr(pc[k]) - Fr(pd[k]) for k in range(3)]
    d = (a[2] * (b[0] * c[1] - c[0] * b[1]) + b[2] * (c[0] * a[1] - a[0] * c[1])
         + c[2] * (a[0] * b[1] - b[0] * a[1]))
    return (d > 0) - (d < 0)


def test_orient3d_random_and_degenerate(O):
    rng = np.random.default_rng(0)
    before = O.orient3d_exact_calls()
    cases = []
    for _ in range(300):  # generic
        cases.append(rng.normal(size=(4, 3)))
    for _ in range(300):  # exactly coplanar: z = a x + b y + c, all dyadic
        a, b, c = (rng.integers(-8, 9, 3) / 8.0)
        P = rng.integers(-64, 65, size=(4, 2)) / 16.0
        pts = np.column_stack([P, a * P[:, 0] + b * P[:, 1] + c])
        cases.append(pts)
        q = pts.copy()  # pd one ulp off the plane
        q[3, 2] = np.nextafter(q[3, 2], np.inf if rng.random() < 0.5 else -np.inf)
        cases.append(q)
    for _ in range(200):  # mixed magnitudes: inexact differences
        pts = rng.normal(size=(4, 3)) * 10.0 ** rng.integers(-20, 4, size=(4, 3))
        pts[3] = pts[0] + (pts[1] - pts[0]) * 0.5 + (pts[2] - pts[0]) * 0.25
        cases.append(pts)
    n_zero = 0
    for pts in cases:
        s = O.orient3d(*pts)
        assert s == det3_sign(*pts), pts
        n_zero += s == 0
    assert n_zero >= 300  # the exactly coplanar set
    assert O.orient3d_exact_calls() - before >= 600  # the filter deferred these


def test_fullpivlu_well_conditioned(O):
    rng = np.random.default_rng(1)
    for _ in range(500):
        A = rng.normal(size=(3, 3))
        if np.linalg.cond(A) > 1e6:
            continue
        b = rng.normal(size=3)
        x = O.fullpivlu_solve3(A, b)
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(x - ref) <= 1e-12 * np.linalg.cond(A) * np.linalg.norm(ref)


def test_fullpivlu_exact_cases(O):
    rng = np.random.default_rng(2)
    b = np.array([5.0, 9.0, 13.0])
    # diagonal: one correctly rounded division per unknown
    x = O.fullpivlu_solve3(np.diag([3.0, 7.0, 11.0]), b)
    assert list(x) == [5.0 / 3.0, 9.0 / 7.0, 13.0 / 11.0]
    for _ in range(20):  # permutations (with signs): exact
        P = np.eye(3)[rng.permutation(3)] * rng.choice([-1.0, 1.0], 3)[:, None]
        assert np.array_equal(O.fullpivlu_solve3(P, b), P.T @ b)
    assert np.array_equal(O.fullpivlu_solve3(np.zeros((3, 3)), b), np.zeros(3))
    # rank 2, worked by hand through FullPivLU: pivots (0,2) then (1,1);
    # the unknown in the last pivot column (u) is set to zero
    A = np.array([[1.0, 0.0, -3.0], [0.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
    assert list(O.fullpivlu_solve3(A, [-1.0, 0.25, 0.0])) == [0.0, 0.25, 1.0 / 3.0]
    # rank 1 consistent system: basic solution with two zeros, A x = b
    A = np.outer([1.0, 2.0, -1.0], [0.5, -4.0, 2.0])
    b1 = A @ np.array([0.0, 1.0, 0.0])
    x = O.fullpivlu_solve3(A, b1)
    assert np.count_nonzero(x) == 1 and np.allclose(A @ x, b1, rtol=0, atol=1e-15)


def exact_edge_tri(e0, e1, t0, t1, t2):
    """(intersects, margin, degenerate) by Cramer's rule in Fractions"""
    e0, e1, t0, t1, t2 = ([Fr(float(v)) for v in p] for p in (e0, e1, t0, t1, t2))
    c0 = [t1[k] - t0[k] for k in range(3)]
    c1 = [t2[k] - t0[k] for k in range(3)]
    c2 = [e0[k] - e1[k] for k in range(3)]
    r = [e0[k] - t0[k] for k in range(3)]

    def det(a, b, c):
        return (a[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (a[1] * c[2] - a[2] * c[1])
                + c[0] * (a[1] * b[2] - a[2] * b[1]))
    d = det(c0, c1, c2)
    if d == 0:
        return None, 0, True
    u, v, t = det(r, c1, c2) / d, det(c0, r, c2) / d, det(c0, c1, r) / d
    m = min(u, v, 1 - u - v, t, 1 - t)
    return m >= 0, abs(m), False


def float_edge_tri_margin(X, E, F, ei, fi):
    """vectorised Cramer in float64: (inside-margin, |det|) per pair"""
    e0, e1 = X[E[ei, 0]], X[E[ei, 1]]
    t0, t1, t2 = X[F[fi, 0]], X[F[fi, 1]], X[F[fi, 2]]
    c0, c1, c2, r = t1 - t0, t2 - t0, e0 - e1, e0 - t0
    d = np.einsum("ij,ij->i", c0, np.cross(c1, c2))
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        u = np.einsum("ij,ij->i", r, np.cross(c1, c2)) / d
        v = np.einsum("ij,ij->i", c0, np.cross(r, c2)) / d
        t = np.einsum("ij,ij->i", c0, np.cross(c1, r)) / d
    m = np.minimum.reduce([u, v, 1 - u - v, t, 1 - t])
    return m, np.abs(d)


def sphere_pair(rng, scenes, subdiv, gap):
    """two rotated icospheres of radius 1, centres 2 + gap apart (gap < 0:
    overlapping)"""
    Vs, Fs = scenes.icosphere(subdiv, 1.0)
    V0 = Vs @ scenes.random_rotation(rng).T
    V1 = Vs @ scenes.random_rotation(rng).T + np.array([2.0 + gap, 0.0, 0.0])
    return scenes.merge([(V0, Fs), (V1, Fs)])


def nonadjacent_pairs(E, F):
    ei, fi = np.meshgrid(np.arange(len(E)), np.arange(len(F)), indexing="ij")
    ei, fi = ei.ravel(), fi.ravel()
    share = (E[ei][:, :, None] == F[fi][:, None, :]).any(axis=(1, 2))
    return ei[~share], fi[~share]


def truth_any(V, E, F):
    """exact has-intersection over all non-adjacent edge-face pairs"""
    ei, fi = nonadjacent_pairs(E, F)
    m, ad = float_edge_tri_margin(V, E, F, ei, fi)
    clear_in = (ad > 1e-9) & (m > 1e-9)
    if clear_in.any():
        return True
    unsure = ~((ad > 1e-9) & (m < -1e-9))
    for k in np.nonzero(unsure)[0]:
        args = (V[E[ei[k], 0]], V[E[ei[k], 1]], V[F[fi[k], 0]], V[F[fi[k], 1]],
                V[F[fi[k], 2]])
        hit, _, degen = exact_edge_tri(*args)
        if degen:  # edge parallel to the plane: crossing needs coplanarity
            assert det3_sign(args[2], args[3], args[4], args[0]) != 0, "coplanar pair"
            continue
        if hit:
            return True
    return False


def test_edge_face_candidates_brute(O, scenes):
    rng = np.random.default_rng(3)
    for gap in (-0.3, 0.05):
        V, E, F = sphere_pair(rng, scenes, 1, gap)
        r = 1e-3
        ef = O.edge_face_candidates(V, E, F, r)
        _, eb, fb = np_boxes(V, E, F, r)
        M = overlap(eb, fb)
        ei, fi = np.nonzero(M)
        adj = (E[ei][:, :, None] == F[fi][:, None, :]).any(axis=(1, 2))
        want = {(int(a), int(b)) for a, b in zip(ei[~adj], fi[~adj])}
        got = {(int(a), int(b)) for a, b in ef}
        assert len(got) == len(ef) and got == want
        # patch filter: the two spheres as groups keeps only cross pairs
        g = np.repeat([0, 1], len(V) // 2).astype(np.int32)
        efg = O.edge_face_candidates(V, E, F, r, vertex_group=g)
        cross = {(a, b) for a, b in want if g[E[a, 0]] != g[F[b, 0]]}
        assert {(int(a), int(b)) for a, b in efg} == cross


@pytest.mark.parametrize("seed", range(4))
def test_edge_triangle_vs_exact(O, scenes, seed):
    rng = np.random.default_rng(10 + seed)
    V, E, F = sphere_pair(rng, scenes, 1, -0.5)
    ef = O.edge_face_candidates(V, E, F, 1e-6 * O.world_bbox_diagonal(V))
    n_hit = 0
    for e, f in ef:
        args = (V[E[e, 0]], V[E[e, 1]], V[F[f, 0]], V[F[f, 1]], V[F[f, 2]])
        hit, margin, degen = exact_edge_tri(*args)
        if degen or margin < 1e-12:
            continue
        assert O.is_edge_intersecting_triangle(*args) == hit
        n_hit += hit
    assert n_hit > 0


def test_edge_triangle_hand_cases(O):
    t = (np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0]))
    f = O.is_edge_intersecting_triangle
    assert f([0.25, 0.25, -1], [0.25, 0.25, 1], *t)       # crosses the interior
    assert not f([2.0, 2.0, -1], [2.0, 2.0, 1], *t)       # crosses the plane outside
    assert not f([0, 0, 1.0], [1, 1, 1.0], *t)            # parallel, off the plane
    assert f([0.25, 0.25, 0], [0.25, 0.25, 1], *t)        # touches: t = 0
    assert f([0.5, 0.5, -1], [0.5, 0.5, 1], *t)           # through the hypotenuse
    # coplanar: the singular solve sets u = 0 (worked above), so a crossing
    # segment counts when its line meets the triangle's v range
    assert f([-1, 0.25, 0], [2, 0.25, 0], *t)
    assert not f([-1, 2.0, 0], [2, 2.0, 0], *t)


def test_world_bbox_diagonal(O):
    V = np.array([[0.0, 0, 0], [3.0, -4, 1], [1.0, 2, 13]])
    assert O.world_bbox_diagonal(V) == np.sqrt((3.0 * 3 + 6.0 * 6) + 13.0 * 13)


def test_has_intersections_cubes(O, scenes):
    rng = np.random.default_rng(4)
    Vc, Fc = scenes.cube()
    # two cubes, each rotated about its centre, overlapping (the reference's
    # "Has intersections" case with coincident centres)
    for k in range(4):
        R1 = np.eye(3) if k == 0 else scenes.random_rotation(rng)
        R2 = np.eye(3) if k < 2 else scenes.random_rotation(rng)
        V, E, F = scenes.merge([(Vc @ R1.T, Fc), (Vc @ R2.T, Fc)])
        assert O.has_intersections(V, E, F), k
    V, E, F = scenes.merge([(Vc, Fc), (Vc + [1.5, 0.2, 0.1], Fc)])  # apart
    assert not O.has_intersections(V, E, F)
    V, E, F = scenes.merge([(Vc, Fc), (Vc + [0.6, 0.2, 0.1], Fc)])  # overlapping
    assert O.has_intersections(V, E, F)
    V, E, F = scenes.merge([(Vc, Fc), (0.3 * Vc, Fc)])  # nested, no crossing
    assert not O.has_intersections(V, E, F)


@pytest.mark.parametrize("gap", [-0.6, -0.02, 0.001, 0.3])
def test_has_intersections_vs_exact(O, scenes, gap):
    rng = np.random.default_rng(int(1000 * abs(gap)))
    V, E, F = sphere_pair(rng, scenes, 1, gap)
    assert O.has_intersections(V, E, F) == truth_any(V, E, F)
    # a patch filter over the two spheres: only cross pairs count
    g = np.repeat([0, 1], len(V) // 2).astype(np.int32)
    assert O.has_intersections(V, E, F, vertex_group=g) == truth_any(V, E, F)
