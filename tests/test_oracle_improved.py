"""Oracle IMPROVED_MAX_APPROX collision sets against known answers.

normal_collisions.cpp:84-128 converts the candidates to sub-element pairs
(candidates.cpp:583-692) and adds signed vertex-vertex, edge-vertex and
mollified edge-edge contributions (normal_collisions_builder.cpp:340-543)
so that each contact is counted once, by inclusion-exclusion over the mesh
adjacencies (collision_mesh.cpp:247-291).  Pinned here by:

* the reference's own case (test_normal_collisions.cpp:120-200): the codim
  star gives 2 vertex-vertex + 4 edge-vertex collisions in IPC mode and one
  more vertex-vertex collision in IMPROVED mode, with or without area
  weighting; the codim vertex cube gives 12 vertex-vertex collisions in both;
* a point over the interior of a flat sheet with every primitive within
  dhat: IPC counts hundreds of face-, edge- and vertex-reduced pairs, the
  improved set keeps exactly one face-vertex collision of weight 1 (the
  signed edge and vertex terms cancel by the Euler count of the fan around
  each vertex), 0.25 a_p with area weighting;
* a vertex of valence k under an edge: IPC sums k edge-vertex contributions,
  the improved set one (weight 1; 0.25 (A_ea + k A_eb) with area
  weighting);
* a star with one edge parallel to the edge above it: the mollified
  edge-edge corrections and their distance types, worked by hand from
  normal_collisions_builder.cpp:463-543.
"""
import numpy as np
import pytest

Z2 = np.zeros((0, 2), np.int32)
Z3 = np.zeros((0, 3), np.int32)


def codim_star():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 2, 0],
                  [0, 3, 0]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], np.int32)
    return V, E, Z3


@pytest.mark.parametrize("aw", [False, True])
def test_reference_codim_star(O, aw):
    V, E, F = codim_star()
    dhat, dmin = 0.25, 0.8
    vv, ev = O.codim_candidates(V, E, F, 0.5 * (dhat + dmin))
    ipc = O.build_collisions(V, E, F, Z2, Z2, dhat, dmin, vv_cand=vv, ev_cand=ev,
                             area_weighting=aw)
    imp = O.build_collisions(V, E, F, Z2, Z2, dhat, dmin, vv_cand=vv, ev_cand=ev,
                             area_weighting=aw, improved=True)
    assert (len(ipc.vv), len(ipc.ev), len(ipc.ee), len(ipc.fv)) == (2, 4, 0, 0)
    assert (len(imp.vv), len(imp.ev), len(imp.ee), len(imp.fv)) == (3, 4, 0, 0)
    # the added pair is the centre with the hovering codim vertex: four
    # incident edges, none to vertex 5, so 1 - 4 times its (unit) weight
    k = [tuple(p) for p in imp.vv].index((0, 5))
    assert imp.vv_w[k] == (-1.5 if aw else -3.0)


def test_reference_codim_vertices(O):
    V = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    V -= V.mean(0)
    dhat, dmin = 0.25, 0.8
    vv, ev = O.codim_candidates(V, Z2, Z3, 0.5 * (dhat + dmin))
    for imp in (False, True):
        co = O.build_collisions(V, Z2, Z3, Z2, Z2, dhat, dmin, vv_cand=vv, ev_cand=ev,
                                improved=imp)
        assert len(co) == len(co.vv) == 12


def point_over_sheet(scenes, rng):
    Vs, Fs = scenes.grid_sheet(33)
    V, E, F = scenes.merge([(Vs, Fs)])
    p = np.array([0.5 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05),
                  rng.uniform(0.01, 0.05)])
    return np.vstack([V, p[None]]), E, F


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("aw", [False, True])
def test_point_over_flat_sheet_counts_once(O, scenes, seed, aw):
    V, E, F = point_over_sheet(scenes, np.random.default_rng(seed))
    p = len(V) - 1
    dhat = 0.3  # reaches hundreds of faces, stays 0.15 inside the rim
    ee, fv = O.candidates(V, E, F, 0.5 * dhat)
    fv = fv[fv[:, 1] == p]  # the point's pairs only (the flat sheet is not self-tested)
    ipc = O.build_collisions(V, E, F, Z2, fv, dhat, area_weighting=aw)
    imp = O.build_collisions(V, E, F, Z2, fv, dhat, area_weighting=aw, improved=True)
    assert len(ipc.vv) > 100 and len(ipc.ev) > 20 and len(ipc.fv) == 1
    assert (len(imp.vv), len(imp.ev), len(imp.ee), len(imp.fv)) == (0, 0, 0, 1)
    assert np.array_equal(imp.fv, ipc.fv)
    # the lone point's vertex area is 1
    assert imp.fv_w[0] == (0.25 if aw else 1.0)


def star_under_edge(k, rng, h=0.01):
    ang = np.sort(rng.uniform(0, 2 * np.pi, k))
    star = np.column_stack([0.3 * np.cos(ang), 0.3 * np.sin(ang), np.zeros(k)])
    d = rng.normal(size=2)
    d /= np.linalg.norm(d)
    a0 = np.array([-0.4 * d[0], -0.4 * d[1], h])
    a1 = np.array([0.4 * d[0], 0.4 * d[1], h])
    V = np.vstack([[0.0, 0, 0], star, a0, a1])
    E = np.array([[0, i + 1] for i in range(k)] + [[k + 1, k + 2]], np.int32)
    return V, E, Z3


@pytest.mark.parametrize("k", [2, 3, 4, 6])
def test_vertex_under_edge_counts_once(O, k):
    V, E, F = star_under_edge(k, np.random.default_rng(k))
    dhat = 0.05
    ee, fv = O.candidates(V, E, F, 0.5 * dhat)
    assert len(ee) == k and len(fv) == 0
    ea = k  # the hovering edge
    for aw in (False, True):
        ipc = O.build_collisions(V, E, F, ee, fv, dhat, area_weighting=aw)
        imp = O.build_collisions(V, E, F, ee, fv, dhat, area_weighting=aw, improved=True)
        for co in (ipc, imp):
            assert len(co.vv) == len(co.ee) == len(co.fv) == 0
            assert co.ev.tolist() == [[ea, 0]]
        if not aw:
            assert ipc.ev_w[0] == k and imp.ev_w[0] == 1.0
        else:  # edge areas are the codim edge lengths
            A_ea, A_eb = 0.8, 0.3
            assert abs(ipc.ev_w[0] - 0.25 * k * (A_ea + A_eb)) <= 1e-15
            assert abs(imp.ev_w[0] - 0.25 * (A_ea + k * A_eb)) <= 1e-15


def test_mollified_star_by_hand(O):
    """p = 0 with edges to q1 (along x, parallel to the edge 3 = (4, 5)
    above), q2, q3; every pair reduces to (3, 0) except the parallel one,
    which is mollified (cross product 0 < eps_x)."""
    h = 0.01
    V = np.array([[0, 0, 0], [0.25, 0, 0], [0, 0.3, 0], [-0.2, -0.25, 0], [-0.5, 0, h],
                  [0.5, 0, h]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [4, 5]], np.int32)
    dhat = 0.05
    ee, fv = O.candidates(V, E, Z3, 0.5 * dhat)
    assert sorted(map(sorted, ee.tolist())) == [[0, 3], [1, 3], [2, 3]]
    ipc = O.build_collisions(V, E, Z3, ee, fv, dhat)
    imp = O.build_collisions(V, E, Z3, ee, fv, dhat, improved=True)
    eps = 1e-3 * 1.0 * 0.25 ** 2  # 1e-3 |ea|^2 |eb|^2 at rest
    # IPC: (ea, e0) mollified -> edge-edge with its actual type; the other
    # two reduce to edge-vertex (3, 0)
    assert ipc.ev.tolist() == [[3, 0]] and ipc.ev_w.tolist() == [2.0]
    assert len(ipc.ee) == 1 and ipc.ee_w[0] == 1.0 and ipc.ee_eps_x[0] == eps
    actual = int(ipc.ee_dtype[0])
    # IMPROVED: (3, 0): -1 x (2 nonmollified - 1) on (3, 0) and a mollified
    # (3, e0) of type EA_EB0 (p = e0's first vertex) weighted -1;
    # (3, 1): no nonmollified incident edge -> +1 on (3, 1) and a mollified
    # (3, e0) of type EA_EB1, -1
    assert imp.ev.tolist() == [[3, 0], [3, 1]] and imp.ev_w.tolist() == [1.0, 1.0]
    want = {(3, 0, 4): -1.0, (3, 0, 5): -1.0}
    want[tuple(ipc.ee[0]) + (actual,)] = want.get(tuple(ipc.ee[0]) + (actual,), 0.0) + 1.0
    want = {k: v for k, v in want.items() if v != 0}
    got = {(int(a), int(b), int(t)): w
           for (a, b), t, w in zip(imp.ee, imp.ee_dtype, imp.ee_w)}
    assert got == want
    assert np.all(imp.ee_eps_x == eps)
    assert len(imp.vv) == len(imp.fv) == 0
