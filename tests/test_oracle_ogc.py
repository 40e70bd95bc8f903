"""Oracle OGC collision sets against known answers.

CollisionSetType::OGC (normal_collisions_builder.cpp:56-336 with use_ogc)
drops edge-vertex, edge-edge and face-vertex pairs whose closest point lies
outside the feasible region of the feature it reduces to
(ogc/feasible_region.cpp) and divides a pair reduced to a vertex or an edge
among that feature's incident elements.  Pinned by:

* the reference's feasible-region cases (test_feasible_region.cpp:15-118),
  transcribed;
* the reference's collision counts for the codim cases under OGC
  (test_normal_collisions.cpp: 12 vertex-vertex; 2 + 4);
* closed forms: a point over the interior of a flat sheet keeps only the
  face-vertex collision of the face below it (every edge and vertex region
  excludes it); a point above the centre of a 4-face fan reduces to one
  vertex-vertex collision whose four quarter weights sum to 1 (IPC: 4); a
  point nearest the corner of a codim polyline of valence 2 gets weight
  1/2 + 1/2 = 1 (IPC: 2), and one over an edge's interior keeps that
  edge-vertex pair while the corner's region excludes it.
"""
import numpy as np
import pytest

Z2 = np.zeros((0, 2), np.int32)
Z3 = np.zeros((0, 3), np.int32)


def test_reference_vertex_region(O):
    V = np.array([[-1.0, 0, 0], [0, 0, 0], [1, 0, 0]])
    E = np.array([[0, 1], [1, 2]], np.int32)
    want = {-1.1: (True, False, False), -1e-8: (False, False, False), 0.0: (False, True, False),
            1e-8: (False, False, False), 1.1: (False, False, True)}
    for x0, w in want.items():
        x = np.array([x0, 0.1, 0.0])  # only x.x() changes between the checks
        got = tuple(O.ogc_vertex_region(V, E, Z3, x, vi) for vi in range(3))
        assert got == w, (x0, got)


def edge_region_mesh():
    V = np.array([[0, 0, 1], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [1e-8, 1, 1e-8]], float)
    E = np.array([[0, 2], [0, 1], [1, 2], [2, 3], [3, 0]], np.int32)
    F = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return V, E, F


def test_reference_edge_region(O):
    V, E, F = edge_region_mesh()
    assert not O.ogc_edge_region(V, E, F, 4, 0)
    V[0, 1] = V[2, 1] = 0.25
    assert O.ogc_edge_region(V, E, F, 4, 0)
    V[4, 2] = -1.1
    assert not O.ogc_edge_region(V, E, F, 4, 0)


def test_reference_edge_vertex_feasible(O):
    V, E, F = edge_region_mesh()
    assert O.ogc_feasible(V, E, F, "ev", 0, 4)
    V[0, 1] = V[2, 1] = 0.25
    assert O.ogc_feasible(V, E, F, "ev", 0, 4)
    V[4, 2] = -1.1
    assert O.ogc_feasible(V, E, F, "ev", 0, 4)
    V[4, 2] = 1.1
    assert O.ogc_feasible(V, E, F, "ev", 0, 4)


def test_pyramid_peak_feasible_for_every_face(O):
    # test_feasible_region.cpp:156-175 on a pentagonal pyramid of our own
    # (the reference's pentagon-dome.obj is not in the tree): the peak is a
    # vertex of every side face, so each pair reduces to the peak itself
    ang = 2 * np.pi * np.arange(5) / 5
    V = np.vstack([np.column_stack([np.cos(ang), np.zeros(5), np.sin(ang)]), [[0, 0.8, 0]],
                   [[0, 1, 0]]])
    F = np.array([[i, (i + 1) % 5, 5] for i in range(5)] + [[0, 2, 1], [0, 3, 2], [0, 4, 3]],
                 np.int32)
    E = np.unique(np.sort(np.vstack([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]]), axis=1), axis=0)
    for fi in range(5):
        assert O.ogc_feasible(V, E.astype(np.int32), F, "fv", fi, 5)


def test_reference_codim_counts(O):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 2, 0],
                  [0, 3, 0]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], np.int32)
    vv, ev = O.codim_candidates(V, E, Z3, 0.5 * (0.25 + 0.8))
    for aw in (False, True):
        co = O.build_collisions(V, E, Z3, Z2, Z2, 0.25, 0.8, vv_cand=vv, ev_cand=ev,
                                area_weighting=aw, ogc=True)
        assert (len(co.vv), len(co.ev), len(co.ee), len(co.fv)) == (2, 4, 0, 0)
    P = np.array([[x, y, z] for x in (0, 1) for y in (0, 1) for z in (0, 1)], float)
    P -= P.mean(0)
    vv, ev = O.codim_candidates(P, Z2, Z3, 0.5 * (0.25 + 0.8))
    co = O.build_collisions(P, Z2, Z3, Z2, Z2, 0.25, 0.8, vv_cand=vv, ev_cand=ev, ogc=True)
    assert len(co) == len(co.vv) == 12


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("aw", [False, True])
def test_point_over_flat_sheet(O, scenes, seed, aw):
    rng = np.random.default_rng(seed)
    Vs, Fs = scenes.grid_sheet(33)
    V, E, F = scenes.merge([(Vs, Fs)])
    p = np.array([0.5 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05), 0.02])
    V = np.vstack([V, p[None]])
    ee, fv = O.candidates(V, E, F, 0.15)
    fv = fv[fv[:, 1] == len(V) - 1]
    ipc = O.build_collisions(V, E, F, Z2, fv, 0.3, area_weighting=aw)
    co = O.build_collisions(V, E, F, Z2, fv, 0.3, area_weighting=aw, ogc=True)
    assert len(ipc) > 100
    assert (len(co.vv), len(co.ev), len(co.ee), len(co.fv)) == (0, 0, 0, 1)
    assert np.array_equal(co.fv, ipc.fv) and co.fv_w[0] == (0.25 if aw else 1.0)


def fan4(h=0.02, off=(0.0, 0.0)):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [off[0], off[1], h]])
    F = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 1]], np.int32)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4], [1, 2], [2, 3], [3, 4], [4, 1]], np.int32)
    return V, E, F


def test_point_above_fan_centre(O):
    V, E, F = fan4()
    fv = np.array([[f, 5] for f in range(4)], np.int32)
    ipc = O.build_collisions(V, E, F, Z2, fv, 0.05)
    co = O.build_collisions(V, E, F, Z2, fv, 0.05, ogc=True)
    # every face reduces to the centre vertex: IPC sums 4, OGC 4 x 1/4
    assert ipc.vv.tolist() == [[0, 5]] and ipc.vv_w.tolist() == [4.0]
    assert co.vv.tolist() == [[0, 5]] and co.vv_w.tolist() == [1.0]
    assert len(co.ev) == len(co.fv) == 0


def test_codim_corner(O):
    # polyline a - c - b with the corner c at the origin, a point near c
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [-0.1, 0.1, -0.1]], float)
    E = np.array([[0, 1], [0, 2]], np.int32)
    ev = np.array([[0, 3], [1, 3]], np.int32)
    ipc = O.build_collisions(V, E, Z3, Z2, Z2, 0.5, ev_cand=ev)
    co = O.build_collisions(V, E, Z3, Z2, Z2, 0.5, ev_cand=ev, ogc=True)
    assert ipc.vv.tolist() == [[0, 3]] and ipc.vv_w.tolist() == [2.0]
    assert co.vv.tolist() == [[0, 3]] and co.vv_w.tolist() == [1.0]
    # over the first edge's interior: that pair stays an edge-vertex
    # collision; the second edge reduces to the corner, whose region
    # excludes the point (it lies beyond the first edge's side)
    V2 = V.copy()
    V2[3] = [0.05, 0.1, -0.2]
    co2 = O.build_collisions(V2, E, Z3, Z2, Z2, 0.5, ev_cand=ev, ogc=True)
    ipc2 = O.build_collisions(V2, E, Z3, Z2, Z2, 0.5, ev_cand=ev)
    assert ipc2.vv.tolist() == [[0, 3]] and ipc2.ev.tolist() == [[0, 3]]
    assert len(co2.vv) == 0 and co2.ev.tolist() == [[0, 3]] and co2.ev_w.tolist() == [1.0]
