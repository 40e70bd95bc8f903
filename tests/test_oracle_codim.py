"""Codim candidates (free vertices and free edges), pinned by the
reference's codim known-answer tests (tests/src/tests/collisions/
test_normal_collisions.cpp:14-206)."""
import numpy as np
import pytest


def cube():
    V = np.array([[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1], [1, 0, 0], [1, 0, 1],
                  [1, 1, 0], [1, 1, 1]], dtype=np.float64)
    return V - V.mean(axis=0)


E0 = np.zeros((0, 2), np.int32)
F0 = np.zeros((0, 3), np.int32)


def all_candidates(O, V, E, F, r, V1=None):
    ee, fv = O.candidates(V, E, F, r, V1=V1)
    vv, ev = O.codim_candidates(V, E, F, r, V1=V1)
    return ee, fv, vv, ev


def test_codim_lists(O):
    V = np.zeros((8, 3))
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], dtype=np.int32)
    cv, ce = O.codim_lists(8, E, F0)
    assert cv.tolist() == [5, 6, 7] and ce.tolist() == [0, 1, 2, 3]
    from ipctk_amd.scenes import edges_from_faces
    F = np.array([[0, 1, 2]], dtype=np.int32)
    E2 = np.vstack([edges_from_faces(F), [[3, 4]]]).astype(np.int32)
    cv, ce = O.codim_lists(6, E2, F)
    assert cv.tolist() == [5] and ce.tolist() == [3]


def test_codim_vertex_vertex_kat(O):
    """test_normal_collisions.cpp:14-107"""
    thickness, md = 0.4, 0.8
    V = cube()
    V1 = V.copy()
    V1[:, 1] *= 0.5
    ee, fv, vv, ev = all_candidates(O, V, E0, F0, thickness, V1)
    assert len(vv) > 0 and len(ee) == len(fv) == len(ev) == 0
    assert not O.is_step_collision_free(V, V1, E0, F0, min_distance=md)
    # candidates.compute_collision_free_stepsize with inflation = thickness
    expected = (1 - (md + 1e-4)) / 2.0 / 0.25
    t = O.candidates_stepsize(V, V1, E0, F0, ee, fv, min_distance=md, vv=vv, ev=ev)
    assert t == pytest.approx(expected, rel=1.2e-5)
    # collisions: the 12 cube edges (dhat 0.25, dmin 0.8)
    dhat = 0.25
    ee, fv, vv, ev = all_candidates(O, V, E0, F0, 0.5 * (dhat + md))
    c = O.build_collisions(V, E0, F0, ee, fv, dhat, dmin=md, vv_cand=vv, ev_cand=ev)
    assert len(c) == 12 and len(c.vv) == 12
    assert O.barrier_energy(c, V, E0, F0, dhat, 1.0) > 0
    g = O.barrier_gradient(c, V, E0, F0, dhat, 1.0).reshape(-1, 3)
    f = -g / np.linalg.norm(g, axis=1, keepdims=True)
    assert np.allclose(f, V / np.linalg.norm(V, axis=1, keepdims=True), atol=1e-12)


def test_codim_edge_vertex_kat(O):
    """test_normal_collisions.cpp:108-206"""
    thickness, md = 1e-3, 2e-3
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [0, 1, 0],
                  [0, 2, 0], [0, 3, 0]], dtype=np.float64)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], dtype=np.int32)
    V1 = V.copy()
    V1[-3:, 1] -= 4
    ee, fv, vv, ev = all_candidates(O, V, E, F0, thickness, V1)
    assert (len(vv), len(ev), len(ee), len(fv)) == (3, 12, 0, 0)
    assert not O.is_step_collision_free(V, V1, E, F0, min_distance=md)
    expected = (1 - (md + 1e-4)) / 4
    t = O.candidates_stepsize(V, V1, E, F0, ee, fv, min_distance=md, vv=vv, ev=ev)
    assert t == pytest.approx(expected, rel=1.2e-5)
    dhat, dmin = 0.25, 0.8
    ee, fv, vv, ev = all_candidates(O, V, E, F0, 0.5 * (dhat + dmin))
    c = O.build_collisions(V, E, F0, ee, fv, dhat, dmin=dmin, vv_cand=vv, ev_cand=ev)
    assert (len(c), len(c.vv), len(c.ev), len(c.ee), len(c.fv)) == (6, 2, 4, 0, 0)
    assert O.barrier_energy(c, V, E, F0, dhat, 1.0) > 0


def test_codim_candidates_brute_force(O):
    """random free points and free segments plus a triangle: the codim lists
    are the FP32 box overlaps among free vertices and between free edges and
    free vertices"""
    rng = np.random.default_rng(3)
    V = rng.uniform(0, 1, (40, 3))
    F = np.array([[0, 1, 2]], dtype=np.int32)
    from ipctk_amd.scenes import edges_from_faces
    seg = np.array([[3 + 2 * i, 4 + 2 * i] for i in range(8)], dtype=np.int32)
    E = np.vstack([edges_from_faces(F), seg]).astype(np.int32)
    r = 0.15
    vb, eb, _ = O.boxes_f32(V, E, F, r)
    cv, ce = O.codim_lists(len(V), E, F)
    assert len(cv) == 40 - 3 - 16 and len(ce) == 8
    want_vv = {(int(min(a, b)), int(max(a, b))) for i, a in enumerate(cv) for b in cv[i + 1:]
               if np.all(vb[a, :3] <= vb[b, 3:]) and np.all(vb[b, :3] <= vb[a, 3:])}
    want_ev = {(int(e), int(v)) for e in ce for v in cv
               if np.all(eb[e, :3] <= vb[v, 3:]) and np.all(vb[v, :3] <= eb[e, 3:])}
    vv, ev = O.codim_candidates(V, E, F, r)
    assert {tuple(sorted(p)) for p in vv.tolist()} == want_vv and len(vv) == len(want_vv)
    assert set(map(tuple, ev.tolist())) == want_ev and len(ev) == len(want_ev)
    # codim point-point / point-edge CCD kinds agree with the stencil CCD
    ee, fv, vv, ev = all_candidates(O, V, E, F, 0.0, V + 0.05)
    te = O.candidates_toi(V, V + 0.05, E, F, ev, "ev")
    assert te.shape == (len(ev),)
