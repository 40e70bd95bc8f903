"""Oracle mesh areas and area-weighted collision sets.

CollisionMesh::init_areas (collision_mesh.cpp:309-376): vertex areas are a
third of the incident faces' areas (else half the incident edges' lengths,
else 1), edge areas a third of the incident faces' areas (else the edge
length).  Pins: on a flat grid both sum to the sheet's area and take the
closed values h^2, h^2/3 and h^2/6; a polyline and isolated points take
lengths and 1.

Area weighting (normal_collisions_builder.cpp:26-336): every active
candidate contributes 0.5 (a_i + a_j) (codim vertex pair), 0.5 a_v (codim
edge-vertex), 0.25 (A_ea + A_eb) (edge-edge) or 0.25 a_v (face-vertex) to
the collision it reduces to.  Pins: face-vertex and edge-edge collisions
(one contribution each) carry exactly that weight, and the weights of
every type sum to the contributions of the active candidates, found with
the oracle's distance functions."""
import numpy as np
import pytest


def test_areas_flat_grid(O, scenes):
    n = 17
    h = 1.0 / (n - 1)
    V, F = scenes.grid_sheet(n)
    E = scenes.edges_from_faces(F)
    va, ea = O.mesh_areas(V, E, F)
    assert abs(va.sum() - 1.0) <= 1e-14 and abs(ea.sum() - 1.0) <= 1e-14
    interior = [(i * n + j) for i in range(1, n - 1) for j in range(1, n - 1)]
    assert np.allclose(va[interior], h * h, rtol=1e-14, atol=0)
    # interior edges touch two faces, boundary edges one
    # an edge lies on the boundary when both ends share x = 0, x = 1, y = 0 or y = 1
    on_boundary = np.array([any(V[a, k] == V[b, k] and V[a, k] in (0.0, 1.0) for k in range(2))
                            for a, b in E])
    assert np.allclose(ea[~on_boundary], h * h / 3, rtol=1e-14, atol=0)
    assert np.allclose(ea[on_boundary], h * h / 6, rtol=1e-14, atol=0)


def test_areas_codim():
    import oracle as O
    V = np.array([[0.0, 0, 0], [3.0, 0, 0], [3.0, 4.0, 0], [9.0, 9.0, 9.0]])
    E = np.array([[0, 1], [1, 2]], np.int32)
    F = np.zeros((0, 3), np.int32)
    va, ea = O.mesh_areas(V, E, F)
    assert list(ea) == [3.0, 4.0]
    assert list(va) == [1.5, 3.5, 2.0, 1.0]


def candidate_contributions(O, V, E, F, ee, fv, dhat, va, ea):
    """sum over active candidates of their contribution weights"""
    total = 0.0
    act2 = dhat * dhat
    for a, b in ee:
        x = V[[E[a, 0], E[a, 1], E[b, 0], E[b, 1]]]
        if O.edge_edge_distance(*x, -1) < act2:
            total += 0.25 * (ea[a] + ea[b])
    for f, v in fv:
        x = V[[v, F[f, 0], F[f, 1], F[f, 2]]]
        if O.point_triangle_distance(*x, -1) < act2:
            total += 0.25 * va[v]
    return total


@pytest.mark.parametrize("name", ["sheets", "spheres"])
def test_area_weighted_collisions(O, scenes, name):
    s = scenes.two_sheets(12, seed=2) if name == "sheets" else scenes.sphere_chain(3, 2, seed=0)
    dhat = s.dhat if name == "sheets" else 3 * s.dhat
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * dhat)
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, area_weighting=True)
    cu = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat)
    va, ea = O.mesh_areas(s.V, s.E, s.F)
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(getattr(co, k), getattr(cu, k))
    for (f, v), w in zip(co.fv, co.fv_w):
        assert w == 0.25 * va[v]
    for (a, b), w in zip(co.ee, co.ee_w):
        assert w == 0.25 * (ea[a] + ea[b])
    total = sum(float(np.sum(getattr(co, k + "_w"))) for k in ("vv", "ev", "ee", "fv"))
    want = candidate_contributions(O, s.V, s.E, s.F, ee, fv, dhat, va, ea)
    assert abs(total - want) <= 1e-12 * want
    assert len(co) > 0
