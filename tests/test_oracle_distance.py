"""Oracle distance types and distances, pinned by the reference's known-answer
tests (tests/src/tests/distance/test_distance_type.cpp,
test_point_triangle.cpp, test_edge_edge.cpp) and closed forms."""
import numpy as np
import pytest

P_T0, P_T1, P_T2, P_E0, P_E1, P_E2, P_T = range(7)
EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB = range(9)

T0 = np.array([-1.0, 0, 1])
T1 = np.array([1.0, 0, 1])
T2 = np.array([0.0, 0, -1])


def edge_normal_2d(e0, e1):
    # tests/src/tests/utils.hpp:93-99
    e = np.asarray(e1) - np.asarray(e0)
    n = np.array([-e[1], e[0]])
    return n / np.linalg.norm(n)


def _pt_cases(rng):
    """(p without y, expected) following test_distance_type.cpp:95-211."""
    cases = []
    for pz in (0.0, -1 + 1e-12, 1 - 1e-12):
        cases.append(((0.0, pz), P_T))
    for _ in range(100):
        bc = rng.random(3)
        bc /= bc.sum()
        q = bc[0] * T0 + bc[1] * T1 + bc[2] * T2
        cases.append(((q[0], q[2]), P_T))
    for d in (1e-8, 1e-4, 0.1, 11):
        cases.append(((T0[0] - d, T0[2] + d), P_T0))
        cases.append(((T1[0] + d, T1[2] + d), P_T1))
    for pz in (-1 - 1e-12, -1.1, -11):
        cases.append(((0.0, pz), P_T2))
    for (a, b, dt) in ((T0, T1, P_E0), (T1, T2, P_E1), (T2, T0, P_E2)):
        perp = edge_normal_2d((a[0], a[2]), (b[0], b[2]))
        alphas = [1e-4, 0.5, 1 - 1e-4]
        scales = [1e-12, 1e-4, 1, 2, 11, 1000]
        pts = [(al, s) for al in alphas for s in scales]
        pts += list(zip(rng.uniform(1e-8, 1 - 1e-8, 100), rng.uniform(1e-12, 1e4, 100)))
        for al, s in pts:
            c = (b - a) * al + a
            cases.append(((c[0] + s * perp[0], c[2] + s * perp[1]), dt))
    return cases


def test_point_triangle_distance_type_kat(O):
    rng = np.random.default_rng(0)
    bad = []
    for py in (-10, -1, -1e-12, 0, 1e-12, 1, 10):
        for (px, pz), want in _pt_cases(rng):
            p = np.array([px, py, pz], dtype=np.float64)
            got = O.point_triangle_distance_type(p, T0, T1, T2)
            if got != want:
                bad.append((p.tolist(), want, got))
    assert not bad, bad[:5]


def test_point_triangle_distance_type_gh_issue(O):
    # test_distance_type.cpp:213-225
    p = [0.488166, 0.0132623, 0.289055]
    v0 = [0.456476, 0.0526442, 0.260834]
    v1 = [0.609111, 0.0595969, 0.275928]
    v2 = [0.431262, 0.0508414, 0.255831]
    assert O.point_triangle_distance_type(p, v0, v1, v2) == P_E0


def test_edge_edge_distance_type_kat(O):
    """test_distance_type.cpp:227-340 (closest-feature sweep)."""
    rng = np.random.default_rng(1)
    sign = lambda x: -1 if x < 0 else 1
    bad = 0
    total = 0
    alphas = np.arange(-1.0, 2.0 - 1e-9, 0.1)
    for alpha0 in alphas:
        for s0 in np.arange(-10.0, 10.0, 1.0):
            if s0 == 0:
                continue
            for swap_ea in (False, True):
                for swap_eb in (False, True):
                    for swap_edges in (False, True):
                        for _ in range(4):
                            alpha, s = alpha0, s0
                            ea0 = rng.uniform(-1, 1, 3)
                            ea1 = rng.uniform(-1, 1, 3)
                            n = np.cross(ea1 - ea0, [1.0, 0, 0])
                            n /= np.linalg.norm(n)
                            eb0 = ((ea1 - ea0) * alpha + ea0) + s * n
                            if alpha < 0:
                                alpha = 0
                                n = eb0 - ea0
                                s = np.linalg.norm(n)
                                n /= s
                            elif alpha > 1:
                                alpha = 1
                                n = eb0 - ea1
                                s = np.linalg.norm(n)
                                n /= s
                            eb1 = ((ea1 - ea0) * alpha + ea0) + (s + sign(s) * 1) * n
                            if swap_ea:
                                ea0, ea1 = ea1, ea0
                            if swap_eb:
                                eb0, eb1 = eb1, eb0
                            if swap_edges:
                                ea0, eb0 = eb0, ea0
                                ea1, eb1 = eb1, ea1
                            if not swap_edges:
                                if swap_ea and swap_eb:
                                    t00, t10, tm = EA1_EB1, EA0_EB1, EA_EB1
                                elif swap_ea:
                                    t00, t10, tm = EA1_EB0, EA0_EB0, EA_EB0
                                elif swap_eb:
                                    t00, t10, tm = EA0_EB1, EA1_EB1, EA_EB1
                                else:
                                    t00, t10, tm = EA0_EB0, EA1_EB0, EA_EB0
                            else:
                                if swap_ea and swap_eb:
                                    t00, t10, tm = EA1_EB1, EA1_EB0, EA1_EB
                                elif swap_ea:
                                    t00, t10, tm = EA0_EB1, EA0_EB0, EA0_EB
                                elif swap_eb:
                                    t00, t10, tm = EA1_EB0, EA1_EB1, EA1_EB
                                else:
                                    t00, t10, tm = EA0_EB0, EA0_EB1, EA0_EB
                            dt = O.edge_edge_distance_type(ea0, ea1, eb0, eb1)
                            total += 1
                            if abs(alpha) <= 1e-15:
                                ok = dt in (t00, tm)
                            elif abs(alpha - 1) <= 1e-15:
                                ok = dt in (t10, tm)
                            elif alpha < 0:
                                ok = dt == t00
                            elif alpha > 1:
                                ok = dt == t10
                            else:
                                ok = dt == tm
                            bad += not ok
    assert total > 5000
    assert bad == 0


def test_edge_edge_distance_type_interior(O):
    """test_distance_type.cpp:342-380: interior closest points -> EA_EB."""
    rng = np.random.default_rng(2)
    for _ in range(400):
        alpha, beta = rng.uniform(0.01, 0.99, 2)
        s = rng.choice(np.arange(-5.0, 5.0, 1.0))
        if s == 0:
            continue
        ea1 = rng.uniform(-1, 1, 3)
        eb1 = rng.uniform(-1, 1, 3)
        ea0 = alpha / (alpha - 1) * ea1
        eb0 = beta / (beta - 1) * eb1
        n = np.cross(ea1, eb1)
        n /= np.linalg.norm(n)
        eb0 = eb0 + s * n
        eb1 = eb1 + s * n
        assert O.edge_edge_distance_type(ea0, ea1, eb0, eb1) == EA_EB


def test_point_edge_distance_type(O):
    e0, e1 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0])
    assert O.point_edge_distance_type([-0.5, 1, 0], e0, e1) == 0
    assert O.point_edge_distance_type([1.5, 1, 0], e0, e1) == 1
    assert O.point_edge_distance_type([0.5, 1, 0], e0, e1) == 2


def test_point_triangle_distance_closed_form(O):
    """test_point_triangle.cpp:31-111: point above the plane -> y^2."""
    rng = np.random.default_rng(3)
    t0, t1, t2 = np.array([-1.0, 0, 1]), np.array([1.0, 0, 1]), np.array([0.0, 0, -1])
    for _ in range(200):
        bc = rng.random(3)
        bc /= bc.sum()
        q = bc[0] * t0 + bc[1] * t1 + bc[2] * t2
        y = rng.uniform(-10, 10)
        p = q + np.array([0, y, 0])
        assert O.point_triangle_distance(p, t0, t1, t2) == pytest.approx(y * y, abs=1e-10)
        assert O.point_plane_distance(p, t0, t1, t2) == pytest.approx(y * y, abs=1e-10)


def test_edge_edge_distance_closed_form(O):
    """test_edge_edge.cpp:31-80: perpendicular crossing edges at height d."""
    for d in (0.0, 1e-8, 1e-3, 0.5, 3.0):
        for ea_shift in (-0.5, 0.0, 0.3):
            ea0, ea1 = np.array([-1.0, 0, 0]), np.array([1.0, 0, 0])
            eb0, eb1 = np.array([ea_shift, -1.0, d]), np.array([ea_shift, 1.0, d])
            assert O.edge_edge_distance(ea0, ea1, eb0, eb1) == pytest.approx(d * d, abs=1e-12)
            assert O.line_line_distance(ea0, ea1, eb0, eb1) == pytest.approx(d * d, abs=1e-12)
    # parallel edges, endpoint-to-endpoint closest
    ea0, ea1 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0])
    eb0, eb1 = np.array([2.0, 1.0, 0]), np.array([3.0, 1.0, 0])
    assert O.edge_edge_distance(ea0, ea1, eb0, eb1) == pytest.approx(2.0, abs=1e-14)


def test_point_point_and_line(O):
    assert O.point_point_distance([0, 0, 0], [1, 2, 2]) == 9.0
    assert O.point_line_distance([0, 2, 0], [-1, 0, 0], [3, 0, 0]) == pytest.approx(4.0)
    assert O.point_edge_distance([-2, 1, 0], [0, 0, 0], [1, 0, 0]) == pytest.approx(5.0)


def test_mollifier_threshold_kat(O):
    """test_edge_edge_mollifier.cpp:17-26, 264-270: perpendicular unit-2
    edges: cross^2 = 16, eps_x = 1e-3 * 4 * 4 = 0.016."""
    ea0, ea1 = [-1.0, 0, 0], [1.0, 0, 0]
    eb0, eb1 = [0.0, -1, 0], [0.0, 1, 0]
    assert O.edge_edge_cross_squarednorm(ea0, ea1, eb0, eb1) == pytest.approx(16)
    assert O.edge_edge_mollifier_threshold(ea0, ea1, eb0, eb1) == pytest.approx(0.016)


def test_mollifier_value_range(O):
    """m in [0, 1]; m = 1 at or above the threshold; derivatives by central
    differences (edge_edge_mollifier.cpp:16-70)."""
    rng = np.random.default_rng(4)
    for _ in range(50):
        x = rng.uniform(-1, 1, 12)
        c = O.edge_edge_cross_squarednorm(x[0:3], x[3:6], x[6:9], x[9:12])
        for eps in (0.5 * c, 2 * c, 10 * c):
            m, g, H = O.edge_edge_mollifier(x, eps)
            assert 0 <= m <= 1
            if c >= eps:
                assert m == 1 and not g.any() and not H.any()
                continue
            h = 1e-7
            fd = np.array([(O.edge_edge_mollifier(x + h * e, eps)[0]
                            - O.edge_edge_mollifier(x - h * e, eps)[0]) / (2 * h)
                           for e in np.eye(12)])
            assert np.allclose(g, fd, rtol=1e-5, atol=1e-6 * max(1, np.abs(g).max()))
            fdH = np.array([(O.edge_edge_mollifier(x + h * e, eps)[1]
                             - O.edge_edge_mollifier(x - h * e, eps)[1]) / (2 * h)
                            for e in np.eye(12)])
            assert np.allclose(H, fdH, rtol=1e-4, atol=1e-5 * max(1, np.abs(H).max()))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_stencil_distance_derivatives_fd(O, kind):
    """Closed-form distance gradients / Hessians against central differences
    (the reference checks them with finite-diff: test_point_triangle.cpp,
    test_edge_edge.cpp)."""
    rng = np.random.default_rng(10 + kind)
    n = O.STENCIL_NDOF[kind]
    for _ in range(20):
        x = rng.uniform(-1, 1, n)
        v, g, H = O.stencil_distance_derivs(kind, x)
        h = 1e-6
        E = np.eye(n)
        fd = np.array([(O.stencil_distance_derivs(kind, x + h * e)[0]
                        - O.stencil_distance_derivs(kind, x - h * e)[0]) / (2 * h) for e in E])
        assert np.allclose(g, fd, rtol=1e-5, atol=1e-7)
        fdH = np.array([(O.stencil_distance_derivs(kind, x + h * e)[1]
                         - O.stencil_distance_derivs(kind, x - h * e)[1]) / (2 * h) for e in E])
        assert np.allclose(H, fdH, rtol=1e-4, atol=1e-6)
        assert np.allclose(H, H.T, atol=1e-12)
