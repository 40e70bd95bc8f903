"""Oracle friction (TangentialCollisions + FrictionPotential) against the
mathematics.

* smooth_friction_mollifier.cpp / smooth_mu.cpp: f1 = f0', f2 = f1',
  f1_over_x = f1 / y, f2_x_minus_f1_over_x3 = (f2 y - f1) / y^3, and the same
  for the mu-weighted forms with mu_s != mu_k in each of the three pieces
  (|y| < eps_v / 2, eps_v / 2 <= |y| < eps_v, |y| >= eps_v), continuity
  across the joins, mu(0) = mu_s and mu(eps_v) = mu_k (as in
  test_smooth_friction_mollifier.cpp and test_smooth_mu.cpp);
* tangent bases (tangent_basis.cpp): orthonormal columns orthogonal to the
  contact normal of each stencil kind; closest-point coordinates: the
  closest-point difference is orthogonal to the edges it was projected on;
* a point pair along z with a known normal force: energy
  mu N f0(|(v_x, v_y)|) exactly;
* gradient = finite differences of the energy, Hessian (unprojected) =
  finite differences of the gradient, in the static and the sliding
  regimes; the projected Hessian is positive semi-definite.
"""
import numpy as np
import pytest

Z2 = np.zeros((0, 2), np.int32)
Z3 = np.zeros((0, 3), np.int32)

EPS_V = 1e-3


def fd(f, y, h):
    return (f(y + h) - f(y - h)) / (2 * h)


@pytest.mark.parametrize("mu_s,mu_k", [(0.5, 0.5), (0.6, 0.3), (0.2, 0.5)])
def test_smooth_mu_calculus(O, mu_s, mu_k):
    g = lambda name: (lambda y: O.smooth_mu_fn(name, y, mu_s, mu_k, EPS_V))
    h = 1e-9
    for y in np.concatenate([np.linspace(0.02, 0.48, 7), np.linspace(0.52, 0.98, 7),
                             [1.2, 3.0]]) * EPS_V:
        assert abs(fd(g("mu_f0"), y, h) - g("mu_f1")(y)) <= 1e-6 * max(abs(g("mu_f1")(y)), 1e-3)
        assert abs(fd(g("mu_f1"), y, h) - g("mu_f2")(y)) <= 1e-5 * max(abs(g("mu_f2")(y)), 1.0)
        assert abs(fd(g("mu"), y, h) - g("dmu")(y)) <= 1e-5 * max(abs(g("dmu")(y)), 1.0)
        assert abs(g("mu_f1_over_x")(y) - g("mu_f1")(y) / y) <= 1e-12 * abs(g("mu_f1")(y) / y)
        want = (g("mu_f2")(y) * y - g("mu_f1")(y)) / y ** 3
        assert abs(g("mu_f2_x_minus_f1_over_x3")(y) - want) <= 1e-9 * abs(want)
    # the joins and the ends
    for y in (0.5 * EPS_V, EPS_V):
        for name in ("mu", "mu_f0", "mu_f1"):
            a, b = g(name)(y * (1 - 1e-12)), g(name)(y * (1 + 1e-12))
            assert abs(a - b) <= 1e-9 * max(abs(a), 1e-12), (name, y)
    assert g("mu")(0.0) == mu_s and g("mu")(EPS_V) == mu_k
    assert g("mu_f0")(2 * EPS_V) == mu_k * 2 * EPS_V


def test_mollifier_pieces(O):
    f = lambda name: (lambda y: O.smooth_mu_fn(name, y, 1, 1, EPS_V))
    assert f("f0")(0.0) == EPS_V / 3 and f("f0")(2 * EPS_V) == 2 * EPS_V
    assert f("f1")(EPS_V) == 1 and f("f2")(EPS_V) == 0
    for y in np.linspace(0.05, 0.95, 9) * EPS_V:
        assert abs(fd(f("f0"), y, 1e-9) - f("f1")(y)) <= 1e-6
        assert abs(fd(f("f1"), y, 1e-9) - f("f2")(y)) <= 1e-5 * f("f2")(y)
        assert abs(f("f1_over_x")(y) - f("f1")(y) / y) <= 1e-12 * f("f1")(y) / y


def four_kinds(rng):
    """one collision of each kind: a codim vertex pair, a codim vertex over a
    codim edge, two crossing codim edges, a codim vertex over a triangle"""
    d = 0.01
    V = np.array([
        [0, 0, 0], [0.003, -0.002, d],                        # VV: 0, 1
        [5, 0, 0], [6, 0, 0], [5.4, 0.003, d],                # EV: edge 2-3, vertex 4
        [10, -0.5, 0], [10, 0.5, 0], [9.5, 0.1, d], [10.6, -0.2, d],  # EE: 5-6, 7-8
        [15, 0, 0], [16, 0, 0], [15, 1, 0], [15.3, 0.2, d],   # FV: face 9-10-11, vertex 12
    ], float)
    V += rng.normal(0, 1e-4, V.shape) * [1, 1, 0]
    E = np.array([[2, 3], [5, 6], [7, 8], [9, 10], [10, 11], [11, 9]], np.int32)
    F = np.array([[9, 10, 11]], np.int32)
    return V, E, F


def kinds_collisions(O, V, E, F, dhat):
    ee, fv = O.candidates(V, E, F, 0.5 * dhat)
    vv, ev = O.codim_candidates(V, E, F, 0.5 * dhat)
    return O.build_collisions(V, E, F, ee, fv, dhat, vv_cand=vv, ev_cand=ev)


def test_bases_and_closest_points(O):
    V, E, F = four_kinds(np.random.default_rng(0))
    co = kinds_collisions(O, V, E, F, 0.02)
    assert (len(co.vv), len(co.ev), len(co.ee), len(co.fv)) == (1, 1, 1, 1)
    fr = O.friction_build(co, V, E, F, 0.02, 1.0, 0.5)
    assert fr.keep.all()
    for i, P in enumerate(fr.basis):
        assert np.allclose(P.T @ P, np.eye(2), atol=1e-14)
    # contact normals (closest-point differences) are orthogonal to the bases
    vv = co.vv[0]
    n_vv = V[vv[1]] - V[vv[0]]
    e, v = co.ev[0]
    a = fr.beta[1, 0]
    q = V[E[e, 0]] + a * (V[E[e, 1]] - V[E[e, 0]])
    n_ev = V[v] - q
    ea, eb = co.ee[0]
    al, be = fr.beta[2]
    pa = V[E[ea, 0]] + al * (V[E[ea, 1]] - V[E[ea, 0]])
    pb = V[E[eb, 0]] + be * (V[E[eb, 1]] - V[E[eb, 0]])
    n_ee = pa - pb
    f, v2 = co.fv[0]
    t = V[F[f]]
    b0, b1 = fr.beta[3]
    n_fv = V[v2] - (t[0] + b0 * (t[1] - t[0]) + b1 * (t[2] - t[0]))
    for i, nrm in enumerate([n_vv, n_ev, n_ee, n_fv]):
        P = fr.basis[i]
        assert np.all(np.abs(P.T @ nrm) <= 1e-12 * np.linalg.norm(nrm)), i
    # the closest-point differences are orthogonal to what they project on
    assert abs(n_ev @ (V[E[e, 1]] - V[E[e, 0]])) <= 1e-14
    for u in (V[E[ea, 1]] - V[E[ea, 0]], V[E[eb, 1]] - V[E[eb, 0]]):
        assert abs(n_ee @ u) <= 1e-14
    for u in (t[1] - t[0], t[2] - t[0]):
        assert abs(n_fv @ u) <= 1e-14
    # the first basis vector runs along the first edge (EV, EE, FV)
    for i, u in ((1, V[E[e, 1]] - V[E[e, 0]]), (2, V[E[ea, 1]] - V[E[ea, 0]]), (3, t[1] - t[0])):
        assert np.allclose(fr.basis[i][:, 0], u / np.linalg.norm(u), atol=1e-15)


def test_point_pair_closed_form(O):
    d, dhat, kappa, mu = 0.004, 0.01, 2.0, 0.4
    V = np.array([[0, 0, 0], [0, 0, d]], float)
    vv = np.array([[0, 1]], np.int32)
    co = O.build_collisions(V, Z2, Z3, Z2, Z2, dhat, vv_cand=vv)
    fr = O.friction_build(co, V, Z2, Z3, dhat, kappa, mu)
    # N = -kappa b'(d^2) 2 d with b(x) = -(x - dhat^2)^2 log(x / dhat^2)
    x, t = d * d, dhat * dhat
    db = -2 * (x - t) * np.log(x / t) - (x - t) ** 2 / x
    N = -kappa * db * 2 * d
    assert abs(fr.normal_force[0] - N) <= 1e-13 * N
    # basis along x then y: u = (v0x - v1x, v0y - v1y)
    assert np.allclose(fr.basis[0], [[1, 0], [0, 1], [0, 0]])
    for vx, vy in ((3e-4, -2e-4), (2e-3, 1e-3)):
        Vel = np.array([[vx, vy, 0.7], [0, 0, -0.2]])
        y = np.hypot(vx, vy)
        f0 = y if y >= EPS_V else y * y * (1 - y / (3 * EPS_V)) / EPS_V + EPS_V / 3
        assert abs(O.friction_energy(fr, Vel, EPS_V) - mu * N * f0) <= 1e-13 * mu * N * f0


@pytest.mark.parametrize("regime", ["static", "sliding"])
def test_derivatives(O, scenes, regime):
    s = scenes.two_sheets(10, seed=2)
    co = O.build_collisions(s.V, s.E, s.F, *O.candidates(s.V, s.E, s.F, 1.5 * s.dhat), 3 * s.dhat)
    assert len(co) > 20
    fr = O.friction_build(co, s.V, s.E, s.F, 3 * s.dhat, s.kappa, 0.6, 0.3)
    rng = np.random.default_rng(3)
    scale = 0.3 * EPS_V if regime == "static" else 20 * EPS_V
    Vel = rng.normal(0, scale, s.V.shape)
    g = O.friction_gradient(fr, Vel, EPS_V)
    h = 1e-7 * scale
    for k in rng.choice(np.nonzero(g)[0], 20, replace=False):
        Vp, Vm = Vel.copy(), Vel.copy()
        Vp.flat[k] += h
        Vm.flat[k] -= h
        num = (O.friction_energy(fr, Vp, EPS_V) - O.friction_energy(fr, Vm, EPS_V)) / (2 * h)
        assert abs(num - g[k]) <= 1e-5 * np.abs(g).max(), k
    H = O.friction_hessian(fr, Vel, EPS_V, "none").toarray()
    assert np.allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    for k in rng.choice(np.nonzero(g)[0], 10, replace=False):
        Vp, Vm = Vel.copy(), Vel.copy()
        Vp.flat[k] += h
        Vm.flat[k] -= h
        num = (O.friction_gradient(fr, Vp, EPS_V) - O.friction_gradient(fr, Vm, EPS_V)) / (2 * h)
        assert np.linalg.norm(num - H[:, k]) <= 1e-4 * np.abs(H).max(), k
    Hp = O.friction_hessian(fr, Vel, EPS_V, "clamp").toarray()
    assert np.linalg.eigvalsh(Hp).min() >= -1e-9 * np.abs(Hp).max()


def test_near_parallel_edges_are_not_kept(O):
    # two parallel codim edges 0.005 apart: mollified edge-edge collision
    V = np.array([[0, 0, 0], [1, 0, 0], [0.2, 0, 0.005], [1.2, 0, 0.005]], float)
    E = np.array([[0, 1], [2, 3]], np.int32)
    ee = np.array([[0, 1]], np.int32)
    co = O.build_collisions(V, E, Z3, ee, Z2, 0.01)
    assert len(co.ee) == 1
    fr = O.friction_build(co, V, E, Z3, 0.01, 1.0, 0.5)
    assert not fr.keep[0]
    assert O.friction_energy(fr, np.ones_like(V), EPS_V) == 0.0


def test_per_vertex_coefficients(O):
    """tangential_collisions.cpp:86-166 with default_blend_mu (the average):
    each side's coefficient at its closest point, averaged"""
    V, E, F = four_kinds(np.random.default_rng(1))
    co = kinds_collisions(O, V, E, F, 0.02)
    m = np.random.default_rng(2).uniform(0.1, 0.9, len(V))
    fr = O.friction_build(co, V, E, F, 0.02, 1.0, m, 0.5 * m)
    b = fr.beta
    a, c = co.vv[0]
    want = [(m[a] + m[c]) / 2]
    e, v = co.ev[0]
    want.append(((m[E[e, 1]] - m[E[e, 0]]) * b[1, 0] + m[E[e, 0]] + m[v]) / 2)
    ea, eb = co.ee[0]
    want.append((((m[E[ea, 1]] - m[E[ea, 0]]) * b[2, 0] + m[E[ea, 0]])
                 + ((m[E[eb, 1]] - m[E[eb, 0]]) * b[2, 1] + m[E[eb, 0]])) / 2)
    f, v = co.fv[0]
    t = F[f]
    want.append((m[t[0]] + b[3, 0] * (m[t[1]] - m[t[0]]) + b[3, 1] * (m[t[2]] - m[t[0]]) + m[v]) / 2)
    assert np.allclose(fr.mu_s, want, rtol=1e-15)
    assert np.allclose(fr.mu_k, 0.5 * np.array(want), rtol=1e-15)
    # equal coefficients everywhere reproduce the scalar build exactly
    f1 = O.friction_build(co, V, E, F, 0.02, 1.0, 0.4, 0.3)
    f2 = O.friction_build(co, V, E, F, 0.02, 1.0, np.full(len(V), 0.4), np.full(len(V), 0.3))
    assert np.array_equal(f1.mu_s, f2.mu_s) and np.all(f1.mu_s == 0.4)
    Vel = np.random.default_rng(4).normal(0, 1e-3, V.shape)
    assert O.friction_energy(f1, Vel, EPS_V) == O.friction_energy(f2, Vel, EPS_V)


def test_anisotropic_mu_eff_reference_case(O):
    # test_anisotropic_friction.cpp:23-79: the L2 (elliptical) projection
    # sqrt((mu0 t0)^2 + (mu1 t1)^2) over directions, the principal axes
    rng = np.random.default_rng(0)
    for _ in range(50):
        ms, mk = rng.uniform(0.1, 0.9, 2), rng.uniform(0.1, 0.9, 2)
        for ang in np.arange(0, 2 * np.pi, np.pi / 4):
            t = np.array([np.cos(ang), np.sin(ang)])
            got = O.anisotropic_mu_eff(3.7 * t, ms, mk, 0.5, 0.4)
            assert np.allclose(got, [np.hypot(*(ms * t)), np.hypot(*(mk * t))], rtol=1e-14)
        assert np.allclose(O.anisotropic_mu_eff([2.0, 0], ms, mk, 0.5, 0.4), [ms[0], mk[0]])
        assert np.allclose(O.anisotropic_mu_eff([0, -2.0], ms, mk, 0.5, 0.4), [ms[1], mk[1]])
    # zero tangential velocity: the first axis; no anisotropy: the isotropic pair
    assert np.allclose(O.anisotropic_mu_eff([0, 0], [0.3, 0.7], [0.2, 0.6], 0.5, 0.4), [0.3, 0.2])
    assert np.array_equal(O.anisotropic_mu_eff([1, 1], [0, 0], [0, 0], 0.5, 0.4), [0.5, 0.4])


def test_anisotropic_friction(O, scenes):
    s = scenes.two_sheets(10, seed=2)
    co = O.build_collisions(s.V, s.E, s.F, *O.candidates(s.V, s.E, s.F, 1.5 * s.dhat), 3 * s.dhat)
    fr = O.friction_build(co, s.V, s.E, s.F, 3 * s.dhat, s.kappa, 0.6, 0.3)
    n = len(fr.keep)
    rng = np.random.default_rng(7)
    Vel = rng.normal(0, 2 * EPS_V, s.V.shape)
    e_iso = O.friction_energy(fr, Vel, EPS_V)
    # defaults reproduce the isotropic potential exactly
    O.friction_set_anisotropy(fr, np.ones((n, 2)), np.zeros((n, 2)), np.zeros((n, 2)))
    O.friction_update_lagged_mu(fr, Vel)
    assert O.friction_energy(fr, Vel, EPS_V) == e_iso
    # velocity scaling and elliptical coefficients from the lagged velocities
    a = rng.uniform(0.5, 2.0, (n, 2))
    msa, mka = rng.uniform(0.2, 0.8, (n, 2)), rng.uniform(0.1, 0.5, (n, 2))
    O.friction_set_anisotropy(fr, a, msa, mka)
    O.friction_update_lagged_mu(fr, Vel)
    ms_eff, mk_eff = O.friction_effective_mu(fr)
    k = fr.keep
    assert np.all(ms_eff[k] <= np.linalg.norm(msa[k], axis=1) + 1e-15)
    assert np.all(ms_eff[k] >= msa[k].min(axis=1) - 1e-15)
    # derivatives with the coefficients held (lagged)
    g = O.friction_gradient(fr, Vel, EPS_V)
    h = 1e-7 * 2 * EPS_V
    for j in rng.choice(np.nonzero(g)[0], 15, replace=False):
        Vp, Vm = Vel.copy(), Vel.copy()
        Vp.flat[j] += h
        Vm.flat[j] -= h
        num = (O.friction_energy(fr, Vp, EPS_V) - O.friction_energy(fr, Vm, EPS_V)) / (2 * h)
        assert abs(num - g[j]) <= 1e-5 * np.abs(g).max(), j
    H = O.friction_hessian(fr, Vel, EPS_V, "none").toarray()
    for j in rng.choice(np.nonzero(g)[0], 8, replace=False):
        Vp, Vm = Vel.copy(), Vel.copy()
        Vp.flat[j] += h
        Vm.flat[j] -= h
        num = (O.friction_gradient(fr, Vp, EPS_V) - O.friction_gradient(fr, Vm, EPS_V)) / (2 * h)
        assert np.linalg.norm(num - H[:, j]) <= 1e-4 * np.abs(H).max(), j
