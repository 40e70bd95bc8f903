"""Device friction (tangential collisions + friction potential) against the
oracle.

The oracle's friction is pinned on the CPU (calculus identities of the
smooth-mu polynomials, tangent-basis geometry, a closed-form point pair,
finite differences).  Here the device records built over the device's own
collision set must match the oracle's on the same set: keep flags
identical, closest-point coordinates, bases and normal forces within
1e-13 relative (the same formulas in the same order); energy within 1e-12,
gradient within 1e-12 and the Hessian (projected and not) within 1e-10
relative, in the static and the sliding regimes, with mu_s = mu_k and
mu_s != mu_k."""
import sys

import numpy as np
import pytest

from conftest import rel_err, sparse_rel_err
from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu

EPS_V = 1e-3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mus", [(0.5, 0.5), (0.6, 0.3)])
@pytest.mark.parametrize("name", ["sheets:16:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"])
def test_friction(ctx, O, scenes, name, mus):
    s = make(scenes, name)
    dhat = 3 * s.dhat
    say("case", name, mus)
    upload(ctx, s)
    ctx.build_collisions(dhat)
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, vv_cand=vv, ev_cand=ev)
    assert len(co) == ctx.collision_counts().sum() > 0
    ctx.friction_build(dhat, s.kappa, *mus)
    fr = O.friction_build(co, s.V, s.E, s.F, dhat, s.kappa, *mus)
    keep, beta, basis, force = ctx.friction_records()
    assert np.array_equal(keep, fr.keep)
    k = fr.keep
    assert np.allclose(beta[k], fr.beta[k], rtol=1e-13, atol=1e-15)
    assert np.allclose(basis[k], fr.basis[k], rtol=1e-13, atol=1e-15)
    assert np.allclose(force[k], fr.normal_force[k], rtol=1e-13, atol=0)
    rng = np.random.default_rng(1)
    for scale in (0.3 * EPS_V, 20 * EPS_V):
        Vel = rng.normal(0, scale, s.V.shape)
        ctx.upload_velocities(Vel)
        e_o = O.friction_energy(fr, Vel, EPS_V)
        assert abs(ctx.friction_energy(EPS_V) - e_o) <= 1e-12 * abs(e_o)
        g_o = O.friction_gradient(fr, Vel, EPS_V)
        assert rel_err(ctx.friction_gradient(EPS_V), g_o) <= 1e-12
        for psd in ("clamp", "none"):
            H_o = O.friction_hessian(fr, Vel, EPS_V, psd)
            assert sparse_rel_err(ctx.friction_hessian(EPS_V, psd), H_o) <= 1e-10, psd


@pytest.mark.parametrize("name", ["spheres:4:3:0", "stack:32:0"])
def test_per_vertex_coefficients(ctx, O, scenes, name):
    s = make(scenes, name)
    dhat = 3 * s.dhat
    say("case per-vertex", name)
    upload(ctx, s)
    ctx.build_collisions(dhat)
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, vv_cand=vv, ev_cand=ev)
    m = np.random.default_rng(5).uniform(0.1, 0.9, len(s.V))
    ctx.friction_build(dhat, s.kappa, m, 0.6 * m)
    fr = O.friction_build(co, s.V, s.E, s.F, dhat, s.kappa, m, 0.6 * m)
    Vel = np.random.default_rng(6).normal(0, 2 * EPS_V, s.V.shape)
    ctx.upload_velocities(Vel)
    e_o = O.friction_energy(fr, Vel, EPS_V)
    assert abs(ctx.friction_energy(EPS_V) - e_o) <= 1e-12 * abs(e_o)
    assert rel_err(ctx.friction_gradient(EPS_V), O.friction_gradient(fr, Vel, EPS_V)) <= 1e-12
    H_o = O.friction_hessian(fr, Vel, EPS_V, "clamp")
    assert sparse_rel_err(ctx.friction_hessian(EPS_V, "clamp"), H_o) <= 1e-10


@pytest.mark.parametrize("name", ["twisted:12:1", "stack:32:0"])
def test_anisotropic(ctx, O, scenes, name):
    s = make(scenes, name)
    dhat = 3 * s.dhat
    say("case anisotropic", name)
    upload(ctx, s)
    ctx.build_collisions(dhat)
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, vv_cand=vv, ev_cand=ev)
    ctx.friction_build(dhat, s.kappa, 0.6, 0.3)
    fr = O.friction_build(co, s.V, s.E, s.F, dhat, s.kappa, 0.6, 0.3)
    n = len(fr.keep)
    rng = np.random.default_rng(8)
    a = rng.uniform(0.5, 2.0, (n, 2))
    msa, mka = rng.uniform(0.2, 0.8, (n, 2)), rng.uniform(0.1, 0.5, (n, 2))
    msa[::3] = 0  # some collisions stay isotropic
    ctx.friction_set_anisotropy(a, msa, mka)
    O.friction_set_anisotropy(fr, a, msa, mka)
    Vel = rng.normal(0, 2 * EPS_V, s.V.shape)
    ctx.upload_velocities(Vel)
    ctx.friction_update_lagged_mu()
    O.friction_update_lagged_mu(fr, Vel)
    e_o = O.friction_energy(fr, Vel, EPS_V)
    assert abs(ctx.friction_energy(EPS_V) - e_o) <= 1e-12 * abs(e_o)
    assert rel_err(ctx.friction_gradient(EPS_V), O.friction_gradient(fr, Vel, EPS_V)) <= 1e-12
    for psd in ("clamp", "none"):
        H_o = O.friction_hessian(fr, Vel, EPS_V, psd)
        assert sparse_rel_err(ctx.friction_hessian(EPS_V, psd), H_o) <= 1e-10, psd


def test_signed_weights(ctx, capi, O, scenes):
    """friction over an IMPROVED_MAX_APPROX set: negative weights give
    negative scales, which the projected Hessian drops (sliding and zero
    velocity) or projects out (static)"""
    s = make(scenes, "sheets:16:0")
    dhat = 3 * s.dhat
    say("case signed weights")
    upload(ctx, s)
    ctx.set_area_weighting(True)
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(dhat)
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, vv_cand=vv, ev_cand=ev,
                            area_weighting=True, improved=True)
    assert (co.ev_w < 0).any() or (co.ee_w < 0).any()
    ctx.friction_build(dhat, s.kappa, 0.5)
    fr = O.friction_build(co, s.V, s.E, s.F, dhat, s.kappa, 0.5)
    rng = np.random.default_rng(9)
    for scale in (0.0, 0.3 * EPS_V, 20 * EPS_V):
        Vel = rng.normal(0, scale, s.V.shape) if scale else np.zeros_like(s.V)
        ctx.upload_velocities(Vel)
        e_o = O.friction_energy(fr, Vel, EPS_V)
        assert abs(ctx.friction_energy(EPS_V) - e_o) <= 1e-12 * abs(e_o)
        g_o = O.friction_gradient(fr, Vel, EPS_V)
        if scale:
            assert rel_err(ctx.friction_gradient(EPS_V), g_o) <= 1e-12
        for psd in ("clamp", "abs", "none"):
            H_o = O.friction_hessian(fr, Vel, EPS_V, psd)
            assert sparse_rel_err(ctx.friction_hessian(EPS_V, psd), H_o) <= 1e-10, (scale, psd)


def test_requires_a_current_set(ctx, capi, scenes):
    s = make(scenes, "sheets:16:0")
    upload(ctx, s)
    ctx.build_collisions(3 * s.dhat)
    ctx.friction_build(3 * s.dhat, s.kappa, 0.5)
    with pytest.raises(capi.IPCGError):
        ctx.friction_energy(EPS_V)  # no velocities yet
    ctx.upload_velocities(np.zeros_like(s.V))
    ctx.friction_energy(EPS_V)
    ctx.build_collisions(3 * s.dhat)  # a new collision set discards the friction set
    with pytest.raises(capi.IPCGError):
        ctx.friction_energy(EPS_V)


def test_benchmark_size_properties(ctx, scenes):
    """stacked_sheets_8x512: zero velocities give the constant energy
    sum mu N eps_v / 3 and a zero gradient; uniform translation too"""
    s = scenes.stacked_sheets(512, 8, seed=0)
    say("case stacked_sheets_8x512", s.sizes)
    upload(ctx, s)
    ctx.build_collisions(s.dhat)
    ctx.friction_build(s.dhat, s.kappa, 0.5)
    keep, beta, basis, force = ctx.friction_records()
    cw = ctx.collisions()
    w = np.concatenate([cw["vv_w"], cw["ev_w"], cw["ee_w"], cw["fv_w"]])
    ctx.upload_velocities(np.zeros_like(s.V))
    e0 = ctx.friction_energy(EPS_V)
    want = 0.5 * EPS_V / 3 * np.sum((w * force)[keep])
    assert abs(e0 - want) <= 1e-9 * abs(want)
    assert np.all(ctx.friction_gradient(EPS_V) == 0)
    ctx.upload_velocities(np.tile([0.3, -0.2, 0.1], (len(s.V), 1)))
    assert abs(ctx.friction_energy(EPS_V) - e0) <= 1e-12 * abs(e0)
    assert np.abs(ctx.friction_gradient(EPS_V)).max() <= 1e-12 * np.abs(force).max()
