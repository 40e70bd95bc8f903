"""BarrierPotential use_physical_barrier on the device
(barrier_potential.cpp:33-99): the barrier and its derivatives scaled by
dhat / units with units = ((2 dmin + dhat) dhat)^2 (barrier.hpp:133-137),
so energy, gradient, Hessian and the Gauss-Newton diagonal equal the
oracle's at stiffness kappa * dhat / units (1e-12 / 1e-12 / 1e-10; the
device multiplies the factor into b before kappa, as the reference does),
and turning it off restores the plain values exactly."""
import numpy as np
import pytest

from conftest import rel_err, sparse_rel_err
from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,dmin", [("sheets:16:0", 0.0), ("spheres:4:3:0", 0.0),
                                       ("twisted:12:1", 2e-4)])
def test_physical_barrier(ctx, O, scenes, name, dmin):
    s = make(scenes, name)
    print("case", name, dmin, flush=True)
    upload(ctx, s)
    dhat = 2 * s.dhat
    ctx.build_collisions(dhat, dmin)
    e0 = ctx.energy(dhat, s.kappa)
    ee, fv = ctx.candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, dmin)
    t = (2 * dmin + dhat) * dhat
    k_eff = s.kappa * dhat / (t * t)
    ctx.set_physical_barrier(True)
    e_o = O.barrier_energy(co, s.V, s.E, s.F, dhat, k_eff)
    assert abs(ctx.energy(dhat, s.kappa) - e_o) <= 1e-12 * abs(e_o)
    assert rel_err(ctx.gradient(dhat, s.kappa), O.barrier_gradient(co, s.V, s.E, s.F, dhat, k_eff)) <= 1e-12
    H_o = O.barrier_hessian(co, s.V, s.E, s.F, dhat, k_eff, "clamp")
    assert sparse_rel_err(ctx.hessian(dhat, s.kappa, "clamp"), H_o) <= 1e-10
    d_o = O.gauss_newton_diagonal(co, s.V, s.E, s.F, dhat, k_eff)
    d_g = ctx.gauss_newton_diagonal(dhat, s.kappa)
    assert np.linalg.norm(d_g - d_o) <= 1e-12 * np.linalg.norm(d_o)
    ctx.set_physical_barrier(False)
    assert ctx.energy(dhat, s.kappa) == e0


@pytest.mark.parametrize("name,dmin", [("sheets:16:0", 0.0), ("twisted:12:1", 2e-4)])
def test_physical_barrier_friction_and_quadratic_form(ctx, O, scenes, name, dmin):
    """with the physical barrier the friction build's normal forces and the
    friction potential equal the oracle's at stiffness kappa * dhat / units,
    and so does the Gauss-Newton quadratic form"""
    s = make(scenes, name)
    print("case", name, dmin, flush=True)
    upload(ctx, s)
    dhat = 2 * s.dhat
    ctx.build_collisions(dhat, dmin)
    ee, fv = ctx.candidates()
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, dhat, dmin)
    t = (2 * dmin + dhat) * dhat
    k_eff = s.kappa * dhat / (t * t)
    ctx.set_physical_barrier(True)
    p = np.random.default_rng(3).normal(size=3 * len(s.V))
    q_o = O.gauss_newton_quadratic_form(co, s.V, s.E, s.F, dhat, k_eff, p)
    assert abs(ctx.gauss_newton_quadratic_form(dhat, s.kappa, p) - q_o) <= 1e-12 * abs(q_o)
    ctx.friction_build(dhat, s.kappa, 0.5, 0.3)
    fr = O.friction_build(co, s.V, s.E, s.F, dhat, k_eff, 0.5, 0.3)
    keep, beta, basis, force = ctx.friction_records()
    assert np.array_equal(keep, fr.keep) and keep.any()
    assert np.allclose(force[keep], fr.normal_force[keep], rtol=1e-13, atol=0)
    eps_v = 1e-3
    vel = np.random.default_rng(4).normal(0, 2 * eps_v, s.V.shape)
    ctx.upload_velocities(vel)
    e_o = O.friction_energy(fr, vel, eps_v)
    assert abs(ctx.friction_energy(eps_v) - e_o) <= 1e-12 * abs(e_o)
    assert rel_err(ctx.friction_gradient(eps_v), O.friction_gradient(fr, vel, eps_v)) <= 1e-12
    ctx.set_physical_barrier(False)
