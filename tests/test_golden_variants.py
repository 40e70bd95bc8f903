"""Regression fixtures of the collision-set variants and friction
(tests/golden/variants/*.npz, written by tests/golden/make_golden.py): the
oracle reproduces them, and on a GPU the device path does too (sets and
weights exactly, friction energy / gradient / Hessian within 1e-12 /
1e-12 / 1e-10)."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err, sparse_rel_err

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "variants", "*.npz")))
KEYS = ("vv", "ev", "ee", "fv", "vv_w", "ev_w", "ee_w", "fv_w", "ee_eps_x", "ee_dtype")


def load(path):
    return dict(np.load(path, allow_pickle=False))


def csr(g):
    import scipy.sparse as sp
    n = 3 * len(g["V"])
    return sp.csr_matrix((g["fric_H_data"], g["fric_H_indices"], g["fric_H_indptr"]), shape=(n, n))


def test_fixtures_present():
    assert FIXTURES


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_variants(O, path):
    g = load(path)
    V, E, F, dhat, kappa = g["V"], g["E"], g["F"], float(g["dhat"]), float(g["kappa"])
    ee, fv = O.candidates(V, E, F, 0.5 * dhat)
    for tag, kw in (("imp", dict(area_weighting=True, improved=True)), ("ogc", dict(ogc=True))):
        c = O.build_collisions(V, E, F, ee, fv, dhat, **kw)
        for k in KEYS:
            assert np.array_equal(getattr(c, k), g[f"{tag}_{k}"]), (tag, k)
    c = O.build_collisions(V, E, F, ee, fv, dhat)
    fr = O.friction_build(c, V, E, F, dhat, kappa, 0.6, 0.3)
    O.friction_set_anisotropy(fr, g["mu_aniso"], g["mu_s_aniso"], g["mu_k_aniso"])
    O.friction_update_lagged_mu(fr, g["Vel"])
    ev = float(g["eps_v"])
    assert O.friction_energy(fr, g["Vel"], ev) == pytest.approx(float(g["fric_energy"]), rel=1e-12)
    assert rel_err(O.friction_gradient(fr, g["Vel"], ev), g["fric_gradient"]) <= 1e-12
    assert sparse_rel_err(O.friction_hessian(fr, g["Vel"], ev, "clamp"), csr(g)) <= 1e-10


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_device_reproduces_variants(ctx, capi, path):
    g = load(path)
    V, E, F, dhat, kappa = g["V"], g["E"], g["F"], float(g["dhat"]), float(g["kappa"])
    print("case", os.path.basename(path), flush=True)
    ctx.upload_mesh(V, E, F)
    ctx.upload_positions(0, V)
    for tag, aw, kind in (("imp", True, capi.COLLISIONS_IMPROVED_MAX_APPROX),
                          ("ogc", False, capi.COLLISIONS_OGC)):
        ctx.set_area_weighting(aw)
        ctx.set_collision_set_type(kind)
        ctx.build_collisions(dhat)
        c = ctx.collisions()
        for k in KEYS:
            assert np.array_equal(c[k], g[f"{tag}_{k}"]), (tag, k)
    ctx.set_area_weighting(False)
    ctx.set_collision_set_type(capi.COLLISIONS_IPC)
    ctx.build_collisions(dhat)
    ctx.friction_build(dhat, kappa, 0.6, 0.3)
    ctx.friction_set_anisotropy(g["mu_aniso"], g["mu_s_aniso"], g["mu_k_aniso"])
    ctx.upload_velocities(g["Vel"])
    ctx.friction_update_lagged_mu()
    ev = float(g["eps_v"])
    assert ctx.friction_energy(ev) == pytest.approx(float(g["fric_energy"]), rel=1e-12)
    assert rel_err(ctx.friction_gradient(ev), g["fric_gradient"]) <= 1e-12
    assert sparse_rel_err(ctx.friction_hessian(ev, "clamp"), csr(g)) <= 1e-10
