"""Regression fixtures (tests/golden/*.npz, written by make_golden.py): the
oracle reproduces them, and on a GPU the device path does too."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err, sparse_rel_err

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(path):
    return dict(np.load(path, allow_pickle=False))


def canon(a):
    a = np.asarray(a, dtype=np.int32).reshape(-1, 2)
    return a[np.lexsort((a[:, 1], a[:, 0]))] if len(a) else a


def csr(g):
    import scipy.sparse as sp
    n = 3 * len(g["V"])
    return sp.csr_matrix((g["H_data"], g["H_indices"], g["H_indptr"]), shape=(n, n))


@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_oracle_reproduces_golden(O, path):
    g = load(path)
    V, V1, E, F, dhat, kappa = g["V"], g["V1"], g["E"], g["F"], float(g["dhat"]), float(g["kappa"])
    ee, fv = O.candidates(V, E, F, 0.5 * dhat)
    assert np.array_equal(canon(ee), g["cand_ee"]) and np.array_equal(canon(fv), g["cand_fv"])
    c = O.build_collisions(V, E, F, ee, fv, dhat)
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(getattr(c, k), g[k]) and np.array_equal(getattr(c, k + "_w"), g[k + "_w"])
    assert np.array_equal(c.ee_eps_x, g["ee_eps_x"]) and np.array_equal(c.ee_dtype, g["ee_dtype"])
    assert O.barrier_energy(c, V, E, F, dhat, kappa) == pytest.approx(float(g["energy"]), rel=1e-12)
    assert rel_err(O.barrier_gradient(c, V, E, F, dhat, kappa), g["gradient"]) <= 1e-12
    assert sparse_rel_err(O.barrier_hessian(c, V, E, F, dhat, kappa, "clamp"), csr(g)) <= 1e-10
    see, sfv = O.candidates(V, E, F, 0.0, V1=V1)
    assert np.array_equal(canon(see), g["swept_ee"]) and np.array_equal(canon(sfv), g["swept_fv"])
    assert O.compute_collision_free_stepsize(V, V1, E, F) == float(g["stepsize"])
    if "toi_ee" in g:
        assert np.array_equal(O.candidates_toi(V, V1, E, F, g["swept_ee"], True), g["toi_ee"])
        assert np.array_equal(O.candidates_toi(V, V1, E, F, g["swept_fv"], False), g["toi_fv"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=os.path.basename)
def test_device_reproduces_golden(ctx, path):
    g = load(path)
    V, V1, E, F, dhat, kappa = g["V"], g["V1"], g["E"], g["F"], float(g["dhat"]), float(g["kappa"])
    print("case", os.path.basename(path), flush=True)
    ctx.upload_mesh(V, E, F)
    ctx.upload_positions(0, V)
    ctx.upload_positions(1, V1)
    ctx.broad_phase(False, 0.5 * dhat)
    ee, fv = ctx.candidates()
    assert np.array_equal(canon(ee), g["cand_ee"]) and np.array_equal(canon(fv), g["cand_fv"])
    ctx.build_collisions(dhat)
    c = ctx.collisions()
    for k in ("vv", "ev", "ee", "fv", "vv_w", "ev_w", "ee_w", "fv_w", "ee_eps_x", "ee_dtype"):
        assert np.array_equal(c[k], g[k]), k
    assert ctx.energy(dhat, kappa) == pytest.approx(float(g["energy"]), rel=1e-12)
    assert rel_err(ctx.gradient(dhat, kappa), g["gradient"]) <= 1e-12
    assert sparse_rel_err(ctx.hessian(dhat, kappa, "clamp"), csr(g)) <= 1e-10
    assert ctx.stepsize() == float(g["stepsize"])
    if "toi_ee" in g:
        ctx.broad_phase(True, 0.0)
        see, sfv = ctx.candidates()
        te, tf = ctx.candidates_toi()
        me = {tuple(p): t for p, t in zip(g["swept_ee"].tolist(), g["toi_ee"])}
        mf = {tuple(p): t for p, t in zip(g["swept_fv"].tolist(), g["toi_fv"])}
        assert all(me[tuple(p)] == t for p, t in zip(see.tolist(), te))
        assert all(mf[tuple(p)] == t for p, t in zip(sfv.tolist(), tf))
