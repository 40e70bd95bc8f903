"""OGC collision sets on the device against the oracle.

The oracle's OGC set is pinned on the CPU (the reference's feasible-region
cases and codim counts, and count-once closed forms).  The device set built
from the device's own candidate lists must equal the oracle's on the same
lists: ids, distance types, eps_x and weights identical, with and without
area weighting; energy, gradient and projected Hessian within 1e-12 /
1e-12 / 1e-10 relative."""
import sys

import numpy as np
import pytest

from conftest import rel_err, sparse_rel_err
from test_gpu_improved import assert_same_set
from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu

Z2 = np.zeros((0, 2), np.int32)
Z3 = np.zeros((0, 3), np.int32)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_ogc(ctx, O, V, E, F, dhat, dmin=0.0, aw=False, ogc=True):
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    return O.build_collisions(V, E, F, ee, fv, dhat, dmin, vv_cand=vv, ev_cand=ev,
                              area_weighting=aw, ogc=ogc)


@pytest.mark.parametrize("aw", [False, True])
@pytest.mark.parametrize("name", ["sheets:16:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"])
def test_ogc_collisions(ctx, capi, O, scenes, name, aw):
    s = make(scenes, name)
    dhat = 3 * s.dhat
    say("case", name, "aw", aw)
    upload(ctx, s)
    ctx.set_area_weighting(aw)
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(dhat)
    co = oracle_ogc(ctx, O, s.V, s.E, s.F, dhat, aw=aw)
    assert_same_set(ctx.collisions(), co)
    ci = oracle_ogc(ctx, O, s.V, s.E, s.F, dhat, aw=aw, ogc=False)
    assert len(co) < len(ci)  # the scene drops infeasible pairs
    e_o = O.barrier_energy(co, s.V, s.E, s.F, dhat, s.kappa)
    assert abs(ctx.energy(dhat, s.kappa) - e_o) <= 1e-12 * abs(e_o)
    g_o = O.barrier_gradient(co, s.V, s.E, s.F, dhat, s.kappa)
    assert rel_err(ctx.gradient(dhat, s.kappa), g_o) <= 1e-12
    H_o = O.barrier_hessian(co, s.V, s.E, s.F, dhat, s.kappa, "clamp")
    assert sparse_rel_err(ctx.hessian(dhat, s.kappa, "clamp"), H_o) <= 1e-10


def test_reference_codim_counts(ctx, capi, O):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 2, 0],
                  [0, 3, 0]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], np.int32)
    say("case codim star")
    ctx.upload_mesh(V, E, Z3)
    ctx.upload_positions(0, V)
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(0.25, 0.8)
    assert tuple(ctx.collision_counts()) == (2, 4, 0, 0)
    assert_same_set(ctx.collisions(), oracle_ogc(ctx, O, V, E, Z3, 0.25, 0.8))


@pytest.mark.parametrize("aw", [False, True])
def test_point_over_flat_sheet(ctx, capi, O, scenes, aw):
    rng = np.random.default_rng(4)
    Vs, Fs = scenes.grid_sheet(33)
    V, E, F = scenes.merge([(Vs, Fs)])
    p = np.array([0.5 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05), 0.02])
    V = np.vstack([V, p[None]])
    group = np.zeros(len(V), np.int32)
    group[-1] = 1
    say("case point over sheet aw", aw)
    ctx.upload_mesh(V, E, F, vertex_patch=group)
    ctx.upload_positions(0, V)
    ctx.set_area_weighting(aw)
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(0.3)
    assert tuple(ctx.collision_counts()) == (0, 0, 0, 1)
    cg = ctx.collisions()
    assert cg["fv_w"][0] == (0.25 if aw else 1.0)
    assert_same_set(cg, oracle_ogc(ctx, O, V, E, F, 0.3, aw=aw))


def test_point_above_fan_centre(ctx, capi, O):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 0.02]])
    F = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 4], [0, 4, 1]], np.int32)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4], [1, 2], [2, 3], [3, 4], [4, 1]], np.int32)
    group = np.array([0, 0, 0, 0, 0, 1], np.int32)
    say("case fan centre")
    ctx.upload_mesh(V, E, F, vertex_patch=group)
    ctx.upload_positions(0, V)
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(0.05)
    cg = ctx.collisions()
    assert cg["vv"].tolist() == [[0, 5]] and cg["vv_w"].tolist() == [1.0]
    assert_same_set(cg, oracle_ogc(ctx, O, V, E, F, 0.05))
    ctx.set_collision_set_type(capi.COLLISIONS_IPC)
    ctx.build_collisions(0.05)
    assert ctx.collisions()["vv_w"].tolist() == [4.0]


def test_codim_corner(ctx, capi, O):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1], [-0.1, 0.1, -0.1]], float)
    E = np.array([[0, 1], [0, 2]], np.int32)
    say("case codim corner")
    ctx.upload_mesh(V, E, Z3)
    ctx.upload_positions(0, V)
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(0.5)
    cg = ctx.collisions()
    assert cg["vv"].tolist() == [[0, 3]] and cg["vv_w"].tolist() == [1.0]
    assert_same_set(cg, oracle_ogc(ctx, O, V, E, Z3, 0.5))


def test_benchmark_size_property(ctx, capi, scenes):
    """stacked_sheets_8x512: OGC keeps a subset of the IPC set's pairs"""
    s = scenes.stacked_sheets(512, 8, seed=0)
    say("case stacked_sheets_8x512", s.sizes)
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    ctx.build_collisions(s.dhat)
    ci = ctx.collisions()
    ctx.set_collision_set_type(capi.COLLISIONS_OGC)
    ctx.build_collisions(s.dhat)
    cg = ctx.collisions()
    for k in ("ee", "fv"):
        a = {tuple(p) for p in cg[k].tolist()}
        assert a <= {tuple(p) for p in ci[k].tolist()}, k
    assert np.all(cg["fv_w"] == 1.0) and np.all(cg["ee_w"] == 1.0)
    assert np.all(cg["vv_w"] > 0) and np.all(cg["ev_w"] > 0)
    assert np.isfinite(ctx.energy(s.dhat, s.kappa))
