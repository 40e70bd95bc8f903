"""IMPROVED_MAX_APPROX collision sets on the device against the oracle.

The oracle's improved set is pinned on the CPU (the reference's codim case,
the point-over-a-flat-sheet and vertex-under-an-edge counts, a mollified
case worked by hand).  Here the device set built from the device's own
candidate lists must equal the oracle's on the same lists: ids, distance
types and eps_x identical, weights bit-identical (both sum duplicates in
ascending order of value), with and without area weighting; energy,
gradient and projected Hessian of the signed set within 1e-12 / 1e-12 /
1e-10 relative."""
import sys

import numpy as np
import pytest

from conftest import rel_err, sparse_rel_err
from test_gpu_parity import make, upload

pytestmark = pytest.mark.gpu

Z2 = np.zeros((0, 2), np.int32)
Z3 = np.zeros((0, 3), np.int32)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def oracle_set(ctx, O, V, E, F, dhat, dmin=0.0, aw=False, improved=True):
    """the oracle's set over the candidate lists the device just built"""
    ee, fv = ctx.candidates()
    vv, ev = ctx.codim_candidates()
    return O.build_collisions(V, E, F, ee, fv, dhat, dmin, vv_cand=vv, ev_cand=ev,
                              area_weighting=aw, improved=improved)


def assert_same_set(cg, co):
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(cg[k], getattr(co, k)), k
        assert np.array_equal(cg[k + "_w"], getattr(co, k + "_w")), k + "_w"
    assert np.array_equal(cg["ee_dtype"], co.ee_dtype)
    assert np.array_equal(cg["ee_eps_x"], co.ee_eps_x)


@pytest.mark.parametrize("aw", [False, True])
@pytest.mark.parametrize("name", ["sheets:16:0", "twisted:12:1", "spheres:4:3:0", "stack:32:0"])
def test_improved_collisions(ctx, capi, O, scenes, name, aw):
    s = make(scenes, name)
    dhat = 3 * s.dhat  # wide enough that edges and vertices reduce
    say("case", name, "aw", aw)
    upload(ctx, s)
    ctx.set_area_weighting(aw)
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(dhat)
    co = oracle_set(ctx, O, s.V, s.E, s.F, dhat, aw=aw)
    cg = ctx.collisions()
    assert_same_set(cg, co)
    ci = oracle_set(ctx, O, s.V, s.E, s.F, dhat, aw=aw, improved=False)
    # the scene exercises the signed contributions
    assert any(len(getattr(co, k)) != len(getattr(ci, k))
               or not np.array_equal(getattr(co, k + "_w"), getattr(ci, k + "_w"))
               for k in ("vv", "ev", "ee"))
    assert np.array_equal(co.fv, ci.fv)  # face-vertex pairs are never converted
    e_o = O.barrier_energy(co, s.V, s.E, s.F, dhat, s.kappa)
    assert abs(ctx.energy(dhat, s.kappa) - e_o) <= 1e-12 * abs(e_o)
    g_o = O.barrier_gradient(co, s.V, s.E, s.F, dhat, s.kappa)
    assert rel_err(ctx.gradient(dhat, s.kappa), g_o) <= 1e-12
    for psd in ("clamp", "none"):
        H_o = O.barrier_hessian(co, s.V, s.E, s.F, dhat, s.kappa, psd)
        assert sparse_rel_err(ctx.hessian(dhat, s.kappa, psd), H_o) <= 1e-10, psd
    # back to the IPC set
    ctx.set_collision_set_type(capi.COLLISIONS_IPC)
    ctx.build_collisions(dhat)
    assert_same_set(ctx.collisions(), ci)


@pytest.mark.parametrize("aw", [False, True])
def test_reference_codim_star(ctx, capi, O, aw):
    # test_normal_collisions.cpp:120-200
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 0, -1], [-1, 0, 0], [0, 0, 1], [0, 1, 0], [0, 2, 0],
                  [0, 3, 0]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [0, 4]], np.int32)
    dhat, dmin = 0.25, 0.8
    say("case codim star aw", aw)
    ctx.upload_mesh(V, E, Z3)
    ctx.upload_positions(0, V)
    ctx.set_area_weighting(aw)
    for t, want in ((capi.COLLISIONS_IPC, (2, 4, 0, 0)),
                    (capi.COLLISIONS_IMPROVED_MAX_APPROX, (3, 4, 0, 0))):
        ctx.set_collision_set_type(t)
        ctx.build_collisions(dhat, dmin)
        assert tuple(ctx.collision_counts()) == want
        assert_same_set(ctx.collisions(),
                        oracle_set(ctx, O, V, E, Z3, dhat, dmin, aw=aw, improved=t == 1))


@pytest.mark.parametrize("aw", [False, True])
def test_point_over_flat_sheet(ctx, capi, O, scenes, aw):
    rng = np.random.default_rng(3)
    Vs, Fs = scenes.grid_sheet(33)
    V, E, F = scenes.merge([(Vs, Fs)])
    p = np.array([0.5 + rng.uniform(-0.05, 0.05), 0.5 + rng.uniform(-0.05, 0.05), 0.02])
    V = np.vstack([V, p[None]])
    # the sheet and the point in different patches: only the point's pairs
    group = np.zeros(len(V), np.int32)
    group[-1] = 1
    say("case point over sheet aw", aw)
    ctx.upload_mesh(V, E, F, vertex_patch=group)
    ctx.upload_positions(0, V)
    ctx.set_area_weighting(aw)
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(0.3)
    assert tuple(ctx.collision_counts()) == (0, 0, 0, 1)
    cg = ctx.collisions()
    assert cg["fv"][0, 1] == len(V) - 1 and cg["fv_w"][0] == (0.25 if aw else 1.0)
    assert_same_set(cg, oracle_set(ctx, O, V, E, F, 0.3, aw=aw))
    ctx.set_collision_set_type(capi.COLLISIONS_IPC)
    ctx.build_collisions(0.3)
    assert ctx.collision_counts().sum() > 100


@pytest.mark.parametrize("k", [3, 6])
def test_vertex_under_edge(ctx, capi, O, k):
    rng = np.random.default_rng(k)
    ang = np.sort(rng.uniform(0, 2 * np.pi, k))
    star = np.column_stack([0.3 * np.cos(ang), 0.3 * np.sin(ang), np.zeros(k)])
    V = np.vstack([[0.0, 0, 0], star, [-0.4, 0.05, 0.01], [0.4, -0.05, 0.01]])
    E = np.array([[0, i + 1] for i in range(k)] + [[k + 1, k + 2]], np.int32)
    say("case vertex under edge", k)
    ctx.upload_mesh(V, E, Z3)
    ctx.upload_positions(0, V)
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(0.05)
    cg = ctx.collisions()
    assert cg["ev"].tolist() == [[k, 0]] and cg["ev_w"].tolist() == [1.0]
    assert_same_set(cg, oracle_set(ctx, O, V, E, Z3, 0.05))


def test_mollified_star(ctx, capi, O):
    h = 0.01
    V = np.array([[0, 0, 0], [0.25, 0, 0], [0, 0.3, 0], [-0.2, -0.25, 0], [-0.5, 0, h],
                  [0.5, 0, h]], float)
    E = np.array([[0, 1], [0, 2], [0, 3], [4, 5]], np.int32)
    say("case mollified star")
    ctx.upload_mesh(V, E, Z3)
    ctx.upload_positions(0, V)
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(0.05)
    cg = ctx.collisions()
    co = oracle_set(ctx, O, V, E, Z3, 0.05)
    assert_same_set(cg, co)
    assert cg["ev"].tolist() == [[3, 0], [3, 1]]
    assert sorted(cg["ee_dtype"].tolist()) == sorted([4, 5] + [int(t) for t in co.ee_dtype
                                                               if t not in (4, 5)])
    assert np.sum(cg["ee_w"] == -1.0) == 2
    e_o = O.barrier_energy(co, V, E, Z3, 0.05, 1.0)
    assert abs(ctx.energy(0.05, 1.0) - e_o) <= 1e-12 * abs(e_o)


def test_rejects_unknown_type(ctx, capi):
    with pytest.raises(capi.IPCGError):
        ctx.set_collision_set_type(3)


def test_benchmark_size_property(ctx, capi, scenes):
    """stacked_sheets_8x512: the improved set keeps the IPC set's face-vertex
    collisions, adds signed vertex-vertex / edge-vertex / edge-edge ones and
    drops none with zero weight"""
    s = scenes.stacked_sheets(512, 8, seed=0)
    say("case stacked_sheets_8x512", s.sizes)
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    ctx.set_area_weighting(True)
    ctx.build_collisions(s.dhat)
    ci = ctx.collisions()
    ctx.set_collision_set_type(capi.COLLISIONS_IMPROVED_MAX_APPROX)
    ctx.build_collisions(s.dhat)
    cg = ctx.collisions()
    assert np.array_equal(cg["fv"], ci["fv"]) and np.array_equal(cg["fv_w"], ci["fv_w"])
    for k in ("vv", "ev", "ee"):
        assert np.all(cg[k + "_w"] != 0), k
        assert np.all(np.isfinite(cg[k + "_w"])), k
    assert np.isfinite(ctx.energy(s.dhat, s.kappa))
