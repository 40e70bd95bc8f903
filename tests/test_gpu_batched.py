"""Batched contexts (ipcg_mesh_upload_batched): scenes of uneven size in one
context, every stage one launch over all of them.  Each scene's results are
compared with a context holding that scene alone, through the C ABI:
candidate sets (static and swept, with their orientation), collision sets
(ids, weights, eps_x, distance types), energies, gradients, projected
Hessians and step sizes, all bit-identical (the batch runs the same
arithmetic in the same order per scene).  Also: scenes left inactive report
nothing and do not change the others, and a subset of the batched-form
workload (uneven sheet pairs and icosphere rows)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_batch(scenes):
    return scenes.Batch([
        scenes.two_sheets(12, seed=1),
        scenes.sphere_chain(3, 2, seed=2),
        scenes.stacked_sheets(16, 4, seed=3, h=1.0 / 511),
        scenes.two_sheets(20, seed=4, twist=0.3),
        scenes.sphere_chain(2, 3, seed=5),
        scenes.two_sheets(2, seed=6),  # tiny: single-leaf trees
    ])


def upload_alone(ctx, s):
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    ctx.upload_positions(1, s.V1)


def slice_pairs(a, off0, off1, k):
    """rows of the global pair list `a` that belong to scene k (first id in
    the scene's range of off0), shifted to scene-local ids"""
    a = np.asarray(a).reshape(-1, 2)
    m = (a[:, 0] >= off0[k]) & (a[:, 0] < off0[k + 1])
    out = a[m] - np.array([off0[k], off1[k]])
    return out, m


def check_candidates(ctx, alone, B, dhat):
    for swept in (False, True):
        r = 0.0 if swept else 0.5 * dhat
        ctx.broad_phase(swept, r)
        ee, fv = ctx.candidates()
        n_ee = n_fv = 0
        for k, s in enumerate(B.scenes):
            alone.upload_mesh(s.V, s.E, s.F)
            alone.upload_positions(0, s.V)
            alone.upload_positions(1, s.V1)
            alone.broad_phase(swept, r)
            ea, fa = alone.candidates()
            eb, m = slice_pairs(ee, B.eoff, B.eoff, k)
            assert np.all((eb >= 0) & (eb < len(s.E))), "edge-edge pair across scenes"
            fb, _ = slice_pairs(fv, B.foff, B.voff, k)
            assert np.all((fb[:, 1] >= 0) & (fb[:, 1] < len(s.V))), "face-vertex pair across scenes"
            # same oriented pairs (orientation follows the scene's Morton order)
            assert sorted(map(tuple, eb.tolist())) == sorted(map(tuple, ea.tolist())), (k, swept)
            assert sorted(map(tuple, fb.tolist())) == sorted(map(tuple, fa.tolist())), (k, swept)
            n_ee += len(eb)
            n_fv += len(fb)
        assert n_ee == len(ee) and n_fv == len(fv)


def check_step(ctx, alone, B, dhat, kappa, hessian=True):
    ctx.build_collisions(dhat)
    counts = ctx.collision_counts_batched()
    cg = ctx.collisions()
    energies = ctx.energy_batched(dhat, kappa)
    g = ctx.gradient(dhat, kappa)
    H = ctx.hessian(dhat, kappa, "clamp") if hessian else None
    steps = ctx.stepsize_batched()
    assert np.array_equal(counts.sum(0), [len(cg[k]) for k in ("vv", "ev", "ee", "fv")])
    offs = {"vv": (B.voff, B.voff), "ev": (B.eoff, B.voff), "ee": (B.eoff, B.eoff),
            "fv": (B.foff, B.voff)}
    first = np.zeros(4, np.int64)
    for k, s in enumerate(B.scenes):
        upload_alone(alone, s)
        alone.build_collisions(dhat)
        ca = alone.collisions()
        for j, key in enumerate(("vv", "ev", "ee", "fv")):
            n = counts[k, j]
            o0, o1 = offs[key][0][k], offs[key][1][k]
            ids = cg[key][first[j]:first[j] + n] - np.array([o0, o1])
            assert np.array_equal(ids, ca[key]), (k, key)
            assert np.array_equal(cg[key + "_w"][first[j]:first[j] + n], ca[key + "_w"]), (k, key)
            if key == "ee":
                assert np.array_equal(cg["ee_eps_x"][first[j]:first[j] + n], ca["ee_eps_x"])
                assert np.array_equal(cg["ee_dtype"][first[j]:first[j] + n], ca["ee_dtype"])
            first[j] += n
        e = alone.energy(dhat, kappa)
        assert energies[k] == e, (k, energies[k], e)
        r0, r1 = 3 * B.voff[k], 3 * B.voff[k + 1]
        assert np.array_equal(g[r0:r1], alone.gradient(dhat, kappa)), k
        if hessian:
            Ha = alone.hessian(dhat, kappa, "clamp")
            Hs = H[r0:r1, r0:r1]
            assert H[r0:r1].nnz == Hs.nnz, "Hessian entries across scenes"
            assert np.array_equal(Hs.indptr, Ha.indptr) and np.array_equal(Hs.indices, Ha.indices)
            assert np.array_equal(Hs.data, Ha.data), k
        t = alone.stepsize()
        assert steps[k] == t, (k, steps[k], t)
    return counts, energies, steps


def test_batched_equals_alone(ctx, capi, scenes):
    B = small_batch(scenes)
    say("batch", len(B), B.sizes)
    dhat, kappa = 1e-3, 1e3
    ctx.upload_batch(B)
    assert ctx.batch_size() == len(B)
    ctx.upload_positions(0, B.V)
    ctx.upload_positions(1, B.V1)
    alone = capi.Context(0)
    try:
        check_candidates(ctx, alone, B, dhat)
        counts, energies, steps = check_step(ctx, alone, B, dhat, kappa)
        say("counts", counts.tolist(), "steps", steps.tolist())
        assert counts.sum() > 0 and np.all(steps <= 1.0)
    finally:
        alone.close()


def test_batched_inactive_scenes(ctx, capi, scenes):
    B = small_batch(scenes)
    dhat, kappa = 1e-3, 1e3
    ctx.upload_batch(B)
    ctx.upload_positions(0, B.V)
    ctx.upload_positions(1, B.V1)
    ctx.build_collisions(dhat)
    full_counts = ctx.collision_counts_batched()
    full_e = ctx.energy_batched(dhat, kappa)
    full_t = ctx.stepsize_batched()
    active = np.array([k % 2 == 0 for k in range(len(B))])
    ctx.set_active(active)
    ctx.build_collisions(dhat)
    counts = ctx.collision_counts_batched()
    e = ctx.energy_batched(dhat, kappa)
    t = ctx.stepsize_batched()
    say("active", active.tolist(), counts.tolist(), e.tolist(), t.tolist())
    for k in range(len(B)):
        if active[k]:
            assert np.array_equal(counts[k], full_counts[k]) and e[k] == full_e[k]
            assert t[k] == full_t[k]
        else:
            assert counts[k].sum() == 0 and e[k] == 0.0 and t[k] == 1.0
    ctx.set_active(None)
    ctx.build_collisions(dhat)
    assert np.array_equal(ctx.collision_counts_batched(), full_counts)


def test_batched_rejects_sweep_and_prune(ctx, capi, scenes):
    B = small_batch(scenes)
    ctx.upload_batch(B)
    ctx.upload_positions(0, B.V)
    with pytest.raises(capi.IPCGError):
        ctx.broad_phase(False, 1e-3, capi.BP_SAP)


def test_batched_workload_subset(ctx, capi, scenes):
    """the first 24 scenes of the batched-form workload (sheet pairs of
    uneven side and rows of icospheres, 2.6k to 300k triangles each)"""
    B = scenes.Batch(scenes.batch_scenes(24, seed=0))  # the first 24 of the 256
    say("batch", len(B), B.sizes)
    dhat, kappa = 1e-3, 1e3
    ctx.upload_batch(B)
    ctx.upload_positions(0, B.V)
    ctx.upload_positions(1, B.V1)
    alone = capi.Context(0)
    try:
        counts, energies, steps = check_step(ctx, alone, B, dhat, kappa, hessian=False)
        say("collisions", int(counts.sum()), "steps", np.round(steps, 4).tolist())
    finally:
        alone.close()


def test_autotuned_launch_shapes(ctx, capi, scenes):
    """the autotuner times each traversal at 64, 128 and 256 lanes per
    workgroup on the first broad phase and keeps the fastest; the results
    are those of the default shape"""
    B = small_batch(scenes)
    dhat, kappa = 1e-3, 1e3
    ref = capi.Context(0)
    try:
        for c in (ref, ctx):
            c.upload_batch(B)
            c.upload_positions(0, B.V)
            c.upload_positions(1, B.V1)
        ctx.set_autotune(True)
        for c in (ref, ctx):
            c.build_collisions(dhat)
        for stage in ("bp_traverse_ee.static", "bp_traverse_fv.static"):
            b, ms = ctx.launch_shape(stage)
            say(stage, b, ms)
            assert b in (64, 128, 256) and ms > 0
            assert ref.launch_shape(stage)[0] == 0
        a, r = ctx.collisions(), ref.collisions()
        for k in a:
            assert np.array_equal(a[k], r[k]), k
        assert np.array_equal(ctx.energy_batched(dhat, kappa), ref.energy_batched(dhat, kappa))
        assert np.array_equal(ctx.stepsize_batched(), ref.stepsize_batched())
        assert ctx.launch_shape("bp_traverse_ee.swept")[0] in (64, 128, 256)
    finally:
        ref.close()
