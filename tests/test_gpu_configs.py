"""The full contact step on the benchmark configurations at their real sizes.

* Configuration 1: two stacked 64 x 64 sheets (two_sheets(64), seeds 0-2):
  every result of the step against the oracle.
* Configuration 2: 8 interleaved 512 x 512 sheets (stacked_sheets(512, 8)):
  a 96 x 96 tile of it (same spacing) against the oracle in full, and the
  full-size scene against the oracle for the collision set, energy,
  gradient, step size and sampled Hessian rows.
* Configuration 3: configuration 1 with the FP32 broad phase: the FP32
  candidate superset against the FP64-box candidate set (the oracle's sweep
  and prune, pinned against an FP64 brute force), the identical collision
  set and the error of every quantity against the FP64 oracle.

Bars: candidate and collision sets, weights, eps_x, distance types,
per-candidate TOIs and step sizes bit-identical; energy and gradient within
1e-12 relative, Hessians within 1e-10 relative (Frobenius norm over the
union pattern).

The step size of the reference (candidates.cpp:269-287) passes the running
minimum as each query's tmax, so its outcome depends on the thread
schedule.  The device searches every candidate on [0, 1], which is the
outcome of the schedule in which every query starts before any finishes;
it equals the oracle's restatement of that schedule bit for bit.  Against
the serial running-minimum schedule (forward and reversed candidate order)
the gap is bounded by Tight-Inclusion's time resolution; both outcomes are
conservative.
"""
import sys
import time

import numpy as np
import pytest

from conftest import pair_set, rel_err, sparse_rel_err, unordered_set

pytestmark = pytest.mark.gpu

TOL_SUM = 1e-12
TOL_HESS = 1e-10
# Tight-Inclusion resolves time only to about tolerance / (3 |displacement|)
# (1e-6 / 3e-3 at these steps); schedules differ by at most a few of those.
TOL_SCHEDULE = 5e-3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(scenes, name):
    kind, *a = name.split(":")
    if kind == "config1":
        return scenes.two_sheets(64, seed=int(a[0]))
    if kind == "config2_tile":
        return scenes.stacked_sheets(96, 8, seed=int(a[0]), h=1.0 / 511)
    raise ValueError(name)


def upload(ctx, s):
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    ctx.upload_positions(1, s.V1)


def compare_sets(cg, co):
    for k in ("vv", "ev", "ee", "fv"):
        assert np.array_equal(getattr(co, k), cg[k]), k
        assert np.array_equal(getattr(co, k + "_w"), cg[k + "_w"]), k
    assert np.array_equal(co.ee_eps_x, cg["ee_eps_x"])
    assert np.array_equal(co.ee_dtype, cg["ee_dtype"])


@pytest.mark.parametrize("name", ["config1:0", "config2_tile:0"])
def test_config_step(ctx, O, scenes, name):
    """candidates (static and swept), collision set, E, grad, Hessian under
    every projection and the step size of one configuration"""
    s = make(scenes, name)
    say("case", name, s.sizes)
    upload(ctx, s)
    for swept in (False, True):
        r = 0.0 if swept else 0.5 * s.dhat
        ctx.broad_phase(swept, r)
        ee_g, fv_g = ctx.candidates()
        ee_o, fv_o = O.candidates(s.V, s.E, s.F, r, V1=s.V1 if swept else None)
        assert len(ee_g) == len(ee_o) and pair_set(ee_g) == pair_set(ee_o)
        assert len(fv_g) == len(fv_o) and pair_set(fv_g) == pair_set(fv_o)
        say("  candidates", "swept" if swept else "static", len(ee_g), len(fv_g))
    ctx.build_collisions(s.dhat)
    cg = ctx.collisions()
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat)
    compare_sets(cg, co)
    say("  collisions", [len(co.vv), len(co.ev), len(co.ee), len(co.fv)])
    Eo = O.barrier_energy(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert abs(ctx.energy(s.dhat, s.kappa) - Eo) <= TOL_SUM * abs(Eo)
    go = O.barrier_gradient(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert rel_err(ctx.gradient(s.dhat, s.kappa), go) <= TOL_SUM
    for psd in ("none", "clamp", "abs"):
        Ho = O.barrier_hessian(co, s.V, s.E, s.F, s.dhat, s.kappa, psd)
        Hg = ctx.hessian(s.dhat, s.kappa, psd)
        err = sparse_rel_err(Hg, Ho)
        say("  hessian", psd, "nnz", Hg.nnz, "rel err", err)
        assert err <= TOL_HESS, psd
    t = ctx.stepsize()
    assert t == O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F)
    assert O.is_step_collision_free(s.V, s.V + t * (s.V1 - s.V), s.E, s.F)
    say("  step", t)


@pytest.mark.parametrize("name", ["config1:0", "config1:1", "config1:2", "config2_tile:0"])
def test_ccd_schedules(ctx, O, scenes, name):
    """device step against the reference's running-minimum schedule"""
    s = make(scenes, name)
    say("case", name, s.sizes)
    upload(ctx, s)
    t_dev = ctx.stepsize()
    ctx.broad_phase(True, 0.0)
    ee, fv = ctx.candidates()
    threads = O.get_max_threads()
    O.set_num_threads(1)  # serial: a deterministic schedule
    try:
        t_all = O.candidates_stepsize(s.V, s.V1, s.E, s.F, ee, fv)
        t_fw = O.candidates_stepsize(s.V, s.V1, s.E, s.F, ee, fv, running_tmax=True)
        t_rv = O.candidates_stepsize(s.V, s.V1, s.E, s.F, ee[::-1].copy(), fv[::-1].copy(),
                                     running_tmax=True)
    finally:
        O.set_num_threads(threads)
    say(f"  device {t_dev!r} serial forward {t_fw!r} reversed {t_rv!r}; "
        f"rel gap {(t_dev - t_fw) / t_fw:+.3e} / {(t_dev - t_rv) / t_rv:+.3e}; "
        f"reference spread {(t_fw - t_rv) / t_rv:+.3e}")
    assert t_dev == t_all
    for t_ref in (t_fw, t_rv):
        assert abs(t_dev - t_ref) <= TOL_SCHEDULE * t_ref
    for t in (t_dev, t_fw, t_rv):
        assert 0.0 < t <= 1.0
        assert O.is_step_collision_free(s.V, s.V + t * (s.V1 - s.V), s.E, s.F)


def orient_like(ee, ref):
    """the pairs of `ee` oriented as `ref` holds them (every pair of `ee`
    is in `ref` up to orientation)"""
    def key(a):
        a = np.asarray(a, dtype=np.int64).reshape(-1, 2)
        return (np.minimum(a[:, 0], a[:, 1]) << 32) | np.maximum(a[:, 0], a[:, 1])
    kr = key(ref)
    order = np.argsort(kr)
    pos = np.searchsorted(kr[order], key(ee))
    assert np.array_equal(kr[order][pos], key(ee))
    return np.asarray(ref)[order[pos]]


@pytest.mark.parametrize("name", ["config1:0", "config2_tile:0"])
def test_config3_precision_form(ctx, O, scenes, capi, name):
    """FP32 outward-rounded LBVH boxes report a superset of the FP64 box
    overlaps; the FP64 narrow phase reduces both to the same collision set,
    and every quantity is compared with the FP64 oracle"""
    s = make(scenes, name)
    say("case config3", name, s.sizes)
    upload(ctx, s)
    r = 0.5 * s.dhat
    ctx.broad_phase(False, r, capi.BP_LBVH)
    ee32, fv32 = ctx.candidates()
    ctx.broad_phase(False, r, capi.BP_SAP)
    ee64_g, fv64_g = ctx.candidates()
    ee64, fv64 = O.candidates(s.V, s.E, s.F, r, method="sap")
    # the device's FP64 sweep equals the oracle's FP64 box set
    assert unordered_set(ee64_g) == unordered_set(ee64) and len(ee64_g) == len(ee64)
    assert pair_set(fv64_g) == pair_set(fv64) and len(fv64_g) == len(fv64)
    # superset
    assert unordered_set(ee64) <= unordered_set(ee32)
    assert pair_set(fv64) <= pair_set(fv32)
    extra_ee, extra_fv = len(ee32) - len(ee64), len(fv32) - len(fv64)
    say(f"  FP32 candidates ee {len(ee32)} fv {len(fv32)}; FP64 ee {len(ee64)} fv {len(fv64)}; "
        f"superset +{extra_ee} ee, +{extra_fv} fv "
        f"({(extra_ee + extra_fv) / max(1, len(ee64) + len(fv64)):.3%})")
    # the collision set from the FP32 candidates is the one from the FP64
    # candidates (edge-edge pairs oriented as the LBVH orients them: an
    # edge-edge collision keeps its candidate's orientation)
    ctx.build_collisions(s.dhat, method=capi.BP_LBVH)
    cg = ctx.collisions()
    co = O.build_collisions(s.V, s.E, s.F, orient_like(ee64, ee32), fv64, s.dhat)
    compare_sets(cg, co)
    # per-quantity error against FP64
    Eo = O.barrier_energy(co, s.V, s.E, s.F, s.dhat, s.kappa)
    e_E = abs(ctx.energy(s.dhat, s.kappa) - Eo) / abs(Eo)
    e_g = rel_err(ctx.gradient(s.dhat, s.kappa), O.barrier_gradient(co, s.V, s.E, s.F, s.dhat,
                                                                    s.kappa))
    e_H = sparse_rel_err(ctx.hessian(s.dhat, s.kappa, "clamp"),
                         O.barrier_hessian(co, s.V, s.E, s.F, s.dhat, s.kappa, "clamp"))
    t_g = ctx.stepsize()
    t_o = O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F)
    say(f"  rel err vs FP64: energy {e_E:.2e} gradient {e_g:.2e} hessian {e_H:.2e} "
        f"step {abs(t_g - t_o) / t_o:.2e}")
    assert e_E <= TOL_SUM and e_g <= TOL_SUM and e_H <= TOL_HESS and t_g == t_o


def test_config2_full_size(ctx, O, scenes):
    """8 x 512^2 sheets: collision set, energy, gradient and step size against
    the oracle over the whole scene; Hessian rows of 48 sampled vertices
    against the oracle's Hessian of the collisions incident to them"""
    t0 = time.time()
    s = scenes.stacked_sheets(512, 8, seed=0)
    say("case config2", s.sizes)
    upload(ctx, s)
    ctx.build_collisions(s.dhat)
    cg = ctx.collisions()
    Eg = ctx.energy(s.dhat, s.kappa)
    gg = ctx.gradient(s.dhat, s.kappa)
    Hg = ctx.hessian(s.dhat, s.kappa, "clamp")
    t_g = ctx.stepsize()
    say(f"  device done {time.time() - t0:.1f}s", [len(cg[k]) for k in ("vv", "ev", "ee", "fv")])
    ee, fv = O.candidates(s.V, s.E, s.F, 0.5 * s.dhat)
    say(f"  oracle candidates {len(ee)} {len(fv)} {time.time() - t0:.1f}s")
    co = O.build_collisions(s.V, s.E, s.F, ee, fv, s.dhat)
    del ee, fv
    say(f"  oracle collisions {len(co)} {time.time() - t0:.1f}s")
    compare_sets(cg, co)
    del cg
    Eo = O.barrier_energy(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert abs(Eg - Eo) <= TOL_SUM * abs(Eo)
    go = O.barrier_gradient(co, s.V, s.E, s.F, s.dhat, s.kappa)
    assert rel_err(gg, go) <= TOL_SUM
    del go, gg
    say(f"  energy, gradient {time.time() - t0:.1f}s")
    # sampled Hessian rows: a vertex's rows are the sum over the collisions
    # incident to it, so the oracle's Hessian of exactly those collisions
    # has the same rows
    rng = np.random.default_rng(7)
    sample = rng.choice(len(s.V), 48, replace=False)
    mark = np.zeros(len(s.V), bool)
    mark[sample] = True
    E, F = s.E, s.F
    k_vv = mark[co.vv].any(1)
    k_ev = mark[E[co.ev[:, 0]]].any(1) | mark[co.ev[:, 1]]
    k_ee = mark[E[co.ee[:, 0]]].any(1) | mark[E[co.ee[:, 1]]].any(1)
    k_fv = mark[F[co.fv[:, 0]]].any(1) | mark[co.fv[:, 1]]
    sub = O.collisions_subset(co, k_vv, k_ev, k_ee, k_fv)
    say("  sampled collisions", len(sub))
    Ho = O.barrier_hessian(sub, s.V, s.E, s.F, s.dhat, s.kappa, "clamp")
    rows = (3 * sample[:, None] + np.arange(3)[None, :]).reshape(-1)
    err = sparse_rel_err(Hg[rows], Ho[rows])
    say(f"  hessian sampled rows rel err {err:.2e}")
    assert err <= TOL_HESS
    del Hg, Ho
    t_o = O.compute_collision_free_stepsize(s.V, s.V1, s.E, s.F)
    say(f"  step device {t_g!r} oracle {t_o!r} {time.time() - t0:.1f}s")
    assert t_g == t_o


def test_ccd_tiers_on_config2_tile(ctx, O, scenes):
    """the configuration-2 tile sends searches through all three CCD tiers
    (LDS lists, per-lane HBM slabs, the wavefront): per-candidate TOIs
    bit-identical to the oracle, and the step with the running minimum"""
    s = make(scenes, "config2_tile:0")
    say("case config2_tile tiers", s.sizes)
    upload(ctx, s)
    t = ctx.stepsize()
    bounded = {k: ctx.counter("ccd." + k) for k in ("searched", "spilled", "wavefront")}
    ctx.broad_phase(True, 0.0)
    ee, fv = ctx.candidates()
    te, tf = ctx.candidates_toi()
    full = {k: ctx.counter("ccd." + k) for k in ("searched", "spilled", "wavefront")}
    say("  step", t, "bounded", bounded, "per candidate", full)
    assert bounded["spilled"] > 0 and bounded["wavefront"] > 0
    assert full["spilled"] > 0 and full["wavefront"] > 0
    oe = O.candidates_toi(s.V, s.V1, s.E, s.F, ee, True)
    of = O.candidates_toi(s.V, s.V1, s.E, s.F, fv, False)
    assert np.array_equal(te, oe) and np.array_equal(tf, of)
    hits = np.concatenate([te, tf])
    assert t == hits[np.isfinite(hits)].min()
