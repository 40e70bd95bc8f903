"""The serving loop (ipcg_server_*): simulated scenes send energy, gradient,
Hessian and step requests at Poisson arrival times; the server batches the
requests of all admitted scenes into shared launches.  Each request's
result must be the same when its scene is served alone as in the mix:
energies, gradients, projected Hessian blocks, step sizes, acceptance and
the scene's final positions, bit for bit.  Admission under a small HBM
budget serves the scenes in several batches with the same results; a scene
larger than the budget is refused."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DHAT, KAPPA = 1e-3, 1e3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def scene_set(scenes):
    return [scenes.two_sheets(12, seed=1), scenes.sphere_chain(3, 2, seed=2),
            scenes.two_sheets(16, seed=3, twist=0.2), scenes.sphere_chain(2, 3, seed=4),
            scenes.two_sheets(10, seed=5)]


def requests(n_scenes, count, rate, seed):
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    return [(float(t[i]), int(rng.integers(n_scenes)), int(rng.integers(4)))
            for i in range(count)]


def serve(capi, scs, reqs, budget=64 << 30, only=None):
    srv = capi.Server(hbm_budget_bytes=budget, keep_results=True)
    ids = {}
    for k, s in enumerate(scs):
        if only is None or k == only:
            ids[k] = srv.add_scene(s.V, s.E, s.F, DHAT, KAPPA, 0.1 * DHAT, seed=1000 + k)
    tickets = []
    for arrival, k, kind in reqs:
        if k in ids:
            tickets.append((k, kind, srv.submit(ids[k], kind, arrival)))
    srv.run()
    out = []
    for k, kind, t in tickets:
        r = srv.result(t)
        item = {"scene": k, "kind": kind, "value": r.value, "pairs": r.pairs,
                "accepted": r.accepted, "latency": r.finish - r.arrival,
                "iteration": r.iteration}
        if kind == capi.REQ_GRADIENT:
            item["grad"] = srv.gradient(t)
        if kind == capi.REQ_HESSIAN:
            item["hess"] = srv.hessian(t)
        out.append(item)
    pos = {k: srv.positions(ids[k]) for k in ids}
    st = srv.stats()
    srv.close()
    return out, pos, st


def same(a, b):
    assert a["kind"] == b["kind"] and a["scene"] == b["scene"]
    assert a["value"] == b["value"], (a["kind"], a["value"], b["value"])
    assert a["pairs"] == b["pairs"] and a["accepted"] == b["accepted"]
    if "grad" in a:
        assert np.array_equal(a["grad"], b["grad"])
    if "hess" in a:
        H, G = a["hess"], b["hess"]
        assert np.array_equal(H.indptr, G.indptr) and np.array_equal(H.indices, G.indices)
        assert np.array_equal(H.data, G.data)


def test_results_alone_and_in_the_mix(capi, scenes):
    scs = scene_set(scenes)
    reqs = requests(len(scs), 60, 2000.0, seed=7)
    mix, pos_mix, st = serve(capi, scs, reqs)
    say("mix", st, "kinds", np.bincount([r["kind"] for r in mix], minlength=4).tolist(),
        "p50 latency ms", round(1e3 * float(np.median([r["latency"] for r in mix])), 3))
    assert st["requests"] == len(reqs) and st["epochs"] == 1
    # several requests shared iterations
    assert st["iterations"] < len(reqs)
    assert any(r["kind"] == capi.REQ_STEP and r["accepted"] for r in mix)
    assert any(r["pairs"] > 0 for r in mix)
    for k in range(len(scs)):
        alone, pos, _ = serve(capi, scs, reqs, only=k)
        mine = [r for r in mix if r["scene"] == k]
        assert len(alone) == len(mine)
        for a, b in zip(alone, mine):
            same(a, b)
        assert np.array_equal(pos[k], pos_mix[k])


def test_admission_under_a_small_budget(capi, scenes):
    scs = scene_set(scenes)
    reqs = requests(len(scs), 40, 2000.0, seed=11)
    full, pos_full, _ = serve(capi, scs, reqs)
    # room for about two of these scenes at a time
    budget = 2 * max(4096 * len(s.F) + 1024 * len(s.V) for s in scs)
    small, pos_small, st = serve(capi, scs, reqs, budget=budget)
    say("small budget", budget, st)
    assert st["epochs"] > 1 and st["estimated_bytes"] <= budget
    assert st["requests"] == len(reqs)
    by_scene = {}
    for r in small:
        by_scene.setdefault(r["scene"], []).append(r)
    for k in by_scene:
        mine = [r for r in full if r["scene"] == k]
        for a, b in zip(by_scene[k], mine):
            same(a, b)
        assert np.array_equal(pos_small[k], pos_full[k])


def test_scene_beyond_the_budget_is_refused(capi, scenes):
    s = scenes.two_sheets(12, seed=1)
    srv = capi.Server(hbm_budget_bytes=1024)
    sid = srv.add_scene(s.V, s.E, s.F, DHAT, KAPPA, 0.1 * DHAT, seed=1)
    srv.submit(sid, capi.REQ_ENERGY, 0.0)
    with pytest.raises(capi.IPCGError) as e:
        srv.run()
    assert e.value.code == capi.ERR_OOM
    srv.close()
