"""Contact-step benchmark on one MI355X.

One step = the five calls of IPC Toolkit's contact hot path on one scene,
every input already resident in HBM:
  NormalCollisions::build (LBVH broad phase + narrow phase),
  BarrierPotential energy, gradient, PSD-projected Hessian (CSR),
  compute_collision_free_stepsize (swept LBVH + Tight-Inclusion CCD).
Metric: collision pairs per second = |C| / step time, |C| = vv+ev+ee+fv
collisions after deduplication; p50/p99 step latency beside it.

Workload: 8 stacked, interleaved 512x512 sheets (2.1M vertices, 4.2M
triangles), sheet spacing 0.6 dhat, z-jitter U(-0.2,0.2) dhat, dhat 1e-3,
kappa 1e4; step N(0,(0.5 dhat)^2) in-plane and -0.8 dhat in z.

cpu_baseline: the CPU restatement (oracle/, OpenMP) running the same five
calls on a bounded tile of the same workload (same grid spacing and sheet
stack, fewer vertices per sheet), timed on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ipc-toolkit_amd"))

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix), TFLOP/s

# Algorithmic work models, per unit of each stage (bytes a kernel must move
# and FP64 flops it must execute, counted from the algorithm).
# Hessian record of one collision: ids, weight (+ eps, dtype for EE), the
# stencil's points, and the packed projected Hessian written out.
REC_BYTES = {"ee": 8 + 8 + 8 + 4 + 4 * 24 + 45 * 8,
             "fv": 8 + 8 + 4 * 24 + 45 * 8,
             "ev": 8 + 8 + 3 * 24 + 21 * 8}
# Jacobi PSD projection of an n x n matrix: SWEEPS[n] sweeps of n(n-1)/2
# rotations (12n + 7 flops each), a convergence sum per sweep (n(n+1)
# flops) and the reconstruction V diag(lambda) V^T (3n flops per packed
# entry); SWEEPS is the mean sweep count measured on contact stencils.
SWEEPS = {3: 4, 6: 6, 9: 7}
PSD_FLOPS = {n: SWEEPS[n] * (n * (n - 1) // 2 * (12 * n + 7) + n * (n + 1))
             + 3 * n * n * (n + 1) // 2 for n in (3, 6, 9)}
# Rank-5 path for a 9 x 9 matrix: 5 Lanczos steps (matrix-vector product,
# two reorthogonalization passes), T = Q^T A Q (6 products + 21 dots), the
# residual A - Q T Q^T (45 entries x 36 terms), a 6 x 6 Jacobi (5 sweeps),
# the Ritz vectors and the reconstruction.
KRYLOV9_FLOPS = (5 * (162 + 2 * 3 * 36 + 20) + (6 * 162 + 21 * 18) + 45 * 36 * 3
                 + 5 * (15 * (12 * 6 + 7) + 42) + 9 * 6 * 6 * 2 + 45 * 6 * 3)
# Tight-Inclusion: one box = the 8 corners of F along 3 axes (about 70 flops
# per axis); a candidate that survives the root pass visits 5.0 boxes on
# average on the sheet-stack workload (counted in-kernel).
TI_BOX_FLOPS = 210
TI_BOXES_PER_SEARCH = 5.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=512, help="vertices per sheet side")
    p.add_argument("--sheets", type=int, default=8)
    p.add_argument("--seed", type=int, default=0)
    p.add_argument("--cpu-tile", type=int, default=64,
                   help="sheet side of the CPU-baseline tile (0: skip)")
    p.add_argument("--cpu-steps", type=int, default=10,
                   help="timed CPU-baseline steps on all host cores (after 2 warm-ups)")
    p.add_argument("--cpu-steps-1core", type=int, default=1,
                   help="timed CPU-baseline steps on one core")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only copy a MIB MiB device buffer five times (torch), the known-size "
                        "reference for scaling the HBM byte counters of a profiler")
    p.add_argument("--forms", type=int, default=1,
                   help="also measure the batched form (--batch-scenes scenes of uneven size in "
                        "one context, beside the same scenes one context each) and the served "
                        "form (Poisson requests); 0: skip")
    p.add_argument("--batch-scenes", type=int, default=256)
    p.add_argument("--batch-steps", type=int, default=2)
    p.add_argument("--serve-scenes", type=int, default=64)
    p.add_argument("--serve-rate", type=float, default=2000.0, help="requests per second")
    p.add_argument("--serve-seconds", type=float, default=4.0,
                   help="span of the request arrivals")
    p.add_argument("--rows", action="store_true",
                   help="time the collision-set variants, helpers and friction at the same "
                        "workload instead (one JSON line per entry point)")
    p.add_argument("--verbose", action="store_true")
    return p.parse_args()


def log(args, *a):
    if args.verbose:
        print(*a, file=sys.stderr, flush=True)


def progress(*a):
    """a line on stderr for the long phases (the JSON result is on stdout)"""
    print("[bench]", *a, file=sys.stderr, flush=True)


def gpu_step(ctx, sc):
    ctx.build_collisions(sc.dhat, sc.dmin)
    ctx.energy(sc.dhat, sc.kappa)
    ctx.gradient(sc.dhat, sc.kappa, download=False)
    ctx.hessian(sc.dhat, sc.kappa, "clamp", download=False)
    step = ctx.stepsize(min_distance=0.0)
    ctx.synchronize()
    return step


def algorithmic_bytes(sc, counts, n_static, n_swept, nnz):
    """Bytes a step must move at least (inputs read once, outputs written once)."""
    nV, nE, nF = sc.sizes
    C = int(sum(counts))
    return (24 * nV + 24 * nV + 8 * nE + 12 * nF + 16 * n_static + 16 * n_swept + 16 * C
            + 24 * nV + 12 * nnz + 4 * (3 * nV + 1) + 8)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_step(O, tile):
    """The five calls on the CPU restatement; returns (pairs, results)."""
    ee, fv = O.candidates(tile.V, tile.E, tile.F, 0.5 * tile.dhat)
    coll = O.build_collisions(tile.V, tile.E, tile.F, ee, fv, tile.dhat)
    e = O.barrier_energy(coll, tile.V, tile.E, tile.F, tile.dhat, tile.kappa)
    g = O.barrier_gradient(coll, tile.V, tile.E, tile.F, tile.dhat, tile.kappa)
    H = O.barrier_hessian(coll, tile.V, tile.E, tile.F, tile.dhat, tile.kappa, "clamp")
    # the reference's own schedule: running-minimum tmax (candidates.cpp:269-287)
    t = O.compute_collision_free_stepsize(tile.V, tile.V1, tile.E, tile.F, running_tmax=True)
    return len(coll), (coll, e, g, H, t)


def tile_parity(O, capi, tile, ref):
    """The device's results on the CPU tile against the oracle's (the
    checker): set equality and the relative error of every quantity."""
    coll, e_o, g_o, H_o, _ = ref
    ctx = capi.Context(0)
    try:
        ctx.upload_mesh(tile.V, tile.E, tile.F)
        ctx.upload_positions(0, tile.V)
        ctx.upload_positions(1, tile.V1)
        ctx.build_collisions(tile.dhat)
        cg = ctx.collisions()
        same = all(np.array_equal(cg[k], getattr(coll, k)) and
                   np.array_equal(cg[k + "_w"], getattr(coll, k + "_w"))
                   for k in ("vv", "ev", "ee", "fv"))
        same = same and np.array_equal(cg["ee_eps_x"], coll.ee_eps_x) and \
            np.array_equal(cg["ee_dtype"], coll.ee_dtype)
        e_g = ctx.energy(tile.dhat, tile.kappa)
        g_g = ctx.gradient(tile.dhat, tile.kappa)
        H_g = ctx.hessian(tile.dhat, tile.kappa, "clamp")
        t_g = ctx.stepsize()
    finally:
        ctx.close()
    D = H_g - H_o
    t_all = O.compute_collision_free_stepsize(tile.V, tile.V1, tile.E, tile.F)
    return {"collision_set_identical": bool(same),
            "energy_rel_err": abs(e_g - e_o) / abs(e_o),
            "gradient_rel_err": float(np.linalg.norm(g_g - g_o) / np.linalg.norm(g_o)),
            "hessian_rel_err": float(np.sqrt(D.multiply(D).sum()) /
                                     np.sqrt(H_o.multiply(H_o).sum())),
            "step_equal_to_oracle": bool(t_g == t_all)}


def cpu_baseline(args, scenes, capi):
    """The oracle on a bounded tile of the workload: all host cores (median
    of --cpu-steps after two warm-ups) and one core (fewer steps)."""
    import oracle as O
    tile = scenes.stacked_sheets(args.cpu_tile, args.sheets, seed=args.seed,
                                 h=1.0 / (args.n - 1))
    threads = O.get_max_threads()

    def timed(steps, warm):
        ts, pairs, res = [], 0, None
        for i in range(warm + steps):
            t0 = time.perf_counter()
            pairs, res = cpu_step(O, tile)
            if i >= warm:
                ts.append(time.perf_counter() - t0)
        return pairs, res, ts

    pairs, res, ts = timed(max(1, args.cpu_steps), 2)
    t, t99 = float(np.median(ts)), float(np.percentile(ts, 99))
    O.set_num_threads(1)
    try:
        _, _, ts1 = timed(max(1, args.cpu_steps_1core), 0)
    finally:
        O.set_num_threads(threads)
    t1 = float(np.median(ts1))
    out = {"value": pairs / t, "unit": "pairs/s", "cores": threads, "kind": "port",
           "cpu": cpu_model(), "nproc": os.cpu_count(),
           "p50_ms": round(1e3 * t, 1), "p99_ms": round(1e3 * t99, 1),
           "value_1core": pairs / t1, "ms_1core": round(1e3 * t1, 1),
           "sample": f"{args.sheets} stacked {args.cpu_tile}x{args.cpu_tile} tile of the "
                     f"workload (same spacing), {pairs} collision pairs; {len(ts)} timed steps "
                     f"on {threads} threads after 2 warm-ups, {len(ts1)} on one thread; "
                     f"the oracle (C, OpenMP, -O3 -ffp-contract=off), not a reference build"}
    try:
        out["device_vs_oracle"] = tile_parity(O, capi, tile, res)
    except Exception as e:  # the check never blocks the line
        out["device_vs_oracle"] = {"error": str(e)}
    return out


def batched_step(ctx, dhat, kappa):
    ctx.build_collisions(dhat)
    ctx.energy_batched(dhat, kappa)
    ctx.gradient(dhat, kappa, download=False)
    ctx.hessian(dhat, kappa, "clamp", download=False)
    steps = ctx.stepsize_batched()
    ctx.synchronize()
    return steps


def batched_form(args, capi, scenes):
    """Many independent scenes of uneven size (sheet pairs of side
    log-uniform in [32, 384] and rows of 2-16 icospheres of subdivision 3-5,
    2.6k to 300k triangles each): the five calls once per step over all of
    them in one batched context, against the same scenes each in its own
    context stepped one after the other (the plain single-scene path)."""
    dhat, kappa = 1e-3, 1e3
    sc = scenes.batch_scenes(args.batch_scenes, seed=args.seed)
    B = scenes.Batch(sc)
    progress("batched form:", len(sc), "scenes generated", B.sizes)
    ctx = capi.Context(0)
    ctx.set_autotune(True)  # traversal workgroup sizes timed on the first step
    ctx.upload_batch(B)
    ctx.upload_positions(0, B.V)
    ctx.upload_positions(1, B.V1)
    progress("batched form: uploaded")
    batched_step(ctx, dhat, kappa)  # warm-up: buffers reach their size, shapes are tuned
    progress("batched form: warm-up step done")
    shapes = {st: ctx.launch_shape(st)[0] for st in
              ("bp_traverse_ee.static", "bp_traverse_fv.static", "bp_traverse_ee.swept",
               "bp_traverse_fv.swept")}
    lat = []
    for _ in range(max(1, args.batch_steps)):
        t0 = time.perf_counter()
        batched_step(ctx, dhat, kappa)
        lat.append(time.perf_counter() - t0)
    pairs = int(ctx.collision_counts_batched().sum())
    dev_bytes = ctx.device_bytes()
    ctx.close()
    t_b = float(np.median(lat))
    # plain path: one resident context per scene, the step called per scene
    progress("batched form: %d timed steps, %.1f ms median" % (len(lat), 1e3 * t_b))
    ctxs = []
    for k, s in enumerate(sc):
        c = capi.Context(0)
        c.upload_mesh(s.V, s.E, s.F)
        c.upload_positions(0, s.V)
        c.upload_positions(1, s.V1)
        ctxs.append(c)
        if k % 64 == 63:
            progress("plain path: %d contexts uploaded" % (k + 1))

    def plain_pass():
        for s, c in zip(sc, ctxs):
            c.build_collisions(dhat)
            c.energy(dhat, kappa)
            c.gradient(dhat, kappa, download=False)
            c.hessian(dhat, kappa, "clamp", download=False)
            c.stepsize()
            c.synchronize()

    plain_pass()
    progress("plain path: warm-up pass done")
    lat_p = []
    for _ in range(max(1, args.batch_steps)):
        t0 = time.perf_counter()
        plain_pass()
        lat_p.append(time.perf_counter() - t0)
        progress("plain path: pass %.1f ms" % (1e3 * lat_p[-1]))
    pairs_p = int(sum(int(c.collision_counts().sum()) for c in ctxs))
    for c in ctxs:
        c.close()
    t_p = float(np.median(lat_p))
    nV, nE, nF = B.sizes
    return {"workload": f"batch_{args.batch_scenes}_scenes", "scenes": len(sc),
            "vertices": nV, "faces": nF, "collisions": pairs, "collisions_plain": pairs_p,
            "ms_per_step": round(1e3 * t_b, 3), "p50_ms": round(1e3 * t_b, 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "value": round(pairs / t_b, 1), "unit": "pairs/s",
            "plain_ms_per_step": round(1e3 * t_p, 3), "plain_value": round(pairs_p / t_p, 1),
            "speedup_vs_plain": round(t_p / t_b, 3), "device_bytes": dev_bytes,
            "steps": len(lat), "tuned_workgroup_sizes": shapes}


def served_form(args, capi, scenes):
    """Scenes drawn as for the batched form send energy, gradient, Hessian
    and step requests (uniform kind, uniform scene) with Poisson arrivals;
    a step perturbs the scene's positions by N(0, (0.1 dhat)^2) and applies
    them when the CCD step is 1.  Latency = finish - arrival."""
    dhat, kappa = 1e-3, 1e3
    sc = scenes.batch_scenes(args.serve_scenes, seed=args.seed + 1)
    progress("served form:", len(sc), "scenes generated")
    srv = capi.Server(hbm_budget_bytes=64 << 30)
    ids = [srv.add_scene(s.V, s.E, s.F, dhat, kappa, 0.1 * dhat, seed=k)
           for k, s in enumerate(sc)]
    rng = np.random.default_rng(args.seed)
    n = max(1, int(args.serve_rate * args.serve_seconds))
    arrivals = np.cumsum(rng.exponential(1.0 / args.serve_rate, n))
    who = rng.integers(len(sc), size=n)
    kinds = rng.integers(4, size=n)
    tickets = [srv.submit(ids[int(w)], int(k), float(a)) for a, w, k in zip(arrivals, who, kinds)]
    progress("served form: %d requests over %.1f s submitted" % (n, float(arrivals[-1])))
    srv.run()
    progress("served form: done")
    res = [srv.result(t) for t in tickets]
    st = srv.stats()
    srv.close()
    lat = np.array([r.finish - r.arrival for r in res])
    pairs = sum(r.pairs for r in res)
    wall = st["wall_seconds"]
    per_kind = {name: round(1e3 * float(np.percentile(lat[kinds == k], 99)), 3)
                for k, name in enumerate(("energy", "gradient", "hessian", "step"))
                if np.any(kinds == k)}
    return {"workload": f"serve_{args.serve_scenes}_scenes_poisson_{args.serve_rate:g}",
            "requests": n, "rate_offered": args.serve_rate,
            "rate_served": round(n / wall, 1), "wall_s": round(wall, 3),
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "p99_ms_by_kind": per_kind,
            "value": round(pairs / wall, 1), "unit": "pairs/s",
            "collision_pairs_served": int(pairs),
            "iterations": st["iterations"], "requests_per_iteration": round(n / st["iterations"], 2),
            "busy_fraction": round(st["busy_seconds"] / wall, 3),
            "steps_accepted": int(sum(r.accepted for r in res)),
            "peak_device_bytes": st["peak_device_bytes"],
            "estimated_bytes": st["estimated_bytes"], "hbm_budget_bytes": 64 << 30}


def copy_calibration(mib):
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_bytes": n, "launches": 5, "ms_per_copy": round(1e3 * dt, 4),
                      "GBps": round(2 * n / dt / 1e9, 1)}))


def rows(args, ctx, sc, capi):
    """The entry points beside the step on the same workload: per entry
    point, the median wall time of `args.steps` synchronous calls after one
    untimed call, and the HIP-event time of its stages per call."""
    import numpy as np
    dhat, kappa, eps_v = sc.dhat, sc.kappa, 1e-3

    def timed(name, fn, **work):
        fn()
        ctx.timing(True)
        ts = []
        for _ in range(args.steps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        st = ctx.stage_times()
        ctx.timing(False)
        out = {"row": name, "ms_per_call": round(1e3 * float(np.median(ts)), 3),
               "calls": args.steps, "workload": sc.name,
               "stages_ms_per_launch": {k: round(v[0] / max(v[1], 1), 4) for k, v in st.items()}}
        out.update(work)
        print(json.dumps(out), flush=True)

    ctx.build_collisions(dhat)
    base = [int(x) for x in ctx.collision_counts()]
    timed("has_intersections", ctx.has_intersections)
    for label, aw, kind in (("collisions_ipc", False, capi.COLLISIONS_IPC),
                            ("collisions_area_weighted", True, capi.COLLISIONS_IPC),
                            ("collisions_improved_max_approx", True,
                             capi.COLLISIONS_IMPROVED_MAX_APPROX),
                            ("collisions_ogc", False, capi.COLLISIONS_OGC)):
        ctx.set_area_weighting(aw)
        ctx.set_collision_set_type(kind)
        timed(label, lambda: ctx.build_collisions_from_candidates(dhat),
              collisions=[int(x) for x in ctx.collision_counts()])
    ctx.set_area_weighting(False)
    ctx.set_collision_set_type(capi.COLLISIONS_IPC)
    ctx.build_collisions_from_candidates(dhat)
    p = np.random.default_rng(0).normal(size=3 * len(sc.V))
    timed("gauss_newton_diagonal", lambda: ctx.gauss_newton_diagonal(dhat, kappa),
          collisions=base, note="includes the 3 nV download")
    timed("gauss_newton_quadratic_form", lambda: ctx.gauss_newton_quadratic_form(dhat, kappa, p),
          collisions=base, note="includes the 3 nV upload")
    r = 0.5 * dhat
    timed("per_vertex_safe_distances", lambda: ctx.per_vertex_safe_distances(r, 0.0),
          note="includes the nV download")
    timed("noncandidate_conservative_stepsize",
          lambda: ctx.noncandidate_conservative_stepsize(dhat))
    timed("friction_build", lambda: ctx.friction_build(dhat, kappa, 0.5, 0.3), collisions=base)
    ctx.upload_velocities(np.random.default_rng(1).normal(0, 2 * eps_v, sc.V.shape))
    timed("friction_energy", lambda: ctx.friction_energy(eps_v), collisions=base)
    timed("friction_gradient", lambda: ctx.friction_gradient(eps_v), collisions=base,
          note="includes the 3 nV download")
    timed("friction_hessian", lambda: ctx.friction_hessian(eps_v, "clamp", download=False),
          collisions=base)


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(several GPUs would each take whole scenes or a Morton range of one scene)",
              file=sys.stderr)
        sys.exit(2)
    from ipctk_amd import capi, scenes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    sc = scenes.stacked_sheets(args.n, args.sheets, seed=args.seed)
    nV, nE, nF = sc.sizes
    log(args, "scene", sc.name, sc.sizes)
    ctx = capi.Context(0)
    ctx.upload_mesh(sc.V, sc.E, sc.F)
    ctx.upload_positions(0, sc.V)
    ctx.upload_positions(1, sc.V1)
    ctx.synchronize()

    if args.rows:
        rows(args, ctx, sc, capi)
        return

    for i in range(args.warmup):
        gpu_step(ctx, sc)
        log(args, "warmup", i)

    # timed steps (per-step latency) with stage timing on the context stream
    ctx.timing(True)
    lat = []
    step = None
    t_all = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        step = gpu_step(ctx, sc)
        lat.append(time.perf_counter() - t0)
    total = time.perf_counter() - t_all
    stages = ctx.stage_times()
    ctx.timing(False)

    ccd_work = {k: ctx.counter(k) for k in ("ccd.candidates", "ccd.searched", "ccd.spilled",
                                              "ccd.spilled_deep", "ccd.wavefront")}
    ee_other = ctx.counter("hessian.ee_other")
    counts = ctx.collision_counts()
    ctx.broad_phase(False, 0.5 * sc.dhat)
    ne_s, nf_s = [len(x) for x in ctx.candidates()]
    nnz = ctx.hessian(sc.dhat, sc.kappa, "clamp", download=False)
    ctx.broad_phase(True, 0.0)
    ne_d, nf_d = [len(x) for x in ctx.candidates()]
    pairs = int(counts.sum())
    ms_step = 1e3 * total / args.steps
    value = pairs * args.steps / total

    # Roofline of the dominant kernel.  Each stage below is one kernel
    # launch per call, timed with HIP events on the context stream.  Its
    # algorithmic bytes and FP64 flops per launch follow the models above;
    # the binding roof is the larger of bytes / HBM peak and flops / FP64 peak.
    per_launch = {k: v[0] / v[1] for k, v in stages.items()}
    per_step = {k: v[0] / args.steps for k, v in stages.items()}
    n_vv, n_ev, n_ee, n_fv = [int(x) for x in counts]
    searched = int(ccd_work["ccd.searched"])
    model = {  # stage: (bytes, flops) per launch
        "bp_traverse_ee.static": (24 * nE + 8 * ne_s, 0),
        "bp_traverse_fv.static": (24 * nV + 8 * nf_s, 0),
        "bp_traverse_ee.swept": (24 * nE + 8 * ne_d, 0),
        "bp_traverse_fv.swept": (24 * nV + 8 * nf_d, 0),
        "np_classify": (8 * (ne_s + nf_s) + 16 * pairs, 0),
        "energy": (16 * pairs + 8, 0),
        "hessian_records.ee": (REC_BYTES["ee"] * (n_ee - ee_other),
                               KRYLOV9_FLOPS * (n_ee - ee_other)),
        "hessian_records.ee_other": (REC_BYTES["ee"] * ee_other, PSD_FLOPS[9] * ee_other),
        "hessian_records.fv": (REC_BYTES["fv"] * n_fv, KRYLOV9_FLOPS * n_fv),
        "hessian_records.ev": (REC_BYTES["ev"] * n_ev, PSD_FLOPS[6] * n_ev),
        "hessian_assemble": (12 * nnz + 8 * (3 * nV + 1), 0),
        "ccd_root": (200 * (ne_d + nf_d), 0),
        "ccd_lds": (204 * searched, TI_BOX_FLOPS * TI_BOXES_PER_SEARCH * searched),
    }

    def roof(k):
        byt, flo = model[k]
        t = per_launch[k] * 1e-3
        gbs, tfs = byt / t / 1e9, flo / t / 1e12
        if flo and tfs / FP64_PEAK_TFLOPS > gbs / HBM_PEAK_GBS:
            return {"bound": "fp64_valu", "achieved": round(tfs, 3), "peak": FP64_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(tfs / FP64_PEAK_TFLOPS, 5),
                    "flops_per_launch": int(flo)}
        return {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 5), "bytes_per_launch": int(byt)}

    known = [k for k in per_step if k in model]
    dom = max(known, key=per_step.get)
    roof_dom = roof(dom)
    roof_dom.update({"traffic": None, "kernel": dom,
                     "ms_per_launch": round(per_launch[dom], 4)})
    # HBM traffic of the dominant kernel: the rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes of this workload, scaled by a copy kernel of known size (the
    # summary and the commands are under profiles/)
    tpath = os.path.join(ROOT, "profiles", f"traffic__{sc.name}.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            t = json.load(f)
        if t.get("kernel") == dom:
            roof_dom["traffic"] = int(t["traffic_bytes_per_launch"])
            roof_dom["traffic_source"] = os.path.relpath(tpath, ROOT)
    if roof_dom["bound"] == "fp64_valu":
        roof_dom["note"] = ("FP64 on the vector ALUs (no step is a dense contraction, so "
                            "nothing runs on MFMA); the MI355X FP64 vector peak is 78.6 TFLOP/s")
    roof_dom["others"] = {k: {"bound": r["bound"], "frac": r["frac"]}
                          for k in sorted(known, key=per_step.get, reverse=True)[1:6]
                          for r in [roof(k)]}

    # host transfers (outside `value`): both position slots up, gradient down
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.upload_positions(0, sc.V)
        ctx.upload_positions(1, sc.V1)
    ctx.synchronize()
    up_ms = (time.perf_counter() - t0) / 3 * 1e3
    ctx.build_collisions(sc.dhat, sc.dmin)
    ctx.gradient(sc.dhat, sc.kappa, download=False)  # incidences of the new collision set
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        ctx.gradient(sc.dhat, sc.kappa, download=True)
    grad_ms = (time.perf_counter() - t0) / 3 * 1e3
    host_transfer = {"positions_upload_ms": round(up_ms, 3),
                     "gradient_with_download_ms": round(grad_ms, 3),
                     "gradient_device_ms": round(per_step.get("gradient", 0.0), 3)}

    cpu = None
    progress("step: %.3f ms per step" % ms_step)
    if args.cpu_tile > 0:
        try:
            progress("cpu baseline")
            cpu = cpu_baseline(args, scenes, capi)
        except Exception as e:  # the baseline never blocks the GPU line
            cpu = {"value": None, "error": str(e)}

    forms = None
    if args.forms:
        forms = {}
        for name, fn in (("batched", batched_form), ("served", served_form)):
            try:
                forms[name] = fn(args, capi, scenes)
            except Exception as e:  # a form never blocks the line
                forms[name] = {"error": str(e)}
            log(args, name, forms[name])

    out = {
        "metric": "collision pairs/sec (build+E/grad/Hess+CCD) at 1e-10 rel err; p50/p99 latency",
        "value": round(value, 1),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": sc.name, "vertices": nV, "edges": nE, "faces": nF,
                   "dhat": sc.dhat, "kappa": sc.kappa, "seed": args.seed,
                   "collisions": pairs, "collisions_by_type": [int(x) for x in counts],
                   "candidates_static": ne_s + nf_s, "candidates_swept": ne_d + nf_d,
                   "hessian_nnz": int(nnz), "step_size": step, "ccd_work": ccd_work,
                   "stage_ms_per_step": {k: round(v, 4) for k, v in sorted(per_step.items())},
                   "bytes_algorithmic": algorithmic_bytes(sc, counts, ne_s + nf_s,
                                                          ne_d + nf_d, nnz),
                   "host_transfer": host_transfer},
        "roofline": roof_dom,
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
