"""Device edge-face candidates and has_intersections against the oracle.

Candidate sets must be identical (ids, orientation (edge, face)); the
has_intersections answer must equal the oracle's, which is pinned against
exact rational arithmetic on the CPU.  Scenes cover the reference's
"Has intersections" case (two cubes rotated about a common centre), apart,
overlapping and nested solids, flat sheets (every edge-face pair coplanar,
so every test goes through the exact orient3d), a vertex pushed through a
sheet, and the patch filter.  At the benchmark size the oracle is too slow,
so the property checked there is size independent: the stacked sheets are
intersection-free by construction, and pushing one vertex through the sheet
below makes them intersect."""
import sys

import numpy as np
import pytest

from conftest import pair_set

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def two_solids(scenes, kind, rng):
    Vc, Fc = scenes.cube()
    if kind == "cubes_coincident":
        return scenes.merge([(Vc, Fc), (Vc.copy(), Fc)])
    if kind == "cubes_rotated":
        return scenes.merge([(Vc @ scenes.random_rotation(rng).T, Fc),
                             (Vc @ scenes.random_rotation(rng).T, Fc)])
    if kind == "cubes_apart":
        return scenes.merge([(Vc, Fc), (Vc + [1.5, 0.2, 0.1], Fc)])
    if kind == "cubes_overlap":
        return scenes.merge([(Vc, Fc), (Vc + [0.6, 0.2, 0.1], Fc)])
    if kind == "cubes_nested":
        return scenes.merge([(Vc, Fc), (0.3 * Vc, Fc)])
    gap = {"spheres_overlap": -0.3, "spheres_touching": -0.01, "spheres_apart": 0.05}[kind]
    Vs, Fs = scenes.icosphere(2, 1.0)
    return scenes.merge([(Vs @ scenes.random_rotation(rng).T, Fs),
                         (Vs @ scenes.random_rotation(rng).T + [2.0 + gap, 0.0, 0.0], Fs)])


SOLIDS = ["cubes_coincident", "cubes_rotated", "cubes_apart", "cubes_overlap", "cubes_nested",
          "spheres_overlap", "spheres_touching", "spheres_apart"]


@pytest.mark.parametrize("kind", SOLIDS)
def test_has_intersections_solids(ctx, O, scenes, kind):
    rng = np.random.default_rng(SOLIDS.index(kind))
    V, E, F = two_solids(scenes, kind, rng)
    say("case", kind, V.shape)
    ctx.upload_mesh(V, E, F)
    ctx.upload_positions(0, V)
    want = O.has_intersections(V, E, F)
    assert ctx.has_intersections() == want
    if kind in ("cubes_coincident", "cubes_rotated", "cubes_overlap", "spheres_overlap"):
        assert want
    if kind in ("cubes_apart", "cubes_nested", "spheres_apart"):
        assert not want
    # the candidate list it used
    r = 1e-6 * O.world_bbox_diagonal(V)
    assert ctx.counter("intersection.candidates") == len(O.edge_face_candidates(V, E, F, r))
    # cube faces are coplanar with their neighbours' edges: exact path used
    if kind.startswith("cubes"):
        assert ctx.counter("intersection.deferred") > 0


@pytest.mark.parametrize("name,r", [("sheets:16:0", 1e-3), ("spheres:4:3:0", 2e-3),
                                    ("stack:24:1", 5e-4), ("twisted:12:1", 0.0)])
def test_edge_face_candidates(ctx, O, scenes, name, r):
    from test_gpu_parity import make
    s = make(scenes, name)
    say("case", name, r)
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    got = ctx.edge_face_candidates(r)
    want = O.edge_face_candidates(s.V, s.E, s.F, r)
    assert len(got) == len(want) and pair_set(got) == pair_set(want)
    # the earlier broad phase lists survive the edge-face build
    ctx.broad_phase(False, 0.5 * s.dhat)
    ee0, fv0 = ctx.candidates()
    ctx.edge_face_candidates(r)
    ee1, fv1 = ctx.candidates()
    assert np.array_equal(ee0, ee1) and np.array_equal(fv0, fv1)


def test_patch_filter(capi, O, scenes):
    rng = np.random.default_rng(7)
    V, E, F = two_solids(scenes, "cubes_overlap", rng)
    ctx = capi.Context(0)
    try:
        for group in (np.zeros(len(V), np.int32), np.repeat([0, 1], 8).astype(np.int32)):
            say("case patch", group.max())
            ctx.upload_mesh(V, E, F, vertex_patch=group)
            ctx.upload_positions(0, V)
            assert ctx.has_intersections() == O.has_intersections(V, E, F, vertex_group=group)
            got = ctx.edge_face_candidates(1e-3)
            want = O.edge_face_candidates(V, E, F, 1e-3, vertex_group=group)
            assert pair_set(got) == pair_set(want)
        # one patch: no pair may collide, so no intersection is reported
        ctx.upload_mesh(V, E, F, vertex_patch=np.zeros(len(V), np.int32))
        ctx.upload_positions(0, V)
        assert not ctx.has_intersections()
    finally:
        ctx.close()


def test_flat_sheets_and_pushed_vertex(ctx, O, scenes):
    s = scenes.stacked_sheets(24, 4, seed=3, h=1.0 / 511)
    V = s.V.copy()
    n = len(V) // 4
    for k in range(4):  # each sheet flat at its nominal height
        V[k * n:(k + 1) * n, 2] = k * 0.6 * s.dhat
    say("case flat sheets", V.shape)
    ctx.upload_mesh(V, s.E, s.F)
    ctx.upload_positions(0, V)
    assert ctx.has_intersections() == O.has_intersections(V, s.E, s.F) == False  # noqa: E712
    assert ctx.counter("intersection.deferred") > 0
    k = n + 12 * 24 + 11  # an interior vertex of sheet 1, pushed below sheet 0
    V[k, 2] = -0.5 * s.dhat
    ctx.upload_positions(0, V)
    assert O.has_intersections(V, s.E, s.F)
    assert ctx.has_intersections()


def test_benchmark_size_property(ctx, scenes):
    s = scenes.stacked_sheets(512, 8, seed=0)
    say("case stacked_sheets_8x512", s.sizes)
    ctx.upload_mesh(s.V, s.E, s.F)
    ctx.upload_positions(0, s.V)
    assert not ctx.has_intersections()
    n0 = ctx.counter("intersection.candidates")
    assert n0 > 0
    V = s.V.copy()
    n = len(V) // 8
    # an interior vertex of sheet 3 pushed below the lowest point sheet 2 can
    # reach (1.2 - 0.2 dhat); its neighbours stay above sheet 2's highest
    k = 3 * n + 200 * 512 + 100
    V[k, 2] = (2 * 0.6 - 0.5) * s.dhat
    ctx.upload_positions(0, V)
    assert ctx.has_intersections()
