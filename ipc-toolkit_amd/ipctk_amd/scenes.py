"""Seeded synthetic contact scenes (the workloads the benchmarks quote).

Every generator is deterministic given its seed and returns float64
positions (nV, 3), int32 edges (nE, 2) and faces (nF, 3), plus the scene
parameters.  Edges are the unique undirected face edges, each stored as
(min, max) and sorted lexicographically (the set ``igl::edges`` returns).

Sheets are n x n vertex grids over the unit square, vertex (i, j) at
(j h, i h) with h = 1/(n-1); each grid cell is split along its
(i,j)-(i+1,j+1) diagonal into two counter-clockwise triangles.
"""
import functools
import math

import numpy as np


def grid_sheet(n, z=0.0, offset=(0.0, 0.0), size=1.0):
    h = size / (n - 1)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n))
    V = np.zeros((n * n, 3))
    V[:, 0] = jj.reshape(-1) * h + offset[0]
    V[:, 1] = ii.reshape(-1) * h + offset[1]
    V[:, 2] = z
    i, j = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    v00 = (i * n + j).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + n
    v11 = v10 + 1
    F = np.empty((2 * (n - 1) ** 2, 3), dtype=np.int64)
    F[0::2] = np.stack([v00, v01, v11], 1)
    F[1::2] = np.stack([v00, v11, v10], 1)
    return V, F.astype(np.int32)


def edges_from_faces(F):
    F = np.asarray(F, dtype=np.int64)
    a = np.concatenate([F[:, 0], F[:, 1], F[:, 2]])
    b = np.concatenate([F[:, 1], F[:, 2], F[:, 0]])
    key = np.unique((np.minimum(a, b) << 32) | np.maximum(a, b))
    return np.stack([key >> 32, key & 0xFFFFFFFF], 1).astype(np.int32)


def merge(meshes):
    Vs, Fs, off = [], [], 0
    for V, F in meshes:
        Vs.append(V)
        Fs.append(F + off)
        off += len(V)
    V = np.concatenate(Vs)
    F = np.concatenate(Fs).astype(np.int32)
    return V, edges_from_faces(F), F


class Scene:
    """Positions at t0 (V), the step target V1, topology and parameters."""

    def __init__(self, name, V, E, F, V1, dhat, kappa, dmin=0.0, rest=None, **meta):
        self.name = name
        self.V = np.ascontiguousarray(V, dtype=np.float64)
        self.E = np.ascontiguousarray(E, dtype=np.int32)
        self.F = np.ascontiguousarray(F, dtype=np.int32)
        self.V1 = np.ascontiguousarray(V1, dtype=np.float64)
        self.rest = self.V if rest is None else np.ascontiguousarray(rest, np.float64)
        self.dhat = float(dhat)
        self.kappa = float(kappa)
        self.dmin = float(dmin)
        self.meta = meta

    @property
    def sizes(self):
        return len(self.V), len(self.E), len(self.F)


def two_sheets(n=64, seed=0, dhat=1e-3, kappa=1e3, twist=0.0):
    """Two stacked n x n sheets: bottom at z = 0, top at
    z = 0.5 dhat + U(-0.25, 0.25) dhat per vertex.  Step: the top sheet
    moves by (0, 0, -1e-3) + U(-1e-4, 1e-4)^3 per vertex.  `twist` rotates
    the top sheet about the vertical axis through the square's centre (radians;
    0 keeps its edges parallel to the bottom sheet's)."""
    rng = np.random.default_rng(seed)
    Vb, Fb = grid_sheet(n, 0.0)
    Vt, Ft = grid_sheet(n, 0.0)
    if twist:
        c, s_ = math.cos(twist), math.sin(twist)
        x, y = Vt[:, 0] - 0.5, Vt[:, 1] - 0.5
        Vt[:, 0], Vt[:, 1] = 0.5 + c * x - s_ * y, 0.5 + s_ * x + c * y
    Vt[:, 2] = 0.5 * dhat + rng.uniform(-0.25, 0.25, len(Vt)) * dhat
    V, E, F = merge([(Vb, Fb), (Vt, Ft)])
    V1 = V.copy()
    nb = len(Vb)
    disp = np.zeros((len(Vt), 3))
    disp[:, 2] = -1e-3
    disp += rng.uniform(-1e-4, 1e-4, (len(Vt), 3))
    V1[nb:] += disp
    name = f"two_sheets_{n}" + (f"_twist{twist:g}" if twist else "")
    return Scene(name, V, E, F, V1, dhat, kappa, seed=seed)


def stacked_sheets(n=512, sheets=8, seed=0, dhat=1e-3, kappa=1e4, spacing=0.6,
                   jitter=0.2, inplane_sigma=0.5, dz=-0.8, h=None):
    """`sheets` n x n sheets stacked in z, sheet k at k * spacing * dhat plus a
    per-vertex z-jitter U(-jitter, jitter) dhat.  "Interleaved": odd sheets
    are shifted by half a grid cell in x and y, so no vertex sits exactly
    above a vertex of the neighbouring sheets.  Step: per-vertex in-plane
    N(0, (inplane_sigma dhat)^2) and a uniform dz * dhat in z.  `h` is the
    grid spacing (default 1/(n-1): unit-square sheets); a smaller n with the
    default spacing of a larger scene is a tile of that scene's workload."""
    rng = np.random.default_rng(seed)
    h = 1.0 / (n - 1) if h is None else float(h)
    meshes = []
    for k in range(sheets):
        off = (0.5 * h, 0.5 * h) if (k % 2) else (0.0, 0.0)
        Vk, Fk = grid_sheet(n, 0.0, offset=off, size=h * (n - 1))
        Vk[:, 2] = k * spacing * dhat + rng.uniform(-jitter, jitter, len(Vk)) * dhat
        meshes.append((Vk, Fk))
    V, E, F = merge(meshes)
    V1 = V.copy()
    V1[:, :2] += rng.normal(0.0, inplane_sigma * dhat, (len(V), 2))
    V1[:, 2] += dz * dhat
    return Scene(f"stacked_sheets_{sheets}x{n}", V, E, F, V1, dhat, kappa, seed=seed,
                 interleaved="odd sheets shifted by half a cell in x and y")


def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0)):
    V, F = _unit_icosphere(subdiv)
    return V * radius + np.asarray(center, dtype=np.float64), F.copy()


@functools.lru_cache(maxsize=None)
def _unit_icosphere(subdiv):
    t = (1.0 + math.sqrt(5.0)) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9),
         (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2),
         (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(subdiv):
        cache = {}
        newF = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            newF += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = newF
    V = np.array(V)
    V.flags.writeable = False
    F = np.array(F, dtype=np.int32)
    F.flags.writeable = False
    return V, F


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q *= np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


random_rotation = _rotation


def cube(size=1.0, center=(0.0, 0.0, 0.0)):
    """Axis-aligned cube surface: 8 vertices, 12 outward triangles."""
    V = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (0, 1)], dtype=np.float64)
    V = (V - 0.5) * size + np.asarray(center, dtype=np.float64)
    F = np.array([[0, 2, 1], [1, 2, 3], [4, 5, 6], [5, 7, 6], [0, 1, 4], [1, 5, 4],
                  [2, 6, 3], [3, 6, 7], [0, 4, 2], [2, 4, 6], [1, 3, 5], [3, 7, 5]],
                 dtype=np.int32)
    return V, F


def sphere_chain(count=2, subdiv=3, radius=0.1, seed=0, dhat=1e-3, kappa=1e3,
                 approach=1.5, noise=0.1):
    """`count` randomly rotated icospheres of one subdivision level in a row
    along x, neighbouring surfaces U(0.2, 1.5) dhat apart.  Step: sphere k
    moves by (c - k) * approach * dhat in x towards the chain centre c, plus a
    per-vertex N(0, (noise dhat)^2) perturbation, so neighbours close their
    gap and pass into each other."""
    rng = np.random.default_rng(seed)
    Vs, Fs = icosphere(subdiv, radius)
    # the surface gap uses the polyhedron's extent along x, not the sphere's
    meshes, x = [], 0.0
    shifts = []
    for k in range(count):
        R = _rotation(rng)
        Vk = Vs @ R.T
        lo, hi = Vk[:, 0].min(), Vk[:, 0].max()
        if k:
            x += rng.uniform(0.2, 1.5) * dhat
        Vk = Vk + np.array([x - lo, 0.0, 0.0])
        x += hi - lo
        meshes.append((Vk, Fs))
        shifts.append(((count - 1) / 2.0 - k) * approach * dhat)
    V, E, F = merge(meshes)
    V1 = V.copy()
    nv = len(Vs)
    for k in range(count):
        V1[k * nv:(k + 1) * nv, 0] += shifts[k]
    V1 += rng.normal(0.0, noise * dhat, V1.shape)
    return Scene(f"sphere_chain_{count}x{subdiv}", V, E, F, V1, dhat, kappa, seed=seed)


MAX_BATCH_TRIS = 300_000


def batch_scenes(count=256, seed=0, dhat=1e-3, kappa=1e3):
    """`count` independent scenes of uneven size: each, with probability 1/2,
    a sheet pair (two_sheets with side drawn log-uniform in [32, 384]
    vertices) or a row of icospheres (sphere_chain: subdivision U{3,4,5},
    radius 0.1, count U{2..16}, neighbouring surface gaps U(0.2, 1.5) dhat).
    Draws whose triangle count exceeds MAX_BATCH_TRIS are drawn again, so
    scene sizes run from 3.8k to 300k triangles (a sheet pair of side 384
    would have 586k: the triangle range is the rule kept)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        sub_seed = int(rng.integers(1 << 31))
        while True:
            if rng.random() < 0.5:
                n = int(round(math.exp(rng.uniform(math.log(32), math.log(384)))))
                tris = 4 * (n - 1) ** 2
                if tris > MAX_BATCH_TRIS:
                    continue
                s = two_sheets(n, seed=sub_seed, dhat=dhat, kappa=kappa)
            else:
                sub = int(rng.integers(3, 6))
                cnt = int(rng.integers(2, 17))
                tris = cnt * 20 * 4 ** sub
                if tris > MAX_BATCH_TRIS:
                    continue
                s = sphere_chain(cnt, sub, radius=0.1, seed=sub_seed, dhat=dhat, kappa=kappa)
            break
        s.name = f"batch{k}_{s.name}"
        out.append(s)
    return out


class Batch:
    """Scenes concatenated into one vertex / edge / face numbering: scene s
    owns vertices [voff[s], voff[s+1]) and likewise edges and faces; the
    topology arrays hold global vertex ids."""

    def __init__(self, scenes):
        self.scenes = list(scenes)
        nv = [len(s.V) for s in self.scenes]
        ne = [len(s.E) for s in self.scenes]
        nf = [len(s.F) for s in self.scenes]
        self.voff = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
        self.eoff = np.concatenate([[0], np.cumsum(ne)]).astype(np.int64)
        self.foff = np.concatenate([[0], np.cumsum(nf)]).astype(np.int64)
        self.V = np.concatenate([s.V for s in self.scenes])
        self.V1 = np.concatenate([s.V1 for s in self.scenes])
        self.rest = np.concatenate([s.rest for s in self.scenes])
        # scene-local topology (what the batched upload takes)
        self.E_local = np.concatenate([s.E for s in self.scenes]).astype(np.int32)
        self.F_local = np.concatenate([s.F for s in self.scenes]).astype(np.int32)
        self.E = self.E_local + np.repeat(self.voff[:-1], ne)[:, None].astype(np.int32)
        self.F = self.F_local + np.repeat(self.voff[:-1], nf)[:, None].astype(np.int32)

    def __len__(self):
        return len(self.scenes)

    @property
    def sizes(self):
        return len(self.V), len(self.E), len(self.F)
