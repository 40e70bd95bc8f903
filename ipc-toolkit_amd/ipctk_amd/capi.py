"""ctypes binding of the C ABI in include/ipcg.h (libipcg.so).

This is the thinnest host layer: one `Context` per collision mesh, numpy
arrays in and out.  It raises on any non-zero status with the context's
error message.  The library must be built (``make -C ipc-toolkit_amd/csrc``
or ``__graft_entry__.build()``); importing fails loudly if it is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libipcg.so")

OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_STATE, ERR_NUMERIC, ERR_NO_DEVICE = range(7)
LAYOUT_AOS, LAYOUT_SOA = 0, 1
BP_LBVH, BP_SAP = 0, 1
COLLISIONS_IPC, COLLISIONS_IMPROVED_MAX_APPROX, COLLISIONS_OGC = 0, 1, 2
PSD = {"none": 0, "clamp": 1, "abs": 2}

# every symbol the header declares
EXPORTS = [
    "ipcg_ctx_create", "ipcg_ctx_destroy", "ipcg_last_error", "ipcg_ctx_synchronize",
    "ipcg_ctx_stream", "ipcg_mesh_upload", "ipcg_positions_upload",
    "ipcg_positions_upload_device", "ipcg_broad_phase", "ipcg_candidates_count",
    "ipcg_candidates_download", "ipcg_collisions_build",
    "ipcg_collisions_build_from_candidates", "ipcg_collisions_count",
    "ipcg_collisions_download", "ipcg_barrier_energy", "ipcg_barrier_gradient",
    "ipcg_barrier_hessian", "ipcg_hessian_download", "ipcg_device_results",
    "ipcg_compute_collision_free_stepsize", "ipcg_candidates_stepsize",
    "ipcg_is_step_collision_free", "ipcg_candidates_toi", "ipcg_stage_times",
    "ipcg_stage_times_reset", "ipcg_mesh_codim", "ipcg_codim_candidates_count",
    "ipcg_codim_candidates_download", "ipcg_codim_candidates_toi",
    "ipcg_candidates_is_step_collision_free", "ipcg_counter", "ipcg_mesh_filter",
    "ipcg_edge_face_candidates", "ipcg_edge_face_candidates_download", "ipcg_has_intersections",
    "ipcg_barrier_gauss_newton_diagonal", "ipcg_barrier_gauss_newton_quadratic_form",
    "ipcg_candidates_per_vertex_safe_distances",
    "ipcg_candidates_noncandidate_conservative_stepsize", "ipcg_candidates_cfl_stepsize",
    "ipcg_collisions_set_area_weighting", "ipcg_collisions_set_type",
    "ipcg_friction_build", "ipcg_friction_build_per_vertex", "ipcg_friction_download", "ipcg_velocities_upload",
    "ipcg_friction_energy", "ipcg_friction_gradient", "ipcg_friction_hessian",
    "ipcg_friction_set_anisotropy", "ipcg_friction_update_lagged_mu", "ipcg_barrier_set_physical",
    "ipcg_barrier_energy_noalloc", "ipcg_ctx_device_bytes", "ipcg_ctx_set_autotune", "ipcg_launch_shape",
    "ipcg_mesh_upload_batched", "ipcg_batch_size", "ipcg_batch_set_active",
    "ipcg_collisions_count_batched", "ipcg_barrier_energy_batched",
    "ipcg_compute_collision_free_stepsize_batched",
    "ipcg_server_create", "ipcg_server_destroy", "ipcg_server_last_error", "ipcg_server_add_scene",
    "ipcg_server_submit", "ipcg_server_run", "ipcg_server_result", "ipcg_server_gradient",
    "ipcg_server_hessian", "ipcg_server_positions", "ipcg_server_get_stats", "ipcg_server_context",
]

REQ_ENERGY, REQ_GRADIENT, REQ_HESSIAN, REQ_STEP = range(4)


class RequestResult(C.Structure):
    _fields_ = [("arrival", C.c_double), ("start", C.c_double), ("finish", C.c_double),
                ("value", C.c_double), ("pairs", C.c_int64), ("accepted", C.c_int32),
                ("iteration", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("epochs", C.c_int64), ("requests", C.c_int64),
                ("peak_device_bytes", C.c_int64), ("estimated_bytes", C.c_int64),
                ("busy_seconds", C.c_double), ("wall_seconds", C.c_double)]

_lib = None
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_L = C.POINTER(C.c_int64)


class IPCGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ipcg error {code}: {msg}")
        self.code = code


def load(path=LIB_PATH):
    """Load libipcg.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} not built: run `make -C ipc-toolkit_amd/csrc`")
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ipcg_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.ipcg_ctx_destroy.argtypes = [vp]
    L.ipcg_last_error.argtypes = [vp]
    L.ipcg_last_error.restype = C.c_char_p
    L.ipcg_ctx_synchronize.argtypes = [vp]
    L.ipcg_ctx_stream.argtypes = [vp]
    L.ipcg_ctx_stream.restype = vp
    L.ipcg_mesh_upload.argtypes = [vp, C.c_int32, _D, C.c_int, C.c_int32, _I, C.c_int32, _I, _I]
    L.ipcg_positions_upload.argtypes = [vp, C.c_int, _D, C.c_int]
    L.ipcg_positions_upload_device.argtypes = [vp, C.c_int, vp, C.c_int]
    L.ipcg_broad_phase.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.ipcg_candidates_count.argtypes = [vp, _L, _L]
    L.ipcg_candidates_download.argtypes = [vp, _I, _I]
    L.ipcg_collisions_build.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.ipcg_collisions_build_from_candidates.argtypes = [vp, C.c_double, C.c_double]
    L.ipcg_collisions_count.argtypes = [vp, _L]
    L.ipcg_collisions_download.argtypes = [vp, _I, _D, _I, _D, _I, _D, _D, _I, _I, _D]
    L.ipcg_barrier_energy.argtypes = [vp, C.c_double, C.c_double, _D]
    L.ipcg_barrier_energy_noalloc.argtypes = [vp, C.c_double, C.c_double, vp]
    L.ipcg_barrier_gradient.argtypes = [vp, C.c_double, C.c_double, _D]
    L.ipcg_barrier_hessian.argtypes = [vp, C.c_double, C.c_double, C.c_int, _L]
    L.ipcg_hessian_download.argtypes = [vp, _L, _I, _D]
    L.ipcg_device_results.argtypes = [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                                      C.POINTER(vp)]
    L.ipcg_compute_collision_free_stepsize.argtypes = [vp, C.c_double, C.c_double, C.c_int64,
                                                       C.c_double, C.c_int, _D]
    L.ipcg_candidates_stepsize.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_double, _D]
    L.ipcg_is_step_collision_free.argtypes = [vp, C.c_double, C.c_double, C.c_int64,
                                              C.c_double, C.c_int, C.POINTER(C.c_int)]
    L.ipcg_candidates_toi.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_double, _D, _D]
    L.ipcg_mesh_codim.argtypes = [vp, _I, _I]
    L.ipcg_mesh_filter.argtypes = [vp, C.c_int32, C.c_int32]
    L.ipcg_counter.argtypes = [vp, C.c_char_p, _L]
    L.ipcg_candidates_is_step_collision_free.argtypes = [vp, C.c_double, C.c_double, C.c_int64,
                                                         C.c_double, C.POINTER(C.c_int)]
    L.ipcg_codim_candidates_count.argtypes = [vp, _L, _L]
    L.ipcg_codim_candidates_download.argtypes = [vp, _I, _I]
    L.ipcg_codim_candidates_toi.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_double,
                                            _D, _D]
    L.ipcg_edge_face_candidates.argtypes = [vp, C.c_double, _L]
    L.ipcg_edge_face_candidates_download.argtypes = [vp, _I]
    L.ipcg_has_intersections.argtypes = [vp, C.POINTER(C.c_int)]
    L.ipcg_collisions_set_area_weighting.argtypes = [vp, C.c_int]
    L.ipcg_collisions_set_type.argtypes = [vp, C.c_int]
    L.ipcg_friction_build.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
    L.ipcg_friction_download.argtypes = [vp, _I, _D, _D, _D]
    L.ipcg_friction_build_per_vertex.argtypes = [vp, C.c_double, C.c_double, _D, _D]
    L.ipcg_friction_set_anisotropy.argtypes = [vp, _D, _D, _D]
    L.ipcg_friction_update_lagged_mu.argtypes = [vp]
    L.ipcg_barrier_set_physical.argtypes = [vp, C.c_int]
    L.ipcg_velocities_upload.argtypes = [vp, _D, C.c_int]
    L.ipcg_friction_energy.argtypes = [vp, C.c_double, _D]
    L.ipcg_friction_gradient.argtypes = [vp, C.c_double, _D]
    L.ipcg_friction_hessian.argtypes = [vp, C.c_double, C.c_int, _L]
    L.ipcg_barrier_gauss_newton_diagonal.argtypes = [vp, C.c_double, C.c_double, _D]
    L.ipcg_barrier_gauss_newton_quadratic_form.argtypes = [vp, C.c_double, C.c_double, _D, _D]
    L.ipcg_candidates_per_vertex_safe_distances.argtypes = [vp, C.c_double, C.c_double, _D]
    L.ipcg_candidates_noncandidate_conservative_stepsize.argtypes = [vp, C.c_double, _D]
    L.ipcg_candidates_cfl_stepsize.argtypes = [vp, C.c_double, C.c_double, C.c_double,
                                               C.c_int64, C.c_double, C.c_int, _D]
    L.ipcg_ctx_device_bytes.argtypes = [vp, _L]
    L.ipcg_ctx_set_autotune.argtypes = [vp, C.c_int]
    L.ipcg_launch_shape.argtypes = [vp, C.c_char_p, _I, _D]
    L.ipcg_mesh_upload_batched.argtypes = [vp, C.c_int32, _L, _D, C.c_int, _L, _I, _L, _I]
    L.ipcg_batch_size.argtypes = [vp, _I]
    L.ipcg_batch_set_active.argtypes = [vp, C.POINTER(C.c_uint8)]
    L.ipcg_collisions_count_batched.argtypes = [vp, _L]
    L.ipcg_barrier_energy_batched.argtypes = [vp, C.c_double, C.c_double, _D]
    L.ipcg_compute_collision_free_stepsize_batched.argtypes = [vp, C.c_double, C.c_double,
                                                               C.c_int64, C.c_double, _D]
    L.ipcg_server_create.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(vp)]
    L.ipcg_server_destroy.argtypes = [vp]
    L.ipcg_server_last_error.argtypes = [vp]
    L.ipcg_server_last_error.restype = C.c_char_p
    L.ipcg_server_add_scene.argtypes = [vp, C.c_int32, _D, C.c_int32, _I, C.c_int32, _I,
                                        C.c_double, C.c_double, C.c_double, C.c_uint64, _I]
    L.ipcg_server_submit.argtypes = [vp, C.c_int32, C.c_int, C.c_double, _L]
    L.ipcg_server_run.argtypes = [vp]
    L.ipcg_server_result.argtypes = [vp, C.c_int64, C.POINTER(RequestResult)]
    L.ipcg_server_gradient.argtypes = [vp, C.c_int64, _D]
    L.ipcg_server_hessian.argtypes = [vp, C.c_int64, _L, _L, _I, _D]
    L.ipcg_server_positions.argtypes = [vp, C.c_int32, _D]
    L.ipcg_server_get_stats.argtypes = [vp, C.POINTER(ServerStats)]
    L.ipcg_server_context.argtypes = [vp]
    L.ipcg_server_context.restype = vp
    L.ipcg_stage_times.argtypes = [vp, C.c_int, C.POINTER(C.c_char_p), _D, _L]
    L.ipcg_stage_times_reset.argtypes = [vp, C.c_int]
    for name in EXPORTS:
        if name not in ("ipcg_last_error", "ipcg_ctx_stream", "ipcg_server_last_error",
                        "ipcg_server_context"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


class Context:
    """One device context: a collision mesh and its device-resident state."""

    def __init__(self, device=0):
        self._L = load()
        self._p = C.c_void_p()
        rc = self._L.ipcg_ctx_create(device, C.byref(self._p))
        if rc != OK:
            raise IPCGError(rc, "ipcg_ctx_create failed (no HIP device?)")
        self.nV = 0

    def close(self):
        if self._p:
            self._L.ipcg_ctx_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise IPCGError(rc, self._L.ipcg_last_error(self._p).decode())

    # ---- mesh / positions --------------------------------------------------
    def upload_mesh(self, rest, E, F, vertex_patch=None):
        rest, pr = _d(rest)
        E, pE = _i(np.asarray(E).reshape(-1, 2))
        F, pF = _i(np.asarray(F).reshape(-1, 3))
        pp = None
        if vertex_patch is not None:
            vertex_patch, pp = _i(vertex_patch)
        self._check(self._L.ipcg_mesh_upload(self._p, len(rest), pr, LAYOUT_AOS, len(E), pE,
                                             len(F), pF, pp))
        self.nV, self.E, self.F = len(rest), E, F

    def upload_positions(self, slot, V, layout="aos"):
        """V: (nV, 3) positions; layout "soa" sends them as x[], y[], z[]."""
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        if layout == "soa":
            V = np.ascontiguousarray(V.T)
        V, pV = _d(V)
        self._check(self._L.ipcg_positions_upload(self._p, slot, pV,
                                                  LAYOUT_SOA if layout == "soa" else LAYOUT_AOS))

    def upload_positions_device(self, slot, ptr, layout="aos"):
        """ptr: device address of nV x 3 float64 positions (e.g. a torch
        tensor's data_ptr()), AoS or SoA."""
        self._check(self._L.ipcg_positions_upload_device(
            self._p, slot, C.c_void_p(ptr),
            LAYOUT_SOA if layout == "soa" else LAYOUT_AOS))

    # ---- broad phase -------------------------------------------------------
    def broad_phase(self, swept=False, inflation=0.0, method=BP_LBVH):
        self._check(self._L.ipcg_broad_phase(self._p, int(swept), inflation, method))

    def candidates(self):
        ne, nf = C.c_int64(), C.c_int64()
        self._check(self._L.ipcg_candidates_count(self._p, C.byref(ne), C.byref(nf)))
        ee = np.zeros((ne.value, 2), np.int32)
        fv = np.zeros((nf.value, 2), np.int32)
        self._check(self._L.ipcg_candidates_download(self._p, ee.ctypes.data_as(_I),
                                                     fv.ctypes.data_as(_I)))
        return ee, fv

    def set_filter(self, n_dynamic=-1, n_codim_vertices=-1):
        """Static obstacle / codim cross filters, intersected with the
        vertex patches given to upload_mesh (-1: off)."""
        self._check(self._L.ipcg_mesh_filter(self._p, int(n_dynamic), int(n_codim_vertices)))

    def codim_counts(self):
        """(codim vertices, codim edges) of the uploaded mesh"""
        a, b = C.c_int32(), C.c_int32()
        self._check(self._L.ipcg_mesh_codim(self._p, C.byref(a), C.byref(b)))
        return a.value, b.value

    def codim_candidates(self):
        """(vv, ev) codim candidate lists of the last broad phase"""
        nv, ne = C.c_int64(), C.c_int64()
        self._check(self._L.ipcg_codim_candidates_count(self._p, C.byref(nv), C.byref(ne)))
        vv = np.zeros((nv.value, 2), np.int32)
        ev = np.zeros((ne.value, 2), np.int32)
        self._check(self._L.ipcg_codim_candidates_download(self._p, vv.ctypes.data_as(_I),
                                                           ev.ctypes.data_as(_I)))
        return vv, ev

    def edge_face_candidates(self, inflation=0.0):
        """(edge, face) pairs over static boxes of positions slot 0
        (BroadPhase::detect_edge_face_candidates)"""
        n = C.c_int64()
        self._check(self._L.ipcg_edge_face_candidates(self._p, inflation, C.byref(n)))
        ef = np.zeros((n.value, 2), np.int32)
        self._check(self._L.ipcg_edge_face_candidates_download(self._p, ef.ctypes.data_as(_I)))
        return ef

    def has_intersections(self):
        """has_intersections(mesh, V) over positions slot 0"""
        r = C.c_int()
        self._check(self._L.ipcg_has_intersections(self._p, C.byref(r)))
        return bool(r.value)

    def per_vertex_safe_distances(self, inflation_radius, min_distance=0.0):
        """Candidates::compute_per_vertex_safe_distances over the current lists"""
        out = np.zeros(self.nV)
        self._check(self._L.ipcg_candidates_per_vertex_safe_distances(
            self._p, inflation_radius, min_distance, out.ctypes.data_as(_D)))
        return out

    def noncandidate_conservative_stepsize(self, dhat):
        r = C.c_double()
        self._check(self._L.ipcg_candidates_noncandidate_conservative_stepsize(
            self._p, dhat, C.byref(r)))
        return r.value

    def cfl_stepsize(self, dhat, min_distance=0.0, tolerance=1e-6, max_iterations=10_000_000,
                     rescale=0.8, method=BP_LBVH):
        r = C.c_double()
        self._check(self._L.ipcg_candidates_cfl_stepsize(self._p, dhat, min_distance, tolerance,
                                                         max_iterations, rescale, method,
                                                         C.byref(r)))
        return r.value

    # ---- collisions --------------------------------------------------------
    def build_collisions(self, dhat, dmin=0.0, method=BP_LBVH):
        self._check(self._L.ipcg_collisions_build(self._p, dhat, dmin, method))

    def set_area_weighting(self, use=True):
        """NormalCollisions::set_use_area_weighting for the builds that follow"""
        self._check(self._L.ipcg_collisions_set_area_weighting(self._p, int(bool(use))))

    def set_collision_set_type(self, collision_set_type):
        """NormalCollisions::set_collision_set_type: COLLISIONS_IPC,
        COLLISIONS_IMPROVED_MAX_APPROX or COLLISIONS_OGC, for the builds that
        follow"""
        self._check(self._L.ipcg_collisions_set_type(self._p, int(collision_set_type)))

    def set_physical_barrier(self, use=True):
        """BarrierPotential use_physical_barrier for the calls that follow"""
        self._check(self._L.ipcg_barrier_set_physical(self._p, int(bool(use))))

    def build_collisions_from_candidates(self, dhat, dmin=0.0):
        self._check(self._L.ipcg_collisions_build_from_candidates(self._p, dhat, dmin))

    def collision_counts(self):
        c = np.zeros(4, np.int64)
        self._check(self._L.ipcg_collisions_count(self._p, c.ctypes.data_as(_L)))
        return c

    def collisions(self):
        n = self.collision_counts()
        out = {
            "vv": np.zeros((n[0], 2), np.int32), "vv_w": np.zeros(n[0]),
            "ev": np.zeros((n[1], 2), np.int32), "ev_w": np.zeros(n[1]),
            "ee": np.zeros((n[2], 2), np.int32), "ee_w": np.zeros(n[2]),
            "ee_eps_x": np.zeros(n[2]), "ee_dtype": np.zeros(n[2], np.int32),
            "fv": np.zeros((n[3], 2), np.int32), "fv_w": np.zeros(n[3]),
        }
        a = out
        self._check(self._L.ipcg_collisions_download(
            self._p, a["vv"].ctypes.data_as(_I), a["vv_w"].ctypes.data_as(_D),
            a["ev"].ctypes.data_as(_I), a["ev_w"].ctypes.data_as(_D),
            a["ee"].ctypes.data_as(_I), a["ee_w"].ctypes.data_as(_D),
            a["ee_eps_x"].ctypes.data_as(_D), a["ee_dtype"].ctypes.data_as(_I),
            a["fv"].ctypes.data_as(_I), a["fv_w"].ctypes.data_as(_D)))
        return out

    # ---- barrier potential -------------------------------------------------
    def energy(self, dhat, kappa):
        e = C.c_double()
        self._check(self._L.ipcg_barrier_energy(self._p, dhat, kappa, C.byref(e)))
        return e.value

    def gradient(self, dhat, kappa, download=True):
        g = np.zeros(3 * self.nV) if download else None
        self._check(self._L.ipcg_barrier_gradient(self._p, dhat, kappa,
                                                  g.ctypes.data_as(_D) if download else None))
        return g

    def gauss_newton_diagonal(self, dhat, kappa):
        """NormalPotential::gauss_newton_hessian_diagonal (3 nV, RowMajor)"""
        d = np.zeros(3 * self.nV)
        self._check(self._L.ipcg_barrier_gauss_newton_diagonal(self._p, dhat, kappa,
                                                               d.ctypes.data_as(_D)))
        return d

    def gauss_newton_quadratic_form(self, dhat, kappa, p):
        """NormalPotential::gauss_newton_hessian_quadratic_form; p: 3 nV RowMajor"""
        p = np.ascontiguousarray(p, dtype=np.float64).reshape(-1)
        if len(p) != 3 * self.nV:
            raise ValueError("p must hold 3 * nV values")
        r = C.c_double()
        self._check(self._L.ipcg_barrier_gauss_newton_quadratic_form(
            self._p, dhat, kappa, p.ctypes.data_as(_D), C.byref(r)))
        return r.value

    def hessian(self, dhat, kappa, psd="clamp", download=True):
        nnz = C.c_int64()
        self._check(self._L.ipcg_barrier_hessian(self._p, dhat, kappa, PSD[psd], C.byref(nnz)))
        if not download:
            return nnz.value
        import scipy.sparse as sp
        n = 3 * self.nV
        rowptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        self._check(self._L.ipcg_hessian_download(self._p, rowptr.ctypes.data_as(_L),
                                                  col.ctypes.data_as(_I), val.ctypes.data_as(_D)))
        return sp.csr_matrix((val, col, rowptr), shape=(n, n))

    # ---- friction ----------------------------------------------------------
    def friction_build(self, dhat, kappa, mu_s, mu_k=None):
        """TangentialCollisions::build over the current collision set at the
        slot-0 positions; mu_s, mu_k (default mu_s): scalars or per-vertex
        arrays"""
        mu_k = mu_s if mu_k is None else mu_k
        if np.ndim(mu_s) == 0 and np.ndim(mu_k) == 0:
            self._check(self._L.ipcg_friction_build(self._p, dhat, kappa, mu_s, mu_k))
            return
        ms, pms = _d(np.broadcast_to(np.asarray(mu_s, float), (self.nV,)))
        mk, pmk = _d(np.broadcast_to(np.asarray(mu_k, float), (self.nV,)))
        self._check(self._L.ipcg_friction_build_per_vertex(self._p, dhat, kappa, pms, pmk))

    def friction_records(self):
        """keep (bool), closest-point coordinates (n, 2), tangent bases
        (n, 3, 2), normal force magnitudes (n,), in collision order"""
        n = int(self.collision_counts().sum())
        keep = np.zeros(n, np.int32)
        beta = np.zeros((n, 2))
        basis = np.zeros((n, 6))
        force = np.zeros(n)
        self._check(self._L.ipcg_friction_download(
            self._p, keep.ctypes.data_as(_I), beta.ctypes.data_as(_D),
            basis.ctypes.data_as(_D), force.ctypes.data_as(_D)))
        return keep.astype(bool), beta, basis.reshape(n, 2, 3).transpose(0, 2, 1), force

    def friction_set_anisotropy(self, mu_aniso=None, mu_s_aniso=None, mu_k_aniso=None):
        """per-collision (n, 2) arrays in collision order; None keeps"""
        keep = []

        def arg(a):
            if a is None:
                return None
            a, p = _d(np.asarray(a, float).reshape(-1, 2))
            keep.append(a)
            return p
        self._check(self._L.ipcg_friction_set_anisotropy(self._p, arg(mu_aniso), arg(mu_s_aniso),
                                                         arg(mu_k_aniso)))

    def friction_update_lagged_mu(self):
        self._check(self._L.ipcg_friction_update_lagged_mu(self._p))

    def upload_velocities(self, V):
        V = np.ascontiguousarray(V, dtype=np.float64).reshape(-1, 3)
        V, pV = _d(V)
        self._check(self._L.ipcg_velocities_upload(self._p, pV, LAYOUT_AOS))

    def friction_energy(self, eps_v):
        e = C.c_double()
        self._check(self._L.ipcg_friction_energy(self._p, eps_v, C.byref(e)))
        return e.value

    def friction_gradient(self, eps_v):
        g = np.zeros(3 * self.nV)
        self._check(self._L.ipcg_friction_gradient(self._p, eps_v, g.ctypes.data_as(_D)))
        return g

    def friction_hessian(self, eps_v, psd="clamp", download=True):
        nnz = C.c_int64()
        self._check(self._L.ipcg_friction_hessian(self._p, eps_v, PSD[psd], C.byref(nnz)))
        if not download:
            return nnz.value
        import scipy.sparse as sp
        n = 3 * self.nV
        rowptr = np.zeros(n + 1, np.int64)
        col = np.zeros(nnz.value, np.int32)
        val = np.zeros(nnz.value)
        self._check(self._L.ipcg_hessian_download(self._p, rowptr.ctypes.data_as(_L),
                                                  col.ctypes.data_as(_I), val.ctypes.data_as(_D)))
        return sp.csr_matrix((val, col, rowptr), shape=(n, n))

    # ---- CCD ---------------------------------------------------------------
    def stepsize(self, min_distance=0.0, tolerance=1e-6, max_iterations=10_000_000,
                 rescale=0.8, method=BP_LBVH):
        s = C.c_double()
        self._check(self._L.ipcg_compute_collision_free_stepsize(
            self._p, min_distance, tolerance, max_iterations, rescale, method, C.byref(s)))
        return s.value

    def candidates_stepsize(self, min_distance=0.0, tolerance=1e-6, max_iterations=10_000_000,
                            rescale=0.8):
        s = C.c_double()
        self._check(self._L.ipcg_candidates_stepsize(self._p, min_distance, tolerance,
                                                     max_iterations, rescale, C.byref(s)))
        return s.value

    def is_step_collision_free(self, min_distance=0.0, tolerance=1e-6,
                               max_iterations=10_000_000, rescale=0.8, method=BP_LBVH):
        r = C.c_int()
        self._check(self._L.ipcg_is_step_collision_free(self._p, min_distance, tolerance,
                                                        max_iterations, rescale, method,
                                                        C.byref(r)))
        return bool(r.value)

    def candidates_toi(self, min_distance=0.0, tolerance=1e-6, max_iterations=10_000_000,
                       rescale=0.8):
        ne, nf = C.c_int64(), C.c_int64()
        self._check(self._L.ipcg_candidates_count(self._p, C.byref(ne), C.byref(nf)))
        te = np.zeros(ne.value)
        tf = np.zeros(nf.value)
        self._check(self._L.ipcg_candidates_toi(self._p, min_distance, tolerance, max_iterations,
                                                rescale, te.ctypes.data_as(_D),
                                                tf.ctypes.data_as(_D)))
        return te, tf

    def codim_candidates_toi(self, min_distance=0.0, tolerance=1e-6,
                             max_iterations=10_000_000, rescale=0.8):
        nv, ne = C.c_int64(), C.c_int64()
        self._check(self._L.ipcg_codim_candidates_count(self._p, C.byref(nv), C.byref(ne)))
        tv = np.zeros(nv.value)
        te = np.zeros(ne.value)
        self._check(self._L.ipcg_codim_candidates_toi(self._p, min_distance, tolerance,
                                                      max_iterations, rescale,
                                                      tv.ctypes.data_as(_D),
                                                      te.ctypes.data_as(_D)))
        return tv, te

    # ---- timing ------------------------------------------------------------
    def timing(self, enable=True):
        self._check(self._L.ipcg_stage_times_reset(self._p, int(enable)))

    def stage_times(self):
        names = (C.c_char_p * 64)()
        ms = np.zeros(64)
        calls = np.zeros(64, np.int64)
        n = self._L.ipcg_stage_times(self._p, 64, names, ms.ctypes.data_as(_D),
                                     calls.ctypes.data_as(_L))
        if n < 0:
            raise IPCGError(-n, self._L.ipcg_last_error(self._p).decode())
        return {names[i].decode(): (ms[i], int(calls[i])) for i in range(n)}

    def counter(self, name):
        v = C.c_int64()
        self._check(self._L.ipcg_counter(self._p, name.encode(), C.byref(v)))
        return v.value

    def synchronize(self):
        self._check(self._L.ipcg_ctx_synchronize(self._p))

    def set_autotune(self, enable=True):
        self._check(self._L.ipcg_ctx_set_autotune(self._p, int(bool(enable))))

    def launch_shape(self, stage):
        """(workgroup size, ms) the autotuner chose for a stage (0: untuned)"""
        b = C.c_int32()
        ms = C.c_double()
        self._check(self._L.ipcg_launch_shape(self._p, stage.encode(), C.byref(b), C.byref(ms)))
        return b.value, ms.value

    def device_bytes(self):
        v = C.c_int64()
        self._check(self._L.ipcg_ctx_device_bytes(self._p, C.byref(v)))
        return v.value

    # ---- batched scenes ----------------------------------------------------
    def upload_batch(self, batch):
        """batch: ipctk_amd.scenes.Batch (scene-local topology with ragged
        vertex / edge / face offsets)."""
        voff = np.ascontiguousarray(batch.voff, dtype=np.int64)
        eoff = np.ascontiguousarray(batch.eoff, dtype=np.int64)
        foff = np.ascontiguousarray(batch.foff, dtype=np.int64)
        rest, pr = _d(batch.rest)
        E, pE = _i(batch.E_local)
        F, pF = _i(batch.F_local)
        self._check(self._L.ipcg_mesh_upload_batched(
            self._p, len(voff) - 1, voff.ctypes.data_as(_L), pr, LAYOUT_AOS,
            eoff.ctypes.data_as(_L), pE, foff.ctypes.data_as(_L), pF))
        self.nV = int(voff[-1])
        self.E, self.F = batch.E, batch.F
        self.n_scenes = len(voff) - 1

    def batch_size(self):
        n = C.c_int32()
        self._check(self._L.ipcg_batch_size(self._p, C.byref(n)))
        return n.value

    def set_active(self, active=None):
        """active: n_scenes booleans (None: every scene); the scenes left out
        report no candidates or collisions until the next call."""
        if active is None:
            self._check(self._L.ipcg_batch_set_active(self._p, None))
            return
        a = np.ascontiguousarray(np.asarray(active, dtype=bool).astype(np.uint8))
        self._check(self._L.ipcg_batch_set_active(self._p, a.ctypes.data_as(
            C.POINTER(C.c_uint8))))

    def collision_counts_batched(self):
        out = np.zeros((self.n_scenes, 4), np.int64)
        self._check(self._L.ipcg_collisions_count_batched(self._p, out.ctypes.data_as(_L)))
        return out

    def energy_batched(self, dhat, kappa):
        out = np.zeros(self.n_scenes)
        self._check(self._L.ipcg_barrier_energy_batched(self._p, dhat, kappa,
                                                        out.ctypes.data_as(_D)))
        return out

    def stepsize_batched(self, min_distance=0.0, tolerance=1e-6, max_iterations=10_000_000,
                         rescale=0.8):
        out = np.zeros(self.n_scenes)
        self._check(self._L.ipcg_compute_collision_free_stepsize_batched(
            self._p, min_distance, tolerance, max_iterations, rescale, out.ctypes.data_as(_D)))
        return out


class Server:
    """The serving loop (ipcg_server_*): scenes, timed requests, one run."""

    def __init__(self, hbm_budget_bytes=64 << 30, keep_results=False, device=0):
        self._L = load()
        self._p = C.c_void_p()
        rc = self._L.ipcg_server_create(device, int(hbm_budget_bytes), int(bool(keep_results)),
                                        C.byref(self._p))
        if rc != OK:
            raise IPCGError(rc, "ipcg_server_create failed")
        self.nV = []
        self.kinds = []

    def close(self):
        if self._p:
            self._L.ipcg_server_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise IPCGError(rc, self._L.ipcg_server_last_error(self._p).decode())

    def add_scene(self, V, E, F, dhat, kappa, sigma, seed):
        V, pV = _d(np.asarray(V, dtype=np.float64).reshape(-1, 3))
        E, pE = _i(np.asarray(E).reshape(-1, 2))
        F, pF = _i(np.asarray(F).reshape(-1, 3))
        sid = C.c_int32()
        self._check(self._L.ipcg_server_add_scene(self._p, len(V), pV, len(E), pE, len(F), pF,
                                                  dhat, kappa, sigma, C.c_uint64(seed),
                                                  C.byref(sid)))
        self.nV.append(len(V))
        return sid.value

    def submit(self, scene, kind, arrival):
        t = C.c_int64()
        self._check(self._L.ipcg_server_submit(self._p, scene, kind, float(arrival), C.byref(t)))
        self.kinds.append((scene, kind))
        return t.value

    def run(self):
        self._check(self._L.ipcg_server_run(self._p))

    def result(self, ticket):
        r = RequestResult()
        self._check(self._L.ipcg_server_result(self._p, ticket, C.byref(r)))
        return r

    def gradient(self, ticket):
        scene = self.kinds[ticket][0]
        g = np.zeros(3 * self.nV[scene])
        self._check(self._L.ipcg_server_gradient(self._p, ticket, g.ctypes.data_as(_D)))
        return g

    def hessian(self, ticket):
        import scipy.sparse as sp
        scene = self.kinds[ticket][0]
        n = 3 * self.nV[scene]
        nnz = C.c_int64()
        self._check(self._L.ipcg_server_hessian(self._p, ticket, C.byref(nnz), None, None, None))
        rowptr = np.zeros(n + 1, np.int64)
        col = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1))
        self._check(self._L.ipcg_server_hessian(self._p, ticket, None,
                                                rowptr.ctypes.data_as(_L),
                                                col.ctypes.data_as(_I), val.ctypes.data_as(_D)))
        return sp.csr_matrix((val[:nnz.value], col[:nnz.value], rowptr), shape=(n, n))

    def positions(self, scene):
        V = np.zeros((self.nV[scene], 3))
        self._check(self._L.ipcg_server_positions(self._p, scene, V.ctypes.data_as(_D)))
        return V

    def context(self):
        """the server's batched context as a Context view (stage times,
        counters); it stays owned by the server"""
        view = Context.__new__(Context)
        view._L = self._L
        view._p = C.c_void_p(self._L.ipcg_server_context(self._p))
        view.nV = 0
        view.close = lambda: None
        return view

    def stats(self):
        st = ServerStats()
        self._check(self._L.ipcg_server_get_stats(self._p, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}
