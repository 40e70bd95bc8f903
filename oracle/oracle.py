"""ctypes front end of the CPU restatement (oracle/libipc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py, always as the
checker or the timed CPU baseline, never as the product path.

Positions are (nV, 3) float64 C-contiguous arrays (AoS), edges (nE, 2) and
faces (nF, 3) int32.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libipc_oracle.so")
_lib = None

PSD = {"none": 0, "clamp": 1, "abs": 2}


def build(force=False):
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s"], cwd=_HERE, check=True)
    return _LIB_PATH


class _Pairs(C.Structure):
    _fields_ = [("n", C.c_int64), ("ij", C.POINTER(C.c_int32))]


class _Collisions(C.Structure):
    _fields_ = [
        ("nvv", C.c_int64), ("nev", C.c_int64), ("nee", C.c_int64), ("nfv", C.c_int64),
        ("vv", C.POINTER(C.c_int32)), ("vv_w", C.POINTER(C.c_double)),
        ("ev", C.POINTER(C.c_int32)), ("ev_w", C.POINTER(C.c_double)),
        ("ee", C.POINTER(C.c_int32)), ("ee_w", C.POINTER(C.c_double)),
        ("ee_eps_x", C.POINTER(C.c_double)), ("ee_dtype", C.POINTER(C.c_int32)),
        ("fv", C.POINTER(C.c_int32)), ("fv_w", C.POINTER(C.c_double)),
        ("dmin", C.c_double),
    ]


class _CSR(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double))]


class _Friction(C.Structure):
    _fields_ = [("n", C.c_int64), ("keep", C.POINTER(C.c_uint8)), ("beta", C.POINTER(C.c_double)),
                ("basis", C.POINTER(C.c_double)), ("normal_force", C.POINTER(C.c_double)),
                ("weight", C.POINTER(C.c_double)), ("mu_s", C.POINTER(C.c_double)),
                ("mu_k", C.POINTER(C.c_double)), ("mu_aniso", C.POINTER(C.c_double)),
                ("mu_s_aniso", C.POINTER(C.c_double)), ("mu_k_aniso", C.POINTER(C.c_double)),
                ("mu_s_eff", C.POINTER(C.c_double)), ("mu_k_eff", C.POINTER(C.c_double))]


class CCDParams(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("max_iterations", C.c_long),
                ("conservative_rescaling", C.c_double)]


DEFAULT_CCD = CCDParams(1e-6, 10_000_000, 0.8)

_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int32)
_F = C.POINTER(C.c_float)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        for name in ("ipco_point_edge_distance_type",):
            getattr(L, name).argtypes = [_D, _D, _D]
            getattr(L, name).restype = C.c_int
        for name in ("ipco_point_triangle_distance_type", "ipco_edge_edge_distance_type"):
            getattr(L, name).argtypes = [_D, _D, _D, _D]
            getattr(L, name).restype = C.c_int
        L.ipco_point_point_distance.argtypes = [_D, _D]
        L.ipco_point_line_distance.argtypes = [_D, _D, _D]
        L.ipco_point_plane_distance.argtypes = [_D, _D, _D, _D]
        L.ipco_line_line_distance.argtypes = [_D, _D, _D, _D]
        L.ipco_point_edge_distance.argtypes = [_D, _D, _D, C.c_int]
        L.ipco_point_triangle_distance.argtypes = [_D, _D, _D, _D, C.c_int]
        L.ipco_edge_edge_distance.argtypes = [_D, _D, _D, _D, C.c_int]
        L.ipco_stencil_distance_derivs.argtypes = [C.c_int, C.c_int, _D, _D, _D]
        L.ipco_edge_edge_cross_squarednorm.argtypes = [_D, _D, _D, _D]
        L.ipco_edge_edge_mollifier_threshold.argtypes = [_D, _D, _D, _D]
        L.ipco_edge_edge_mollifier.argtypes = [_D, C.c_double, _D, _D]
        for name in ("ipco_point_point_distance", "ipco_point_line_distance",
                     "ipco_point_plane_distance", "ipco_line_line_distance",
                     "ipco_point_edge_distance", "ipco_point_triangle_distance",
                     "ipco_edge_edge_distance", "ipco_stencil_distance_derivs",
                     "ipco_edge_edge_cross_squarednorm",
                     "ipco_edge_edge_mollifier_threshold", "ipco_edge_edge_mollifier",
                     "ipco_barrier", "ipco_barrier_first_derivative",
                     "ipco_barrier_second_derivative"):
            getattr(L, name).restype = C.c_double
        for name in ("ipco_barrier", "ipco_barrier_first_derivative",
                     "ipco_barrier_second_derivative"):
            getattr(L, name).argtypes = [C.c_double, C.c_double]
        L.ipco_project_to_psd.argtypes = [_D, C.c_int, C.c_int]
        L.ipco_symmetric_eigenvalues.argtypes = [_D, C.c_int, _D]
        L.ipco_faces_to_edges.argtypes = [_I, C.c_int, _I, C.c_int, _I]
        L.ipco_build_boxes_f32.argtypes = [_D, _D, C.c_int, _I, C.c_int, _I, C.c_int,
                                           C.c_double, _F, _F, _F]
        L.ipco_candidates.argtypes = [_D, _D, C.c_int, _I, C.c_int, _I, C.c_int,
                                      C.c_double, _I, C.POINTER(_Pairs), C.POINTER(_Pairs)]
        L.ipco_candidates_sap.argtypes = L.ipco_candidates.argtypes
        L.ipco_sap_axis.argtypes = [_D, _D, C.c_int, C.c_double]
        L.ipco_sap_axis.restype = C.c_int
        L.ipco_pairs_free.argtypes = [C.POINTER(_Pairs)]
        L.ipco_build_collisions.argtypes = [_D, C.c_int, _D, _I, C.c_int, _I, C.c_int, _I,
                                            C.POINTER(_Pairs), C.POINTER(_Pairs),
                                            C.POINTER(_Pairs), C.POINTER(_Pairs),
                                            C.c_double, C.c_double, C.POINTER(_Collisions)]
        L.ipco_codim_candidates.argtypes = [_D, _D, C.c_int, _I, C.c_int, _I, C.c_int, _I,
                                            C.c_double, _I, C.POINTER(_Pairs),
                                            C.POINTER(_Pairs)]
        L.ipco_codim_lists.argtypes = [C.c_int, _I, C.c_int, _I, C.c_int, _I, _I, _I]
        L.ipco_collisions_free.argtypes = [C.POINTER(_Collisions)]
        L.ipco_barrier_energy.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I,
                                          C.c_double, C.c_double]
        L.ipco_barrier_energy.restype = C.c_double
        L.ipco_barrier_gradient.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I,
                                            C.c_double, C.c_double, _D]
        L.ipco_barrier_hessian.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I,
                                           C.c_double, C.c_double, C.c_int, C.POINTER(_CSR)]
        L.ipco_csr_free.argtypes = [C.POINTER(_CSR)]
        L.ipco_ticcd.argtypes = [C.c_int, _D, _D, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_long, C.c_int, _D, _D]
        L.ipco_stencil_ccd.argtypes = [C.c_int, _D, _D, C.c_double, C.c_double,
                                       C.POINTER(CCDParams), _D]
        L.ipco_candidates_stepsize.argtypes = [_D, _D, C.c_int, _I, _I, C.POINTER(_Pairs),
                                               C.POINTER(_Pairs), C.POINTER(_Pairs),
                                               C.POINTER(_Pairs), C.c_double,
                                               C.POINTER(CCDParams), C.c_int]
        L.ipco_candidates_stepsize.restype = C.c_double
        L.ipco_compute_collision_free_stepsize.argtypes = [
            _D, _D, C.c_int, _I, C.c_int, _I, C.c_int, C.c_double, C.POINTER(CCDParams),
            C.c_int]
        L.ipco_compute_collision_free_stepsize.restype = C.c_double
        L.ipco_is_step_collision_free.argtypes = [_D, _D, C.c_int, _I, C.c_int, _I, C.c_int,
                                                  C.c_double, C.POINTER(CCDParams)]
        L.ipco_candidates_toi.argtypes = [_D, _D, _I, _I, C.POINTER(_Pairs), C.c_int,
                                          C.c_double, C.POINTER(CCDParams), _D]
        L.ipco_edge_face_candidates.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int,
                                                 C.c_double, _I, C.POINTER(_Pairs)]
        L.ipco_orient3d.argtypes = [_D, _D, _D, _D]
        L.ipco_orient3d.restype = C.c_int
        L.ipco_orient3d_exact_calls.restype = C.c_longlong
        L.ipco_fullpivlu_solve3.argtypes = [_D, _D, _D]
        L.ipco_is_edge_intersecting_triangle.argtypes = [_D, _D, _D, _D, _D]
        L.ipco_is_edge_intersecting_triangle.restype = C.c_int
        L.ipco_world_bbox_diagonal.argtypes = [_D, C.c_int]
        L.ipco_world_bbox_diagonal.restype = C.c_double
        L.ipco_has_intersections.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int, _I,
                                             C.POINTER(C.c_int)]
        L.ipco_distance_vector.argtypes = [C.POINTER(_Collisions), _D, _I, _I, C.c_int64, _D,
                                           _D]
        L.ipco_distance_vector.restype = C.c_double
        L.ipco_gauss_newton_diagonal.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I,
                                                 C.c_double, C.c_double, _D]
        L.ipco_gauss_newton_quadratic_form.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I,
                                                       _I, C.c_double, C.c_double, _D]
        L.ipco_gauss_newton_quadratic_form.restype = C.c_double
        L.ipco_per_vertex_safe_distances.argtypes = [
            _D, C.c_int, _I, _I, C.POINTER(_Pairs), C.POINTER(_Pairs), C.POINTER(_Pairs),
            C.POINTER(_Pairs), C.c_double, C.c_double, _D]
        L.ipco_noncandidate_conservative_stepsize.argtypes = [
            _D, _D, C.c_int, _I, _I, C.POINTER(_Pairs), C.POINTER(_Pairs), C.POINTER(_Pairs),
            C.POINTER(_Pairs), C.c_double]
        L.ipco_noncandidate_conservative_stepsize.restype = C.c_double
        L.ipco_build_collisions_weighted.argtypes = [
            _D, C.c_int, _D, _I, C.c_int, _I, C.c_int, _I, C.POINTER(_Pairs), C.POINTER(_Pairs),
            C.POINTER(_Pairs), C.POINTER(_Pairs), C.c_double, C.c_double, _D, _D,
            C.POINTER(_Collisions)]
        L.ipco_build_collisions_ex.argtypes = [
            _D, C.c_int, _D, _I, C.c_int, _I, C.c_int, _I, C.POINTER(_Pairs), C.POINTER(_Pairs),
            C.POINTER(_Pairs), C.POINTER(_Pairs), C.c_double, C.c_double, _D, _D, C.c_int,
            C.POINTER(_Collisions)]
        L.ipco_ogc_vertex_region.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int, _I, _D,
                                             C.c_int]
        L.ipco_ogc_edge_region.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int, _I, C.c_int,
                                           C.c_int]
        L.ipco_ogc_feasible.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int, _I, C.c_int,
                                        C.c_int, C.c_int]
        L.ipco_friction_build.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.POINTER(_Friction)]
        L.ipco_friction_free.argtypes = [C.POINTER(_Friction)]
        L.ipco_friction_set_anisotropy.argtypes = [C.POINTER(_Friction), _D, _D, _D]
        L.ipco_friction_update_lagged_mu.argtypes = [C.POINTER(_Friction), C.POINTER(_Collisions),
                                                     _D, _I, _I]
        L.ipco_anisotropic_mu_eff.argtypes = [_D, _D, _D, C.c_double, C.c_double, _D]
        L.ipco_friction_build_pv.argtypes = [C.POINTER(_Collisions), _D, C.c_int, _I, _I,
                                             C.c_double, C.c_double, _D, _D, C.POINTER(_Friction)]
        L.ipco_friction_energy.argtypes = [C.POINTER(_Friction), C.POINTER(_Collisions), _D, C.c_int,
                                           _I, _I, C.c_double]
        L.ipco_friction_energy.restype = C.c_double
        L.ipco_friction_gradient.argtypes = [C.POINTER(_Friction), C.POINTER(_Collisions), _D,
                                             C.c_int, _I, _I, C.c_double, _D]
        L.ipco_friction_hessian.argtypes = [C.POINTER(_Friction), C.POINTER(_Collisions), _D,
                                            C.c_int, _I, _I, C.c_double, C.c_int, C.POINTER(_CSR)]
        L.ipco_smooth_mu_fn.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double]
        L.ipco_smooth_mu_fn.restype = C.c_double
        L.ipco_mesh_areas.argtypes = [_D, C.c_int, _I, C.c_int, _I, C.c_int, _I, _D, _D]
        L.ipco_set_num_threads.argtypes = [C.c_int]
        L.ipco_get_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_D)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I)


def set_num_threads(n):
    lib().ipco_set_num_threads(int(n))


def get_max_threads():
    return lib().ipco_get_max_threads()


# ---- primitives ------------------------------------------------------------

def point_edge_distance_type(p, e0, e1):
    return lib().ipco_point_edge_distance_type(_d(p)[1], _d(e0)[1], _d(e1)[1])


def point_triangle_distance_type(p, t0, t1, t2):
    a = [_d(x) for x in (p, t0, t1, t2)]
    return lib().ipco_point_triangle_distance_type(*[x[1] for x in a])


def edge_edge_distance_type(ea0, ea1, eb0, eb1):
    a = [_d(x) for x in (ea0, ea1, eb0, eb1)]
    return lib().ipco_edge_edge_distance_type(*[x[1] for x in a])


def point_point_distance(p0, p1):
    return lib().ipco_point_point_distance(_d(p0)[1], _d(p1)[1])


def point_line_distance(p, e0, e1):
    return lib().ipco_point_line_distance(_d(p)[1], _d(e0)[1], _d(e1)[1])


def point_plane_distance(p, t0, t1, t2):
    a = [_d(x) for x in (p, t0, t1, t2)]
    return lib().ipco_point_plane_distance(*[x[1] for x in a])


def line_line_distance(ea0, ea1, eb0, eb1):
    a = [_d(x) for x in (ea0, ea1, eb0, eb1)]
    return lib().ipco_line_line_distance(*[x[1] for x in a])


def point_edge_distance(p, e0, e1, dtype=-1):
    return lib().ipco_point_edge_distance(_d(p)[1], _d(e0)[1], _d(e1)[1], dtype)


def point_triangle_distance(p, t0, t1, t2, dtype=-1):
    a = [_d(x) for x in (p, t0, t1, t2)]
    return lib().ipco_point_triangle_distance(*[x[1] for x in a], dtype)


def edge_edge_distance(ea0, ea1, eb0, eb1, dtype=-1):
    a = [_d(x) for x in (ea0, ea1, eb0, eb1)]
    return lib().ipco_edge_edge_distance(*[x[1] for x in a], dtype)


STENCIL_NDOF = {0: 12, 1: 12, 2: 9, 3: 6}


def stencil_distance_derivs(kind, x, dtype=-1):
    """kind 0 point-triangle, 1 edge-edge, 2 point-edge, 3 point-point.
    Returns (value, gradient, hessian)."""
    n = STENCIL_NDOF[kind]
    x, px = _d(np.asarray(x, dtype=np.float64).reshape(-1))
    g = np.zeros(n)
    H = np.zeros((n, n))
    v = lib().ipco_stencil_distance_derivs(kind, dtype, px, g.ctypes.data_as(_D),
                                           H.ctypes.data_as(_D))
    return v, g, H


def edge_edge_cross_squarednorm(ea0, ea1, eb0, eb1):
    a = [_d(x) for x in (ea0, ea1, eb0, eb1)]
    return lib().ipco_edge_edge_cross_squarednorm(*[x[1] for x in a])


def edge_edge_mollifier_threshold(ea0, ea1, eb0, eb1):
    a = [_d(x) for x in (ea0, ea1, eb0, eb1)]
    return lib().ipco_edge_edge_mollifier_threshold(*[x[1] for x in a])


def edge_edge_mollifier(x12, eps_x):
    x, px = _d(np.asarray(x12, dtype=np.float64).reshape(12))
    g = np.zeros(12)
    H = np.zeros((12, 12))
    m = lib().ipco_edge_edge_mollifier(px, eps_x, g.ctypes.data_as(_D), H.ctypes.data_as(_D))
    return m, g, H


def barrier(d, dhat):
    return lib().ipco_barrier(d, dhat)


def barrier_first_derivative(d, dhat):
    return lib().ipco_barrier_first_derivative(d, dhat)


def barrier_second_derivative(d, dhat):
    return lib().ipco_barrier_second_derivative(d, dhat)


def project_to_psd(A, method="clamp"):
    A = np.array(A, dtype=np.float64, order="C")
    rc = lib().ipco_project_to_psd(A.ctypes.data_as(_D), A.shape[0], PSD[method])
    if rc != 0:
        raise RuntimeError("unable to project matrix onto positive semi-definite cone")
    return A


def symmetric_eigenvalues(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    lam = np.zeros(A.shape[0])
    lib().ipco_symmetric_eigenvalues(A.ctypes.data_as(_D), A.shape[0], lam.ctypes.data_as(_D))
    return lam


def faces_to_edges(E, F):
    E, pE = _i(E)
    F, pF = _i(F)
    F2E = np.zeros((len(F), 3), dtype=np.int32)
    if lib().ipco_faces_to_edges(pE, len(E), pF, len(F), F2E.ctypes.data_as(_I)) != 0:
        raise RuntimeError("Unable to find edge!")
    return F2E


# ---- broad phase -----------------------------------------------------------

def boxes_f32(V0, E, F, inflation=0.0, V1=None):
    V0, p0 = _d(V0)
    p1 = None
    if V1 is not None:
        V1, p1 = _d(V1)
    E, pE = _i(E)
    F, pF = _i(F)
    vb = np.zeros((len(V0), 6), np.float32)
    eb = np.zeros((len(E), 6), np.float32)
    fb = np.zeros((len(F), 6), np.float32)
    lib().ipco_build_boxes_f32(p0, p1, len(V0), pE, len(E), pF, len(F), inflation,
                               vb.ctypes.data_as(_F), eb.ctypes.data_as(_F),
                               fb.ctypes.data_as(_F))
    return vb, eb, fb


def _take_pairs(p):
    n = p.n
    out = np.ctypeslib.as_array(p.ij, shape=(n * 2,)).reshape(n, 2).copy() if n else \
        np.zeros((0, 2), np.int32)
    lib().ipco_pairs_free(C.byref(p))
    return out


def _pairs_struct(a):
    a = np.ascontiguousarray(a, dtype=np.int32).reshape(-1, 2)
    s = _Pairs(len(a), a.ctypes.data_as(_I))
    return s, a


def candidates(V0, E, F, inflation=0.0, V1=None, vertex_group=None, method="lbvh"):
    """(ee, fv) candidate arrays.  method "lbvh": FP32 boxes, ee rows
    (query, other) in Morton order; "sap": FP64 boxes, ee rows (earlier,
    later) in sweep order."""
    V0, p0 = _d(V0)
    p1 = None
    if V1 is not None:
        V1, p1 = _d(V1)
    E, pE = _i(E)
    F, pF = _i(F)
    pg = None
    if vertex_group is not None:
        vertex_group, pg = _i(vertex_group)
    ee, fv = _Pairs(), _Pairs()
    fn = lib().ipco_candidates_sap if method == "sap" else lib().ipco_candidates
    fn(p0, p1, len(V0), pE, len(E), pF, len(F), inflation, pg, C.byref(ee), C.byref(fv))
    return _take_pairs(ee), _take_pairs(fv)


def codim_lists(nV, E, F):
    """(codim vertex ids, codim edge ids): vertices in no edge, edges in no
    face (collision_mesh.cpp:145-182)."""
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F2E, pF2E = _i(faces_to_edges(E, F) if len(F) else np.zeros((0, 3), np.int32))
    cnt = np.zeros(2, np.int32)
    lib().ipco_codim_lists(nV, pE, len(E), pF2E, len(F2E), None, None,
                           cnt.ctypes.data_as(_I))
    cv = np.zeros(max(cnt[0], 1), np.int32)
    ce = np.zeros(max(cnt[1], 1), np.int32)
    lib().ipco_codim_lists(nV, pE, len(E), pF2E, len(F2E), cv.ctypes.data_as(_I),
                           ce.ctypes.data_as(_I), cnt.ctypes.data_as(_I))
    return cv[:cnt[0]], ce[:cnt[1]]


def codim_candidates(V0, E, F, inflation=0.0, V1=None, vertex_group=None):
    """(vv, ev) codim candidate arrays (candidates.cpp:65-118): vv pairs of
    codim vertices, ev rows (codim edge, codim vertex)."""
    V0, p0 = _d(V0)
    p1 = None
    if V1 is not None:
        V1, p1 = _d(V1)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    F2E, pF2E = _i(faces_to_edges(E, F) if len(F) else np.zeros((0, 3), np.int32))
    pg = None
    if vertex_group is not None:
        vertex_group, pg = _i(vertex_group)
    vv, ev = _Pairs(), _Pairs()
    lib().ipco_codim_candidates(p0, p1, len(V0), pE, len(E), pF, len(F), pF2E, inflation,
                                pg, C.byref(vv), C.byref(ev))
    return _take_pairs(vv), _take_pairs(ev)


def edge_face_candidates(V, E, F, inflation=0.0, vertex_group=None):
    """(edge, face) candidate array over static boxes (lbvh.cpp:771-784)."""
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    pg = None
    if vertex_group is not None:
        vertex_group, pg = _i(vertex_group)
    ef = _Pairs()
    lib().ipco_edge_face_candidates(pV, len(V), pE, len(E), pF, len(F), inflation, pg,
                                    C.byref(ef))
    return _take_pairs(ef)


def orient3d(pa, pb, pc, pd):
    """exact sign of Shewchuk's orient3d: +1, -1 or 0"""
    return lib().ipco_orient3d(_d(pa)[1], _d(pb)[1], _d(pc)[1], _d(pd)[1])


def orient3d_exact_calls():
    return int(lib().ipco_orient3d_exact_calls())


def fullpivlu_solve3(A, b):
    """Eigen FullPivLU<Matrix3d>(A).solve(b)"""
    A = np.asarray(A, dtype=np.float64).reshape(3, 3)
    Acm, pA = _d(A.T.copy())  # column-major
    b, pb = _d(b)
    x = np.zeros(3)
    lib().ipco_fullpivlu_solve3(pA, pb, x.ctypes.data_as(_D))
    return x


def is_edge_intersecting_triangle(e0, e1, t0, t1, t2):
    return bool(lib().ipco_is_edge_intersecting_triangle(
        _d(e0)[1], _d(e1)[1], _d(t0)[1], _d(t1)[1], _d(t2)[1]))


def world_bbox_diagonal(V):
    V, pV = _d(V)
    return lib().ipco_world_bbox_diagonal(pV, len(V))


def has_intersections(V, E, F, vertex_group=None):
    """ipc.cpp:105-166 (3D)"""
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    pg = None
    if vertex_group is not None:
        vertex_group, pg = _i(vertex_group)
    r = C.c_int(0)
    lib().ipco_has_intersections(pV, len(V), pE, len(E), pF, len(F), pg, C.byref(r))
    return bool(r.value)


def sap_axis(V0, inflation=0.0, V1=None):
    V0, p0 = _d(V0)
    p1 = None
    if V1 is not None:
        V1, p1 = _d(V1)
    return lib().ipco_sap_axis(p0, p1, len(V0), inflation)


# ---- narrow phase + potentials ---------------------------------------------

class Collisions:
    """Collision set (canonical order) holding the C struct for potentials."""

    def __init__(self, s, keep):
        self._s = s
        self._keep = keep

        def arr(p, n, k=1, dt=np.int32):
            if n == 0:
                return np.zeros((0, k) if k > 1 else (0,), dt)
            a = np.ctypeslib.as_array(p, shape=(n * k,)).copy()
            return a.reshape(n, k) if k > 1 else a

        self.vv = arr(s.vv, s.nvv, 2)
        self.vv_w = arr(s.vv_w, s.nvv, 1, np.float64)
        self.ev = arr(s.ev, s.nev, 2)
        self.ev_w = arr(s.ev_w, s.nev, 1, np.float64)
        self.ee = arr(s.ee, s.nee, 2)
        self.ee_w = arr(s.ee_w, s.nee, 1, np.float64)
        self.ee_eps_x = arr(s.ee_eps_x, s.nee, 1, np.float64)
        self.ee_dtype = arr(s.ee_dtype, s.nee, 1)
        self.fv = arr(s.fv, s.nfv, 2)
        self.fv_w = arr(s.fv_w, s.nfv, 1, np.float64)

    def __len__(self):
        return len(self.vv) + len(self.ev) + len(self.ee) + len(self.fv)

    def __del__(self):
        try:
            lib().ipco_collisions_free(C.byref(self._s))
        except Exception:
            pass


def mesh_areas(V_rest, E, F):
    """(vertex areas, edge areas) of CollisionMesh::init_areas"""
    V, pV = _d(V_rest)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    F2E, pF2E = _i(faces_to_edges(E, F) if len(F) else np.zeros((0, 3), np.int32))
    va = np.zeros(len(V))
    ea = np.zeros(max(len(E), 1))
    lib().ipco_mesh_areas(pV, len(V), pE, len(E), pF, len(F), pF2E, va.ctypes.data_as(_D),
                          ea.ctypes.data_as(_D))
    return va, ea[:len(E)]


def build_collisions(V, E, F, ee_cand, fv_cand, dhat, dmin=0.0, V_rest=None, vv_cand=None,
                     ev_cand=None, area_weighting=False, improved=False, ogc=False):
    """NormalCollisions::build over the given candidates; area_weighting,
    improved (CollisionSetType::IMPROVED_MAX_APPROX) and ogc
    (CollisionSetType::OGC) select the variants."""
    V, pV = _d(V)
    if V_rest is None:
        V_rest = V
    V_rest, pR = _d(V_rest)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    F2E = faces_to_edges(E, F) if len(F) else np.zeros((0, 3), np.int32)
    F2E, pF2E = _i(F2E)
    pva = pea = None
    if area_weighting:
        va, ea = mesh_areas(V_rest, E, F)
        va, pva = _d(va)
        ea, pea = _d(ea if len(ea) else np.zeros(1))
    ee_s, ee_a = _pairs_struct(ee_cand)
    fv_s, fv_a = _pairs_struct(fv_cand)
    vv_s, vv_a = _pairs_struct(np.zeros((0, 2)) if vv_cand is None else vv_cand)
    ev_s, ev_a = _pairs_struct(np.zeros((0, 2)) if ev_cand is None else ev_cand)
    out = _Collisions()
    lib().ipco_build_collisions_ex(pV, len(V), pR, pE, len(E), pF, len(F), pF2E, C.byref(vv_s),
                                   C.byref(ev_s), C.byref(ee_s), C.byref(fv_s), dhat, dmin, pva,
                                   pea, 2 if ogc else int(bool(improved)), C.byref(out))
    return Collisions(out, (E, F))


class CollisionArrays:
    """A collision set given by explicit arrays (canonical order and layout
    of Collisions), e.g. a subset of another set, for the potentials.  The
    struct points into numpy arrays that this object keeps alive."""

    def __init__(self, vv, vv_w, ev, ev_w, ee, ee_w, ee_eps_x, ee_dtype, fv, fv_w, dmin=0.0):
        def ids(a):
            return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1, 2))

        def vals(a, dt=np.float64):
            return np.ascontiguousarray(np.asarray(a, dtype=dt).reshape(-1))

        self.vv, self.vv_w = ids(vv), vals(vv_w)
        self.ev, self.ev_w = ids(ev), vals(ev_w)
        self.ee, self.ee_w = ids(ee), vals(ee_w)
        self.ee_eps_x, self.ee_dtype = vals(ee_eps_x), vals(ee_dtype, np.int32)
        self.fv, self.fv_w = ids(fv), vals(fv_w)
        s = _Collisions()
        s.nvv, s.nev, s.nee, s.nfv = len(self.vv), len(self.ev), len(self.ee), len(self.fv)
        s.vv, s.vv_w = self.vv.ctypes.data_as(_I), self.vv_w.ctypes.data_as(_D)
        s.ev, s.ev_w = self.ev.ctypes.data_as(_I), self.ev_w.ctypes.data_as(_D)
        s.ee, s.ee_w = self.ee.ctypes.data_as(_I), self.ee_w.ctypes.data_as(_D)
        s.ee_eps_x = self.ee_eps_x.ctypes.data_as(_D)
        s.ee_dtype = self.ee_dtype.ctypes.data_as(_I)
        s.fv, s.fv_w = self.fv.ctypes.data_as(_I), self.fv_w.ctypes.data_as(_D)
        s.dmin = dmin
        self._s = s

    def __len__(self):
        return len(self.vv) + len(self.ev) + len(self.ee) + len(self.fv)


def collisions_subset(coll, keep_vv, keep_ev, keep_ee, keep_fv, dmin=0.0):
    """The collisions of `coll` selected by one boolean mask per kind."""
    return CollisionArrays(coll.vv[keep_vv], coll.vv_w[keep_vv], coll.ev[keep_ev],
                           coll.ev_w[keep_ev], coll.ee[keep_ee], coll.ee_w[keep_ee],
                           coll.ee_eps_x[keep_ee], coll.ee_dtype[keep_ee], coll.fv[keep_fv],
                           coll.fv_w[keep_fv], dmin)


def barrier_energy(coll, V, E, F, dhat, kappa):
    V, pV = _d(V)
    E, pE = _i(E)
    F, pF = _i(F)
    return lib().ipco_barrier_energy(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa)


def barrier_gradient(coll, V, E, F, dhat, kappa):
    V, pV = _d(V)
    E, pE = _i(E)
    F, pF = _i(F)
    g = np.zeros(3 * len(V))
    lib().ipco_barrier_gradient(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa,
                                g.ctypes.data_as(_D))
    return g


def barrier_hessian(coll, V, E, F, dhat, kappa, psd="clamp"):
    """scipy.sparse.csr_matrix (3nV x 3nV)."""
    import scipy.sparse as sp
    V, pV = _d(V)
    E, pE = _i(E)
    F, pF = _i(F)
    m = _CSR()
    rc = lib().ipco_barrier_hessian(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa,
                                    PSD[psd], C.byref(m))
    n, nnz = m.n, m.nnz
    rowptr = np.ctypeslib.as_array(m.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(m.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().ipco_csr_free(C.byref(m))
    if rc != 0:
        raise RuntimeError("unable to project matrix onto positive semi-definite cone")
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


# ---- CCD -------------------------------------------------------------------

def ticcd(is_vf, x_t0, x_t1, ms=0.0, tolerance=1e-6, t_max=1.0, max_itr=10_000_000,
          no_zero_toi=False, err=-1.0):
    x0, p0 = _d(np.asarray(x_t0, dtype=np.float64).reshape(12))
    x1, p1 = _d(np.asarray(x_t1, dtype=np.float64).reshape(12))
    toi = C.c_double(0)
    otol = C.c_double(0)
    hit = lib().ipco_ticcd(int(is_vf), p0, p1, err, ms, tolerance, t_max, max_itr,
                           int(no_zero_toi), C.byref(toi), C.byref(otol))
    return bool(hit), toi.value, otol.value


def stencil_ccd(kind, x_t0, x_t1, min_distance=0.0, tmax=1.0, params=DEFAULT_CCD):
    """kind 0 point-triangle, 1 edge-edge, 2 point-edge, 3 point-point.
    Returns (hit, toi)."""
    x0, p0 = _d(np.asarray(x_t0, dtype=np.float64).reshape(-1))
    x1, p1 = _d(np.asarray(x_t1, dtype=np.float64).reshape(-1))
    toi = C.c_double(np.inf)
    hit = lib().ipco_stencil_ccd(kind, p0, p1, min_distance, tmax, C.byref(params),
                                 C.byref(toi))
    return bool(hit), toi.value


CAND_KIND = {"fv": 0, "ee": 1, "ev": 2, "vv": 3}


def candidates_toi(V0, V1, E, F, cand, is_ee, min_distance=0.0, params=DEFAULT_CCD):
    """is_ee: True (ee list), False (fv list) or a kind name ("ee", "fv",
    "ev", "vv")."""
    V0, p0 = _d(V0)
    V1, p1 = _d(V1)
    E, pE = _i(E)
    F, pF = _i(F)
    s, a = _pairs_struct(cand)
    out = np.zeros(len(a))
    kind = CAND_KIND[is_ee] if isinstance(is_ee, str) else int(bool(is_ee))
    lib().ipco_candidates_toi(p0, p1, pE, pF, C.byref(s), kind, min_distance,
                              C.byref(params), out.ctypes.data_as(_D))
    return out


def candidates_stepsize(V0, V1, E, F, ee, fv, min_distance=0.0, params=DEFAULT_CCD,
                        running_tmax=False, vv=None, ev=None):
    V0, p0 = _d(V0)
    V1, p1 = _d(V1)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    es, ea = _pairs_struct(ee)
    fs, fa = _pairs_struct(fv)
    vs, va = _pairs_struct(np.zeros((0, 2)) if vv is None else vv)
    ws, wa = _pairs_struct(np.zeros((0, 2)) if ev is None else ev)
    return lib().ipco_candidates_stepsize(p0, p1, len(V0), pE, pF, C.byref(es),
                                          C.byref(fs), C.byref(vs), C.byref(ws),
                                          min_distance, C.byref(params), int(running_tmax))


def compute_collision_free_stepsize(V0, V1, E, F, min_distance=0.0, params=DEFAULT_CCD,
                                    running_tmax=False):
    V0, p0 = _d(V0)
    V1, p1 = _d(V1)
    E, pE = _i(E)
    F, pF = _i(F)
    return lib().ipco_compute_collision_free_stepsize(p0, p1, len(V0), pE, len(E), pF,
                                                      len(F), min_distance,
                                                      C.byref(params), int(running_tmax))


def distance_vector(coll, V, E, F, i):
    """(t, coefficients) of collision i (order vv, ev, ee, fv)"""
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    t = np.zeros(3)
    c = np.zeros(4)
    n = int(lib().ipco_distance_vector(C.byref(coll._s), pV, pE, pF, i, t.ctypes.data_as(_D),
                                       c.ctypes.data_as(_D)))
    return t, c[:n]


def gauss_newton_diagonal(coll, V, E, F, dhat, kappa):
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    g = np.zeros(3 * len(V))
    lib().ipco_gauss_newton_diagonal(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa,
                                     g.ctypes.data_as(_D))
    return g


def gauss_newton_quadratic_form(coll, V, E, F, dhat, kappa, p):
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    p, pp = _d(np.asarray(p, dtype=np.float64).reshape(-1))
    return lib().ipco_gauss_newton_quadratic_form(C.byref(coll._s), pV, len(V), pE, pF, dhat,
                                                  kappa, pp)


def _lists(ee, fv, vv, ev):
    z = np.zeros((0, 2), np.int32)
    return [_pairs_struct(z if a is None else a) for a in (ee, fv, vv, ev)]


def per_vertex_safe_distances(V, E, F, ee, fv, inflation_radius, min_distance=0.0, vv=None,
                              ev=None):
    """Candidates::compute_per_vertex_safe_distances (candidates.cpp:372-408)"""
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    L = _lists(ee, fv, vv, ev)
    out = np.zeros(len(V))
    lib().ipco_per_vertex_safe_distances(pV, len(V), pE, pF, *[C.byref(x[0]) for x in L],
                                         inflation_radius, min_distance,
                                         out.ctypes.data_as(_D))
    return out


def noncandidate_conservative_stepsize(V0, V1, E, F, ee, fv, dhat, vv=None, ev=None):
    """Candidates::compute_noncandidate_conservative_stepsize (candidates.cpp:294-333)
    with displacements V1 - V0"""
    V0, p0 = _d(V0)
    V1, p1 = _d(V1)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    L = _lists(ee, fv, vv, ev)
    return lib().ipco_noncandidate_conservative_stepsize(p0, p1, len(V0), pE, pF,
                                                         *[C.byref(x[0]) for x in L], dhat)


def cfl_stepsize(V0, V1, E, F, ee, fv, dhat, min_distance=0.0, params=DEFAULT_CCD, vv=None,
                 ev=None):
    """Candidates::compute_cfl_stepsize (candidates.cpp:335-360)"""
    a_c = candidates_stepsize(V0, V1, E, F, ee, fv, min_distance, params, vv=vv, ev=ev)
    a_f = noncandidate_conservative_stepsize(V0, V1, E, F, ee, fv, dhat, vv=vv, ev=ev)
    if a_f < 0.5 * a_c:
        return compute_collision_free_stepsize(V0, V1, E, F, min_distance, params)
    return min(a_c, a_f)


def is_step_collision_free(V0, V1, E, F, min_distance=0.0, params=DEFAULT_CCD):
    V0, p0 = _d(V0)
    V1, p1 = _d(V1)
    E, pE = _i(E)
    F, pF = _i(F)
    return bool(lib().ipco_is_step_collision_free(p0, p1, len(V0), pE, len(E), pF, len(F),
                                                  min_distance, C.byref(params)))


def _mesh_args(V, E, F):
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    F2E = faces_to_edges(E, F) if len(F) else np.zeros((0, 3), np.int32)
    F2E, pF2E = _i(F2E)
    return (V, E, F, F2E), (pV, len(V), pE, len(E), pF, len(F), pF2E)


def ogc_vertex_region(V, E, F, x, vi):
    """ogc::check_vertex_feasible_region(mesh, V, x, vi)"""
    keep, args = _mesh_args(V, E, F)
    x, px = _d(np.asarray(x, float).reshape(3))
    return bool(lib().ipco_ogc_vertex_region(*args, px, int(vi)))


def ogc_edge_region(V, E, F, xi, ei):
    """ogc::check_edge_feasible_region(mesh, V, xi, ei)"""
    keep, args = _mesh_args(V, E, F)
    return bool(lib().ipco_ogc_edge_region(*args, int(xi), int(ei)))


def ogc_feasible(V, E, F, kind, a, b):
    """ogc::is_edge_vertex_feasible ("ev": edge a, vertex b),
    is_edge_edge_feasible ("ee") or is_face_vertex_feasible ("fv": face a,
    vertex b) with the distance type classified"""
    keep, args = _mesh_args(V, E, F)
    k = {"ev": 1, "ee": 2, "fv": 3}[kind]
    return bool(lib().ipco_ogc_feasible(*args, k, int(a), int(b)))


# ---- friction ----------------------------------------------------------------
SMOOTH_MU_FN = {"mu": 0, "dmu": 1, "mu_f0": 2, "mu_f1": 3, "mu_f2": 4, "mu_f1_over_x": 5,
                "mu_f2_x_minus_f1_over_x3": 6, "f0": 7, "f1": 8, "f2": 9, "f1_over_x": 10,
                "f2_x_minus_f1_over_x3": 11}


def smooth_mu_fn(name, y, mu_s, mu_k, eps_v):
    """smooth_mu.cpp / smooth_friction_mollifier.cpp scalar functions"""
    return lib().ipco_smooth_mu_fn(SMOOTH_MU_FN[name], float(y), float(mu_s), float(mu_k),
                                   float(eps_v))


class Friction:
    """TangentialCollisions over a Collisions set (one record per collision,
    `keep` false for skipped near-parallel edge pairs)."""

    def __init__(self, s, coll, E, F):
        self._s, self._coll, self._E, self._F = s, coll, E, F
        n = s.n

        def arr(p, k, dt=np.float64):
            if n == 0:
                return np.zeros((0, k) if k > 1 else (0,), dt)
            a = np.ctypeslib.as_array(p, shape=(n * k,)).copy()
            return a.reshape(n, k) if k > 1 else a

        self.keep = arr(s.keep, 1, np.uint8).astype(bool)
        self.beta = arr(s.beta, 2)
        self.basis = arr(s.basis, 6).reshape(n, 2, 3).transpose(0, 2, 1)  # n x 3 x 2
        self.normal_force = arr(s.normal_force, 1)
        self.weight = arr(s.weight, 1)
        self.mu_s = arr(s.mu_s, 1)
        self.mu_k = arr(s.mu_k, 1)

    def __del__(self):
        try:
            lib().ipco_friction_free(C.byref(self._s))
        except Exception:
            pass


def friction_build(coll, V, E, F, dhat, kappa, mu_s, mu_k=None):
    """mu_s / mu_k: scalars, or per-vertex arrays (blended per collision)"""
    V, pV = _d(V)
    E, pE = _i(np.asarray(E).reshape(-1, 2))
    F, pF = _i(np.asarray(F).reshape(-1, 3))
    s = _Friction()
    mu_k = mu_s if mu_k is None else mu_k
    if np.ndim(mu_s) == 0 and np.ndim(mu_k) == 0:
        lib().ipco_friction_build(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa, mu_s, mu_k,
                                  C.byref(s))
    else:
        ms, pms = _d(np.broadcast_to(np.asarray(mu_s, float), (len(V),)))
        mk, pmk = _d(np.broadcast_to(np.asarray(mu_k, float), (len(V),)))
        lib().ipco_friction_build_pv(C.byref(coll._s), pV, len(V), pE, pF, dhat, kappa, pms, pmk,
                                     C.byref(s))
    return Friction(s, coll, E, F)


def _fargs(fr, Vel):
    Vel, pVel = _d(Vel)
    _, pE = _i(fr._E)
    _, pF = _i(fr._F)
    return Vel, (C.byref(fr._s), C.byref(fr._coll._s), pVel, len(Vel), pE, pF)


def friction_energy(fr, Vel, eps_v):
    keep, a = _fargs(fr, Vel)
    return lib().ipco_friction_energy(*a, eps_v)


def friction_gradient(fr, Vel, eps_v):
    keep, a = _fargs(fr, Vel)
    g = np.zeros(3 * len(keep))
    lib().ipco_friction_gradient(*a, eps_v, g.ctypes.data_as(_D))
    return g


def friction_hessian(fr, Vel, eps_v, psd="clamp"):
    import scipy.sparse as sp
    keep, a = _fargs(fr, Vel)
    m = _CSR()
    lib().ipco_friction_hessian(*a, eps_v, PSD[psd], C.byref(m))
    n, nnz = m.n, m.nnz
    rowptr = np.ctypeslib.as_array(m.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(m.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(m.val, shape=(max(nnz, 1),))[:nnz].copy()
    lib().ipco_csr_free(C.byref(m))
    return sp.csr_matrix((val, col, rowptr), shape=(n, n))


def friction_set_anisotropy(fr, mu_aniso=None, mu_s_aniso=None, mu_k_aniso=None):
    """per-collision (n, 2) arrays; None keeps the current values; resets the
    effective coefficients"""
    keep = []

    def arg(a):
        if a is None:
            return None
        a, p = _d(np.asarray(a, float).reshape(-1, 2))
        keep.append(a)
        return p
    lib().ipco_friction_set_anisotropy(C.byref(fr._s), arg(mu_aniso), arg(mu_s_aniso),
                                       arg(mu_k_aniso))


def friction_update_lagged_mu(fr, Vel):
    Vel, a = _fargs(fr, Vel)
    lib().ipco_friction_update_lagged_mu(a[0], a[1], a[2], a[4], a[5])


def friction_effective_mu(fr):
    n = fr._s.n
    if n == 0:
        return np.zeros(0), np.zeros(0)
    return (np.ctypeslib.as_array(fr._s.mu_s_eff, shape=(n,)).copy(),
            np.ctypeslib.as_array(fr._s.mu_k_eff, shape=(n,)).copy())


def anisotropic_mu_eff(tau_aniso, mu_s_aniso, mu_k_aniso, mu_s, mu_k):
    t, pt = _d(np.asarray(tau_aniso, float))
    a, pa = _d(np.asarray(mu_s_aniso, float))
    b, pb = _d(np.asarray(mu_k_aniso, float))
    out = np.zeros(2)
    lib().ipco_anisotropic_mu_eff(pt, pa, pb, mu_s, mu_k, out.ctypes.data_as(_D))
    return out
