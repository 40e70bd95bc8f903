/*
 * ipcg.h -- C ABI of the MI355X (gfx950) contact hot path of IPC Toolkit.
 *
 * One context owns one HIP stream and all device memory of one collision
 * mesh: topology, positions at t0 (slot 0) and t1 (slot 1), broad-phase
 * trees, candidate lists, the collision set, the gradient and the Hessian.
 * Every call returns an ipcg_status (0 = success) and never throws; the last
 * error message of a context is available from ipcg_last_error().  Calls on
 * one context must not overlap; distinct contexts are independent.
 *
 * Layouts: positions are float64 with IPCG_LAYOUT_AOS (row i = x,y,z, a
 * C-contiguous numpy (nV,3) array) or IPCG_LAYOUT_SOA (all x, then all y,
 * then all z: the column-major Eigen::MatrixXd the reference takes).
 * Edges (nE,2) and faces (nF,3) are int32 row-major.  Gradients are RowMajor
 * [x0,y0,z0,x1,...] (reference config.hpp.in:45).  Hessians are CSR with
 * int64 row pointers, int32 columns, float64 values; the matrix is symmetric,
 * so CSR equals the reference's column-major CSC.
 *
 * Citations name the reference interface (IPC Toolkit v1.6.0) each entry
 * point stands behind.
 */
#ifndef IPCG_H
#define IPCG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum ipcg_status {
    IPCG_OK = 0,
    IPCG_ERR_ARG = 1,       /* invalid argument (sizes, ids, layout)        */
    IPCG_ERR_HIP = 2,       /* HIP runtime error                            */
    IPCG_ERR_OOM = 3,       /* device allocation failed                     */
    IPCG_ERR_STATE = 4,     /* call out of order (e.g. no mesh uploaded)     */
    IPCG_ERR_NUMERIC = 5,   /* PSD eigensolver failed / CCD work overflow   */
    IPCG_ERR_NO_DEVICE = 6  /* no HIP device                                */
} ipcg_status;

enum { IPCG_LAYOUT_AOS = 0, IPCG_LAYOUT_SOA = 1 };
enum { IPCG_BP_LBVH = 0, IPCG_BP_SAP = 1 };
/* utils/eigen_ext.hpp:202-206 PSDProjectionMethod */
enum { IPCG_PSD_NONE = 0, IPCG_PSD_CLAMP = 1, IPCG_PSD_ABS = 2 };
/* candidate / collision kinds */
enum { IPCG_VV = 0, IPCG_EV = 1, IPCG_EE = 2, IPCG_FV = 3 };

typedef struct ipcg_ctx ipcg_ctx;

/* ---- context ------------------------------------------------------------ */
int ipcg_ctx_create(int device, ipcg_ctx** out);
int ipcg_ctx_destroy(ipcg_ctx* ctx);
const char* ipcg_last_error(const ipcg_ctx* ctx);
/* wait for all work of the context's stream */
int ipcg_ctx_synchronize(ipcg_ctx* ctx);
/* the context's hipStream_t (as void*), for callers that order their own
 * work against the library's */
void* ipcg_ctx_stream(ipcg_ctx* ctx);
/* bytes of device memory the context holds (its buffers only grow; the CCD
 * slab tier's lists, one buffer per device shared by its contexts, are not
 * counted) */
int ipcg_ctx_device_bytes(ipcg_ctx* ctx, int64_t* bytes);
/* Launch-shape autotuning (no reference counterpart): with enable != 0,
 * the first broad phase of each kind ("bp_traverse_ee.static",
 * "bp_traverse_fv.static", "bp_traverse_ee.swept", "bp_traverse_fv.swept")
 * times the LBVH traversal at 64, 128 and 256 lanes per workgroup on the
 * real input and keeps the fastest for the context; results do not depend
 * on the choice.  Off (the default, and after turning it off): 128.
 * ipcg_launch_shape reports a stage's chosen size (0: not tuned) and its
 * measured time in milliseconds (ms may be NULL). */
int ipcg_ctx_set_autotune(ipcg_ctx* ctx, int enable);
int ipcg_launch_shape(ipcg_ctx* ctx, const char* stage, int32_t* block, double* ms);

/* ---- collision mesh (collision_mesh.hpp:14-126, collision_filter.hpp) ---
 * rest_positions: nV x 3 in `layout`; used for the edge-edge mollifier
 * threshold (edge_edge_mollifier.cpp:194-202).  faces_to_edges is built
 * on upload (collision_mesh.cpp:510-541; IPCG_ERR_ARG if a face edge is
 * missing from `edges`).  vertex_patch may be NULL (CollisionMesh's default
 * filter: every vertex pair may collide) or nV patch ids, giving
 * can_collide(i,j) = patch[i] != patch[j] (make_vertex_patches_filter,
 * collision_filter.hpp:113-118). */
int ipcg_mesh_upload(ipcg_ctx* ctx, int32_t nV, const double* rest_positions,
                     int layout, int32_t nE, const int32_t* edges, int32_t nF,
                     const int32_t* faces, const int32_t* vertex_patch);

/* Further vertex-pair filters of CollisionMesh::can_collide
 * (collision_mesh.hpp:338), intersected with the patch filter
 * (CollisionFilter operator&, collision_filter.hpp:74-81): n_dynamic >= 0
 * adds make_static_obstacle_filter(n_dynamic) (collision_filter.hpp:126-
 * 131: i < n_dynamic || j < n_dynamic), n_codim_vertices >= 0 adds
 * make_codim_cross_filter (collision_filter.hpp:138-143: exactly one of i, j
 * below n_codim_vertices); -1 turns a filter off.
 * make_connected_components_filter (collision_filter.hpp:152) is the patch
 * filter with per-vertex component labels.  A new mesh upload
 * turns both off; a change invalidates candidates and collisions. */
int ipcg_mesh_filter(ipcg_ctx* ctx, int32_t n_dynamic, int32_t n_codim_vertices);

/* ---- positions ----------------------------------------------------------
 * slot 0: the configuration x (static queries, potentials, CCD start);
 * slot 1: the CCD end configuration.  _device takes a device pointer
 * (same layout rules) and copies on the context stream. */
int ipcg_positions_upload(ipcg_ctx* ctx, int slot, const double* V, int layout);
int ipcg_positions_upload_device(ipcg_ctx* ctx, int slot, const double* dV,
                                 int layout);

/* ---- broad phase (broad_phase.hpp:19-133, lbvh.cpp, sweep_and_prune.cpp) -
 * swept = 0: boxes of slot 0 inflated by `inflation`
 *            (BroadPhase::build(vertices, ...), Candidates::build static);
 * swept = 1: boxes of the segment slot0 -> slot1
 *            (BroadPhase::build(vertices_t0, vertices_t1, ...)).
 * method IPCG_BP_LBVH (default, lbvh.cpp) or IPCG_BP_SAP (sweep and prune).
 * Produces the 3D candidate lists (broad_phase.cpp:79-91): edge-edge pairs,
 * each unordered pair once as (lower Morton position, higher), and
 * face-vertex pairs (face, vertex). */
int ipcg_broad_phase(ipcg_ctx* ctx, int swept, double inflation, int method);
int ipcg_candidates_count(ipcg_ctx* ctx, int64_t* n_ee, int64_t* n_fv);
/* ee: 2*n_ee int32, fv: 2*n_fv int32 (either may be NULL) */
int ipcg_candidates_download(ipcg_ctx* ctx, int32_t* ee, int32_t* fv);

/* ---- codim candidates (candidates.cpp:65-118) ----------------------------
 * Meshes with free vertices (in no edge) or free edges (in no face,
 * collision_mesh.cpp:145-182) get two more lists from every broad phase
 * call: codim vertex pairs (each unordered pair once) and (codim edge, codim
 * vertex) pairs, FP32-box (LBVH) semantics with either method.  The
 * collision build and the CCD entry points use them together with the
 * edge-edge and face-vertex lists. */
int ipcg_mesh_codim(ipcg_ctx* ctx, int32_t* n_codim_vertices, int32_t* n_codim_edges);
int ipcg_codim_candidates_count(ipcg_ctx* ctx, int64_t* n_vv, int64_t* n_ev);
/* vv: 2*n_vv int32, ev: 2*n_ev int32 (either may be NULL) */
int ipcg_codim_candidates_download(ipcg_ctx* ctx, int32_t* vv, int32_t* ev);

/* ---- edge-face candidates and intersections ------------------------------
 * BroadPhase::detect_edge_face_candidates (broad_phase.hpp:91-92,
 * lbvh.cpp:771-784, filter can_edge_face_collide lbvh.cpp:836-849) after a
 * static build over positions slot 0 with the given inflation radius:
 * (edge, face) pairs whose FP32 boxes overlap, sharing no vertex, passing
 * the mesh filter.  Rebuilds the edge and face trees; the edge-edge and
 * face-vertex lists of an earlier broad phase call stay valid. */
int ipcg_edge_face_candidates(ipcg_ctx* ctx, double inflation, int64_t* n_ef);
/* ef: 2*n_ef int32 */
int ipcg_edge_face_candidates_download(ipcg_ctx* ctx, int32_t* ef);
/* has_intersections (ipc.cpp:105-166, 3D) over positions slot 0: edge-face
 * candidates at 1e-6 x the world bounding-box diagonal
 * (world_bbox_diagonal_length.hpp:11-15), each tested with
 * is_edge_intersecting_triangle (intersection.cpp:116-138, the default
 * non-rational build: exact orient3d, then Eigen's FullPivLU solve).
 * *intersecting = 1 when some edge crosses some triangle.  The answer
 * does not depend on the broad phase method (every method reports a
 * superset of the overlapping pairs), so none is taken. */
int ipcg_has_intersections(ipcg_ctx* ctx, int* intersecting);

/* ---- collision set (normal_collisions.cpp:20-158) ------------------------
 * build: static broad phase with inflation 0.5*(dhat+dmin) followed by the
 * narrow phase (NormalCollisions::build(mesh, V, dhat, dmin, broad_phase)).
 * build_from_candidates: narrow phase over the current static candidate
 * lists (NormalCollisions::build(candidates, mesh, V, dhat, dmin)). */
int ipcg_collisions_build(ipcg_ctx* ctx, double dhat, double dmin, int method);
int ipcg_collisions_build_from_candidates(ipcg_ctx* ctx, double dhat, double dmin);
/* NormalCollisions::set_use_area_weighting (normal_collisions.hpp,
 * normal_collisions_builder.cpp:26-336) for the builds that follow: each
 * candidate's contribution weighs 0.5 (a_i + a_j) (codim vertex pair),
 * 0.5 a_v (codim edge-vertex), 0.25 (A_ea + A_eb) (edge-edge) or 0.25 a_v
 * (face-vertex) instead of 1, summed over duplicates, with the mesh's
 * vertex areas a and edge areas A (CollisionMesh::init_areas,
 * collision_mesh.cpp:309-376, over the rest positions).  Default off. */
int ipcg_collisions_set_area_weighting(ipcg_ctx* ctx, int use_area_weighting);
/* NormalCollisions::set_collision_set_type (normal_collisions.cpp:175-191;
 * CollisionSetType in normal_collisions.hpp) for the builds that follow.
 * IPCG_COLLISIONS_IMPROVED_MAX_APPROX (normal_collisions.cpp:84-128) adds,
 * from the candidates converted to sub-element pairs (candidates.cpp:
 * 583-692), signed vertex-vertex, edge-vertex and mollified edge-edge
 * contributions (normal_collisions_builder.cpp:340-543) weighted through
 * the mesh adjacencies (collision_mesh.cpp:247-291).  Duplicates merge by
 * (vv, ev) pair or (ee) edge pair and distance type, summing weights in
 * ascending order of value; an edge-edge collision keeps the orientation
 * and eps_x of its candidate's contribution, else of the contribution that
 * lists the lower edge id first; vertex-vertex, edge-vertex and edge-edge
 * collisions whose weights sum to zero are dropped
 * (normal_collisions_builder.cpp:660-684).  The reference warns that the
 * approximation is meant to be used with area weighting.
 * IPCG_COLLISIONS_OGC drops edge-vertex, edge-edge and face-vertex pairs
 * whose closest point lies outside the feasible region of the feature it
 * reduces to (ogc/feasible_region.cpp) and divides a pair reduced to a
 * vertex or an edge among the vertex's incident edges (codim edge-vertex)
 * or faces, or the edge's opposite vertices (face-vertex)
 * (normal_collisions_builder.cpp:56-336 with use_ogc).  Unknown types:
 * IPCG_ERR_ARG.  Default IPCG_COLLISIONS_IPC. */
enum {
    IPCG_COLLISIONS_IPC = 0,
    IPCG_COLLISIONS_IMPROVED_MAX_APPROX = 1,
    IPCG_COLLISIONS_OGC = 2
};
int ipcg_collisions_set_type(ipcg_ctx* ctx, int collision_set_type);
/* counts[4] = {vv, ev, ee, fv} */
int ipcg_collisions_count(ipcg_ctx* ctx, int64_t counts[4]);
/* Canonical order: vv by (min,max); ev by (edge,vertex); ee by
 * (min edge, max edge, distance type) with ids as oriented by the candidate; fv by
 * (face,vertex).  Any pointer may be NULL. */
int ipcg_collisions_download(ipcg_ctx* ctx, int32_t* vv, double* vv_w,
                             int32_t* ev, double* ev_w, int32_t* ee,
                             double* ee_w, double* ee_eps_x, int32_t* ee_dtype,
                             int32_t* fv, double* fv_w);

/* ---- barrier potential (potential.cpp:36-222, barrier_potential.cpp) ----
 * Over the current collision set at slot-0 positions. */
/* BarrierPotential use_physical_barrier (barrier_potential.cpp:33-99) for
 * the calls that follow (energy, gradient, Hessian, Gauss-Newton helpers,
 * the friction build's normal forces): the barrier and its derivatives
 * scaled by dhat / ((2 dmin + dhat) dhat)^2 (barrier.hpp:133-137).
 * Default off. */
int ipcg_barrier_set_physical(ipcg_ctx* ctx, int use_physical_barrier);
int ipcg_barrier_energy(ipcg_ctx* ctx, double dhat, double kappa, double* out);
/* Allocation-free variant of ipcg_barrier_energy.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops).  Same kernels and the same value as
 * ipcg_barrier_energy, but the block partial sums go to a file-scope device
 * array of 2^20 doubles (8 MiB on every GPU the library runs on, held until
 * the process ends) and the energy is written to energy_device (caller-owned
 * device memory) on the context stream, without synchronizing: it is there
 * once the stream reaches it (ipcg_ctx_synchronize).  Limit: the collision
 * set may span at most 2^20 blocks of 128 collisions per kind
 * (IPCG_ERR_ARG beyond).  Calls that overlap on one GPU share the array
 * and race. */
int ipcg_barrier_energy_noalloc(ipcg_ctx* ctx, double dhat, double kappa,
                                double* energy_device);
/* grad: host array of 3*nV doubles (RowMajor), or NULL to keep it on device */
int ipcg_barrier_gradient(ipcg_ctx* ctx, double dhat, double kappa, double* grad);
/* assembles the CSR on device; *nnz receives the entry count */
int ipcg_barrier_hessian(ipcg_ctx* ctx, double dhat, double kappa, int psd,
                         int64_t* nnz);
/* rowptr: 3nV+1 int64, col: nnz int32, val: nnz float64 */
int ipcg_hessian_download(ipcg_ctx* ctx, int64_t* rowptr, int32_t* col, double* val);
/* device pointers of the last gradient / CSR (valid until the next call) */
int ipcg_device_results(ipcg_ctx* ctx, const double** grad, const int64_t** rowptr,
                        const int32_t** col, const double** val);
/* NormalPotential::gauss_newton_hessian_diagonal (normal_potential.cpp:
 * 15-46, per collision :236-270): per collision the distance vector
 * t = sum_i c_i x_i of its stored distance type (closest-point
 * coefficients), w (4 f''(|t|^2) c_i^2 t_k^2 + 2 f'(|t|^2) c_i^2) summed
 * per vertex coordinate.  diag: host array of 3*nV doubles (RowMajor). */
int ipcg_barrier_gauss_newton_diagonal(ipcg_ctx* ctx, double dhat, double kappa,
                                       double* diag);
/* NormalPotential::gauss_newton_hessian_quadratic_form (normal_potential.cpp:
 * 48-79, :272-308): sum over collisions of
 * w (4 f'' (q.t)^2 + 2 f' |q|^2), q = sum_i c_i p_i.  p: host array of
 * 3*nV doubles (RowMajor). */
int ipcg_barrier_gauss_newton_quadratic_form(ipcg_ctx* ctx, double dhat, double kappa,
                                             const double* p, double* value);

/* ---- friction (tangential_collisions.cpp, tangential_potential.cpp) ------
 * TangentialCollisions::build(mesh, V, collisions, BarrierPotential(dhat,
 * kappa), mu_s, mu_k) with the same coefficients at every vertex
 * (tangential_collisions.cpp:61-186; every blend of equal coefficients
 * returns them): over
 * the current collision set at the slot-0 positions (the lagged
 * positions), one record per collision: closest-point coordinates
 * (closest_point.cpp), tangent basis (tangent_basis.cpp), normal force
 * magnitude -kappa b'(d^2) 2 d (barrier_force_magnitude.cpp:5-15) times the
 * collision's weight; near-parallel edge-edge collisions (cross product
 * below eps_x) are not kept (tangential_collisions.cpp:118-121).  A new
 * collision build or mesh upload discards the friction set. */
int ipcg_friction_build(ipcg_ctx* ctx, double dhat, double kappa, double mu_s, double mu_k);
/* the same with per-vertex coefficients (host arrays of nV doubles each):
 * each collision takes the average (default_blend_mu,
 * tangential_collisions.hpp:143-149) of its two sides' coefficients at
 * their closest points (tangential_collisions.cpp:86-166) */
int ipcg_friction_build_per_vertex(ipcg_ctx* ctx, double dhat, double kappa, const double* mu_s,
                                   const double* mu_k);
/* Anisotropic friction (tangential_collision.hpp:108-139): per collision in
 * download order, 2 values each: mu_aniso scales the tangential velocity
 * (u_aniso = mu_aniso * u, default 1), mu_s_aniso / mu_k_aniso give
 * elliptical coefficients (default 0: isotropic).  Any pointer may be NULL
 * to keep the current values; the effective coefficients reset to mu_s,
 * mu_k (reset_lagged_anisotropic_friction_coefficients). */
int ipcg_friction_set_anisotropy(ipcg_ctx* ctx, const double* mu_aniso, const double* mu_s_aniso,
                                 const double* mu_k_aniso);
/* update_lagged_anisotropic_friction_coefficients (tangential_collisions.cpp:
 * 19-57) at the uploaded velocities, with the lagged positions the set was
 * built at: tau = T^T v, the effective coefficients sqrt((mu0 t0)^2 +
 * (mu1 t1)^2) along t = tau_aniso / |tau_aniso| (smooth_mu.cpp:115-171)
 * where mu_s_aniso is nonzero */
int ipcg_friction_update_lagged_mu(ipcg_ctx* ctx);
/* per collision in download order (vv, ev, ee, fv): keep flag,
 * closest-point coordinates (2), tangent basis (3 x 2, column major),
 * normal force magnitude (0 when not kept); any pointer may be NULL */
int ipcg_friction_download(ipcg_ctx* ctx, int32_t* keep, double* closest_point,
                           double* tangent_basis, double* normal_force);
/* velocities for the friction potential (nV x 3, layout as positions) */
int ipcg_velocities_upload(ipcg_ctx* ctx, const double* V, int layout);
/* FrictionPotential(eps_v) (friction_potential.cpp, smooth_mu.cpp,
 * smooth_friction_mollifier.cpp) at the uploaded velocities
 * (tangential_potential.cpp:156-331) with the lagged effective
 * coefficients and the velocity scaling above: energy, gradient (host
 * array of 3*nV doubles, or NULL to keep it on device) and Hessian with
 * the reference's projection of the inner 2 x 2 matrix (psd as for the
 * barrier; CSR read with ipcg_hessian_download). */
int ipcg_friction_energy(ipcg_ctx* ctx, double eps_v, double* out);
int ipcg_friction_gradient(ipcg_ctx* ctx, double eps_v, double* grad);
int ipcg_friction_hessian(ipcg_ctx* ctx, double eps_v, int psd, int64_t* nnz);

/* ---- CCD (ipc.cpp:45-101, candidates.cpp:252-292, tight_inclusion_ccd) --
 * compute_collision_free_stepsize: swept broad phase (inflation
 * 0.5*min_distance) + Tight-Inclusion CCD of every candidate, slot 0 ->
 * slot 1.  Every candidate is solved on [0,1] and outcomes at or above the
 * running minimum are left unexplored, which yields the minimum of the
 * per-candidate times of impact independently of scheduling. */
int ipcg_compute_collision_free_stepsize(ipcg_ctx* ctx, double min_distance,
                                         double tolerance, int64_t max_iterations,
                                         double conservative_rescaling, int method,
                                         double* step);
/* same narrow phase over the current candidate lists, however they were
 * built (Candidates::compute_collision_free_stepsize) */
int ipcg_candidates_stepsize(ipcg_ctx* ctx, double min_distance, double tolerance,
                             int64_t max_iterations, double conservative_rescaling,
                             double* step);
/* Candidates::is_step_collision_free (candidates.cpp:224-250) over the
 * current swept lists: 1 if no candidate reports a hit */
int ipcg_candidates_is_step_collision_free(ipcg_ctx* ctx, double min_distance,
                                           double tolerance, int64_t max_iterations,
                                           double conservative_rescaling,
                                           int* collision_free);
/* is_step_collision_free (ipc.cpp:20-41): 1 if no candidate reports a hit */
int ipcg_is_step_collision_free(ipcg_ctx* ctx, double min_distance, double tolerance,
                                int64_t max_iterations, double conservative_rescaling,
                                int method, int* collision_free);
/* Candidates::compute_per_vertex_safe_distances (candidates.cpp:372-408)
 * over the current lists at slot-0 positions: per vertex, the smallest
 * sqrt(distance) - min_distance over the candidates it belongs to (distance
 * type classified per candidate), at most inflation_radius - min_distance.
 * out: nV doubles. */
int ipcg_candidates_per_vertex_safe_distances(ipcg_ctx* ctx, double inflation_radius,
                                              double min_distance, double* out);
/* Candidates::compute_noncandidate_conservative_stepsize (candidates.cpp:
 * 294-333) with displacements slot 1 - slot 0: 0.5 dhat / the largest
 * displacement norm of a vertex in some current candidate (1 when there is
 * no candidate). */
int ipcg_candidates_noncandidate_conservative_stepsize(ipcg_ctx* ctx, double dhat,
                                                       double* step);
/* Candidates::compute_cfl_stepsize (candidates.cpp:335-360): the CCD step of
 * the current lists and the noncandidate step; when the latter is below
 * half the former, the full compute_collision_free_stepsize (swept broad
 * phase with `method`, which replaces the current lists). */
int ipcg_candidates_cfl_stepsize(ipcg_ctx* ctx, double dhat, double min_distance,
                                 double tolerance, int64_t max_iterations,
                                 double conservative_rescaling, int method, double* step);
/* per-candidate times of impact of the current swept lists, each on
 * [0,1] with no bound (+inf when no hit); toi_ee: n_ee, toi_fv: n_fv */
int ipcg_candidates_toi(ipcg_ctx* ctx, double min_distance, double tolerance,
                        int64_t max_iterations, double conservative_rescaling,
                        double* toi_ee, double* toi_fv);

/* the same for the codim lists (point-point and point-edge CCD,
 * tight_inclusion_ccd.cpp:77-200); toi_vv: n_vv, toi_ev: n_ev */
int ipcg_codim_candidates_toi(ipcg_ctx* ctx, double min_distance, double tolerance,
                              int64_t max_iterations, double conservative_rescaling,
                              double* toi_vv, double* toi_ev);

/* ---- batched scenes -------------------------------------------------------
 * Many independent scenes in one context, each stage one launch over all of
 * them.  No reference counterpart: each scene is what one CollisionMesh /
 * NormalCollisions / BarrierPotential / compute_collision_free_stepsize
 * call sequence (normal_collisions.cpp:20-158, potential.cpp:36-222,
 * ipc.cpp:45-101) computes for that scene alone, and every per-scene result
 * equals the one a context holding the scene alone returns.
 *
 * Scene s owns vertices [vertex_offsets[s], vertex_offsets[s+1]), edges
 * [edge_offsets[s], ...) and faces [face_offsets[s], ...) of the
 * concatenated arrays (n_scenes + 1 offsets each, starting at 0); edges and
 * faces hold scene-local vertex ids.  Positions (both slots) are the
 * concatenated nV x 3 arrays.  Pairs of different scenes never collide: the
 * broad phase builds one LBVH per scene (each scene's Morton codes mapped
 * by its own box) and every query walks its own scene's tree.  The
 * ordinary entry points then work on the batch: collision lists,
 * candidates and the gradient use global ids (scene s's gradient is rows
 * [3 vertex_offsets[s], 3 vertex_offsets[s+1])), the Hessian is block
 * diagonal with one block per scene.  LBVH broad phase only
 * (IPCG_BP_SAP: IPCG_ERR_ARG). */
int ipcg_mesh_upload_batched(ipcg_ctx* ctx, int32_t n_scenes, const int64_t* vertex_offsets,
                             const double* rest_positions, int layout,
                             const int64_t* edge_offsets, const int32_t* edges,
                             const int64_t* face_offsets, const int32_t* faces);
/* number of scenes of a batched context (0: an ordinary context) */
int ipcg_batch_size(ipcg_ctx* ctx, int32_t* n_scenes);
/* active[n_scenes] (NULL: all): the broad phase skips the queries of
 * inactive scenes, so the calls that follow leave them without candidates
 * and collisions (energy 0, step size 1); the results of active scenes do
 * not depend on which others are active.  Invalidates the candidate and
 * collision sets. */
int ipcg_batch_set_active(ipcg_ctx* ctx, const uint8_t* active);
/* counts[4 * n_scenes]: {vv, ev, ee, fv} of every scene of the current
 * collision set */
int ipcg_collisions_count_batched(ipcg_ctx* ctx, int64_t* counts);
/* energies[n_scenes]: each scene's barrier energy, summed in the order a
 * context holding the scene alone uses (the same value bit for bit) */
int ipcg_barrier_energy_batched(ipcg_ctx* ctx, double dhat, double kappa, double* energies);
/* steps[n_scenes]: compute_collision_free_stepsize of every scene (one
 * running minimum per scene) */
int ipcg_compute_collision_free_stepsize_batched(ipcg_ctx* ctx, double min_distance,
                                                 double tolerance, int64_t max_iterations,
                                                 double conservative_rescaling, double* steps);

/* ---- serving --------------------------------------------------------------
 * No reference counterpart: a request loop for callers that serve many
 * simulations at once.  Simulated scenes send energy, gradient, Hessian and
 * step requests with arrival times (seconds after ipcg_server_run starts).
 * The server keeps every admitted scene's topology and positions resident
 * on the device in one batched context and serves, per iteration, the
 * earliest arrived request of every admitted scene (continuous batching:
 * requests that arrive while an iteration runs join the next one).  Energy,
 * gradient and Hessian requests share one collision build and evaluation
 * (BarrierPotential with PSD clamp, inflation 0.5 dhat) over the scenes that
 * asked; step requests share one CCD: the scene's positions are perturbed by
 * N(0, sigma^2) per coordinate (drawn on the device from a counter-based
 * generator keyed by the scene's seed, its step count and the coordinate,
 * so a draw depends on nothing else) and the
 * perturbed positions replace the current ones only when
 * compute_collision_free_stepsize returns 1 (min_distance 0, tolerance 1e-6,
 * 1e7 iterations, rescaling 0.8).  A scene's results do not depend on the
 * other scenes it shares launches with: each equals the result of serving
 * the scene alone.  Admission: scenes enter the batch in submission order
 * while the sum of their estimated device footprints stays within
 * hbm_budget_bytes; when the admitted scenes have no request left, the
 * batch is rebuilt with the scenes still waiting.  Every scene of a server
 * shares dhat and kappa.  keep_results != 0 keeps every gradient and the
 * Hessian block of every Hessian request on the host (for checks). */
typedef struct ipcg_server ipcg_server;
enum { IPCG_REQ_ENERGY = 0, IPCG_REQ_GRADIENT = 1, IPCG_REQ_HESSIAN = 2, IPCG_REQ_STEP = 3 };
typedef struct ipcg_request_result {
    double arrival, start, finish; /* seconds after the run started */
    double value;     /* energy | gradient 2-norm | Hessian block nonzeros | step size */
    int64_t pairs;    /* collisions of the scene (energy, gradient, Hessian requests) */
    int32_t accepted; /* step requests: the perturbation was applied */
    int32_t iteration;
} ipcg_request_result;
typedef struct ipcg_server_stats {
    int64_t iterations, epochs, requests;
    int64_t peak_device_bytes; /* the batched context's device memory, largest seen */
    int64_t estimated_bytes;   /* sum of the admitted scenes' estimates, largest seen */
    double busy_seconds, wall_seconds;
} ipcg_server_stats;
int ipcg_server_create(int device, int64_t hbm_budget_bytes, int keep_results,
                       ipcg_server** out);
int ipcg_server_destroy(ipcg_server* srv);
const char* ipcg_server_last_error(const ipcg_server* srv);
/* a scene: nV x 3 positions (AoS, also its rest positions), scene-local
 * edges and faces, barrier parameters, step perturbation sigma and seed */
int ipcg_server_add_scene(ipcg_server* srv, int32_t nV, const double* V, int32_t nE,
                          const int32_t* E, int32_t nF, const int32_t* F, double dhat,
                          double kappa, double sigma, uint64_t seed, int32_t* scene);
int ipcg_server_submit(ipcg_server* srv, int32_t scene, int kind, double arrival,
                       int64_t* ticket);
/* serves every submitted request; returns when all are done */
int ipcg_server_run(ipcg_server* srv);
int ipcg_server_result(ipcg_server* srv, int64_t ticket, ipcg_request_result* out);
/* kept results: the scene's gradient (3 nV), its Hessian block (CSR with
 * 3 nV + 1 row pointers; nnz first, any pointer may be NULL) */
int ipcg_server_gradient(ipcg_server* srv, int64_t ticket, double* grad);
int ipcg_server_hessian(ipcg_server* srv, int64_t ticket, int64_t* nnz, int64_t* rowptr,
                        int32_t* col, double* val);
/* a scene's current positions (nV x 3) */
int ipcg_server_positions(ipcg_server* srv, int32_t scene, double* V);
int ipcg_server_get_stats(ipcg_server* srv, ipcg_server_stats* out);
/* the server's batched context (stage times, counters; do not call other
 * entry points on it while the server runs) */
ipcg_ctx* ipcg_server_context(ipcg_server* srv);

/* ---- timing of the last call's kernels (for benchmarks) ----------------- */
/* Kernel time of named stages recorded with HIP events on the context
 * stream since the last reset; names are stage names, ms in milliseconds.
 * Returns the number of stages (up to max). */
int ipcg_stage_times(ipcg_ctx* ctx, int max, const char** names, double* ms,
                     int64_t* calls);
int ipcg_stage_times_reset(ipcg_ctx* ctx, int enable);
/* work counts of the last calls, by name: "ccd.candidates" (swept
 * candidates of the last CCD), "ccd.searched" (those whose root box
 * includes the origin), "ccd.spilled" (searches that outgrew the LDS
 * lists), "ccd.wavefront" (searches that outgrew a lane's HBM slab),
 * "hessian.ee_other" (edge-edge collisions of the last Hessian whose type
 * is not the interior one; their projection takes the 9 x 9 path),
 * "hessian.pattern_sorted" (1 when a row of the last Hessian pattern had
 * more than 256 distinct block columns and the pattern was built by a
 * sort of all stencil pairs instead of the per-row hash),
 * "intersection.candidates" (edge-face candidates of the last
 * has_intersections), "intersection.deferred" (their orient3d calls the
 * floating-point filter left to the exact evaluation) */
int ipcg_counter(ipcg_ctx* ctx, const char* name, int64_t* value);

#ifdef __cplusplus
}
#endif
#endif
