/*
 * ipc_oracle.h -- CPU restatement of the IPC Toolkit contact hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation and the timed CPU baseline.  Only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it; the
 * product path (libipcg.so and its C++/Python front ends) never does.
 *
 * It restates, in plain C99 + OpenMP, the reference files (IPC Toolkit
 * v1.6.0) listed beside each declaration.  Arithmetic that decides discrete
 * outcomes (box bounds, distance types, squared distances, activity tests,
 * the mollifier threshold) follows the reference operation by operation,
 * with Eigen's fixed-size evaluation order made explicit:
 *   dot(a,b)     = a0*b0 + (a1*b1 + a2*b2)   (Eigen unrolled redux halves)
 *   cross(a,b)   = (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
 *   normalized() = n / sqrt(n.squaredNorm())
 *   2x2 LDLT     = Eigen's diagonal-pivoting LDLT, zero pivot -> 0 solution
 * and is compiled with -ffp-contract=off.
 *
 * Derivatives (gradients, Hessians) are computed with forward-mode
 * second-order automatic differentiation (hyper-dual numbers), an approach
 * independent of both the reference's generated code and the HIP kernels'
 * hand-derived closed forms.
 *
 * Tight-Inclusion CCD (third party, Tight-Inclusion v1.0.6, not vendored in
 * the reference) is restated from its published algorithm: Wang, Ferguson,
 * Schneider, Panozzo, "A Large Scale Benchmark and an Inclusion-Based
 * Algorithm for Continuous Collision Detection", ACM TOG 40(5), 2021.
 *
 * Layout conventions of this C API: positions are AoS (row i = x,y,z of
 * vertex i, i.e. a C-contiguous numpy (nV,3) array); edges (nE,2) and faces
 * (nF,3) int32 row-major.
 */
#ifndef IPC_ORACLE_H
#define IPC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- enums (values equal the reference's enum values) ------------------ */
/* distance/distance_type.hpp */
enum { IPCO_PE_P_E0 = 0, IPCO_PE_P_E1 = 1, IPCO_PE_P_E = 2 };
enum {
    IPCO_PT_P_T0 = 0, IPCO_PT_P_T1, IPCO_PT_P_T2, IPCO_PT_P_E0,
    IPCO_PT_P_E1, IPCO_PT_P_E2, IPCO_PT_P_T
};
enum {
    IPCO_EE_EA0_EB0 = 0, IPCO_EE_EA0_EB1, IPCO_EE_EA1_EB0, IPCO_EE_EA1_EB1,
    IPCO_EE_EA_EB0, IPCO_EE_EA_EB1, IPCO_EE_EA0_EB, IPCO_EE_EA1_EB,
    IPCO_EE_EA_EB
};
/* utils/eigen_ext.hpp:202-206 */
enum { IPCO_PSD_NONE = 0, IPCO_PSD_CLAMP = 1, IPCO_PSD_ABS = 2 };

/* ---- primitives ---------------------------------------------------------- */
/* distance_type.cpp:10-34 */
int ipco_point_edge_distance_type(const double p[3], const double e0[3],
                                  const double e1[3]);
/* distance_type.cpp:36-79 */
int ipco_point_triangle_distance_type(const double p[3], const double t0[3],
                                      const double t1[3], const double t2[3]);
/* distance_type.cpp:82-174 */
int ipco_edge_edge_distance_type(const double ea0[3], const double ea1[3],
                                 const double eb0[3], const double eb1[3]);

/* squared distances; dtype < 0 means AUTO
 * point_point.cpp:5-9, point_line.cpp:5-23, point_plane.cpp:10-26,
 * line_line.cpp:7-16, point_edge.cpp:10-37, point_triangle.cpp:11-48,
 * edge_edge.cpp:13-56 */
double ipco_point_point_distance(const double p0[3], const double p1[3]);
double ipco_point_line_distance(const double p[3], const double e0[3],
                                const double e1[3]);
double ipco_point_plane_distance(const double p[3], const double t0[3],
                                 const double t1[3], const double t2[3]);
double ipco_line_line_distance(const double ea0[3], const double ea1[3],
                               const double eb0[3], const double eb1[3]);
double ipco_point_edge_distance(const double p[3], const double e0[3],
                                const double e1[3], int dtype);
double ipco_point_triangle_distance(const double p[3], const double t0[3],
                                    const double t1[3], const double t2[3],
                                    int dtype);
double ipco_edge_edge_distance(const double ea0[3], const double ea1[3],
                               const double eb0[3], const double eb1[3],
                               int dtype);

/* Gradient (12) and Hessian (12x12 row-major) of the squared distance of a
 * stencil with the given dtype.  kind: 0 = point-triangle (p,t0,t1,t2),
 * 1 = edge-edge (ea0,ea1,eb0,eb1), 2 = point-edge (p,e0,e1) [9 dofs],
 * 3 = point-point (p0,p1) [6 dofs].  x holds the stencil dofs
 * (3 * n_vertices doubles).  grad/hess may be NULL.  Returns the value. */
double ipco_stencil_distance_derivs(int kind, int dtype, const double *x,
                                    double *grad, double *hess);

/* edge_edge_mollifier.cpp:7-14, 194-202, 43-67 (scalar m and derivatives
 * in x = cross squared norm), 83-136 (m(x) with gradient/Hessian wrt the 12
 * edge-edge dofs; grad/hess may be NULL). */
double ipco_edge_edge_cross_squarednorm(const double ea0[3],
                                        const double ea1[3],
                                        const double eb0[3],
                                        const double eb1[3]);
double ipco_edge_edge_mollifier_threshold(const double ea0[3],
                                          const double ea1[3],
                                          const double eb0[3],
                                          const double eb1[3]);
double ipco_edge_edge_mollifier(const double *x12, double eps_x, double *grad,
                                double *hess);

/* barrier.cpp:11-43 */
double ipco_barrier(double d, double dhat);
double ipco_barrier_first_derivative(double d, double dhat);
double ipco_barrier_second_derivative(double d, double dhat);

/* eigen_ext.tpp:63-108: in-place PSD projection of an n x n symmetric
 * row-major matrix (n <= 12).  Returns 0, or -1 if the eigensolver failed. */
int ipco_project_to_psd(double *A, int n, int method);
/* Eigenvalues (ascending) of an n x n symmetric matrix (Jacobi). */
int ipco_symmetric_eigenvalues(const double *A, int n, double *lambda);

/* ---- mesh helpers ---------------------------------------------------------*/
/* collision_mesh.cpp:510-541: faces_to_edges (nF x 3), edge j of face f is
 * (f[j], f[(j+1)%3]).  Returns 0 or -1 if an edge is missing. */
int ipco_faces_to_edges(const int32_t *E, int nE, const int32_t *F, int nF,
                        int32_t *F2E);

/* ---- broad phase ----------------------------------------------------------*/
/* aabb.cpp:35-55,57-126 + lbvh.cpp:29-41: FP32 outward-rounded boxes of
 * vertices (static when V1 == NULL, swept otherwise), edges and faces.
 * Each box is 6 floats (min xyz, max xyz). Any output may be NULL. */
void ipco_build_boxes_f32(const double *V0, const double *V1, int nV,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          double inflation, float *vbox, float *ebox,
                          float *fbox);

typedef struct ipco_pairs {
    int64_t n;
    int32_t *ij; /* 2*n: (a0,b0,a1,b1,...) */
} ipco_pairs;

/* candidates.cpp:43-222 with the LBVH of lbvh.cpp:53-873: edge-edge pairs
 * (triangular, each unordered pair once, query edge first) and
 * face-vertex pairs (face, vertex).  V1 == NULL: static boxes.
 * vertex_group may be NULL (default filter: all vertex pairs may collide),
 * else can_collide(i,j) = group[i] != group[j]
 * (collision_filter.hpp:113-118). */
int ipco_candidates(const double *V0, const double *V1, int nV,
                    const int32_t *E, int nE, const int32_t *F, int nF,
                    double inflation, const int32_t *vertex_group,
                    ipco_pairs *ee, ipco_pairs *fv);
/* candidates.cpp:65-118: codim vertex-vertex pairs (each unordered pair
 * once) and (codim edge, codim vertex) pairs, LBVH semantics (FP32 boxes).
 * F2E: faces_to_edges (ipco_faces_to_edges).  ipco_codim_lists fills the
 * codim vertex / edge ids (either may be NULL) and counts[2]. */
int ipco_codim_lists(int nV, const int32_t *E, int nE, const int32_t *F2E,
                     int nF, int32_t *cv, int32_t *ce, int32_t *counts);
int ipco_codim_candidates(const double *V0, const double *V1, int nV,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          const int32_t *F2E, double inflation,
                          const int32_t *vertex_group, ipco_pairs *vv,
                          ipco_pairs *ev);

/* Sweep and prune over FP64 boxes (sweep_and_prune.cpp:22-249; sort and
 * sweep restated from scalable-ccd): same filters, edge-edge pairs as
 * (earlier, later) in (min along the sweep axis, id) order.  The sweep axis
 * is the largest extent of the mesh box (ipco_sap_axis). */
int ipco_candidates_sap(const double *V0, const double *V1, int nV,
                        const int32_t *E, int nE, const int32_t *F, int nF,
                        double inflation, const int32_t *vertex_group,
                        ipco_pairs *ee, ipco_pairs *fv);
int ipco_sap_axis(const double *V0, const double *V1, int nV, double inflation);
void ipco_pairs_free(ipco_pairs *p);
/* detect_edge_face_candidates (lbvh.cpp:771-784, filter lbvh.cpp:836-849)
 * over static boxes: (edge, face) pairs, LBVH semantics (FP32 boxes). */
int ipco_edge_face_candidates(const double *V0, int nV, const int32_t *E, int nE,
                              const int32_t *F, int nF, double inflation,
                              const int32_t *vertex_group, ipco_pairs *ef);

/* ---- intersections ------------------------------------------------------- */
/* exact sign of orient3d (igl::predicates::orient3d over Shewchuk's
 * predicate): +1, -1 or 0 (coplanar) */
int ipco_orient3d(const double *pa, const double *pb, const double *pc, const double *pd);
/* orient3d calls the floating-point filter could not decide (process total) */
long long ipco_orient3d_exact_calls(void);
/* Eigen FullPivLU<Matrix3d>(A).solve(b), A column-major */
void ipco_fullpivlu_solve3(const double *A, const double *b, double *x);
/* intersection.cpp:116-138 (non-rational build) */
int ipco_is_edge_intersecting_triangle(const double *e0, const double *e1, const double *t0,
                                       const double *t1, const double *t2);
/* world_bbox_diagonal_length.hpp:11-15 */
double ipco_world_bbox_diagonal(const double *V, int nV);
/* ipc.cpp:105-166 (3D branch); *result = 1 when some edge crosses a face */
int ipco_has_intersections(const double *V, int nV, const int32_t *E, int nE,
                           const int32_t *F, int nF, const int32_t *vertex_group,
                           int *result);

/* ---- narrow phase -------------------------------------------------------- */
typedef struct ipco_collisions {
    int64_t nvv, nev, nee, nfv;
    int32_t *vv;     /* 2*nvv (min, max) */
    double *vv_w;
    int32_t *ev;     /* 2*nev (edge, vertex) */
    double *ev_w;
    int32_t *ee;     /* 2*nee (edge0, edge1) as first found */
    double *ee_w;
    double *ee_eps_x;
    int32_t *ee_dtype; /* actual distance type stored at build */
    int32_t *fv;     /* 2*nfv (face, vertex) */
    double *fv_w;
    double dmin;
} ipco_collisions;

/* normal_collisions.cpp:38-158 + normal_collisions_builder.cpp:26-336,
 * 547-689 (IPC collision-set type, no area weighting); any candidate list
 * may be NULL.  Output is sorted
 * canonically: vv by (min,max), ev by (edge,vertex), ee by
 * ((min,max),dtype), fv by (face,vertex). */
int ipco_build_collisions(const double *V, int nV, const double *V_rest,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          const int32_t *F2E, const ipco_pairs *vv_cand,
                          const ipco_pairs *ev_cand, const ipco_pairs *ee_cand,
                          const ipco_pairs *fv_cand, double dhat, double dmin,
                          ipco_collisions *out);
void ipco_collisions_free(ipco_collisions *c);
/* NormalCollisions::build with collision_set_type IMPROVED_MAX_APPROX when
 * `improved` (normal_collisions.cpp:84-128): the candidate conversions of
 * candidates.cpp:583-692 and the signed vertex-vertex / edge-vertex /
 * mollified edge-edge contributions of normal_collisions_builder.cpp:340-543
 * over the mesh adjacencies of collision_mesh.cpp:247-291, merged with the
 * ordinary ones.  Duplicate weights are summed in ascending order of value;
 * an edge-edge collision keeps the orientation and eps_x of the candidate's
 * contribution, else of the one listing the lower edge id first; collisions
 * (other than face-vertex) whose weights sum to zero are dropped.  NULL
 * areas: unit weights.  set_type: 0 IPC (ipco_build_collisions_weighted),
 * 1 IMPROVED_MAX_APPROX, 2 OGC (the feasible-region filter of
 * ogc/feasible_region.cpp and the weight division of
 * normal_collisions_builder.cpp:87-101, 269-289). */
int ipco_build_collisions_ex(const double *V, int nV, const double *V_rest, const int32_t *E,
                             int nE, const int32_t *F, int nF, const int32_t *F2E,
                             const ipco_pairs *vv_cand, const ipco_pairs *ev_cand,
                             const ipco_pairs *ee_cand, const ipco_pairs *fv_cand, double dhat,
                             double dmin, const double *vertex_area, const double *edge_area,
                             int set_type, ipco_collisions *out);
/* ogc/feasible_region.cpp: check_vertex_feasible_region (point x against
 * vertex vi), check_edge_feasible_region (vertex xi against edge ei), and
 * is_{edge_vertex,edge_edge,face_vertex}_feasible with the distance type
 * classified (kind 1: edge a / vertex b; 2: edges a, b; 3: face a / vertex
 * b).  Return 1 if feasible. */
int ipco_ogc_vertex_region(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                           int nF, const int32_t *F2E, const double x[3], int vi);
int ipco_ogc_edge_region(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                         int nF, const int32_t *F2E, int xi, int ei);
int ipco_ogc_feasible(const double *V, int nV, const int32_t *E, int nE, const int32_t *F, int nF,
                      const int32_t *F2E, int kind, int a, int b);

/* NormalCollisions::set_use_area_weighting(true): the same build with
 * weights 0.5 (a_i + a_j) (codim vertex pairs), 0.5 a_v (codim edge-vertex),
 * 0.25 (A_ea + A_eb) (edge-edge) and 0.25 a_v (face-vertex) per candidate,
 * summed over duplicates (normal_collisions_builder.cpp:26-336); areas from
 * ipco_mesh_areas.  NULL areas: unit weights. */
int ipco_build_collisions_weighted(const double *V, int nV, const double *V_rest,
                                   const int32_t *E, int nE, const int32_t *F, int nF,
                                   const int32_t *F2E, const ipco_pairs *vv_cand,
                                   const ipco_pairs *ev_cand, const ipco_pairs *ee_cand,
                                   const ipco_pairs *fv_cand, double dhat, double dmin,
                                   const double *vertex_area, const double *edge_area,
                                   ipco_collisions *out);
/* CollisionMesh::init_areas (collision_mesh.cpp:309-376): vertex areas (a
 * third of the incident faces' areas, else half the incident edges'
 * lengths, else 1) and edge areas (a third of the incident faces' areas,
 * else the edge length) over rest positions V */
void ipco_mesh_areas(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                     int nF, const int32_t *F2E, double *vertex_area, double *edge_area);

/* ---- barrier potential (potential.cpp:36-222, normal_potential.cpp:127-232,
 *      barrier_potential.cpp:62-99) --------------------------------------- */
double ipco_barrier_energy(const ipco_collisions *c, const double *V, int nV,
                           const int32_t *E, const int32_t *F, double dhat,
                           double kappa);
/* grad: 3*nV, RowMajor [x0,y0,z0,x1,...] (config.hpp.in:45) */
void ipco_barrier_gradient(const ipco_collisions *c, const double *V, int nV,
                           const int32_t *E, const int32_t *F, double dhat,
                           double kappa, double *grad);

typedef struct ipco_csr {
    int32_t n;      /* rows = cols = 3 nV */
    int64_t nnz;
    int64_t *rowptr; /* n+1 */
    int32_t *col;   /* nnz */
    double *val;    /* nnz */
} ipco_csr;
/* Triplets with exact zeros skipped (local_to_global.hpp:267-305),
 * duplicates summed (Eigen setFromTriplets). */
int ipco_barrier_hessian(const ipco_collisions *c, const double *V, int nV,
                         const int32_t *E, const int32_t *F, double dhat,
                         double kappa, int psd, ipco_csr *out);
void ipco_csr_free(ipco_csr *m);
/* ---- friction ------------------------------------------------------------
 * TangentialCollisions::build(mesh, V, collisions, BarrierPotential(dhat,
 * kappa), mu_s, mu_k) with coefficients equal at every vertex (tangential_collisions.cpp:
 * 61-186; tangential_collision.cpp:8-29): one record per collision of `c`
 * (its order); near-parallel edge-edge collisions are not kept.  beta: the
 * closest-point coordinates (2), basis: the tangent basis (3 x 2 column
 * major), normal_force: barrier_force_magnitude at V, weight: the
 * collision's weight.  Energy / gradient / Hessian of FrictionPotential(eps_v)
 * at velocities Vel (tangential_potential.cpp:156-331), isotropic. */
typedef struct ipco_friction {
    int64_t n;
    uint8_t *keep;
    double *beta;
    double *basis;
    double *normal_force;
    double *weight;
    double *mu_s, *mu_k; /* per collision, blended from the vertices' */
    /* anisotropy per collision (tangential_collision.hpp:108-139): velocity
     * scaling mu_aniso (2, default 1), elliptical mu_s_aniso / mu_k_aniso
     * (2 each, default 0) and the lagged effective coefficients used by the
     * potential (default mu_s, mu_k) */
    double *mu_aniso, *mu_s_aniso, *mu_k_aniso, *mu_s_eff, *mu_k_eff;
} ipco_friction;
int ipco_friction_build(const ipco_collisions *c, const double *V, int nV, const int32_t *E,
                        const int32_t *F, double dhat, double kappa, double mu_s, double mu_k,
                        ipco_friction *out);
/* the same with per-vertex coefficients (nV each) blended per collision:
 * the value at each side's closest point, averaged
 * (tangential_collisions.cpp:86-166, default_blend_mu) */
int ipco_friction_build_pv(const ipco_collisions *c, const double *V, int nV, const int32_t *E,
                           const int32_t *F, double dhat, double kappa, const double *mu_s,
                           const double *mu_k, ipco_friction *out);
void ipco_friction_free(ipco_friction *f);
/* set the per-collision anisotropy (2 values per collision each; NULL keeps
 * the current values) and reset the effective coefficients to mu_s, mu_k
 * (reset_lagged_anisotropic_friction_coefficients) */
void ipco_friction_set_anisotropy(ipco_friction *f, const double *mu_aniso,
                                  const double *mu_s_aniso, const double *mu_k_aniso);
/* update_lagged_anisotropic_friction_coefficients (tangential_collisions.cpp:
 * 19-57) with the lagged positions the set was built at: tau = T^T v at the
 * lagged basis, tau_aniso = mu_aniso * tau, effective coefficients from the
 * elliptical model (smooth_mu.cpp anisotropic_mu_eff_from_tau_aniso) when
 * mu_s_aniso is nonzero, else mu_s, mu_k */
void ipco_friction_update_lagged_mu(ipco_friction *f, const ipco_collisions *c, const double *Vel,
                                    const int32_t *E, const int32_t *F);
/* smooth_mu.cpp:115-171: (mu_s_eff, mu_k_eff) for tau_aniso */
void ipco_anisotropic_mu_eff(const double tau_aniso[2], const double mu_s_aniso[2],
                             const double mu_k_aniso[2], double mu_s, double mu_k, double out[2]);
double ipco_friction_energy(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                            int nV, const int32_t *E, const int32_t *F, double eps_v);
void ipco_friction_gradient(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                            int nV, const int32_t *E, const int32_t *F, double eps_v, double *grad);
int ipco_friction_hessian(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                          int nV, const int32_t *E, const int32_t *F, double eps_v, int psd,
                          ipco_csr *out);
/* smooth_mu.cpp / smooth_friction_mollifier.cpp scalar functions: which =
 * 0 mu, 1 mu', 2 mu f0, 3 mu f1, 4 (mu f1)', 5 mu f1/x, 6 (mu f2 x - mu f1)/x^3,
 * 7 f0, 8 f1, 9 f2, 10 f1/x, 11 (f2 x - f1)/x^3 */
double ipco_smooth_mu_fn(int which, double y, double mu_s, double mu_k, double eps_v);


/* ---- Gauss-Newton helpers (normal_potential.cpp:15-83, 236-308) ---------- */
/* distance vector t of collision i (collision order vv, ev, ee, fv) and its
 * coefficients (collision_stencil.cpp:14-25); returns the point count */
double ipco_distance_vector(const ipco_collisions *c, const double *V, const int32_t *E,
                            const int32_t *F, int64_t i, double *t, double *coeffs);
/* diag: 3*nV RowMajor */
void ipco_gauss_newton_diagonal(const ipco_collisions *c, const double *V, int nV,
                                const int32_t *E, const int32_t *F, double dhat,
                                double kappa, double *diag);
/* p: 3*nV RowMajor */
double ipco_gauss_newton_quadratic_form(const ipco_collisions *c, const double *V, int nV,
                                        const int32_t *E, const int32_t *F, double dhat,
                                        double kappa, const double *p);

/* ---- candidate-set helpers (candidates.cpp:294-408) ----------------------- */
/* any list may be NULL; out: nV */
void ipco_per_vertex_safe_distances(const double *V, int nV, const int32_t *E,
                                    const int32_t *F, const ipco_pairs *ee,
                                    const ipco_pairs *fv, const ipco_pairs *vv,
                                    const ipco_pairs *ev, double inflation_radius,
                                    double min_distance, double *out);
double ipco_noncandidate_conservative_stepsize(const double *V0, const double *V1, int nV,
                                               const int32_t *E, const int32_t *F,
                                               const ipco_pairs *ee, const ipco_pairs *fv,
                                               const ipco_pairs *vv, const ipco_pairs *ev,
                                               double dhat);

/* ---- CCD (tight_inclusion_ccd.cpp:23-336 + Tight-Inclusion BFS) --------- */
typedef struct ipco_ccd_params {
    double tolerance;        /* 1e-6 */
    long max_iterations;     /* 1e7 */
    double conservative_rescaling; /* 0.8 */
} ipco_ccd_params;

/* Raw Tight-Inclusion root finders.  is_vf: 1 vertex-face
 * (p, t0, t1, t2), 0 edge-edge (ea0, ea1, eb0, eb1).  x_t0/x_t1 hold 12
 * doubles.  err < 0 selects the automatic error filter.  Returns hit. */
int ipco_ticcd(int is_vf, const double *x_t0, const double *x_t1,
               double err_in, double ms, double tolerance, double t_max,
               long max_itr, int no_zero_toi, double *toi,
               double *output_tolerance);

/* Boxes processed and the largest queue of the last ipco_ticcd call made
 * on the calling thread (work probes for tests). */
void ipco_ticcd_last_stats(long *refine, long *max_queue);

/* TightInclusionCCD::{point_point,point_edge,edge_edge,point_triangle}_ccd.
 * kind as in ipco_stencil_distance_derivs.  Returns hit; toi written. */
int ipco_stencil_ccd(int kind, const double *x_t0, const double *x_t1,
                     double min_distance, double tmax,
                     const ipco_ccd_params *prm, double *toi);

/* candidates.cpp:252-292 over the ee, fv, vv and ev lists (vv / ev may be
 * NULL).  running_tmax != 0 follows the reference's running earliest-toi
 * (order dependent); 0 searches every candidate on [0,1] hiding outcomes at
 * or above the running minimum (deterministic: the minimum of the
 * per-candidate tois).  Returns the step in [0,1]. */
double ipco_candidates_stepsize(const double *V0, const double *V1, int nV,
                                const int32_t *E, const int32_t *F,
                                const ipco_pairs *ee, const ipco_pairs *fv,
                                const ipco_pairs *vv, const ipco_pairs *ev,
                                double min_distance,
                                const ipco_ccd_params *prm, int running_tmax);
/* ipc.cpp:45-101 default (LBVH) path: swept candidates (surface and codim
 * lists) with inflation 0.5*min_distance then the above. */
double ipco_compute_collision_free_stepsize(
    const double *V0, const double *V1, int nV, const int32_t *E, int nE,
    const int32_t *F, int nF, double min_distance,
    const ipco_ccd_params *prm, int running_tmax);
/* ipc.cpp:20-41 + candidates.cpp:224-250 */
int ipco_is_step_collision_free(const double *V0, const double *V1, int nV,
                                const int32_t *E, int nE, const int32_t *F,
                                int nF, double min_distance,
                                const ipco_ccd_params *prm);

/* per-candidate toi for a candidate list (tmax = 1 each); kind 0 = fv
 * list, 1 = ee, 2 = ev (edge, vertex), 3 = vv; toi[i] = +inf when no hit. */
void ipco_candidates_toi(const double *V0, const double *V1, const int32_t *E,
                         const int32_t *F, const ipco_pairs *cand, int kind,
                         double min_distance, const ipco_ccd_params *prm,
                         double *toi);

void ipco_set_num_threads(int n);
int ipco_get_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
