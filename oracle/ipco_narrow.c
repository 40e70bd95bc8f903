/* Narrow phase (collision-set build) and barrier potential restatement.
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h). */
#include "ipco_internal.h"

enum { T_VV = 0, T_EV = 1, T_EE = 2, T_FV = 3 };

typedef struct rec {
    int32_t type; /* -1 inactive */
    int32_t a, b; /* VV (min,max); EV (edge,vertex); EE (ea,eb); FV (f,v) */
    int32_t dtype;
    double eps_x;
    double w; /* weight of this contribution */
    int32_t rank; /* EE representative order: 0 candidate, 1 (min,max), 2 (max,min) */
} rec;

/* ---- mesh areas (collision_mesh.cpp:309-376 init_areas) ------------------
 * edge_length = |e1 - e0| over runtime-size vectors (left-to-right sum);
 * triangle_area = 0.5 |(t1 - t0) x (t2 - t0)| (fixed-size, unrolled sum). */
static double edge_len(const double *V, const int32_t *E, int i)
{
    const double *a = V + 3 * (size_t)E[2 * (size_t)i], *b = V + 3 * (size_t)E[2 * (size_t)i + 1];
    const double d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
    return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

void ipco_mesh_areas(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                     int nF, const int32_t *F2E, double *vertex_area, double *edge_area)
{
    double *ve = (double *)malloc(sizeof(double) * (size_t)(nV > 0 ? nV : 1));
    double *vf = (double *)malloc(sizeof(double) * (size_t)(nV > 0 ? nV : 1));
    for (int i = 0; i < nV; i++) ve[i] = vf[i] = -1.0;
    for (int i = 0; i < nE; i++) {
        const double l = edge_len(V, E, i);
        for (int j = 0; j < 2; j++) {
            const int v = E[2 * (size_t)i + j];
            ve[v] = ve[v] > 0.0 ? ve[v] : 0.0;
            ve[v] += 0.5 * l;
        }
    }
    for (int i = 0; i < nE; i++) edge_area[i] = -1.0;
    for (int i = 0; i < nF; i++) {
        double e0[3], e1[3], c[3];
        const double *t0 = V + 3 * (size_t)F[3 * (size_t)i];
        v3_sub(V + 3 * (size_t)F[3 * (size_t)i + 1], t0, e0);
        v3_sub(V + 3 * (size_t)F[3 * (size_t)i + 2], t0, e1);
        v3_cross(e0, e1, c);
        const double area = 0.5 * sqrt(v3_sqnorm(c));
        for (int j = 0; j < 3; j++) {
            const int v = F[3 * (size_t)i + j];
            vf[v] = vf[v] > 0.0 ? vf[v] : 0.0;
            vf[v] += area / 3.0;
            const int e = F2E[3 * (size_t)i + j];
            edge_area[e] = edge_area[e] > 0.0 ? edge_area[e] : 0.0;
            edge_area[e] += area / 3.0;
        }
    }
    for (int i = 0; i < nV; i++)
        vertex_area[i] = vf[i] < 0 ? (ve[i] < 0 ? 1.0 : ve[i]) : vf[i];
    for (int i = 0; i < nE; i++)
        if (edge_area[i] < 0) edge_area[i] = edge_len(V, E, i);
    free(ve);
    free(vf);
}

static void vv_rec(rec *r, int32_t u, int32_t v)
{
    r->type = T_VV;
    r->a = u < v ? u : v;
    r->b = u < v ? v : u;
}

static void ev_rec(rec *r, int32_t e, int32_t v)
{
    r->type = T_EV;
    r->a = e;
    r->b = v;
}

static inline const double *vx(const double *V, int i)
{
    return V + 3 * (size_t)i;
}

/* normal_collisions_builder.cpp:26-54 add_vertex_vertex_collision */
static void vv_candidate(const double *V, int32_t vi, int32_t vj, double active2, rec *r)
{
    r->type = -1;
    const double d = ipco_point_point_distance(vx(V, vi), vx(V, vj));
    if (d < active2) vv_rec(r, vi, vj);
}

/* normal_collisions_builder.cpp:56-138 add_edge_vertex_collision */
static void ev_candidate(const double *V, const int32_t *E, int32_t ei, int32_t vi,
                         double active2, rec *r)
{
    r->type = -1;
    const int32_t e0 = E[2 * (size_t)ei], e1 = E[2 * (size_t)ei + 1];
    const double *p = vx(V, vi), *q0 = vx(V, e0), *q1 = vx(V, e1);
    const int dt = ipco_point_edge_distance_type(p, q0, q1);
    const double d = ipco_point_edge_distance(p, q0, q1, dt);
    if (!(d < active2)) return;
    switch (dt) {
    case IPCO_PE_P_E0: vv_rec(r, vi, e0); break;
    case IPCO_PE_P_E1: vv_rec(r, vi, e1); break;
    default: ev_rec(r, ei, vi);
    }
}

/* normal_collisions_builder.cpp:239-336 add_face_vertex_collision */
static void fv_candidate(const double *V, const int32_t *F, const int32_t *F2E,
                         int32_t fi, int32_t vi, double active2, rec *r)
{
    r->type = -1;
    const int32_t f0 = F[3 * (size_t)fi], f1 = F[3 * (size_t)fi + 1],
                  f2 = F[3 * (size_t)fi + 2];
    const double *p = vx(V, vi), *t0 = vx(V, f0), *t1 = vx(V, f1),
                 *t2 = vx(V, f2);
    const int dt = ipco_point_triangle_distance_type(p, t0, t1, t2);
    const double d = ipco_point_triangle_distance(p, t0, t1, t2, dt);
    if (!(d < active2)) return;
    switch (dt) {
    case IPCO_PT_P_T0: vv_rec(r, vi, f0); break;
    case IPCO_PT_P_T1: vv_rec(r, vi, f1); break;
    case IPCO_PT_P_T2: vv_rec(r, vi, f2); break;
    case IPCO_PT_P_E0: ev_rec(r, F2E[3 * (size_t)fi], vi); break;
    case IPCO_PT_P_E1: ev_rec(r, F2E[3 * (size_t)fi + 1], vi); break;
    case IPCO_PT_P_E2: ev_rec(r, F2E[3 * (size_t)fi + 2], vi); break;
    default:
        r->type = T_FV;
        r->a = fi;
        r->b = vi;
    }
}

/* normal_collisions_builder.cpp:140-237 add_edge_edge_collision */
static void ee_candidate(const double *V, const double *Vr, const int32_t *E,
                         int32_t ea, int32_t eb, double active2, rec *r)
{
    r->type = -1;
    const int32_t a0 = E[2 * (size_t)ea], a1 = E[2 * (size_t)ea + 1];
    const int32_t b0 = E[2 * (size_t)eb], b1 = E[2 * (size_t)eb + 1];
    const double *pa0 = vx(V, a0), *pa1 = vx(V, a1), *pb0 = vx(V, b0),
                 *pb1 = vx(V, b1);
    const int actual = ipco_edge_edge_distance_type(pa0, pa1, pb0, pb1);
    const double d = ipco_edge_edge_distance(pa0, pa1, pb0, pb1, actual);
    if (!(d < active2)) return;
    const double eps_x = ipco_edge_edge_mollifier_threshold(
        vx(Vr, a0), vx(Vr, a1), vx(Vr, b0), vx(Vr, b1));
    const double x = ipco_edge_edge_cross_squarednorm(pa0, pa1, pb0, pb1);
    const int dt = x < eps_x ? IPCO_EE_EA_EB : actual;
    switch (dt) {
    case IPCO_EE_EA0_EB0: vv_rec(r, a0, b0); break;
    case IPCO_EE_EA0_EB1: vv_rec(r, a0, b1); break;
    case IPCO_EE_EA1_EB0: vv_rec(r, a1, b0); break;
    case IPCO_EE_EA1_EB1: vv_rec(r, a1, b1); break;
    case IPCO_EE_EA_EB0: ev_rec(r, ea, b0); break;
    case IPCO_EE_EA_EB1: ev_rec(r, ea, b1); break;
    case IPCO_EE_EA0_EB: ev_rec(r, eb, a0); break;
    case IPCO_EE_EA1_EB: ev_rec(r, eb, a1); break;
    default:
        r->type = T_EE;
        r->a = ea;
        r->b = eb;
        r->dtype = actual;
        r->eps_x = eps_x;
    }
}

static uint64_t rec_key(const rec *r)
{
    if (r->type == T_EE) {
        const uint32_t lo = (uint32_t)(r->a < r->b ? r->a : r->b);
        const uint32_t hi = (uint32_t)(r->a < r->b ? r->b : r->a);
        /* edge ids < 2^28 assumed by this packing; dtype in the low 4 bits */
        return ((uint64_t)lo << 36) | ((uint64_t)hi << 4) | (uint64_t)r->dtype;
    }
    return ((uint64_t)(uint32_t)r->a << 32) | (uint32_t)r->b;
}

int ipco_build_collisions(const double *V, int nV, const double *V_rest,
                          const int32_t *E, int nE, const int32_t *F, int nF,
                          const int32_t *F2E, const ipco_pairs *vv_cand,
                          const ipco_pairs *ev_cand, const ipco_pairs *ee_cand,
                          const ipco_pairs *fv_cand, double dhat, double dmin,
                          ipco_collisions *out)
{
    return ipco_build_collisions_weighted(V, nV, V_rest, E, nE, F, nF, F2E, vv_cand, ev_cand,
                                          ee_cand, fv_cand, dhat, dmin, NULL, NULL, out);
}

int ipco_build_collisions_weighted(const double *V, int nV, const double *V_rest,
                                   const int32_t *E, int nE, const int32_t *F, int nF,
                                   const int32_t *F2E, const ipco_pairs *vv_cand,
                                   const ipco_pairs *ev_cand, const ipco_pairs *ee_cand,
                                   const ipco_pairs *fv_cand, double dhat, double dmin,
                                   const double *vertex_area, const double *edge_area,
                                   ipco_collisions *out)
{
    return ipco_build_collisions_ex(V, nV, V_rest, E, nE, F, nF, F2E, vv_cand, ev_cand, ee_cand,
                                    fv_cand, dhat, dmin, vertex_area, edge_area, 0, out);
}

/* ---- mesh adjacencies (collision_mesh.cpp:247-291 init_adjacencies) -------
 * Sorted, duplicate-free lists in CSR form. */
typedef struct adjl {
    int64_t *off;
    int32_t *id;
} adjl;

static int cmp_i32(const void *a, const void *b)
{
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* rows[k] -> vals[k] for k < m, n rows */
static void adj_build(int n, int64_t m, const int32_t *rows, const int32_t *vals, adjl *A)
{
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t k = 0; k < m; k++) cnt[rows[k] + 1]++;
    for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    int32_t *tmp = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; i++) fill[i] = cnt[i];
    for (int64_t k = 0; k < m; k++) tmp[fill[rows[k]]++] = vals[k];
    A->off = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    A->id = (int32_t *)malloc(sizeof(int32_t) * (size_t)(m > 0 ? m : 1));
    int64_t at = 0;
    for (int i = 0; i < n; i++) {
        A->off[i] = at;
        const int64_t s0 = cnt[i], s1 = cnt[i + 1];
        qsort(tmp + s0, (size_t)(s1 - s0), sizeof(int32_t), cmp_i32);
        for (int64_t k = s0; k < s1; k++)
            if (k == s0 || tmp[k] != tmp[k - 1]) A->id[at++] = tmp[k];
    }
    A->off[n] = at;
    free(cnt);
    free(tmp);
    free(fill);
}

static int64_t adj_deg(const adjl *A, int i)
{
    return A->off[i + 1] - A->off[i];
}

static int adj_has(const adjl *A, int i, int32_t v)
{
    int64_t lo = A->off[i], hi = A->off[i + 1];
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (A->id[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < A->off[i + 1] && A->id[lo] == v;
}

typedef struct mesh_adj {
    adjl vv;      /* vertex -> vertices sharing an edge */
    adjl ve;      /* vertex -> incident edges */
    adjl ev;      /* edge -> vertices opposite it in its faces */
    adjl vf;      /* vertex -> incident faces */
    uint8_t *bnd; /* is_vertex_on_boundary (3D rule) */
} mesh_adj;

static void mesh_adj_build(int nV, const int32_t *E, int nE, const int32_t *F, int nF,
                           const int32_t *F2E, mesh_adj *M)
{
    const int64_t m = 2 * (int64_t)nE;
    int32_t *r = (int32_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(int32_t));
    int32_t *v = (int32_t *)calloc((size_t)(m > 0 ? m : 1), sizeof(int32_t));
    for (int i = 0; i < nE; i++) {
        r[2 * i] = E[2 * i];
        v[2 * i] = E[2 * i + 1];
        r[2 * i + 1] = E[2 * i + 1];
        v[2 * i + 1] = E[2 * i];
    }
    adj_build(nV, m, r, v, &M->vv);
    for (int i = 0; i < nE; i++) v[2 * i] = v[2 * i + 1] = i;
    adj_build(nV, m, r, v, &M->ve);
    free(r);
    free(v);
    const int64_t mf = 3 * (int64_t)nF;
    r = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mf > 0 ? mf : 1));
    v = (int32_t *)malloc(sizeof(int32_t) * (size_t)(mf > 0 ? mf : 1));
    for (int i = 0; i < nF; i++)
        for (int j = 0; j < 3; j++) {
            r[3 * i + j] = F2E[3 * i + j];
            v[3 * i + j] = F[3 * i + (j + 2) % 3];
        }
    adj_build(nE, mf, r, v, &M->ev);
    for (int i = 0; i < nF; i++)
        for (int j = 0; j < 3; j++) {
            r[3 * i + j] = F[3 * i + j];
            v[3 * i + j] = i;
        }
    adj_build(nV, mf, r, v, &M->vf);
    free(r);
    free(v);
    M->bnd = (uint8_t *)malloc((size_t)(nV > 0 ? nV : 1));
    memset(M->bnd, 1, (size_t)(nV > 0 ? nV : 1));
    for (int i = 0; i < nE; i++)
        if (adj_deg(&M->ev, i) >= 2) M->bnd[E[2 * i]] = M->bnd[E[2 * i + 1]] = 0;
}

static void mesh_adj_free(mesh_adj *M)
{
    free(M->vv.off);
    free(M->vv.id);
    free(M->ve.off);
    free(M->ve.id);
    free(M->ev.off);
    free(M->ev.id);
    free(M->vf.off);
    free(M->vf.id);
    free(M->bnd);
}

/* ---- OGC feasible regions (ogc/feasible_region.cpp) ----------------------
 * Dot products over runtime-size vectors sum left to right, over fixed
 * 3-vectors as a0 b0 + (a1 b1 + a2 b2). */
static double dot_lr(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

static double dot_fx(const double *a, const double *b)
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}

/* closest_point.cpp:11-18 point_edge_closest_point */
static double pe_param(const double *p, const double *e0, const double *e1)
{
    const double e[3] = { e1[0] - e0[0], e1[1] - e0[1], e1[2] - e0[2] };
    const double d[3] = { p[0] - e0[0], p[1] - e0[1], p[2] - e0[2] };
    return dot_lr(d, e) / dot_lr(e, e);
}

/* feasible_region.cpp:7-24: x behind vertex vi against every neighbour */
static int ogc_vertex_region(const double *V, const mesh_adj *M, const double *x, int vi)
{
    const double *xv = vx(V, vi);
    const double d[3] = { x[0] - xv[0], x[1] - xv[1], x[2] - xv[2] };
    for (int64_t t = M->vv.off[vi]; t < M->vv.off[vi + 1]; t++) {
        const double *xj = vx(V, M->vv.id[t]);
        const double u[3] = { xv[0] - xj[0], xv[1] - xj[1], xv[2] - xj[2] };
        if (!(dot_lr(d, u) >= 0)) return 0;
    }
    return 1;
}

/* feasible_region.cpp:26-54: vertex xi over edge ei, beyond the edge from
 * every opposite vertex */
static int ogc_edge_region(const double *V, const int32_t *E, const mesh_adj *M, int xi, int ei)
{
    const double *x = vx(V, xi), *x0 = vx(V, E[2 * (size_t)ei]), *x1 = vx(V, E[2 * (size_t)ei + 1]);
    const double alpha = pe_param(x, x0, x1);
    if (alpha <= 0 || alpha >= 1) return 0;
    for (int64_t t = M->ev.off[ei]; t < M->ev.off[ei + 1]; t++) {
        const double *x2 = vx(V, M->ev.id[t]);
        const double a2 = pe_param(x2, x0, x1);
        double p[3], u[3], w[3];
        for (int k = 0; k < 3; k++) {
            p[k] = (x1[k] - x0[k]) * a2 + x0[k];
            u[k] = x[k] - p[k];
            w[k] = p[k] - x2[k];
        }
        if (!(dot_fx(u, w) >= 0)) return 0;
    }
    return 1;
}

static int ogc_vv_pair(const double *V, const mesh_adj *M, int a, int b)
{
    return ogc_vertex_region(V, M, vx(V, a), b) && ogc_vertex_region(V, M, vx(V, b), a);
}

/* feasible_region.cpp:56-89 is_edge_vertex_feasible */
static int ogc_ev_feasible(const double *V, const int32_t *E, const mesh_adj *M, int ei, int vi,
                           int dt)
{
    if (dt == IPCO_PE_P_E0) return ogc_vv_pair(V, M, vi, E[2 * (size_t)ei]);
    if (dt == IPCO_PE_P_E1) return ogc_vv_pair(V, M, vi, E[2 * (size_t)ei + 1]);
    return 1;
}

/* feasible_region.cpp:91-167 is_edge_edge_feasible */
static int ogc_ee_feasible(const double *V, const int32_t *E, const mesh_adj *M, int ea, int eb,
                           int dt)
{
    const int a[2] = { E[2 * (size_t)ea], E[2 * (size_t)ea + 1] };
    const int b[2] = { E[2 * (size_t)eb], E[2 * (size_t)eb + 1] };
    switch (dt) {
    case IPCO_EE_EA0_EB0: return ogc_vv_pair(V, M, a[0], b[0]);
    case IPCO_EE_EA0_EB1: return ogc_vv_pair(V, M, a[0], b[1]);
    case IPCO_EE_EA1_EB0: return ogc_vv_pair(V, M, a[1], b[0]);
    case IPCO_EE_EA1_EB1: return ogc_vv_pair(V, M, a[1], b[1]);
    case IPCO_EE_EA_EB: return 1;
    default: break;
    }
    /* EA_EB0 / EA_EB1: the end of eb against its closest point on ea, and
     * EA0_EB / EA1_EB the other way round */
    const int on_a = dt == IPCO_EE_EA_EB0 || dt == IPCO_EE_EA_EB1;
    const int v = on_a ? b[dt == IPCO_EE_EA_EB1] : a[dt == IPCO_EE_EA1_EB];
    const int *e = on_a ? a : b;
    const double *q0 = vx(V, e[0]), *q1 = vx(V, e[1]);
    const double alpha = pe_param(vx(V, v), q0, q1);
    double xc[3];
    for (int k = 0; k < 3; k++) xc[k] = (q1[k] - q0[k]) * alpha + q0[k];
    return ogc_vertex_region(V, M, xc, v);
}

/* feasible_region.cpp:169-224 is_face_vertex_feasible */
static int ogc_fv_feasible(const double *V, const int32_t *E, const int32_t *F, const int32_t *F2E,
                           const mesh_adj *M, int fi, int vi, int dt)
{
    if (dt <= IPCO_PT_P_T2) return ogc_vv_pair(V, M, vi, F[3 * (size_t)fi + dt]);
    if (dt <= IPCO_PT_P_E2) return ogc_edge_region(V, E, M, vi, F2E[3 * (size_t)fi + dt - 3]);
    return 1;
}

/* normal_collisions_builder.cpp:56-114 (use_ogc): infeasible edge-vertex
 * pairs are dropped; a pair reduced to a vertex is divided among the
 * vertex's incident edges */
static void ogc_ev(const double *V, const int32_t *E, const mesh_adj *M, int32_t ei, int32_t vi,
                   rec *r)
{
    if (r->type < 0) return;
    const int32_t e[2] = { E[2 * (size_t)ei], E[2 * (size_t)ei + 1] };
    const int dt = ipco_point_edge_distance_type(vx(V, vi), vx(V, e[0]), vx(V, e[1]));
    if (!ogc_ev_feasible(V, E, M, ei, vi, dt)) {
        r->type = -1;
        return;
    }
    if (dt < IPCO_PE_P_E) {
        const int64_t n = adj_deg(&M->ve, e[dt]);
        if (n > 1) r->w /= (double)n;
    }
}

/* normal_collisions_builder.cpp:140-182 (use_ogc): the feasibility test
 * uses the actual distance type, before the mollifier's override */
static void ogc_ee(const double *V, const int32_t *E, const mesh_adj *M, int32_t ea, int32_t eb,
                   rec *r)
{
    if (r->type < 0) return;
    const int dt = ipco_edge_edge_distance_type(vx(V, E[2 * (size_t)ea]), vx(V, E[2 * (size_t)ea + 1]),
                                                vx(V, E[2 * (size_t)eb]), vx(V, E[2 * (size_t)eb + 1]));
    if (!ogc_ee_feasible(V, E, M, ea, eb, dt)) r->type = -1;
}

/* normal_collisions_builder.cpp:239-296 (use_ogc): divided among the faces
 * of the vertex or the opposite vertices of the edge it reduces to */
static void ogc_fv(const double *V, const int32_t *E, const int32_t *F, const int32_t *F2E,
                   const mesh_adj *M, int32_t fi, int32_t vi, rec *r)
{
    if (r->type < 0) return;
    const int32_t *f = F + 3 * (size_t)fi;
    const int dt = ipco_point_triangle_distance_type(vx(V, vi), vx(V, f[0]), vx(V, f[1]), vx(V, f[2]));
    if (!ogc_fv_feasible(V, E, F, F2E, M, fi, vi, dt)) {
        r->type = -1;
        return;
    }
    int64_t n = 1;
    if (dt <= IPCO_PT_P_T2) n = adj_deg(&M->vf, f[dt]);
    else if (dt <= IPCO_PT_P_E2) n = adj_deg(&M->ev, F2E[3 * (size_t)fi + dt - 3]);
    if (n > 1) r->w /= (double)n;
}

int ipco_ogc_vertex_region(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                           int nF, const int32_t *F2E, const double x[3], int vi)
{
    mesh_adj M;
    mesh_adj_build(nV, E, nE, F, nF, F2E, &M);
    const int r = ogc_vertex_region(V, &M, x, vi);
    mesh_adj_free(&M);
    return r;
}

int ipco_ogc_edge_region(const double *V, int nV, const int32_t *E, int nE, const int32_t *F,
                         int nF, const int32_t *F2E, int xi, int ei)
{
    mesh_adj M;
    mesh_adj_build(nV, E, nE, F, nF, F2E, &M);
    const int r = ogc_edge_region(V, E, &M, xi, ei);
    mesh_adj_free(&M);
    return r;
}

int ipco_ogc_feasible(const double *V, int nV, const int32_t *E, int nE, const int32_t *F, int nF,
                      const int32_t *F2E, int kind, int a, int b)
{
    mesh_adj M;
    mesh_adj_build(nV, E, nE, F, nF, F2E, &M);
    int r;
    if (kind == 1) { /* (edge a, vertex b) */
        const int dt = ipco_point_edge_distance_type(vx(V, b), vx(V, E[2 * (size_t)a]),
                                                     vx(V, E[2 * (size_t)a + 1]));
        r = ogc_ev_feasible(V, E, &M, a, b, dt);
    } else if (kind == 2) { /* (edge a, edge b) */
        const int dt = ipco_edge_edge_distance_type(
            vx(V, E[2 * (size_t)a]), vx(V, E[2 * (size_t)a + 1]), vx(V, E[2 * (size_t)b]),
            vx(V, E[2 * (size_t)b + 1]));
        r = ogc_ee_feasible(V, E, &M, a, b, dt);
    } else { /* (face a, vertex b) */
        const int32_t *f = F + 3 * (size_t)a;
        const int dt = ipco_point_triangle_distance_type(vx(V, b), vx(V, f[0]), vx(V, f[1]),
                                                         vx(V, f[2]));
        r = ogc_fv_feasible(V, E, F, F2E, &M, a, b, dt);
    }
    mesh_adj_free(&M);
    return r;
}

/* ---- IMPROVED_MAX_APPROX (normal_collisions.cpp:84-128) ------------------ */

typedef struct recvec {
    rec *r;
    int64_t n, cap;
} recvec;

static rec *recvec_push(recvec *R)
{
    if (R->n == R->cap) {
        R->cap = R->cap ? 2 * R->cap : 64;
        R->r = (rec *)realloc(R->r, sizeof(rec) * (size_t)R->cap);
    }
    rec *x = &R->r[R->n++];
    memset(x, 0, sizeof(*x));
    return x;
}

static void push_vv(recvec *R, int32_t u, int32_t v, double w)
{
    rec *x = recvec_push(R);
    vv_rec(x, u, v);
    x->w = w;
}

static void push_ev(recvec *R, int32_t e, int32_t v, double w)
{
    rec *x = recvec_push(R);
    ev_rec(x, e, v);
    x->w = w;
}

/* normal_collisions_builder.cpp:116-138: an edge-vertex pair reduced by its
 * point-edge distance type */
static void push_ev_typed(recvec *R, const int32_t *E, int32_t ei, int32_t vi, int dt, double w)
{
    if (dt == IPCO_PE_P_E0) push_vv(R, vi, E[2 * (size_t)ei], w);
    else if (dt == IPCO_PE_P_E1) push_vv(R, vi, E[2 * (size_t)ei + 1], w);
    else push_ev(R, ei, vi, w);
}

static uint64_t pk(int32_t a, int32_t b)
{
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

static uint64_t vvk(int32_t a, int32_t b)
{
    return a < b ? pk(a, b) : pk(b, a);
}

/* sort + unique of a key list (candidates.cpp:604-612, 655-657, 686-688) */
static int64_t unique_keys(uint64_t *k, int64_t n)
{
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) idx[i] = i;
    ipco_sort_u64_pairs(k, idx, n);
    free(idx);
    int64_t u = 0;
    for (int64_t i = 0; i < n; i++)
        if (i == 0 || k[i] != k[i - 1]) k[u++] = k[i];
    return u;
}

static void improved_records(const double *V, const double *Vr, const int32_t *E, int nE,
                             const int32_t *F, int nF, const int32_t *F2E, int nV,
                             const ipco_pairs *ev_cand, const ipco_pairs *ee_cand,
                             const ipco_pairs *fv_cand, double active2, const double *va,
                             const double *ea, recvec *R)
{
    mesh_adj M;
    mesh_adj_build(nV, E, nE, F, nF, F2E, &M);
    const int aw = va != NULL;
    const int64_t nev = ev_cand ? ev_cand->n : 0, nee = ee_cand ? ee_cand->n : 0,
                  nfv = fv_cand ? fv_cand->n : 0;
    uint64_t *k;
    int64_t m;

    /* edge-vertex -> vertex-vertex (candidates.cpp:622-629), negative
     * vertex-vertex weights (normal_collisions_builder.cpp:340-383) */
    k = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(2 * nev + 1));
    m = 0;
    for (int64_t i = 0; i < nev; i++) {
        const int32_t ei = ev_cand->ij[2 * i], vi = ev_cand->ij[2 * i + 1];
        for (int j = 0; j < 2; j++) {
            const int32_t vj = E[2 * (size_t)ei + j];
            if (ipco_point_point_distance(vx(V, vi), vx(V, vj)) < active2) k[m++] = vvk(vi, vj);
        }
    }
    m = unique_keys(k, m);
    for (int64_t i = 0; i < m; i++) {
        const int32_t u[2] = { (int32_t)(k[i] >> 32), (int32_t)(k[i] & 0xFFFFFFFFu) };
        double w = 0.0;
        for (int s = 0; s < 2; s++) {
            const int32_t a = u[s], b = u[1 - s];
            const int64_t inc = adj_deg(&M.vv, b) - adj_has(&M.vv, b, a);
            if (inc > 1) w += (double)(1 - inc) * (aw ? 0.5 * va[a] : 1.0);
        }
        if (w != 0) push_vv(R, u[0], u[1], w);
    }
    free(k);

    /* edge-edge -> edge-vertex (candidates.cpp:665-692), negative
     * edge-vertex weights with mollified edge-edge corrections
     * (normal_collisions_builder.cpp:463-543) */
    k = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(4 * nee + 1));
    m = 0;
    for (int64_t i = 0; i < nee; i++) {
        for (int s = 0; s < 2; s++) {
            const int32_t ei = ee_cand->ij[2 * i + s], ej = ee_cand->ij[2 * i + 1 - s];
            const int32_t ei0 = E[2 * (size_t)ei], ei1 = E[2 * (size_t)ei + 1];
            for (int j = 0; j < 2; j++) {
                const int32_t vj = E[2 * (size_t)ej + j];
                const double *p = vx(V, vj), *q0 = vx(V, ei0), *q1 = vx(V, ei1);
                const int dt = ipco_point_edge_distance_type(p, q0, q1);
                if (ipco_point_edge_distance(p, q0, q1, dt) < active2) k[m++] = pk(ei, vj);
            }
        }
    }
    m = unique_keys(k, m);
    for (int64_t i = 0; i < m; i++) {
        const int32_t a = (int32_t)(k[i] >> 32), p = (int32_t)(k[i] & 0xFFFFFFFFu);
        const int32_t a0 = E[2 * (size_t)a], a1 = E[2 * (size_t)a + 1];
        const double w = aw ? -0.25 * ea[a] : -1.0;
        const int dt = ipco_point_edge_distance_type(vx(V, p), vx(V, a0), vx(V, a1));
        int64_t nonmollified = 0;
        for (int64_t t = M.ve.off[p]; t < M.ve.off[p + 1]; t++) {
            const int32_t b = M.ve.id[t];
            const int32_t b0 = E[2 * (size_t)b], b1 = E[2 * (size_t)b + 1];
            const int32_t q = p == b0 ? b1 : b0;
            if (q == a0 || q == a1) continue;
            const double eps_x = ipco_edge_edge_mollifier_threshold(vx(Vr, a0), vx(Vr, a1),
                                                                    vx(Vr, b0), vx(Vr, b1));
            const double x = ipco_edge_edge_cross_squarednorm(vx(V, a0), vx(V, a1), vx(V, b0),
                                                              vx(V, b1));
            if (x >= eps_x) {
                nonmollified++;
                continue;
            }
            int edt;
            if (dt == IPCO_PE_P_E0) edt = p == b0 ? IPCO_EE_EA0_EB0 : IPCO_EE_EA0_EB1;
            else if (dt == IPCO_PE_P_E1) edt = p == b0 ? IPCO_EE_EA1_EB0 : IPCO_EE_EA1_EB1;
            else edt = p == b0 ? IPCO_EE_EA_EB0 : IPCO_EE_EA_EB1;
            rec *x_ = recvec_push(R);
            x_->type = T_EE;
            x_->a = a;
            x_->b = b;
            x_->dtype = edt;
            x_->eps_x = eps_x;
            x_->w = w;
            x_->rank = a < b ? 1 : 2;
        }
        if (nonmollified != 1) push_ev_typed(R, E, a, p, dt, (double)(nonmollified - 1) * w);
    }
    free(k);

    /* face-vertex -> edge-vertex (candidates.cpp:640-663), negative
     * edge-vertex weights (normal_collisions_builder.cpp:427-461) */
    k = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(3 * nfv + 1));
    m = 0;
    for (int64_t i = 0; i < nfv; i++) {
        const int32_t fi = fv_cand->ij[2 * i], vi = fv_cand->ij[2 * i + 1];
        for (int j = 0; j < 3; j++) {
            const int32_t ei = F2E[3 * (size_t)fi + j];
            const double *p = vx(V, vi), *q0 = vx(V, E[2 * (size_t)ei]),
                         *q1 = vx(V, E[2 * (size_t)ei + 1]);
            const int dt = ipco_point_edge_distance_type(p, q0, q1);
            if (ipco_point_edge_distance(p, q0, q1, dt) < active2) k[m++] = pk(ei, vi);
        }
    }
    m = unique_keys(k, m);
    for (int64_t i = 0; i < m; i++) {
        const int32_t ei = (int32_t)(k[i] >> 32), vi = (int32_t)(k[i] & 0xFFFFFFFFu);
        const int64_t inc = adj_deg(&M.ev, ei) - adj_has(&M.ev, ei, vi);
        if (inc > 1) {
            const double w = (double)(1 - inc) * (aw ? 0.25 * va[vi] : 1.0);
            const int dt = ipco_point_edge_distance_type(vx(V, vi), vx(V, E[2 * (size_t)ei]),
                                                         vx(V, E[2 * (size_t)ei + 1]));
            push_ev_typed(R, E, ei, vi, dt, w);
        }
    }

    /* face-vertex -> vertex-vertex (candidates.cpp:631-638), positive
     * vertex-vertex weights (normal_collisions_builder.cpp:385-425) */
    k = (uint64_t *)realloc(k, sizeof(uint64_t) * (size_t)(3 * nfv + 1));
    m = 0;
    for (int64_t i = 0; i < nfv; i++) {
        const int32_t fi = fv_cand->ij[2 * i], vi = fv_cand->ij[2 * i + 1];
        for (int j = 0; j < 3; j++) {
            const int32_t vj = F[3 * (size_t)fi + j];
            if (ipco_point_point_distance(vx(V, vi), vx(V, vj)) < active2) k[m++] = vvk(vi, vj);
        }
    }
    m = unique_keys(k, m);
    for (int64_t i = 0; i < m; i++) {
        const int32_t u[2] = { (int32_t)(k[i] >> 32), (int32_t)(k[i] & 0xFFFFFFFFu) };
        double w = 0.0;
        for (int s = 0; s < 2; s++) {
            const int32_t a = u[s], b = u[1 - s];
            if (M.bnd[b] || adj_has(&M.vv, b, a)) continue;
            w += aw ? 0.25 * va[a] : 1.0;
        }
        if (w != 0) push_vv(R, u[0], u[1], w);
    }
    free(k);
    mesh_adj_free(&M);
}

static int cmp_f64(const void *a, const void *b)
{
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

int ipco_build_collisions_ex(const double *V, int nV, const double *V_rest, const int32_t *E,
                             int nE, const int32_t *F, int nF, const int32_t *F2E,
                             const ipco_pairs *vv_cand, const ipco_pairs *ev_cand,
                             const ipco_pairs *ee_cand, const ipco_pairs *fv_cand, double dhat,
                             double dmin, const double *vertex_area, const double *edge_area,
                             int set_type, ipco_collisions *out)
{
    const int improved = set_type == 1, ogc = set_type == 2;
    memset(out, 0, sizeof(*out));
    out->dmin = dmin;
    const double off = dmin + dhat;
    const double active2 = off * off; /* normal_collisions.cpp:52-54 */
    const int64_t nfv = fv_cand ? fv_cand->n : 0;
    const int64_t nee = ee_cand ? ee_cand->n : 0;
    const int64_t nvv = vv_cand ? vv_cand->n : 0;
    const int64_t nev = ev_cand ? ev_cand->n : 0;
    const int64_t n0 = nfv + nee + nvv + nev;
    rec *R = (rec *)calloc((size_t)(n0 > 0 ? n0 : 1), sizeof(rec));
    rec *Rc = R + nee + nfv; /* codim candidates after the surface ones */
    const double *va = vertex_area, *ea = edge_area;
    const int aw = va != NULL;
    mesh_adj M;
    if (ogc) mesh_adj_build(nV, E, nE, F, nF, F2E, &M);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nvv; i++) {
        const int32_t a = vv_cand->ij[2 * i], b = vv_cand->ij[2 * i + 1];
        vv_candidate(V, a, b, active2, &Rc[i]);
        Rc[i].w = aw ? 0.5 * (va[a] + va[b]) : 1.0; /* builder.cpp:42-44 */
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nev; i++) {
        ev_candidate(V, E, ev_cand->ij[2 * i], ev_cand->ij[2 * i + 1], active2,
                     &Rc[nvv + i]);
        Rc[nvv + i].w = aw ? 0.5 * va[ev_cand->ij[2 * i + 1]] : 1.0; /* :81 */
        if (ogc) ogc_ev(V, E, &M, ev_cand->ij[2 * i], ev_cand->ij[2 * i + 1], &Rc[nvv + i]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nee; i++) {
        const int32_t a = ee_cand->ij[2 * i], b = ee_cand->ij[2 * i + 1];
        ee_candidate(V, V_rest, E, a, b, active2, &R[i]);
        R[i].w = aw ? 0.25 * (ea[a] + ea[b]) : 1.0; /* :181-183 */
        if (ogc) ogc_ee(V, E, &M, a, b, &R[i]);
    }
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nfv; i++) {
        fv_candidate(V, F, F2E, fv_cand->ij[2 * i], fv_cand->ij[2 * i + 1],
                     active2, &R[nee + i]);
        R[nee + i].w = aw ? 0.25 * va[fv_cand->ij[2 * i + 1]] : 1.0; /* :265 */
        if (ogc) ogc_fv(V, E, F, F2E, &M, fv_cand->ij[2 * i], fv_cand->ij[2 * i + 1], &R[nee + i]);
    }
    if (ogc) mesh_adj_free(&M);
    int64_t n = n0;
    if (improved) {
        recvec X = { NULL, 0, 0 };
        improved_records(V, V_rest, E, nE, F, nF, F2E, nV, ev_cand, ee_cand, fv_cand, active2,
                         va, ea, &X);
        R = (rec *)realloc(R, sizeof(rec) * (size_t)(n0 + X.n > 0 ? n0 + X.n : 1));
        if (X.n) memcpy(R + n0, X.r, sizeof(rec) * (size_t)X.n);
        n = n0 + X.n;
        free(X.r);
    }

    /* Per-type sort by key, then merge duplicates (normal_collisions_builder.cpp:
     * 547-600) summing their weights in ascending order of value; an
     * edge-edge collision keeps the ids, eps_x and orientation of its
     * lowest-rank contribution.  Collisions whose weights sum to zero are
     * dropped (normal_collisions_builder.cpp:660-684). */
    double *run = NULL;
    int64_t run_cap = 0;
    for (int t = 0; t < 4; t++) {
        int64_t m = 0;
        for (int64_t i = 0; i < n; i++) m += R[i].type == t;
        uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(m > 0 ? m : 1));
        int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
        int64_t j = 0;
        for (int64_t i = 0; i < n; i++)
            if (R[i].type == t) {
                key[j] = rec_key(&R[i]);
                idx[j++] = i;
            }
        ipco_sort_u64_pairs(key, idx, m);
        int64_t u = 0;
        for (int64_t i = 0; i < m; i++)
            if (i == 0 || key[i] != key[i - 1] || t == T_FV) u++;
        int32_t *ids = (int32_t *)malloc(sizeof(int32_t) * 2 * (size_t)(u > 0 ? u : 1));
        double *w = (double *)malloc(sizeof(double) * (size_t)(u > 0 ? u : 1));
        double *eps = NULL;
        int32_t *dts = NULL;
        if (t == T_EE) {
            eps = (double *)malloc(sizeof(double) * (size_t)(u > 0 ? u : 1));
            dts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(u > 0 ? u : 1));
        }
        int64_t k = 0;
        for (int64_t s = 0; s < m;) {
            int64_t e = s + 1;
            if (t != T_FV)
                while (e < m && key[e] == key[s]) e++;
            if (e - s > run_cap) {
                run_cap = 2 * (e - s);
                run = (double *)realloc(run, sizeof(double) * (size_t)run_cap);
            }
            const rec *rep = &R[idx[s]];
            for (int64_t i = s; i < e; i++) {
                run[i - s] = R[idx[i]].w;
                if (R[idx[i]].rank < rep->rank) rep = &R[idx[i]];
            }
            qsort(run, (size_t)(e - s), sizeof(double), cmp_f64);
            double sum = 0.0;
            for (int64_t i = 0; i < e - s; i++) sum += run[i];
            s = e;
            if (t != T_FV && sum == 0) continue;
            ids[2 * k] = rep->a;
            ids[2 * k + 1] = rep->b;
            w[k] = sum;
            if (t == T_EE) {
                eps[k] = rep->eps_x;
                dts[k] = rep->dtype;
            }
            k++;
        }
        u = k;
        free(key);
        free(idx);
        switch (t) {
        case T_VV: out->nvv = u; out->vv = ids; out->vv_w = w; break;
        case T_EV: out->nev = u; out->ev = ids; out->ev_w = w; break;
        case T_EE:
            out->nee = u;
            out->ee = ids;
            out->ee_w = w;
            out->ee_eps_x = eps;
            out->ee_dtype = dts;
            break;
        default: out->nfv = u; out->fv = ids; out->fv_w = w; break;
        }
    }
    free(run);
    free(R);
    return 0;
}

void ipco_collisions_free(ipco_collisions *c)
{
    free(c->vv); free(c->vv_w);
    free(c->ev); free(c->ev_w);
    free(c->ee); free(c->ee_w); free(c->ee_eps_x); free(c->ee_dtype);
    free(c->fv); free(c->fv_w);
    memset(c, 0, sizeof(*c));
}

/* ========================================================================= */
/* Barrier potential                                                         */
/* ========================================================================= */

typedef struct stencil {
    int kind;  /* ipco_stencil_distance_derivs kind */
    int dtype; /* fixed dtype used for E/grad/hess */
    int nv;
    int32_t ids[4];
    double w, eps_x; /* eps_x < 0: not mollified */
} stencil;

static int64_t n_collisions(const ipco_collisions *c)
{
    return c->nvv + c->nev + c->nee + c->nfv;
}

/* CollisionStencil::vertex_ids and known_dtype: VV (v0,v1) point-point;
 * EV (v, e0, e1) P_E; EE (ea0,ea1,eb0,eb1) stored dtype; FV (v,f0,f1,f2) P_T
 * (face_vertex.hpp:28-32, edge_vertex.hpp:28-32, edge_edge.hpp:96). */
static void get_stencil(const ipco_collisions *c, const int32_t *E,
                        const int32_t *F, int64_t i, stencil *s)
{
    s->eps_x = -1;
    if (i < c->nvv) {
        s->kind = 3; s->dtype = 0; s->nv = 2;
        s->ids[0] = c->vv[2 * i]; s->ids[1] = c->vv[2 * i + 1];
        s->w = c->vv_w[i];
        return;
    }
    i -= c->nvv;
    if (i < c->nev) {
        const int32_t e = c->ev[2 * i];
        s->kind = 2; s->dtype = IPCO_PE_P_E; s->nv = 3;
        s->ids[0] = c->ev[2 * i + 1];
        s->ids[1] = E[2 * (size_t)e]; s->ids[2] = E[2 * (size_t)e + 1];
        s->w = c->ev_w[i];
        return;
    }
    i -= c->nev;
    if (i < c->nee) {
        const int32_t ea = c->ee[2 * i], eb = c->ee[2 * i + 1];
        s->kind = 1; s->dtype = c->ee_dtype[i]; s->nv = 4;
        s->ids[0] = E[2 * (size_t)ea]; s->ids[1] = E[2 * (size_t)ea + 1];
        s->ids[2] = E[2 * (size_t)eb]; s->ids[3] = E[2 * (size_t)eb + 1];
        s->w = c->ee_w[i];
        s->eps_x = c->ee_eps_x[i];
        return;
    }
    i -= c->nee;
    const int32_t f = c->fv[2 * i];
    s->kind = 0; s->dtype = IPCO_PT_P_T; s->nv = 4;
    s->ids[0] = c->fv[2 * i + 1];
    s->ids[1] = F[3 * (size_t)f]; s->ids[2] = F[3 * (size_t)f + 1];
    s->ids[3] = F[3 * (size_t)f + 2];
    s->w = c->fv_w[i];
}

static void gather(const double *V, const stencil *s, double *x)
{
    for (int i = 0; i < s->nv; i++)
        for (int k = 0; k < 3; k++) x[3 * i + k] = V[3 * (size_t)s->ids[i] + k];
}

/* barrier_potential.cpp:62-99 */
static double f0(double d2, double dmin, double dhat, double kappa)
{
    return kappa * ipco_barrier(d2 - dmin * dmin, (2 * dmin + dhat) * dhat);
}
static double f1(double d2, double dmin, double dhat, double kappa)
{
    return kappa
        * ipco_barrier_first_derivative(d2 - dmin * dmin, (2 * dmin + dhat) * dhat);
}
static double f2(double d2, double dmin, double dhat, double kappa)
{
    return kappa
        * ipco_barrier_second_derivative(d2 - dmin * dmin, (2 * dmin + dhat) * dhat);
}

/* normal_potential.cpp:127-136 */
static double local_energy(const stencil *s, const double *x, double dmin,
                           double dhat, double kappa)
{
    const double d = ipco_stencil_distance_derivs(s->kind, s->dtype, x, NULL, NULL);
    const double m = s->eps_x >= 0 ? ipco_edge_edge_mollifier(x, s->eps_x, NULL, NULL) : 1.0;
    return s->w * m * f0(d, dmin, dhat, kappa);
}

double ipco_barrier_energy(const ipco_collisions *c, const double *V, int nV,
                           const int32_t *E, const int32_t *F, double dhat,
                           double kappa)
{
    (void)nV;
    const int64_t n = n_collisions(c);
    /* per-collision terms in parallel, summed in collision order: the
       result does not depend on the thread count */
    double *e = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!e) return NAN;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        e[i] = local_energy(&s, x, c->dmin, dhat, kappa);
    }
    double sum = 0;
    for (int64_t i = 0; i < n; i++) sum += e[i];
    free(e);
    return sum;
}

/* normal_potential.cpp:138-172 */
static void local_gradient(const stencil *s, const double *x, double dmin,
                           double dhat, double kappa, double *g)
{
    const int nd = 3 * s->nv;
    double gm[12];
    const double m = s->eps_x >= 0 ? ipco_edge_edge_mollifier(x, s->eps_x, gm, NULL) : 1.0;
    for (int i = 0; i < nd; i++) g[i] = 0;
    if (m <= 0) return;
    double gd[12];
    const double d = ipco_stencil_distance_derivs(s->kind, s->dtype, x, gd, NULL);
    const double f = f0(d, dmin, dhat, kappa), fp = f1(d, dmin, dhat, kappa);
    if (s->eps_x < 0) {
        for (int i = 0; i < nd; i++) g[i] = (s->w * fp) * gd[i];
        return;
    }
    for (int i = 0; i < nd; i++)
        g[i] = (s->w * f) * gm[i] + (s->w * m * fp) * gd[i];
}

void ipco_barrier_gradient(const ipco_collisions *c, const double *V, int nV,
                           const int32_t *E, const int32_t *F, double dhat,
                           double kappa, double *grad)
{
    const int64_t n = n_collisions(c);
    memset(grad, 0, sizeof(double) * 3 * (size_t)nV);
    /* serial scatter in collision order (the reference sums per-thread
     * copies; sums agree to rounding) */
    double *loc = (double *)malloc(sizeof(double) * 12 * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        local_gradient(&s, x, c->dmin, dhat, kappa, loc + 12 * i);
    }
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        get_stencil(c, E, F, i, &s);
        for (int a = 0; a < s.nv; a++)
            for (int k = 0; k < 3; k++)
                grad[3 * (size_t)s.ids[a] + k] += loc[12 * i + 3 * a + k];
    }
    free(loc);
}

/* normal_potential.cpp:174-232; H row-major nd x nd */
static int local_hessian(const stencil *s, const double *x, double dmin,
                         double dhat, double kappa, int psd, double *H)
{
    const int nd = 3 * s->nv;
    double gd[12], hdd[144], gm[12], hm[144];
    const double d = ipco_stencil_distance_derivs(s->kind, s->dtype, x, gd, hdd);
    const int moll = s->eps_x >= 0;
    const double m = moll ? ipco_edge_edge_mollifier(x, s->eps_x, gm, hm) : 1.0;
    const double w = s->w;
    if (moll && m <= 0) {
        const double f = f0(d, dmin, dhat, kappa);
        for (int i = 0; i < nd * nd; i++) H[i] = (w * f) * hm[i];
        return 0;
    }
    const double fp = f1(d, dmin, dhat, kappa), fpp = f2(d, dmin, dhat, kappa);
    if (!moll) {
        for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++)
                H[i * nd + j] = (w * fpp) * gd[i] * gd[j] + (w * fp) * hdd[i * nd + j];
    } else {
        const double f = f0(d, dmin, dhat, kappa);
        const double wm = w * m;
        for (int i = 0; i < nd; i++)
            for (int j = 0; j < nd; j++) {
                const double gfgm_ij = (w * fp) * gd[i] * gm[j];
                const double gfgm_ji = (w * fp) * gd[j] * gm[i];
                H[i * nd + j] = (w * f) * hm[i * nd + j] + gfgm_ij + gfgm_ji
                    + (wm * fpp) * gd[i] * gd[j] + (wm * fp) * hdd[i * nd + j];
            }
    }
    return ipco_project_to_psd(H, nd, psd);
}

/* local 3nv x 3nv matrices (144 doubles per collision) -> CSR: triplets,
 * skipping exact zeros (local_to_global.hpp:289-301), summed per entry;
 * keep (optional) selects the collisions assembled */
static void csr_from_local(const ipco_collisions *c, const int32_t *E, const int32_t *F, int nV,
                           const double *loc, const uint8_t *keep, ipco_csr *out)
{
    const int64_t n = n_collisions(c);
    const int64_t ndof = 3 * (int64_t)nV;
    memset(out, 0, sizeof(*out));
    out->n = (int32_t)ndof;
    int64_t nt = 0;
    for (int64_t i = 0; i < n; i++) {
        if (keep && !keep[i]) continue;
        stencil s;
        get_stencil(c, E, F, i, &s);
        const int nd = 3 * s.nv;
        for (int q = 0; q < nd * nd; q++) nt += loc[144 * i + q] != 0;
    }
    uint64_t *key = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(nt > 0 ? nt : 1));
    int64_t *src = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nt > 0 ? nt : 1));
    int64_t t = 0;
    for (int64_t i = 0; i < n; i++) {
        if (keep && !keep[i]) continue;
        stencil s;
        get_stencil(c, E, F, i, &s);
        const int nd = 3 * s.nv;
        for (int a = 0; a < s.nv; a++)
            for (int b = 0; b < s.nv; b++)
                for (int k = 0; k < 3; k++)
                    for (int l = 0; l < 3; l++) {
                        const int q = (3 * a + k) * nd + 3 * b + l;
                        if (loc[144 * i + q] == 0) continue;
                        const uint64_t row = 3 * (uint64_t)s.ids[a] + k;
                        const uint64_t col = 3 * (uint64_t)s.ids[b] + l;
                        key[t] = row * (uint64_t)ndof + col;
                        src[t++] = 144 * i + q;
                    }
    }
    ipco_sort_u64_pairs(key, src, nt);
    int64_t nnz = 0;
    for (int64_t i = 0; i < nt; i++) nnz += (i == 0 || key[i] != key[i - 1]);
    out->nnz = nnz;
    out->rowptr = (int64_t *)calloc((size_t)ndof + 1, sizeof(int64_t));
    out->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    out->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    int64_t k = -1;
    for (int64_t i = 0; i < nt; i++) {
        if (i == 0 || key[i] != key[i - 1]) {
            k++;
            const uint64_t row = key[i] / (uint64_t)ndof;
            out->col[k] = (int32_t)(key[i] % (uint64_t)ndof);
            out->val[k] = loc[src[i]];
            out->rowptr[row + 1]++;
        } else {
            out->val[k] += loc[src[i]];
        }
    }
    for (int64_t r = 0; r < ndof; r++) out->rowptr[r + 1] += out->rowptr[r];
    free(key);
    free(src);
}


int ipco_barrier_hessian(const ipco_collisions *c, const double *V, int nV,
                         const int32_t *E, const int32_t *F, double dhat,
                         double kappa, int psd, ipco_csr *out)
{
    const int64_t n = n_collisions(c);
    const int64_t ndof = 3 * (int64_t)nV;
    memset(out, 0, sizeof(*out));
    out->n = (int32_t)ndof;
    double *loc = (double *)malloc(sizeof(double) * 144 * (size_t)(n > 0 ? n : 1));
    int rc = 0;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        if (local_hessian(&s, x, c->dmin, dhat, kappa, psd, loc + 144 * i) != 0)
            rc = -1;
    }
    csr_from_local(c, E, F, nV, loc, NULL, out);
    free(loc);
    return rc;
}

void ipco_csr_free(ipco_csr *m)
{
    free(m->rowptr);
    free(m->col);
    free(m->val);
    memset(m, 0, sizeof(*m));
}

/* ---- Gauss-Newton helpers (normal_potential.cpp:15-83, 236-308) ----------
 * The distance vector t = sum_i c_i x_i of a collision, with c the
 * closest-point coefficients of its known distance type
 * (candidates/{vertex_vertex,edge_vertex,edge_edge,face_vertex}.cpp
 * compute_coefficients, collision_stencil.cpp:14-25).  Vectors of runtime
 * size (VectorMax3d) reduce left to right; fixed Vector3d dots use the
 * unrolled order a0 b0 + (a1 b1 + a2 b2). */
static double dot_seq(const double *a, const double *b)
{
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

/* closest_point.cpp:11-18 */
static double pe_closest(const double *p, const double *e0, const double *e1)
{
    double e[3], r[3];
    v3_sub(e1, e0, e);
    v3_sub(p, e0, r);
    return dot_seq(r, e) / dot_seq(e, e);
}

/* closest_point.cpp:65-87 */
static void ee_closest(const double *ea0, const double *ea1, const double *eb0,
                       const double *eb1, double x[2])
{
    double d[3], ea[3], eb[3];
    v3_sub(ea0, eb0, d);
    v3_sub(ea1, ea0, ea);
    v3_sub(eb1, eb0, eb);
    const double a01 = -v3_dot(eb, ea);
    ipco_ldlt2(v3_sqnorm(ea), a01, v3_sqnorm(eb), -v3_dot(d, ea), v3_dot(d, eb), x);
}

/* closest_point.cpp:123-137 */
static void pt_closest(const double *p, const double *t0, const double *t1,
                       const double *t2, double x[2])
{
    double b0[3], b1[3], r[3];
    v3_sub(t1, t0, b0);
    v3_sub(t2, t0, b1);
    v3_sub(p, t0, r);
    ipco_ldlt2(v3_dot(b0, b0), v3_dot(b1, b0), v3_dot(b1, b1), v3_dot(b0, r),
               v3_dot(b1, r), x);
}

static void coefficients(const stencil *s, const double *x, double *c)
{
    const double *p0 = x, *p1 = x + 3, *p2 = x + 6, *p3 = x + 9;
    double a, ab[2];
    switch (s->kind) {
    case 3: /* VV */
        c[0] = 1.0; c[1] = -1.0;
        return;
    case 2: /* EV, P_E */
        a = pe_closest(p0, p1, p2);
        c[0] = 1.0; c[1] = -1.0 + a; c[2] = -a;
        return;
    case 0: /* FV, P_T */
        pt_closest(p0, p1, p2, p3, ab);
        c[0] = 1.0; c[1] = -1.0 + ab[0] + ab[1]; c[2] = -ab[0]; c[3] = -ab[1];
        return;
    default: /* EE by stored type */
        switch (s->dtype) {
        case IPCO_EE_EA0_EB0: c[0] = 1; c[1] = 0; c[2] = -1; c[3] = 0; return;
        case IPCO_EE_EA0_EB1: c[0] = 1; c[1] = 0; c[2] = 0; c[3] = -1; return;
        case IPCO_EE_EA1_EB0: c[0] = 0; c[1] = 1; c[2] = -1; c[3] = 0; return;
        case IPCO_EE_EA1_EB1: c[0] = 0; c[1] = 1; c[2] = 0; c[3] = -1; return;
        case IPCO_EE_EA_EB0:
            a = pe_closest(p2, p0, p1);
            c[0] = 1.0 - a; c[1] = a; c[2] = -1.0; c[3] = 0;
            return;
        case IPCO_EE_EA_EB1:
            a = pe_closest(p3, p0, p1);
            c[0] = 1.0 - a; c[1] = a; c[2] = 0; c[3] = -1.0;
            return;
        case IPCO_EE_EA0_EB:
            a = pe_closest(p0, p2, p3);
            c[0] = 1.0; c[1] = 0; c[2] = -1.0 + a; c[3] = -a;
            return;
        case IPCO_EE_EA1_EB:
            a = pe_closest(p1, p2, p3);
            c[0] = 0; c[1] = 1.0; c[2] = -1.0 + a; c[3] = -a;
            return;
        default:
            ee_closest(p0, p1, p2, p3, ab);
            c[0] = 1.0 - ab[0]; c[1] = ab[0]; c[2] = -1.0 + ab[1]; c[3] = -ab[1];
        }
    }
}

/* sum_i c_i v_i, accumulated in point order */
static void combine(const double *c, const double *v, int nv, double *t)
{
    t[0] = t[1] = t[2] = 0.0;
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) t[k] += c[i] * v[3 * i + k];
}

double ipco_distance_vector(const ipco_collisions *c, const double *V, const int32_t *E,
                            const int32_t *F, int64_t i, double *t, double *coeffs)
{
    stencil s;
    double x[12], cc[4];
    get_stencil(c, E, F, i, &s);
    gather(V, &s, x);
    coefficients(&s, x, cc);
    combine(cc, x, s.nv, t);
    if (coeffs)
        for (int j = 0; j < s.nv; j++) coeffs[j] = cc[j];
    return s.nv;
}

void ipco_gauss_newton_diagonal(const ipco_collisions *c, const double *V, int nV,
                                const int32_t *E, const int32_t *F, double dhat,
                                double kappa, double *diag)
{
    const int64_t n = n_collisions(c);
    memset(diag, 0, sizeof(double) * 3 * (size_t)nV);
    double *loc = (double *)malloc(sizeof(double) * 12 * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12], cc[4], t[3];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        coefficients(&s, x, cc);
        combine(cc, x, s.nv, t);
        const double d2 = dot_seq(t, t);
        const double gf = f1(d2, c->dmin, dhat, kappa), hf = f2(d2, c->dmin, dhat, kappa);
        const double s1 = s.w * 4.0 * hf, s2 = s.w * 2.0 * gf;
        for (int a = 0; a < s.nv; a++)
            for (int k = 0; k < 3; k++) {
                const double c2 = cc[a] * cc[a];
                loc[12 * i + 3 * a + k] = s1 * (c2 * (t[k] * t[k])) + s2 * c2;
            }
    }
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        get_stencil(c, E, F, i, &s);
        for (int a = 0; a < s.nv; a++)
            for (int k = 0; k < 3; k++)
                diag[3 * (size_t)s.ids[a] + k] += loc[12 * i + 3 * a + k];
    }
    free(loc);
}

double ipco_gauss_newton_quadratic_form(const ipco_collisions *c, const double *V, int nV,
                                        const int32_t *E, const int32_t *F, double dhat,
                                        double kappa, const double *p)
{
    (void)nV;
    const int64_t n = n_collisions(c);
    /* summed in collision order, as the energy */
    double *e = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!e) return NAN;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12], px[12], cc[4], t[3], q[3];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        gather(p, &s, px);
        coefficients(&s, x, cc);
        combine(cc, x, s.nv, t);
        combine(cc, px, s.nv, q);
        const double d2 = dot_seq(t, t);
        const double gf = f1(d2, c->dmin, dhat, kappa), hf = f2(d2, c->dmin, dhat, kappa);
        const double qt = dot_seq(q, t);
        e[i] = s.w * (4.0 * hf * (qt * qt) + 2.0 * gf * dot_seq(q, q));
    }
    double sum = 0;
    for (int64_t i = 0; i < n; i++) sum += e[i];
    free(e);
    return sum;
}

/* ---- candidate-set helpers (candidates.cpp:294-408) --------------------- */
typedef struct cand_view {
    const ipco_pairs *l[4]; /* ee, fv, vv, ev */
} cand_view;

/* vertex ids of candidate j of list k: EE (ea0, ea1, eb0, eb1), FV
 * (v, f0, f1, f2), VV (a, b), EV (v, e0, e1) */
static int cand_vertices(const int32_t *E, const int32_t *F, int k, const int32_t *ij, int *v)
{
    switch (k) {
    case 0:
        v[0] = E[2 * (size_t)ij[0]]; v[1] = E[2 * (size_t)ij[0] + 1];
        v[2] = E[2 * (size_t)ij[1]]; v[3] = E[2 * (size_t)ij[1] + 1];
        return 4;
    case 1:
        v[0] = ij[1];
        v[1] = F[3 * (size_t)ij[0]]; v[2] = F[3 * (size_t)ij[0] + 1];
        v[3] = F[3 * (size_t)ij[0] + 2];
        return 4;
    case 2:
        v[0] = ij[0]; v[1] = ij[1];
        return 2;
    default:
        v[0] = ij[1];
        v[1] = E[2 * (size_t)ij[0]]; v[2] = E[2 * (size_t)ij[0] + 1];
        return 3;
    }
}

/* CollisionStencil::compute_distance with the type left to be classified */
static double cand_dist2(const double *V, int k, const int *v)
{
    const double *a = V + 3 * (size_t)v[0], *b = V + 3 * (size_t)v[1];
    switch (k) {
    case 0:
        return ipco_edge_edge_distance(a, b, V + 3 * (size_t)v[2], V + 3 * (size_t)v[3], -1);
    case 1:
        return ipco_point_triangle_distance(a, b, V + 3 * (size_t)v[2], V + 3 * (size_t)v[3], -1);
    case 2:
        return ipco_point_point_distance(a, b);
    default:
        return ipco_point_edge_distance(a, b, V + 3 * (size_t)v[2], -1);
    }
}

void ipco_per_vertex_safe_distances(const double *V, int nV, const int32_t *E,
                                    const int32_t *F, const ipco_pairs *ee,
                                    const ipco_pairs *fv, const ipco_pairs *vv,
                                    const ipco_pairs *ev, double inflation_radius,
                                    double min_distance, double *out)
{
    const ipco_pairs *L[4] = { ee, fv, vv, ev };
    for (int i = 0; i < nV; i++) out[i] = inflation_radius - min_distance;
    for (int k = 0; k < 4; k++) {
        if (!L[k]) continue;
        for (int64_t j = 0; j < L[k]->n; j++) {
            int v[4];
            const int nv = cand_vertices(E, F, k, L[k]->ij + 2 * j, v);
            const double d = sqrt(cand_dist2(V, k, v)) - min_distance;
            for (int a = 0; a < nv; a++)
                if (d < out[v[a]]) out[v[a]] = d;
        }
    }
}

double ipco_noncandidate_conservative_stepsize(const double *V0, const double *V1, int nV,
                                               const int32_t *E, const int32_t *F,
                                               const ipco_pairs *ee, const ipco_pairs *fv,
                                               const ipco_pairs *vv, const ipco_pairs *ev,
                                               double dhat)
{
    const ipco_pairs *L[4] = { ee, fv, vv, ev };
    int64_t total = 0;
    for (int k = 0; k < 4; k++) total += L[k] ? L[k]->n : 0;
    if (total == 0) return 1.0;
    unsigned char *mark = (unsigned char *)calloc((size_t)(nV > 0 ? nV : 1), 1);
    for (int k = 0; k < 4; k++) {
        if (!L[k]) continue;
        for (int64_t j = 0; j < L[k]->n; j++) {
            int v[4];
            const int nv = cand_vertices(E, F, k, L[k]->ij + 2 * j, v);
            for (int a = 0; a < nv; a++) mark[v[a]] = 1;
        }
    }
    double mx = 0.0;
    for (int i = 0; i < nV; i++) {
        if (!mark[i]) continue;
        double d[3];
        v3_sub(V1 + 3 * (size_t)i, V0 + 3 * (size_t)i, d);
        const double nrm = sqrt(dot_seq(d, d));
        mx = mx < nrm ? nrm : mx;
    }
    free(mark);
    return 0.5 * dhat / mx;
}


/* ======================================================================
 * Friction: TangentialCollisions::build over a normal collision set with
 * global friction coefficients (tangential_collisions.cpp:61-186 with
 * uniform per-vertex mu, so every blend returns it) and FrictionPotential
 * energy / gradient / Hessian (tangential_potential.cpp:156-331,
 * friction_potential.cpp, smooth_mu.cpp, smooth_friction_mollifier.cpp),
 * isotropic (mu_aniso = 1).
 * ==================================================================== */

/* smooth_friction_mollifier.cpp:8-53 */
static double sf_f0(double y, double ev)
{
    if (fabs(y) >= ev) return y;
    return y * y * (1 - y / (3 * ev)) / ev + ev / 3;
}
static double sf_f1(double y, double ev)
{
    if (fabs(y) >= ev) return 1;
    const double t = y / ev;
    return t * (2 - t);
}
static double sf_f2(double y, double ev)
{
    if (fabs(y) >= ev) return 0;
    return (2 - 2 * y / ev) / ev;
}
static double sf_f1_over_x(double y, double ev)
{
    if (fabs(y) >= ev) return 1 / y;
    return (2 - y / ev) / ev;
}
static double sf_f2_x_minus_f1_over_x3(double y, double ev)
{
    if (fabs(y) >= ev) return -1 / (y * y * y);
    return -1 / (y * ev * ev);
}

/* smooth_mu.cpp:10-112 */
static double smooth_mu(double y, double ms, double mk, double ev)
{
    if (ms == mk || fabs(y) >= ev) return mk;
    const double z = fabs(y) / ev;
    if (fabs(y) < 0.5 * ev) return 2 * (mk - ms) * z * z + ms;
    return -2 * (mk - ms) * (z * (z - 2) + 1) + mk;
}
static double smooth_mu_derivative(double y, double ms, double mk, double ev)
{
    if (ms == mk || fabs(y) >= ev) return 0;
    const double z = fabs(y) / ev;
    if (fabs(y) < 0.5 * ev) return 4 * (mk - ms) * z / ev;
    return -4 * (mk - ms) * (z - 1) / ev;
}
static double smooth_mu_f0(double y, double ms, double mk, double ev)
{
    if (ms == mk || fabs(y) >= ev) return mk * sf_f0(y, ev);
    const double dm = mk - ms;
    const double z = fabs(y) / ev;
    if (fabs(y) < 0.5 * ev)
        return y * z * (z * (z * (1 - 0.4 * z) * dm - ms / 3.0) + ms) + (9.0 / 16.0) * ev * mk
            - (11.0 / 48.0) * ev * ms;
    return y * z * (z * (z * (0.4 * z - 2) * dm + (3 * mk - (10.0 / 3.0) * ms)) + (2 * ms - mk))
        + 0.6 * ev * mk - (4.0 / 15.0) * ev * ms;
}
static double smooth_mu_f1(double y, double ms, double mk, double ev)
{
    return smooth_mu(y, ms, mk, ev) * sf_f1(y, ev);
}
static double smooth_mu_f2(double y, double ms, double mk, double ev)
{
    return smooth_mu_derivative(y, ms, mk, ev) * sf_f1(y, ev) + smooth_mu(y, ms, mk, ev) * sf_f2(y, ev);
}
static double smooth_mu_f1_over_x(double y, double ms, double mk, double ev)
{
    return smooth_mu(y, ms, mk, ev) * sf_f1_over_x(y, ev);
}
static double smooth_mu_f2_x_minus_f1_over_x3(double y, double ms, double mk, double ev)
{
    if (ms == mk || fabs(y) >= ev) return mk * sf_f2_x_minus_f1_over_x3(y, ev);
    const double dm = mk - ms;
    const double z = 1 / ev;
    if (fabs(y) < 0.5 * ev) return z * z * (z * (8 - 6 * y * z) * dm - ms / y);
    return z * z * (z * (6 * y * z - 16) * dm + (9 * mk - 10 * ms) / y);
}

double ipco_smooth_mu_fn(int which, double y, double ms, double mk, double ev)
{
    switch (which) {
    case 0: return smooth_mu(y, ms, mk, ev);
    case 1: return smooth_mu_derivative(y, ms, mk, ev);
    case 2: return smooth_mu_f0(y, ms, mk, ev);
    case 3: return smooth_mu_f1(y, ms, mk, ev);
    case 4: return smooth_mu_f2(y, ms, mk, ev);
    case 5: return smooth_mu_f1_over_x(y, ms, mk, ev);
    case 6: return smooth_mu_f2_x_minus_f1_over_x3(y, ms, mk, ev);
    case 7: return sf_f0(y, ev);
    case 8: return sf_f1(y, ev);
    case 9: return sf_f2(y, ev);
    case 10: return sf_f1_over_x(y, ev);
    default: return sf_f2_x_minus_f1_over_x3(y, ev);
    }
}

/* Eigen normalized(): divide by the norm when it is positive */
static void v3_normalize(const double *a, double *o)
{
    const double z = v3_sqnorm(a);
    const double n = z > 0 ? sqrt(z) : 1.0;
    for (int k = 0; k < 3; k++) o[k] = z > 0 ? a[k] / n : a[k];
}

/* tangent_basis.cpp:17-175; P column-major 3 x 2 */
static void tangent_basis(int kind, const double *x, double *P)
{
    double a[3], b[3], c[3];
    if (kind == 3) { /* point-point */
        const double *p0 = x, *p1 = x + 3;
        double d[3], cx[3], cy[3];
        const double ux[3] = { 1, 0, 0 }, uy[3] = { 0, 1, 0 };
        v3_sub(p1, p0, d);
        v3_cross(ux, d, cx);
        v3_cross(uy, d, cy);
        const double *w = v3_sqnorm(cx) > v3_sqnorm(cy) ? cx : cy;
        v3_normalize(w, P);
        v3_cross(d, w, c);
        v3_normalize(c, P + 3);
        return;
    }
    if (kind == 2) { /* point-edge: p, e0, e1 */
        v3_sub(x + 6, x + 3, a);
        v3_sub(x, x + 3, b);
        v3_normalize(a, P);
        v3_cross(a, b, c);
        v3_normalize(c, P + 3);
        return;
    }
    /* edge-edge (ea0, ea1, eb0, eb1) or point-triangle (p, t0, t1, t2):
     * along the first edge, then normal x edge */
    if (kind == 1) {
        v3_sub(x + 3, x, a);
        v3_sub(x + 9, x + 6, b);
    } else {
        v3_sub(x + 6, x + 3, a);
        v3_sub(x + 9, x + 3, b);
    }
    double nrm[3];
    v3_cross(a, b, nrm);
    v3_normalize(a, P);
    v3_cross(nrm, a, c);
    v3_normalize(c, P + 3);
}

/* closest-point coordinates (closest_point.cpp:11-137) and the
 * relative-velocity Jacobian coefficients Gamma (relative_velocity.cpp) */
static void friction_coords(int kind, const double *x, double *beta, double *coef)
{
    beta[0] = beta[1] = 0;
    coef[0] = coef[1] = coef[2] = coef[3] = 0;
    switch (kind) {
    case 3: coef[0] = 1; coef[1] = -1; break;
    case 2:
        beta[0] = pe_closest(x, x + 3, x + 6);
        coef[0] = 1; coef[1] = beta[0] - 1; coef[2] = -beta[0];
        break;
    case 1:
        ee_closest(x, x + 3, x + 6, x + 9, beta);
        coef[0] = 1 - beta[0]; coef[1] = beta[0]; coef[2] = beta[1] - 1; coef[3] = -beta[1];
        break;
    default:
        pt_closest(x, x + 3, x + 6, x + 9, beta);
        coef[0] = 1; coef[1] = beta[0] + beta[1] - 1; coef[2] = -beta[0]; coef[3] = -beta[1];
    }
}

int ipco_friction_build(const ipco_collisions *c, const double *V, int nV, const int32_t *E,
                        const int32_t *F, double dhat, double kappa, double mu_s, double mu_k,
                        ipco_friction *out)
{
    double *ms = (double *)malloc(sizeof(double) * (size_t)(nV > 0 ? nV : 1));
    double *mk = (double *)malloc(sizeof(double) * (size_t)(nV > 0 ? nV : 1));
    for (int i = 0; i < nV; i++) {
        ms[i] = mu_s;
        mk[i] = mu_k;
    }
    const int rc = ipco_friction_build_pv(c, V, nV, E, F, dhat, kappa, ms, mk, out);
    free(ms);
    free(mk);
    return rc;
}

/* tangential_collisions.hpp:143-149 default_blend_mu */
static double blend_mu(double a, double b)
{
    return (a + b) / 2;
}

/* tangential_collisions.cpp:86-166: the coefficient at the closest point
 * of each side, blended */
static double collision_mu(const stencil *s, const double *beta, const double *m)
{
    const int32_t *v = s->ids;
    switch (s->kind) {
    case 3: return blend_mu(m[v[0]], m[v[1]]);
    case 2: return blend_mu((m[v[2]] - m[v[1]]) * beta[0] + m[v[1]], m[v[0]]);
    case 1:
        return blend_mu((m[v[1]] - m[v[0]]) * beta[0] + m[v[0]],
                        (m[v[3]] - m[v[2]]) * beta[1] + m[v[2]]);
    default:
        return blend_mu(m[v[1]] + beta[0] * (m[v[2]] - m[v[1]]) + beta[1] * (m[v[3]] - m[v[1]]),
                        m[v[0]]);
    }
}

int ipco_friction_build_pv(const ipco_collisions *c, const double *V, int nV, const int32_t *E,
                           const int32_t *F, double dhat, double kappa, const double *mu_s,
                           const double *mu_k, ipco_friction *out)
{
    (void)nV;
    const int64_t n = n_collisions(c);
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->mu_s = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    out->mu_k = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    out->mu_aniso = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
    out->mu_s_aniso = (double *)calloc(2 * (size_t)(n > 0 ? n : 1), sizeof(double));
    out->mu_k_aniso = (double *)calloc(2 * (size_t)(n > 0 ? n : 1), sizeof(double));
    out->mu_s_eff = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    out->mu_k_eff = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    for (int64_t i = 0; i < 2 * n; i++) out->mu_aniso[i] = 1.0;
    out->keep = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
    out->beta = (double *)calloc(2 * (size_t)(n > 0 ? n : 1), sizeof(double));
    out->basis = (double *)calloc(6 * (size_t)(n > 0 ? n : 1), sizeof(double));
    out->normal_force = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
    out->weight = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        stencil s;
        double x[12], coef[4];
        get_stencil(c, E, F, i, &s);
        gather(V, &s, x);
        double d2;
        if (s.kind == 1) {
            /* near-parallel edge pairs are skipped (tangential_collisions.cpp:
             * 118-121); the rest are line-line (edge_edge.hpp:27-31) */
            if (ipco_edge_edge_cross_squarednorm(x, x + 3, x + 6, x + 9) < s.eps_x) continue;
            d2 = ipco_edge_edge_distance(x, x + 3, x + 6, x + 9, IPCO_EE_EA_EB);
        } else if (s.kind == 3) {
            d2 = ipco_point_point_distance(x, x + 3);
        } else if (s.kind == 2) {
            d2 = ipco_point_edge_distance(x, x + 3, x + 6,
                                          ipco_point_edge_distance_type(x, x + 3, x + 6));
        } else {
            d2 = ipco_point_triangle_distance(
                x, x + 3, x + 6, x + 9, ipco_point_triangle_distance_type(x, x + 3, x + 6, x + 9));
        }
        out->keep[i] = 1;
        friction_coords(s.kind, x, out->beta + 2 * i, coef);
        tangent_basis(s.kind, x, out->basis + 6 * i);
        /* barrier_force_magnitude.cpp:5-15 */
        out->normal_force[i] = -f1(d2, c->dmin, dhat, kappa) * 2 * sqrt(d2);
        out->weight[i] = s.w;
        out->mu_s[i] = collision_mu(&s, out->beta + 2 * i, mu_s);
        out->mu_k[i] = collision_mu(&s, out->beta + 2 * i, mu_k);
        out->mu_s_eff[i] = out->mu_s[i];
        out->mu_k_eff[i] = out->mu_k[i];
    }
    return 0;
}

/* smooth_mu.cpp:115-171 */
void ipco_anisotropic_mu_eff(const double ta[2], const double msa[2], const double mka[2],
                             double mu_s, double mu_k, double out[2])
{
    const int use_s = msa[0] * msa[0] + msa[1] * msa[1] > 0;
    const int use_k = mka[0] * mka[0] + mka[1] * mka[1] > 0;
    out[0] = mu_s;
    out[1] = mu_k;
    if (!use_s && !use_k) return;
    double d[2] = { 1.0, 0.0 };
    const double n = sqrt(ta[0] * ta[0] + ta[1] * ta[1]);
    if (!(n < 1e-10)) {
        d[0] = ta[0] / n;
        d[1] = ta[1] / n;
    }
    if (use_s) {
        const double a = msa[0] * d[0], b = msa[1] * d[1];
        out[0] = sqrt(a * a + b * b);
    }
    if (use_k) {
        const double a = mka[0] * d[0], b = mka[1] * d[1];
        out[1] = sqrt(a * a + b * b);
    }
}

void ipco_friction_set_anisotropy(ipco_friction *f, const double *mu_aniso,
                                  const double *mu_s_aniso, const double *mu_k_aniso)
{
    for (int64_t i = 0; i < 2 * f->n; i++) {
        if (mu_aniso) f->mu_aniso[i] = mu_aniso[i];
        if (mu_s_aniso) f->mu_s_aniso[i] = mu_s_aniso[i];
        if (mu_k_aniso) f->mu_k_aniso[i] = mu_k_aniso[i];
    }
    for (int64_t i = 0; i < f->n; i++) {
        f->mu_s_eff[i] = f->mu_s[i];
        f->mu_k_eff[i] = f->mu_k[i];
    }
}

void ipco_friction_update_lagged_mu(ipco_friction *f, const ipco_collisions *c, const double *Vel,
                                    const int32_t *E, const int32_t *F)
{
    for (int64_t i = 0; i < f->n; i++) {
        f->mu_s_eff[i] = f->mu_s[i];
        f->mu_k_eff[i] = f->mu_k[i];
        if (!f->keep[i]) continue;
        const double *msa = f->mu_s_aniso + 2 * i;
        if (msa[0] * msa[0] + msa[1] * msa[1] == 0) continue;
        stencil s;
        double v[12], coef[4], tau[2] = { 0, 0 };
        get_stencil(c, E, F, i, &s);
        for (int a = 0; a < 4; a++)
            for (int k = 0; k < 3; k++) v[3 * a + k] = a < s.nv ? Vel[3 * (size_t)s.ids[a] + k] : 0.0;
        const double *beta = f->beta + 2 * i, *P = f->basis + 6 * i;
        switch (s.kind) {
        case 3: coef[0] = 1; coef[1] = -1; coef[2] = coef[3] = 0; break;
        case 2: coef[0] = 1; coef[1] = beta[0] - 1; coef[2] = -beta[0]; coef[3] = 0; break;
        case 1: coef[0] = 1 - beta[0]; coef[1] = beta[0]; coef[2] = beta[1] - 1; coef[3] = -beta[1]; break;
        default: coef[0] = 1; coef[1] = beta[0] + beta[1] - 1; coef[2] = -beta[0]; coef[3] = -beta[1];
        }
        /* tau = T^T v with T = Gamma^T P */
        for (int a = 0; a < s.nv; a++)
            for (int k = 0; k < 3; k++) {
                tau[0] += coef[a] * P[k] * v[3 * a + k];
                tau[1] += coef[a] * P[3 + k] * v[3 * a + k];
            }
        const double ta[2] = { f->mu_aniso[2 * i] * tau[0], f->mu_aniso[2 * i + 1] * tau[1] };
        double out[2];
        ipco_anisotropic_mu_eff(ta, msa, f->mu_k_aniso + 2 * i, f->mu_s[i], f->mu_k[i], out);
        f->mu_s_eff[i] = out[0];
        f->mu_k_eff[i] = out[1];
    }
}

void ipco_friction_free(ipco_friction *f)
{
    free(f->keep);
    free(f->beta);
    free(f->basis);
    free(f->normal_force);
    free(f->weight);
    free(f->mu_s);
    free(f->mu_k);
    free(f->mu_aniso);
    free(f->mu_s_aniso);
    free(f->mu_k_aniso);
    free(f->mu_s_eff);
    free(f->mu_k_eff);
    memset(f, 0, sizeof(*f));
}

/* relative velocity of the closest points (relative_velocity.cpp) */
static void relative_velocity(int kind, const double *v, const double *beta, double *r)
{
    for (int k = 0; k < 3; k++) {
        switch (kind) {
        case 3: r[k] = v[k] - v[3 + k]; break;
        case 2: r[k] = v[k] - ((v[6 + k] - v[3 + k]) * beta[0] + v[3 + k]); break;
        case 1:
            r[k] = ((v[3 + k] - v[k]) * beta[0] + v[k]) - ((v[9 + k] - v[6 + k]) * beta[1] + v[6 + k]);
            break;
        default:
            r[k] = v[k] - ((v[3 + k] + beta[0] * (v[6 + k] - v[3 + k])) + beta[1] * (v[9 + k] - v[3 + k]));
        }
    }
}

/* u = P^T r */
static void tangent_project(const double *P, const double *r, double *u)
{
    u[0] = v3_dot(P, r);
    u[1] = v3_dot(P + 3, r);
}

static void friction_local(const ipco_friction *fr, const stencil *s, int64_t i, const double *v,
                           double eps_v, double *energy, double *grad, double *hess, int psd)
{
    double r[3], u[2], coef[4], beta2[2];
    const double *P = fr->basis + 6 * i, *beta = fr->beta + 2 * i;
    relative_velocity(s->kind, v, beta, r);
    tangent_project(P, r, u);
    /* tangential_potential.cpp:170-181: u_aniso = mu_aniso (.) u, T_aniso =
     * T diag(mu_aniso), the lagged effective coefficients */
    const double an[2] = { fr->mu_aniso[2 * i], fr->mu_aniso[2 * i + 1] };
    u[0] = an[0] * u[0];
    u[1] = an[1] * u[1];
    const double norm = sqrt(u[0] * u[0] + u[1] * u[1]);
    const double scale = fr->weight[i] * fr->normal_force[i];
    const double ms = fr->mu_s_eff[i], mk = fr->mu_k_eff[i];
    (void)beta2;
    /* Gamma coefficients from the stored coordinates */
    switch (s->kind) {
    case 3: coef[0] = 1; coef[1] = -1; coef[2] = coef[3] = 0; break;
    case 2: coef[0] = 1; coef[1] = beta[0] - 1; coef[2] = -beta[0]; coef[3] = 0; break;
    case 1: coef[0] = 1 - beta[0]; coef[1] = beta[0]; coef[2] = beta[1] - 1; coef[3] = -beta[1]; break;
    default: coef[0] = 1; coef[1] = beta[0] + beta[1] - 1; coef[2] = -beta[0]; coef[3] = -beta[1];
    }
    if (energy) *energy = scale * smooth_mu_f0(norm, ms, mk, eps_v);
    const int nd = 3 * s->nv;
    if (grad) {
        const double f1x = smooth_mu_f1_over_x(norm, ms, mk, eps_v);
        const double g0 = f1x * u[0] * scale, g1 = f1x * u[1] * scale;
        for (int a = 0; a < s->nv; a++)
            for (int k = 0; k < 3; k++)
                grad[3 * a + k] = (coef[a] * P[k] * an[0]) * g0 + (coef[a] * P[3 + k] * an[1]) * g1;
    }
    if (hess) {
        /* inner 2 x 2 matrix M, H = T M T^T with T[3a + k][c] = coef_a P[k][c] */
        double M[4] = { 0, 0, 0, 0 };
        const double f1x = smooth_mu_f1_over_x(norm, ms, mk, eps_v);
        if (norm > eps_v) { /* dynamic: f2 = 0 */
            if (!(psd != 0 && scale <= 0)) {
                const double up[2] = { -u[1], u[0] };
                const double cc = scale * f1x / (norm * norm);
                M[0] = cc * up[0] * up[0];
                M[1] = M[2] = cc * up[0] * up[1];
                M[3] = cc * up[1] * up[1];
            }
        } else if (norm == 0) {
            if (!(psd != 0 && scale <= 0)) M[0] = M[3] = scale * f1x;
        } else {
            const double f2 = smooth_mu_f2_x_minus_f1_over_x3(norm, ms, mk, eps_v);
            M[0] = (f2 * u[0] * u[0] + f1x) * scale;
            M[1] = M[2] = (f2 * u[0] * u[1]) * scale;
            M[3] = (f2 * u[1] * u[1] + f1x) * scale;
            ipco_project_to_psd(M, 2, psd);
        }
        for (int a = 0; a < nd; a++)
            for (int b = 0; b < nd; b++) {
                const double ta0 = coef[a / 3] * P[a % 3] * an[0], ta1 = coef[a / 3] * P[3 + a % 3] * an[1];
                const double tb0 = coef[b / 3] * P[b % 3] * an[0], tb1 = coef[b / 3] * P[3 + b % 3] * an[1];
                hess[a * nd + b] = ta0 * (M[0] * tb0 + M[1] * tb1) + ta1 * (M[2] * tb0 + M[3] * tb1);
            }
    }
}

static void gather_velocities(const double *Vel, const stencil *s, double *v)
{
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) v[3 * i + k] = i < s->nv ? Vel[3 * (size_t)s->ids[i] + k] : 0.0;
}

double ipco_friction_energy(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                            int nV, const int32_t *E, const int32_t *F, double eps_v)
{
    (void)nV;
    double e = 0;
    for (int64_t i = 0; i < fr->n; i++) {
        if (!fr->keep[i]) continue;
        stencil s;
        double v[12], ei;
        get_stencil(c, E, F, i, &s);
        gather_velocities(Vel, &s, v);
        friction_local(fr, &s, i, v, eps_v, &ei, NULL, NULL, 0);
        e += ei;
    }
    return e;
}

void ipco_friction_gradient(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                            int nV, const int32_t *E, const int32_t *F, double eps_v, double *grad)
{
    memset(grad, 0, sizeof(double) * 3 * (size_t)nV);
    for (int64_t i = 0; i < fr->n; i++) {
        if (!fr->keep[i]) continue;
        stencil s;
        double v[12], g[12];
        get_stencil(c, E, F, i, &s);
        gather_velocities(Vel, &s, v);
        friction_local(fr, &s, i, v, eps_v, NULL, g, NULL, 0);
        for (int a = 0; a < s.nv; a++)
            for (int k = 0; k < 3; k++) grad[3 * (size_t)s.ids[a] + k] += g[3 * a + k];
    }
}

int ipco_friction_hessian(const ipco_friction *fr, const ipco_collisions *c, const double *Vel,
                          int nV, const int32_t *E, const int32_t *F, double eps_v, int psd,
                          ipco_csr *out)
{
    const int64_t n = fr->n;
    double *loc = (double *)calloc(144 * (size_t)(n > 0 ? n : 1), sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        if (!fr->keep[i]) continue;
        stencil s;
        double v[12];
        get_stencil(c, E, F, i, &s);
        gather_velocities(Vel, &s, v);
        friction_local(fr, &s, i, v, eps_v, NULL, NULL, loc + 144 * i, psd);
    }
    csr_from_local(c, E, F, nV, loc, fr->keep, out);
    free(loc);
    return 0;
}
