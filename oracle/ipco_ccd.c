/* Tight-Inclusion CCD and the IPC Toolkit CCD strategy, restated.
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h).
 *
 * Tight-Inclusion v1.0.6 (third party, absent from the reference tree) is
 * restated from Wang, Ferguson, Schneider, Panozzo, ACM TOG 40(5) 2021:
 * breadth-first interval bisection over (t,u,v), an inclusion function
 * evaluated at the 8 corners of each (t,u,v) box (the functions are
 * multilinear, so the corner box is tight), an epsilon box from the
 * published floating-point filter, and the "level" termination rule.
 * Within a level boxes are visited by increasing lower t; equal keys are
 * visited first-in first-out. */
#include "ipco_internal.h"

#include <limits.h>
#include <stdio.h>

/* The no_zero_toi retry loop is capped at 64 halvings here and in the HIP
 * kernels (Tight-Inclusion allows INT_MAX); a level cap of 1024 keeps every
 * search finite. */

/* work counters of the last root finder call on this thread (test probes) */
static __thread long tl_refine, tl_max_queue;

typedef struct tbox {
    double t0, t1, u0, u1, v0, v1;
    int level;
    int64_t seq;
} tbox;

typedef struct heap {
    tbox *a;
    int64_t n, cap, seq;
} heap;

static int before(const tbox *x, const tbox *y)
{
    if (x->level != y->level) return x->level < y->level;
    if (x->t0 != y->t0) return x->t0 < y->t0;
    return x->seq < y->seq;
}

static void heap_push(heap *h, tbox b)
{
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 64;
        h->a = (tbox *)realloc(h->a, sizeof(tbox) * (size_t)h->cap);
    }
    b.seq = h->seq++;
    int64_t i = h->n++;
    h->a[i] = b;
    while (i > 0) {
        const int64_t p = (i - 1) / 2;
        if (!before(&h->a[i], &h->a[p])) break;
        tbox t = h->a[i];
        h->a[i] = h->a[p];
        h->a[p] = t;
        i = p;
    }
}

static tbox heap_pop(heap *h)
{
    tbox top = h->a[0];
    h->a[0] = h->a[--h->n];
    int64_t i = 0;
    for (;;) {
        const int64_t l = 2 * i + 1, r = l + 1;
        int64_t m = i;
        if (l < h->n && before(&h->a[l], &h->a[m])) m = l;
        if (r < h->n && before(&h->a[r], &h->a[m])) m = r;
        if (m == i) break;
        tbox t = h->a[i];
        h->a[i] = h->a[m];
        h->a[m] = t;
        i = m;
    }
    return top;
}

/* F(t,u,v) for one coordinate at one corner.
 * edge-edge: ea(t,u) - eb(t,v); vertex-face: p(t) - tri(t,u,v). */
static double f_ee(const double *s, const double *e, int d, double t, double u,
                   double v)
{
    const double a0 = (e[0 + d] - s[0 + d]) * t + s[0 + d];
    const double a1 = (e[3 + d] - s[3 + d]) * t + s[3 + d];
    const double b0 = (e[6 + d] - s[6 + d]) * t + s[6 + d];
    const double b1 = (e[9 + d] - s[9 + d]) * t + s[9 + d];
    const double va = (a1 - a0) * u + a0;
    const double vb = (b1 - b0) * v + b0;
    return va - vb;
}

static double f_vf(const double *s, const double *e, int d, double t, double u,
                   double v)
{
    const double p = (e[0 + d] - s[0 + d]) * t + s[0 + d];
    const double t0 = (e[3 + d] - s[3 + d]) * t + s[3 + d];
    const double t1 = (e[6 + d] - s[6 + d]) * t + s[6 + d];
    const double t2 = (e[9 + d] - s[9 + d]) * t + s[9 + d];
    const double pt = (t1 - t0) * u + (t2 - t0) * v + t0;
    return p - pt;
}

/* Inclusion test of the origin in the corner box of F over b, inflated by
 * err (per axis) and ms.  Returns zero_in; box_in set when the whole box
 * lies within the inflated epsilon box; true_tol = per-axis box widths. */
static int origin_in_inclusion(int is_vf, const double *s, const double *e,
                               const tbox *b, const double *err, double ms,
                               int *box_in, double *true_tol)
{
    *box_in = 1;
    for (int d = 0; d < 3; d++) {
        double mn = 0, mx = 0;
        for (int c = 0; c < 8; c++) {
            const double t = (c & 4) ? b->t1 : b->t0;
            const double u = (c & 2) ? b->u1 : b->u0;
            const double v = (c & 1) ? b->v1 : b->v0;
            const double val = is_vf ? f_vf(s, e, d, t, u, v) : f_ee(s, e, d, t, u, v);
            if (c == 0 || val < mn) mn = val;
            if (c == 0 || val > mx) mx = val;
        }
        true_tol[d] = mx - mn;
        const double eps = err[d];
        if (mn - ms > eps || mx + ms < -eps) return 0;
        if (!(mn + ms >= -eps && mx - ms <= eps)) *box_in = 0;
    }
    return 1;
}

static double linf_diff(const double *a, const double *b)
{
    double m = 0;
    for (int k = 0; k < 3; k++) {
        const double v = fabs(b[k] - a[k]);
        if (v > m) m = v;
    }
    return m;
}

/* parameter-space tolerances: delta / (3 max|dF/dparam| in L-inf over the
 * corners of the unit (t,u,v) cube). */
static void param_tolerances(int is_vf, const double *s, const double *e,
                             double delta, double tol[3])
{
    double p[8][3]; /* p[tuv bits] */
    for (int k = 0; k < 3; k++) {
        if (is_vf) {
            const double vs = s[k], f0s = s[3 + k], f1s = s[6 + k], f2s = s[9 + k];
            const double ve = e[k], f0e = e[3 + k], f1e = e[6 + k], f2e = e[9 + k];
            p[0][k] = vs - f0s;                 /* 000 */
            p[1][k] = vs - f2s;                 /* 001 */
            p[3][k] = vs - (f1s + f2s - f0s);   /* 011 */
            p[2][k] = vs - f1s;                 /* 010 */
            p[4][k] = ve - f0e;                 /* 100 */
            p[5][k] = ve - f2e;                 /* 101 */
            p[7][k] = ve - (f1e + f2e - f0e);   /* 111 */
            p[6][k] = ve - f1e;                 /* 110 */
        } else {
            const double a0s = s[k], a1s = s[3 + k], b0s = s[6 + k], b1s = s[9 + k];
            const double a0e = e[k], a1e = e[3 + k], b0e = e[6 + k], b1e = e[9 + k];
            p[0][k] = a0s - b0s;
            p[1][k] = a0s - b1s;
            p[3][k] = a1s - b1s;
            p[2][k] = a1s - b0s;
            p[4][k] = a0e - b0e;
            p[5][k] = a0e - b1e;
            p[7][k] = a1e - b1e;
            p[6][k] = a1e - b0e;
        }
    }
    /* along t: (000,100) (001,101) (011,111) (010,110) */
    double dl = linf_diff(p[0], p[4]);
    double x;
    if ((x = linf_diff(p[1], p[5])) > dl) dl = x;
    if ((x = linf_diff(p[3], p[7])) > dl) dl = x;
    if ((x = linf_diff(p[2], p[6])) > dl) dl = x;
    /* along u: (000,010) (100,110) (101,111) (001,011) */
    double lu = linf_diff(p[0], p[2]);
    if ((x = linf_diff(p[4], p[6])) > lu) lu = x;
    if ((x = linf_diff(p[5], p[7])) > lu) lu = x;
    if ((x = linf_diff(p[1], p[3])) > lu) lu = x;
    /* along v: (000,001) (100,101) (110,111) (010,011) */
    double lv = linf_diff(p[0], p[1]);
    if ((x = linf_diff(p[4], p[5])) > lv) lv = x;
    if ((x = linf_diff(p[6], p[7])) > lv) lv = x;
    if ((x = linf_diff(p[2], p[3])) > lv) lv = x;
    tol[0] = delta / (3 * dl);
    tol[1] = delta / (3 * lu);
    tol[2] = delta / (3 * lv);
}

/* floating-point filter: max(|x|,1)^3 * c per axis over the 8 points */
static void numerical_error(int is_vf, const double *s, const double *e,
                            int use_ms, double err[3])
{
    const double c = is_vf ? (use_ms ? 7.549516567451064e-15 : 6.661338147750939e-15)
                           : (use_ms ? 7.105427357601002e-15 : 6.217248937900877e-15);
    for (int k = 0; k < 3; k++) {
        double m = fabs(s[k]);
        for (int i = 0; i < 4; i++) {
            if (fabs(s[3 * i + k]) > m) m = fabs(s[3 * i + k]);
            if (fabs(e[3 * i + k]) > m) m = fabs(e[3 * i + k]);
        }
        const double dlt = m > 1 ? m : 1;
        err[k] = dlt * dlt * dlt * c;
    }
}

/* split the axis with the largest width/tol among those above tol; first
 * index on ties */
static int next_split(const double w[3], const double tol[3])
{
    int best = 0;
    double bv = -INFINITY;
    for (int i = 0; i < 3; i++) {
        const double val = w[i] > tol[i] ? w[i] / tol[i] : -INFINITY;
        if (val > bv) {
            bv = val;
            best = i;
        }
    }
    return best;
}

/* skip_from: boxes whose lower t is >= skip_from are dropped unvisited.
 * Because boxes of a level are visited by increasing lower t, this leaves
 * every outcome below skip_from unchanged (it only hides outcomes >=
 * skip_from); +inf reproduces the plain algorithm. */
static int ti_bfs(int is_vf, const double *s, const double *e,
                  const double tol[3], double co_tol, const double err[3],
                  double ms, double max_time, long max_itr, double skip_from,
                  double *toi, double *out_tol)
{
    heap h = { 0 };
    *out_tol = co_tol;
    double temp_out_tol = co_tol, temp_toi = INFINITY;
    double toi_skip = 4;
    int use_skip = 0, this_level_less_tol = 1, find_level_root = 0;
    int current_level = -2, box_in_level = -2;
    long refine = 0;
    *toi = INFINITY;
    tbox root = { 0, max_time, 0, 1, 0, 1, -1, 0 };
    heap_push(&h, root);
    int ret = 0;
    while (h.n > 0) {
        if (h.n > tl_max_queue) tl_max_queue = h.n;
        if (h.a[0].level >= 1023) break; /* level cap (see top of file) */
        const tbox b = heap_pop(&h);
        if (b.t0 >= toi_skip || b.t0 >= skip_from) continue;
        if (box_in_level != b.level) {
            box_in_level = b.level;
            this_level_less_tol = 1;
        }
        refine++;
        tl_refine++;
        int box_in;
        double tt[3];
        if (!origin_in_inclusion(is_vf, s, e, &b, err, ms, &box_in, tt)) continue;
        const double w[3] = { b.t1 - b.t0, b.u1 - b.u0, b.v1 - b.v0 };
        const int tol_cond = tt[0] <= co_tol && tt[1] <= co_tol && tt[2] <= co_tol;
        const int c1 = w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2];
        const int c2 = box_in && this_level_less_tol;
        if (!tol_cond) this_level_less_tol = 0;
        const int c3 = this_level_less_tol;
        if (c1 || c2 || c3) {
            *toi = b.t0;
            ret = 1;
            goto done;
        }
        if (max_itr > 0) {
            if (current_level != b.level) {
                current_level = b.level;
                find_level_root = 0;
            }
            if (!find_level_root) {
                temp_toi = b.t0;
                double m = co_tol;
                for (int k = 0; k < 3; k++) if (tt[k] > m) m = tt[k];
                temp_out_tol = m;
                find_level_root = 1;
            }
            if (refine > max_itr) {
                *toi = temp_toi;
                *out_tol = temp_out_tol;
                ret = 1;
                goto done;
            }
        }
        if (tol_cond || box_in) {
            if (b.t0 < toi_skip) toi_skip = b.t0;
            use_skip = 1;
            continue;
        }
        const int sp = next_split(w, tol);
        tbox lo = b, hi = b;
        lo.level = hi.level = b.level + 1;
        if (sp == 0) {
            const double mid = (b.t0 + b.t1) / 2;
            lo.t1 = mid;
            hi.t0 = mid;
            heap_push(&h, hi);
            heap_push(&h, lo);
        } else if (sp == 1) {
            const double mid = (b.u0 + b.u1) / 2;
            lo.u1 = mid;
            hi.u0 = mid;
            if (!is_vf || hi.u0 + hi.v0 <= 1) heap_push(&h, hi);
            if (!is_vf || lo.u0 + lo.v0 <= 1) heap_push(&h, lo);
        } else {
            const double mid = (b.v0 + b.v1) / 2;
            lo.v1 = mid;
            hi.v0 = mid;
            if (!is_vf || hi.u0 + hi.v0 <= 1) heap_push(&h, hi);
            if (!is_vf || lo.u0 + lo.v0 <= 1) heap_push(&h, lo);
        }
    }
    if (use_skip) {
        *toi = toi_skip;
        ret = 1;
    }
done:
    free(h.a);
    return ret;
}

static int ticcd_skip(int is_vf, const double *x_t0, const double *x_t1,
                      double err_in, double ms, double tolerance, double t_max,
                      long max_itr, int no_zero_toi, double skip_from,
                      double *toi, double *output_tolerance)
{
    double tol[3], err[3];
    param_tolerances(is_vf, x_t0, x_t1, tolerance, tol);
    if (err_in < 0)
        numerical_error(is_vf, x_t0, x_t1, ms > 0, err);
    else
        err[0] = err[1] = err[2] = err_in;
    int hit;
    long iter = 0;
    tl_refine = tl_max_queue = 0;
    do {
        hit = ti_bfs(is_vf, x_t0, x_t1, tol, tolerance, err, ms, t_max, max_itr,
                     skip_from, toi, output_tolerance);
        if (no_zero_toi && hit && *toi == 0) {
            if (*output_tolerance > tolerance) {
                t_max *= 0.9;
            } else if (10 * tolerance < ms) {
                ms *= 0.5;
            } else {
                tolerance *= 0.5;
                param_tolerances(is_vf, x_t0, x_t1, tolerance, tol);
            }
        }
        iter++;
    } while (no_zero_toi && iter < 64 && hit && *toi == 0);
    return hit;
}

int ipco_ticcd(int is_vf, const double *x_t0, const double *x_t1,
               double err_in, double ms, double tolerance, double t_max,
               long max_itr, int no_zero_toi, double *toi,
               double *output_tolerance)
{
    return ticcd_skip(is_vf, x_t0, x_t1, err_in, ms, tolerance, t_max, max_itr,
                      no_zero_toi, INFINITY, toi, output_tolerance);
}

/* ---- tight_inclusion_ccd.cpp ---------------------------------------------- */

/* ccd_strategy (tight_inclusion_ccd.cpp:33-75).  `bound` only hides
 * outcomes whose final toi would be >= bound (see ti_bfs skip_from): the
 * first query is pruned at max(bound, SMALL_TOI) and the rescaled retry at
 * bound / rescale.  bound = +inf is the plain strategy. */
static int ccd_strategy(int is_vf, const double *s, const double *e,
                        double min_distance, double initial_distance,
                        double adjusted_tolerance, double tmax,
                        const ipco_ccd_params *prm, double bound, double *toi)
{
    if (!(initial_distance > min_distance)) { /* check_initial_distance */
        *toi = 0;
        return 1;
    }
    double eff = (1.0 - prm->conservative_rescaling) * (initial_distance - min_distance);
    eff = eff < 1e-4 ? eff : 1e-4;
    eff += min_distance;
    double out_tol;
    const double small_toi = 1e-6;
    const double skip1 = bound > small_toi ? bound : small_toi;
    int hit = ticcd_skip(is_vf, s, e, -1, eff, adjusted_tolerance, tmax,
                         prm->max_iterations, 0, skip1, toi, &out_tol);
    if (hit && *toi < small_toi) {
        const double skip2 = isinf(bound) ? bound
            : bound / prm->conservative_rescaling * (1 + 1e-12);
        hit = ticcd_skip(is_vf, s, e, -1, min_distance, adjusted_tolerance, tmax,
                         -1, 1, skip2, toi, &out_tol);
        if (hit) *toi *= prm->conservative_rescaling;
    }
    return hit;
}

static int stencil_ccd_bounded(int kind, const double *x0, const double *x1,
                               double min_distance, double tmax,
                               const ipco_ccd_params *prm, double bound,
                               double *toi);

int ipco_stencil_ccd(int kind, const double *x0, const double *x1,
                     double min_distance, double tmax,
                     const ipco_ccd_params *prm, double *toi)
{
    return stencil_ccd_bounded(kind, x0, x1, min_distance, tmax, prm, INFINITY, toi);
}

static int stencil_ccd_bounded(int kind, const double *x0, const double *x1,
                               double min_distance, double tmax,
                               const ipco_ccd_params *prm, double bound,
                               double *toi)
{
    double s[12], e[12];
    double d2;
    int nd, is_vf;
    switch (kind) {
    case 0: /* point-triangle: p, t0, t1, t2 */
        d2 = ipco_point_triangle_distance(x0, x0 + 3, x0 + 6, x0 + 9, -1);
        nd = 12;
        is_vf = 1;
        memcpy(s, x0, sizeof(s));
        memcpy(e, x1, sizeof(e));
        break;
    case 1: /* edge-edge */
        d2 = ipco_edge_edge_distance(x0, x0 + 3, x0 + 6, x0 + 9, -1);
        nd = 12;
        is_vf = 0;
        memcpy(s, x0, sizeof(s));
        memcpy(e, x1, sizeof(e));
        break;
    case 2: /* point-edge as degenerate edge-edge (p, p, e0, e1) */
        d2 = ipco_point_edge_distance(x0, x0 + 3, x0 + 6, -1);
        nd = 9;
        is_vf = 0;
        memcpy(s, x0, 3 * sizeof(double));
        memcpy(s + 3, x0, 3 * sizeof(double));
        memcpy(s + 6, x0 + 3, 6 * sizeof(double));
        memcpy(e, x1, 3 * sizeof(double));
        memcpy(e + 3, x1, 3 * sizeof(double));
        memcpy(e + 6, x1 + 3, 6 * sizeof(double));
        break;
    default: /* point-point as degenerate edge-edge (p0, p0, p1, p1) */
        d2 = ipco_point_point_distance(x0, x0 + 3);
        nd = 6;
        is_vf = 0;
        memcpy(s, x0, 3 * sizeof(double));
        memcpy(s + 3, x0, 3 * sizeof(double));
        memcpy(s + 6, x0 + 3, 3 * sizeof(double));
        memcpy(s + 9, x0 + 3, 3 * sizeof(double));
        memcpy(e, x1, 3 * sizeof(double));
        memcpy(e + 3, x1, 3 * sizeof(double));
        memcpy(e + 6, x1 + 3, 3 * sizeof(double));
        memcpy(e + 9, x1 + 3, 3 * sizeof(double));
        break;
    }
    const double d0 = sqrt(d2);
    int moving = 0;
    for (int i = 0; i < nd; i++) moving |= x0[i] != x1[i];
    if (!moving) {
        if (d0 > min_distance) return 0;
        *toi = 0;
        return 1;
    }
    const double adj = 0.5 * d0 < prm->tolerance ? 0.5 * d0 : prm->tolerance;
    return ccd_strategy(is_vf, s, e, min_distance, d0, adj, tmax, prm, bound, toi);
}

/* candidate kinds: 0 face-vertex (face, vertex), 1 edge-edge, 2 edge-vertex
 * (edge, vertex), 3 vertex-vertex; dofs in the order of ipco_stencil_ccd */
static void cand_dofs(const double *V, const int32_t *E, const int32_t *F,
                      int kind, int32_t a, int32_t b, double *x)
{
    int32_t ids[4];
    int np = 4;
    switch (kind) {
    case 1:
        ids[0] = E[2 * (size_t)a];
        ids[1] = E[2 * (size_t)a + 1];
        ids[2] = E[2 * (size_t)b];
        ids[3] = E[2 * (size_t)b + 1];
        break;
    case 0: /* (face, vertex) -> (v, f0, f1, f2) */
        ids[0] = b;
        ids[1] = F[3 * (size_t)a];
        ids[2] = F[3 * (size_t)a + 1];
        ids[3] = F[3 * (size_t)a + 2];
        break;
    case 2: /* (edge, vertex) -> (v, e0, e1) */
        ids[0] = b;
        ids[1] = E[2 * (size_t)a];
        ids[2] = E[2 * (size_t)a + 1];
        np = 3;
        break;
    default: /* (v0, v1) */
        ids[0] = a;
        ids[1] = b;
        np = 2;
    }
    for (int i = 0; i < np; i++)
        for (int k = 0; k < 3; k++) x[3 * i + k] = V[3 * (size_t)ids[i] + k];
}

void ipco_candidates_toi(const double *V0, const double *V1, const int32_t *E,
                         const int32_t *F, const ipco_pairs *cand, int kind,
                         double min_distance, const ipco_ccd_params *prm,
                         double *toi)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < cand->n; i++) {
        double x0[12], x1[12], t = INFINITY;
        cand_dofs(V0, E, F, kind, cand->ij[2 * i], cand->ij[2 * i + 1], x0);
        cand_dofs(V1, E, F, kind, cand->ij[2 * i], cand->ij[2 * i + 1], x1);
        if (!ipco_stencil_ccd(kind, x0, x1, min_distance, 1.0, prm, &t)) t = INFINITY;
        toi[i] = t;
    }
}

/* the four candidate lists as one index range: ee, fv, vv, ev (any NULL) */
typedef struct clist {
    const ipco_pairs *p[4];
    int kind[4];
    int64_t off[5];
} clist;

static void clist_init(clist *c, const ipco_pairs *ee, const ipco_pairs *fv,
                       const ipco_pairs *vv, const ipco_pairs *ev)
{
    const ipco_pairs *p[4] = { ee, fv, vv, ev };
    const int kind[4] = { 1, 0, 3, 2 };
    c->off[0] = 0;
    for (int k = 0; k < 4; k++) {
        c->p[k] = p[k];
        c->kind[k] = kind[k];
        c->off[k + 1] = c->off[k] + (p[k] ? p[k]->n : 0);
    }
}

static void clist_get(const clist *c, int64_t i, int *kind, int32_t *a, int32_t *b)
{
    int k = 0;
    while (i >= c->off[k + 1]) k++;
    const int64_t j = i - c->off[k];
    *kind = c->kind[k];
    *a = c->p[k]->ij[2 * j];
    *b = c->p[k]->ij[2 * j + 1];
}

double ipco_candidates_stepsize(const double *V0, const double *V1, int nV,
                                const int32_t *E, const int32_t *F,
                                const ipco_pairs *ee, const ipco_pairs *fv,
                                const ipco_pairs *vv, const ipco_pairs *ev,
                                double min_distance,
                                const ipco_ccd_params *prm, int running_tmax)
{
    (void)nV;
    clist L;
    clist_init(&L, ee, fv, vv, ev);
    const int64_t n = L.off[4];
    if (n == 0) return 1.0;
    double earliest = 1.0;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        int kind;
        int32_t a, b;
        clist_get(&L, i, &kind, &a, &b);
        double x0[12], x1[12], t = INFINITY, cur;
#pragma omp atomic read
        cur = earliest;
        cand_dofs(V0, E, F, kind, a, b, x0);
        cand_dofs(V1, E, F, kind, a, b, x1);
        int hit;
        if (running_tmax) /* candidates.cpp:269-287: tmax = running min */
            hit = ipco_stencil_ccd(kind, x0, x1, min_distance, cur, prm, &t);
        else /* every candidate on [0,1]; outcomes >= the running min hidden */
            hit = stencil_ccd_bounded(kind, x0, x1, min_distance, 1.0, prm, cur, &t);
        if (hit) {
#pragma omp critical(ipco_toi)
            {
                if (t < earliest) earliest = t;
            }
        }
    }
    return earliest;
}

/* Candidates::build (dynamic) of candidates.cpp:120-222: surface lists and
 * the codim lists */
static void all_candidates(const double *V0, const double *V1, int nV, const int32_t *E,
                           int nE, const int32_t *F, int nF, double inflation,
                           ipco_pairs *ee, ipco_pairs *fv, ipco_pairs *vv, ipco_pairs *ev)
{
    ipco_candidates(V0, V1, nV, E, nE, F, nF, inflation, NULL, ee, fv);
    int32_t *F2E = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)(nF > 0 ? nF : 1));
    if (nF > 0 && ipco_faces_to_edges(E, nE, F, nF, F2E) != 0) {
        fprintf(stderr, "ipco: a face edge is missing from the edge list\n");
        abort();
    }
    ipco_codim_candidates(V0, V1, nV, E, nE, F, nF, F2E, inflation, NULL, vv, ev);
    free(F2E);
}

double ipco_compute_collision_free_stepsize(
    const double *V0, const double *V1, int nV, const int32_t *E, int nE,
    const int32_t *F, int nF, double min_distance,
    const ipco_ccd_params *prm, int running_tmax)
{
    ipco_pairs ee, fv, vv, ev;
    all_candidates(V0, V1, nV, E, nE, F, nF, 0.5 * min_distance, &ee, &fv, &vv, &ev);
    const double r = ipco_candidates_stepsize(V0, V1, nV, E, F, &ee, &fv, &vv, &ev,
                                              min_distance, prm, running_tmax);
    ipco_pairs_free(&ee);
    ipco_pairs_free(&fv);
    ipco_pairs_free(&vv);
    ipco_pairs_free(&ev);
    return r;
}

int ipco_is_step_collision_free(const double *V0, const double *V1, int nV,
                                const int32_t *E, int nE, const int32_t *F,
                                int nF, double min_distance,
                                const ipco_ccd_params *prm)
{
    ipco_pairs ee, fv, vv, ev;
    all_candidates(V0, V1, nV, E, nE, F, nF, 0.5 * min_distance, &ee, &fv, &vv, &ev);
    clist L;
    clist_init(&L, &ee, &fv, &vv, &ev);
    int free_ = 1;
    const int64_t n = L.off[4];
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        int done;
#pragma omp atomic read
        done = free_;
        if (!done) continue;
        int kind;
        int32_t a, b;
        clist_get(&L, i, &kind, &a, &b);
        double x0[12], x1[12], t;
        cand_dofs(V0, E, F, kind, a, b, x0);
        cand_dofs(V1, E, F, kind, a, b, x1);
        if (ipco_stencil_ccd(kind, x0, x1, min_distance, 1.0, prm, &t)) {
#pragma omp atomic write
            free_ = 0;
        }
    }
    ipco_pairs_free(&ee);
    ipco_pairs_free(&fv);
    ipco_pairs_free(&vv);
    ipco_pairs_free(&ev);
    return free_;
}

void ipco_ticcd_last_stats(long *refine, long *max_queue)
{
    *refine = tl_refine;
    *max_queue = tl_max_queue;
}
