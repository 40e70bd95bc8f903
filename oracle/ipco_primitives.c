/* Primitive geometry of the contact path: distance types, squared
 * distances and their derivatives, EE mollifier, barrier, PSD projection.
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h). */
#include "ipco_internal.h"

#include <float.h>

/* ========================================================================= */
/* Hyper-dual numbers: value, gradient and Hessian propagated together.      */
/* ========================================================================= */

void hd_const(hd *r, double c)
{
    memset(r, 0, sizeof(*r));
    r->v = c;
}

void hd_var(hd *r, double x, int i)
{
    memset(r, 0, sizeof(*r));
    r->v = x;
    r->g[i] = 1.0;
}

void hd_add(const hd *a, const hd *b, hd *r)
{
    r->v = a->v + b->v;
    for (int i = 0; i < HD_N; i++) r->g[i] = a->g[i] + b->g[i];
    for (int i = 0; i < HD_H; i++) r->h[i] = a->h[i] + b->h[i];
}

void hd_sub(const hd *a, const hd *b, hd *r)
{
    r->v = a->v - b->v;
    for (int i = 0; i < HD_N; i++) r->g[i] = a->g[i] - b->g[i];
    for (int i = 0; i < HD_H; i++) r->h[i] = a->h[i] - b->h[i];
}

void hd_mul(const hd *a, const hd *b, hd *r)
{
    hd t;
    t.v = a->v * b->v;
    for (int i = 0; i < HD_N; i++) t.g[i] = a->v * b->g[i] + b->v * a->g[i];
    int k = 0;
    for (int i = 0; i < HD_N; i++)
        for (int j = 0; j <= i; j++, k++)
            t.h[k] = a->v * b->h[k] + b->v * a->h[k] + a->g[i] * b->g[j]
                + b->g[i] * a->g[j];
    *r = t;
}

void hd_div(const hd *a, const hd *b, hd *r)
{
    /* c = a / b  <=>  a = c b: c'' = (a'' - c b'' - c' b'^T - b' c'^T) / b */
    hd t;
    t.v = a->v / b->v;
    for (int i = 0; i < HD_N; i++) t.g[i] = (a->g[i] - t.v * b->g[i]) / b->v;
    int k = 0;
    for (int i = 0; i < HD_N; i++)
        for (int j = 0; j <= i; j++, k++)
            t.h[k] = (a->h[k] - t.v * b->h[k] - t.g[i] * b->g[j]
                      - b->g[i] * t.g[j])
                / b->v;
    *r = t;
}

void hd_scale(const hd *a, double s, hd *r)
{
    r->v = a->v * s;
    for (int i = 0; i < HD_N; i++) r->g[i] = a->g[i] * s;
    for (int i = 0; i < HD_H; i++) r->h[i] = a->h[i] * s;
}

void hd_add_const(const hd *a, double c, hd *r)
{
    *r = *a;
    r->v += c;
}

void hd_dot3(const hd *a, const hd *b, hd *r)
{
    hd t0, t1, t2;
    hd_mul(&a[0], &b[0], &t0);
    hd_mul(&a[1], &b[1], &t1);
    hd_mul(&a[2], &b[2], &t2);
    hd_add(&t1, &t2, &t1);
    hd_add(&t0, &t1, r);
}

void hd_cross3(const hd *a, const hd *b, hd *r)
{
    hd p, q, out[3];
    hd_mul(&a[1], &b[2], &p);
    hd_mul(&a[2], &b[1], &q);
    hd_sub(&p, &q, &out[0]);
    hd_mul(&a[2], &b[0], &p);
    hd_mul(&a[0], &b[2], &q);
    hd_sub(&p, &q, &out[1]);
    hd_mul(&a[0], &b[1], &p);
    hd_mul(&a[1], &b[0], &q);
    hd_sub(&p, &q, &out[2]);
    r[0] = out[0];
    r[1] = out[1];
    r[2] = out[2];
}

void hd_unpack(const hd *a, int n, double *grad, double *hess)
{
    if (grad)
        for (int i = 0; i < n; i++) grad[i] = a->g[i];
    if (hess)
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const int ii = i > j ? i : j, jj = i > j ? j : i;
                hess[i * n + j] = a->h[ii * (ii + 1) / 2 + jj];
            }
}

/* ========================================================================= */
/* Distance types (distance_type.cpp)                                        */
/* ========================================================================= */

/* distance_type.cpp:10-34 */
int ipco_point_edge_distance_type(const double p[3], const double e0[3],
                                  const double e1[3])
{
    double e[3], pe0[3];
    v3_sub(e1, e0, e);
    const double len2 = v3_sqnorm(e);
    if (len2 == 0) return IPCO_PE_P_E0; /* degenerate edge: arbitrary end */
    v3_sub(p, e0, pe0);
    const double ratio = v3_dot(e, pe0) / len2;
    if (ratio < 0) return IPCO_PE_P_E0;
    if (ratio > 1) return IPCO_PE_P_E1;
    return IPCO_PE_P_E;
}

/* Eigen LDLT<Matrix2d> (lower, diagonal pivoting) compute + solve, restated
 * from Eigen's LDLT.h: pivot = largest |diagonal| (first on ties); an
 * all-zero diagonal leaves the identity permutation; D^+ zeroes pivots with
 * |d| <= numeric_limits<double>::min(). */
static void ldlt2_solve(double a00, double a10, double a11, double b0,
                        double b1, double x[2])
{
    int swap = fabs(a11) > fabs(a00);
    double m00 = swap ? a11 : a00, m11 = swap ? a00 : a11, l = a10;
    double d0 = m00, d1;
    if (!(fabs(d0) > 0.0)) {
        swap = 0;
        d0 = a00;
        d1 = a11;
    } else {
        l = a10 / d0;
        d1 = m11 - l * (d0 * l);
    }
    double y0 = swap ? b1 : b0, y1 = swap ? b0 : b1;
    y1 = y1 - l * y0;
    y0 = fabs(d0) > DBL_MIN ? y0 / d0 : 0.0;
    y1 = fabs(d1) > DBL_MIN ? y1 / d1 : 0.0;
    y0 = y0 - l * y1;
    x[0] = swap ? y1 : y0;
    x[1] = swap ? y0 : y1;
}

void ipco_ldlt2(double a00, double a10, double a11, double b0, double b1, double x[2])
{
    ldlt2_solve(a00, a10, a11, b0, b1, x);
}

/* One edge of point_triangle_distance_type: basis rows (e, e x n),
 * param = (B B^T).ldlt().solve(B (p - a)). */
static void pt_edge_param(const double p[3], const double a[3],
                          const double b[3], const double n[3], double x[2])
{
    double r0[3], r1[3], pa[3];
    v3_sub(b, a, r0);
    v3_cross(r0, n, r1);
    v3_sub(p, a, pa);
    const double m00 = v3_dot(r0, r0);
    const double m10 = v3_dot(r1, r0);
    const double m11 = v3_dot(r1, r1);
    ldlt2_solve(m00, m10, m11, v3_dot(r0, pa), v3_dot(r1, pa), x);
}

/* distance_type.cpp:36-79 */
int ipco_point_triangle_distance_type(const double p[3], const double t0[3],
                                      const double t1[3], const double t2[3])
{
    double e0[3], e1[3], n[3], x0[2], x1[2], x2[2];
    v3_sub(t1, t0, e0);
    v3_sub(t2, t0, e1);
    v3_cross(e0, e1, n);

    pt_edge_param(p, t0, t1, n, x0);
    if (x0[0] > 0.0 && x0[0] < 1.0 && x0[1] >= 0.0) return IPCO_PT_P_E0;
    pt_edge_param(p, t1, t2, n, x1);
    if (x1[0] > 0.0 && x1[0] < 1.0 && x1[1] >= 0.0) return IPCO_PT_P_E1;
    pt_edge_param(p, t2, t0, n, x2);
    if (x2[0] > 0.0 && x2[0] < 1.0 && x2[1] >= 0.0) return IPCO_PT_P_E2;

    if (x0[0] <= 0.0 && x2[0] >= 1.0) return IPCO_PT_P_T0;
    if (x1[0] <= 0.0 && x0[0] >= 1.0) return IPCO_PT_P_T1;
    if (x2[0] <= 0.0 && x1[0] >= 1.0) return IPCO_PT_P_T2;
    return IPCO_PT_P_T;
}

/* distance_type.cpp:176-211 */
static int edge_edge_parallel_distance_type(const double ea0[3],
                                            const double ea1[3],
                                            const double eb0[3],
                                            const double eb1[3])
{
    double ea[3], d0[3], d1[3];
    v3_sub(ea1, ea0, ea);
    v3_sub(eb0, ea0, d0);
    v3_sub(eb1, ea0, d1);
    const double alpha = v3_dot(d0, ea) / v3_sqnorm(ea);
    const double beta = v3_dot(d1, ea) / v3_sqnorm(ea);
    int eac, ebc; /* 0: *0, 1: *1, 2: interior */
    if (alpha < 0) {
        eac = (0 <= beta && beta <= 1) ? 2 : 0;
        ebc = (beta <= alpha) ? 0 : (beta <= 1 ? 1 : 2);
    } else if (alpha > 1) {
        eac = (0 <= beta && beta <= 1) ? 2 : 1;
        ebc = (beta >= alpha) ? 0 : (0 <= beta ? 1 : 2);
    } else {
        eac = 2;
        ebc = 0;
    }
    return ebc < 2 ? (eac << 1 | ebc) : (6 + eac);
}

/* distance_type.cpp:82-174 */
int ipco_edge_edge_distance_type(const double ea0[3], const double ea1[3],
                                 const double eb0[3], const double eb1[3])
{
    const double PARALLEL_THRESHOLD = 2.5e-16;
    double u[3], v[3], w[3], uxv[3];
    v3_sub(ea1, ea0, u);
    v3_sub(eb1, eb0, v);
    v3_sub(ea0, eb0, w);
    const double a = v3_sqnorm(u);
    const double b = v3_dot(u, v);
    const double c = v3_sqnorm(v);
    const double d = v3_dot(u, w);
    const double e = v3_dot(v, w);
    const double D = a * c - b * b;

    if (a == 0.0 && c == 0.0) return IPCO_EE_EA0_EB0;
    if (a == 0.0) return IPCO_EE_EA0_EB;
    if (c == 0.0) return IPCO_EE_EA_EB0;

    const double parallel_tolerance = PARALLEL_THRESHOLD * a * c;
    v3_cross(u, v, uxv);
    const double cross2 = v3_sqnorm(uxv);
    if (cross2 < parallel_tolerance)
        return edge_edge_parallel_distance_type(ea0, ea1, eb0, eb1);

    int default_case = IPCO_EE_EA_EB;
    const double sN = (b * e - c * d);
    double tN, tD;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        default_case = IPCO_EE_EA0_EB;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        default_case = IPCO_EE_EA1_EB;
    } else {
        tN = (a * e - b * d);
        tD = D;
        if (tN > 0.0 && tN < tD && cross2 < parallel_tolerance) {
            if (sN < D / 2) {
                tN = e;
                tD = c;
                default_case = IPCO_EE_EA0_EB;
            } else {
                tN = e + b;
                tD = c;
                default_case = IPCO_EE_EA1_EB;
            }
        }
    }

    if (tN <= 0.0) {
        if (-d <= 0.0) return IPCO_EE_EA0_EB0;
        if (-d >= a) return IPCO_EE_EA1_EB0;
        return IPCO_EE_EA_EB0;
    } else if (tN >= tD) {
        if ((-d + b) <= 0.0) return IPCO_EE_EA0_EB1;
        if ((-d + b) >= a) return IPCO_EE_EA1_EB1;
        return IPCO_EE_EA_EB1;
    }
    return default_case;
}

/* ========================================================================= */
/* Squared distances (values follow the reference op by op)                  */
/* ========================================================================= */

/* point_point.cpp:5-9 */
double ipco_point_point_distance(const double p0[3], const double p1[3])
{
    double d[3];
    v3_sub(p1, p0, d);
    return v3_sqnorm(d);
}

/* point_line.cpp:5-23 (3D branch) */
double ipco_point_line_distance(const double p[3], const double e0[3],
                                const double e1[3])
{
    double a[3], b[3], c[3], e[3];
    v3_sub(e0, p, a);
    v3_sub(e1, p, b);
    v3_cross(a, b, c);
    v3_sub(e1, e0, e);
    return v3_sqnorm(c) / v3_sqnorm(e);
}

/* point_plane.cpp:10-26 with geometry/normal.hpp:149-168 (normalized
 * normal: n / sqrt(|n|^2) when |n|^2 > 0) */
double ipco_point_plane_distance(const double p[3], const double t0[3],
                                 const double t1[3], const double t2[3])
{
    double a[3], b[3], n[3], r[3];
    v3_sub(t1, t0, a);
    v3_sub(t2, t0, b);
    v3_cross(a, b, n);
    const double z = v3_sqnorm(n);
    if (z > 0) {
        const double s = sqrt(z);
        n[0] = n[0] / s;
        n[1] = n[1] / s;
        n[2] = n[2] / s;
    }
    v3_sub(p, t0, r);
    const double ptp = v3_dot(r, n);
    return ptp * ptp / v3_sqnorm(n);
}

/* line_line.cpp:7-16 */
double ipco_line_line_distance(const double ea0[3], const double ea1[3],
                               const double eb0[3], const double eb1[3])
{
    double u[3], v[3], n[3], w[3];
    v3_sub(ea1, ea0, u);
    v3_sub(eb1, eb0, v);
    v3_cross(u, v, n);
    v3_sub(eb0, ea0, w);
    const double l = v3_dot(w, n);
    return l * l / v3_sqnorm(n);
}

/* point_edge.cpp:10-37 */
double ipco_point_edge_distance(const double p[3], const double e0[3],
                                const double e1[3], int dtype)
{
    if (dtype < 0) dtype = ipco_point_edge_distance_type(p, e0, e1);
    switch (dtype) {
    case IPCO_PE_P_E0: return ipco_point_point_distance(p, e0);
    case IPCO_PE_P_E1: return ipco_point_point_distance(p, e1);
    default: return ipco_point_line_distance(p, e0, e1);
    }
}

/* point_triangle.cpp:11-48 */
double ipco_point_triangle_distance(const double p[3], const double t0[3],
                                    const double t1[3], const double t2[3],
                                    int dtype)
{
    if (dtype < 0) dtype = ipco_point_triangle_distance_type(p, t0, t1, t2);
    switch (dtype) {
    case IPCO_PT_P_T0: return ipco_point_point_distance(p, t0);
    case IPCO_PT_P_T1: return ipco_point_point_distance(p, t1);
    case IPCO_PT_P_T2: return ipco_point_point_distance(p, t2);
    case IPCO_PT_P_E0: return ipco_point_line_distance(p, t0, t1);
    case IPCO_PT_P_E1: return ipco_point_line_distance(p, t1, t2);
    case IPCO_PT_P_E2: return ipco_point_line_distance(p, t2, t0);
    default: return ipco_point_plane_distance(p, t0, t1, t2);
    }
}

/* edge_edge.cpp:13-56 */
double ipco_edge_edge_distance(const double ea0[3], const double ea1[3],
                               const double eb0[3], const double eb1[3],
                               int dtype)
{
    if (dtype < 0) dtype = ipco_edge_edge_distance_type(ea0, ea1, eb0, eb1);
    switch (dtype) {
    case IPCO_EE_EA0_EB0: return ipco_point_point_distance(ea0, eb0);
    case IPCO_EE_EA0_EB1: return ipco_point_point_distance(ea0, eb1);
    case IPCO_EE_EA1_EB0: return ipco_point_point_distance(ea1, eb0);
    case IPCO_EE_EA1_EB1: return ipco_point_point_distance(ea1, eb1);
    case IPCO_EE_EA_EB0: return ipco_point_line_distance(eb0, ea0, ea1);
    case IPCO_EE_EA_EB1: return ipco_point_line_distance(eb1, ea0, ea1);
    case IPCO_EE_EA0_EB: return ipco_point_line_distance(ea0, eb0, eb1);
    case IPCO_EE_EA1_EB: return ipco_point_line_distance(ea1, eb0, eb1);
    default: return ipco_line_line_distance(ea0, ea1, eb0, eb1);
    }
}

/* ========================================================================= */
/* Derivatives by hyper-dual AD.  The derivative of each squared distance is */
/* that of the mathematical function the reference differentiates           */
/* (point_plane: ((p-t0).n)^2/|n|^2 with the unnormalised normal).           */
/* ========================================================================= */

static void hd_vec(const double *x, int vi, hd out[3])
{
    for (int k = 0; k < 3; k++) hd_var(&out[k], x[3 * vi + k], 3 * vi + k);
}

static void hd_v3_sub(const hd *a, const hd *b, hd *r)
{
    for (int k = 0; k < 3; k++) hd_sub(&a[k], &b[k], &r[k]);
}

/* |a - b|^2 */
static void ad_pp(const hd *a, const hd *b, hd *r)
{
    hd d[3];
    hd_v3_sub(b, a, d);
    hd_dot3(d, d, r);
}

/* |(e0-p) x (e1-p)|^2 / |e1-e0|^2 */
static void ad_pl(const hd *p, const hd *e0, const hd *e1, hd *r)
{
    hd a[3], b[3], c[3], e[3], num, den;
    hd_v3_sub(e0, p, a);
    hd_v3_sub(e1, p, b);
    hd_cross3(a, b, c);
    hd_dot3(c, c, &num);
    hd_v3_sub(e1, e0, e);
    hd_dot3(e, e, &den);
    hd_div(&num, &den, r);
}

/* ((p - t0).n)^2 / |n|^2, n = (t1-t0)x(t2-t0) */
static void ad_pplane(const hd *p, const hd *t0, const hd *t1, const hd *t2,
                      hd *r)
{
    hd a[3], b[3], n[3], w[3], s, s2, den;
    hd_v3_sub(t1, t0, a);
    hd_v3_sub(t2, t0, b);
    hd_cross3(a, b, n);
    hd_v3_sub(p, t0, w);
    hd_dot3(w, n, &s);
    hd_mul(&s, &s, &s2);
    hd_dot3(n, n, &den);
    hd_div(&s2, &den, r);
}

/* ((eb0-ea0).n)^2/|n|^2, n = (ea1-ea0)x(eb1-eb0) */
static void ad_ll(const hd *ea0, const hd *ea1, const hd *eb0, const hd *eb1,
                  hd *r)
{
    hd u[3], v[3], n[3], w[3], s, s2, den;
    hd_v3_sub(ea1, ea0, u);
    hd_v3_sub(eb1, eb0, v);
    hd_cross3(u, v, n);
    hd_v3_sub(eb0, ea0, w);
    hd_dot3(w, n, &s);
    hd_mul(&s, &s, &s2);
    hd_dot3(n, n, &den);
    hd_div(&s2, &den, r);
}

double ipco_stencil_distance_derivs(int kind, int dtype, const double *x,
                                    double *grad, double *hess)
{
    hd v[4][3];
    hd r;
    const int nv = kind <= 1 ? 4 : (kind == 2 ? 3 : 2);
    for (int i = 0; i < nv; i++) hd_vec(x, i, v[i]);
    double value;
    if (kind == 0) { /* p, t0, t1, t2 */
        const double *p = x, *t0 = x + 3, *t1 = x + 6, *t2 = x + 9;
        if (dtype < 0) dtype = ipco_point_triangle_distance_type(p, t0, t1, t2);
        value = ipco_point_triangle_distance(p, t0, t1, t2, dtype);
        switch (dtype) {
        case IPCO_PT_P_T0: ad_pp(v[0], v[1], &r); break;
        case IPCO_PT_P_T1: ad_pp(v[0], v[2], &r); break;
        case IPCO_PT_P_T2: ad_pp(v[0], v[3], &r); break;
        case IPCO_PT_P_E0: ad_pl(v[0], v[1], v[2], &r); break;
        case IPCO_PT_P_E1: ad_pl(v[0], v[2], v[3], &r); break;
        case IPCO_PT_P_E2: ad_pl(v[0], v[3], v[1], &r); break;
        default: ad_pplane(v[0], v[1], v[2], v[3], &r); break;
        }
    } else if (kind == 1) { /* ea0, ea1, eb0, eb1 */
        const double *a0 = x, *a1 = x + 3, *b0 = x + 6, *b1 = x + 9;
        if (dtype < 0) dtype = ipco_edge_edge_distance_type(a0, a1, b0, b1);
        value = ipco_edge_edge_distance(a0, a1, b0, b1, dtype);
        switch (dtype) {
        case IPCO_EE_EA0_EB0: ad_pp(v[0], v[2], &r); break;
        case IPCO_EE_EA0_EB1: ad_pp(v[0], v[3], &r); break;
        case IPCO_EE_EA1_EB0: ad_pp(v[1], v[2], &r); break;
        case IPCO_EE_EA1_EB1: ad_pp(v[1], v[3], &r); break;
        case IPCO_EE_EA_EB0: ad_pl(v[2], v[0], v[1], &r); break;
        case IPCO_EE_EA_EB1: ad_pl(v[3], v[0], v[1], &r); break;
        case IPCO_EE_EA0_EB: ad_pl(v[0], v[2], v[3], &r); break;
        case IPCO_EE_EA1_EB: ad_pl(v[1], v[2], v[3], &r); break;
        default: ad_ll(v[0], v[1], v[2], v[3], &r); break;
        }
    } else if (kind == 2) { /* p, e0, e1 */
        const double *p = x, *e0 = x + 3, *e1 = x + 6;
        if (dtype < 0) dtype = ipco_point_edge_distance_type(p, e0, e1);
        value = ipco_point_edge_distance(p, e0, e1, dtype);
        switch (dtype) {
        case IPCO_PE_P_E0: ad_pp(v[0], v[1], &r); break;
        case IPCO_PE_P_E1: ad_pp(v[0], v[2], &r); break;
        default: ad_pl(v[0], v[1], v[2], &r); break;
        }
    } else {
        value = ipco_point_point_distance(x, x + 3);
        ad_pp(v[0], v[1], &r);
    }
    hd_unpack(&r, 3 * nv, grad, hess);
    return value;
}

/* ========================================================================= */
/* Edge-edge mollifier (edge_edge_mollifier.cpp)                             */
/* ========================================================================= */

/* :7-14 */
double ipco_edge_edge_cross_squarednorm(const double ea0[3],
                                        const double ea1[3],
                                        const double eb0[3],
                                        const double eb1[3])
{
    double u[3], v[3], c[3];
    v3_sub(ea1, ea0, u);
    v3_sub(eb1, eb0, v);
    v3_cross(u, v, c);
    return v3_sqnorm(c);
}

/* :194-202  1e-3 * |ea0-ea1|^2 * |eb0-eb1|^2 (left to right) */
double ipco_edge_edge_mollifier_threshold(const double ea0[3],
                                          const double ea1[3],
                                          const double eb0[3],
                                          const double eb1[3])
{
    double a[3], b[3];
    v3_sub(ea0, ea1, a);
    v3_sub(eb0, eb1, b);
    return 1e-3 * v3_sqnorm(a) * v3_sqnorm(b);
}

/* :83-136: m(x) = (-x/eps + 2)(x/eps) for x = |u x v|^2 < eps, else 1. */
double ipco_edge_edge_mollifier(const double *x12, double eps_x, double *grad,
                                double *hess)
{
    const double s =
        ipco_edge_edge_cross_squarednorm(x12, x12 + 3, x12 + 6, x12 + 9);
    if (!(s < eps_x)) {
        if (grad) memset(grad, 0, 12 * sizeof(double));
        if (hess) memset(hess, 0, 144 * sizeof(double));
        return 1.0;
    }
    const double q = s / eps_x;
    const double m = (-q + 2) * q;
    if (grad || hess) {
        hd v[4][3], u[3], w[3], c[3], x, t, two_minus, r;
        for (int i = 0; i < 4; i++) hd_vec(x12, i, v[i]);
        hd_v3_sub(v[1], v[0], u);
        hd_v3_sub(v[3], v[2], w);
        hd_cross3(u, w, c);
        hd_dot3(c, c, &x);
        hd_scale(&x, 1.0 / eps_x, &t);
        hd_scale(&t, -1.0, &two_minus);
        hd_add_const(&two_minus, 2.0, &two_minus);
        hd_mul(&two_minus, &t, &r);
        hd_unpack(&r, 12, grad, hess);
    }
    return m;
}

/* ========================================================================= */
/* Barrier (barrier.cpp:11-43)                                               */
/* ========================================================================= */

double ipco_barrier(double d, double dhat)
{
    if (d <= 0.0) return INFINITY;
    if (d >= dhat) return 0;
    const double dm = (d - dhat);
    return -dm * dm * log(d / dhat);
}

double ipco_barrier_first_derivative(double d, double dhat)
{
    if (d <= 0.0 || d >= dhat) return 0.0;
    return (dhat - d) * (2 * log(d / dhat) - dhat / d + 1);
}

double ipco_barrier_second_derivative(double d, double dhat)
{
    if (d <= 0.0 || d >= dhat) return 0.0;
    const double dhat_d = dhat / d;
    return (dhat_d + 2) * dhat_d - 2 * log(d / dhat) - 3;
}

/* ========================================================================= */
/* PSD projection (eigen_ext.tpp:63-108) with a cyclic Jacobi eigensolver    */
/* standing in for Eigen's SelfAdjointEigenSolver (agreement is to          */
/* O(eps |A|), not bitwise).                                                 */
/* ========================================================================= */

static int jacobi_eigen(const double *A, int n, double *lambda, double *Vout)
{
    double a[144], v[144];
    memcpy(a, A, sizeof(double) * n * n);
    for (int i = 0; i < n * n; i++) v[i] = 0;
    for (int i = 0; i < n; i++) v[i * n + i] = 1;
    int converged = 0;
    for (int sweep = 0; sweep <= 60; sweep++) {
        double off = 0, diag = 0;
        for (int p = 0; p < n; p++) {
            diag += a[p * n + p] * a[p * n + p];
            for (int q = p + 1; q < n; q++) off += a[p * n + q] * a[p * n + q];
        }
        const double fro2 = diag + 2 * off;
        converged = off <= 1e-28 * fro2;
        if (off == 0 || off <= 1e-34 * fro2 || sweep == 60) break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a[p * n + q];
                if (apq == 0) continue;
                const double app = a[p * n + p], aqq = a[q * n + q];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0)
                    / (fabs(theta) + sqrt(theta * theta + 1));
                const double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; k++) { /* columns p,q */
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - s * akq;
                    a[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++) { /* rows p,q */
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - s * aqk;
                    a[q * n + k] = s * apk + c * aqk;
                }
                a[p * n + q] = a[q * n + p] = 0;
                for (int k = 0; k < n; k++) {
                    const double vkp = v[k * n + p], vkq = v[k * n + q];
                    v[k * n + p] = c * vkp - s * vkq;
                    v[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort ascending */
    int idx[12];
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) {
        int j = i;
        while (j > 0 && a[idx[j - 1] * n + idx[j - 1]] > a[idx[j] * n + idx[j]]) {
            int t = idx[j];
            idx[j] = idx[j - 1];
            idx[j - 1] = t;
            j--;
        }
    }
    for (int i = 0; i < n; i++) {
        lambda[i] = a[idx[i] * n + idx[i]];
        if (Vout)
            for (int k = 0; k < n; k++) Vout[k * n + i] = v[k * n + idx[i]];
    }
    return converged ? 0 : -1;
}

int ipco_symmetric_eigenvalues(const double *A, int n, double *lambda)
{
    return jacobi_eigen(A, n, lambda, NULL);
}

int ipco_project_to_psd(double *A, int n, int method)
{
    if (method == IPCO_PSD_NONE) return 0;
    double lambda[12], V[144];
    if (jacobi_eigen(A, n, lambda, V) != 0) return -1;
    if (lambda[0] >= 0.0) return 0; /* already PSD: return A unchanged */
    for (int i = 0; i < n; i++) {
        if (lambda[i] < 0.0)
            lambda[i] = method == IPCO_PSD_CLAMP ? 0.0 : fabs(lambda[i]);
        else
            break;
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double s = 0;
            for (int k = 0; k < n; k++) s += V[i * n + k] * lambda[k] * V[j * n + k];
            A[i * n + j] = s;
        }
    return 0;
}
