/* Intersection queries: has_intersections (ipc.cpp:105-166, 3D branch) and
 * is_edge_intersecting_triangle (intersection.cpp:116-138, the default
 * build without IPC_TOOLKIT_WITH_RATIONAL_INTERSECTION).
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h).
 *
 * The orientation predicate is exact: igl::predicates::orient3d wraps
 * Shewchuk's adaptive orient3d, whose sign is exact for every input free of
 * overflow and underflow.  This restatement gets the same sign with a
 * floating-point filter (Shewchuk 1997, "Adaptive Precision Floating-Point
 * Arithmetic and Fast Robust Geometric Predicates", the orient3d error bound
 * (7 + 56 eps) eps * permanent) and, when the filter cannot decide, an exact
 * sum of the determinant's expanded terms in nonoverlapping expansion
 * arithmetic (the same paper's Two-Sum and Grow-Expansion with zero
 * elimination; products split exactly with fma).
 *
 * The linear solve restates Eigen's FullPivLU (compute, rank, solve) for a
 * 3x3 column-major matrix with vectorisation off, operation by operation. */
#include "ipco_internal.h"

#include <stdio.h>

/* ---- expansion arithmetic ------------------------------------------------ */
static inline void two_sum(double a, double b, double *x, double *y)
{
    const double s = a + b;
    const double bv = s - a;
    const double av = s - bv;
    *x = s;
    *y = (a - av) + (b - bv);
}

/* h := e + b for a nonoverlapping expansion e (increasing magnitude) */
static int grow_expansion(int elen, const double *e, double b, double *h)
{
    double q = b;
    int hl = 0;
    for (int i = 0; i < elen; i++) {
        double s, t;
        two_sum(q, e[i], &s, &t);
        q = s;
        if (t != 0.0) h[hl++] = t;
    }
    if (q != 0.0 || hl == 0) h[hl++] = q;
    return hl;
}

#define EXP_CAP 192 /* 6 terms x 8 component products x 4 parts */

typedef struct expansion {
    int n;
    double v[EXP_CAP];
} expansion;

static void exp_add(expansion *x, double b)
{
    if (b == 0.0) return;
    double h[EXP_CAP];
    const int n = grow_expansion(x->n, x->v, b, h);
    memcpy(x->v, h, sizeof(double) * (size_t)n);
    x->n = n;
}

/* adds sign * a * b * c exactly */
static void exp_add_triple(expansion *x, double a, double b, double c, int sign)
{
    if (a == 0.0 || b == 0.0 || c == 0.0) return;
    const double p = a * b;
    const double pe = fma(a, b, -p); /* a*b = p + pe exactly */
    const double q = p * c, qe = fma(p, c, -q);
    const double r = pe * c, re = fma(pe, c, -r);
    exp_add(x, sign * qe);
    exp_add(x, sign * q);
    exp_add(x, sign * re);
    exp_add(x, sign * r);
}

static int exp_sign(const expansion *x)
{
    for (int i = x->n - 1; i >= 0; i--)
        if (x->v[i] != 0.0) return x->v[i] > 0 ? 1 : -1;
    return 0;
}

/* a - b = hi + lo exactly */
static inline void two_diff(double a, double b, double d[2])
{
    const double x = a - b;
    const double bv = a - x;
    const double av = x + bv;
    d[0] = x;
    d[1] = (a - av) + (bv - b);
}

static long long g_exact_calls = 0;

/* sign of det[pa-pd; pb-pd; pc-pd] (Shewchuk's orient3d convention:
 * positive when pd lies below the plane of pa, pb, pc seen so that they
 * turn counterclockwise) */
int ipco_orient3d(const double *pa, const double *pb, const double *pc, const double *pd)
{
    const double adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    const double ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    const double adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy)
        + cdz * (adxbdy - bdxady);
    const double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz)
        + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz)
        + (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    const double eps = 0x1p-53;
    const double errbound = (7.0 + 56.0 * eps) * eps * permanent;
    if (det > errbound) return 1;
    if (-det > errbound) return -1;
#pragma omp atomic
    g_exact_calls++;
    /* exact: each difference is a two-part expansion; the determinant
     * expands into 6 signed triple products of them */
    double a[3][2], b[3][2], c[3][2];
    for (int k = 0; k < 3; k++) {
        two_diff(pa[k], pd[k], a[k]);
        two_diff(pb[k], pd[k], b[k]);
        two_diff(pc[k], pd[k], c[k]);
    }
    /* det = az(bx cy - cx by) + bz(cx ay - ax cy) + cz(ax by - bx ay) */
    static const int terms[6][4] = {
        /* u, v, w: 0 = a, 1 = b, 2 = c; axis per factor below; sign */
        { 0, 1, 2, +1 }, { 0, 2, 1, -1 }, { 1, 2, 0, +1 },
        { 1, 0, 2, -1 }, { 2, 0, 1, +1 }, { 2, 1, 0, -1 },
    };
    /* factor 1 is the z of point u, factor 2 the x of point v, factor 3 the
     * y of point w */
    expansion x;
    x.n = 0;
    for (int t = 0; t < 6; t++) {
        const double(*P[3])[2] = { a, b, c };
        const double *z = P[terms[t][0]][2];
        const double *xx = P[terms[t][1]][0];
        const double *yy = P[terms[t][2]][1];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++)
                    exp_add_triple(&x, z[i], xx[j], yy[k], terms[t][3]);
    }
    return exp_sign(&x);
}

long long ipco_orient3d_exact_calls(void) { return g_exact_calls; }

/* ---- Eigen FullPivLU<Matrix3d>::solve ------------------------------------ */
/* A column-major: A[r + 3 c] */
void ipco_fullpivlu_solve3(const double *A_in, const double *b, double *x)
{
    double A[9];
    memcpy(A, A_in, sizeof(A));
    int rowt[3], colt[3];
    int nonzero = 3;
    double maxpivot = 0.0;
    for (int k = 0; k < 3; k++) {
        /* maxCoeff over |bottom-right corner| in column-major order, the
         * first strict maximum wins */
        int br = k, bc = k;
        double big = fabs(A[k + 3 * k]);
        for (int c = k; c < 3; c++)
            for (int r = k; r < 3; r++) {
                if (r == k && c == k) continue;
                const double s = fabs(A[r + 3 * c]);
                if (s > big) {
                    big = s;
                    br = r;
                    bc = c;
                }
            }
        if (big == 0.0) {
            nonzero = k;
            for (int i = k; i < 3; i++) rowt[i] = colt[i] = i;
            break;
        }
        if (big > maxpivot) maxpivot = big;
        rowt[k] = br;
        colt[k] = bc;
        if (k != br)
            for (int c = 0; c < 3; c++) {
                const double t = A[k + 3 * c];
                A[k + 3 * c] = A[br + 3 * c];
                A[br + 3 * c] = t;
            }
        if (k != bc)
            for (int r = 0; r < 3; r++) {
                const double t = A[r + 3 * k];
                A[r + 3 * k] = A[r + 3 * bc];
                A[r + 3 * bc] = t;
            }
        for (int r = k + 1; r < 3; r++) A[r + 3 * k] = A[r + 3 * k] / A[k + 3 * k];
        for (int c = k + 1; c < 3; c++)
            for (int r = k + 1; r < 3; r++)
                A[r + 3 * c] = A[r + 3 * c] - A[r + 3 * k] * A[k + 3 * c];
    }
    /* rank(): pivots above |maxpivot| * (3 * epsilon) */
    const double thr = maxpivot * (2.220446049250313e-16 * 3.0);
    int rank = 0;
    for (int i = 0; i < nonzero; i++) rank += fabs(A[i + 3 * i]) > thr;
    if (rank == 0) {
        x[0] = x[1] = x[2] = 0.0;
        return;
    }
    /* c = P b: the row transpositions in order */
    double c[3] = { b[0], b[1], b[2] };
    for (int k = 0; k < 3; k++) {
        const double t = c[k];
        c[k] = c[rowt[k]];
        c[rowt[k]] = t;
    }
    /* unit lower solve, column oriented, skipping zero entries */
    for (int i = 0; i < 3; i++)
        if (c[i] != 0.0)
            for (int r = i + 1; r < 3; r++) c[r] = c[r] - c[i] * A[r + 3 * i];
    /* upper solve on the leading rank x rank block */
    for (int i = rank - 1; i >= 0; i--)
        if (c[i] != 0.0) {
            c[i] = c[i] / A[i + 3 * i];
            for (int r = 0; r < i; r++) c[r] = c[r] - c[i] * A[r + 3 * i];
        }
    /* x = Q c: the column transpositions composed in order */
    int q[3] = { 0, 1, 2 };
    for (int k = 0; k < 3; k++) {
        const int t = q[k];
        q[k] = q[colt[k]];
        q[colt[k]] = t;
    }
    for (int i = 0; i < rank; i++) x[q[i]] = c[i];
    for (int i = rank; i < 3; i++) x[q[i]] = 0.0;
}

/* intersection.cpp:116-138 */
int ipco_is_edge_intersecting_triangle(const double *e0, const double *e1, const double *t0,
                                       const double *t1, const double *t2)
{
    const int o1 = ipco_orient3d(t0, t1, t2, e0);
    const int o2 = ipco_orient3d(t0, t1, t2, e1);
    if (o1 != 0 && o2 != 0 && o1 == o2) return 0;
    double M[9], rhs[3], uvt[3];
    for (int k = 0; k < 3; k++) {
        M[k] = t1[k] - t0[k];
        M[3 + k] = t2[k] - t0[k];
        M[6 + k] = e0[k] - e1[k];
        rhs[k] = e0[k] - t0[k];
    }
    ipco_fullpivlu_solve3(M, rhs, uvt);
    return uvt[0] >= 0.0 && uvt[1] >= 0.0 && uvt[0] + uvt[1] <= 1.0 && uvt[2] >= 0.0
        && uvt[2] <= 1.0;
}

/* world_bbox_diagonal_length.hpp:11-15: |colwise max - colwise min|, the
 * squared norm summed left to right (unvectorised redux) */
double ipco_world_bbox_diagonal(const double *V, int nV)
{
    if (nV <= 0) return 0.0;
    double mn[3], mx[3];
    for (int k = 0; k < 3; k++) mn[k] = mx[k] = V[k];
    for (int i = 1; i < nV; i++)
        for (int k = 0; k < 3; k++) {
            const double v = V[3 * (size_t)i + k];
            if (v < mn[k]) mn[k] = v;
            if (v > mx[k]) mx[k] = v;
        }
    const double d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

/* ipc.cpp:105-166 (3D): edge-face candidates at 1e-6 x the world diagonal,
 * then any edge crossing a triangle */
int ipco_has_intersections(const double *V, int nV, const int32_t *E, int nE,
                           const int32_t *F, int nF, const int32_t *vertex_group,
                           int *result)
{
    *result = 0;
    const double r = 1e-6 * ipco_world_bbox_diagonal(V, nV);
    ipco_pairs ef;
    ipco_edge_face_candidates(V, nV, E, nE, F, nF, r, vertex_group, &ef);
    int hit = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : hit)
    for (int64_t i = 0; i < ef.n; i++) {
        const int ei = ef.ij[2 * i], fi = ef.ij[2 * i + 1];
        const double *e0 = V + 3 * (size_t)E[2 * (size_t)ei];
        const double *e1 = V + 3 * (size_t)E[2 * (size_t)ei + 1];
        const double *t0 = V + 3 * (size_t)F[3 * (size_t)fi];
        const double *t1 = V + 3 * (size_t)F[3 * (size_t)fi + 1];
        const double *t2 = V + 3 * (size_t)F[3 * (size_t)fi + 2];
        hit |= ipco_is_edge_intersecting_triangle(e0, e1, t0, t1, t2);
    }
    ipco_pairs_free(&ef);
    *result = hit;
    return 0;
}
