/* Internal helpers shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see ipc_oracle.h). */
#ifndef IPCO_INTERNAL_H
#define IPCO_INTERNAL_H

#include "ipc_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* ---- 3-vectors with Eigen's fixed-size evaluation order ------------------ */
static inline void v3_sub(const double *a, const double *b, double *r)
{
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
/* Eigen's unrolled redux of a length-3 expression splits it in halves:
 * x0 + (x1 + x2) (Redux.h redux_novec_unroller). */
static inline double v3_dot(const double *a, const double *b)
{
    return a[0] * b[0] + (a[1] * b[1] + a[2] * b[2]);
}
static inline double v3_sqnorm(const double *a) { return v3_dot(a, a); }
static inline void v3_cross(const double *a, const double *b, double *r)
{
    const double x = a[1] * b[2] - a[2] * b[1];
    const double y = a[2] * b[0] - a[0] * b[2];
    const double z = a[0] * b[1] - a[1] * b[0];
    r[0] = x;
    r[1] = y;
    r[2] = z;
}

/* ---- second-order forward AD ("hyper-dual") over N <= 12 dofs ------------ */
#define HD_N 12
#define HD_H (HD_N * (HD_N + 1) / 2)
typedef struct hd {
    double v;
    double g[HD_N];
    double h[HD_H]; /* packed lower triangle, (i,j) i>=j at i*(i+1)/2+j */
} hd;

void hd_const(hd *r, double c);
void hd_var(hd *r, double x, int i);
void hd_add(const hd *a, const hd *b, hd *r);
void hd_sub(const hd *a, const hd *b, hd *r);
void hd_mul(const hd *a, const hd *b, hd *r);
void hd_div(const hd *a, const hd *b, hd *r);
void hd_scale(const hd *a, double s, hd *r);
void hd_add_const(const hd *a, double c, hd *r);
void hd_dot3(const hd *a, const hd *b, hd *r);
void hd_cross3(const hd *a, const hd *b, hd *r); /* a,b,r arrays of 3 */
void hd_unpack(const hd *a, int n, double *grad, double *hess);

/* Eigen LDLT<Matrix2d> solve (ipco_primitives.c) */
void ipco_ldlt2(double a00, double a10, double a11, double b0, double b1, double x[2]);

/* ---- misc ---------------------------------------------------------------- */
void ipco_sort_u64_pairs(uint64_t *key, int64_t *val, int64_t n); /* stable */

#endif
