// Edge-triangle intersection predicate of has_intersections
// (intersection.cpp:116-138, the default build without
// IPC_TOOLKIT_WITH_RATIONAL_INTERSECTION):
//   1. orient3d of both edge endpoints against the triangle; same nonzero
//      sign -> no intersection.  igl::predicates::orient3d is Shewchuk's
//      adaptive predicate, whose sign is exact.
//   2. otherwise Eigen's FullPivLU<Matrix3d> solve of
//      [t1-t0, t2-t0, e0-e1] (u, v, t) = e0 - t0 and the test
//      u, v >= 0, u + v <= 1, 0 <= t <= 1.
//
// orient3d here is split in two for the GPU: a floating-point filter with
// Shewchuk's published error bound (Shewchuk 1997, "Adaptive Precision
// Floating-Point Arithmetic and Fast Robust Geometric Predicates", orient3d
// bound (7 + 56 eps) eps * permanent) that decides almost every call in
// registers, and an exact evaluation for the undecided rest: the
// determinant of the exact coordinate differences (Two-Diff, two parts
// each) expands into 6 x 8 triple products, each split exactly into 4
// doubles with fma, summed into a nonoverlapping expansion (Two-Sum,
// Grow-Expansion with zero elimination, same paper) whose largest
// component carries the sign.  Both give the exact sign whenever no
// product overflows or underflows, as the reference's predicate does.
#pragma once

#include "geometry.hpp"

namespace ipcg {

// sign of det[pa-pd; pb-pd; pc-pd] when the filter can certify it (+1 / -1),
// else 0 (undecided: the exact path decides)
__device__ __forceinline__ int orient3d_filter(const d3& pa, const d3& pb, const d3& pc,
                                               const d3& pd)
{
    const double adx = pa.x - pd.x, bdx = pb.x - pd.x, cdx = pc.x - pd.x;
    const double ady = pa.y - pd.y, bdy = pb.y - pd.y, cdy = pc.y - pd.y;
    const double adz = pa.z - pd.z, bdz = pb.z - pd.z, cdz = pc.z - pd.z;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const double permanent = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz)
        + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) + (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    constexpr double eps = 0x1p-53;
    constexpr double bound_a = (7.0 + 56.0 * eps) * eps;
    const double errbound = bound_a * permanent;
    if (det > errbound) return 1;
    if (-det > errbound) return -1;
    return 0;
}

constexpr int EXPANSION_CAP = 192; // 6 terms x 8 part products x 4 parts

// e := e + b in place (e nonoverlapping, increasing magnitude; the output
// index never passes the input index, so one array suffices)
__device__ __forceinline__ int expansion_grow(double* e, int n, double b)
{
    double q = b;
    int h = 0;
    for (int i = 0; i < n; i++) {
        const double ei = e[i];
        const double s = q + ei;
        const double bv = s - q;
        const double av = s - bv;
        const double t = (q - av) + (ei - bv);
        q = s;
        if (t != 0.0) e[h++] = t;
    }
    if (q != 0.0 || h == 0) e[h++] = q;
    return h;
}

__device__ int orient3d_exact(const d3& pa, const d3& pb, const d3& pc, const d3& pd)
{
    // exact differences hi + lo
    double d[3][3][2]; // point (a, b, c), axis, part
    const double P[3][3] = { { pa.x, pa.y, pa.z }, { pb.x, pb.y, pb.z }, { pc.x, pc.y, pc.z } };
    const double D[3] = { pd.x, pd.y, pd.z };
    for (int p = 0; p < 3; p++)
        for (int k = 0; k < 3; k++) {
            const double a = P[p][k], b = D[k];
            const double x = a - b;
            const double bv = a - x;
            const double av = x + bv;
            d[p][k][0] = x;
            d[p][k][1] = (a - av) + (bv - b);
        }
    // det = az(bx cy - cx by) + bz(cx ay - ax cy) + cz(ax by - bx ay):
    // (point of the z factor, of the x factor, of the y factor, sign)
    constexpr int T[6][4] = { { 0, 1, 2, 1 }, { 0, 2, 1, -1 }, { 1, 2, 0, 1 },
                              { 1, 0, 2, -1 }, { 2, 0, 1, 1 }, { 2, 1, 0, -1 } };
    double e[EXPANSION_CAP];
    int n = 0;
    for (int t = 0; t < 6; t++) {
        const double sg = (double)T[t][3];
        for (int i = 0; i < 2; i++)
            for (int j = 0; j < 2; j++)
                for (int k = 0; k < 2; k++) {
                    const double a = d[T[t][0]][2][i], b = d[T[t][1]][0][j], c = d[T[t][2]][1][k];
                    if (a == 0.0 || b == 0.0 || c == 0.0) continue;
                    const double p = a * b, pe = fma(a, b, -p);
                    const double q = p * c, qe = fma(p, c, -q);
                    const double r = pe * c, re = fma(pe, c, -r);
                    const double parts[4] = { qe, q, re, r };
                    for (int m = 0; m < 4; m++)
                        if (parts[m] != 0.0) n = expansion_grow(e, n, sg * parts[m]);
                }
    }
    for (int i = n - 1; i >= 0; i--)
        if (e[i] != 0.0) return e[i] > 0.0 ? 1 : -1;
    return 0;
}

// Eigen FullPivLU<Matrix3d>(A).solve(b) with A column-major (A[r + 3c]):
// full pivoting on the first strict maximum of |corner| in column-major
// order, rank = pivots above |max pivot| * 3 eps, unit-lower then upper
// column-oriented substitution skipping zero entries, free unknowns zero.
__device__ __forceinline__ void fullpivlu_solve3(double A[9], const double b[3], double x[3])
{
    int rowt[3] = { 0, 1, 2 }, colt[3] = { 0, 1, 2 };
    int nonzero = 3;
    double maxpivot = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int br = k, bc = k;
        double big = fabs(A[k + 3 * k]);
#pragma unroll
        for (int c = k; c < 3; c++)
#pragma unroll
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
            break;
        }
        maxpivot = big > maxpivot ? big : maxpivot;
        rowt[k] = br;
        colt[k] = bc;
        // swaps with selects keep A in registers
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double tk = A[k + 3 * c];
            double tb = tk;
#pragma unroll
            for (int r = k + 1; r < 3; r++) tb = br == r ? A[r + 3 * c] : tb;
#pragma unroll
            for (int r = k + 1; r < 3; r++) A[r + 3 * c] = br == r ? tk : A[r + 3 * c];
            A[k + 3 * c] = tb;
        }
#pragma unroll
        for (int r = 0; r < 3; r++) {
            const double tk = A[r + 3 * k];
            double tb = tk;
#pragma unroll
            for (int c = k + 1; c < 3; c++) tb = bc == c ? A[r + 3 * c] : tb;
#pragma unroll
            for (int c = k + 1; c < 3; c++) A[r + 3 * c] = bc == c ? tk : A[r + 3 * c];
            A[r + 3 * k] = tb;
        }
#pragma unroll
        for (int r = k + 1; r < 3; r++) A[r + 3 * k] = A[r + 3 * k] / A[k + 3 * k];
#pragma unroll
        for (int c = k + 1; c < 3; c++)
#pragma unroll
            for (int r = k + 1; r < 3; r++) A[r + 3 * c] = A[r + 3 * c] - A[r + 3 * k] * A[k + 3 * c];
    }
    const double thr = maxpivot * (2.220446049250313e-16 * 3.0);
    int rank = 0;
#pragma unroll
    for (int i = 0; i < 3; i++) rank += (i < nonzero && fabs(A[i + 3 * i]) > thr) ? 1 : 0;
    if (rank == 0) {
        x[0] = x[1] = x[2] = 0.0;
        return;
    }
    double c[3] = { b[0], b[1], b[2] };
#pragma unroll
    for (int k = 0; k < 3; k++) { // row transpositions in order
        const double tk = c[k];
        double tb = tk;
#pragma unroll
        for (int r = k + 1; r < 3; r++) tb = rowt[k] == r ? c[r] : tb;
#pragma unroll
        for (int r = k + 1; r < 3; r++) c[r] = rowt[k] == r ? tk : c[r];
        c[k] = tb;
    }
#pragma unroll
    for (int i = 0; i < 3; i++)
        if (c[i] != 0.0)
#pragma unroll
            for (int r = i + 1; r < 3; r++) c[r] = c[r] - c[i] * A[r + 3 * i];
#pragma unroll
    for (int i = 2; i >= 0; i--)
        if (i < rank && c[i] != 0.0) {
            c[i] = c[i] / A[i + 3 * i];
#pragma unroll
            for (int r = 0; r < i; r++) c[r] = c[r] - c[i] * A[r + 3 * i];
        }
    // x[q[i]] = c[i] (i < rank), 0 otherwise; q = column transpositions
    int q[3] = { 0, 1, 2 };
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int tk = q[k];
        int tb = tk;
#pragma unroll
        for (int r = k + 1; r < 3; r++) tb = colt[k] == r ? q[r] : tb;
#pragma unroll
        for (int r = k + 1; r < 3; r++) q[r] = colt[k] == r ? tk : q[r];
        q[k] = tb;
    }
#pragma unroll
    for (int j = 0; j < 3; j++) {
        double v = 0.0;
#pragma unroll
        for (int i = 0; i < 3; i++) v = (q[i] == j && i < rank) ? c[i] : v;
        x[j] = v;
    }
}

// the solve-and-test half of intersection.cpp:127-137
__device__ __forceinline__ bool edge_tri_solve(const d3& e0, const d3& e1, const d3& t0,
                                               const d3& t1, const d3& t2)
{
    double M[9] = { t1.x - t0.x, t1.y - t0.y, t1.z - t0.z, t2.x - t0.x, t2.y - t0.y,
                    t2.z - t0.z, e0.x - e1.x, e0.y - e1.y, e0.z - e1.z };
    const double rhs[3] = { e0.x - t0.x, e0.y - t0.y, e0.z - t0.z };
    double uvt[3];
    fullpivlu_solve3(M, rhs, uvt);
    return uvt[0] >= 0.0 && uvt[1] >= 0.0 && uvt[0] + uvt[1] <= 1.0 && uvt[2] >= 0.0
        && uvt[2] <= 1.0;
}

} // namespace ipcg
