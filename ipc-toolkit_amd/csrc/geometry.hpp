// Device geometry of the contact path: distance classification and squared
// distances (bit-reproducible), closed-form derivatives, mollifier, barrier.
//
// Every function that decides a discrete outcome (distance type, activity
// d^2 < (dhat+dmin)^2, mollified or not) evaluates the same sequence of
// IEEE-754 double operations as the reference's Eigen expressions with
// EIGEN_DONT_VECTORIZE (CMakeLists.txt:279):
//   - length-3 reductions are unrolled as x0 + (x1 + x2)
//   - cross products as (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0)
//   - normalized() divides by sqrt(squaredNorm())
// and the library is compiled with -ffp-contract=off so no a*b+c is fused.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#define IPCG_HD __host__ __device__ __forceinline__

namespace ipcg {

// ---- enums: same numeric values as the reference (distance_type.hpp) -----
enum PEType : int { PE_P_E0 = 0, PE_P_E1 = 1, PE_P_E = 2 };
enum PTType : int {
    PT_P_T0 = 0, PT_P_T1, PT_P_T2, PT_P_E0, PT_P_E1, PT_P_E2, PT_P_T
};
enum EEType : int {
    EE_EA0_EB0 = 0, EE_EA0_EB1, EE_EA1_EB0, EE_EA1_EB1, EE_EA_EB0,
    EE_EA_EB1, EE_EA0_EB, EE_EA1_EB, EE_EA_EB
};

struct d3 {
    double x, y, z;
};

// Positions live in HBM as one 32-byte record per vertex (x, y, z, unused):
// a gathered vertex is one aligned access to one cache line.
constexpr int PSTRIDE = 4;

__device__ __forceinline__ d3 ldx(const double* X, int /*nV*/, int i)
{
    const double* p = X + PSTRIDE * (int64_t)i;
    const double2 a = *reinterpret_cast<const double2*>(p);
    return { a.x, a.y, p[2] };
}

IPCG_HD d3 sub(const d3& a, const d3& b) { return { a.x - b.x, a.y - b.y, a.z - b.z }; }
IPCG_HD d3 add(const d3& a, const d3& b) { return { a.x + b.x, a.y + b.y, a.z + b.z }; }
IPCG_HD d3 scale(const d3& a, double s) { return { a.x * s, a.y * s, a.z * s }; }
IPCG_HD double dot(const d3& a, const d3& b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
IPCG_HD double sqn(const d3& a) { return dot(a, a); }
IPCG_HD d3 cross(const d3& a, const d3& b)
{
    return { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}

// ---------------------------------------------------------------------------
// Distance classification (distance_type.cpp)
// ---------------------------------------------------------------------------

// Solve the 2x2 symmetric system the way Eigen's LDLT<Matrix2d> does:
// pivot on the larger |diagonal| (the first on ties), all-zero diagonal
// keeps the identity permutation, pivots with |d| <= DBL_MIN give 0.
IPCG_HD void solve_sym2_ldlt(double a00, double a10, double a11, double b0,
                             double b1, double& x0, double& x1)
{
    bool perm = fabs(a11) > fabs(a00);
    const double p00 = perm ? a11 : a00;
    const double p11 = perm ? a00 : a11;
    double d0 = p00, d1, l = a10;
    if (fabs(d0) > 0.0) {
        l = a10 / d0;
        d1 = p11 - l * (d0 * l);
    } else {
        perm = false;
        d0 = a00;
        d1 = a11;
    }
    double y0 = perm ? b1 : b0;
    double y1 = perm ? b0 : b1;
    y1 = y1 - l * y0;
    y0 = fabs(d0) > DBL_MIN ? y0 / d0 : 0.0;
    y1 = fabs(d1) > DBL_MIN ? y1 / d1 : 0.0;
    y0 = y0 - l * y1;
    x0 = perm ? y1 : y0;
    x1 = perm ? y0 : y1;
}

// Coordinates of p relative to edge (a -> b) of a triangle with normal n in
// the basis (b - a, (b - a) x n): point_triangle_distance_type's param column.
IPCG_HD void edge_frame_coords(const d3& p, const d3& a, const d3& b,
                               const d3& n, double& s, double& h)
{
    const d3 e = sub(b, a);
    const d3 f = cross(e, n);
    const d3 r = sub(p, a);
    solve_sym2_ldlt(dot(e, e), dot(f, e), dot(f, f), dot(e, r), dot(f, r), s, h);
}

// distance_type.cpp:36-79
IPCG_HD int point_triangle_type(const d3& p, const d3& t0, const d3& t1,
                                const d3& t2)
{
    const d3 n = cross(sub(t1, t0), sub(t2, t0));
    double s0, h0, s1, h1, s2, h2;
    edge_frame_coords(p, t0, t1, n, s0, h0);
    if (s0 > 0.0 && s0 < 1.0 && h0 >= 0.0) return PT_P_E0;
    edge_frame_coords(p, t1, t2, n, s1, h1);
    if (s1 > 0.0 && s1 < 1.0 && h1 >= 0.0) return PT_P_E1;
    edge_frame_coords(p, t2, t0, n, s2, h2);
    if (s2 > 0.0 && s2 < 1.0 && h2 >= 0.0) return PT_P_E2;
    if (s0 <= 0.0 && s2 >= 1.0) return PT_P_T0;
    if (s1 <= 0.0 && s0 >= 1.0) return PT_P_T1;
    if (s2 <= 0.0 && s1 >= 1.0) return PT_P_T2;
    return PT_P_T;
}

// distance_type.cpp:10-34
IPCG_HD int point_edge_type(const d3& p, const d3& e0, const d3& e1)
{
    const d3 e = sub(e1, e0);
    const double len2 = sqn(e);
    if (len2 == 0) return PE_P_E0;
    const double ratio = dot(e, sub(p, e0)) / len2;
    return ratio < 0 ? PE_P_E0 : (ratio > 1 ? PE_P_E1 : PE_P_E);
}

// distance_type.cpp:176-211 (near-parallel edges)
IPCG_HD int edge_edge_parallel_type(const d3& ea0, const d3& ea1, const d3& eb0,
                                    const d3& eb1)
{
    const d3 ea = sub(ea1, ea0);
    const double alpha = dot(sub(eb0, ea0), ea) / sqn(ea);
    const double beta = dot(sub(eb1, ea0), ea) / sqn(ea);
    int ca, cb; // 0 / 1 endpoint, 2 interior
    if (alpha < 0) {
        ca = (0 <= beta && beta <= 1) ? 2 : 0;
        cb = (beta <= alpha) ? 0 : (beta <= 1 ? 1 : 2);
    } else if (alpha > 1) {
        ca = (0 <= beta && beta <= 1) ? 2 : 1;
        cb = (beta >= alpha) ? 0 : (0 <= beta ? 1 : 2);
    } else {
        ca = 2;
        cb = 0;
    }
    return cb < 2 ? (ca << 1 | cb) : (6 + ca);
}

// distance_type.cpp:82-174
IPCG_HD int edge_edge_type(const d3& ea0, const d3& ea1, const d3& eb0,
                           const d3& eb1)
{
    const d3 u = sub(ea1, ea0), v = sub(eb1, eb0), w = sub(ea0, eb0);
    const double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w),
                 e = dot(v, w);
    const double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EE_EA0_EB0;
    if (a == 0.0) return EE_EA0_EB;
    if (c == 0.0) return EE_EA_EB0;
    const double ptol = 2.5e-16 * a * c;
    const double s2 = sqn(cross(u, v));
    if (s2 < ptol) return edge_edge_parallel_type(ea0, ea1, eb0, eb1);

    int def = EE_EA_EB;
    const double sN = b * e - c * d;
    double tN, tD;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        def = EE_EA0_EB;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        def = EE_EA1_EB;
    } else {
        tN = a * e - b * d;
        tD = D;
        if (tN > 0.0 && tN < tD && s2 < ptol) {
            if (sN < D / 2) {
                tN = e;
                tD = c;
                def = EE_EA0_EB;
            } else {
                tN = e + b;
                tD = c;
                def = EE_EA1_EB;
            }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return EE_EA0_EB0;
        if (-d >= a) return EE_EA1_EB0;
        return EE_EA_EB0;
    }
    if (tN >= tD) {
        if ((-d + b) <= 0.0) return EE_EA0_EB1;
        if ((-d + b) >= a) return EE_EA1_EB1;
        return EE_EA_EB1;
    }
    return def;
}

// ---------------------------------------------------------------------------
// Squared distances (point_point.cpp, point_line.cpp, point_plane.cpp,
// line_line.cpp)
// ---------------------------------------------------------------------------

IPCG_HD double pp_dist2(const d3& p0, const d3& p1) { return sqn(sub(p1, p0)); }

IPCG_HD double pl_dist2(const d3& p, const d3& e0, const d3& e1)
{
    return sqn(cross(sub(e0, p), sub(e1, p))) / sqn(sub(e1, e0));
}

IPCG_HD double pplane_dist2(const d3& p, const d3& t0, const d3& t1, const d3& t2)
{
    d3 n = cross(sub(t1, t0), sub(t2, t0));
    const double z = sqn(n);
    if (z > 0) {
        const double r = sqrt(z);
        n = { n.x / r, n.y / r, n.z / r };
    }
    const double s = dot(sub(p, t0), n);
    return s * s / sqn(n);
}

IPCG_HD double ll_dist2(const d3& ea0, const d3& ea1, const d3& eb0, const d3& eb1)
{
    const d3 n = cross(sub(ea1, ea0), sub(eb1, eb0));
    const double s = dot(sub(eb0, ea0), n);
    return s * s / sqn(n);
}

IPCG_HD double pt_dist2(const d3& p, const d3& t0, const d3& t1, const d3& t2, int dt)
{
    switch (dt) {
    case PT_P_T0: return pp_dist2(p, t0);
    case PT_P_T1: return pp_dist2(p, t1);
    case PT_P_T2: return pp_dist2(p, t2);
    case PT_P_E0: return pl_dist2(p, t0, t1);
    case PT_P_E1: return pl_dist2(p, t1, t2);
    case PT_P_E2: return pl_dist2(p, t2, t0);
    default: return pplane_dist2(p, t0, t1, t2);
    }
}

IPCG_HD double ee_dist2(const d3& ea0, const d3& ea1, const d3& eb0, const d3& eb1,
                        int dt)
{
    switch (dt) {
    case EE_EA0_EB0: return pp_dist2(ea0, eb0);
    case EE_EA0_EB1: return pp_dist2(ea0, eb1);
    case EE_EA1_EB0: return pp_dist2(ea1, eb0);
    case EE_EA1_EB1: return pp_dist2(ea1, eb1);
    case EE_EA_EB0: return pl_dist2(eb0, ea0, ea1);
    case EE_EA_EB1: return pl_dist2(eb1, ea0, ea1);
    case EE_EA0_EB: return pl_dist2(ea0, eb0, eb1);
    case EE_EA1_EB: return pl_dist2(ea1, eb0, eb1);
    default: return ll_dist2(ea0, ea1, eb0, eb1);
    }
}

IPCG_HD double pe_dist2(const d3& p, const d3& e0, const d3& e1, int dt)
{
    return dt == PE_P_E0 ? pp_dist2(p, e0) : (dt == PE_P_E1 ? pp_dist2(p, e1) : pl_dist2(p, e0, e1));
}

// edge_edge_mollifier.cpp:7-14 and :194-202
IPCG_HD double ee_cross2(const d3& ea0, const d3& ea1, const d3& eb0, const d3& eb1)
{
    return sqn(cross(sub(ea1, ea0), sub(eb1, eb0)));
}

IPCG_HD double ee_mollifier_threshold(const d3& ra0, const d3& ra1, const d3& rb0,
                                      const d3& rb1)
{
    return 1e-3 * sqn(sub(ra0, ra1)) * sqn(sub(rb0, rb1));
}

// ---------------------------------------------------------------------------
// Barrier (barrier.cpp:11-43, barrier_potential.cpp:62-99)
// ---------------------------------------------------------------------------

IPCG_HD double barrier_b(double d, double dh)
{
    if (d <= 0.0) return INFINITY;
    if (d >= dh) return 0;
    const double dm = d - dh;
    return -dm * dm * log(d / dh);
}

IPCG_HD double barrier_db(double d, double dh)
{
    if (d <= 0.0 || d >= dh) return 0.0;
    return (dh - d) * (2 * log(d / dh) - dh / d + 1);
}

IPCG_HD double barrier_d2b(double d, double dh)
{
    if (d <= 0.0 || d >= dh) return 0.0;
    const double q = dh / d;
    return (q + 2) * q - 2 * log(d / dh) - 3;
}

} // namespace ipcg
