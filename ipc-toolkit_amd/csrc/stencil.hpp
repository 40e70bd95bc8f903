// Per-collision barrier energy, gradient and PSD-projected Hessian
// (normal_potential.cpp:127-232, barrier_potential.cpp:62-99).
//
// Stencils: VV (v0, v1) point-point; EV (v, e0, e1) point-line (P_E);
// EE (ea0, ea1, eb0, eb1) with the distance type stored at build and the
// edge-edge mollifier; FV (v, t0, t1, t2) point-plane (P_T).
// (candidates/*.hpp vertex_ids, known_dtype)
#pragma once

#include "derivs.hpp"

namespace ipcg {

enum StencilKind : int { K_VV = 0, K_EV = 1, K_EE = 2, K_FV = 3 };

struct BarrierParams {
    double dhat, kappa, dmin;
    int psd;
    // use_physical_barrier: b, b' and b'' scaled by dhat / units(dhat') with
    // units = dhat'^2 for the clamped log barrier (barrier_potential.cpp:
    // 62-99, barrier.hpp:133-137); 1 otherwise (an exact product)
    double phys = 1.0;
};

IPCG_HD double bar_f(double d2, const BarrierParams& b)
{
    return b.kappa * (barrier_b(d2 - b.dmin * b.dmin, (2 * b.dmin + b.dhat) * b.dhat) * b.phys);
}
IPCG_HD double bar_f1(double d2, const BarrierParams& b)
{
    return b.kappa * (barrier_db(d2 - b.dmin * b.dmin, (2 * b.dmin + b.dhat) * b.dhat) * b.phys);
}
IPCG_HD double bar_f2(double d2, const BarrierParams& b)
{
    return b.kappa * (barrier_d2b(d2 - b.dmin * b.dmin, (2 * b.dmin + b.dhat) * b.dhat) * b.phys);
}

// Squared distance of a stencil with its stored type (bit-exact path).
IPCG_HD double stencil_dist2(int kind, int dtype, const d3* P)
{
    switch (kind) {
    case K_VV: return pp_dist2(P[0], P[1]);
    case K_EV: return pl_dist2(P[0], P[1], P[2]);
    case K_FV: return pplane_dist2(P[0], P[1], P[2], P[3]);
    default: return ee_dist2(P[0], P[1], P[2], P[3], dtype);
    }
}

// Reduced derivatives of the stencil's squared distance; calls
// f(red, plus, minus) with the reduced variables P[plus[i]] - P[minus[i]].
template <int KIND> struct StencilTraits;
template <> struct StencilTraits<K_VV> { static constexpr int NP = 2; };
template <> struct StencilTraits<K_EV> { static constexpr int NP = 3; };
template <> struct StencilTraits<K_EE> { static constexpr int NP = 4; };
template <> struct StencilTraits<K_FV> { static constexpr int NP = 4; };

template <int KIND, class Fn>
IPCG_HD void with_distance_derivs(int dtype, const d3* P, Fn&& f)
{
    if constexpr (KIND == K_VV) {
        Red<1> r;
        pp_derivs(sub(P[0], P[1]), r);
        const int pl[1] = { 0 }, mi[1] = { 1 };
        f(r, pl, mi);
        return;
    } else if constexpr (KIND == K_EV) {
        Red<2> r;
        pl_derivs(sub(P[1], P[0]), sub(P[2], P[0]), r);
        const int pl[2] = { 1, 2 }, mi[2] = { 0, 0 };
        f(r, pl, mi);
        return;
    } else if constexpr (KIND == K_FV) {
        Red<3> r;
        triple_derivs(sub(P[0], P[1]), sub(P[2], P[1]), sub(P[3], P[1]), r);
        const int pl[3] = { 0, 2, 3 }, mi[3] = { 1, 1, 1 };
        f(r, pl, mi);
        return;
    } else {
    // edge-edge stencil (ea0, ea1, eb0, eb1) by stored distance type
    switch (dtype) {
    case EE_EA0_EB0:
    case EE_EA0_EB1:
    case EE_EA1_EB0:
    case EE_EA1_EB1: {
        const int a = (dtype >> 1) & 1, b = 2 + (dtype & 1);
        Red<1> r; // the points picked by selects: P stays in registers
        pp_derivs(sub(a ? P[1] : P[0], b == 3 ? P[3] : P[2]), r);
        const int pl[1] = { a }, mi[1] = { b };
        f(r, pl, mi);
        return;
    }
    case EE_EA_EB0:
    case EE_EA_EB1: { // point eb0/eb1 against line ea
        const int p = dtype == EE_EA_EB0 ? 2 : 3;
        const d3 Pp = p == 3 ? P[3] : P[2];
        Red<2> r;
        pl_derivs(sub(P[0], Pp), sub(P[1], Pp), r);
        const int pl[2] = { 0, 1 }, mi[2] = { p, p };
        f(r, pl, mi);
        return;
    }
    case EE_EA0_EB:
    case EE_EA1_EB: { // point ea0/ea1 against line eb
        const int p = dtype == EE_EA0_EB ? 0 : 1;
        const d3 Pp = p ? P[1] : P[0];
        Red<2> r;
        pl_derivs(sub(P[2], Pp), sub(P[3], Pp), r);
        const int pl[2] = { 2, 3 }, mi[2] = { p, p };
        f(r, pl, mi);
        return;
    }
    default: {
        Red<3> r;
        triple_derivs(sub(P[2], P[0]), sub(P[1], P[0]), sub(P[3], P[2]), r);
        const int pl[3] = { 2, 1, 3 }, mi[3] = { 0, 0, 2 };
        f(r, pl, mi);
        return;
    }
    }
    }
}

// Edge-edge mollifier m(x), x = |(ea1-ea0) x (eb1-eb0)|^2 < eps_x
// (edge_edge_mollifier.cpp:43-97): m, dm/dx, d2m/dx2.
struct Moll {
    double m, m1, m2;
    bool active;
};

IPCG_HD Moll mollifier(const d3* P, double eps_x)
{
    const double x = ee_cross2(P[0], P[1], P[2], P[3]);
    Moll r { 1.0, 0.0, 0.0, false };
    if (x < eps_x) {
        const double q = x / eps_x;
        r.m = (-q + 2) * q;
        const double ie = 1 / eps_x;
        r.m1 = 2 * ie * fma(-ie, x, 1.0);
        r.m2 = -2 / (eps_x * eps_x);
        r.active = true;
    }
    return r;
}

IPCG_HD double stencil_energy(int kind, int dtype, const d3* P, double w, double eps_x,
                              const BarrierParams& bp)
{
    const double d2 = stencil_dist2(kind, dtype, P);
    const double m = kind == K_EE ? mollifier(P, eps_x).m : 1.0;
    return w * m * bar_f(d2, bp);
}

// Point-space gradient (NP points) of w m f(d).
template <int KIND, int NP = StencilTraits<KIND>::NP>
IPCG_HD void stencil_gradient(int dtype, const d3* P, double w, double eps_x,
                              const BarrierParams& bp, d3* gp)
{
    constexpr int kind = KIND;
#pragma unroll
    for (int i = 0; i < NP; i++) gp[i] = { 0, 0, 0 };
    const Moll M = kind == K_EE ? mollifier(P, eps_x) : Moll { 1.0, 0.0, 0.0, false };
    if (M.m <= 0) return;
    const double d2 = stencil_dist2(kind, dtype, P);
    const double f1 = bar_f1(d2, bp);
    with_distance_derivs<KIND>(dtype, P, [&](auto& red, const int* pl, const int* mi) {
        to_points<NP>(red, pl, mi, w * M.m * f1, gp);
    });
    if constexpr (KIND == K_EE) {
        if (M.active) {
            Red<2> X;
            cross2_derivs(sub(P[1], P[0]), sub(P[3], P[2]), X);
            const int pl[2] = { 1, 3 }, mi[2] = { 0, 2 };
            to_points<NP>(X, pl, mi, w * bar_f(d2, bp) * M.m1, gp);
        }
    }
}

// PSD-projected energy Hessian in the Q basis (packed 3(NP-1) square).
// defer_psd: leave the projection to the caller; the return value says
// whether the reference projects this stencil's Hessian at all.
template <int KIND, int NP = StencilTraits<KIND>::NP>
IPCG_HD bool stencil_hessian_q(int dtype, const d3* P, double w, double eps_x,
                               const BarrierParams& bp, double* Hq, bool defer_psd = false)
{
    constexpr int kind = KIND;
    constexpr int K = NP - 1, N = 3 * K;
#pragma unroll
    for (int i = 0; i < N * (N + 1) / 2; i++) Hq[i] = 0;
    const double d2 = stencil_dist2(kind, dtype, P);
    const Moll M = kind == K_EE ? mollifier(P, eps_x) : Moll { 1.0, 0.0, 0.0, false };
    const int mpl[2] = { 1, 3 }, mmi[2] = { 0, 2 };
    if (kind == K_EE && M.m <= 0) { // normal_potential.cpp:188-192: no projection
        if constexpr (KIND == K_EE) {
            const double f = bar_f(d2, bp);
            Red<2> X;
            cross2_derivs(sub(P[1], P[0]), sub(P[3], P[2]), X);
            d3 gx[K];
            to_qspace<NP>(X, mpl, mmi, w * f * M.m1, gx, Hq);
            add_sym_outer<K>(Hq, 0.5 * w * f * M.m2, gx, gx);
        }
        return false;
    }
    const double f1 = bar_f1(d2, bp), f2 = bar_f2(d2, bp);
    d3 gq[K];
    with_distance_derivs<KIND>(dtype, P, [&](auto& red, const int* pl, const int* mi) {
        to_qspace<NP>(red, pl, mi, w * M.m * f1, gq, Hq);
    });
    add_sym_outer<K>(Hq, 0.5 * w * M.m * f2, gq, gq);
    if constexpr (KIND == K_EE) {
        if (M.active) {
            const double f = bar_f(d2, bp);
            Red<2> X;
            cross2_derivs(sub(P[1], P[0]), sub(P[3], P[2]), X);
            d3 gx[K];
            to_qspace<NP>(X, mpl, mmi, w * f * M.m1, gx, Hq);
            add_sym_outer<K>(Hq, 0.5 * w * f * M.m2, gx, gx);
            add_sym_outer<K>(Hq, w * f1 * M.m1, gq, gx);
        }
    }
    if (!defer_psd) project_psd<N>(Hq, bp.psd);
    return true;
}

} // namespace ipcg
