// Tight-Inclusion continuous collision detection for one candidate, run by
// one lane over a private slab of the CCD work buffer, plus the IPC Toolkit
// strategy around it (tight_inclusion_ccd.cpp:33-336).
//
// Tight-Inclusion (v1.0.6, third party) is restated from its published
// algorithm (Wang, Ferguson, Schneider, Panozzo, ACM TOG 40(5) 2021):
// breadth-first bisection of the (t,u,v) box [0,1]^3, the corner inclusion
// of the multilinear F(t,u,v), the floating-point filter epsilon box, the
// per-level termination rule.  Within a level boxes are visited in
// increasing lower t, ties first-in first-out.
//
// GPU form: all boxes of a level share their widths (the split axis is a
// function of the widths only), so a box is its lower corner (t0,u0,v0) and
// a level is a sorted array.  A t-split level is built by merging the lower
// and upper halves (each already sorted); a u/v-split level is the children
// in push order.  The lower t of every box is dyadic, so all corners are
// exact and the corner evaluations reproduce the CPU restatement bit for
// bit.  `bound` hides outcomes >= bound (boxes starting at t >= bound are
// dropped), which leaves every outcome below bound unchanged because a
// level is visited in increasing t.
#pragma once

#include "geometry.hpp"

namespace ipcg {

struct TIBox {
    double t, u, v;
};

struct TISlab {
    TIBox* cur;  // level being visited
    TIBox* lo;   // next level, lower halves (t split) or all children
    TIBox* hi;   // next level, upper halves (t split)
    int cap;     // capacity of each list
};

// TI_OVERFLOW: a level outgrew the lists; TI_DEEP: a box needs more
// halvings than the packed LDS form holds (both: search again on the slab)
enum TIResult : int { TI_MISS = 0, TI_HIT = 1, TI_OVERFLOW = 2, TI_DEEP = 3 };

// Guards that keep every search finite on the GPU: a level cap (each level
// halves one width, so a search that resolves needs far fewer) and a cap on
// the no_zero_toi retries (the reference allows INT_MAX; both restatements
// stop after this many halvings).
constexpr int TI_MAX_LEVELS = 1024;
constexpr int TI_MAX_NO_ZERO_TOI_ITER = 64;

// F(t,u,v) along one axis.  x0/x1: 12 doubles at t=0 / t=1 in the order
// (ea0, ea1, eb0, eb1) or (p, t0, t1, t2).
IPCG_HD double ti_f_ee(const double* s, const double* e, int d, double t, double u, double v)
{
    const double a0 = (e[0 + d] - s[0 + d]) * t + s[0 + d];
    const double a1 = (e[3 + d] - s[3 + d]) * t + s[3 + d];
    const double b0 = (e[6 + d] - s[6 + d]) * t + s[6 + d];
    const double b1 = (e[9 + d] - s[9 + d]) * t + s[9 + d];
    const double va = (a1 - a0) * u + a0;
    const double vb = (b1 - b0) * v + b0;
    return va - vb;
}

IPCG_HD double ti_f_vf(const double* s, const double* e, int d, double t, double u, double v)
{
    const double p = (e[0 + d] - s[0 + d]) * t + s[0 + d];
    const double t0 = (e[3 + d] - s[3 + d]) * t + s[3 + d];
    const double t1 = (e[6 + d] - s[6 + d]) * t + s[6 + d];
    const double t2 = (e[9 + d] - s[9 + d]) * t + s[9 + d];
    const double pt = (t1 - t0) * u + (t2 - t0) * v + t0;
    return p - pt;
}

struct TIQuery {
    double s[12], e[12];
    bool vf;
    double err[3];
    double tol[3];
};

IPCG_HD double linf3(const double* a, const double* b)
{
    double m = 0;
    for (int k = 0; k < 3; k++) {
        const double x = fabs(b[k] - a[k]);
        m = x > m ? x : m;
    }
    return m;
}

// Parameter tolerances delta / (3 max |dF| along t, u, v over cube corners)
IPCG_HD void ti_param_tol(TIQuery& q, double delta)
{
    double p[8][3];
    for (int k = 0; k < 3; k++) {
        const double* s = q.s;
        const double* e = q.e;
        if (q.vf) {
            p[0][k] = s[k] - s[3 + k];
            p[1][k] = s[k] - s[9 + k];
            p[3][k] = s[k] - (s[6 + k] + s[9 + k] - s[3 + k]);
            p[2][k] = s[k] - s[6 + k];
            p[4][k] = e[k] - e[3 + k];
            p[5][k] = e[k] - e[9 + k];
            p[7][k] = e[k] - (e[6 + k] + e[9 + k] - e[3 + k]);
            p[6][k] = e[k] - e[6 + k];
        } else {
            p[0][k] = s[k] - s[6 + k];
            p[1][k] = s[k] - s[9 + k];
            p[3][k] = s[3 + k] - s[9 + k];
            p[2][k] = s[3 + k] - s[6 + k];
            p[4][k] = e[k] - e[6 + k];
            p[5][k] = e[k] - e[9 + k];
            p[7][k] = e[3 + k] - e[9 + k];
            p[6][k] = e[3 + k] - e[6 + k];
        }
    }
    double dl = linf3(p[0], p[4]), x;
    if ((x = linf3(p[1], p[5])) > dl) dl = x;
    if ((x = linf3(p[3], p[7])) > dl) dl = x;
    if ((x = linf3(p[2], p[6])) > dl) dl = x;
    double lu = linf3(p[0], p[2]);
    if ((x = linf3(p[4], p[6])) > lu) lu = x;
    if ((x = linf3(p[5], p[7])) > lu) lu = x;
    if ((x = linf3(p[1], p[3])) > lu) lu = x;
    double lv = linf3(p[0], p[1]);
    if ((x = linf3(p[4], p[5])) > lv) lv = x;
    if ((x = linf3(p[6], p[7])) > lv) lv = x;
    if ((x = linf3(p[2], p[3])) > lv) lv = x;
    q.tol[0] = delta / (3 * dl);
    q.tol[1] = delta / (3 * lu);
    q.tol[2] = delta / (3 * lv);
}

// Published floating-point filter: max(|x|, 1)^3 * c per axis.
IPCG_HD void ti_filter(TIQuery& q, bool use_ms)
{
    const double c = q.vf ? (use_ms ? 7.549516567451064e-15 : 6.661338147750939e-15)
                          : (use_ms ? 7.105427357601002e-15 : 6.217248937900877e-15);
    for (int k = 0; k < 3; k++) {
        double m = fabs(q.s[k]);
        for (int i = 0; i < 4; i++) {
            const double a = fabs(q.s[3 * i + k]), b = fabs(q.e[3 * i + k]);
            m = a > m ? a : m;
            m = b > m ? b : m;
        }
        const double dl = m > 1 ? m : 1;
        q.err[k] = dl * dl * dl * c;
    }
}

// Inclusion of the origin in the corner box of F over [t,t+wt]x[u,u+wu]x
// [v,v+wv], inflated by err and ms.  The kind is a template argument so the
// eight corner evaluations of an axis are one straight block (shared
// subexpressions computed once), not eight branches on q.vf.
template <bool VF>
IPCG_HD bool ti_inclusion_k(const TIQuery& q, const TIBox& b, double wt, double wu, double wv,
                            double ms, bool& box_in, double* width)
{
    box_in = true;
    const double t1 = b.t + wt, u1 = b.u + wu, v1 = b.v + wv;
    for (int d = 0; d < 3; d++) {
        double mn = 0, mx = 0;
#pragma unroll
        for (int c = 0; c < 8; c++) {
            const double t = (c & 4) ? t1 : b.t;
            const double u = (c & 2) ? u1 : b.u;
            const double v = (c & 1) ? v1 : b.v;
            const double f = VF ? ti_f_vf(q.s, q.e, d, t, u, v) : ti_f_ee(q.s, q.e, d, t, u, v);
            mn = (c == 0 || f < mn) ? f : mn;
            mx = (c == 0 || f > mx) ? f : mx;
        }
        width[d] = mx - mn;
        const double eps = q.err[d];
        if (mn - ms > eps || mx + ms < -eps) return false;
        if (!(mn + ms >= -eps && mx - ms <= eps)) box_in = false;
    }
    return true;
}

IPCG_HD bool ti_inclusion(const TIQuery& q, const TIBox& b, double wt, double wu, double wv,
                          double ms, bool& box_in, double* width)
{
    return q.vf ? ti_inclusion_k<true>(q, b, wt, wu, wv, ms, box_in, width)
                : ti_inclusion_k<false>(q, b, wt, wu, wv, ms, box_in, width);
}

// One breadth-first root search.  Returns TI_HIT (toi set), TI_MISS, or
// TI_OVERFLOW when a level does not fit the slab.
// Outcomes at or above max(bound_scale * (*bound_ptr), bound_floor) are
// hidden (bound_ptr null: no bound).
struct TIBound {
    const double* ptr;
    double scale, floor;
    __device__ double get() const
    {
        if (!ptr) return INFINITY;
        const double g = __hip_atomic_load(ptr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const double b = g * scale;
        return b > floor ? b : floor;
    }
};

__device__ inline int ti_bfs(const TIQuery& q, double co_tol, double ms, long max_itr,
                             const TIBound& B, const TISlab& S, double& toi, double& out_tol)
{
    out_tol = co_tol;
    double temp_out_tol = co_tol, temp_toi = INFINITY;
    double toi_skip = 4;
    bool use_skip = false;
    long refine = 0;
    double wt = 1, wu = 1, wv = 1; // widths of the current level
    int ncur = 1;
    S.cur[0] = { 0.0, 0.0, 0.0 };
    toi = INFINITY;
    int levels = 0;
    while (ncur > 0) {
        if (++levels > TI_MAX_LEVELS) return TI_OVERFLOW;
        const double bound = B.get();
        bool level_less_tol = true, level_root = false;
        const bool c1 = wt <= q.tol[0] && wu <= q.tol[1] && wv <= q.tol[2];
        // axis to split this level's boxes along (first max of width/tol)
        int axis = 0;
        {
            double best = -INFINITY;
            const double w3[3] = { wt, wu, wv };
            for (int i = 0; i < 3; i++) {
                const double r = w3[i] > q.tol[i] ? w3[i] / q.tol[i] : -INFINITY;
                if (r > best) {
                    best = r;
                    axis = i;
                }
            }
        }
        int nlo = 0, nhi = 0;
        for (int i = 0; i < ncur; i++) {
            const TIBox b = S.cur[i];
            if (b.t >= toi_skip || b.t >= bound) continue;
            refine++;
            bool box_in;
            double tw[3];
            if (!ti_inclusion(q, b, wt, wu, wv, ms, box_in, tw)) continue;
            const bool tol_cond = tw[0] <= co_tol && tw[1] <= co_tol && tw[2] <= co_tol;
            const bool c2 = box_in && level_less_tol;
            if (!tol_cond) level_less_tol = false;
            if (c1 || c2 || level_less_tol) {
                toi = b.t;
                return TI_HIT;
            }
            if (max_itr > 0) {
                if (!level_root) {
                    temp_toi = b.t;
                    double m = co_tol;
                    for (int k = 0; k < 3; k++) m = tw[k] > m ? tw[k] : m;
                    temp_out_tol = m;
                    level_root = true;
                }
                if (refine > max_itr) {
                    toi = temp_toi;
                    out_tol = temp_out_tol;
                    return TI_HIT;
                }
            }
            if (tol_cond || box_in) {
                toi_skip = b.t < toi_skip ? b.t : toi_skip;
                use_skip = true;
                continue;
            }
            if (axis == 0) {
                if (nlo >= S.cap || nhi >= S.cap) return TI_OVERFLOW;
                S.hi[nhi++] = { b.t + 0.5 * wt, b.u, b.v };
                S.lo[nlo++] = b;
            } else {
                TIBox h = b;
                if (axis == 1) h.u = b.u + 0.5 * wu;
                else h.v = b.v + 0.5 * wv;
                if (!q.vf || h.u + h.v <= 1) {
                    if (nlo >= S.cap) return TI_OVERFLOW;
                    S.lo[nlo++] = h;
                }
                if (!q.vf || b.u + b.v <= 1) {
                    if (nlo >= S.cap) return TI_OVERFLOW;
                    S.lo[nlo++] = b;
                }
            }
        }
        // next level
        if (axis == 0) {
            wt *= 0.5;
            if (nlo + nhi > S.cap) return TI_OVERFLOW;
            int i = 0, j = 0, k = 0;
            while (i < nlo && j < nhi) {
                if (S.hi[j].t <= S.lo[i].t) S.cur[k++] = S.hi[j++];
                else S.cur[k++] = S.lo[i++];
            }
            while (i < nlo) S.cur[k++] = S.lo[i++];
            while (j < nhi) S.cur[k++] = S.hi[j++];
            ncur = k;
        } else {
            if (axis == 1) wu *= 0.5;
            else wv *= 0.5;
            for (int i = 0; i < nlo; i++) S.cur[i] = S.lo[i];
            ncur = nlo;
        }
    }
    if (use_skip) {
        toi = toi_skip;
        return TI_HIT;
    }
    return TI_MISS;
}

// ---------------------------------------------------------------------------
// Level lists in LDS.  Every box of a level is its lower corner counted in
// the level widths 2^-kt, 2^-ku, 2^-kv, so TI_LDS_BITS bits per axis hold
// it exactly while no axis has been halved more often than that; a search
// that needs more halvings, or more than `cap` boxes in a list, returns
// TI_OVERFLOW (the caller runs it again on the HBM slab).  The decoded
// corners are the same doubles the slab search visits, so both searches
// give the same outcome.
// ---------------------------------------------------------------------------
constexpr int TI_LDS_BITS = 21;
constexpr unsigned TI_LDS_MASK = (1u << TI_LDS_BITS) - 1;

struct TILdsLists {
    unsigned long long* base; // this lane's slot 0 of list 0
    int stride;               // elements between a lane's consecutive slots
    int cap;                  // boxes per list
    __device__ unsigned long long& at(int list, int j) const
    {
        return base[(list * cap + j) * stride];
    }
};

__device__ __forceinline__ unsigned long long ti_pack(unsigned it, unsigned iu, unsigned iv)
{
    return ((unsigned long long)it << (2 * TI_LDS_BITS)) |
           ((unsigned long long)iu << TI_LDS_BITS) | iv;
}

__device__ inline int ti_bfs_lds(const TIQuery& q, double co_tol, double ms, long max_itr,
                                 const TIBound& B, const TILdsLists& L, double& toi,
                                 double& out_tol)
{
    out_tol = co_tol;
    double temp_out_tol = co_tol, temp_toi = INFINITY;
    double toi_skip = 4;
    bool use_skip = false;
    long refine = 0;
    int kt = 0, ku = 0, kv = 0; // the level's widths are 2^-k
    int ncur = 1;
    L.at(0, 0) = 0;
    toi = INFINITY;
    // The running minimum only decreases; a value read once per search
    // hides a subset of what a fresher read hides, all at or above the final
    // minimum.
    const double bound = B.get();
    while (ncur > 0) {
        const double wt = ldexp(1.0, -kt), wu = ldexp(1.0, -ku), wv = ldexp(1.0, -kv);
        bool level_less_tol = true, level_root = false;
        const bool c1 = wt <= q.tol[0] && wu <= q.tol[1] && wv <= q.tol[2];
        int axis = 0;
        {
            double best = -INFINITY;
            const double w3[3] = { wt, wu, wv };
            for (int i = 0; i < 3; i++) {
                const double r = w3[i] > q.tol[i] ? w3[i] / q.tol[i] : -INFINITY;
                if (r > best) {
                    best = r;
                    axis = i;
                }
            }
        }
        const int nk = (axis == 0 ? kt : axis == 1 ? ku : kv) + 1;
        int nlo = 0, nhi = 0;
        for (int i = 0; i < ncur; i++) {
            const unsigned long long pb = L.at(0, i);
            const unsigned it = (unsigned)(pb >> (2 * TI_LDS_BITS));
            const unsigned iu = (unsigned)(pb >> TI_LDS_BITS) & TI_LDS_MASK;
            const unsigned iv = (unsigned)pb & TI_LDS_MASK;
            const TIBox b { ldexp((double)it, -kt), ldexp((double)iu, -ku),
                            ldexp((double)iv, -kv) };
            if (b.t >= toi_skip || b.t >= bound) continue;
            refine++;
            bool box_in;
            double tw[3];
            if (!ti_inclusion(q, b, wt, wu, wv, ms, box_in, tw)) continue;
            const bool tol_cond = tw[0] <= co_tol && tw[1] <= co_tol && tw[2] <= co_tol;
            const bool c2 = box_in && level_less_tol;
            if (!tol_cond) level_less_tol = false;
            if (c1 || c2 || level_less_tol) {
                toi = b.t;
                return TI_HIT;
            }
            if (max_itr > 0) {
                if (!level_root) {
                    temp_toi = b.t;
                    double m = co_tol;
                    for (int k = 0; k < 3; k++) m = tw[k] > m ? tw[k] : m;
                    temp_out_tol = m;
                    level_root = true;
                }
                if (refine > max_itr) {
                    toi = temp_toi;
                    out_tol = temp_out_tol;
                    return TI_HIT;
                }
            }
            if (tol_cond || box_in) {
                toi_skip = b.t < toi_skip ? b.t : toi_skip;
                use_skip = true;
                continue;
            }
            if (nk > TI_LDS_BITS) return TI_DEEP;
            if (axis == 0) {
                if (nlo >= L.cap || nhi >= L.cap) return TI_OVERFLOW;
                L.at(2, nhi++) = ti_pack(2 * it + 1, iu, iv);
                L.at(1, nlo++) = ti_pack(2 * it, iu, iv);
            } else {
                // upper half first, each kept when it can hold a point of the
                // triangle (vertex-face)
                const unsigned hu = axis == 1 ? 2 * iu + 1 : iu, hv = axis == 1 ? iv : 2 * iv + 1;
                const unsigned lu = axis == 1 ? 2 * iu : iu, lv = axis == 1 ? iv : 2 * iv;
                const int nku = axis == 1 ? nk : ku, nkv = axis == 1 ? kv : nk;
                if (!q.vf || ldexp((double)hu, -nku) + ldexp((double)hv, -nkv) <= 1) {
                    if (nlo >= L.cap) return TI_OVERFLOW;
                    L.at(1, nlo++) = ti_pack(it, hu, hv);
                }
                if (!q.vf || b.u + b.v <= 1) {
                    if (nlo >= L.cap) return TI_OVERFLOW;
                    L.at(1, nlo++) = ti_pack(it, lu, lv);
                }
            }
        }
        // next level
        if (axis == 0) {
            kt = nk;
            if (nlo + nhi > L.cap) return TI_OVERFLOW;
            int i = 0, j = 0, k = 0;
            while (i < nlo && j < nhi) {
                const unsigned long long a = L.at(1, i), h = L.at(2, j);
                if ((h >> (2 * TI_LDS_BITS)) <= (a >> (2 * TI_LDS_BITS))) {
                    L.at(0, k++) = h;
                    j++;
                } else {
                    L.at(0, k++) = a;
                    i++;
                }
            }
            while (i < nlo) L.at(0, k++) = L.at(1, i++);
            while (j < nhi) L.at(0, k++) = L.at(2, j++);
            ncur = k;
        } else {
            if (axis == 1) ku = nk;
            else kv = nk;
            for (int i = 0; i < nlo; i++) L.at(0, i) = L.at(1, i);
            ncur = nlo;
        }
    }
    if (use_skip) {
        toi = toi_skip;
        return TI_HIT;
    }
    return TI_MISS;
}

// ticcd::{edgeEdge,vertexFace}CCD with the automatic error filter and the
// no_zero_toi loop around one of the searches above.
template <class Search>
__device__ inline int ti_query_with(TIQuery& q, double ms, double tolerance, long max_itr,
                                    bool no_zero_toi, const Search& bfs, double& toi)
{
    ti_param_tol(q, tolerance);
    ti_filter(q, ms > 0);
    int r;
    double out_tol;
    for (int it = 0;; it++) {
        r = bfs(q, tolerance, ms, max_itr, toi, out_tol);
        if (it + 1 >= TI_MAX_NO_ZERO_TOI_ITER) break;
        if (!(no_zero_toi && r == TI_HIT && toi == 0)) break;
        if (out_tol > tolerance) {
            break; // unreachable with unlimited iterations (out_tol == tolerance)
        } else if (10 * tolerance < ms) {
            ms *= 0.5;
        } else {
            tolerance *= 0.5;
            ti_param_tol(q, tolerance);
        }
    }
    return r;
}

__device__ inline int ti_query(TIQuery& q, double ms, double tolerance, long max_itr,
                               bool no_zero_toi, const TIBound& B, const TISlab& S, double& toi)
{
    return ti_query_with(
        q, ms, tolerance, max_itr, no_zero_toi,
        [&](const TIQuery& qq, double co_tol, double m, long mi, double& t, double& ot) {
            return ti_bfs(qq, co_tol, m, mi, B, S, t, ot);
        },
        toi);
}

__device__ inline int ti_query_lds(TIQuery& q, double ms, double tolerance, long max_itr,
                                   bool no_zero_toi, const TIBound& B, const TILdsLists& L,
                                   double& toi)
{
    return ti_query_with(
        q, ms, tolerance, max_itr, no_zero_toi,
        [&](const TIQuery& qq, double co_tol, double m, long mi, double& t, double& ot) {
            return ti_bfs_lds(qq, co_tol, m, mi, B, L, t, ot);
        },
        toi);
}

} // namespace ipcg
