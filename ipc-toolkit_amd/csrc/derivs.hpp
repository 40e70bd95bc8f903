// Closed-form derivatives of the barrier energy of one collision stencil, in
// a translation-free orthonormal basis, with PSD projection.
//
// Every squared distance on the contact path is a function of point
// differences ("reduced variables", each P_plus - P_minus).  Derivatives are
// formed in those variables (quotient rule on d = N / D with N, D low-degree
// polynomials) and mapped into the n-1 dimensional subspace orthogonal to
// rigid translations, spanned by an orthonormal basis Q of the stencil's
// point space (Hadamard rows for 4 points, Helmert rows for 3, (1,-1)/sqrt2
// for 2).  The energy Hessian H_x of a stencil satisfies H_x = Q B Q^T with
// B = Q^T H_x Q (3(n-1) x 3(n-1)), and the translation directions are
// exact null vectors, so PSD(H_x) = Q PSD(B) Q^T: the eigen-problem of the
// reference's project_to_psd (eigen_ext.tpp:63-108) on 12x12 matrices is
// solved here on 9x9 ones, in registers, with cyclic Jacobi.
#pragma once

#include "geometry.hpp"

namespace ipcg {

struct m33 {
    double a[9];
};

IPCG_HD m33 m33_zero()
{
    m33 r;
#pragma unroll
    for (int i = 0; i < 9; i++) r.a[i] = 0;
    return r;
}
IPCG_HD m33 outer(const d3& u, const d3& v)
{
    return { { u.x * v.x, u.x * v.y, u.x * v.z, u.y * v.x, u.y * v.y, u.y * v.z,
               u.z * v.x, u.z * v.y, u.z * v.z } };
}
IPCG_HD m33 skew(const d3& c) // [c]x : [c]x w = c x w
{
    return { { 0, -c.z, c.y, c.z, 0, -c.x, -c.y, c.x, 0 } };
}
IPCG_HD m33 diag3(double s) { return { { s, 0, 0, 0, s, 0, 0, 0, s } }; }
IPCG_HD m33 transpose(const m33& m)
{
    return { { m.a[0], m.a[3], m.a[6], m.a[1], m.a[4], m.a[7], m.a[2], m.a[5], m.a[8] } };
}
IPCG_HD void axpy(m33& y, double s, const m33& x)
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int i = 0; i < 9; i++) y.a[i] += s * x.a[i];
}
IPCG_HD d3 axpy3(const d3& y, double s, const d3& x)
{
#pragma clang fp contract(fast)
    return { y.x + s * x.x, y.y + s * x.y, y.z + s * x.z };
}

// Reduced derivatives of a scalar: value, R gradient blocks, upper Hessian
// blocks H[A][B] (A <= B) stored at tri(A,B).
template <int R> struct Red {
    static constexpr int NB = R * (R + 1) / 2;
    double v;
    d3 g[R];
    m33 H[NB];
    static IPCG_HD constexpr int tri(int a, int b) { return a * R - a * (a - 1) / 2 + (b - a); }
};

// d = N / D by the quotient rule.
template <int R>
IPCG_HD void quotient(const Red<R>& N, const Red<R>& D, Red<R>& out)
{
#pragma clang fp contract(fast)
    const double iD = 1.0 / D.v, iD2 = iD * iD, iD3 = iD2 * iD;
    out.v = N.v * iD;
#pragma unroll
    for (int a = 0; a < R; a++)
        out.g[a] = axpy3(scale(N.g[a], iD), -N.v * iD2, D.g[a]);
#pragma unroll
    for (int a = 0; a < R; a++)
#pragma unroll
        for (int b = a; b < R; b++) {
            const int t = Red<R>::tri(a, b);
            m33 h = N.H[t];
#pragma unroll
            for (int i = 0; i < 9; i++) h.a[i] = h.a[i] * iD - N.v * iD2 * D.H[t].a[i];
            axpy(h, -iD2, outer(N.g[a], D.g[b]));
            axpy(h, -iD2, outer(D.g[a], N.g[b]));
            axpy(h, 2 * N.v * iD3, outer(D.g[a], D.g[b]));
            out.H[t] = h;
        }
}

// |u x v|^2 over reduced (u, v)
IPCG_HD void cross2_derivs(const d3& u, const d3& v, Red<2>& X)
{
    const d3 n = cross(u, v);
    X.v = sqn(n);
    X.g[0] = scale(cross(v, n), 2.0);
    X.g[1] = scale(cross(n, u), 2.0);
    m33 uu = diag3(2 * sqn(v));
    axpy(uu, -2.0, outer(v, v));
    m33 vv = diag3(2 * sqn(u));
    axpy(vv, -2.0, outer(u, u));
    m33 uv = diag3(-2 * dot(u, v));
    axpy(uv, 2.0, outer(u, v));
    axpy(uv, -2.0, skew(n));
    X.H[Red<2>::tri(0, 0)] = uu;
    X.H[Red<2>::tri(0, 1)] = uv;
    X.H[Red<2>::tri(1, 1)] = vv;
}

// point-point: d = |delta|^2
IPCG_HD void pp_derivs(const d3& delta, Red<1>& out)
{
    out.v = sqn(delta);
    out.g[0] = scale(delta, 2.0);
    out.H[0] = diag3(2.0);
}

// point-line: a = e0 - p, b = e1 - p, d = |a x b|^2 / |b - a|^2
IPCG_HD void pl_derivs(const d3& a, const d3& b, Red<2>& out)
{
    Red<2> N, D;
    cross2_derivs(a, b, N);
    const d3 e = sub(b, a);
    D.v = sqn(e);
    D.g[0] = scale(e, -2.0);
    D.g[1] = scale(e, 2.0);
    D.H[0] = diag3(2.0);
    D.H[1] = diag3(-2.0);
    D.H[2] = diag3(2.0);
    quotient<2>(N, D, out);
}

// point-plane / line-line: s = w . (u x v), d = s^2 / |u x v|^2
IPCG_HD void triple_derivs(const d3& w, const d3& u, const d3& v, Red<3>& out)
{
    const d3 n = cross(u, v);
    const double s = dot(w, n);
    const d3 gs[3] = { n, cross(v, w), cross(w, u) };
    Red<3> N, D;
    N.v = s * s;
#pragma unroll
    for (int a = 0; a < 3; a++) N.g[a] = scale(gs[a], 2 * s);
    // Hessian of s: (w,u) = -[v]x, (w,v) = [u]x, (u,v) = -[w]x, diagonal 0
    m33 hs[6];
    hs[Red<3>::tri(0, 0)] = m33_zero();
    hs[Red<3>::tri(1, 1)] = m33_zero();
    hs[Red<3>::tri(2, 2)] = m33_zero();
    hs[Red<3>::tri(0, 1)] = skew(scale(v, -1.0));
    hs[Red<3>::tri(0, 2)] = skew(u);
    hs[Red<3>::tri(1, 2)] = skew(scale(w, -1.0));
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = a; b < 3; b++) {
            const int t = Red<3>::tri(a, b);
            m33 h = outer(gs[a], gs[b]);
#pragma unroll
            for (int i = 0; i < 9; i++) h.a[i] = 2 * h.a[i] + 2 * s * hs[t].a[i];
            N.H[t] = h;
        }
    Red<2> X;
    cross2_derivs(u, v, X);
    D.v = X.v;
    D.g[0] = { 0, 0, 0 };
    D.g[1] = X.g[0];
    D.g[2] = X.g[1];
    D.H[Red<3>::tri(0, 0)] = m33_zero();
    D.H[Red<3>::tri(0, 1)] = m33_zero();
    D.H[Red<3>::tri(0, 2)] = m33_zero();
    D.H[Red<3>::tri(1, 1)] = X.H[0];
    D.H[Red<3>::tri(1, 2)] = X.H[1];
    D.H[Red<3>::tri(2, 2)] = X.H[2];
    quotient<3>(N, D, out);
}

// ---------------------------------------------------------------------------
// Orthonormal translation-free bases Q (n points x K = n-1 columns)
// ---------------------------------------------------------------------------
template <int NP> struct QBasis;
template <> struct QBasis<2> {
    static IPCG_HD double q(int x, int k) { return x == 0 ? M_SQRT1_2 : -M_SQRT1_2; }
};
template <> struct QBasis<3> {
    static IPCG_HD double q(int x, int k)
    {
        // (1,-1,0)/sqrt2, (1,1,-2)/sqrt6
        constexpr double r6 = 0.40824829046386301637; // 1/sqrt(6)
        if (k == 0) return x == 0 ? M_SQRT1_2 : (x == 1 ? -M_SQRT1_2 : 0.0);
        return x == 2 ? -2 * r6 : r6;
    }
};
template <> struct QBasis<4> {
    static IPCG_HD double q(int x, int k)
    {
        // rows of the 4x4 Hadamard matrix / 2 without the all-ones row
        const int s = (k == 0) ? (x & 1) : (k == 1 ? (x >> 1) : ((x ^ (x >> 1)) & 1));
        return s ? -0.5 : 0.5;
    }
};

// Packed symmetric N x N (lower), index of (i, j), i >= j
IPCG_HD constexpr int sp(int i, int j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// Q-space image of reduced derivatives: reduced variable i is
// P[plus[i]] - P[minus[i]], so S[i][k] = Q[plus][k] - Q[minus][k].
// gq = S^T g;  Hq += alpha * S^T H S  (packed 3K x 3K).
template <int NP, int R>
IPCG_HD void to_qspace(const Red<R>& d, const int* plus, const int* minus,
                       double alpha, d3* gq, double* Hq)
{
#pragma clang fp contract(fast)
    constexpr int K = NP - 1;
    double S[R][K];
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int k = 0; k < K; k++)
            S[i][k] = QBasis<NP>::q(plus[i], k) - QBasis<NP>::q(minus[i], k);
#pragma unroll
    for (int k = 0; k < K; k++) {
        d3 acc = { 0, 0, 0 };
#pragma unroll
        for (int i = 0; i < R; i++) acc = axpy3(acc, S[i][k], d.g[i]);
        gq[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < K; k++)
#pragma unroll
        for (int l = k; l < K; l++) {
            m33 blk = m33_zero();
#pragma unroll
            for (int i = 0; i < R; i++)
#pragma unroll
                for (int j = 0; j < R; j++) {
                    const double c = alpha * S[i][k] * S[j][l];
                    const m33 h = i <= j ? d.H[Red<R>::tri(i, j)]
                                         : transpose(d.H[Red<R>::tri(j, i)]);
                    axpy(blk, c, h);
                }
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const int I = 3 * k + r, J = 3 * l + c;
                    if (k < l || I >= J) Hq[sp(I, J)] += blk.a[3 * r + c];
                }
        }
}

// point-space gradient: g[plus[i]] += alpha g_i, g[minus[i]] -= alpha g_i
// (gp holds NP points; every point is addressed by a compile-time index so
// the array stays in registers, each point updated in the same order)
template <int NP, int R>
IPCG_HD void to_points(const Red<R>& d, const int* plus, const int* minus,
                       double alpha, d3* gp)
{
#pragma unroll
    for (int i = 0; i < R; i++)
#pragma unroll
        for (int j = 0; j < NP; j++) {
            if (plus[i] == j) gp[j] = axpy3(gp[j], alpha, d.g[i]);
            if (minus[i] == j) gp[j] = axpy3(gp[j], -alpha, d.g[i]);
        }
}

// Hq += alpha * (a b^T + b a^T) / (sym ? 2 : 1) on packed 3K storage, with
// a, b given as K blocks.
template <int K>
IPCG_HD void add_sym_outer(double* Hq, double alpha, const d3* a, const d3* b)
{
#pragma clang fp contract(fast)
    constexpr int N = 3 * K;
    double av[N], bv[N];
#pragma unroll
    for (int k = 0; k < K; k++) {
        av[3 * k] = a[k].x; av[3 * k + 1] = a[k].y; av[3 * k + 2] = a[k].z;
        bv[3 * k] = b[k].x; bv[3 * k + 1] = b[k].y; bv[3 * k + 2] = b[k].z;
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) Hq[sp(i, j)] += alpha * (av[i] * bv[j] + bv[i] * av[j]);
}

// ---------------------------------------------------------------------------
// 1/sqrt(x) for x in [1, 2]: the hardware estimate refined by two Newton
// steps (each squares the relative error), a few ulp from 1.0 / sqrt(x) at
// a fraction of the instructions of a correctly rounded sqrt and divide.
// ---------------------------------------------------------------------------
IPCG_HD double rsqrt_nr(double x)
{
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    r = r * fma(-h * r, r, 1.5);
    return r * fma(-h * r, r, 1.5);
#else
    return 1.0 / sqrt(x);
#endif
}

// ---------------------------------------------------------------------------
// PSD projection (cyclic Jacobi, eigenvectors accumulated in registers)
// method: 0 none, 1 clamp, 2 abs (PSDProjectionMethod, eigen_ext.hpp:202-206)
// ---------------------------------------------------------------------------
template <int N>
IPCG_HD void project_psd(double* A /* packed N */, int method)
{
#pragma clang fp contract(fast)
    if (method == 0) return;
    double V[N][N];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j < N; j++) V[i][j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; sweep++) {
        double off = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < N; i++) {
            tot += A[sp(i, i)] * A[sp(i, i)];
#pragma unroll
            for (int j = 0; j < i; j++) off += A[sp(i, j)] * A[sp(i, j)];
        }
        tot += 2 * off;
        if (!(off > 1e-30 * tot)) break;
#pragma unroll
        for (int p = 0; p < N; p++)
#pragma unroll
            for (int q = p + 1; q < N; q++) {
                const double apq = A[sp(p, q)];
                const double app = A[sp(p, p)], aqq = A[sp(q, q)];
                const double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                t = apq == 0.0 ? 0.0 : t;
                const double c = rsqrt_nr(t * t + 1), s = t * c; // |t| <= 1
#pragma unroll
                for (int k = 0; k < N; k++) {
                    if (k == p || k == q) continue;
                    const double akp = A[sp(k, p)], akq = A[sp(k, q)];
                    A[sp(k, p)] = c * akp - s * akq;
                    A[sp(k, q)] = s * akp + c * akq;
                }
                A[sp(p, p)] = app - t * apq;
                A[sp(q, q)] = aqq + t * apq;
                A[sp(p, q)] = 0.0;
#pragma unroll
                for (int k = 0; k < N; k++) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    double lam[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double l = A[sp(i, i)];
        lam[i] = l >= 0 ? l : (method == 1 ? 0.0 : -l);
    }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double acc = 0;
#pragma unroll
            for (int k = 0; k < N; k++) acc += V[i][k] * lam[k] * V[j][k];
            A[sp(i, j)] = acc;
        }
}

// ---------------------------------------------------------------------------
// PSD projection of a 9 x 9 matrix of rank at most KR (contact stencils of
// the point-triangle and edge-edge interior types have rank 5 in the
// translation-free basis: 3 positive and 2 negative eigenvalues).
//
// Lanczos from b = A w (w fixed, b in the range of A) with full
// reorthogonalization builds an orthonormal basis Q of the Krylov space
// until the next vector vanishes; the eigen-pairs of T = Q^T A Q (at most
// KR x KR, cyclic Jacobi) give A = (Q V) diag(lambda) (Q V)^T.  The result
// is accepted only if the residual |A - Q T Q^T|_F is at most 1e-12 |A|_F
// (the Krylov space holds the whole range of A); otherwise A is left
// untouched and false is returned (the caller runs project_psd<9>).
// ---------------------------------------------------------------------------
template <int KR>
IPCG_HD bool project_psd_krylov9(double* A /* packed 9 */, int method)
{
#pragma clang fp contract(fast)
    constexpr int N = 9;
    if (method == 0) return true;
    double nA2 = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) nA2 += (i == j ? 1.0 : 2.0) * A[sp(i, j)] * A[sp(i, j)];
    if (!(nA2 > 0)) return true; // the zero matrix is its own projection
    if (!(nA2 < 1e300)) return false;
    auto matvec = [&](const double* x, double* y) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            double acc = 0;
#pragma unroll
            for (int j = 0; j < N; j++) acc += A[sp(i, j)] * x[j];
            y[i] = acc;
        }
    };
    constexpr double w[N] = { 0.3719, -0.6541, 0.2333, 0.8123, -0.1427,
                              0.5021, -0.7433, 0.0679, 0.4517 };
    double Q[KR][N], al[KR], be[KR];
#pragma unroll
    for (int k = 0; k < KR; k++) {
        al[k] = be[k] = 0;
#pragma unroll
        for (int i = 0; i < N; i++) Q[k][i] = 0;
    }
    double u[N];
    matvec(w, u);
    const double tiny = 1e-13 * sqrt(nA2);
    double nb = 0;
#pragma unroll
    for (int i = 0; i < N; i++) nb += u[i] * u[i];
    nb = sqrt(nb);
    if (!(nb > tiny)) return false;
#pragma unroll
    for (int i = 0; i < N; i++) Q[0][i] = u[i] / nb;
    int m = 0; // Krylov dimension
    bool closed = false;
#pragma unroll
    for (int k = 0; k < KR; k++) {
        if (!closed) {
            matvec(Q[k], u);
            double a = 0;
#pragma unroll
            for (int i = 0; i < N; i++) a += Q[k][i] * u[i];
            al[k] = a;
            // full reorthogonalization, twice
#pragma unroll
            for (int pass = 0; pass < 2; pass++)
#pragma unroll
                for (int j = 0; j <= k; j++) {
                    double d = 0;
#pragma unroll
                    for (int i = 0; i < N; i++) d += Q[j][i] * u[i];
#pragma unroll
                    for (int i = 0; i < N; i++) u[i] -= d * Q[j][i];
                }
            double b = 0;
#pragma unroll
            for (int i = 0; i < N; i++) b += u[i] * u[i];
            b = sqrt(b);
            m = k + 1;
            if (b <= tiny) {
                closed = true;
            } else if (k + 1 < KR) {
                be[k] = b;
#pragma unroll
                for (int i = 0; i < N; i++) Q[k + 1][i] = u[i] / b;
            }
        }
    }
    if (!closed) return false;
    // T = Q^T A Q in the basis (tridiagonal up to rounding: use the exact
    // projection so the residual test sees the same matrix)
    double T[KR * (KR + 1) / 2];
#pragma unroll
    for (int k = 0; k < KR; k++) {
        matvec(Q[k], u);
#pragma unroll
        for (int l = 0; l <= k; l++) {
            double d = 0;
#pragma unroll
            for (int i = 0; i < N; i++) d += Q[l][i] * u[i];
            T[sp(k, l)] = d;
        }
    }
    (void)m;
    // residual |A - Q T Q^T|_F
    double r2 = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double x = A[sp(i, j)];
#pragma unroll
            for (int k = 0; k < KR; k++)
#pragma unroll
                for (int l = 0; l < KR; l++) x -= Q[k][i] * T[sp(k, l)] * Q[l][j];
            r2 += (i == j ? 1.0 : 2.0) * x * x;
        }
    if (!(r2 <= 1e-24 * nA2)) return false;
    // eigen-pairs of T (unused dimensions are zero rows: zero eigenvalues)
    double V[KR][KR];
#pragma unroll
    for (int i = 0; i < KR; i++)
#pragma unroll
        for (int j = 0; j < KR; j++) V[i][j] = i == j ? 1.0 : 0.0;
#pragma clang loop unroll(disable)
    for (int sweep = 0; sweep < 16; sweep++) {
        double off = 0, tot = 0;
#pragma unroll
        for (int i = 0; i < KR; i++) {
            tot += T[sp(i, i)] * T[sp(i, i)];
#pragma unroll
            for (int j = 0; j < i; j++) off += T[sp(i, j)] * T[sp(i, j)];
        }
        tot += 2 * off;
        if (!(off > 1e-30 * tot)) break;
#pragma unroll
        for (int p = 0; p < KR; p++)
#pragma unroll
            for (int q = p + 1; q < KR; q++) {
                const double apq = T[sp(p, q)];
                const double app = T[sp(p, p)], aqq = T[sp(q, q)];
                const double theta = (aqq - app) / (2 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
                t = apq == 0.0 ? 0.0 : t;
                const double c = rsqrt_nr(t * t + 1), s = t * c;
#pragma unroll
                for (int k = 0; k < KR; k++) {
                    if (k == p || k == q) continue;
                    const double akp = T[sp(k, p)], akq = T[sp(k, q)];
                    T[sp(k, p)] = c * akp - s * akq;
                    T[sp(k, q)] = s * akp + c * akq;
                }
                T[sp(p, p)] = app - t * apq;
                T[sp(q, q)] = aqq + t * apq;
                T[sp(p, q)] = 0.0;
#pragma unroll
                for (int k = 0; k < KR; k++) {
                    const double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
    }
    // W = Q^T V (N x KR columns), A = W diag(lambda+) W^T
    double lam[KR];
#pragma unroll
    for (int c = 0; c < KR; c++) {
        const double l = T[sp(c, c)];
        lam[c] = l >= 0 ? l : (method == 1 ? 0.0 : -l);
    }
    double W[N][KR];
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int c = 0; c < KR; c++) {
            double x = 0;
#pragma unroll
            for (int k = 0; k < KR; k++) x += Q[k][i] * V[k][c];
            W[i][c] = x;
        }
#pragma unroll
    for (int i = 0; i < N; i++)
#pragma unroll
        for (int j = 0; j <= i; j++) {
            double acc = 0;
#pragma unroll
            for (int c = 0; c < KR; c++) acc += W[i][c] * lam[c] * W[j][c];
            A[sp(i, j)] = acc;
        }
    return true;
}

} // namespace ipcg
