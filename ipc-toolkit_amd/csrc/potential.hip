// Barrier potential over the device collision set: energy (fixed-order
// reduction), gradient (per-vertex sums in a fixed order, RowMajor), and the
// PSD-projected Hessian assembled to CSR (potential.cpp:36-222).
//
// Hessian assembly (no global atomics).  Symbolic, once per collision set:
// the off-diagonal vertex pairs of all stencils are radix-sorted and
// run-length encoded; with one diagonal block per touched vertex they give a
// block-CSR pattern (bptr, sorted bcol); the (vertex, collision, point)
// incidences are radix-sorted by vertex.  Numeric: every collision writes
// its Hessian projected to PSD in the translation-free basis Q (3K x 3K,
// K = points - 1) to a record; one wave per vertex then expands the records
// of its incident collisions into its block row, summing in LDS with a
// per-entry "some contribution is nonzero" mask that reproduces the
// reference's rule of dropping exact-zero triplets before setFromTriplets
// (local_to_global.hpp:289-301), and writes the row; one wave per vertex
// finally emits its three CSR rows in column order.
#include "context.hpp"
#include "sort.hpp"
#include "stencil.hpp"

#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace ipcg {

namespace {

constexpr int PB = 128;

struct CollView {
    const int2* ids[4];
    const double* w[4];
    const double* eps;
    const int32_t* dtype;
    int64_t n[4];
    const int32_t* E;
    const int32_t* F;
};

// stencil vertex ids (CollisionStencil::vertex_ids)
template <int KIND>
__device__ __forceinline__ void stencil_ids(const CollView& C, int64_t i, int* v)
{
    const int2 p = C.ids[KIND][i];
    if constexpr (KIND == K_VV) {
        v[0] = p.x;
        v[1] = p.y;
    } else if constexpr (KIND == K_EV) {
        v[0] = p.y;
        v[1] = C.E[2 * (int64_t)p.x];
        v[2] = C.E[2 * (int64_t)p.x + 1];
    } else if constexpr (KIND == K_EE) {
        v[0] = C.E[2 * (int64_t)p.x];
        v[1] = C.E[2 * (int64_t)p.x + 1];
        v[2] = C.E[2 * (int64_t)p.y];
        v[3] = C.E[2 * (int64_t)p.y + 1];
    } else {
        v[0] = p.y;
        v[1] = C.F[3 * (int64_t)p.x];
        v[2] = C.F[3 * (int64_t)p.x + 1];
        v[3] = C.F[3 * (int64_t)p.x + 2];
    }
}

template <int KIND>
__device__ __forceinline__ double collision_energy(const CollView& C, const double* X, int nV,
                                                   const BarrierParams& bp, int64_t i)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    return stencil_energy(KIND, KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                          KIND == K_EE ? C.eps[i] : 0.0, bp);
}

// sum of the block's lane values (waves reduce by shuffles, then in order)
__device__ __forceinline__ void block_partial(double e, double* partial)
{
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, WAVE);
    __shared__ double s[PB / WAVE];
    if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int j = 0; j < PB / WAVE; j++) t += s[j];
        partial[blockIdx.x] = t;
    }
}

template <int KIND>
__global__ __launch_bounds__(PB) void k_energy(CollView C, const double* X, int nV,
                                               BarrierParams bp, double* partial)
{
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    double e = 0;
    if (i < C.n[KIND]) e = collision_energy<KIND>(C, X, nV, bp, i);
    block_partial(e, partial);
}

// Batched contexts: the blocks of every (scene, kind) range, scene-major
// (boff: first block of each of the 4 n_scenes ranges, then the total), so
// each scene's partials are the ones a context holding the scene alone
// computes, in the same order.
__global__ __launch_bounds__(PB) void k_energy_scenes(CollView C, const double* X, int nV,
                                                      BarrierParams bp, const int64_t* coff,
                                                      const int64_t* boff, int S, double* partial)
{
    __shared__ int64_t first, last;
    __shared__ int kind;
    const int64_t b = blockIdx.x;
    if (threadIdx.x == 0) {
        int lo = 0, hi = 4 * S; // the last range whose first block is <= b
        while (hi - lo > 1) {
            const int m = (lo + hi) >> 1;
            if (boff[m] <= b) lo = m;
            else hi = m;
        }
        const int sc = lo / 4, k = lo % 4;
        kind = k;
        first = coff[k * (int64_t)(S + 1) + sc] + (b - boff[lo]) * PB;
        last = coff[k * (int64_t)(S + 1) + sc + 1];
    }
    __syncthreads();
    const int64_t i = first + threadIdx.x;
    double e = 0;
    if (i < last) {
        switch (kind) {
        case K_VV: e = collision_energy<K_VV>(C, X, nV, bp, i); break;
        case K_EV: e = collision_energy<K_EV>(C, X, nV, bp, i); break;
        case K_EE: e = collision_energy<K_EE>(C, X, nV, bp, i); break;
        default: e = collision_energy<K_FV>(C, X, nV, bp, i); break;
        }
    }
    block_partial(e, partial);
}

// first collision of every scene per kind (kind-major, n_scenes + 1 each)
// by binary search of the canonically sorted lists (soff: first vertex,
// edge and face of every scene, n_scenes + 1 each)
__global__ void k_scene_coff(CollView C, const int64_t* soff, int S, int64_t* coff)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 4 * (S + 1)) return;
    const int k = t / (S + 1), sc = t % (S + 1);
    const int64_t target = soff[(k == K_VV ? 0 : (k == K_FV ? 2 : 1)) * (int64_t)(S + 1) + sc];
    const int2* ids = C.ids[k];
    int64_t lo = 0, hi = C.n[k];
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        const int2 p = ids[m];
        const int64_t key = k == K_EE ? (p.x < p.y ? p.x : p.y) : p.x;
        if (key < target) lo = m + 1;
        else hi = m;
    }
    coff[t] = lo;
}

__device__ __forceinline__ void sum_partials(const double* partial, int64_t n, double* out)
{
    // a thread adds partial[tid], partial[tid + 256], ... in that order;
    // eight loads are issued before their adds (one memory wait per eight)
    double t = 0;
    int64_t i = threadIdx.x;
    for (; i + 7 * 256 < n; i += 8 * 256) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = partial[i + 256 * k];
#pragma unroll
        for (int k = 0; k < 8; k++) t += v[k];
    }
    for (; i < n; i += 256) t += partial[i];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, WAVE);
    __shared__ double s[4];
    if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) *out = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ __launch_bounds__(256) void k_sum(const double* partial, int64_t n, double* out)
{
    sum_partials(partial, n, out);
}

// one workgroup per scene over its partials (boff as for k_energy_scenes)
__global__ __launch_bounds__(256) void k_sum_scenes(const double* partial, const int64_t* boff,
                                                    double* out)
{
    const int sc = blockIdx.x;
    sum_partials(partial + boff[4 * (int64_t)sc], boff[4 * (int64_t)sc + 4] - boff[4 * (int64_t)sc],
                 out + sc);
}

// Atomic-free gradient: every collision writes its 3 NP derivatives to a
// record (point-major), then one wave per vertex sums the records of the
// vertex's incidences in a fixed order.
template <int KIND>
__global__ __launch_bounds__(PB) void k_gradient_records(CollView C, const double* X, int nV,
                                                         BarrierParams bp, double* rec)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    d3 g[NP];
    stencil_gradient<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                           KIND == K_EE ? C.eps[i] : 0.0, bp, g);
    double* o = rec + (int64_t)3 * NP * i;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        o[3 * j] = g[j].x;
        o[3 * j + 1] = g[j].y;
        o[3 * j + 2] = g[j].z;
    }
}

// ---- Gauss-Newton helpers (normal_potential.cpp:236-308) -----------------
// The distance vector t = sum_i c_i x_i of a collision, with c the
// closest-point coefficients of its known distance type
// (candidates/{vertex_vertex,edge_vertex,edge_edge,face_vertex}.cpp
// compute_coefficients).  Runtime-size vectors (VectorMax3d) reduce left to
// right; fixed-size dots use dot().
__device__ __forceinline__ double dot_seq(const d3& a, const d3& b)
{
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}

// closest_point.cpp:11-18
__device__ __forceinline__ double pe_closest(const d3& p, const d3& e0, const d3& e1)
{
    const d3 e = sub(e1, e0);
    return dot_seq(sub(p, e0), e) / dot_seq(e, e);
}

template <int KIND>
__device__ __forceinline__ void stencil_coeffs(int dtype, const d3* P, double* c)
{
    if constexpr (KIND == K_VV) {
        c[0] = 1.0;
        c[1] = -1.0;
    } else if constexpr (KIND == K_EV) { // P_E
        const double a = pe_closest(P[0], P[1], P[2]);
        c[0] = 1.0;
        c[1] = -1.0 + a;
        c[2] = -a;
    } else if constexpr (KIND == K_FV) { // P_T: closest_point.cpp:123-137
        const d3 b0 = sub(P[2], P[1]), b1 = sub(P[3], P[1]), r = sub(P[0], P[1]);
        double u, v;
        solve_sym2_ldlt(dot(b0, b0), dot(b1, b0), dot(b1, b1), dot(b0, r), dot(b1, r), u, v);
        c[0] = 1.0;
        c[1] = -1.0 + u + v;
        c[2] = -u;
        c[3] = -v;
    } else {
        double a;
        c[0] = c[1] = c[2] = c[3] = 0.0;
        switch (dtype) {
        case EE_EA0_EB0: c[0] = 1; c[2] = -1; break;
        case EE_EA0_EB1: c[0] = 1; c[3] = -1; break;
        case EE_EA1_EB0: c[1] = 1; c[2] = -1; break;
        case EE_EA1_EB1: c[1] = 1; c[3] = -1; break;
        case EE_EA_EB0:
            a = pe_closest(P[2], P[0], P[1]);
            c[0] = 1.0 - a; c[1] = a; c[2] = -1.0;
            break;
        case EE_EA_EB1:
            a = pe_closest(P[3], P[0], P[1]);
            c[0] = 1.0 - a; c[1] = a; c[3] = -1.0;
            break;
        case EE_EA0_EB:
            a = pe_closest(P[0], P[2], P[3]);
            c[0] = 1.0; c[2] = -1.0 + a; c[3] = -a;
            break;
        case EE_EA1_EB:
            a = pe_closest(P[1], P[2], P[3]);
            c[1] = 1.0; c[2] = -1.0 + a; c[3] = -a;
            break;
        default: { // closest_point.cpp:65-87
            const d3 d = sub(P[0], P[2]), ea = sub(P[1], P[0]), eb = sub(P[3], P[2]);
            double x0, x1;
            solve_sym2_ldlt(sqn(ea), -dot(eb, ea), sqn(eb), -dot(d, ea), dot(d, eb), x0, x1);
            c[0] = 1.0 - x0; c[1] = x0; c[2] = -1.0 + x1; c[3] = -x1;
        }
        }
    }
}

template <int NP>
__device__ __forceinline__ d3 combine(const double* c, const d3* P)
{
    d3 t { 0.0, 0.0, 0.0 };
#pragma unroll
    for (int i = 0; i < NP; i++) t = add(t, scale(P[i], c[i]));
    return t;
}

// normal_potential.cpp:236-270, records in the gradient's layout
template <int KIND>
__global__ __launch_bounds__(PB) void k_gn_diag_records(CollView C, const double* X, int nV,
                                                        BarrierParams bp, double* rec)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    double c[4];
    stencil_coeffs<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, c);
    const d3 t = combine<NP>(c, P);
    const double d2 = dot_seq(t, t);
    const double w = C.w[KIND][i];
    const double s1 = w * 4.0 * bar_f2(d2, bp), s2 = w * 2.0 * bar_f1(d2, bp);
    const double t2[3] = { t.x * t.x, t.y * t.y, t.z * t.z };
    double* o = rec + (int64_t)3 * NP * i;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        const double c2 = c[j] * c[j];
#pragma unroll
        for (int k = 0; k < 3; k++) o[3 * j + k] = s1 * (c2 * t2[k]) + s2 * c2;
    }
}

// normal_potential.cpp:272-308; p: 3 nV RowMajor
template <int KIND>
__global__ __launch_bounds__(PB) void k_gn_quadratic_form(CollView C, const double* X, int nV,
                                                          BarrierParams bp, const double* p,
                                                          double* partial)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    double e = 0;
    if (i < C.n[KIND]) {
        int v[4];
        stencil_ids<KIND>(C, i, v);
        d3 P[4], Q[4];
#pragma unroll
        for (int j = 0; j < NP; j++) {
            P[j] = ldx(X, nV, v[j]);
            Q[j] = { p[3 * (int64_t)v[j]], p[3 * (int64_t)v[j] + 1], p[3 * (int64_t)v[j] + 2] };
        }
        double c[4];
        stencil_coeffs<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, c);
        const d3 t = combine<NP>(c, P), q = combine<NP>(c, Q);
        const double d2 = dot_seq(t, t);
        const double qt = dot_seq(q, t);
        e = C.w[KIND][i] * (4.0 * bar_f2(d2, bp) * (qt * qt) + 2.0 * bar_f1(d2, bp) * dot_seq(q, q));
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, WAVE);
    __shared__ double s[PB / WAVE];
    if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int j = 0; j < PB / WAVE; j++) t += s[j];
        partial[blockIdx.x] = t;
    }
}

template <int KIND> struct HqTraits {
    static constexpr int NP = StencilTraits<KIND>::NP;
    static constexpr int N = 3 * (NP - 1);
    static constexpr int LEN = N * (N + 1) / 2;
};

// ---- symbolic phase: block pattern of the collision set ---------------------
// Off-diagonal vertex pairs (min, max) of every stencil.
template <int KIND>
__global__ __launch_bounds__(256) void k_pair_keys(CollView C, uint64_t* keys, uint32_t* vals,
                                                   int64_t key_off, uint32_t* mark)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    constexpr int NPAIR = NP * (NP - 1) / 2;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    int64_t o = key_off + (int64_t)NPAIR * i;
#pragma unroll
    for (int a = 0; a < NP; a++) {
        mark[v[a]] = 1;
#pragma unroll
        for (int b = a + 1; b < NP; b++) {
            const uint32_t lo = (uint32_t)(v[a] < v[b] ? v[a] : v[b]);
            const uint32_t hi = (uint32_t)(v[a] < v[b] ? v[b] : v[a]);
            keys[o] = ((uint64_t)lo << 32) | hi;
            vals[o] = 0;
            o++;
        }
    }
}

__global__ void k_pair_degree(const uint64_t* upairs, const uint32_t* nu, uint32_t* deg)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)*nu) return;
    const uint64_t k = upairs[i];
    atomicAdd(deg + (k >> 32), 1u);
    atomicAdd(deg + (k & 0xFFFFFFFFu), 1u);
}

__global__ void k_add_mark(const uint32_t* mark, int nV, uint32_t* deg)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < nV) deg[v] += mark[v];
}

__global__ void k_fill_diag(const uint32_t* mark, int nV, const uint32_t* bptr, uint32_t* cursor,
                            int32_t* bcol)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v < nV && mark[v]) bcol[bptr[v] + atomicAdd(cursor + v, 1u)] = v;
}

__global__ void k_fill_pairs(const uint64_t* upairs, const uint32_t* nu, const uint32_t* bptr,
                             uint32_t* cursor, int32_t* bcol)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (int64_t)*nu) return;
    const uint64_t k = upairs[i];
    const uint32_t a = (uint32_t)(k >> 32), b = (uint32_t)(k & 0xFFFFFFFFu);
    bcol[bptr[a] + atomicAdd(cursor + a, 1u)] = (int32_t)b;
    bcol[bptr[b] + atomicAdd(cursor + b, 1u)] = (int32_t)a;
}

// sort each block row's columns (rows are short)
__global__ void k_sort_rows(const uint32_t* bptr, int nV, int32_t* bcol)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nV) return;
    const uint32_t s = bptr[v], e = bptr[v + 1];
    for (uint32_t i = s + 1; i < e; i++) {
        const int32_t x = bcol[i];
        uint32_t j = i;
        while (j > s && bcol[j - 1] > x) {
            bcol[j] = bcol[j - 1];
            j--;
        }
        bcol[j] = x;
    }
}

// ---- symbolic phase from the vertex incidences -------------------------------
// The block columns of vertex v's row are the vertices of every stencil that
// contains v.  One wave per vertex walks its incidences, inserts the stencil
// vertices into a hash set in LDS, sorts the distinct ones and writes them
// to a staging area with room for four per incidence; a second pass packs
// them behind the scanned row lengths.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct CollRanges {
    int64_t goff[5]; // global collision index ranges per kind
};

constexpr int PH_BITS = 9;
constexpr int PH = 1 << PH_BITS; // hash slots per wave (LDS: 8 waves per SIMD fit)
constexpr int PU = 256;          // most distinct columns of a row on this path

__global__ __launch_bounds__(256) void k_pattern_rows(const uint32_t* vptr, const int4* inc_vtx,
                                                      int nV, int32_t* stage, uint32_t* cnt,
                                                      int* overflow)
{
    __shared__ int32_t hs[256 / WAVE][PH];
    __shared__ int32_t uq[256 / WAVE][PU];
    __shared__ unsigned nu[256 / WAVE];
    const int wv = threadIdx.x / WAVE, l = lane_id();
    const int64_t v = ((int64_t)xcd_block() * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (v >= nV) return;
    const uint32_t i0 = vptr[v], i1 = vptr[v + 1];
    if (i0 == i1) {
        if (l == 0) cnt[v] = 0;
        return;
    }
    for (int j = l; j < PH; j += WAVE) hs[wv][j] = -1;
    if (l == 0) nu[wv] = 0;
    wave_sync();
    for (uint32_t j = i0 + l; j < i1; j += WAVE) {
        const int4 vs = inc_vtx[j];
#pragma unroll
        for (int b = 0; b < 4; b++) {
            const int key = b == 0 ? vs.x : b == 1 ? vs.y : b == 2 ? vs.z : vs.w;
            if (key < 0) continue;
            unsigned h = ((unsigned)key * 2654435761u) >> (32 - PH_BITS); // in [0, PH)
            for (int probe = 0; probe < PH; probe++) {
                const int old = atomicCAS(&hs[wv][h], -1, key);
                if (old == -1) {
                    const unsigned k = atomicAdd(&nu[wv], 1u);
                    if (k < PU) uq[wv][k] = key;
                    break;
                }
                if (old == key) break;
                h = (h + 1) & (PH - 1);
            }
        }
    }
    wave_sync();
    const unsigned n = nu[wv];
    if (n > PU) {
        if (l == 0) {
            cnt[v] = 0;
            atomicOr(overflow, 1);
        }
        return;
    }
    // bitonic sort of the n distinct columns (padded to a power of two)
    unsigned P = 1;
    while (P < n) P <<= 1;
    for (unsigned t = n + l; t < P; t += WAVE) uq[wv][t] = INT32_MAX;
    wave_sync();
    for (unsigned k = 2; k <= P; k <<= 1)
        for (unsigned jj = k >> 1; jj > 0; jj >>= 1) {
            for (unsigned t = l; t < P; t += WAVE) {
                const unsigned q = t ^ jj;
                if (q > t) {
                    const int a = uq[wv][t], b = uq[wv][q];
                    if ((a > b) == ((t & k) == 0)) {
                        uq[wv][t] = b;
                        uq[wv][q] = a;
                    }
                }
            }
            wave_sync();
        }
    int32_t* out = stage + 4 * (int64_t)i0;
    for (unsigned t = l; t < n; t += WAVE) out[t] = uq[wv][t];
    if (l == 0) cnt[v] = n;
}

__global__ __launch_bounds__(256) void k_pattern_pack(const uint32_t* vptr, const int32_t* stage,
                                                      const uint32_t* bptr, int nV, int32_t* bcol)
{
    const int64_t v = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (v >= nV) return;
    const uint32_t b0 = bptr[v], n = bptr[v + 1] - b0;
    const int32_t* in = stage + 4 * (int64_t)vptr[v];
    for (uint32_t t = lane_id(); t < n; t += WAVE) bcol[b0 + t] = in[t];
}

struct GradRecs {
    const double* rec[4];
};

// order: vertex visited by each wave (null: vertex order)
__global__ __launch_bounds__(256) void k_gradient_rows(CollRanges G, GradRecs R,
                                                       const uint32_t* vptr, const uint32_t* inc,
                                                       int nV, double* grad,
                                                       const uint32_t* order)
{
    const int64_t w = ((int64_t)xcd_block() * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (w >= nV) return;
    const int64_t v = order ? (int64_t)order[w] : w;
    const int l = lane_id();
    double gx = 0, gy = 0, gz = 0;
    for (uint32_t j = vptr[v] + l; j < vptr[v + 1]; j += WAVE) {
        const uint32_t x = inc[j];
        const int64_t g = x >> 2;
        const int a = (int)(x & 3u);
        const double* r;
        if (g < G.goff[1]) r = R.rec[0] + 6 * g;
        else if (g < G.goff[2]) r = R.rec[1] + 9 * (g - G.goff[1]);
        else if (g < G.goff[3]) r = R.rec[2] + 12 * (g - G.goff[2]);
        else r = R.rec[3] + 12 * (g - G.goff[3]);
        r += 3 * a;
        gx += r[0];
        gy += r[1];
        gz += r[2];
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        gx += __shfl_xor(gx, o, WAVE);
        gy += __shfl_xor(gy, o, WAVE);
        gz += __shfl_xor(gz, o, WAVE);
    }
    if (l == 0) {
        grad[3 * v] = gx;
        grad[3 * v + 1] = gy;
        grad[3 * v + 2] = gz;
    }
}

// ---- numeric phase without global atomics ------------------------------------
// (1) every collision writes its PSD-projected Hessian in the Q basis
//     (packed 3K x 3K) to a record;
// (2) the (vertex, collision, point) incidences are sorted by vertex;
// (3) one wave per vertex owns its block row: the lanes walk the vertex's
//     incidences, expand each record to the row's 3x3 blocks and add them
//     into the row's accumulators in LDS, then the wave writes the row.
// Every record holds the projected Hessian in the 4-point Q basis (packed
// 9 x 9, REC_LEN values), so one code path expands all kinds.  A stencil of
// NP < 4 points is projected in its own NP-point basis and mapped exactly:
// with T[k][m] = sum_{x<NP} q4(x,k) qNP(x,m), the 4-point matrix
// T B T^T expands to the same point blocks (Q4 Q4^T = I - 11^T/4 and the
// columns of Q_NP sum to zero) and to zero blocks for the unused points.
constexpr int REC_LEN = 45;

// PATH (4-point stencils): REC_ALL projects every collision with the 9 x 9
// Jacobi; REC_LOW projects by the rank-5 Krylov path (falling back to the
// Jacobi for a matrix it does not capture) and, for edge-edge, takes only
// the interior (EA_EB) type, whose Hessians have rank 5; REC_LIST takes the
// collisions listed in `only` (the other edge-edge types) with the Jacobi.
enum { REC_ALL = 0, REC_LOW = 1, REC_LIST = 2 };

template <int KIND, int PATH = REC_ALL>
__global__ __launch_bounds__(64) void k_hessian_records(CollView C, const double* X, int nV,
                                                        BarrierParams bp, double* rec,
                                                        const uint32_t* only = nullptr,
                                                        int64_t n_only = 0)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    constexpr int K = NP - 1;
    constexpr int LEN = HqTraits<KIND>::LEN;
    const int64_t t = (int64_t)xcd_block() * 64 + threadIdx.x;
    int64_t i = t;
    if constexpr (PATH == REC_LIST) {
        if (t >= n_only) return;
        i = only[t];
    } else {
        if (i >= C.n[KIND]) return;
    }
    if constexpr (PATH == REC_LOW && KIND == K_EE)
        if (C.dtype[i] != EE_EA_EB) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    double H[LEN];
    if constexpr (PATH == REC_LOW && NP == 4) {
        const bool psd = stencil_hessian_q<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                                                 KIND == K_EE ? C.eps[i] : 0.0, bp, H, true);
        if (psd && !project_psd_krylov9<6>(H, bp.psd)) project_psd<9>(H, bp.psd);
    } else {
        stencil_hessian_q<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                                KIND == K_EE ? C.eps[i] : 0.0, bp, H);
    }
    double* o = rec + (int64_t)REC_LEN * i;
    if constexpr (NP == 4) {
#pragma unroll
        for (int e = 0; e < LEN; e++) o[e] = H[e];
    } else {
        double T[3][K];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int m = 0; m < K; m++) {
                double t = 0;
#pragma unroll
                for (int x = 0; x < NP; x++) t += QBasis<4>::q(x, k) * QBasis<NP>::q(x, m);
                T[k][m] = t;
            }
#pragma unroll
        for (int p = 0; p < 9; p++)
#pragma unroll
            for (int q = 0; q <= p; q++) {
                const int k = p / 3, r = p % 3, l = q / 3, t = q % 3;
                double acc = 0;
#pragma unroll
                for (int m = 0; m < K; m++)
#pragma unroll
                    for (int n = 0; n < K; n++)
                        acc += T[k][m] * T[l][n] * H[sp(3 * m + r, 3 * n + t)];
                o[sp(p, q)] = acc;
            }
    }
}

// (vertex, global collision g, point a): key vertex, value g << 2 | a
template <int KIND>
__global__ __launch_bounds__(256) void k_incidence(CollView C, int64_t goff, int64_t ioff,
                                                   uint32_t* keys, uint32_t* vals)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
#pragma unroll
    for (int a = 0; a < NP; a++) {
        keys[ioff + (int64_t)NP * i + a] = (uint32_t)v[a];
        vals[ioff + (int64_t)NP * i + a] = (uint32_t)(((goff + i) << 2) | a);
    }
}

// First incidence of every vertex in the sorted keys (vptr[nV] = M): lane i
// writes the vertices (keys[i-1], keys[i]] when they are few; the vertices
// of a longer gap (no incidence: a masked scene of a batch, a free part of
// the mesh) keep the fill value and k_gap_ptr finds them by binary search.
constexpr int RUN_PTR_GAP = 32;
constexpr uint32_t RUN_PTR_NONE = 0xFFFFFFFFu;

__global__ __launch_bounds__(256) void k_run_ptr(const uint32_t* keys, int64_t M, int nV,
                                                 uint32_t* vptr)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i > M) return;
    const int64_t prev = i > 0 ? (int64_t)keys[i - 1] : -1;
    const int64_t cur = i < M ? (int64_t)keys[i] : (int64_t)nV;
    const int64_t from = cur - prev <= RUN_PTR_GAP ? prev + 1 : cur;
    for (int64_t v = from; v <= cur; v++) vptr[v] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_gap_ptr(const uint32_t* keys, int64_t M, int nV,
                                                 uint32_t* vptr)
{
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= nV || vptr[v] != RUN_PTR_NONE) return;
    int64_t lo = 0, hi = M; // first i with keys[i] >= v
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    vptr[v] = (uint32_t)lo;
}

// edge-edge collisions of the other (non-interior) types, whose Hessians
// have higher rank: flags, then the list at the scanned positions
__global__ __launch_bounds__(256) void k_ee_other_flags(const int32_t* dtype, int64_t n,
                                                        uint32_t* flag)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) flag[i] = dtype[i] != EE_EA_EB ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_ee_other_list(const uint32_t* flag, const uint32_t* pos,
                                                       int64_t n, uint32_t* list)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && flag[i]) list[pos[i]] = (uint32_t)i;
}

// stencil vertices of every sorted incidence (-1 for unused points)
__global__ __launch_bounds__(256) void k_inc_vertices(CollView C, CollRanges G,
                                                      const uint32_t* inc, int64_t M, int4* out)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= M) return;
    const int64_t g = inc[j] >> 2;
    int v[4] = { -1, -1, -1, -1 };
    if (g < G.goff[1]) stencil_ids<K_VV>(C, g, v);
    else if (g < G.goff[2]) stencil_ids<K_EV>(C, g - G.goff[1], v);
    else if (g < G.goff[3]) stencil_ids<K_EE>(C, g - G.goff[2], v);
    else stencil_ids<K_FV>(C, g - G.goff[3], v);
    out[j] = make_int4(v[0], v[1], v[2], v[3]);
}

constexpr int RW = 64; // block columns per accumulation window (one per lane)

// Adds the contributions of one collision (record h in the 4-point basis,
// stencil vertices vs, the row vertex is point a) to the window of the row's
// blocks whose columns are cols[0, RW) (sorted, padded with INT32_MAX).  The
// record stays in memory (read through the cache) and is expanded one block
// row r at a time, so few values are live at once.
__device__ __forceinline__ void row_accumulate(const double* h, int a, int4 vs,
                                               const int32_t* cols, double (*acc)[9],
                                               uint32_t* msk)
{
    constexpr int K = 3;
    // window slot of each stencil point's column (-1: unused point or outside
    // this window)
    int slot[4];
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const int col = b == 0 ? vs.x : b == 1 ? vs.y : b == 2 ? vs.z : vs.w;
        slot[b] = -1;
        if (col < 0 || col < cols[0] || col > cols[RW - 1]) continue;
        int lo = 0, hi = RW;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cols[mid] < col) lo = mid + 1;
            else hi = mid;
        }
        if (lo < RW && cols[lo] == col) slot[b] = lo;
    }
    uint32_t mask[4] = { 0, 0, 0, 0 };
#pragma unroll
    for (int r = 0; r < 3; r++) {
        // G[l][t] = sum_k q(a,k) H[3k+r][3l+t]
        double G[K][3];
#pragma unroll
        for (int l = 0; l < K; l++)
#pragma unroll
            for (int t = 0; t < 3; t++) {
                double x = 0;
#pragma unroll
                for (int k = 0; k < K; k++) x += QBasis<4>::q(a, k) * h[sp(3 * k + r, 3 * l + t)];
                G[l][t] = x;
            }
#pragma unroll
        for (int b = 0; b < 4; b++) {
            if (slot[b] < 0) continue;
#pragma unroll
            for (int t = 0; t < 3; t++) {
                double x = 0;
#pragma unroll
                for (int l = 0; l < K; l++) x += QBasis<4>::q(b, l) * G[l][t];
                if (x != 0) { // exact zeros are not triplets (local_to_global.hpp:291)
                    atomicAdd(&acc[slot[b]][3 * r + t], x);
                    mask[b] |= 1u << (3 * r + t);
                }
            }
        }
    }
#pragma unroll
    for (int b = 0; b < 4; b++)
        if (mask[b]) atomicOr(&msk[slot[b]], mask[b]);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_hessian_rows(const double* rec, const uint32_t* vptr,
                                                      const uint32_t* inc, const int4* inc_vtx,
                                                      const uint32_t* bptr,
                                                      const int32_t* bcol, int nV, double* bval,
                                                      uint32_t* bmask, const uint32_t* order)
{
    __shared__ double acc[256 / WAVE][RW][9];
    __shared__ uint32_t msk[256 / WAVE][RW];
    __shared__ int32_t cols[256 / WAVE][RW];
    const int wv = threadIdx.x / WAVE, l = lane_id();
    const int64_t w = ((int64_t)xcd_block() * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (w >= nV) return;
    const int64_t v = order ? (int64_t)order[w] : w; // vertex of this wave
    const uint32_t row_s = bptr[v], m = bptr[v + 1] - row_s;
    const uint32_t i0 = vptr[v], i1 = vptr[v + 1];
    for (uint32_t w0 = 0; w0 < m; w0 += RW) {
#pragma unroll
        for (int e = 0; e < 9; e++) acc[wv][l][e] = 0;
        msk[wv][l] = 0;
        cols[wv][l] = w0 + l < m ? bcol[row_s + w0 + l] : INT32_MAX;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t j = i0 + l; j < i1; j += WAVE) {
            const uint32_t x = inc[j];
            row_accumulate(rec + (int64_t)REC_LEN * (x >> 2), (int)(x & 3u), inc_vtx[j], cols[wv],
                           acc[wv], msk[wv]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (w0 + l < m) {
            const int64_t u = (int64_t)row_s + w0 + l;
#pragma unroll
            for (int e = 0; e < 9; e++) bval[9 * u + e] = acc[wv][l][e];
            bmask[u] = msk[wv][l];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// One wave per vertex for the two CSR passes: the lanes walk the vertex's
// block row together, so the reads of the pattern and the writes of its
// three matrix rows are contiguous across the wave.
__device__ __forceinline__ unsigned wave_exclusive_sum(unsigned x, unsigned& total)
{
    const int l = lane_id();
    unsigned inc = x;
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        const unsigned y = __shfl_up(inc, o, WAVE);
        if (l >= o) inc += y;
    }
    total = __shfl(inc, WAVE - 1, WAVE);
    return inc - x;
}

__global__ __launch_bounds__(256) void k_row_counts(const uint32_t* bptr, const uint32_t* bmask,
                                                    int nV, uint64_t* cnt)
{
    const int64_t v = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (v >= nV) return;
    const int l = lane_id();
    unsigned c[3] = { 0, 0, 0 };
    for (uint32_t u = bptr[v] + l; u < bptr[v + 1]; u += WAVE) {
        const uint32_t m = bmask[u];
        c[0] += __popc(m & 7u);
        c[1] += __popc((m >> 3) & 7u);
        c[2] += __popc((m >> 6) & 7u);
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int o = WAVE / 2; o > 0; o >>= 1) c[k] += __shfl_xor(c[k], o, WAVE);
    if (l < 3) cnt[3 * v + l] = c[l == 0 ? 0 : (l == 1 ? 1 : 2)];
}

__global__ __launch_bounds__(256) void k_fill_rows(const uint32_t* bptr, const int32_t* bcol,
                                                   const uint32_t* bmask, const double* bval,
                                                   int nV, const int64_t* rowptr, int32_t* col,
                                                   double* val)
{
    const int64_t v = ((int64_t)blockIdx.x * 256 + threadIdx.x) / WAVE; // uniform per wave
    if (v >= nV) return;
    const int l = lane_id();
    const uint32_t s = bptr[v], e = bptr[v + 1];
    const int m = (int)(e - s);
    if (m > 0 && m <= WAVE) {
        // R of the three rows per pass, m lanes each (lane l: row k0 + l / m,
        // block l % m)
        const int R = WAVE / m < 3 ? WAVE / m : 3;
        const int kk = l / m, b = l - kk * m;
        for (int k0 = 0; k0 < 3; k0 += R) {
            const int k = k0 + kk;
            const bool on = kk < R && k < 3;
            const uint32_t u = s + (uint32_t)b;
            const uint32_t msk = on ? (bmask[u] >> (3 * k)) & 7u : 0u;
            unsigned total;
            const unsigned excl = wave_exclusive_sum(__popc(msk), total);
            const unsigned first = __shfl(excl, on ? kk * m : 0, WAVE); // the row's first lane
            if (msk) {
                const int cb = bcol[u];
                int64_t at = rowptr[3 * v + k] + (excl - first);
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (msk & (1u << j)) {
                        col[at] = 3 * cb + j;
                        val[at] = bval[9 * (int64_t)u + 3 * k + j];
                        at++;
                    }
            }
        }
        return;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int64_t pos = rowptr[3 * v + k];
        for (uint32_t base = s; base < e; base += WAVE) {
            const uint32_t u = base + l;
            const uint32_t m = u < e ? (bmask[u] >> (3 * k)) & 7u : 0u;
            unsigned total;
            const unsigned off = wave_exclusive_sum(__popc(m), total);
            if (m) {
                const int cb = bcol[u];
                int64_t at = pos + off;
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (m & (1u << j)) {
                        col[at] = 3 * cb + j;
                        val[at] = bval[9 * (int64_t)u + 3 * k + j];
                        at++;
                    }
            }
            pos += total;
        }
    }
}

CollView view(Ctx& c)
{
    CollView C;
    for (int k = 0; k < 4; k++) {
        C.ids[k] = c.c_ids[k].as<int2>();
        C.w[k] = c.c_w[k].as<double>();
        C.n[k] = c.ncoll[k];
    }
    C.eps = c.c_eps.as<double>();
    C.dtype = c.c_dtype.as<int32_t>();
    C.E = c.E.as<int32_t>();
    C.F = c.F.as<int32_t>();
    return C;
}

BarrierParams barrier_params(const Ctx& c, double dhat, double kappa, int psd)
{
    BarrierParams bp { dhat, kappa, c.coll_dmin, psd };
    if (c.physical_barrier) {
        const double t = (2 * c.coll_dmin + dhat) * dhat;
        bp.phys = dhat / (t * t);
    }
    return bp;
}

void require_collisions(Ctx& c)
{
    c.require_mesh();
    c.require_pos(0);
    if (!c.coll_valid) throw Error(IPCG_ERR_STATE, "no collision set (build collisions first)");
}

} // namespace

double pot_energy(Ctx& c, double dhat, double kappa)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    int64_t nb[4], tot = 0;
    for (int k = 0; k < 4; k++) {
        nb[k] = (C.n[k] + PB - 1) / PB;
        tot += nb[k];
    }
    c.red_partial.reserve(sizeof(double) * (size_t)std::max<int64_t>(tot, 1));
    c.red_out.reserve(sizeof(double) * 2);
    double* part = c.red_partial.as<double>();
    const double* X = c.X[0].as<double>();
    c.stage_begin("energy");
    int64_t o = 0;
    if (nb[0]) hipLaunchKernelGGL(k_energy<K_VV>, dim3(nb[0]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[0];
    if (nb[1]) hipLaunchKernelGGL(k_energy<K_EV>, dim3(nb[1]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[1];
    if (nb[2]) hipLaunchKernelGGL(k_energy<K_EE>, dim3(nb[2]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[2];
    if (nb[3]) hipLaunchKernelGGL(k_energy<K_FV>, dim3(nb[3]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, part, tot, c.red_out.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    double* h = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, c.red_out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    return h[0];
}

// ---- allocation-free energy -------------------------------------------------
// The energy's block partial sums in a file-scope device array instead of a
// context buffer, for callers that cannot allocate device memory (code under
// graph capture, serving loops).  Every GPU the library runs on holds its own
// copy until the process ends; its device address is looked up once per
// device (noalloc_init, from context creation with the device current).
// Calls on one GPU share the array: calls that overlap on one GPU race.
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[MAX_DEVICES] = {};
std::string g_scratch_err[MAX_DEVICES];
} // namespace

void noalloc_init(int device)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES || g_scratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return;
    }
    g_scratch[device] = static_cast<double*>(p);
}

void pot_energy_noalloc(Ctx& c, double dhat, double kappa, double* d_out)
{
    require_collisions(c);
    if (!d_out) throw Error(IPCG_ERR_ARG, "null output pointer");
    double* part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c.device >= 0 && c.device < MAX_DEVICES) part = g_scratch[c.device];
        if (!part)
            throw Error(IPCG_ERR_STATE,
                        "no scratch array on this device"
                            + (c.device >= 0 && c.device < MAX_DEVICES
                                   ? ": " + g_scratch_err[c.device] : std::string()));
    }
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    int64_t nb[4], tot = 0;
    for (int k = 0; k < 4; k++) {
        nb[k] = (C.n[k] + PB - 1) / PB;
        tot += nb[k];
    }
    if (tot > NOALLOC_PARTIALS)
        throw Error(IPCG_ERR_ARG, "collision set too large for the allocation-free energy: "
                                      + std::to_string(tot) + " blocks of "
                                      + std::to_string(PB) + " collisions, at most "
                                      + std::to_string(NOALLOC_PARTIALS));
    const double* X = c.X[0].as<double>();
    int64_t o = 0;
    if (nb[0]) hipLaunchKernelGGL(k_energy<K_VV>, dim3(nb[0]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[0];
    if (nb[1]) hipLaunchKernelGGL(k_energy<K_EV>, dim3(nb[1]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[1];
    if (nb[2]) hipLaunchKernelGGL(k_energy<K_EE>, dim3(nb[2]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    o += nb[2];
    if (nb[3]) hipLaunchKernelGGL(k_energy<K_FV>, dim3(nb[3]), dim3(PB), 0, s, C, X, c.nV, bp, part + o);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, part, tot, d_out);
    IPCG_CHECK_LAUNCH();
}

// ---- batched contexts ------------------------------------------------------
void scene_collision_offsets(Ctx& c)
{
    if (!c.batched()) throw Error(IPCG_ERR_STATE, "not a batched context");
    require_collisions(c);
    if (c.coff_valid) return;
    const int S = c.n_scenes;
    hipStream_t s = c.stream;
    c.coff_d.reserve(sizeof(int64_t) * 4 * (size_t)(S + 1));
    hipLaunchKernelGGL(k_scene_coff, dim3(grid_for(4 * (S + 1), 256)), dim3(256), 0, s, view(c),
                       c.voff_d.as<int64_t>(), S, c.coff_d.as<int64_t>());
    IPCG_CHECK_LAUNCH();
    c.h_coff.assign(4 * (size_t)(S + 1), 0);
    IPCG_HIP(hipMemcpyAsync(c.h_coff.data(), c.coff_d.p, sizeof(int64_t) * c.h_coff.size(),
                            hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    c.coff_valid = true;
}

void pot_energy_scenes(Ctx& c, double dhat, double kappa, double* out)
{
    scene_collision_offsets(c);
    const int S = c.n_scenes;
    hipStream_t s = c.stream;
    // blocks of PB collisions per (scene, kind), scene-major
    std::vector<int64_t> boff(4 * (size_t)S + 1, 0);
    for (int sc = 0; sc < S; sc++)
        for (int k = 0; k < 4; k++) {
            const int64_t n = c.h_coff[k * (size_t)(S + 1) + sc + 1] - c.h_coff[k * (size_t)(S + 1) + sc];
            boff[4 * (size_t)sc + k + 1] = boff[4 * (size_t)sc + k] + (n + PB - 1) / PB;
        }
    const int64_t tot = boff.back();
    c.red_partial.reserve(sizeof(double) * (size_t)std::max<int64_t>(tot, 1));
    c.scene_part.reserve(sizeof(int64_t) * boff.size() + sizeof(double) * (size_t)S);
    int64_t* boff_d = c.scene_part.as<int64_t>();
    double* out_d = reinterpret_cast<double*>(boff_d + boff.size());
    IPCG_HIP(hipMemcpyAsync(boff_d, boff.data(), sizeof(int64_t) * boff.size(),
                            hipMemcpyHostToDevice, s));
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    c.stage_begin("energy");
    if (tot > 0) {
        if (tot > INT32_MAX) throw Error(IPCG_ERR_ARG, "too many collisions for one launch");
        hipLaunchKernelGGL(k_energy_scenes, dim3((unsigned)tot), dim3(PB), 0, s, view(c),
                           c.X[0].as<double>(), c.nV, bp, c.coff_d.as<int64_t>(), boff_d, S,
                           c.red_partial.as<double>());
    }
    hipLaunchKernelGGL(k_sum_scenes, dim3(S), dim3(256), 0, s, c.red_partial.as<double>(), boff_d,
                       out_d);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    IPCG_HIP(hipMemcpyAsync(out, out_d, sizeof(double) * (size_t)S, hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
}

// Global collision ranges per kind and the number of (vertex, collision)
// incidences of the current collision set.
CollRanges ranges(const CollView& C, int64_t& M)
{
    const int npk[4] = { 2, 3, 4, 4 };
    CollRanges G;
    G.goff[0] = 0;
    M = 0;
    for (int k = 0; k < 4; k++) {
        G.goff[k + 1] = G.goff[k] + C.n[k];
        M += (int64_t)npk[k] * C.n[k];
    }
    if (G.goff[4] >= (1ll << 30)) throw Error(IPCG_ERR_ARG, "too many collisions for one context");
    return G;
}

// The order in which the per-vertex row kernels visit the vertices: Morton
// order of the vertex positions with one scale for the three axes (cubic
// cells of the scene's box, 13 bits per axis), scenes of a batch in turn.
// The vertices of one collision are within dhat of each other, so they are
// visited close together and the records they share are read while they are
// still in L2.  No result depends on the order: every row is computed on
// its own.
__device__ __forceinline__ uint64_t spread3(uint64_t v)
{
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(256) void k_row_keys(const double* X, int nV, const int32_t* vscene,
                                                  const double* aabb, uint64_t* keys,
                                                  uint32_t* ids)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nV) return;
    const int sc = vscene ? vscene[i] : 0;
    const double* b = aabb + 6 * (int64_t)sc;
    double ext = b[3] - b[0];
    ext = b[4] - b[1] > ext ? b[4] - b[1] : ext;
    ext = b[5] - b[2] > ext ? b[5] - b[2] : ext;
    const double inv = ext > 0 ? 8192.0 / ext : 0.0;
    const d3 x = ldx(X, nV, (int)i);
    const double xs[3] = { x.x, x.y, x.z };
    uint64_t code = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        double t = (xs[k] - b[k]) * inv;
        t = t < 0.0 ? 0.0 : (t > 8191.0 ? 8191.0 : t);
        code |= spread3((uint64_t)t) << (2 - k);
    }
    keys[i] = ((uint64_t)sc << 39) | code;
    ids[i] = (uint32_t)i;
}

static void build_row_order(Ctx& c)
{
    c.ro_ready = false;
    if (!c.row_order_valid || c.nV <= 1) return;
    hipStream_t s = c.stream;
    const int64_t n = c.nV;
    c.ro_keys.reserve(sizeof(uint64_t) * (size_t)n);
    c.ro_keys2.reserve(sizeof(uint64_t) * (size_t)n);
    c.ro_ids.reserve(sizeof(uint32_t) * (size_t)n);
    c.ro_ids2.reserve(sizeof(uint32_t) * (size_t)n);
    c.blk_tmp.reserve(radix_tmp_bytes(n));
    hipLaunchKernelGGL(k_row_keys, dim3(grid_for(n, 256)), dim3(256), 0, s, c.X[0].as<double>(),
                       c.nV, c.batched() ? c.vscene.as<int32_t>() : nullptr, c.aabb.as<double>(),
                       c.ro_keys.as<uint64_t>(), c.ro_ids.as<uint32_t>());
    IPCG_CHECK_LAUNCH();
    int bits = 39;
    if (c.batched())
        while ((1ll << (bits - 39)) < c.n_scenes) bits++;
    radix_sort_pairs(c.ro_keys.as<uint64_t>(), c.ro_ids.as<uint32_t>(), c.ro_keys2.as<uint64_t>(),
                     c.ro_ids2.as<uint32_t>(), n, bits, c.blk_tmp.p, s);
    c.ro_ready = true;
}

// the order in which the per-vertex row kernels visit the vertices (null:
// vertex order)
static const uint32_t* row_order(const Ctx& c)
{
    return c.ro_ready ? c.ro_ids.as<uint32_t>() : nullptr;
}

// (vertex, global collision g, point a) incidences sorted by vertex (stable,
// so each vertex's incidences keep the collision order) and the first
// incidence of every vertex; kept until the next collision build.
void require_incidences(Ctx& c, const CollView& C, const CollRanges& G, int64_t M)
{
    if (c.inc_valid) return;
    hipStream_t s = c.stream;
    c.stage_begin("incidence");
    const int64_t mk = std::max<int64_t>(M, 1);
    c.inc_keys.reserve(sizeof(uint32_t) * mk);
    c.inc_keys2.reserve(sizeof(uint32_t) * mk);
    c.inc_vals.reserve(sizeof(uint32_t) * mk);
    c.inc_vals2.reserve(sizeof(uint32_t) * mk);
    c.inc_ptr.reserve(sizeof(uint32_t) * ((size_t)c.nV + 1));
    c.blk_tmp.reserve(radix_tmp_bytes(mk));
    uint32_t* keys = c.inc_keys.as<uint32_t>();
    uint32_t* vals = c.inc_vals.as<uint32_t>();
    int64_t io = 0;
    if (C.n[0]) hipLaunchKernelGGL(k_incidence<K_VV>, dim3(grid_for(C.n[0], 256, INT32_MAX)), dim3(256), 0, s, C, G.goff[0], io, keys, vals);
    io += 2 * C.n[0];
    if (C.n[1]) hipLaunchKernelGGL(k_incidence<K_EV>, dim3(grid_for(C.n[1], 256, INT32_MAX)), dim3(256), 0, s, C, G.goff[1], io, keys, vals);
    io += 3 * C.n[1];
    if (C.n[2]) hipLaunchKernelGGL(k_incidence<K_EE>, dim3(grid_for(C.n[2], 256, INT32_MAX)), dim3(256), 0, s, C, G.goff[2], io, keys, vals);
    io += 4 * C.n[2];
    if (C.n[3]) hipLaunchKernelGGL(k_incidence<K_FV>, dim3(grid_for(C.n[3], 256, INT32_MAX)), dim3(256), 0, s, C, G.goff[3], io, keys, vals);
    IPCG_CHECK_LAUNCH();
    int vb = 1;
    while ((1ll << vb) <= c.nV) vb++;
    if (radix_sort_pairs_all(keys, vals, c.inc_keys2.as<uint32_t>(), c.inc_vals2.as<uint32_t>(),
                             M, vb, c.blk_tmp.p, s)) { // sorted pairs in the second buffers
        std::swap(c.inc_keys, c.inc_keys2);
        std::swap(c.inc_vals, c.inc_vals2);
        keys = c.inc_keys.as<uint32_t>();
        vals = c.inc_vals.as<uint32_t>();
    }
    if (M >= (int64_t)RUN_PTR_NONE) throw Error(IPCG_ERR_ARG, "too many collision incidences");
    IPCG_HIP(hipMemsetAsync(c.inc_ptr.p, 0xFF, sizeof(uint32_t) * (size_t)c.nV, s));
    hipLaunchKernelGGL(k_run_ptr, dim3(grid_for(M + 1, 256, INT32_MAX)), dim3(256), 0, s, keys, M,
                       c.nV, c.inc_ptr.as<uint32_t>());
    hipLaunchKernelGGL(k_gap_ptr, dim3(grid_for(c.nV, 256)), dim3(256), 0, s, keys, M, c.nV,
                       c.inc_ptr.as<uint32_t>());
    c.inc_vtx.reserve(sizeof(int4) * (size_t)mk);
    hipLaunchKernelGGL(k_inc_vertices, dim3(grid_for(M, 256, INT32_MAX)), dim3(256), 0, s, C, G,
                       vals, M, c.inc_vtx.as<int4>());
    IPCG_CHECK_LAUNCH();
    build_row_order(c);
    c.inc_valid = true;
    c.stage_end();
}

void pot_gradient(Ctx& c, double dhat, double kappa)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    int64_t M;
    const CollRanges G = ranges(C, M);
    require_incidences(c, C, G, M);
    c.grad.reserve(sizeof(double) * 3 * (size_t)std::max(c.nV, 1));
    double* g = c.grad.as<double>();
    const double* X = c.X[0].as<double>();
    // records share the Hessian record buffer (rewritten by every Hessian)
    const int npk[4] = { 2, 3, 4, 4 };
    GradRecs R;
    int64_t nrec = 0;
    for (int k = 0; k < 4; k++) nrec += 3 * (int64_t)npk[k] * C.n[k];
    c.hrec.reserve(sizeof(double) * (size_t)std::max<int64_t>(nrec, 1));
    int64_t ro = 0;
    for (int k = 0; k < 4; k++) {
        R.rec[k] = c.hrec.as<double>() + ro;
        ro += 3 * (int64_t)npk[k] * C.n[k];
    }
    c.stage_begin("gradient");
    if (C.n[0]) hipLaunchKernelGGL(k_gradient_records<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[0]));
    if (C.n[1]) hipLaunchKernelGGL(k_gradient_records<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[1]));
    if (C.n[2]) hipLaunchKernelGGL(k_gradient_records<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[2]));
    if (C.n[3]) hipLaunchKernelGGL(k_gradient_records<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[3]));
    const unsigned gw = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_gradient_rows, dim3(gw), dim3(256), 0, s, G, R, c.inc_ptr.as<uint32_t>(),
                       c.inc_vals.as<uint32_t>(), c.nV, g, row_order(c));
    IPCG_CHECK_LAUNCH();
    c.stage_end();
}

// Block sparsity pattern of the collision set (rebuilt after every build):
// bptr (nV+1) block-row pointers, bcol sorted block columns, one diagonal
// block per vertex that appears in a collision.
void hess_pattern(Ctx& c)
{
    hipStream_t s = c.stream;
    CollView C = view(c);
    const int np2[4] = { 1, 3, 6, 6 }; // pairs per stencil
    int64_t npair = 0;
    for (int k = 0; k < 4; k++) npair += (int64_t)np2[k] * C.n[k];
    const int64_t nk = std::max<int64_t>(npair, 1);
    c.blk_keys.reserve(sizeof(uint64_t) * nk);
    c.blk_keys2.reserve(sizeof(uint64_t) * nk);
    c.blk_vals.reserve(sizeof(uint32_t) * nk);
    c.blk_vals2.reserve(sizeof(uint32_t) * nk);
    c.blk_ukeys.reserve(sizeof(uint64_t) * nk);
    c.blk_start.reserve(sizeof(uint32_t) * (nk + 1));
    c.blk_tmp.reserve(std::max(std::max(radix_tmp_bytes(nk), rle_tmp_bytes(nk)),
                               scan_tmp_bytes(3 * (int64_t)c.nV + 1)));
    const size_t nv1 = (size_t)c.nV + 1;
    c.hp_mark.reserve(sizeof(uint32_t) * nv1 * 3);
    uint32_t* mark = c.hp_mark.as<uint32_t>();
    uint32_t* deg = mark + nv1;
    uint32_t* cursor = deg + nv1;
    IPCG_HIP(hipMemsetAsync(mark, 0, sizeof(uint32_t) * nv1 * 3, s));
    uint64_t* keys = c.blk_keys.as<uint64_t>();
    uint32_t* vals = c.blk_vals.as<uint32_t>();
    int64_t ko = 0;
    if (C.n[0]) hipLaunchKernelGGL(k_pair_keys<K_VV>, dim3(grid_for(C.n[0], 256, INT32_MAX)), dim3(256), 0, s, C, keys, vals, ko, mark);
    ko += np2[0] * C.n[0];
    if (C.n[1]) hipLaunchKernelGGL(k_pair_keys<K_EV>, dim3(grid_for(C.n[1], 256, INT32_MAX)), dim3(256), 0, s, C, keys, vals, ko, mark);
    ko += np2[1] * C.n[1];
    if (C.n[2]) hipLaunchKernelGGL(k_pair_keys<K_EE>, dim3(grid_for(C.n[2], 256, INT32_MAX)), dim3(256), 0, s, C, keys, vals, ko, mark);
    ko += np2[2] * C.n[2];
    if (C.n[3]) hipLaunchKernelGGL(k_pair_keys<K_FV>, dim3(grid_for(C.n[3], 256, INT32_MAX)), dim3(256), 0, s, C, keys, vals, ko, mark);
    IPCG_CHECK_LAUNCH();
    int vb = 1;
    while ((1ll << vb) <= c.nV) vb++;
    radix_sort_pairs(keys, vals, c.blk_keys2.as<uint64_t>(), c.blk_vals2.as<uint32_t>(), npair,
                     32 + vb, c.blk_tmp.p, s);
    uint32_t* nu_dev = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 9);
    run_length_encode(keys, npair, c.blk_ukeys.as<uint64_t>(), c.blk_start.as<uint32_t>(), nu_dev,
                      c.blk_tmp.p, s);
    hipLaunchKernelGGL(k_pair_degree, dim3(grid_for(nk, 256, INT32_MAX)), dim3(256), 0, s,
                       c.blk_ukeys.as<uint64_t>(), nu_dev, deg);
    hipLaunchKernelGGL(k_add_mark, dim3(grid_for(c.nV, 256)), dim3(256), 0, s, mark, c.nV, deg);
    c.hp_bptr.reserve(sizeof(uint32_t) * nv1);
    uint32_t* bptr = c.hp_bptr.as<uint32_t>();
    exclusive_scan_u32(deg, bptr, c.nV, bptr + c.nV, c.blk_tmp.p, s);
    IPCG_CHECK_LAUNCH();
    uint32_t* h = static_cast<uint32_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, bptr + c.nV, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    c.nblocks = h[0];
    c.hp_bcol.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(c.nblocks, 1));
    int32_t* bcol = c.hp_bcol.as<int32_t>();
    hipLaunchKernelGGL(k_fill_diag, dim3(grid_for(c.nV, 256)), dim3(256), 0, s, mark, c.nV, bptr,
                       cursor, bcol);
    hipLaunchKernelGGL(k_fill_pairs, dim3(grid_for(nk, 256, INT32_MAX)), dim3(256), 0, s,
                       c.blk_ukeys.as<uint64_t>(), nu_dev, bptr, cursor, bcol);
    hipLaunchKernelGGL(k_sort_rows, dim3(grid_for(c.nV, 256)), dim3(256), 0, s, bptr, c.nV, bcol);
    IPCG_CHECK_LAUNCH();
    c.pattern_valid = true;
}

// Block pattern from the vertex incidences (inc_ptr / inc_vtx valid).
// Returns false, leaving the pattern invalid, when a row has more distinct
// columns than the LDS path holds.
bool hess_pattern_incidences(Ctx& c, int64_t M)
{
    hipStream_t s = c.stream;
    const size_t nv1 = (size_t)c.nV + 1;
    c.blk_keys.reserve(sizeof(int32_t) * 4 * (size_t)std::max<int64_t>(M, 1));
    c.hp_mark.reserve(sizeof(uint32_t) * nv1 * 3);
    c.hp_bptr.reserve(sizeof(uint32_t) * nv1);
    c.blk_tmp.reserve(scan_tmp_bytes(std::max<int64_t>(c.nV, 1)));
    uint32_t* cnt = c.hp_mark.as<uint32_t>();
    uint32_t* bptr = c.hp_bptr.as<uint32_t>();
    int32_t* stage = c.blk_keys.as<int32_t>();
    int* ovf = reinterpret_cast<int*>(c.counters.as<unsigned long long>() + 10);
    IPCG_HIP(hipMemsetAsync(ovf, 0, sizeof(int), s));
    const unsigned gw = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_pattern_rows, dim3(gw), dim3(256), 0, s, c.inc_ptr.as<uint32_t>(),
                       c.inc_vtx.as<int4>(), c.nV, stage, cnt, ovf);
    exclusive_scan_u32(cnt, bptr, c.nV, bptr + c.nV, c.blk_tmp.p, s);
    IPCG_CHECK_LAUNCH();
    uint32_t* h = static_cast<uint32_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, bptr + c.nV, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipMemcpyAsync(h + 1, ovf, sizeof(int), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    if (h[1]) return false;
    c.nblocks = h[0];
    c.hp_bcol.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(c.nblocks, 1));
    hipLaunchKernelGGL(k_pattern_pack, dim3(gw), dim3(256), 0, s, c.inc_ptr.as<uint32_t>(), stage,
                       bptr, c.nV, c.hp_bcol.as<int32_t>());
    IPCG_CHECK_LAUNCH();
    c.pattern_valid = true;
    return true;
}

// incidences and block pattern of the current collision set; the record
// buffer sized for REC_LEN values per collision
static CollRanges hess_prepare(Ctx& c, const CollView& C)
{
    int64_t M;
    const CollRanges G = ranges(C, M);
    require_incidences(c, C, G, M);
    // block pattern: from the incidences, or by sorting the stencil pairs
    // when a row outgrows the LDS path
    c.stage_begin("hessian_pattern");
    if (!c.pattern_valid) {
        const bool hashed = hess_pattern_incidences(c, M);
        if (!hashed) hess_pattern(c); // a row outgrew the LDS hash
        c.stat["hessian.pattern_sorted"] = hashed ? 0 : 1;
    }
    c.stage_end();
    c.hrec.reserve(sizeof(double) * REC_LEN * (size_t)std::max<int64_t>(G.goff[4], 1));
    return G;
}

static void hess_assemble(Ctx& c);

void pot_hessian(Ctx& c, double dhat, double kappa, int psd)
{
    require_collisions(c);
    if (psd < 0 || psd > 2) throw Error(IPCG_ERR_ARG, "invalid PSD projection method");
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, psd);
    const double* X = c.X[0].as<double>();
    const CollRanges G = hess_prepare(c, C);
    // projected Hessian records, REC_LEN values per collision in global
    // collision order
    double* rec = c.hrec.as<double>();
    // one launch per stencil kind, each its own timed stage
    if (C.n[0]) {
        c.stage_begin("hessian_records.vv");
        hipLaunchKernelGGL(k_hessian_records<K_VV>, dim3(grid_for(C.n[0], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, rec + REC_LEN * G.goff[0]);
        c.stage_end();
    }
    if (C.n[1]) {
        c.stage_begin("hessian_records.ev");
        hipLaunchKernelGGL(k_hessian_records<K_EV>, dim3(grid_for(C.n[1], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, rec + REC_LEN * G.goff[1]);
        c.stage_end();
    }
    if (C.n[2]) {
        // interior edge-edge Hessians by the rank-5 path, the other types
        // (listed by a scan of their flags) by the 9 x 9 Jacobi
        const int64_t ne = C.n[2];
        c.ee_other.reserve(sizeof(uint32_t) * 3 * (size_t)ne);
        uint32_t* flag = c.ee_other.as<uint32_t>();
        uint32_t* pos = flag + ne;
        uint32_t* list = pos + ne;
        c.blk_tmp.reserve(scan_tmp_bytes(ne));
        hipLaunchKernelGGL(k_ee_other_flags, dim3(grid_for(ne, 256, INT32_MAX)), dim3(256), 0, s, C.dtype, ne, flag);
        uint32_t* n_other = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 11);
        exclusive_scan_u32(flag, pos, ne, n_other, c.blk_tmp.p, s);
        hipLaunchKernelGGL(k_ee_other_list, dim3(grid_for(ne, 256, INT32_MAX)), dim3(256), 0, s, flag, pos, ne, list);
        IPCG_CHECK_LAUNCH();
        uint32_t* h = static_cast<uint32_t*>(c.pinned);
        IPCG_HIP(hipMemcpyAsync(h, n_other, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        IPCG_HIP(hipStreamSynchronize(s));
        const int64_t no = h[0];
        c.stat["hessian.ee_other"] = no;
        c.stage_begin("hessian_records.ee");
        hipLaunchKernelGGL((k_hessian_records<K_EE, REC_LOW>), dim3(grid_for(ne, 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, rec + REC_LEN * G.goff[2], nullptr, (int64_t)0);
        c.stage_end();
        if (no > 0) {
            c.stage_begin("hessian_records.ee_other");
            hipLaunchKernelGGL((k_hessian_records<K_EE, REC_LIST>), dim3(grid_for(no, 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, rec + REC_LEN * G.goff[2], list, no);
            c.stage_end();
        }
    }
    if (C.n[3]) {
        c.stage_begin("hessian_records.fv");
        hipLaunchKernelGGL((k_hessian_records<K_FV, REC_LOW>), dim3(grid_for(C.n[3], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, rec + REC_LEN * G.goff[3], nullptr, (int64_t)0);
        c.stage_end();
    }
    IPCG_CHECK_LAUNCH();
    hess_assemble(c);
}

// block rows from the records in c.hrec, then the CSR
static void hess_assemble(Ctx& c)
{
    hipStream_t s = c.stream;
    const double* rec = c.hrec.as<double>();
    const int64_t nb = std::max<int64_t>(c.nblocks, 1);
    c.blk_mask.reserve(sizeof(uint32_t) * nb);
    c.hq.reserve(sizeof(double) * 9 * nb);
    uint32_t* bmask = c.blk_mask.as<uint32_t>();
    double* bval = c.hq.as<double>();
    const uint32_t* bptr = c.hp_bptr.as<uint32_t>();
    const int32_t* bcol = c.hp_bcol.as<int32_t>();
    c.stage_begin("hessian_rows");
    const unsigned gwr = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_hessian_rows, dim3(gwr), dim3(256), 0, s, rec, c.inc_ptr.as<uint32_t>(),
                       c.inc_vals.as<uint32_t>(), c.inc_vtx.as<int4>(), bptr, bcol, c.nV, bval,
                       bmask, row_order(c));
    IPCG_CHECK_LAUNCH();
    c.stage_end();

    c.stage_begin("hessian_assemble");
    const int64_t nrows = 3 * (int64_t)c.nV;
    c.row_cnt.reserve(sizeof(uint64_t) * (nrows + 1));
    c.rowptr.reserve(sizeof(int64_t) * (nrows + 1));
    c.blk_tmp.reserve(scan_tmp_bytes(nrows + 1));
    const unsigned gw = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_row_counts, dim3(gw), dim3(256), 0, s, bptr, bmask, c.nV,
                       c.row_cnt.as<uint64_t>());
    uint64_t* rp = reinterpret_cast<uint64_t*>(c.rowptr.p);
    exclusive_scan_u64(c.row_cnt.as<uint64_t>(), rp, nrows, rp + nrows, c.blk_tmp.p, s);
    IPCG_CHECK_LAUNCH();
    int64_t* h = static_cast<int64_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, rp + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    c.nnz = h[0];
    c.col.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(c.nnz, 1));
    c.val.reserve(sizeof(double) * (size_t)std::max<int64_t>(c.nnz, 1));
    hipLaunchKernelGGL(k_fill_rows, dim3(gw), dim3(256), 0, s, bptr, bcol, bmask, bval, c.nV,
                       c.rowptr.as<int64_t>(), c.col.as<int32_t>(), c.val.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    c.hess_valid = true;
}

// NormalPotential::gauss_newton_hessian_diagonal (normal_potential.cpp:
// 15-46) into c.gn_out (3 nV, RowMajor), summed per vertex in incidence
// order like the gradient.
void pot_gn_diagonal(Ctx& c, double dhat, double kappa)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    int64_t M;
    const CollRanges G = ranges(C, M);
    require_incidences(c, C, G, M);
    c.gn_out.reserve(sizeof(double) * 3 * (size_t)std::max(c.nV, 1));
    const double* X = c.X[0].as<double>();
    const int npk[4] = { 2, 3, 4, 4 };
    GradRecs R;
    int64_t nrec = 0;
    for (int k = 0; k < 4; k++) nrec += 3 * (int64_t)npk[k] * C.n[k];
    c.hrec.reserve(sizeof(double) * (size_t)std::max<int64_t>(nrec, 1));
    int64_t ro = 0;
    for (int k = 0; k < 4; k++) {
        R.rec[k] = c.hrec.as<double>() + ro;
        ro += 3 * (int64_t)npk[k] * C.n[k];
    }
    c.stage_begin("gn_diagonal");
    if (C.n[0]) hipLaunchKernelGGL(k_gn_diag_records<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[0]));
    if (C.n[1]) hipLaunchKernelGGL(k_gn_diag_records<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[1]));
    if (C.n[2]) hipLaunchKernelGGL(k_gn_diag_records<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[2]));
    if (C.n[3]) hipLaunchKernelGGL(k_gn_diag_records<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, const_cast<double*>(R.rec[3]));
    const unsigned gw = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_gradient_rows, dim3(gw), dim3(256), 0, s, G, R, c.inc_ptr.as<uint32_t>(),
                       c.inc_vals.as<uint32_t>(), c.nV, c.gn_out.as<double>(), row_order(c));
    IPCG_CHECK_LAUNCH();
    c.stage_end();
}

// NormalPotential::gauss_newton_hessian_quadratic_form (normal_potential.cpp:
// 48-79); p_host: 3 nV RowMajor on the host
double pot_gn_quadratic_form(Ctx& c, double dhat, double kappa, const double* p_host)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    const size_t pb = sizeof(double) * 3 * (size_t)std::max(c.nV, 1);
    c.gn_p.reserve(pb);
    if (c.nV)
        IPCG_HIP(hipMemcpyAsync(c.gn_p.p, p_host, sizeof(double) * 3 * (size_t)c.nV,
                                hipMemcpyHostToDevice, s));
    int64_t nb[4], tot = 0;
    for (int k = 0; k < 4; k++) {
        nb[k] = (C.n[k] + PB - 1) / PB;
        tot += nb[k];
    }
    c.red_partial.reserve(sizeof(double) * (size_t)std::max<int64_t>(tot, 1));
    c.red_out.reserve(sizeof(double) * 2);
    double* part = c.red_partial.as<double>();
    const double* X = c.X[0].as<double>();
    const double* p = c.gn_p.as<double>();
    c.stage_begin("gn_quadratic_form");
    int64_t o = 0;
    if (nb[0]) hipLaunchKernelGGL(k_gn_quadratic_form<K_VV>, dim3(nb[0]), dim3(PB), 0, s, C, X, c.nV, bp, p, part + o);
    o += nb[0];
    if (nb[1]) hipLaunchKernelGGL(k_gn_quadratic_form<K_EV>, dim3(nb[1]), dim3(PB), 0, s, C, X, c.nV, bp, p, part + o);
    o += nb[1];
    if (nb[2]) hipLaunchKernelGGL(k_gn_quadratic_form<K_EE>, dim3(nb[2]), dim3(PB), 0, s, C, X, c.nV, bp, p, part + o);
    o += nb[2];
    if (nb[3]) hipLaunchKernelGGL(k_gn_quadratic_form<K_FV>, dim3(nb[3]), dim3(PB), 0, s, C, X, c.nV, bp, p, part + o);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, part, tot, c.red_out.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    double* h = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, c.red_out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    return h[0];
}


// ==========================================================================
// Friction: TangentialCollisions::build over the current collision set with
// global friction coefficients (tangential_collisions.cpp:61-186 with
// uniform per-vertex mu) and FrictionPotential energy / gradient / Hessian
// at velocities (tangential_potential.cpp:156-331, friction_potential.cpp,
// smooth_mu.cpp, smooth_friction_mollifier.cpp), isotropic.  One record per
// collision in global collision order, so the gradient rows and the
// Hessian rows / CSR of the collision set are reused.  The relative
// velocity is translation invariant, so a collision's Hessian T M T^T
// (T = Gamma^T P, M 2 x 2) lies in the 4-point Q basis, where it is the
// Kronecker product (s s^T) (x) (P M P^T) with s_k = sum_x q4(x, k) Gamma_x.
// ==========================================================================
namespace {

struct FricRec {
    double P[6];      // tangent basis, column major 3 x 2
    double beta[2];   // closest-point coordinates
    double s;         // weight * normal force magnitude; 0: not kept
    double keep;      // 1 kept, 0 skipped (near-parallel edge-edge)
    double mu_s, mu_k; // the collision's static / kinetic coefficients
    double an[2];      // mu_aniso: velocity scaling (default 1)
    double msa[2], mka[2]; // mu_s_aniso / mu_k_aniso: elliptical coefficients (default 0)
    double ms_eff, mk_eff; // lagged effective coefficients used by the potential
};

struct FricParams {
    double eps_v;
    int psd;
};

// tangential_collisions.hpp:143-149 default_blend_mu
__device__ __forceinline__ double blend_mu(double a, double b)
{
    return (a + b) / 2;
}

// tangential_collisions.cpp:86-166: each side's per-vertex coefficient at
// its closest point, blended
template <int KIND>
__device__ __forceinline__ double collision_mu(const int* v, const double* beta, const double* m)
{
    if constexpr (KIND == K_VV) return blend_mu(m[v[0]], m[v[1]]);
    if constexpr (KIND == K_EV) return blend_mu((m[v[2]] - m[v[1]]) * beta[0] + m[v[1]], m[v[0]]);
    if constexpr (KIND == K_EE)
        return blend_mu((m[v[1]] - m[v[0]]) * beta[0] + m[v[0]],
                        (m[v[3]] - m[v[2]]) * beta[1] + m[v[2]]);
    return blend_mu(m[v[1]] + beta[0] * (m[v[2]] - m[v[1]]) + beta[1] * (m[v[3]] - m[v[1]]), m[v[0]]);
}

// smooth_friction_mollifier.cpp:8-53
__device__ __forceinline__ double sf_f0(double y, double ev)
{
    if (fabs(y) >= ev) return y;
    return y * y * (1 - y / (3 * ev)) / ev + ev / 3;
}
__device__ __forceinline__ double sf_f1_over_x(double y, double ev)
{
    if (fabs(y) >= ev) return 1 / y;
    return (2 - y / ev) / ev;
}
__device__ __forceinline__ double sf_f2_x_minus_f1_over_x3(double y, double ev)
{
    if (fabs(y) >= ev) return -1 / (y * y * y);
    return -1 / (y * ev * ev);
}

// smooth_mu.cpp:10-112
__device__ __forceinline__ double smooth_mu(double y, double ms, double mk, double ev)
{
    if (ms == mk || fabs(y) >= ev) return mk;
    const double z = fabs(y) / ev;
    if (fabs(y) < 0.5 * ev) return 2 * (mk - ms) * z * z + ms;
    return -2 * (mk - ms) * (z * (z - 2) + 1) + mk;
}
__device__ __forceinline__ double smooth_mu_f0(double y, double ms, double mk, double ev)
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
__device__ __forceinline__ double smooth_mu_f1_over_x(double y, double ms, double mk, double ev)
{
    return smooth_mu(y, ms, mk, ev) * sf_f1_over_x(y, ev);
}
__device__ __forceinline__ double smooth_mu_f2_x_minus_f1_over_x3(double y, double ms, double mk,
                                                                  double ev)
{
    if (ms == mk || fabs(y) >= ev) return mk * sf_f2_x_minus_f1_over_x3(y, ev);
    const double dm = mk - ms;
    const double z = 1 / ev;
    if (fabs(y) < 0.5 * ev) return z * z * (z * (8 - 6 * y * z) * dm - ms / y);
    return z * z * (z * (6 * y * z - 16) * dm + (9 * mk - 10 * ms) / y);
}

// Eigen normalized(): divide by the norm when it is positive
__device__ __forceinline__ d3 normalized(const d3& a)
{
    const double z = sqn(a);
    if (!(z > 0)) return a;
    const double n = sqrt(z);
    return { a.x / n, a.y / n, a.z / n };
}

// tangent_basis.cpp:17-175 (columns c0, c1) and closest-point coordinates
// (closest_point.cpp:11-137)
template <int KIND>
__device__ void fric_geometry(const d3* P, d3& c0, d3& c1, double* beta)
{
    beta[0] = beta[1] = 0.0;
    if constexpr (KIND == K_VV) {
        const d3 d = sub(P[1], P[0]);
        const d3 cx = cross(d3 { 1, 0, 0 }, d), cy = cross(d3 { 0, 1, 0 }, d);
        const d3 w = sqn(cx) > sqn(cy) ? cx : cy;
        c0 = normalized(w);
        c1 = normalized(cross(d, w));
    } else if constexpr (KIND == K_EV) {
        const d3 e = sub(P[2], P[1]);
        c0 = normalized(e);
        c1 = normalized(cross(e, sub(P[0], P[1])));
        beta[0] = pe_closest(P[0], P[1], P[2]);
    } else if constexpr (KIND == K_EE) {
        const d3 ea = sub(P[1], P[0]), eb = sub(P[3], P[2]);
        const d3 nrm = cross(ea, eb);
        c0 = normalized(ea);
        c1 = normalized(cross(nrm, ea));
        const d3 d = sub(P[0], P[2]);
        solve_sym2_ldlt(sqn(ea), -dot(eb, ea), sqn(eb), -dot(d, ea), dot(d, eb), beta[0], beta[1]);
    } else {
        const d3 e0 = sub(P[2], P[1]), e1 = sub(P[3], P[1]);
        const d3 nrm = cross(e0, e1);
        c0 = normalized(e0);
        c1 = normalized(cross(nrm, e0));
        const d3 r = sub(P[0], P[1]);
        solve_sym2_ldlt(dot(e0, e0), dot(e1, e0), dot(e1, e1), dot(e0, r), dot(e1, r), beta[0],
                        beta[1]);
    }
}

// relative-velocity Jacobian coefficients Gamma (relative_velocity.cpp)
template <int KIND>
__device__ __forceinline__ void fric_coef(const double* beta, double* g)
{
    g[0] = g[1] = g[2] = g[3] = 0.0;
    if constexpr (KIND == K_VV) {
        g[0] = 1;
        g[1] = -1;
    } else if constexpr (KIND == K_EV) {
        g[0] = 1;
        g[1] = beta[0] - 1;
        g[2] = -beta[0];
    } else if constexpr (KIND == K_EE) {
        g[0] = 1 - beta[0];
        g[1] = beta[0];
        g[2] = beta[1] - 1;
        g[3] = -beta[1];
    } else {
        g[0] = 1;
        g[1] = beta[0] + beta[1] - 1;
        g[2] = -beta[0];
        g[3] = -beta[1];
    }
}

// relative velocity of the closest points (relative_velocity.cpp)
template <int KIND>
__device__ __forceinline__ d3 fric_relative_velocity(const d3* v, const double* beta)
{
    if constexpr (KIND == K_VV) return sub(v[0], v[1]);
    auto lerp = [](const d3& a, const d3& b, double t) {
        return d3 { (b.x - a.x) * t + a.x, (b.y - a.y) * t + a.y, (b.z - a.z) * t + a.z };
    };
    if constexpr (KIND == K_EV) return sub(v[0], lerp(v[1], v[2], beta[0]));
    if constexpr (KIND == K_EE) return sub(lerp(v[0], v[1], beta[0]), lerp(v[2], v[3], beta[1]));
    const d3 t = { (v[1].x + beta[0] * (v[2].x - v[1].x)) + beta[1] * (v[3].x - v[1].x),
                   (v[1].y + beta[0] * (v[2].y - v[1].y)) + beta[1] * (v[3].y - v[1].y),
                   (v[1].z + beta[0] * (v[2].z - v[1].z)) + beta[1] * (v[3].z - v[1].z) };
    return sub(v[0], t);
}

// msv / mkv: per-vertex coefficients, or null for the scalars ms / mk
template <int KIND>
__global__ __launch_bounds__(PB) void k_fric_build(CollView C, const double* X, int nV,
                                                   BarrierParams bp, const double* msv,
                                                   const double* mkv, double ms, double mk,
                                                   FricRec* out)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    FricRec r = {};
    double d2;
    if constexpr (KIND == K_EE) {
        // near-parallel pairs are skipped (tangential_collisions.cpp:118-121);
        // the rest are line-line (tangential edge_edge.hpp:27-31)
        if (ee_cross2(P[0], P[1], P[2], P[3]) < C.eps[i]) {
            out[i] = r;
            return;
        }
        d2 = ee_dist2(P[0], P[1], P[2], P[3], EE_EA_EB);
    } else if constexpr (KIND == K_VV) {
        d2 = pp_dist2(P[0], P[1]);
    } else if constexpr (KIND == K_EV) {
        d2 = pe_dist2(P[0], P[1], P[2], point_edge_type(P[0], P[1], P[2]));
    } else {
        d2 = pt_dist2(P[0], P[1], P[2], P[3], point_triangle_type(P[0], P[1], P[2], P[3]));
    }
    d3 c0, c1;
    fric_geometry<KIND>(P, c0, c1, r.beta);
    r.P[0] = c0.x;
    r.P[1] = c0.y;
    r.P[2] = c0.z;
    r.P[3] = c1.x;
    r.P[4] = c1.y;
    r.P[5] = c1.z;
    // barrier_force_magnitude.cpp:5-15
    const double N = -bar_f1(d2, bp) * 2 * sqrt(d2);
    r.s = C.w[KIND][i] * N;
    r.keep = 1.0;
    r.mu_s = msv ? collision_mu<KIND>(v, r.beta, msv) : ms;
    r.mu_k = mkv ? collision_mu<KIND>(v, r.beta, mkv) : mk;
    r.an[0] = r.an[1] = 1.0;
    r.ms_eff = r.mu_s;
    r.mk_eff = r.mu_k;
    out[i] = r;
}

template <int KIND>
__device__ __forceinline__ void fric_local(const FricRec& r, const double* Vel, int nV,
                                           const int* v, double ev, double ms, double mk,
                                           double* u, double& norm)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    d3 w[4];
#pragma unroll
    for (int j = 0; j < NP; j++) w[j] = ldx(Vel, nV, v[j]);
    const d3 rel = fric_relative_velocity<KIND>(w, r.beta);
    // tangential_potential.cpp:170-181: u_aniso = mu_aniso (.) u
    u[0] = r.an[0] * dot(d3 { r.P[0], r.P[1], r.P[2] }, rel);
    u[1] = r.an[1] * dot(d3 { r.P[3], r.P[4], r.P[5] }, rel);
    norm = sqrt(u[0] * u[0] + u[1] * u[1]);
}

template <int KIND>
__global__ __launch_bounds__(PB) void k_fric_energy(CollView C, const FricRec* R, const double* Vel,
                                                    int nV, FricParams fp, double* partial)
{
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    double e = 0;
    if (i < C.n[KIND] && R[i].keep != 0) {
        int v[4];
        stencil_ids<KIND>(C, i, v);
        double u[2], norm;
        fric_local<KIND>(R[i], Vel, nV, v, fp.eps_v, R[i].ms_eff, R[i].mk_eff, u, norm);
        e = R[i].s * smooth_mu_f0(norm, R[i].ms_eff, R[i].mk_eff, fp.eps_v);
    }
    for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o, WAVE);
    __shared__ double sh[PB / WAVE];
    if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = e;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int j = 0; j < PB / WAVE; j++) t += sh[j];
        partial[blockIdx.x] = t;
    }
}

// gradient records: 3 NP values per collision, point major
template <int KIND>
__global__ __launch_bounds__(PB) void k_fric_gradient_records(CollView C, const FricRec* R,
                                                              const double* Vel, int nV,
                                                              FricParams fp, double* rec)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    double* o = rec + (int64_t)3 * NP * i;
    const FricRec r = R[i];
    if (r.keep == 0) {
#pragma unroll
        for (int j = 0; j < 3 * NP; j++) o[j] = 0.0;
        return;
    }
    int v[4];
    stencil_ids<KIND>(C, i, v);
    double u[2], norm, g[4];
    fric_local<KIND>(r, Vel, nV, v, fp.eps_v, r.ms_eff, r.mk_eff, u, norm);
    fric_coef<KIND>(r.beta, g);
    const double f1x = smooth_mu_f1_over_x(norm, r.ms_eff, r.mk_eff, fp.eps_v);
    const double g0 = f1x * u[0] * r.s, g1 = f1x * u[1] * r.s;
#pragma unroll
    for (int j = 0; j < NP; j++)
#pragma unroll
        for (int k = 0; k < 3; k++)
            o[3 * j + k] = (g[j] * r.P[k] * r.an[0]) * g0 + (g[j] * r.P[3 + k] * r.an[1]) * g1;
}

// per-collision anisotropy (n x 2 each, null: keep), effective coefficients
// reset to mu_s, mu_k (reset_lagged_anisotropic_friction_coefficients)
__global__ void k_fric_set_aniso(FricRec* R, int64_t n, const double* an, const double* msa,
                                 const double* mka)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    FricRec& r = R[i];
    for (int k = 0; k < 2; k++) {
        if (an) r.an[k] = an[2 * i + k];
        if (msa) r.msa[k] = msa[2 * i + k];
        if (mka) r.mka[k] = mka[2 * i + k];
    }
    r.ms_eff = r.mu_s;
    r.mk_eff = r.mu_k;
}

// update_lagged_anisotropic_friction_coefficients (tangential_collisions.cpp:
// 19-57) at the build positions: tau = T^T v, tau_aniso = mu_aniso (.) tau,
// the elliptical model of smooth_mu.cpp:115-171 when mu_s_aniso is nonzero
template <int KIND>
__global__ __launch_bounds__(PB) void k_fric_update_lagged(CollView C, FricRec* R, const double* Vel,
                                                           int nV)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    FricRec& r = R[i];
    r.ms_eff = r.mu_s;
    r.mk_eff = r.mu_k;
    if (r.keep == 0 || r.msa[0] * r.msa[0] + r.msa[1] * r.msa[1] == 0) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    double g[4];
    fric_coef<KIND>(r.beta, g);
    double t0 = 0, t1 = 0;
#pragma unroll
    for (int j = 0; j < NP; j++) {
        const d3 w = ldx(Vel, nV, v[j]);
        t0 += g[j] * r.P[0] * w.x;
        t0 += g[j] * r.P[1] * w.y;
        t0 += g[j] * r.P[2] * w.z;
        t1 += g[j] * r.P[3] * w.x;
        t1 += g[j] * r.P[4] * w.y;
        t1 += g[j] * r.P[5] * w.z;
    }
    const double ta0 = r.an[0] * t0, ta1 = r.an[1] * t1;
    const bool use_k = r.mka[0] * r.mka[0] + r.mka[1] * r.mka[1] > 0;
    double d0 = 1.0, d1 = 0.0;
    const double nrm = sqrt(ta0 * ta0 + ta1 * ta1);
    if (!(nrm < 1e-10)) {
        d0 = ta0 / nrm;
        d1 = ta1 / nrm;
    }
    double a = r.msa[0] * d0, b = r.msa[1] * d1;
    r.ms_eff = sqrt(a * a + b * b);
    if (use_k) {
        a = r.mka[0] * d0;
        b = r.mka[1] * d1;
        r.mk_eff = sqrt(a * a + b * b);
    }
}

// 2 x 2 symmetric PSD projection (eigen decomposition in closed form)
__device__ void psd2(double& a, double& b, double& d, int method)
{
    if (method == 0) return;
    const double tr = 0.5 * (a + d), df = 0.5 * (a - d);
    const double rad = sqrt(df * df + b * b);
    const double l0 = tr - rad, l1 = tr + rad;
    if (l0 >= 0) return; // already PSD
    auto fix = [&](double l) { return method == 1 ? (l > 0 ? l : 0.0) : fabs(l); };
    const double m0 = fix(l0), m1 = fix(l1);
    // eigenvector of l1: (b, l1 - a) or (l1 - d, b), the longer of the two
    double x, y;
    if (fabs(l1 - a) >= fabs(l1 - d)) {
        x = b;
        y = l1 - a;
    } else {
        x = l1 - d;
        y = b;
    }
    const double n = sqrt(x * x + y * y);
    if (n == 0) { // a multiple of the identity
        a = d = m1;
        b = 0;
        return;
    }
    x /= n;
    y /= n;
    // A = m1 e1 e1^T + m0 e0 e0^T with e0 = (-y, x)
    a = m1 * x * x + m0 * y * y;
    b = (m1 - m0) * x * y;
    d = m1 * y * y + m0 * x * x;
}

// Hessian records in the 4-point Q basis (REC_LEN packed values)
template <int KIND>
__global__ __launch_bounds__(PB) void k_fric_hessian_records(CollView C, const FricRec* R,
                                                             const double* Vel, int nV,
                                                             FricParams fp, double* rec)
{
    const int64_t i = (int64_t)xcd_block() * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    double* o = rec + (int64_t)REC_LEN * i;
    const FricRec r = R[i];
    double m00 = 0, m01 = 0, m11 = 0;
    double sq[3] = { 0, 0, 0 };
    if (r.keep != 0) {
        int v[4];
        stencil_ids<KIND>(C, i, v);
        double u[2], norm, g[4];
        fric_local<KIND>(r, Vel, nV, v, fp.eps_v, r.ms_eff, r.mk_eff, u, norm);
        fric_coef<KIND>(r.beta, g);
        const double f1x = smooth_mu_f1_over_x(norm, r.ms_eff, r.mk_eff, fp.eps_v);
        const double scale = r.s;
        if (norm > fp.eps_v) { // sliding: f2 = 0, rank one along u_perp
            if (!(fp.psd != 0 && scale <= 0)) {
                const double cc = scale * f1x / (norm * norm);
                m00 = cc * u[1] * u[1];
                m01 = -cc * u[1] * u[0];
                m11 = cc * u[0] * u[0];
            }
        } else if (norm == 0) {
            if (!(fp.psd != 0 && scale <= 0)) m00 = m11 = scale * f1x;
        } else {
            const double f2 = smooth_mu_f2_x_minus_f1_over_x3(norm, r.ms_eff, r.mk_eff, fp.eps_v);
            m00 = (f2 * u[0] * u[0] + f1x) * scale;
            m01 = (f2 * u[0] * u[1]) * scale;
            m11 = (f2 * u[1] * u[1] + f1x) * scale;
            psd2(m00, m01, m11, fp.psd);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
            double t = 0;
#pragma unroll
            for (int x = 0; x < 4; x++) t += QBasis<4>::q(x, k) * g[x];
            sq[k] = t;
        }
    }
    // A = P' M P'^T with P' = P diag(mu_aniso) (T_aniso = Gamma^T P')
    double A[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const double pa0 = r.P[a] * r.an[0], pa1 = r.P[3 + a] * r.an[1];
            const double pb0 = r.P[b] * r.an[0], pb1 = r.P[3 + b] * r.an[1];
            A[a][b] = pa0 * (m00 * pb0 + m01 * pb1) + pa1 * (m01 * pb0 + m11 * pb1);
        }
#pragma unroll
    for (int I = 0; I < 9; I++)
#pragma unroll
        for (int J = 0; J <= I; J++) o[sp(I, J)] = sq[I / 3] * sq[J / 3] * A[I % 3][J % 3];
}

} // namespace

static void require_friction(Ctx& c)
{
    require_collisions(c);
    if (!c.fric_valid) throw Error(IPCG_ERR_STATE, "no friction set (build it after the collisions)");
    if (!c.has_vel) throw Error(IPCG_ERR_STATE, "no velocities uploaded");
}

// tangential_collisions.cpp:61-186 over the current collision set at the
// positions in slot 0; mu_sv / mu_kv: per-vertex coefficients on the host
// (nV each) or null for the scalars
void fric_build(Ctx& c, double dhat, double kappa, double mu_s, double mu_k, const double* mu_sv,
                const double* mu_kv)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    int64_t M;
    const CollRanges G = ranges(C, M);
    c.fric.reserve(sizeof(FricRec) * (size_t)std::max<int64_t>(G.goff[4], 1));
    FricRec* R = c.fric.as<FricRec>();
    const BarrierParams bp = barrier_params(c, dhat, kappa, 0);
    const double* X = c.X[0].as<double>();
    const double *msv = nullptr, *mkv = nullptr;
    if (mu_sv && mu_kv && c.nV) {
        c.fric_mu.reserve(sizeof(double) * 2 * (size_t)c.nV);
        IPCG_HIP(hipMemcpyAsync(c.fric_mu.p, mu_sv, sizeof(double) * c.nV, hipMemcpyHostToDevice, s));
        IPCG_HIP(hipMemcpyAsync(c.fric_mu.as<double>() + c.nV, mu_kv, sizeof(double) * c.nV,
                                hipMemcpyHostToDevice, s));
        msv = c.fric_mu.as<double>();
        mkv = msv + c.nV;
    }
    c.stage_begin("friction_build");
    if (C.n[0]) hipLaunchKernelGGL(k_fric_build<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, msv, mkv, mu_s, mu_k, R + G.goff[0]);
    if (C.n[1]) hipLaunchKernelGGL(k_fric_build<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, msv, mkv, mu_s, mu_k, R + G.goff[1]);
    if (C.n[2]) hipLaunchKernelGGL(k_fric_build<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, msv, mkv, mu_s, mu_k, R + G.goff[2]);
    if (C.n[3]) hipLaunchKernelGGL(k_fric_build<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, msv, mkv, mu_s, mu_k, R + G.goff[3]);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    IPCG_HIP(hipStreamSynchronize(s)); // the host coefficient arrays may go away
    c.fric_valid = true;
}

// per collision (global order): keep flag, coordinates (2), basis (6,
// column major), normal force magnitude; any pointer may be null
void fric_download(Ctx& c, int32_t* keep, double* beta, double* basis, double* normal_force)
{
    require_collisions(c);
    if (!c.fric_valid) throw Error(IPCG_ERR_STATE, "no friction set");
    int64_t n = 0;
    for (int k = 0; k < 4; k++) n += c.ncoll[k];
    std::vector<FricRec> h((size_t)n);
    if (n) IPCG_HIP(hipMemcpyAsync(h.data(), c.fric.p, sizeof(FricRec) * n, hipMemcpyDeviceToHost, c.stream));
    IPCG_HIP(hipStreamSynchronize(c.stream));
    // the weight is the collision's; the normal force is s / weight only
    // when the weight is nonzero, so download the weights to divide
    std::vector<double> w((size_t)n);
    int64_t o = 0;
    for (int k = 0; k < 4; k++) {
        if (c.ncoll[k])
            IPCG_HIP(hipMemcpyAsync(w.data() + o, c.c_w[k].p, sizeof(double) * c.ncoll[k], hipMemcpyDeviceToHost, c.stream));
        o += c.ncoll[k];
    }
    IPCG_HIP(hipStreamSynchronize(c.stream));
    for (int64_t i = 0; i < n; i++) {
        if (keep) keep[i] = h[i].keep != 0;
        if (beta) {
            beta[2 * i] = h[i].beta[0];
            beta[2 * i + 1] = h[i].beta[1];
        }
        if (basis)
            for (int j = 0; j < 6; j++) basis[6 * i + j] = h[i].P[j];
        if (normal_force) normal_force[i] = h[i].keep != 0 && w[i] != 0 ? h[i].s / w[i] : 0.0;
    }
}

double fric_energy(Ctx& c, double eps_v)
{
    require_friction(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    int64_t M;
    const CollRanges G = ranges(C, M);
    const FricRec* R = c.fric.as<FricRec>();
    const FricParams fp { eps_v, 0 };
    int64_t nb[4], tot = 0;
    for (int k = 0; k < 4; k++) {
        nb[k] = (C.n[k] + PB - 1) / PB;
        tot += nb[k];
    }
    c.red_partial.reserve(sizeof(double) * (size_t)std::max<int64_t>(tot, 1));
    c.red_out.reserve(sizeof(double) * 2);
    double* part = c.red_partial.as<double>();
    const double* Vel = c.vel.as<double>();
    c.stage_begin("friction_energy");
    int64_t o = 0;
    if (nb[0]) hipLaunchKernelGGL(k_fric_energy<K_VV>, dim3(nb[0]), dim3(PB), 0, s, C, R + G.goff[0], Vel, c.nV, fp, part + o);
    o += nb[0];
    if (nb[1]) hipLaunchKernelGGL(k_fric_energy<K_EV>, dim3(nb[1]), dim3(PB), 0, s, C, R + G.goff[1], Vel, c.nV, fp, part + o);
    o += nb[1];
    if (nb[2]) hipLaunchKernelGGL(k_fric_energy<K_EE>, dim3(nb[2]), dim3(PB), 0, s, C, R + G.goff[2], Vel, c.nV, fp, part + o);
    o += nb[2];
    if (nb[3]) hipLaunchKernelGGL(k_fric_energy<K_FV>, dim3(nb[3]), dim3(PB), 0, s, C, R + G.goff[3], Vel, c.nV, fp, part + o);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, part, tot, c.red_out.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    double* h = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, c.red_out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    return h[0];
}

void fric_gradient(Ctx& c, double eps_v)
{
    require_friction(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    int64_t M;
    const CollRanges G = ranges(C, M);
    require_incidences(c, C, G, M);
    const FricRec* R = c.fric.as<FricRec>();
    const FricParams fp { eps_v, 0 };
    c.grad.reserve(sizeof(double) * 3 * (size_t)std::max(c.nV, 1));
    const int npk[4] = { 2, 3, 4, 4 };
    GradRecs GR;
    int64_t nrec = 0;
    for (int k = 0; k < 4; k++) nrec += 3 * (int64_t)npk[k] * C.n[k];
    c.hrec.reserve(sizeof(double) * (size_t)std::max<int64_t>(nrec, 1));
    int64_t ro = 0;
    for (int k = 0; k < 4; k++) {
        GR.rec[k] = c.hrec.as<double>() + ro;
        ro += 3 * (int64_t)npk[k] * C.n[k];
    }
    const double* Vel = c.vel.as<double>();
    c.stage_begin("friction_gradient");
    if (C.n[0]) hipLaunchKernelGGL(k_fric_gradient_records<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[0], Vel, c.nV, fp, const_cast<double*>(GR.rec[0]));
    if (C.n[1]) hipLaunchKernelGGL(k_fric_gradient_records<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[1], Vel, c.nV, fp, const_cast<double*>(GR.rec[1]));
    if (C.n[2]) hipLaunchKernelGGL(k_fric_gradient_records<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[2], Vel, c.nV, fp, const_cast<double*>(GR.rec[2]));
    if (C.n[3]) hipLaunchKernelGGL(k_fric_gradient_records<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[3], Vel, c.nV, fp, const_cast<double*>(GR.rec[3]));
    const unsigned gw = grid_for((int64_t)c.nV * WAVE, 256, INT32_MAX); // a wave per vertex
    hipLaunchKernelGGL(k_gradient_rows, dim3(gw), dim3(256), 0, s, G, GR, c.inc_ptr.as<uint32_t>(),
                       c.inc_vals.as<uint32_t>(), c.nV, c.grad.as<double>(), row_order(c));
    IPCG_CHECK_LAUNCH();
    c.stage_end();
}

void fric_hessian(Ctx& c, double eps_v, int psd)
{
    require_friction(c);
    if (psd < 0 || psd > 2) throw Error(IPCG_ERR_ARG, "invalid PSD projection method");
    hipStream_t s = c.stream;
    CollView C = view(c);
    const CollRanges G = hess_prepare(c, C);
    const FricRec* R = c.fric.as<FricRec>();
    const FricParams fp { eps_v, psd };
    double* rec = c.hrec.as<double>();
    const double* Vel = c.vel.as<double>();
    c.stage_begin("friction_hessian_records");
    if (C.n[0]) hipLaunchKernelGGL(k_fric_hessian_records<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[0], Vel, c.nV, fp, rec + REC_LEN * G.goff[0]);
    if (C.n[1]) hipLaunchKernelGGL(k_fric_hessian_records<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[1], Vel, c.nV, fp, rec + REC_LEN * G.goff[1]);
    if (C.n[2]) hipLaunchKernelGGL(k_fric_hessian_records<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[2], Vel, c.nV, fp, rec + REC_LEN * G.goff[2]);
    if (C.n[3]) hipLaunchKernelGGL(k_fric_hessian_records<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[3], Vel, c.nV, fp, rec + REC_LEN * G.goff[3]);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    hess_assemble(c);
}


// per-collision anisotropy from host arrays (2 per collision in global
// order; null keeps the current values)
void fric_set_anisotropy(Ctx& c, const double* an, const double* msa, const double* mka)
{
    require_collisions(c);
    if (!c.fric_valid) throw Error(IPCG_ERR_STATE, "no friction set");
    int64_t n = 0;
    for (int k = 0; k < 4; k++) n += c.ncoll[k];
    if (n == 0) return;
    hipStream_t s = c.stream;
    c.np_tmp.reserve(sizeof(double) * 6 * (size_t)n);
    double* d = c.np_tmp.as<double>();
    const double* src[3] = { an, msa, mka };
    const double* dev[3] = { nullptr, nullptr, nullptr };
    for (int k = 0; k < 3; k++)
        if (src[k]) {
            IPCG_HIP(hipMemcpyAsync(d + 2 * k * n, src[k], sizeof(double) * 2 * n, hipMemcpyHostToDevice, s));
            dev[k] = d + 2 * k * n;
        }
    hipLaunchKernelGGL(k_fric_set_aniso, dim3(grid_for(n, 256, INT32_MAX)), dim3(256), 0, s,
                       c.fric.as<FricRec>(), n, dev[0], dev[1], dev[2]);
    IPCG_CHECK_LAUNCH();
    IPCG_HIP(hipStreamSynchronize(s)); // the host arrays may go away
}

void fric_update_lagged_mu(Ctx& c)
{
    require_friction(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    int64_t M;
    const CollRanges G = ranges(C, M);
    FricRec* R = c.fric.as<FricRec>();
    const double* Vel = c.vel.as<double>();
    if (C.n[0]) hipLaunchKernelGGL(k_fric_update_lagged<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[0], Vel, c.nV);
    if (C.n[1]) hipLaunchKernelGGL(k_fric_update_lagged<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[1], Vel, c.nV);
    if (C.n[2]) hipLaunchKernelGGL(k_fric_update_lagged<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[2], Vel, c.nV);
    if (C.n[3]) hipLaunchKernelGGL(k_fric_update_lagged<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, R + G.goff[3], Vel, c.nV);
    IPCG_CHECK_LAUNCH();
    IPCG_HIP(hipStreamSynchronize(s));
}

} // namespace ipcg

