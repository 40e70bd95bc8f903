// Candidate-set helpers over the current candidate lists
// (candidates.cpp:294-408):
//  * per-vertex safe distances: for every vertex, the smallest
//    sqrt(distance) - min_distance over the candidates it belongs to, at
//    most inflation_radius - min_distance;
//  * the noncandidate conservative step: 0.5 dhat over the largest
//    displacement |x1 - x0| of a vertex that belongs to some candidate.
// One lane per candidate (edge-edge, face-vertex, then the codim lists).
// The per-vertex minimum is taken over squared distances as 64-bit atomic
// minima of their bit patterns (non-negative doubles order like their
// bits); sqrt and the subtraction are monotone, so the minimum of
// sqrt(d2) - m equals sqrt(min d2) - m and the result is the reference's,
// whatever order the candidates arrive in.
#include "context.hpp"
#include "geometry.hpp"

#include <cmath>

namespace ipcg {

namespace {

constexpr int CH_BLOCK = 256;

struct CandLists {
    const int2* l[4]; // ee (edge, edge), fv (face, vertex), vv, ev (edge, vertex)
    int64_t off[5];
    const int32_t* E;
    const int32_t* F;
};

// candidate index -> list and vertex ids (CollisionStencil::vertex_ids)
__device__ __forceinline__ int cand_vertices(const CandLists& L, int64_t i, int* v, int& kind)
{
    kind = i < L.off[1] ? 0 : i < L.off[2] ? 1 : i < L.off[3] ? 2 : 3;
    // list and offset by selects: a runtime index into the argument arrays
    // would copy them to scratch memory
    const int2* lst = kind == 0 ? L.l[0] : kind == 1 ? L.l[1] : kind == 2 ? L.l[2] : L.l[3];
    const int64_t o = kind == 0 ? L.off[0] : kind == 1 ? L.off[1] : kind == 2 ? L.off[2] : L.off[3];
    const int2 p = lst[i - o];
    // every path sets all four entries (unused ones to -1), so the array
    // stays in registers
    int v0, v1, v2 = -1, v3 = -1, n;
    switch (kind) {
    case 0:
        v0 = L.E[2 * (int64_t)p.x];
        v1 = L.E[2 * (int64_t)p.x + 1];
        v2 = L.E[2 * (int64_t)p.y];
        v3 = L.E[2 * (int64_t)p.y + 1];
        n = 4;
        break;
    case 1:
        v0 = p.y;
        v1 = L.F[3 * (int64_t)p.x];
        v2 = L.F[3 * (int64_t)p.x + 1];
        v3 = L.F[3 * (int64_t)p.x + 2];
        n = 4;
        break;
    case 2:
        v0 = p.x;
        v1 = p.y;
        n = 2;
        break;
    default:
        v0 = p.y;
        v1 = L.E[2 * (int64_t)p.x];
        v2 = L.E[2 * (int64_t)p.x + 1];
        n = 3;
    }
    v[0] = v0;
    v[1] = v1;
    v[2] = v2;
    v[3] = v3;
    return n;
}

__global__ __launch_bounds__(CH_BLOCK) void k_fill_u64(unsigned long long* a, int64_t n,
                                                       unsigned long long v)
{
    const int64_t i = (int64_t)blockIdx.x * CH_BLOCK + threadIdx.x;
    if (i < n) a[i] = v;
}

// CollisionStencil::compute_distance with the type classified (AUTO)
__global__ __launch_bounds__(CH_BLOCK) void k_safe_d2(CandLists L, const double* X, int nV,
                                                      unsigned long long* best)
{
    const int64_t i = (int64_t)xcd_block() * CH_BLOCK + threadIdx.x;
    if (i >= L.off[4]) return;
    int v[4], kind;
    const int nv = cand_vertices(L, i, v, kind);
    const d3 a = ldx(X, nV, v[0]), b = ldx(X, nV, v[1]);
    double d2;
    if (kind == 0) {
        const d3 c = ldx(X, nV, v[2]), d = ldx(X, nV, v[3]);
        d2 = ee_dist2(a, b, c, d, edge_edge_type(a, b, c, d));
    } else if (kind == 1) {
        const d3 c = ldx(X, nV, v[2]), d = ldx(X, nV, v[3]);
        d2 = pt_dist2(a, b, c, d, point_triangle_type(a, b, c, d));
    } else if (kind == 2) {
        d2 = pp_dist2(a, b);
    } else {
        const d3 c = ldx(X, nV, v[2]);
        d2 = pe_dist2(a, b, c, point_edge_type(a, b, c));
    }
    // NaN (degenerate) never lowers a minimum: its bits exceed +inf's.  The
    // stored minimum only decreases, so a key not below a value read from it
    // cannot lower it: the atomic is issued only for keys below the read.
    const unsigned long long key = (unsigned long long)__double_as_longlong(d2 + 0.0);
#pragma unroll
    for (int k = 0; k < 4; k++) // compile-time indices keep v in registers
        if (k < nv
            && key < __hip_atomic_load(best + v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMin(best + v[k], key);
}

__global__ __launch_bounds__(CH_BLOCK) void k_mark(CandLists L, uint32_t* mark)
{
    const int64_t i = (int64_t)xcd_block() * CH_BLOCK + threadIdx.x;
    if (i >= L.off[4]) return;
    int v[4], kind;
    const int nv = cand_vertices(L, i, v, kind);
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k < nv && !mark[v[k]]) mark[v[k]] = 1u; // most vertices are marked many times
}

// largest squared displacement over marked vertices (block partials)
__global__ __launch_bounds__(CH_BLOCK) void k_disp2_max(const double* X0, const double* X1,
                                                        int nV, const uint32_t* mark,
                                                        double* partial)
{
    double m = 0.0;
    for (int i = blockIdx.x * CH_BLOCK + threadIdx.x; i < nV; i += gridDim.x * CH_BLOCK) {
        if (!mark[i]) continue;
        const d3 d = sub(ldx(X1, nV, i), ldx(X0, nV, i));
        const double n2 = (d.x * d.x + d.y * d.y) + d.z * d.z;
        m = fmax(m, n2);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, WAVE));
    __shared__ double sh[CH_BLOCK / WAVE];
    if (threadIdx.x % WAVE == 0) sh[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int j = 1; j < CH_BLOCK / WAVE; j++) t = fmax(t, sh[j]);
        partial[blockIdx.x] = t;
    }
}

__global__ void k_max_final(const double* partial, int nb, double* out)
{
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int j = 0; j < nb; j++) t = fmax(t, partial[j]);
        *out = t;
    }
}

CandLists lists(Ctx& c)
{
    if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set (run a broad phase)");
    CandLists L;
    L.l[0] = c.cand_ee.as<int2>();
    L.l[1] = c.cand_fv.as<int2>();
    L.l[2] = c.cand_vv.as<int2>();
    L.l[3] = c.cand_ev.as<int2>();
    const int64_t n[4] = { c.n_ee, c.n_fv, c.n_vv, c.n_ev };
    L.off[0] = 0;
    for (int k = 0; k < 4; k++) L.off[k + 1] = L.off[k] + n[k];
    L.E = c.E.as<int32_t>();
    L.F = c.F.as<int32_t>();
    return L;
}

} // namespace

// candidates.cpp:372-408; out: nV doubles on the host
void cand_safe_distances(Ctx& c, double inflation_radius, double min_distance, double* out)
{
    c.require_mesh();
    c.require_pos(0);
    const CandLists L = lists(c);
    hipStream_t s = c.stream;
    const int64_t n = L.off[4];
    if (c.nV == 0) return;
    c.blk_keys.reserve(sizeof(unsigned long long) * (size_t)c.nV);
    unsigned long long* best = c.blk_keys.as<unsigned long long>();
    c.stage_begin("safe_distances");
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(c.nV, CH_BLOCK)), dim3(CH_BLOCK), 0, s, best,
                       (int64_t)c.nV, 0x7FF0000000000000ull); // +inf
    if (n > 0)
        hipLaunchKernelGGL(k_safe_d2, dim3(grid_for(n, CH_BLOCK, INT32_MAX)), dim3(CH_BLOCK), 0, s,
                           L, c.X[0].as<double>(), c.nV, best);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    IPCG_HIP(hipMemcpyAsync(out, best, sizeof(double) * (size_t)c.nV, hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    const double init = inflation_radius - min_distance;
    for (int i = 0; i < c.nV; i++) { // out holds min d2 per vertex (+inf: none)
        const double d = std::sqrt(out[i]) - min_distance;
        out[i] = d < init ? d : init;
    }
}

// candidates.cpp:294-333 with displacements x1 - x0 (slots 1 and 0)
double cand_noncandidate_stepsize(Ctx& c, double dhat)
{
    c.require_mesh();
    c.require_pos(0);
    c.require_pos(1);
    const CandLists L = lists(c);
    const int64_t n = L.off[4];
    if (n == 0) return 1.0; // no possible collisions
    hipStream_t s = c.stream;
    c.blk_start.reserve(sizeof(uint32_t) * ((size_t)c.nV + 2)); // whole u64 words
    uint32_t* mark = c.blk_start.as<uint32_t>();
    const int nb = (int)grid_for(c.nV, CH_BLOCK, 1024);
    c.red_partial.reserve(sizeof(double) * (size_t)nb);
    c.red_out.reserve(sizeof(double) * 2);
    c.stage_begin("noncandidate_stepsize");
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for((c.nV + 1) / 2, CH_BLOCK)), dim3(CH_BLOCK), 0, s,
                       reinterpret_cast<unsigned long long*>(mark), (int64_t)(c.nV + 1) / 2, 0ull);
    hipLaunchKernelGGL(k_mark, dim3(grid_for(n, CH_BLOCK, INT32_MAX)), dim3(CH_BLOCK), 0, s, L, mark);
    hipLaunchKernelGGL(k_disp2_max, dim3(nb), dim3(CH_BLOCK), 0, s, c.X[0].as<double>(),
                       c.X[1].as<double>(), c.nV, mark, c.red_partial.as<double>());
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, s, c.red_partial.as<double>(), nb,
                       c.red_out.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    double* h = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, c.red_out.p, sizeof(double), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    // the largest norm is the norm of the largest squared norm (sqrt is
    // monotone and correctly rounded)
    return 0.5 * dhat / std::sqrt(h[0]);
}

} // namespace ipcg
