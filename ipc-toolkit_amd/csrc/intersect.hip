// has_intersections (ipc.cpp:105-166, 3D branch): edge-face candidates over
// boxes inflated by 1e-6 x the world bounding-box diagonal, then the
// edge-triangle test of intersection.cpp:116-138 on each.
//
// Device flow: a min/max reduction of the positions (the host finishes the
// diagonal, world_bbox_diagonal_length.hpp:11-15, in the reference's
// operation order), the edge-face broad phase, then k_edge_tri: one lane per
// candidate runs the orient3d filter on both endpoints and, when both
// signs are certified, the rejection test and the LU solve.  Candidates the
// filter cannot decide (coplanar or nearly so) go to a deferred list that
// k_edge_tri_exact finishes with the exact orient3d; that kernel alone
// carries the expansion arrays in scratch memory.
#include "context.hpp"
#include "geometry.hpp"
#include "predicates.hpp"

#include <cmath>

namespace ipcg {

namespace {

constexpr int IS_BLOCK = 256;

__global__ __launch_bounds__(IS_BLOCK) void k_bounds_partial(const double* X, int nV,
                                                             double* partial)
{
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = blockIdx.x * IS_BLOCK + threadIdx.x; i < nV; i += gridDim.x * IS_BLOCK) {
        const d3 p = ldx(X, nV, i);
        lo[0] = fmin(lo[0], p.x);
        lo[1] = fmin(lo[1], p.y);
        lo[2] = fmin(lo[2], p.z);
        hi[0] = fmax(hi[0], p.x);
        hi[1] = fmax(hi[1], p.y);
        hi[2] = fmax(hi[2], p.z);
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, WAVE));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, WAVE));
        }
    __shared__ double sh[IS_BLOCK / WAVE][6];
    const int w = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sh[w][k] = lo[k];
            sh[w][3 + k] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sh[0][threadIdx.x];
        for (int j = 1; j < IS_BLOCK / WAVE; j++)
            v = threadIdx.x < 3 ? fmin(v, sh[j][threadIdx.x]) : fmax(v, sh[j][threadIdx.x]);
        partial[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

__global__ void k_bounds_final(const double* partial, int nb, double* out)
{
    if (threadIdx.x < 6) {
        double v = partial[threadIdx.x];
        for (int j = 1; j < nb; j++)
            v = threadIdx.x < 3 ? fmin(v, partial[j * 6 + threadIdx.x])
                                : fmax(v, partial[j * 6 + threadIdx.x]);
        out[threadIdx.x] = v;
    }
}

struct EdgeTri {
    d3 e0, e1, t0, t1, t2;
};

__device__ __forceinline__ EdgeTri load_edge_tri(const double* X, int nV, const int32_t* E,
                                                 const int32_t* F, int2 c)
{
    EdgeTri q;
    q.e0 = ldx(X, nV, E[2 * (int64_t)c.x]);
    q.e1 = ldx(X, nV, E[2 * (int64_t)c.x + 1]);
    q.t0 = ldx(X, nV, F[3 * (int64_t)c.y]);
    q.t1 = ldx(X, nV, F[3 * (int64_t)c.y + 1]);
    q.t2 = ldx(X, nV, F[3 * (int64_t)c.y + 2]);
    return q;
}

// intersection.cpp:121-126: same nonzero side -> no intersection
__device__ __forceinline__ bool edge_tri_decide(const EdgeTri& q, int o1, int o2)
{
    if (o1 != 0 && o2 != 0 && o1 == o2) return false;
    return edge_tri_solve(q.e0, q.e1, q.t0, q.t1, q.t2);
}

// st[0]: hit flag (int), st[2..3]: deferred count (u64); deferred: indices
__global__ __launch_bounds__(IS_BLOCK) void k_edge_tri(const int2* ef, int64_t n,
                                                       const double* X, int nV,
                                                       const int32_t* E, const int32_t* F,
                                                       int* hit, unsigned long long* n_def,
                                                       uint32_t* deferred)
{
    const int64_t i = (int64_t)xcd_block() * IS_BLOCK + threadIdx.x;
    bool h = false, defer = false;
    if (i < n) {
        const EdgeTri q = load_edge_tri(X, nV, E, F, ef[i]);
        const int o1 = orient3d_filter(q.t0, q.t1, q.t2, q.e0);
        const int o2 = orient3d_filter(q.t0, q.t1, q.t2, q.e1);
        if (o1 == 0 || o2 == 0) defer = true;
        else h = edge_tri_decide(q, o1, o2);
    }
    const unsigned long long slot = wave_append(n_def, defer ? 1u : 0u);
    if (defer) deferred[slot] = (uint32_t)i;
    if (__any(h) && lane_id() == 0) atomicOr(hit, 1);
}

__global__ __launch_bounds__(64) void k_edge_tri_exact(const int2* ef, const uint32_t* deferred,
                                                       const unsigned long long* n_def,
                                                       const double* X, int nV,
                                                       const int32_t* E, const int32_t* F,
                                                       int* hit)
{
    const int64_t n = (int64_t)*n_def;
    for (int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x; j < n; j += (int64_t)gridDim.x * 64) {
        const EdgeTri q = load_edge_tri(X, nV, E, F, ef[deferred[j]]);
        int o1 = orient3d_filter(q.t0, q.t1, q.t2, q.e0);
        if (o1 == 0) o1 = orient3d_exact(q.t0, q.t1, q.t2, q.e0);
        int o2 = orient3d_filter(q.t0, q.t1, q.t2, q.e1);
        if (o2 == 0) o2 = orient3d_exact(q.t0, q.t1, q.t2, q.e1);
        if (edge_tri_decide(q, o1, o2)) atomicOr(hit, 1);
    }
}

} // namespace

bool has_intersections(Ctx& c)
{
    c.require_mesh();
    c.require_pos(0);
    hipStream_t s = c.stream;
    c.stat["intersection.candidates"] = 0;
    c.stat["intersection.deferred"] = 0;
    if (c.nV == 0 || c.nE == 0 || c.nF == 0) return false;
    const double* X = c.X[0].as<double>();
    // world box diagonal: device min/max, host norm in the reference's order
    c.stage_begin("isect_bounds");
    const int nb = (int)grid_for(c.nV, IS_BLOCK, 1024);
    c.aabb_partial.reserve(sizeof(double) * 6 * (size_t)nb);
    c.isect.reserve(256);
    double* bounds = reinterpret_cast<double*>(c.isect.as<char>() + 64);
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(IS_BLOCK), 0, s, X, c.nV,
                       c.aabb_partial.as<double>());
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(64), 0, s, c.aabb_partial.as<double>(), nb,
                       bounds);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    double* hb = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(hb, bounds, sizeof(double) * 6, hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    const double d0 = hb[3] - hb[0], d1 = hb[4] - hb[1], d2 = hb[5] - hb[2];
    const double diag = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    bp_edge_face(c, 1e-6 * diag);
    const int64_t n = c.n_ef;
    c.stat["intersection.candidates"] = n;
    if (n == 0) return false;
    if (n >= (1ll << 32)) throw Error(IPCG_ERR_ARG, "too many edge-face candidates");
    int* hit = reinterpret_cast<int*>(c.isect.as<char>());
    unsigned long long* n_def = reinterpret_cast<unsigned long long*>(c.isect.as<char>() + 8);
    c.ccd_todo.reserve(sizeof(uint32_t) * (size_t)n);
    uint32_t* deferred = c.ccd_todo.as<uint32_t>();
    IPCG_HIP(hipMemsetAsync(c.isect.p, 0, 16, s));
    c.stage_begin("isect_test");
    hipLaunchKernelGGL(k_edge_tri, dim3(grid_for(n, IS_BLOCK, INT32_MAX)), dim3(IS_BLOCK), 0, s,
                       c.cand_ef.as<int2>(), n, X, c.nV, c.E.as<int32_t>(), c.F.as<int32_t>(), hit,
                       n_def, deferred);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    c.stage_begin("isect_exact");
    hipLaunchKernelGGL(k_edge_tri_exact, dim3(256), dim3(64), 0, s, c.cand_ef.as<int2>(), deferred,
                       n_def, X, c.nV, c.E.as<int32_t>(), c.F.as<int32_t>(), hit);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    IPCG_HIP(hipMemcpyAsync(hb, c.isect.p, 16, hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    c.stat["intersection.deferred"] = (int64_t)reinterpret_cast<unsigned long long*>(hb)[1];
    return reinterpret_cast<int*>(hb)[0] != 0;
}

} // namespace ipcg
