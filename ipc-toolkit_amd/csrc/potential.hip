// Barrier potential over the device collision set: energy (fixed-order
// reduction), gradient (FP64 atomic scatter, RowMajor), and the
// PSD-projected Hessian assembled to CSR (potential.cpp:36-222).
//
// Hessian assembly: every collision keeps its projected Hessian in the
// translation-free basis Q (packed 3K x 3K, K = points - 1).  Its 3x3 blocks
// (a, b) are keyed by (vertex a, vertex b); the keys are radix-sorted and
// run-length encoded into unique blocks, each summed from its contributors
// with a per-entry "some contribution is nonzero" mask that reproduces the
// reference's rule of dropping exact-zero triplets before setFromTriplets
// (local_to_global.hpp:289-301).  One lane per vertex then emits its three
// CSR rows in column order.
#include "context.hpp"
#include "sort.hpp"
#include "stencil.hpp"

namespace ipcg {

namespace {

constexpr int PB = 128;

__device__ __forceinline__ d3 ldx(const double* X, int nV, int i)
{
    return { X[i], X[nV + i], X[2 * (int64_t)nV + i] };
}

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
__global__ __launch_bounds__(PB) void k_energy(CollView C, const double* X, int nV,
                                               BarrierParams bp, double* partial)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    double e = 0;
    if (i < C.n[KIND]) {
        int v[4];
        stencil_ids<KIND>(C, i, v);
        d3 P[4];
#pragma unroll
        for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
        e = stencil_energy(KIND, KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                           KIND == K_EE ? C.eps[i] : 0.0, bp);
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

__global__ __launch_bounds__(256) void k_sum(const double* partial, int64_t n, double* out)
{
    double t = 0;
    for (int64_t i = threadIdx.x; i < n; i += 256) t += partial[i];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, WAVE);
    __shared__ double s[4];
    if (threadIdx.x % WAVE == 0) s[threadIdx.x / WAVE] = t;
    __syncthreads();
    if (threadIdx.x == 0) *out = (s[0] + s[1]) + (s[2] + s[3]);
}

template <int KIND>
__global__ __launch_bounds__(PB) void k_gradient(CollView C, const double* X, int nV,
                                                 BarrierParams bp, double* grad)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)blockIdx.x * PB + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    d3 g[NP];
    stencil_gradient<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                           KIND == K_EE ? C.eps[i] : 0.0, bp, g);
#pragma unroll
    for (int j = 0; j < NP; j++) {
        atomicAdd(grad + 3 * (int64_t)v[j], g[j].x);
        atomicAdd(grad + 3 * (int64_t)v[j] + 1, g[j].y);
        atomicAdd(grad + 3 * (int64_t)v[j] + 2, g[j].z);
    }
}

template <int KIND> struct HqTraits {
    static constexpr int NP = StencilTraits<KIND>::NP;
    static constexpr int N = 3 * (NP - 1);
    static constexpr int LEN = N * (N + 1) / 2;
};

template <int KIND>
__global__ __launch_bounds__(64) void k_hessian_q(CollView C, const double* X, int nV,
                                                  BarrierParams bp, double* hq)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    constexpr int LEN = HqTraits<KIND>::LEN;
    const int64_t i = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    d3 P[4];
#pragma unroll
    for (int j = 0; j < NP; j++) P[j] = ldx(X, nV, v[j]);
    double H[LEN];
    stencil_hessian_q<KIND>(KIND == K_EE ? C.dtype[i] : 0, P, C.w[KIND][i],
                            KIND == K_EE ? C.eps[i] : 0.0, bp, H);
    double* o = hq + (int64_t)LEN * i;
#pragma unroll
    for (int j = 0; j < LEN; j++) o[j] = H[j];
}

struct HqLayout {
    int64_t n[4];
    int64_t start[4];   // global collision index of each kind
    int64_t offset[4];  // double offset of each kind in hq
};

template <int KIND>
__global__ __launch_bounds__(256) void k_block_keys(CollView C, HqLayout L, uint64_t* keys,
                                                    uint32_t* vals, int64_t key_off)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= C.n[KIND]) return;
    int v[4];
    stencil_ids<KIND>(C, i, v);
    const int64_t gid = L.start[KIND] + i;
    const int64_t o = key_off + (int64_t)NP * NP * i;
#pragma unroll
    for (int a = 0; a < NP; a++)
#pragma unroll
        for (int b = 0; b < NP; b++) {
            keys[o + a * NP + b] = ((uint64_t)(uint32_t)v[a] << 32) | (uint32_t)v[b];
            vals[o + a * NP + b] = (uint32_t)(gid << 4 | a << 2 | b);
        }
}

template <int NP>
__device__ __forceinline__ void add_block(const double* H, int a, int b, double* acc,
                                          uint32_t& mask)
{
    constexpr int K = NP - 1;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int s = 0; s < 3; s++) {
            double x = 0;
#pragma unroll
            for (int k = 0; k < K; k++)
#pragma unroll
                for (int l = 0; l < K; l++)
                    x += QBasis<NP>::q(a, k) * QBasis<NP>::q(b, l) * H[sp(3 * k + r, 3 * l + s)];
            acc[3 * r + s] += x;
            if (x != 0) mask |= 1u << (3 * r + s);
        }
}

__global__ __launch_bounds__(256) void k_block_values(const uint32_t* vals, const uint32_t* start,
                                                      const uint32_t* nu, HqLayout L,
                                                      const double* hq, double* bval,
                                                      uint32_t* bmask)
{
    const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (u >= (int64_t)*nu) return;
    double acc[9] = { 0, 0, 0, 0, 0, 0, 0, 0, 0 };
    uint32_t mask = 0;
    for (uint32_t j = start[u]; j < start[u + 1]; j++) {
        const uint32_t val = vals[j];
        const int64_t gid = val >> 4;
        const int a = (val >> 2) & 3, b = val & 3;
        if (gid < L.start[1]) {
            add_block<2>(hq + L.offset[0] + 6 * (gid - L.start[0]), a, b, acc, mask);
        } else if (gid < L.start[2]) {
            add_block<3>(hq + L.offset[1] + 21 * (gid - L.start[1]), a, b, acc, mask);
        } else if (gid < L.start[3]) {
            add_block<4>(hq + L.offset[2] + 45 * (gid - L.start[2]), a, b, acc, mask);
        } else {
            add_block<4>(hq + L.offset[3] + 45 * (gid - L.start[3]), a, b, acc, mask);
        }
    }
#pragma unroll
    for (int k = 0; k < 9; k++) bval[9 * u + k] = acc[k];
    bmask[u] = mask;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* a, int64_t n, uint64_t x)
{
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__global__ __launch_bounds__(256) void k_row_counts(const uint64_t* ukeys, const uint32_t* nu,
                                                    const uint32_t* bmask, int nV, uint64_t* cnt)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nV) return;
    const int64_t n = *nu;
    const int64_t b0 = lower_bound_u64(ukeys, n, (uint64_t)v << 32);
    const int64_t b1 = lower_bound_u64(ukeys, n, (uint64_t)(v + 1) << 32);
    uint64_t c[3] = { 0, 0, 0 };
    for (int64_t u = b0; u < b1; u++) {
        const uint32_t m = bmask[u];
        c[0] += __popc(m & 7u);
        c[1] += __popc((m >> 3) & 7u);
        c[2] += __popc((m >> 6) & 7u);
    }
    cnt[3 * (int64_t)v] = c[0];
    cnt[3 * (int64_t)v + 1] = c[1];
    cnt[3 * (int64_t)v + 2] = c[2];
}

__global__ __launch_bounds__(256) void k_fill_rows(const uint64_t* ukeys, const uint32_t* nu,
                                                   const uint32_t* bmask, const double* bval,
                                                   int nV, const int64_t* rowptr, int32_t* col,
                                                   double* val)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nV) return;
    const int64_t n = *nu;
    const int64_t b0 = lower_bound_u64(ukeys, n, (uint64_t)v << 32);
    const int64_t b1 = lower_bound_u64(ukeys, n, (uint64_t)(v + 1) << 32);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        int64_t pos = rowptr[3 * (int64_t)v + k];
        for (int64_t u = b0; u < b1; u++) {
            const uint32_t m = bmask[u];
            const int cb = (int)(ukeys[u] & 0xFFFFFFFFu);
#pragma unroll
            for (int l = 0; l < 3; l++)
                if (m & (1u << (3 * k + l))) {
                    col[pos] = 3 * cb + l;
                    val[pos] = bval[9 * u + 3 * k + l];
                    pos++;
                }
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
    BarrierParams bp { dhat, kappa, c.coll_dmin, 0 };
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

void pot_gradient(Ctx& c, double dhat, double kappa)
{
    require_collisions(c);
    hipStream_t s = c.stream;
    CollView C = view(c);
    BarrierParams bp { dhat, kappa, c.coll_dmin, 0 };
    c.grad.reserve(sizeof(double) * 3 * (size_t)std::max(c.nV, 1));
    double* g = c.grad.as<double>();
    const double* X = c.X[0].as<double>();
    IPCG_HIP(hipMemsetAsync(g, 0, sizeof(double) * 3 * (size_t)c.nV, s));
    c.stage_begin("gradient");
    if (C.n[0]) hipLaunchKernelGGL(k_gradient<K_VV>, dim3(grid_for(C.n[0], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, g);
    if (C.n[1]) hipLaunchKernelGGL(k_gradient<K_EV>, dim3(grid_for(C.n[1], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, g);
    if (C.n[2]) hipLaunchKernelGGL(k_gradient<K_EE>, dim3(grid_for(C.n[2], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, g);
    if (C.n[3]) hipLaunchKernelGGL(k_gradient<K_FV>, dim3(grid_for(C.n[3], PB, INT32_MAX)), dim3(PB), 0, s, C, X, c.nV, bp, g);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
}

void pot_hessian(Ctx& c, double dhat, double kappa, int psd)
{
    require_collisions(c);
    if (psd < 0 || psd > 2) throw Error(IPCG_ERR_ARG, "invalid PSD projection method");
    hipStream_t s = c.stream;
    CollView C = view(c);
    BarrierParams bp { dhat, kappa, c.coll_dmin, psd };
    const double* X = c.X[0].as<double>();
    HqLayout L;
    const int len[4] = { HqTraits<K_VV>::LEN, HqTraits<K_EV>::LEN, HqTraits<K_EE>::LEN,
                         HqTraits<K_FV>::LEN };
    const int np2[4] = { 4, 9, 16, 16 };
    int64_t off = 0, gid = 0, nkeys = 0;
    for (int k = 0; k < 4; k++) {
        L.n[k] = C.n[k];
        L.start[k] = gid;
        L.offset[k] = off;
        gid += C.n[k];
        off += (int64_t)len[k] * C.n[k];
        nkeys += (int64_t)np2[k] * C.n[k];
    }
    if (gid >= (1ll << 28)) throw Error(IPCG_ERR_ARG, "too many collisions for Hessian assembly");
    c.hq.reserve(sizeof(double) * (size_t)std::max<int64_t>(off, 1));
    double* hq = c.hq.as<double>();
    c.stage_begin("hessian_local");
    if (C.n[0]) hipLaunchKernelGGL(k_hessian_q<K_VV>, dim3(grid_for(C.n[0], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, hq + L.offset[0]);
    if (C.n[1]) hipLaunchKernelGGL(k_hessian_q<K_EV>, dim3(grid_for(C.n[1], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, hq + L.offset[1]);
    if (C.n[2]) hipLaunchKernelGGL(k_hessian_q<K_EE>, dim3(grid_for(C.n[2], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, hq + L.offset[2]);
    if (C.n[3]) hipLaunchKernelGGL(k_hessian_q<K_FV>, dim3(grid_for(C.n[3], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, hq + L.offset[3]);
    IPCG_CHECK_LAUNCH();
    c.stage_end();

    c.stage_begin("hessian_assemble");
    const int64_t nk = std::max<int64_t>(nkeys, 1);
    c.blk_keys.reserve(sizeof(uint64_t) * nk);
    c.blk_keys2.reserve(sizeof(uint64_t) * nk);
    c.blk_vals.reserve(sizeof(uint32_t) * nk);
    c.blk_vals2.reserve(sizeof(uint32_t) * nk);
    c.blk_ukeys.reserve(sizeof(uint64_t) * nk);
    c.blk_start.reserve(sizeof(uint32_t) * (nk + 1));
    c.blk_tmp.reserve(std::max(std::max(radix_tmp_bytes(nk), rle_tmp_bytes(nk)),
                               scan_tmp_bytes(3 * (int64_t)c.nV + 1)));
    uint64_t* keys = c.blk_keys.as<uint64_t>();
    uint32_t* vals = c.blk_vals.as<uint32_t>();
    int64_t ko = 0;
    if (C.n[0]) hipLaunchKernelGGL(k_block_keys<K_VV>, dim3(grid_for(C.n[0], 256, INT32_MAX)), dim3(256), 0, s, C, L, keys, vals, ko);
    ko += 4 * C.n[0];
    if (C.n[1]) hipLaunchKernelGGL(k_block_keys<K_EV>, dim3(grid_for(C.n[1], 256, INT32_MAX)), dim3(256), 0, s, C, L, keys, vals, ko);
    ko += 9 * C.n[1];
    if (C.n[2]) hipLaunchKernelGGL(k_block_keys<K_EE>, dim3(grid_for(C.n[2], 256, INT32_MAX)), dim3(256), 0, s, C, L, keys, vals, ko);
    ko += 16 * C.n[2];
    if (C.n[3]) hipLaunchKernelGGL(k_block_keys<K_FV>, dim3(grid_for(C.n[3], 256, INT32_MAX)), dim3(256), 0, s, C, L, keys, vals, ko);
    IPCG_CHECK_LAUNCH();
    int vb = 1;
    while ((1ll << vb) <= c.nV) vb++;
    radix_sort_pairs(keys, vals, c.blk_keys2.as<uint64_t>(), c.blk_vals2.as<uint32_t>(), nkeys,
                     32 + vb, c.blk_tmp.p, s);
    uint32_t* nu_dev = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 7);
    run_length_encode(keys, nkeys, c.blk_ukeys.as<uint64_t>(), c.blk_start.as<uint32_t>(), nu_dev,
                      c.blk_tmp.p, s);
    c.blk_mask.reserve(sizeof(uint32_t) * nk + sizeof(double) * 9 * nk);
    uint32_t* bmask = c.blk_mask.as<uint32_t>();
    double* bval = reinterpret_cast<double*>(c.blk_mask.as<char>() + ((sizeof(uint32_t) * nk + 255) / 256) * 256);
    hipLaunchKernelGGL(k_block_values, dim3(grid_for(nk, 256, INT32_MAX)), dim3(256), 0, s, vals,
                       c.blk_start.as<uint32_t>(), nu_dev, L, hq, bval, bmask);
    const int64_t nrows = 3 * (int64_t)c.nV;
    c.row_cnt.reserve(sizeof(uint64_t) * (nrows + 1));
    c.rowptr.reserve(sizeof(int64_t) * (nrows + 1));
    hipLaunchKernelGGL(k_row_counts, dim3(grid_for(c.nV, 256)), dim3(256), 0, s,
                       c.blk_ukeys.as<uint64_t>(), nu_dev, bmask, c.nV, c.row_cnt.as<uint64_t>());
    uint64_t* rp = reinterpret_cast<uint64_t*>(c.rowptr.p);
    exclusive_scan_u64(c.row_cnt.as<uint64_t>(), rp, nrows, rp + nrows, c.blk_tmp.p, s);
    IPCG_CHECK_LAUNCH();
    int64_t* h = static_cast<int64_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, rp + nrows, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    c.nnz = h[0];
    c.col.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(c.nnz, 1));
    c.val.reserve(sizeof(double) * (size_t)std::max<int64_t>(c.nnz, 1));
    hipLaunchKernelGGL(k_fill_rows, dim3(grid_for(c.nV, 256)), dim3(256), 0, s,
                       c.blk_ukeys.as<uint64_t>(), nu_dev, bmask, bval, c.nV,
                       c.rowptr.as<int64_t>(), c.col.as<int32_t>(), c.val.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    c.hess_valid = true;
}

} // namespace ipcg
