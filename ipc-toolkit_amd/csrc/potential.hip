// Barrier potential over the device collision set: energy (fixed-order
// reduction), gradient (FP64 atomic scatter, RowMajor), and the
// PSD-projected Hessian assembled to CSR (potential.cpp:36-222).
//
// Hessian assembly in two phases.  Symbolic (once per collision set): the
// off-diagonal vertex pairs of all stencils are radix-sorted and run-length
// encoded; with one diagonal block per touched vertex they give a block-CSR
// pattern (bptr, sorted bcol).  Numeric: one lane per collision projects its
// Hessian to PSD in the translation-free basis Q (3K x 3K, K = points - 1),
// expands the point-space 3x3 blocks and adds them with FP64 atomics into the
// pattern slot found by binary search in its block row, OR-ing a per-entry
// "some contribution is nonzero" mask that reproduces the reference's rule of
// dropping exact-zero triplets before setFromTriplets
// (local_to_global.hpp:289-301).  One lane per vertex then emits its three
// CSR rows in column order.
#include "context.hpp"
#include "sort.hpp"
#include "stencil.hpp"

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

__device__ __forceinline__ uint32_t find_block(const uint32_t* bptr, const int32_t* bcol, int r,
                                               int c)
{
    uint32_t lo = bptr[r], hi = bptr[r + 1];
    while (lo < hi) {
        const uint32_t m = (lo + hi) >> 1;
        if (bcol[m] < c) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// ---- numeric phase: projected stencil Hessians scattered into blocks ---------
template <int KIND>
__global__ __launch_bounds__(64) void k_hessian_scatter(CollView C, const double* X, int nV,
                                                        BarrierParams bp, const uint32_t* bptr,
                                                        const int32_t* bcol, double* bval,
                                                        uint32_t* bmask)
{
    constexpr int NP = StencilTraits<KIND>::NP;
    constexpr int K = NP - 1;
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
    // Point-space blocks of Q H Q^T (local_hessian_to_global_triplets).  The
    // global matrix is symmetric, so only blocks (row <= col) are summed: the
    // upper triangle of the diagonal blocks and each off-diagonal block once,
    // stored in the row of its smaller vertex; k_mirror_blocks fills the rest.
#pragma unroll
    for (int a = 0; a < NP; a++)
#pragma unroll
        for (int b = a; b < NP; b++) {
            const bool tr = v[a] > v[b]; // the stored block is the transpose
            const int row = tr ? v[b] : v[a], col = tr ? v[a] : v[b];
            const bool diag = row == col;
            const uint32_t blk = find_block(bptr, bcol, row, col);
            if (blk >= bptr[row + 1] || bcol[blk] != col) continue; // not in pattern
            double* dst = bval + 9 * (int64_t)blk;
            uint32_t mask = 0;
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int s = 0; s < 3; s++) {
                    if (diag && s < r) continue;
                    double x = 0;
#pragma unroll
                    for (int k = 0; k < K; k++)
#pragma unroll
                        for (int l = 0; l < K; l++)
                            x += QBasis<NP>::q(a, k) * QBasis<NP>::q(b, l)
                                * H[sp(3 * k + r, 3 * l + s)];
                    if (diag && a != b) { // two stencil points on one vertex
                        double y = 0;
#pragma unroll
                        for (int k = 0; k < K; k++)
#pragma unroll
                            for (int l = 0; l < K; l++)
                                y += QBasis<NP>::q(a, k) * QBasis<NP>::q(b, l)
                                    * H[sp(3 * k + s, 3 * l + r)];
                        x += y;
                    }
                    if (x != 0) { // exact zeros are not triplets (local_to_global.hpp:291)
                        const int e = tr ? 3 * s + r : 3 * r + s;
                        atomicAdd(dst + e, x);
                        mask |= 1u << e;
                        if (diag) mask |= 1u << (3 * s + r);
                    }
                }
            if (mask) atomicOr(bmask + blk, mask);
        }
}

// Lower blocks from the summed upper ones: block (v, c) with c < v is the
// transpose of block (c, v); diagonal blocks copy their upper triangle.
__global__ __launch_bounds__(256) void k_mirror_blocks(const uint32_t* bptr, const int32_t* bcol,
                                                       int nV, double* bval, uint32_t* bmask)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= nV) return;
    for (uint32_t u = bptr[v]; u < bptr[v + 1]; u++) {
        const int cb = bcol[u];
        if (cb > v) break; // columns are sorted
        double* dst = bval + 9 * (int64_t)u;
        if (cb == v) {
#pragma unroll
            for (int r = 0; r < 3; r++)
#pragma unroll
                for (int s = 0; s < r; s++) dst[3 * r + s] = dst[3 * s + r];
            continue;
        }
        const uint32_t w = find_block(bptr, bcol, cb, v);
        if (w >= bptr[cb + 1] || bcol[w] != v) continue; // symmetric pattern: not reached
        const double* src = bval + 9 * (int64_t)w;
        const uint32_t m = bmask[w];
        uint32_t t = 0;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int s = 0; s < 3; s++) {
                dst[3 * r + s] = src[3 * s + r];
                if (m & (1u << (3 * s + r))) t |= 1u << (3 * r + s);
            }
        bmask[u] = t;
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

void pot_hessian(Ctx& c, double dhat, double kappa, int psd)
{
    require_collisions(c);
    if (psd < 0 || psd > 2) throw Error(IPCG_ERR_ARG, "invalid PSD projection method");
    hipStream_t s = c.stream;
    c.stage_begin("hessian_pattern");
    if (!c.pattern_valid) hess_pattern(c);
    c.stage_end();
    CollView C = view(c);
    BarrierParams bp { dhat, kappa, c.coll_dmin, psd };
    const double* X = c.X[0].as<double>();
    const int64_t nb = std::max<int64_t>(c.nblocks, 1);
    c.blk_mask.reserve(sizeof(uint32_t) * nb);
    c.hq.reserve(sizeof(double) * 9 * nb);
    uint32_t* bmask = c.blk_mask.as<uint32_t>();
    double* bval = c.hq.as<double>();
    const uint32_t* bptr = c.hp_bptr.as<uint32_t>();
    const int32_t* bcol = c.hp_bcol.as<int32_t>();
    IPCG_HIP(hipMemsetAsync(bmask, 0, sizeof(uint32_t) * nb, s));
    IPCG_HIP(hipMemsetAsync(bval, 0, sizeof(double) * 9 * nb, s));
    c.stage_begin("hessian_scatter_vv_ev");
    if (C.n[0]) hipLaunchKernelGGL(k_hessian_scatter<K_VV>, dim3(grid_for(C.n[0], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, bptr, bcol, bval, bmask);
    if (C.n[1]) hipLaunchKernelGGL(k_hessian_scatter<K_EV>, dim3(grid_for(C.n[1], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, bptr, bcol, bval, bmask);
    c.stage_end();
    c.stage_begin("hessian_scatter_ee");
    if (C.n[2]) hipLaunchKernelGGL(k_hessian_scatter<K_EE>, dim3(grid_for(C.n[2], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, bptr, bcol, bval, bmask);
    c.stage_end();
    c.stage_begin("hessian_scatter_fv");
    if (C.n[3]) hipLaunchKernelGGL(k_hessian_scatter<K_FV>, dim3(grid_for(C.n[3], 64, INT32_MAX)), dim3(64), 0, s, C, X, c.nV, bp, bptr, bcol, bval, bmask);
    c.stage_end();
    c.stage_begin("hessian_mirror");
    hipLaunchKernelGGL(k_mirror_blocks, dim3(grid_for(c.nV, 256)), dim3(256), 0, s, bptr, bcol,
                       c.nV, bval, bmask);
    IPCG_CHECK_LAUNCH();
    c.stage_end();

    c.stage_begin("hessian_assemble");
    const int64_t nrows = 3 * (int64_t)c.nV;
    c.row_cnt.reserve(sizeof(uint64_t) * (nrows + 1));
    c.rowptr.reserve(sizeof(int64_t) * (nrows + 1));
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

} // namespace ipcg
