// Broad phase on the GPU: conservative boxes, 63-bit Morton codes, LBVH
// (Apetrei 2014 single bottom-up pass) and candidate traversal.
//
// Boxes follow aabb.cpp:35-126 (nextafter(p -/+ r) outward in FP64, unions
// for edges and faces) and lbvh.cpp:29-41 (FP32 with one more outward ulp).
// Morton codes follow lbvh.cpp:148-169 (FP64 box centre mapped by the mesh
// box of the vertex boxes, morton.hpp:53-63) and are sorted stably, so a
// pair's orientation (query = lower sorted position, lbvh.cpp:413-424) is
// the one the CPU restatement produces.  The candidate set itself is every
// leaf pair whose FP32 boxes overlap and passes the filters, independent of
// the tree shape.
//
// Tree layout: internal node j (splitting sorted keys j | j+1) stores both
// child boxes, child links and the children's rightmost leaf positions in
// one 64-byte line, so a traversal step is one line load.
#include "context.hpp"
#include "shard.hpp"
#include "geometry.hpp"
#include "sort.hpp"

#include <utility>

namespace ipcg {

namespace {

constexpr int BP_BLOCK = 256;
constexpr int TR_BLOCK = 128;
constexpr int TR_STACK = 64;  // lbvh.cpp:378 MAX_STACK_SIZE
constexpr int TR_LSTACK = 16; // entries per lane kept in LDS; the rest in HBM
constexpr int TR_BUF = 4;     // overlapping leaves a lane holds before the wave settles them

// AABB::from_point (+ swept union), aabb.cpp:21-55
__device__ __forceinline__ void vbox(const double* X0, const double* X1, int nV, int i, double r,
                                     double* lo, double* hi)
{
    const d3 p = ldx(X0, nV, i);
    lo[0] = nextafter(p.x - r, -INFINITY);
    lo[1] = nextafter(p.y - r, -INFINITY);
    lo[2] = nextafter(p.z - r, -INFINITY);
    hi[0] = nextafter(p.x + r, INFINITY);
    hi[1] = nextafter(p.y + r, INFINITY);
    hi[2] = nextafter(p.z + r, INFINITY);
    if (X1) {
        const d3 q = ldx(X1, nV, i);
        const double l1[3] = { nextafter(q.x - r, -INFINITY), nextafter(q.y - r, -INFINITY),
                               nextafter(q.z - r, -INFINITY) };
        const double h1[3] = { nextafter(q.x + r, INFINITY), nextafter(q.y + r, INFINITY),
                               nextafter(q.z + r, INFINITY) };
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = l1[k] < lo[k] ? l1[k] : lo[k];
            hi[k] = hi[k] < h1[k] ? h1[k] : hi[k];
        }
    }
}

__global__ __launch_bounds__(BP_BLOCK) void k_mesh_aabb_partial(const double* X0,
                                                                const double* X1, int nV,
                                                                double r, double* partial)
{
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = blockIdx.x * BP_BLOCK + threadIdx.x; i < nV; i += gridDim.x * BP_BLOCK) {
        double l[3], h[3];
        vbox(X0, X1, nV, i, r, l, h);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], l[k]);
            hi[k] = fmax(hi[k], h[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, WAVE));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, WAVE));
        }
    __shared__ double s[BP_BLOCK / WAVE][6];
    const int w = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            s[w][k] = lo[k];
            s[w][3 + k] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = s[0][threadIdx.x];
        for (int j = 1; j < BP_BLOCK / WAVE; j++)
            v = threadIdx.x < 3 ? fmin(v, s[j][threadIdx.x]) : fmax(v, s[j][threadIdx.x]);
        partial[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

__global__ void k_mesh_aabb_final(const double* partial, int nb, double* aabb)
{
    if (threadIdx.x < 6) {
        double v = partial[threadIdx.x];
        for (int j = 1; j < nb; j++)
            v = threadIdx.x < 3 ? fmin(v, partial[j * 6 + threadIdx.x])
                                : fmax(v, partial[j * 6 + threadIdx.x]);
        aabb[threadIdx.x] = v;
    }
}

// Batched contexts: the mesh box of each scene's vertex boxes, one
// workgroup per scene (min / max are exact, so the box equals the one a
// context holding the scene alone computes).
// Scene boxes in two steps: workgroup (scene, chunk) reduces every
// SCENE_CHUNKS-th stretch of the scene's vertices into a partial box, then a
// lane per scene combines its partials (min / max are exact, so the box is
// the one the scene alone gets).
constexpr int SCENE_CHUNKS = 16;

__global__ __launch_bounds__(BP_BLOCK) void k_scene_aabb(const double* X0, const double* X1,
                                                         int nV, const int64_t* voff, double r,
                                                         double* part)
{
    const int sc = blockIdx.x / SCENE_CHUNKS, ch = blockIdx.x % SCENE_CHUNKS;
    double* aabb = part + 6 * (int64_t)ch; // this chunk's partial of scene sc
    double lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int64_t i = voff[sc] + (int64_t)ch * BP_BLOCK + threadIdx.x; i < voff[sc + 1];
         i += (int64_t)SCENE_CHUNKS * BP_BLOCK) {
        double l[3], h[3];
        vbox(X0, X1, nV, (int)i, r, l, h);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], l[k]);
            hi[k] = fmax(hi[k], h[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++)
        for (int o = 32; o > 0; o >>= 1) {
            lo[k] = fmin(lo[k], __shfl_xor(lo[k], o, WAVE));
            hi[k] = fmax(hi[k], __shfl_xor(hi[k], o, WAVE));
        }
    __shared__ double sm[BP_BLOCK / WAVE][6];
    const int w = threadIdx.x / WAVE;
    if (threadIdx.x % WAVE == 0)
#pragma unroll
        for (int k = 0; k < 3; k++) {
            sm[w][k] = lo[k];
            sm[w][3 + k] = hi[k];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sm[0][threadIdx.x];
        for (int j = 1; j < BP_BLOCK / WAVE; j++)
            v = threadIdx.x < 3 ? fmin(v, sm[j][threadIdx.x]) : fmax(v, sm[j][threadIdx.x]);
        aabb[6 * (int64_t)SCENE_CHUNKS * sc + threadIdx.x] = v;
    }
}

__global__ void k_scene_aabb_final(const double* part, int n_scenes, double* aabb)
{
    const int sc = blockIdx.x * 64 + threadIdx.x;
    if (sc >= n_scenes) return;
    const double* p = part + 6 * (int64_t)SCENE_CHUNKS * sc;
    for (int k = 0; k < 6; k++) {
        double v = p[k];
        for (int c = 1; c < SCENE_CHUNKS; c++)
            v = k < 3 ? fmin(v, p[6 * c + k]) : fmax(v, p[6 * c + k]);
        aabb[6 * (int64_t)sc + k] = v;
    }
}

__device__ __forceinline__ uint64_t expand_bits_2(uint64_t v)
{
    v = (v | v << 32) & 0x1F00000000FFFFull;
    v = (v | v << 16) & 0x1F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__device__ __forceinline__ uint64_t morton3(double x, double y, double z)
{
    const double scale = 2097152.0, top = scale - 1;
    x = x * scale;
    y = y * scale;
    z = z * scale;
    x = x < 0.0 ? 0.0 : (top < x ? top : x);
    y = y < 0.0 ? 0.0 : (top < y ? top : y);
    z = z < 0.0 ? 0.0 : (top < z ? top : z);
    return (expand_bits_2((uint64_t)x) << 2) | (expand_bits_2((uint64_t)y) << 1)
        | expand_bits_2((uint64_t)z);
}

// per primitive (ARITY vertices): FP64 box, Morton code, FP32 box
// first vertex of primitive p (I: vertex ids per primitive; arity 1: a
// vertex subset, or null for all vertices)
template <int ARITY>
__device__ __forceinline__ int first_vertex(const int32_t* I, int p)
{
    return (ARITY == 1 && !I) ? p : I[(int64_t)ARITY * p];
}

// vscene (batched contexts): scene per vertex, each primitive mapped by its
// own scene's box; mort then receives a copy of the Morton codes
template <int ARITY>
__global__ __launch_bounds__(BP_BLOCK) void k_prim_boxes(
    const double* X0, const double* X1, int nV, const int32_t* I, int n, double r,
    const double* aabb, const int32_t* vscene, float* box32, uint64_t* keys, uint32_t* ids,
    uint64_t* mort)
{
    const int p = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (p >= n) return;
    double lo[3], hi[3];
    const int v0 = first_vertex<ARITY>(I, p);
    if (vscene) aabb += 6 * (int64_t)vscene[v0];
    vbox(X0, X1, nV, v0, r, lo, hi);
#pragma unroll
    for (int j = 1; j < ARITY; j++) {
        double l[3], h[3];
        vbox(X0, X1, nV, I[(int64_t)ARITY * p + j], r, l, h);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = l[k] < lo[k] ? l[k] : lo[k];
            hi[k] = hi[k] < h[k] ? h[k] : hi[k];
        }
    }
    double m[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double c = 0.5 * (lo[k] + hi[k]);
        m[k] = (c - aabb[k]) / (aabb[3 + k] - aabb[k]);
    }
    const uint64_t code = morton3(m[0], m[1], m[2]);
    keys[p] = code;
    if (mort) mort[p] = code;
    ids[p] = (uint32_t)p;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        box32[(int64_t)6 * p + k] = nextafterf(__double2float_rn(lo[k]), -INFINITY);
        box32[(int64_t)6 * p + 3 + k] = nextafterf(__double2float_rn(hi[k]), INFINITY);
    }
}

// Forest of a batched context: after the Morton sort, the scene of each
// sorted primitive becomes the key of a second stable sort (primitives of
// one scene keep their Morton order); then the Morton codes are gathered
// back in the final order and the scene of every leaf is kept.
template <int ARITY>
__global__ __launch_bounds__(BP_BLOCK) void k_scene_keys(const uint32_t* ids, const int32_t* I,
                                                         const int32_t* vscene, int n,
                                                         uint64_t* keys)
{
    const int i = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (i < n) keys[i] = (uint64_t)vscene[first_vertex<ARITY>(I, (int)ids[i])];
}

__global__ __launch_bounds__(BP_BLOCK) void k_forest_keys(const uint32_t* ids, const uint64_t* mort,
                                                          int n, uint64_t* keys, int32_t* lscene)
{
    const int i = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (i >= n) return;
    lscene[i] = (int32_t)keys[i];
    keys[i] = mort[ids[i]];
}

// Root entry per scene before the refit: none, a single leaf, or to be
// completed (loff: the scene's first sorted leaf position, n_scenes + 1).
__global__ void k_forest_roots(const int64_t* loff, int n_scenes, int32_t* root)
{
    const int sc = blockIdx.x * 256 + threadIdx.x;
    if (sc >= n_scenes) return;
    const int64_t m = loff[sc + 1] - loff[sc];
    root[sc] = m == 0 ? ROOT_NONE : (m == 1 ? ~(int32_t)loff[sc] : ROOT_UNFINISHED);
}

// Common-prefix length of the sorted keys i and j extended by their
// positions, so equal Morton codes still order strictly; -1 outside [0, n)
// and, in a forest, across scenes (lscene: scene per sorted leaf).
__device__ __forceinline__ int delta(const uint64_t* codes, const int32_t* lscene, int n, int i,
                                     int j)
{
    if (j < 0 || j >= n) return -1;
    if (lscene && lscene[i] != lscene[j]) return -1;
    const uint64_t ci = codes[i], cj = codes[j];
    if (ci == cj) return 64 + __clz((unsigned)(i ^ j));
    return __clzll(ci ^ cj);
}

// Hierarchy of the binary radix tree over the sorted keys (Karras 2012):
// internal node i covers a key range that has i as one end and splits it at
// the highest differing bit.  Every node is placed independently, so the
// build needs no inter-thread communication.  The tree shape differs from
// the reference's Apetrei construction (lbvh.cpp:197-302); the candidate set
// does not depend on it.
//
// In a forest (lscene non-null) the prefix is -1 across scenes, so every
// scene's leaves form a tree of their own: internal node i exists when
// leaves i and i + 1 belong to one scene, and the node at a scene's first
// leaf is its root.

// Bottom-up box refit, one level of completed nodes per launch: leaf boxes
// are written into their parents' slots by k_karras, which also starts the
// first work list; then a completed node's box goes into its slot of the
// parent, and the second child to arrive puts the parent on the next work
// list.  Boxes written in one launch are read in the next, so no
// inter-workgroup fences are needed.  Appends to the next list are gathered
// per workgroup (one global atomic per workgroup).
constexpr int RF_BLOCK = 1024;

// Writes the child's box into parent p; returns whether p is now complete.
__device__ __forceinline__ bool refit_child(Node* nodes, int32_t p, int32_t self,
                                            const float* lo, const float* hi, int32_t* visits)
{
    Node* P = nodes + p;
    if (P->c0 == self) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            P->lo0[k] = lo[k];
            P->hi0[k] = hi[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            P->lo1[k] = lo[k];
            P->hi1[k] = hi[k];
        }
    }
    return atomicAdd(visits + p, 1) == 1;
}

// Appends p (when push) to the next work list; every thread of the block
// must call it.
__device__ __forceinline__ void refit_append(bool push, int32_t p, int32_t* next,
                                             uint32_t* next_n)
{
    __shared__ uint32_t n_blk, base;
    if (threadIdx.x == 0) n_blk = 0;
    __syncthreads();
    const uint32_t slot = push ? atomicAdd(&n_blk, 1u) : 0u;
    __syncthreads();
    if (threadIdx.x == 0 && n_blk) base = atomicAdd(next_n, n_blk);
    __syncthreads();
    if (push) next[base + slot] = p;
}

// Also fills the box slot of every leaf child from the leaf's primitive box
// and counts its leaf children in visits[i]; a node whose two children are
// leaves is complete at once and goes on the first refit work list.
__global__ __launch_bounds__(BP_BLOCK) void k_karras(const uint64_t* codes, const int32_t* lscene,
                                                     int n, Node* nodes, int32_t* parent,
                                                     const uint32_t* ids, const float* box32,
                                                     int32_t* visits, int32_t* work,
                                                     uint32_t* work_n)
{
    const int i = blockIdx.x * BP_BLOCK + threadIdx.x;
    bool push = false;
    // no node between scenes
    if (i < n - 1 && !(lscene && lscene[i] != lscene[i + 1])) {
        const int d = delta(codes, lscene, n, i, i + 1) - delta(codes, lscene, n, i, i - 1) >= 0 ? 1 : -1;
        const int dmin = delta(codes, lscene, n, i, i - d);
        int lmax = 2;
        while (delta(codes, lscene, n, i, i + lmax * d) > dmin) lmax <<= 1;
        int l = 0;
        for (int t = lmax >> 1; t >= 1; t >>= 1)
            if (delta(codes, lscene, n, i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(codes, lscene, n, i, j);
        int sp = 0;
        for (int t = (l + 1) >> 1;; t = (t + 1) >> 1) {
            if (delta(codes, lscene, n, i, i + (sp + t) * d) > dnode) sp += t;
            if (t == 1) break;
        }
        const int g = i + sp * d + (d < 0 ? -1 : 0);
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        Node* N = nodes + i;
        int leaves = 0;
        if (lo == g) {
            N->c0 = ~g;
            const uint32_t pid = ids[g];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                N->lo0[k] = box32[(int64_t)6 * pid + k];
                N->hi0[k] = box32[(int64_t)6 * pid + 3 + k];
            }
            leaves++;
        } else {
            N->c0 = g;
            parent[g] = i;
        }
        if (hi == g + 1) {
            N->c1 = ~(g + 1);
            const uint32_t pid = ids[g + 1];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                N->lo1[k] = box32[(int64_t)6 * pid + k];
                N->hi1[k] = box32[(int64_t)6 * pid + 3 + k];
            }
            leaves++;
        } else {
            N->c1 = g + 1;
            parent[g + 1] = i;
        }
        N->r0 = g;
        N->r1 = hi;
        visits[i] = leaves;
        push = leaves == 2;
        if (lscene ? (i == 0 || lscene[i - 1] != lscene[i]) : i == 0) parent[i] = -1; // a root
    }
    refit_append(push, i, work, work_n);
}

// The last levels, once the work list holds at most RF_TOP nodes (128 KB
// of LDS for the two lists): one
// workgroup completes them all, level after level with a barrier between
// levels (a workgroup's global writes are visible to it after the barrier),
// instead of one launch per level.
constexpr int RF_TOP = 16384;

__global__ __launch_bounds__(RF_BLOCK) void k_refit_top(const int32_t* cur, const uint32_t* cur_n,
                                                        const int32_t* parent,
                                                        const int32_t* lscene, Node* nodes,
                                                        int32_t* visits, int32_t* root,
                                                        int64_t max_levels)
{
    __shared__ int32_t list[2][RF_TOP];
    __shared__ uint32_t cnt[2];
    const int t = threadIdx.x;
    if (t == 0) {
        const uint32_t n0 = *cur_n;
        cnt[0] = n0 < (uint32_t)RF_TOP ? n0 : (uint32_t)RF_TOP; // the host checked n0 <= RF_TOP
        cnt[1] = 0;
    }
    __syncthreads();
    for (uint32_t i = t; i < cnt[0]; i += RF_BLOCK) list[0][i] = cur[i];
    __syncthreads();
    // each level completes at least one of the max_levels internal nodes
    int c = 0;
    for (int64_t level = 0; level < max_levels; c ^= 1, level++) {
        const uint32_t m = cnt[c];
        if (m == 0) break;
        for (uint32_t i = t; i < m; i += RF_BLOCK) {
            const int32_t j = list[c][i];
            const Node* N = nodes + j;
            float lo[3], hi[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(N->lo0[k], N->lo1[k]);
                hi[k] = fmaxf(N->hi0[k], N->hi1[k]);
            }
            const int32_t p = parent[j];
            if (p < 0) root[lscene ? lscene[j] : 0] = j; // the tree is complete
            else if (refit_child(nodes, p, j, lo, hi, visits))
                list[c ^ 1][atomicAdd(&cnt[c ^ 1], 1u)] = p;
        }
        __syncthreads();
        if (t == 0) cnt[c] = 0;
        __syncthreads();
    }
}


// One level: the workgroups stride over the level's list (a grid of at
// most RF_GRID workgroups, whatever the list's bound) and clear the counter
// of the level after next (clear_n: read as input by the previous level,
// appended to by the next one, so stream order makes the clear safe).
constexpr unsigned RF_GRID = 512;

__global__ __launch_bounds__(RF_BLOCK) void k_refit_level(const int32_t* cur,
                                                          const uint32_t* cur_n,
                                                          const int32_t* parent,
                                                          const int32_t* lscene, Node* nodes,
                                                          int32_t* visits, int32_t* next,
                                                          uint32_t* next_n, uint32_t* clear_n,
                                                          int32_t* root)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_n = 0;
    const uint32_t m = *cur_n;
    for (uint32_t base = blockIdx.x * RF_BLOCK; base < m; base += gridDim.x * RF_BLOCK) {
        const uint32_t t = base + threadIdx.x; // the loop count is uniform per workgroup
        bool push = false;
        int32_t p = 0;
        if (t < m) {
            const int32_t j = cur[t];
            const Node* N = nodes + j;
            float lo[3], hi[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                lo[k] = fminf(N->lo0[k], N->lo1[k]);
                hi[k] = fmaxf(N->hi0[k], N->hi1[k]);
            }
            p = parent[j];
            if (p < 0) root[lscene ? lscene[j] : 0] = j; // the tree is complete
            else push = refit_child(nodes, p, j, lo, hi, visits);
        }
        refit_append(push, p, next, next_n);
    }
}

__device__ __forceinline__ bool overlap(const float* q, const float* lo, const float* hi)
{
    return q[0] <= hi[0] && q[1] <= hi[1] && q[2] <= hi[2] && lo[0] <= q[3] && lo[1] <= q[4]
        && lo[2] <= q[5];
}

// CollisionMesh::can_collide as data: the intersection of the vertex
// patches filter (patch[i] != patch[j]), the static obstacle filter (i or j
// below n_dynamic) and the codim cross filter (exactly one of i, j below
// n_codim), each optional (collision_filter.hpp:113-143).
struct Filter {
    const int32_t* E;
    const int32_t* F;
    const int32_t* patch; // may be null
    int32_t n_dynamic;    // < 0: no static obstacle filter
    int32_t n_codim;      // < 0: no codim cross filter
    __device__ __forceinline__ bool any() const
    {
        return patch != nullptr || n_dynamic >= 0 || n_codim >= 0;
    }
};

__device__ __forceinline__ bool can_vv(const Filter& f, int a, int b)
{
    return (f.patch == nullptr || f.patch[a] != f.patch[b])
        && (f.n_dynamic < 0 || a < f.n_dynamic || b < f.n_dynamic)
        && (f.n_codim < 0 || ((a < f.n_codim) != (b < f.n_codim)));
}

// lbvh.cpp:826-849 can_edges_collide
__device__ __forceinline__ bool can_ee(const Filter& f, int ea, int eb)
{
    const int a0 = f.E[2 * (int64_t)ea], a1 = f.E[2 * (int64_t)ea + 1];
    const int b0 = f.E[2 * (int64_t)eb], b1 = f.E[2 * (int64_t)eb + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    return can_vv(f, a0, b0) || can_vv(f, a0, b1) || can_vv(f, a1, b0) || can_vv(f, a1, b1);
}

// lbvh.cpp:800-807 can_edge_vertex_collide
__device__ __forceinline__ bool can_ev(const Filter& f, int ei, int vi)
{
    const int e0 = f.E[2 * (int64_t)ei], e1 = f.E[2 * (int64_t)ei + 1];
    if (vi == e0 || vi == e1) return false;
    return can_vv(f, vi, e0) || can_vv(f, vi, e1);
}

// lbvh.cpp:851-858 can_face_vertex_collide
__device__ __forceinline__ bool can_fv(const Filter& f, int fi, int vi)
{
    const int f0 = f.F[3 * (int64_t)fi], f1 = f.F[3 * (int64_t)fi + 1],
              f2 = f.F[3 * (int64_t)fi + 2];
    if (vi == f0 || vi == f1 || vi == f2) return false;
    return can_vv(f, vi, f0) || can_vv(f, vi, f1) || can_vv(f, vi, f2);
}

// One lane per query leaf (in the query set's Morton order).  MODE:
//  TR_EE  edge-edge self traversal: (query, other) for other at a later
//         sorted position;
//  TR_FV  vertex queries against the face tree: (face, vertex);
//  TR_VV  codim vertices, self traversal: (query, other) in mesh ids;
//  TR_EV  codim vertex queries against the codim-edge tree: (edge, vertex)
//         in mesh ids (candidates.cpp:65-118);
//  TR_EF  edge queries against the face tree: (edge, face)
//         (lbvh.cpp:771-784 reports the same set from face queries).
enum { TR_EE = 0, TR_FV = 1, TR_VV = 2, TR_EV = 3, TR_EF = 4 };

// The query primitive: its output id and vertices (-1 padded).
struct Query {
    int id;
    int v[3];
};

// some vertex of a may collide with some vertex of b
__device__ __forceinline__ bool patch_ok(const Filter& f, const int* a, int na, const int* b,
                                         int nb)
{
    if (!f.any()) return true;
    for (int i = 0; i < na; i++)
        for (int j = 0; j < nb; j++)
            if (can_vv(f, a[i], b[j])) return true;
    return false;
}

// lbvh.cpp:800-858 filters and the reported pair for a leaf record
template <int MODE>
__device__ __forceinline__ int2 leaf_pair(const Filter& flt, const Query& q, const int4& r)
{
    if constexpr (MODE == TR_EE) { // edge vs edge: no shared vertex
        if (q.v[0] == r.y || q.v[0] == r.z || q.v[1] == r.y || q.v[1] == r.z) return make_int2(-1, -1);
        const int b[2] = { r.y, r.z };
        return patch_ok(flt, q.v, 2, b, 2) ? make_int2(q.id, r.x) : make_int2(-1, -1);
    } else if constexpr (MODE == TR_FV) { // vertex vs face
        if (q.v[0] == r.y || q.v[0] == r.z || q.v[0] == r.w) return make_int2(-1, -1);
        const int b[3] = { r.y, r.z, r.w };
        return patch_ok(flt, q.v, 1, b, 3) ? make_int2(r.x, q.id) : make_int2(-1, -1);
    } else if constexpr (MODE == TR_VV) { // codim vertex vs codim vertex
        const int b[1] = { r.y };
        return patch_ok(flt, q.v, 1, b, 1) ? make_int2(q.id, r.x) : make_int2(-1, -1);
    } else if constexpr (MODE == TR_EF) { // edge vs face (lbvh.cpp:836-849)
        if (q.v[0] == r.y || q.v[0] == r.z || q.v[0] == r.w || q.v[1] == r.y || q.v[1] == r.z
            || q.v[1] == r.w)
            return make_int2(-1, -1);
        const int b[3] = { r.y, r.z, r.w };
        return patch_ok(flt, q.v, 2, b, 3) ? make_int2(q.id, r.x) : make_int2(-1, -1);
    } else { // codim vertex vs codim edge
        if (q.v[0] == r.y || q.v[0] == r.z) return make_int2(-1, -1);
        const int b[2] = { r.y, r.z };
        return patch_ok(flt, q.v, 1, b, 2) ? make_int2(r.x, q.id) : make_int2(-1, -1);
    }
}

// SWEPT only names the launches of the swept (CCD) broad phase apart from
// the static ones in kernel profiles; the code is the same.
// Forest of a batched context (vscene non-null): every query walks the tree
// of its own scene (root_p: one entry per scene; queries of scenes that
// `active` leaves out are skipped).
struct ForestArgs {
    const int32_t* vscene;   // scene per vertex (null: one tree, root_p[0])
    const uint8_t* active;   // per scene (null: all)
    const uint32_t* tids;    // target tree: primitive per sorted leaf
};

// BLOCK: workgroup size (64, 128 or 256), chosen per launch by the
// context's autotuner or TR_BLOCK.
template <int MODE, int SWEPT, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_traverse(
    const float* qbox, const uint32_t* qids, int nq, const Node* nodes, const int32_t* root_p,
    const int4* leaf, const float* tbox, int nleaves, Filter flt, const int32_t* qmap, int2* out,
    Shards S, int* deep, int* status, ForestArgs fa)
{
    constexpr bool TRI = MODE == TR_EE || MODE == TR_VV;
    // The traversal stack of postponed right children: the first TR_LSTACK
    // entries in LDS (small enough for full occupancy), deeper ones in the
    // query's own HBM slots (deep: TR_STACK - TR_LSTACK ints per query).
    __shared__ int stack[TR_LSTACK][BLOCK];
    const int t = threadIdx.x;
    const int qpos = (int)xcd_block() * BLOCK + t;
    bool active = qpos < nq;
    float q[6] = { 0, 0, 0, 0, 0, 0 };
    Query qr { 0, { -1, -1, -1 } };
    if (active) {
        const int qprim = (int)qids[qpos];
#pragma unroll
        for (int k = 0; k < 6; k++) q[k] = qbox[(int64_t)6 * qprim + k];
        if (MODE == TR_EE || MODE == TR_EF) { // query edge: id, endpoints
            qr.id = qprim;
            qr.v[0] = flt.E[2 * (int64_t)qprim];
            qr.v[1] = flt.E[2 * (int64_t)qprim + 1];
        } else { // query vertex (codim queries map their subset index)
            qr.id = qmap ? qmap[qprim] : qprim;
            qr.v[0] = qr.id;
        }
    }
    const int nin = nleaves - 1; // internal nodes
    const int sh = wave_shard();
    // Overlapping leaves wait in registers (their positions, up to TR_BUF
    // per lane) and the wave settles them together when a lane could
    // overflow or the walks end: the leaf records are loaded for all of them
    // at once (one memory round trip per flush instead of one per iteration
    // with a hit), the filters drop pairs, and one append per wave writes
    // the rest.
    int buf[TR_BUF];
#pragma unroll
    for (int k = 0; k < TR_BUF; k++) buf[k] = -1;
    unsigned nb = 0;
    int2* const reg = out + sh * S.scap;
    auto flush = [&]() {
        int4 rec[TR_BUF]; // every load first, then the filters
#pragma unroll
        for (int k = 0; k < TR_BUF; k++) {
            rec[k] = make_int4(-1, -1, -1, -1);
            if (nb > (unsigned)k) rec[k] = leaf[buf[k]];
        }
        int2 hit[TR_BUF];
#pragma unroll
        for (int k = 0; k < TR_BUF; k++)
            hit[k] = nb > (unsigned)k ? leaf_pair<MODE>(flt, qr, rec[k]) : make_int2(-1, -1);
        unsigned nv = 0;
#pragma unroll
        for (int k = 0; k < TR_BUF; k++) nv += hit[k].x >= 0 ? 1u : 0u;
        const int64_t first = (int64_t)shard_append(S, sh, nv);
        int64_t at = first;
#pragma unroll
        for (int k = 0; k < TR_BUF; k++)
            if (hit[k].x >= 0) {
                if (at < S.scap) reg[at] = hit[k];
                at++;
            }
        nb = 0;
    };
    int node = 0;
    if (active) {
        int root = root_p[0];
        if (fa.vscene) {
            const int sc = fa.vscene[qr.v[0]];
            root = root_p[sc];
            if (root == ROOT_NONE || (fa.active && !fa.active[sc])) active = false;
        }
        if (!active) {
        } else if (root < -nleaves || root >= (nin > 0 ? nin : 1)) {
            // the build did not finish: refuse to walk
            atomicOr(status, 2);
            active = false;
        } else if (root < 0) { // single target leaf (lbvh.cpp:386-396)
            if (!TRI) {
                const int lp = ~root;
                const float* b = tbox + (int64_t)6 * (fa.tids ? fa.tids[lp] : 0);
                if (overlap(q, b, b + 3)) {
                    buf[0] = lp;
                    nb = 1;
                }
            }
            active = false;
        } else {
            node = root;
        }
    }
    int sp = 0, steps = 0;
    const int max_steps = 2 * nleaves + 64; // each node is visited at most once per query
    while (__any(active)) {
        int h0 = -1, h1 = -1;
        bool nd_ok = true;
        if (active) {
            const Node nd = nodes[node]; // node: a valid index (checked when it was read)
            // links must name an internal node or a leaf position, and a walk
            // visits each node at most once: one rarely taken branch for both
            ++steps;
            if ((steps > max_steps) | (nd.c0 >= nin) | (nd.c1 >= nin) | (nd.c0 < -nleaves)
                | (nd.c1 < -nleaves)) {
                atomicOr(status, 2);
                active = false;
                nd_ok = false;
            }
            // branch-free tests (bitwise, no short-circuit)
            bool o0 = nd_ok & overlap(q, nd.lo0, nd.hi0), o1 = nd_ok & overlap(q, nd.lo1, nd.hi1);
            if (TRI) {
                o0 = o0 & (nd.r0 > qpos);
                o1 = o1 & (nd.r1 > qpos);
            }
            const bool l0 = nd.c0 < 0, l1 = nd.c1 < 0;
            h0 = (o0 & l0) ? ~nd.c0 : -1;
            h1 = (o1 & l1) ? ~nd.c1 : -1;
            const bool t0 = o0 & !l0, t1 = o1 & !l1;
            if (!nd_ok) {
            } else if (!t0 && !t1) {
                if (sp == 0) active = false;
                else {
                    --sp;
                    // the LDS entries by an LDS read; the deep ones through a
                    // load of their own (a shared select of the two addresses
                    // would turn every pop into a flat load)
                    if (sp < TR_LSTACK)
                        node = stack[sp][t];
                    else
                        node = __builtin_nontemporal_load(
                            &deep[(int64_t)qpos * (TR_STACK - TR_LSTACK) + sp - TR_LSTACK]);
                }
            } else {
                node = t0 ? nd.c0 : nd.c1;
                if (t0 && t1) {
                    if (sp < TR_LSTACK) stack[sp++][t] = nd.c1;
                    else if (sp < TR_STACK) {
                        deep[(int64_t)qpos * (TR_STACK - TR_LSTACK) + sp - TR_LSTACK] = nd.c1;
                        sp++;
                    } else {
                        atomicOr(status, 1);
                        active = false;
                    }
                }
            }
        }
        // (nb <= TR_BUF - 2 here) queue h0 then h1 without branches:
        // constant indices keep the buffer in registers, a missing hit
        // matches no slot
        {
            const unsigned s0 = h0 >= 0 ? nb : (unsigned)TR_BUF;
            nb += h0 >= 0 ? 1u : 0u;
            const unsigned s1 = h1 >= 0 ? nb : (unsigned)TR_BUF;
            nb += h1 >= 0 ? 1u : 0u;
#pragma unroll
            for (int k = 0; k < TR_BUF; k++)
                buf[k] = s0 == (unsigned)k ? h0 : (s1 == (unsigned)k ? h1 : buf[k]);
        }
        if (__any(nb > TR_BUF - 2)) flush(); // all lanes together
    }
    if (__any(nb > 0)) flush();
}
struct TraverseArgs {
    const float* qbox;
    const uint32_t* qids;
    int nq;
    const Node* nodes;
    const int32_t* root;
    const int4* leaf;
    const float* tbox;
    int nleaves;
    Filter flt;
    const int32_t* qmap;
    int2* out;
    Shards S;
    int* deep;
    int* status;
    ForestArgs fa;
};

template <int MODE, int SWEPT, int BLOCK>
void launch_traverse_b(const TraverseArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL((k_traverse<MODE, SWEPT, BLOCK>), dim3(grid_for(a.nq, BLOCK)), dim3(BLOCK),
                       0, s, a.qbox, a.qids, a.nq, a.nodes, a.root, a.leaf, a.tbox, a.nleaves,
                       a.flt, a.qmap, a.out, a.S, a.deep, a.status, a.fa);
}

template <int MODE, int SWEPT> void launch_traverse(int block, const TraverseArgs& a, hipStream_t s)
{
    if (block == 64) launch_traverse_b<MODE, SWEPT, 64>(a, s);
    else if (block == 256) launch_traverse_b<MODE, SWEPT, 256>(a, s);
    else launch_traverse_b<MODE, SWEPT, TR_BLOCK>(a, s);
}

constexpr int TUNE_BLOCKS[3] = { 64, 128, 256 };

// One traversal launch with the workgroup size tuned for this context: on
// the first call of `stage` with autotuning on, every size runs once on the
// real input (its output is discarded by resetting the list's shard
// counters) and the fastest, by HIP events on the context stream, is kept.
// Candidate sets do not depend on the size.
template <class Launch>
void tuned_traverse(Ctx& c, const std::string& stage, unsigned long long* cnt, Launch&& launch)
{
    hipStream_t s = c.stream;
    auto it = c.launch_block.find(stage);
    int block = it != c.launch_block.end() ? it->second : 0;
    if (block == 0 && c.autotune) {
        hipEvent_t e0, e1;
        IPCG_HIP(hipEventCreate(&e0));
        IPCG_HIP(hipEventCreate(&e1));
        float best = INFINITY;
        for (int b : TUNE_BLOCKS) {
            IPCG_HIP(hipMemsetAsync(cnt, 0, SHARD_CNT_BYTES, s));
            IPCG_HIP(hipEventRecord(e0, s));
            launch(b);
            IPCG_HIP(hipEventRecord(e1, s));
            IPCG_HIP(hipEventSynchronize(e1));
            float ms = 0;
            IPCG_HIP(hipEventElapsedTime(&ms, e0, e1));
            c.tune_ms[stage + "@" + std::to_string(b)] = ms;
            if (ms < best) {
                best = ms;
                block = b;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        c.launch_block[stage] = block;
        IPCG_HIP(hipMemsetAsync(cnt, 0, SHARD_CNT_BYTES, s));
    }
    launch(block ? block : TR_BLOCK);
}

// ---------------------------------------------------------------------------
// Sweep and prune (sweep_and_prune.cpp:22-249 over scalable-ccd's
// sort_and_sweep): FP64 boxes sorted by (min along the sweep axis, id); a
// lane per sorted box scans forward while the next box starts within its
// extent along the axis and reports full 3D overlaps that pass the filters.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ordered_bits(double x)
{
    x = x + 0.0; // -0 -> +0: equal keys stay equal
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

template <int ARITY>
__global__ __launch_bounds__(BP_BLOCK) void k_sap_boxes(const double* X0, const double* X1,
                                                        int nV, const int32_t* I, int n, double r,
                                                        int axis, double* box64, uint64_t* keys,
                                                        uint32_t* ids)
{
    const int p = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (p >= n) return;
    double lo[3], hi[3];
    vbox(X0, X1, nV, ARITY == 1 ? p : I[(int64_t)ARITY * p], r, lo, hi);
#pragma unroll
    for (int j = 1; j < ARITY; j++) {
        double l[3], h[3];
        vbox(X0, X1, nV, I[(int64_t)ARITY * p + j], r, l, h);
#pragma unroll
        for (int k = 0; k < 3; k++) {
            lo[k] = l[k] < lo[k] ? l[k] : lo[k];
            hi[k] = hi[k] < h[k] ? h[k] : hi[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        box64[(int64_t)6 * p + k] = lo[k];
        box64[(int64_t)6 * p + 3 + k] = hi[k];
    }
    keys[p] = ordered_bits(lo[axis]);
    ids[p] = (uint32_t)p;
}

// boxes in sorted order (coalesced sweeps)
__global__ __launch_bounds__(BP_BLOCK) void k_sap_gather(const double* box64, const uint32_t* ids,
                                                         int n, double* sorted)
{
    const int i = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = ids[i];
#pragma unroll
    for (int k = 0; k < 6; k++) sorted[(int64_t)6 * i + k] = box64[(int64_t)6 * p + k];
}

__device__ __forceinline__ bool overlap64(const double* a, const double* b)
{
    return a[0] <= b[3] && a[1] <= b[4] && a[2] <= b[5] && b[0] <= a[3] && b[1] <= a[4]
        && b[2] <= a[5];
}

// Edge-edge: pairs (earlier, later) in sorted order.
__global__ __launch_bounds__(TR_BLOCK) void k_sap_self(const double* sb, const uint32_t* ids, int n,
                                                       int axis, Filter flt, int2* out, Shards S)
{
    const int i = blockIdx.x * TR_BLOCK + threadIdx.x;
    bool active = i < n;
    double a[6] = { 0, 0, 0, 0, 0, 0 };
    int ai = 0;
    if (active) {
#pragma unroll
        for (int k = 0; k < 6; k++) a[k] = sb[(int64_t)6 * i + k];
        ai = (int)ids[i];
    }
    const int sh = wave_shard();
    int j = i + 1;
    while (__any(active)) {
        int hit = -1;
        if (active) {
            if (j >= n || sb[(int64_t)6 * j + axis] > a[3 + axis]) {
                active = false;
            } else {
                const int bj = (int)ids[j];
                if (overlap64(a, sb + (int64_t)6 * j) && can_ee(flt, ai, bj)) hit = bj;
                j++;
            }
        }
        unsigned pre;
        const int64_t slot = (int64_t)shard_append_bits(S, sh, hit >= 0, false, pre) + pre;
        if (hit >= 0 && slot < S.scap) out[sh * S.scap + slot] = make_int2(ai, hit);
    }
}

// Face-vertex from one side: each query box (sorted, `qa`) scans the other
// list (`sb`) from the first box starting at (STRICT: after) its own start.
// FACES: the queries are faces; otherwise vertices.
template <bool FACES, bool STRICT>
__global__ __launch_bounds__(TR_BLOCK) void k_sap_cross(const double* qa, const uint32_t* qids,
                                                        int nq, const double* sb,
                                                        const uint32_t* sids, int ns, int axis,
                                                        Filter flt, int2* out, Shards S)
{
    const int i = blockIdx.x * TR_BLOCK + threadIdx.x;
    bool active = i < nq;
    double a[6] = { 0, 0, 0, 0, 0, 0 };
    int ai = 0, j = 0;
    if (active) {
#pragma unroll
        for (int k = 0; k < 6; k++) a[k] = qa[(int64_t)6 * i + k];
        ai = (int)qids[i];
        int lo = 0, hi = ns;
        while (lo < hi) {
            const int m = (lo + hi) >> 1;
            const double km = sb[(int64_t)6 * m + axis];
            if (STRICT ? km <= a[axis] : km < a[axis]) lo = m + 1;
            else hi = m;
        }
        j = lo;
    }
    const int sh = wave_shard();
    while (__any(active)) {
        int2 hit = make_int2(-1, -1);
        if (active) {
            if (j >= ns || sb[(int64_t)6 * j + axis] > a[3 + axis]) {
                active = false;
            } else {
                const int bj = (int)sids[j];
                const int f = FACES ? ai : bj, v = FACES ? bj : ai;
                if (overlap64(a, sb + (int64_t)6 * j) && can_fv(flt, f, v)) hit = make_int2(f, v);
                j++;
            }
        }
        unsigned pre;
        const int64_t slot = (int64_t)shard_append_bits(S, sh, hit.x >= 0, false, pre) + pre;
        if (hit.x >= 0 && slot < S.scap) out[sh * S.scap + slot] = hit;
    }
}

void sap_list(Ctx& c, Tree& T, const int32_t* I, int arity, int n, const double* X0,
              const double* X1, double r, int axis)
{
    hipStream_t s = c.stream;
    T.n = n;
    if (n <= 0) return;
    T.box.reserve(sizeof(double) * 6 * (size_t)n);
    T.sbox.reserve(sizeof(double) * 6 * (size_t)n);
    T.keys.reserve(sizeof(uint64_t) * (size_t)n);
    T.keys2.reserve(sizeof(uint64_t) * (size_t)n);
    T.ids.reserve(sizeof(uint32_t) * (size_t)n);
    T.ids2.reserve(sizeof(uint32_t) * (size_t)n);
    const unsigned g = grid_for(n, BP_BLOCK);
    if (arity == 1)
        hipLaunchKernelGGL(k_sap_boxes<1>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           axis, T.box.as<double>(), T.keys.as<uint64_t>(), T.ids.as<uint32_t>());
    else if (arity == 2)
        hipLaunchKernelGGL(k_sap_boxes<2>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           axis, T.box.as<double>(), T.keys.as<uint64_t>(), T.ids.as<uint32_t>());
    else
        hipLaunchKernelGGL(k_sap_boxes<3>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           axis, T.box.as<double>(), T.keys.as<uint64_t>(), T.ids.as<uint32_t>());
    IPCG_CHECK_LAUNCH();
    c.sort_tmp.reserve(radix_tmp_bytes(n));
    radix_sort_pairs(T.keys.as<uint64_t>(), T.ids.as<uint32_t>(), T.keys2.as<uint64_t>(),
                     T.ids2.as<uint32_t>(), n, 64, c.sort_tmp.p, s);
    hipLaunchKernelGGL(k_sap_gather, dim3(g), dim3(BP_BLOCK), 0, s, T.box.as<double>(),
                       T.ids.as<uint32_t>(), n, T.sbox.as<double>());
    IPCG_CHECK_LAUNCH();
}

// Leaf records in sorted order: the traversal tests a leaf with one 16-byte
// load (mesh id of the primitive and its vertex ids) instead of two
// dependent gathers (sorted position -> primitive -> vertices).
template <int ARITY>
__global__ __launch_bounds__(BP_BLOCK) void k_leaf_records(const uint32_t* ids, const int32_t* I,
                                                           const int32_t* out_ids, int n,
                                                           int4* rec)
{
    const int i = blockIdx.x * BP_BLOCK + threadIdx.x;
    if (i >= n) return;
    const int p = (int)ids[i];
    int v[3] = { -1, -1, -1 };
#pragma unroll
    for (int j = 0; j < ARITY; j++) v[j] = I ? I[(int64_t)ARITY * p + j] : p;
    rec[i] = make_int4(out_ids ? out_ids[p] : p, v[0], v[1], v[2]);
}

// out_ids: mesh id of each primitive (null: its index).  Batched contexts
// build a forest: `loff` holds each scene's first primitive (n_scenes + 1;
// primitives of a scene are contiguous), c.vscene the scene per vertex.
void build_tree(Ctx& c, Tree& T, const int32_t* I, int arity, int n, const double* X0,
                const double* X1, double r, bool with_nodes, const int32_t* out_ids = nullptr,
                const std::vector<int64_t>* loff = nullptr)
{
    hipStream_t s = c.stream;
    T.n = n;
    const bool forest = loff != nullptr;
    const int S = forest ? (int)loff->size() - 1 : 1;
    T.root.reserve(sizeof(int32_t) * (size_t)S);
    if (forest) { // root entries for scenes with no or one primitive
        T.loff.reserve(sizeof(int64_t) * loff->size());
        IPCG_HIP(hipMemcpyAsync(T.loff.p, loff->data(), sizeof(int64_t) * loff->size(),
                                hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_forest_roots, dim3(grid_for(S, 256)), dim3(256), 0, s,
                           T.loff.as<int64_t>(), S, T.root.as<int32_t>());
        IPCG_CHECK_LAUNCH();
    }
    if (n <= 0) {
        if (forest) IPCG_HIP(hipStreamSynchronize(s)); // the offsets' host copy
        return;
    }
    T.box.reserve(sizeof(float) * 6 * (size_t)n);
    T.keys.reserve(sizeof(uint64_t) * (size_t)n);
    T.keys2.reserve(sizeof(uint64_t) * (size_t)n);
    T.ids.reserve(sizeof(uint32_t) * (size_t)n);
    T.ids2.reserve(sizeof(uint32_t) * (size_t)n);
    const int32_t* vscene = forest ? c.vscene.as<int32_t>() : nullptr;
    if (forest) {
        T.mort.reserve(sizeof(uint64_t) * (size_t)n);
        T.lscene.reserve(sizeof(int32_t) * (size_t)n);
    }
    uint64_t* mort = forest ? T.mort.as<uint64_t>() : nullptr;
    const double* aabb = c.aabb.as<double>();
    const unsigned g = grid_for(n, BP_BLOCK);
    if (arity == 1)
        hipLaunchKernelGGL(k_prim_boxes<1>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           aabb, vscene, T.box.as<float>(), T.keys.as<uint64_t>(),
                           T.ids.as<uint32_t>(), mort);
    else if (arity == 2)
        hipLaunchKernelGGL(k_prim_boxes<2>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           aabb, vscene, T.box.as<float>(), T.keys.as<uint64_t>(),
                           T.ids.as<uint32_t>(), mort);
    else
        hipLaunchKernelGGL(k_prim_boxes<3>, dim3(g), dim3(BP_BLOCK), 0, s, X0, X1, c.nV, I, n, r,
                           aabb, vscene, T.box.as<float>(), T.keys.as<uint64_t>(),
                           T.ids.as<uint32_t>(), mort);
    IPCG_CHECK_LAUNCH();
    c.sort_tmp.reserve(radix_tmp_bytes(n));
    c.stage_begin("bp_sort");
    if (radix_sort_pairs_all(T.keys.as<uint64_t>(), T.ids.as<uint32_t>(), T.keys2.as<uint64_t>(),
                             T.ids2.as<uint32_t>(), n, 63, c.sort_tmp.p, s)) {
        std::swap(T.keys, T.keys2);
        std::swap(T.ids, T.ids2);
    }
    if (forest) { // stable second pass by scene, then the Morton codes in that order
        if (arity == 1)
            hipLaunchKernelGGL(k_scene_keys<1>, dim3(g), dim3(BP_BLOCK), 0, s,
                               T.ids.as<uint32_t>(), I, vscene, n, T.keys.as<uint64_t>());
        else if (arity == 2)
            hipLaunchKernelGGL(k_scene_keys<2>, dim3(g), dim3(BP_BLOCK), 0, s,
                               T.ids.as<uint32_t>(), I, vscene, n, T.keys.as<uint64_t>());
        else
            hipLaunchKernelGGL(k_scene_keys<3>, dim3(g), dim3(BP_BLOCK), 0, s,
                               T.ids.as<uint32_t>(), I, vscene, n, T.keys.as<uint64_t>());
        IPCG_CHECK_LAUNCH();
        int bits = 1;
        while ((1ll << bits) < S) bits++;
        radix_sort_pairs(T.keys.as<uint64_t>(), T.ids.as<uint32_t>(), T.keys2.as<uint64_t>(),
                         T.ids2.as<uint32_t>(), n, bits, c.sort_tmp.p, s);
        hipLaunchKernelGGL(k_forest_keys, dim3(g), dim3(BP_BLOCK), 0, s, T.ids.as<uint32_t>(),
                           mort, n, T.keys.as<uint64_t>(), T.lscene.as<int32_t>());
        IPCG_CHECK_LAUNCH();
    }
    c.stage_end();
    if (!with_nodes) return;
    const int32_t* lscene = forest ? T.lscene.as<int32_t>() : nullptr;
    T.leaf.reserve(sizeof(int4) * (size_t)n);
    if (arity == 1)
        hipLaunchKernelGGL(k_leaf_records<1>, dim3(g), dim3(BP_BLOCK), 0, s, T.ids.as<uint32_t>(), I,
                           out_ids, n, T.leaf.as<int4>());
    else if (arity == 2)
        hipLaunchKernelGGL(k_leaf_records<2>, dim3(g), dim3(BP_BLOCK), 0, s, T.ids.as<uint32_t>(), I,
                           out_ids, n, T.leaf.as<int4>());
    else
        hipLaunchKernelGGL(k_leaf_records<3>, dim3(g), dim3(BP_BLOCK), 0, s, T.ids.as<uint32_t>(), I,
                           out_ids, n, T.leaf.as<int4>());
    IPCG_CHECK_LAUNCH();
    T.nodes.reserve(sizeof(Node) * (size_t)(n > 1 ? n - 1 : 1));
    if (n == 1 && !forest) { // single leaf: no internal node (lbvh.cpp:386-396)
        const int32_t one = -1;
        IPCG_HIP(hipMemcpyAsync(T.root.p, &one, sizeof(int32_t), hipMemcpyHostToDevice, s));
        IPCG_HIP(hipStreamSynchronize(s));
        return;
    }
    // a forest of single-leaf scenes has no internal node either
    bool any_node = !forest;
    if (forest)
        for (size_t k = 0; k + 1 < loff->size(); k++)
            if ((*loff)[k + 1] - (*loff)[k] >= 2) any_node = true;
    if (!any_node) {
        IPCG_HIP(hipStreamSynchronize(s));
        return;
    }
    // int32 arrays: parent[n-1], visits[n-1], two work lists of n, and three
    // counters (level k reads cnt[(k-1) % 3], appends to cnt[k % 3] and
    // clears cnt[(k+1) % 3]; the lists alternate)
    const size_t ni = (size_t)n - 1;
    T.info.reserve(sizeof(int32_t) * (2 * ni + 2 * (size_t)n) + 256);
    int32_t* parent = T.info.as<int32_t>();
    int32_t* visits = parent + ni;
    int32_t* work[2] = { visits + ni, visits + ni + n };
    uint32_t* cnt = reinterpret_cast<uint32_t*>(work[1] + n);
    IPCG_HIP(hipMemsetAsync(cnt, 0, 3 * sizeof(uint32_t), s));
    if (!forest) // "unfinished" sentinel
        IPCG_HIP(hipMemsetAsync(T.root.p, 0x80, sizeof(int32_t), s));
    c.stage_begin("bp_build");
    hipLaunchKernelGGL(k_karras, dim3(grid_for(ni, BP_BLOCK)), dim3(BP_BLOCK), 0, s,
                       T.keys.as<uint64_t>(), lscene, n, T.nodes.as<Node>(), parent,
                       T.ids.as<uint32_t>(), T.box.as<float>(), visits, work[0], cnt);
    IPCG_CHECK_LAUNCH();
    // Round k completes the nodes of height k; their subtrees are disjoint and
    // hold at least k + 1 leaves each, so round k has at most n / (k + 1)
    // nodes.
    // Rounds are launched in batches sized by that bound (a round past the
    // roots finds an empty list) and the host checks for the end once per
    // batch.  Every round completes at least one node, so n - 1 rounds bound
    // the loop.
    uint32_t* h = static_cast<uint32_t*>(c.pinned);
    constexpr size_t BATCH = 8;
    for (size_t level = 1;; level += BATCH) {
        for (size_t k = level; k < level + BATCH; k++) {
            const int64_t bound = std::max<int64_t>((int64_t)n / (int64_t)(k + 1), 1);
            const unsigned grid = std::min<unsigned>(grid_for(bound, RF_BLOCK), RF_GRID);
            hipLaunchKernelGGL(k_refit_level, dim3(grid), dim3(RF_BLOCK), 0, s,
                               work[(k - 1) & 1], cnt + (k - 1) % 3, parent, lscene,
                               T.nodes.as<Node>(), visits, work[k & 1], cnt + k % 3,
                               cnt + (k + 1) % 3, T.root.as<int32_t>());
        }
        IPCG_CHECK_LAUNCH();
        IPCG_HIP(hipMemcpyAsync(h, cnt + (level + BATCH - 1) % 3, sizeof(uint32_t),
                                hipMemcpyDeviceToHost, s));
        IPCG_HIP(hipStreamSynchronize(s));
        if (h[0] == 0) break;
        if (h[0] <= (uint32_t)RF_TOP) { // the rest in one workgroup
            const size_t k = level + BATCH; // the next round
            hipLaunchKernelGGL(k_refit_top, dim3(1), dim3(RF_BLOCK), 0, s, work[(k - 1) & 1],
                               cnt + (k - 1) % 3, parent, lscene, T.nodes.as<Node>(), visits,
                               T.root.as<int32_t>(), (int64_t)ni);
            IPCG_CHECK_LAUNCH();
            break;
        }
        if (level >= ni) throw Error(IPCG_ERR_NUMERIC, "LBVH refit did not converge");
    }
    c.stage_end();
}

ForestArgs forest_args(const Ctx& c, const Tree& target)
{
    if (!c.batched()) return { nullptr, nullptr, target.ids.as<uint32_t>() };
    return { c.vscene.as<int32_t>(), c.all_active ? nullptr : c.scene_active.as<uint8_t>(),
             target.ids.as<uint32_t>() };
}

// mesh box of the vertex boxes (broad_phase.cpp:93-125) into c.aabb
void mesh_bounds(Ctx& c, const double* X0, const double* X1, double inflation)
{
    hipStream_t s = c.stream;
    c.stage_begin("bp_boxes");
    if (c.batched()) { // one box per scene
        c.aabb.reserve(sizeof(double) * 6 * (size_t)c.n_scenes);
        c.aabb_partial.reserve(sizeof(double) * 6 * SCENE_CHUNKS * (size_t)c.n_scenes);
        hipLaunchKernelGGL(k_scene_aabb, dim3(SCENE_CHUNKS * c.n_scenes), dim3(BP_BLOCK), 0, s, X0,
                           X1, c.nV, c.voff_d.as<int64_t>(), inflation,
                           c.aabb_partial.as<double>());
        hipLaunchKernelGGL(k_scene_aabb_final, dim3((c.n_scenes + 63) / 64), dim3(64), 0, s,
                           c.aabb_partial.as<double>(), c.n_scenes, c.aabb.as<double>());
        IPCG_CHECK_LAUNCH();
        c.stage_end();
        return;
    }
    const int nb = (int)grid_for(c.nV, BP_BLOCK, 1024);
    c.aabb_partial.reserve(sizeof(double) * 6 * nb);
    c.aabb.reserve(sizeof(double) * 6);
    hipLaunchKernelGGL(k_mesh_aabb_partial, dim3(nb), dim3(BP_BLOCK), 0, s, X0, X1, c.nV,
                       inflation, c.aabb_partial.as<double>());
    hipLaunchKernelGGL(k_mesh_aabb_final, dim3(1), dim3(64), 0, s, c.aabb_partial.as<double>(),
                       nb, c.aabb.as<double>());
    IPCG_CHECK_LAUNCH();
    c.stage_end();
}

} // namespace

void bp_build(Ctx& c, bool swept, double inflation, int method)
{
    c.require_mesh();
    c.require_pos(0);
    if (swept) c.require_pos(1);
    if (method != IPCG_BP_LBVH && method != IPCG_BP_SAP)
        throw Error(IPCG_ERR_ARG, "unsupported broad phase method");
    const bool sap = method == IPCG_BP_SAP;
    if (sap && c.batched())
        throw Error(IPCG_ERR_ARG, "a batched context supports the LBVH broad phase only");
    hipStream_t s = c.stream;
    const double* X0 = c.X[0].as<double>();
    const double* X1 = swept ? c.X[1].as<double>() : nullptr;
    c.cand_valid = false;
    c.n_ee = c.n_fv = c.n_vv = c.n_ev = 0;
    struct Tag { // stage names carry the query kind while this call runs
        Ctx& c;
        Tag(Ctx& c, bool swept) : c(c) { c.stage_tag = swept ? ".swept" : ".static"; }
        ~Tag() { c.stage_tag.clear(); }
    } tag(c, swept);
    if (c.nV == 0) {
        c.cand_valid = true;
        c.cand_swept = swept;
        return;
    }
    mesh_bounds(c, X0, X1, inflation);
    int axis = 0;
    if (sap) {
        // sweep axis: largest extent of the mesh box
        double* hb = static_cast<double*>(c.pinned);
        IPCG_HIP(hipMemcpyAsync(hb, c.aabb.p, sizeof(double) * 6, hipMemcpyDeviceToHost, s));
        IPCG_HIP(hipStreamSynchronize(s));
        for (int k = 1; k < 3; k++)
            if (hb[3 + k] - hb[k] > hb[3 + axis] - hb[axis]) axis = k;
        sap_list(c, c.tv, nullptr, 1, c.nV, X0, X1, inflation, axis);
        sap_list(c, c.te, c.E.as<int32_t>(), 2, c.nE, X0, X1, inflation, axis);
        sap_list(c, c.tf, c.F.as<int32_t>(), 3, c.nF, X0, X1, inflation, axis);
    } else {
        // vertex order (queries of the face tree), edge tree, face tree
        const bool fo = c.batched();
        build_tree(c, c.tv, nullptr, 1, c.nV, X0, X1, inflation, false, nullptr,
                   fo ? &c.h_voff : nullptr);
        build_tree(c, c.te, c.E.as<int32_t>(), 2, c.nE, X0, X1, inflation, true, nullptr,
                   fo ? &c.h_eoff : nullptr);
        build_tree(c, c.tf, c.F.as<int32_t>(), 3, c.nF, X0, X1, inflation, true, nullptr,
                   fo ? &c.h_foff : nullptr);
    }
    c.row_order_valid = true; // c.aabb: the box of every scene
    // codim lists (candidates.cpp:65-118): a tree over the codim vertices
    // (its own queries) and one over the codim edges; LBVH semantics with
    // either method
    const bool codim = c.n_cv > 0;
    if (codim) {
        const bool fo = c.batched();
        build_tree(c, c.tcv, c.cv.as<int32_t>(), 1, c.n_cv, X0, X1, inflation, true,
                   c.cv.as<int32_t>(), fo ? &c.h_cvoff : nullptr);
        build_tree(c, c.tce, c.cev.as<int32_t>(), 2, c.n_ce, X0, X1, inflation, true,
                   c.ce.as<int32_t>(), fo ? &c.h_ceoff : nullptr);
    }

    c.counters.reserve(sizeof(unsigned long long) * 4);
    Filter flt { c.E.as<int32_t>(), c.F.as<int32_t>(), c.has_patch ? c.patch.as<int32_t>() : nullptr,
                 c.n_dynamic, c.n_codim_filter };
    int* status = reinterpret_cast<int*>(c.counters.as<unsigned long long>() + 2);
    c.shard_cnt.reserve(4 * SHARD_CNT_BYTES);
    c.shard_off.reserve(sizeof(int64_t) * 4 * (SHARDS + 1));
    unsigned long long* cnt[4];
    for (int k = 0; k < 4; k++) cnt[k] = c.shard_cnt.as<unsigned long long>() + k * SHARDS * SHARD_PAD;
    int64_t* scap[4] = { &c.scap_ee, &c.scap_fv, &c.scap_vv, &c.scap_ev };
    const int64_t guess[4] = { 16 * (int64_t)c.nE, 8 * (int64_t)c.nV, 4 * (int64_t)c.n_cv,
                               4 * (int64_t)c.n_cv };
    DevBuf* final_list[4] = { &c.cand_ee, &c.cand_fv, &c.cand_vv, &c.cand_ev };
    int64_t* final_n[4] = { &c.n_ee, &c.n_fv, &c.n_vv, &c.n_ev };
    const int lists = codim ? 4 : 2;
    // deep traversal stack entries, per query of the largest query set
    const int64_t nq_max = std::max<int64_t>(std::max<int64_t>(c.nE, c.nV), 1);
    c.tr_deep.reserve(sizeof(int) * (size_t)nq_max * (TR_STACK - TR_LSTACK));
    int* deep = c.tr_deep.as<int>();
    for (int attempt = 0; attempt < 3; attempt++) {
        // per-shard regions; their sum bounds a list until a retry grows it
        for (int k = 0; k < lists; k++) {
            if (*scap[k] == 0) *scap[k] = std::max<int64_t>(1024, guess[k] / SHARDS);
            c.cand_stage[k].reserve(sizeof(int2) * (size_t)(SHARDS * *scap[k]));
        }
        IPCG_HIP(hipMemsetAsync(c.shard_cnt.p, 0, 4 * SHARD_CNT_BYTES, s));
        IPCG_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
        const Shards S0 { cnt[0], c.scap_ee }, S1 { cnt[1], c.scap_fv };
        if (sap) {
            c.stage_begin("bp_sweep");
            if (c.nE >= 2)
                hipLaunchKernelGGL(k_sap_self, dim3(grid_for(c.nE, TR_BLOCK)), dim3(TR_BLOCK), 0, s,
                                   c.te.sbox.as<double>(), c.te.ids.as<uint32_t>(), c.nE, axis,
                                   flt, c.cand_stage[0].as<int2>(), S0);
            if (c.nF >= 1) {
                hipLaunchKernelGGL((k_sap_cross<true, false>), dim3(grid_for(c.nF, TR_BLOCK)),
                                   dim3(TR_BLOCK), 0, s, c.tf.sbox.as<double>(),
                                   c.tf.ids.as<uint32_t>(), c.nF, c.tv.sbox.as<double>(),
                                   c.tv.ids.as<uint32_t>(), c.nV, axis, flt,
                                   c.cand_stage[1].as<int2>(), S1);
                hipLaunchKernelGGL((k_sap_cross<false, true>), dim3(grid_for(c.nV, TR_BLOCK)),
                                   dim3(TR_BLOCK), 0, s, c.tv.sbox.as<double>(),
                                   c.tv.ids.as<uint32_t>(), c.nV, c.tf.sbox.as<double>(),
                                   c.tf.ids.as<uint32_t>(), c.nF, axis, flt,
                                   c.cand_stage[1].as<int2>(), S1);
            }
            c.stage_end();
        } else {
            const std::string tag = swept ? ".swept" : ".static";
            if (c.nE >= 2) {
                const TraverseArgs a { c.te.box.as<float>(), c.te.ids.as<uint32_t>(), c.nE,
                                       c.te.nodes.as<Node>(), c.te.root.as<int32_t>(),
                                       c.te.leaf.as<int4>(), c.te.box.as<float>(), c.nE, flt,
                                       nullptr, c.cand_stage[0].as<int2>(), S0, deep, status,
                                       forest_args(c, c.te) };
                c.stage_begin("bp_traverse_ee");
                tuned_traverse(c, "bp_traverse_ee" + tag, cnt[0], [&](int b) {
                    if (swept) launch_traverse<TR_EE, 1>(b, a, s);
                    else launch_traverse<TR_EE, 0>(b, a, s);
                });
                c.stage_end();
            }
            if (c.nF >= 1) {
                const TraverseArgs a { c.tv.box.as<float>(), c.tv.ids.as<uint32_t>(), c.nV,
                                       c.tf.nodes.as<Node>(), c.tf.root.as<int32_t>(),
                                       c.tf.leaf.as<int4>(), c.tf.box.as<float>(), c.nF, flt,
                                       nullptr, c.cand_stage[1].as<int2>(), S1, deep, status,
                                       forest_args(c, c.tf) };
                c.stage_begin("bp_traverse_fv");
                tuned_traverse(c, "bp_traverse_fv" + tag, cnt[1], [&](int b) {
                    if (swept) launch_traverse<TR_FV, 1>(b, a, s);
                    else launch_traverse<TR_FV, 0>(b, a, s);
                });
                c.stage_end();
            }
        }
        if (codim) {
            c.stage_begin("bp_traverse_codim");
            const Shards S2 { cnt[2], c.scap_vv }, S3 { cnt[3], c.scap_ev };
            if (c.n_cv >= 2) {
                const TraverseArgs a { c.tcv.box.as<float>(), c.tcv.ids.as<uint32_t>(), c.n_cv,
                                       c.tcv.nodes.as<Node>(), c.tcv.root.as<int32_t>(),
                                       c.tcv.leaf.as<int4>(), c.tcv.box.as<float>(), c.n_cv, flt,
                                       c.cv.as<int32_t>(), c.cand_stage[2].as<int2>(), S2, deep,
                                       status, forest_args(c, c.tcv) };
                if (swept) launch_traverse<TR_VV, 1>(TR_BLOCK, a, s);
                else launch_traverse<TR_VV, 0>(TR_BLOCK, a, s);
            }
            if (c.n_ce >= 1) {
                const TraverseArgs a { c.tcv.box.as<float>(), c.tcv.ids.as<uint32_t>(), c.n_cv,
                                       c.tce.nodes.as<Node>(), c.tce.root.as<int32_t>(),
                                       c.tce.leaf.as<int4>(), c.tce.box.as<float>(), c.n_ce, flt,
                                       c.cv.as<int32_t>(), c.cand_stage[3].as<int2>(), S3, deep,
                                       status, forest_args(c, c.tce) };
                if (swept) launch_traverse<TR_EV, 1>(TR_BLOCK, a, s);
                else launch_traverse<TR_EV, 0>(TR_BLOCK, a, s);
            }
            c.stage_end();
        }
        IPCG_CHECK_LAUNCH();
        // the lists' counts and the walk status in one read-back
        ShardCounts kc[4];
        int st = 0;
        c.shard_sum.reserve(256);
        shard_counts_n(cnt[0], lists, c.shard_sum.as<unsigned long long>(), c.pinned, s, kc,
                       status, &st);
        if (st & 2) throw Error(IPCG_ERR_NUMERIC, "LBVH build produced an invalid tree");
        if (st & 1) throw Error(IPCG_ERR_NUMERIC, "LBVH traversal stack overflow (depth > 64)");
        bool again = false;
        for (int k = 0; k < lists; k++) {
            if (kc[k].max > *scap[k]) {
                *scap[k] = kc[k].max + kc[k].max / 4 + 1024;
                again = true;
            }
        }
        if (again) continue;
        c.stage_begin("bp_compact");
        for (int k = 0; k < 4; k++) {
            const int64_t m = k < lists ? kc[k].total : 0;
            final_list[k]->reserve(sizeof(int2) * (size_t)std::max<int64_t>(m, 1));
            if (m > 0)
                shard_compact(c.cand_stage[k].as<int2>(), *scap[k], cnt[k], kc[k].max,
                              final_list[k]->as<int2>(), c.shard_off.as<int64_t>() + k * (SHARDS + 1),
                              s);
            *final_n[k] = m;
        }
        c.stage_end();
        c.cand_valid = true;
        c.cand_swept = swept;
        return;
    }
    throw Error(IPCG_ERR_NUMERIC, "broad phase candidate lists kept outgrowing their regions");
}

// detect_edge_face_candidates (lbvh.cpp:771-784) over static boxes of
// positions slot 0: edge queries against a face LBVH, (edge, face) pairs
// into c.cand_ef.  The edge and face trees of an earlier broad phase are
// rebuilt here; its candidate lists stay valid.
void bp_edge_face(Ctx& c, double inflation)
{
    c.require_mesh();
    c.require_pos(0);
    c.n_ef = 0;
    if (c.nE == 0 || c.nF == 0) return;
    hipStream_t s = c.stream;
    const double* X0 = c.X[0].as<double>();
    struct Tag {
        Ctx& c;
        Tag(Ctx& c) : c(c) { c.stage_tag = ".ef"; }
        ~Tag() { c.stage_tag.clear(); }
    } tag(c);
    mesh_bounds(c, X0, nullptr, inflation);
    const bool fo = c.batched();
    build_tree(c, c.te, c.E.as<int32_t>(), 2, c.nE, X0, nullptr, inflation, false, nullptr,
               fo ? &c.h_eoff : nullptr);
    build_tree(c, c.tf, c.F.as<int32_t>(), 3, c.nF, X0, nullptr, inflation, true, nullptr,
               fo ? &c.h_foff : nullptr);
    Filter flt { c.E.as<int32_t>(), c.F.as<int32_t>(), c.has_patch ? c.patch.as<int32_t>() : nullptr,
                 c.n_dynamic, c.n_codim_filter };
    c.counters.reserve(sizeof(unsigned long long) * 4);
    int* status = reinterpret_cast<int*>(c.counters.as<unsigned long long>() + 2);
    c.shard_cnt.reserve(4 * SHARD_CNT_BYTES);
    c.shard_off.reserve(sizeof(int64_t) * 4 * (SHARDS + 1));
    unsigned long long* cnt = c.shard_cnt.as<unsigned long long>();
    c.tr_deep.reserve(sizeof(int) * (size_t)c.nE * (TR_STACK - TR_LSTACK));
    for (int attempt = 0; attempt < 3; attempt++) {
        if (c.scap_ef == 0) c.scap_ef = std::max<int64_t>(1024, 8 * (int64_t)c.nE / SHARDS);
        c.cand_stage[0].reserve(sizeof(int2) * (size_t)(SHARDS * c.scap_ef));
        IPCG_HIP(hipMemsetAsync(cnt, 0, SHARD_CNT_BYTES, s));
        IPCG_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
        const Shards S0 { cnt, c.scap_ef };
        c.stage_begin("bp_traverse_ef");
        const TraverseArgs a { c.te.box.as<float>(), c.te.ids.as<uint32_t>(), c.nE,
                               c.tf.nodes.as<Node>(), c.tf.root.as<int32_t>(), c.tf.leaf.as<int4>(),
                               c.tf.box.as<float>(), c.nF, flt, nullptr, c.cand_stage[0].as<int2>(),
                               S0, c.tr_deep.as<int>(), status, forest_args(c, c.tf) };
        launch_traverse<TR_EF, 0>(TR_BLOCK, a, s);
        IPCG_CHECK_LAUNCH();
        c.stage_end();
        ShardCounts kc;
        int st = 0;
        c.shard_sum.reserve(256);
        shard_counts_n(cnt, 1, c.shard_sum.as<unsigned long long>(), c.pinned, s, &kc, status,
                       &st);
        if (st & 2) throw Error(IPCG_ERR_NUMERIC, "LBVH build produced an invalid tree");
        if (st & 1) throw Error(IPCG_ERR_NUMERIC, "LBVH traversal stack overflow (depth > 64)");
        if (kc.max > c.scap_ef) {
            c.scap_ef = kc.max + kc.max / 4 + 1024;
            continue;
        }
        c.cand_ef.reserve(sizeof(int2) * (size_t)std::max<int64_t>(kc.total, 1));
        if (kc.total > 0)
            shard_compact(c.cand_stage[0].as<int2>(), c.scap_ef, cnt, kc.max, c.cand_ef.as<int2>(),
                          c.shard_off.as<int64_t>(), s);
        c.n_ef = kc.total;
        return;
    }
    throw Error(IPCG_ERR_NUMERIC, "edge-face candidate list kept outgrowing its regions");
}

} // namespace ipcg
