// Narrow phase: classify every candidate's distance type, test activity
// d^2 < (dhat + dmin)^2 and reduce it to a VV / EV / EE / FV collision
// (normal_collisions_builder.cpp:140-336), then merge duplicates summing
// their weights (normal_collisions_builder.cpp:547-684) by sorting the
// collision keys and run-length encoding them.  Output order is canonical:
// vv by (min, max), ev by (edge, vertex), ee by (min edge, max edge, distance
// type), fv by (face, vertex).
//
// With area weighting, the IMPROVED_MAX_APPROX collision set
// (normal_collisions.cpp:84-128) or the OGC set the weights are real
// numbers: a run's
// weight is the sum of its contributions in ascending order of value (the
// reference's own order follows its thread schedule), and vertex-vertex,
// edge-vertex and edge-edge collisions whose weights sum to zero are dropped.
// The improved set converts the candidates to sub-element pairs
// (candidates.cpp:583-692): one slot per (candidate, sub-element) holding a
// tagged pair key or a sentinel, sorted and run-length encoded to the
// unique pairs; one lane per unique pair then counts, and after a scan
// writes, its signed contributions (normal_collisions_builder.cpp:340-543).
#include "context.hpp"
#include "geometry.hpp"
#include "shard.hpp"
#include "sort.hpp"

#include <utility>

namespace ipcg {

namespace {

constexpr int NP_BLOCK = 256;

struct EERec {
    int32_t ea, eb, dtype;
    int32_t rank; // representative order: 0 candidate, 1 lower edge first, 2 otherwise
    double eps_x;
    double w;
};

// Sharded staging of the four collision streams: stream k's shard sh owns
// slots [sh * scap, (sh + 1) * scap).  A wave appends at most 64 entries per
// stream and shard sh serves every SHARDS-th wave, so scap =
// 64 * ceil(waves / SHARDS) never overflows.
struct Streams {
    uint64_t* vv;
    uint64_t* ev;
    EERec* ee;
    uint64_t* fv;
    double* vvw; // weights of the vertex-vertex / edge-vertex contributions
    double* evw; // (real-valued weights only, else null)
    unsigned long long* cnt; // 4 x SHARDS x SHARD_PAD
    int64_t scap;
};

// NormalCollisions::use_area_weighting: per-candidate weights
// (normal_collisions_builder.cpp:42-44, 81, 181-183, 265) from the mesh's
// vertex areas va and edge areas ea; null va: unit weights.
struct Areas {
    const double* va;
    const double* ea;
};

__device__ __forceinline__ uint64_t vv_key(int a, int b)
{
    const uint32_t lo = (uint32_t)(a < b ? a : b), hi = (uint32_t)(a < b ? b : a);
    return ((uint64_t)lo << 32) | hi;
}

__device__ __forceinline__ uint64_t pair_key(int a, int b)
{
    return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
}

// CollisionMesh adjacencies (rows sorted, duplicate-free)
struct Adj {
    const int32_t *vv_off, *vv; // vertex -> vertices sharing an edge
    const int32_t *ve_off, *ve; // vertex -> incident edges
    const int32_t *ev_off, *ev; // edge -> vertices opposite it in its faces
    const int32_t* vf_off;      // vertex -> incident faces (counts only)
    const uint8_t* bnd;         // is_vertex_on_boundary
};

__device__ __forceinline__ bool row_has(const int32_t* off, const int32_t* ids, int row, int v)
{
    int lo = off[row], hi = off[row + 1];
    while (lo < hi) {
        const int mid = lo + (hi - lo) / 2;
        if (ids[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo < off[row + 1] && ids[lo] == v;
}

// ---- OGC feasible regions (ogc/feasible_region.cpp) ----------------------
// Dot products over runtime-size vectors sum left to right, over fixed
// 3-vectors as a0 b0 + (a1 b1 + a2 b2), as the reference's Eigen types do.
__device__ __forceinline__ double dot_lr(const d3& a, const d3& b)
{
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}

__device__ __forceinline__ double dot_fx(const d3& a, const d3& b)
{
    return a.x * b.x + (a.y * b.y + a.z * b.z);
}

// closest_point.cpp:11-18 point_edge_closest_point
__device__ __forceinline__ double pe_param(const d3& p, const d3& e0, const d3& e1)
{
    const d3 e = sub(e1, e0);
    return dot_lr(sub(p, e0), e) / dot_lr(e, e);
}

__device__ __forceinline__ d3 on_edge(const d3& q0, const d3& q1, double a)
{
    return { (q1.x - q0.x) * a + q0.x, (q1.y - q0.y) * a + q0.y, (q1.z - q0.z) * a + q0.z };
}

// feasible_region.cpp:7-24: x behind vertex vi against every neighbour
__device__ bool ogc_vertex_region(const double* X, int nV, const Adj& J, const d3& x, int vi)
{
    const d3 xv = ldx(X, nV, vi);
    const d3 d = sub(x, xv);
    for (int t = J.vv_off[vi]; t < J.vv_off[vi + 1]; t++)
        if (!(dot_lr(d, sub(xv, ldx(X, nV, J.vv[t]))) >= 0)) return false;
    return true;
}

__device__ __forceinline__ bool ogc_vv_pair(const double* X, int nV, const Adj& J, int a, int b)
{
    return ogc_vertex_region(X, nV, J, ldx(X, nV, a), b)
        && ogc_vertex_region(X, nV, J, ldx(X, nV, b), a);
}

// feasible_region.cpp:26-54: vertex xi over edge ei, beyond the edge from
// every opposite vertex
__device__ bool ogc_edge_region(const double* X, int nV, const int32_t* E, const Adj& J, int xi,
                                int ei)
{
    const d3 x = ldx(X, nV, xi), x0 = ldx(X, nV, E[2 * (int64_t)ei]),
             x1 = ldx(X, nV, E[2 * (int64_t)ei + 1]);
    const double alpha = pe_param(x, x0, x1);
    if (alpha <= 0 || alpha >= 1) return false;
    for (int t = J.ev_off[ei]; t < J.ev_off[ei + 1]; t++) {
        const d3 x2 = ldx(X, nV, J.ev[t]);
        const d3 p = on_edge(x0, x1, pe_param(x2, x0, x1));
        if (!(dot_fx(sub(x, p), sub(p, x2)) >= 0)) return false;
    }
    return true;
}

// feasible_region.cpp:91-167 is_edge_edge_feasible (actual distance type)
__device__ bool ogc_ee_feasible(const double* X, int nV, const Adj& J, int a0, int a1, int b0,
                                int b1, int dt)
{
    switch (dt) {
    case EE_EA0_EB0: return ogc_vv_pair(X, nV, J, a0, b0);
    case EE_EA0_EB1: return ogc_vv_pair(X, nV, J, a0, b1);
    case EE_EA1_EB0: return ogc_vv_pair(X, nV, J, a1, b0);
    case EE_EA1_EB1: return ogc_vv_pair(X, nV, J, a1, b1);
    case EE_EA_EB: return true;
    default: break;
    }
    // an end of one edge against its closest point on the other
    const bool on_a = dt == EE_EA_EB0 || dt == EE_EA_EB1;
    const int v = on_a ? (dt == EE_EA_EB1 ? b1 : b0) : (dt == EE_EA1_EB ? a1 : a0);
    const d3 q0 = ldx(X, nV, on_a ? a0 : b0), q1 = ldx(X, nV, on_a ? a1 : b1);
    const d3 xc = on_edge(q0, q1, pe_param(ldx(X, nV, v), q0, q1));
    return ogc_vertex_region(X, nV, J, xc, v);
}

// candidates: [0, n_ee) edge-edge, then face-vertex, codim vertex-vertex,
// codim edge-vertex
// OGC: CollisionSetType::OGC (normal_collisions_builder.cpp:56-336 with
// use_ogc): infeasible edge-vertex / edge-edge / face-vertex pairs are
// dropped and pairs reduced to a vertex or edge are divided among its
// incident edges / faces / opposite vertices.
template <bool OGC>
__global__ __launch_bounds__(NP_BLOCK) void k_classify(
    const double* X, const double* R, int nV, const int32_t* E, const int32_t* F,
    const int32_t* F2E, const int2* ee, int64_t n_ee, const int2* fv, int64_t n_fv,
    const int2* vv, int64_t n_vv, const int2* ev, int64_t n_ev, double active2, Streams S,
    Areas A, Adj J)
{
    const int64_t i = (int64_t)xcd_block() * NP_BLOCK + threadIdx.x;
    int kind = -1; // 0 vv, 1 ev, 2 ee, 3 fv
    uint64_t key = 0;
    double w = 1.0; // weight of this candidate's contribution
    EERec er { 0, 0, 0, 0, 0.0, 0.0 };
    if (i < n_ee) {
        const int2 c = ee[i];
        if (A.va) w = 0.25 * (A.ea[c.x] + A.ea[c.y]);
        const int a0 = E[2 * (int64_t)c.x], a1 = E[2 * (int64_t)c.x + 1];
        const int b0 = E[2 * (int64_t)c.y], b1 = E[2 * (int64_t)c.y + 1];
        const d3 pa0 = ldx(X, nV, a0), pa1 = ldx(X, nV, a1), pb0 = ldx(X, nV, b0),
                 pb1 = ldx(X, nV, b1);
        const int actual = edge_edge_type(pa0, pa1, pb0, pb1);
        const double d = ee_dist2(pa0, pa1, pb0, pb1, actual);
        if (d < active2 && (!OGC || ogc_ee_feasible(X, nV, J, a0, a1, b0, b1, actual))) {
            const double eps_x = ee_mollifier_threshold(ldx(R, nV, a0), ldx(R, nV, a1),
                                                        ldx(R, nV, b0), ldx(R, nV, b1));
            const double x = ee_cross2(pa0, pa1, pb0, pb1);
            const int dt = x < eps_x ? (int)EE_EA_EB : actual;
            switch (dt) {
            case EE_EA0_EB0: kind = 0; key = vv_key(a0, b0); break;
            case EE_EA0_EB1: kind = 0; key = vv_key(a0, b1); break;
            case EE_EA1_EB0: kind = 0; key = vv_key(a1, b0); break;
            case EE_EA1_EB1: kind = 0; key = vv_key(a1, b1); break;
            case EE_EA_EB0: kind = 1; key = pair_key(c.x, b0); break;
            case EE_EA_EB1: kind = 1; key = pair_key(c.x, b1); break;
            case EE_EA0_EB: kind = 1; key = pair_key(c.y, a0); break;
            case EE_EA1_EB: kind = 1; key = pair_key(c.y, a1); break;
            default:
                kind = 2;
                er = { c.x, c.y, actual, 0, eps_x, w };
            }
        }
    } else if (i < n_ee + n_fv) {
        const int2 c = fv[i - n_ee];
        const int fi = c.x, vi = c.y;
        if (A.va) w = 0.25 * A.va[vi];
        const int f0 = F[3 * (int64_t)fi], f1 = F[3 * (int64_t)fi + 1],
                  f2 = F[3 * (int64_t)fi + 2];
        const d3 p = ldx(X, nV, vi), t0 = ldx(X, nV, f0), t1 = ldx(X, nV, f1),
                 t2 = ldx(X, nV, f2);
        const int dt = point_triangle_type(p, t0, t1, t2);
        const double d = pt_dist2(p, t0, t1, t2, dt);
        bool keep = d < active2;
        if (OGC && keep) { // feasible_region.cpp:169-224; builder.cpp:269-289
            const int fk = dt <= PT_P_T2 ? F[3 * (int64_t)fi + dt] : 0;
            const int ek = dt >= PT_P_E0 && dt <= PT_P_E2 ? F2E[3 * (int64_t)fi + dt - 3] : 0;
            keep = dt <= PT_P_T2 ? ogc_vv_pair(X, nV, J, vi, fk)
                : dt <= PT_P_E2  ? ogc_edge_region(X, nV, E, J, vi, ek)
                                 : true;
            const int n = dt <= PT_P_T2 ? J.vf_off[fk + 1] - J.vf_off[fk]
                : dt <= PT_P_E2         ? J.ev_off[ek + 1] - J.ev_off[ek]
                                        : 1;
            if (n > 1) w /= (double)n;
        }
        if (keep) {
            switch (dt) {
            case PT_P_T0: kind = 0; key = vv_key(vi, f0); break;
            case PT_P_T1: kind = 0; key = vv_key(vi, f1); break;
            case PT_P_T2: kind = 0; key = vv_key(vi, f2); break;
            case PT_P_E0: kind = 1; key = pair_key(F2E[3 * (int64_t)fi], vi); break;
            case PT_P_E1: kind = 1; key = pair_key(F2E[3 * (int64_t)fi + 1], vi); break;
            case PT_P_E2: kind = 1; key = pair_key(F2E[3 * (int64_t)fi + 2], vi); break;
            default: kind = 3; key = pair_key(fi, vi);
            }
        }
    } else if (i < n_ee + n_fv + n_vv) { // normal_collisions_builder.cpp:26-54
        const int2 c = vv[i - n_ee - n_fv];
        if (A.va) w = 0.5 * (A.va[c.x] + A.va[c.y]);
        if (sqn(sub(ldx(X, nV, c.x), ldx(X, nV, c.y))) < active2) {
            kind = 0;
            key = vv_key(c.x, c.y);
        }
    } else if (i < n_ee + n_fv + n_vv + n_ev) { // normal_collisions_builder.cpp:56-138
        const int2 c = ev[i - n_ee - n_fv - n_vv];
        const int ei = c.x, vi = c.y;
        if (A.va) w = 0.5 * A.va[vi];
        const int e0 = E[2 * (int64_t)ei], e1 = E[2 * (int64_t)ei + 1];
        const d3 p = ldx(X, nV, vi), q0 = ldx(X, nV, e0), q1 = ldx(X, nV, e1);
        const int dt = point_edge_type(p, q0, q1);
        bool keep = pe_dist2(p, q0, q1, dt) < active2;
        if (OGC && keep && dt != PE_P_E) { // feasible_region.cpp:56-89; builder.cpp:87-101
            const int vj = dt == PE_P_E0 ? e0 : e1;
            keep = ogc_vv_pair(X, nV, J, vi, vj);
            const int n = J.ve_off[vj + 1] - J.ve_off[vj];
            if (n > 1) w /= (double)n;
        }
        if (keep) {
            switch (dt) {
            case PE_P_E0: kind = 0; key = vv_key(vi, e0); break;
            case PE_P_E1: kind = 0; key = vv_key(vi, e1); break;
            default: kind = 1; key = pair_key(ei, vi);
            }
        }
    }
    // one wave-aggregated append per stream into this wave's shard
    const int sh = wave_shard();
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const Shards Sk { S.cnt + (int64_t)k * SHARDS * SHARD_PAD, S.scap };
        unsigned pre;
        const int64_t slot = (int64_t)shard_append_bits(Sk, sh, kind == k, false, pre) + pre;
        if (kind == k && slot < S.scap) {
            const int64_t at = sh * S.scap + slot;
            if (k == 0) {
                S.vv[at] = key;
                if (S.vvw) S.vvw[at] = w;
            } else if (k == 1) {
                S.ev[at] = key;
                if (S.evw) S.evw[at] = w;
            }
            else if (k == 2) S.ee[at] = er;
            else S.fv[at] = key;
        }
    }
}

__global__ void k_iota(uint32_t* v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

__global__ void k_ee_keys(const EERec* r, int64_t n, uint64_t* keys, uint32_t* vals)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        keys[i] = vv_key(r[i].ea, r[i].eb);
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_ee_gather(const EERec* r, const uint32_t* perm, int64_t n, int2* ids, double* w,
                            double* eps, int32_t* dtype)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const EERec e = r[perm[i]];
        ids[i] = make_int2(e.ea, e.eb);
        w[i] = e.w;
        eps[i] = e.eps_x;
        dtype[i] = e.dtype;
    }
}

__global__ void k_keys_to_pairs(const uint64_t* keys, int64_t n, Areas A, int2* ids, double* w)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const int vi = (int)(keys[i] & 0xFFFFFFFFu);
        ids[i] = make_int2((int)(keys[i] >> 32), vi);
        w[i] = A.va ? 0.25 * A.va[vi] : 1.0;
    }
}

__global__ void k_runs_to_weights(const uint64_t* ukeys, const uint32_t* start, const uint32_t* nu,
                                  int2* ids, double* w)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (int64_t)*nu) {
        ids[i] = make_int2((int)(ukeys[i] >> 32), (int)(ukeys[i] & 0xFFFFFFFFu));
        w[i] = (double)(start[i + 1] - start[i]);
    }
}

// Sum of wt(j), j in [s, e), taken in ascending order of value (equal
// values in any order): the same double whatever order the contributions
// arrived in.  Runs are a handful of entries, so a repeated minimum scan.
template <class W>
__device__ double sorted_sum(uint32_t s, uint32_t e, W wt)
{
    double sum = 0.0, last = 0.0;
    uint32_t done = 0;
    bool first = true;
    while (done < e - s) {
        double v = INFINITY;
        uint32_t c = 0;
        for (uint32_t j = s; j < e; j++) {
            const double x = wt(j);
            if (!first && !(x > last)) continue;
            if (c == 0 || x < v) {
                v = x;
                c = 1;
            } else if (x == v) {
                c++;
            }
        }
        if (c == 0) return NAN; // NaN contributions have no place in the order
        for (uint32_t t = 0; t < c; t++) sum += v;
        done += c;
        last = v;
        first = false;
    }
    return sum;
}

// real-valued merge of a vertex-vertex / edge-vertex run
__global__ void k_run_sums(const uint32_t* start, int64_t nu, const uint32_t* perm,
                           const double* cw, double* rw, uint32_t* keep)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nu) {
        const double w = sorted_sum(start[i], start[i + 1], [&](uint32_t j) { return cw[perm[j]]; });
        rw[i] = w;
        keep[i] = w != 0.0;
    }
}

__global__ void k_run_scatter(const uint64_t* ukeys, int64_t nu, const double* rw,
                              const uint32_t* keep, const uint32_t* pos, int2* ids, double* w)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nu && keep[i]) {
        ids[pos[i]] = make_int2((int)(ukeys[i] >> 32), (int)(ukeys[i] & 0xFFFFFFFFu));
        w[pos[i]] = rw[i];
    }
}

// edge-edge merge key: (min edge, max edge, distance type); edge ids < 2^30
__global__ void k_ee_typed_keys(const EERec* r, int64_t n, uint64_t* keys, uint32_t* vals)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint32_t a = (uint32_t)r[i].ea, b = (uint32_t)r[i].eb;
        const uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
        keys[i] = (lo << 34) | (hi << 4) | (uint64_t)r[i].dtype;
        vals[i] = (uint32_t)i;
    }
}

__global__ void k_ee_runs(const uint32_t* start, int64_t nu, const uint32_t* perm, const EERec* r,
                          double* rw, uint32_t* rep, uint32_t* keep)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nu) {
        const uint32_t s = start[i], e = start[i + 1];
        uint32_t best = perm[s];
        for (uint32_t j = s + 1; j < e; j++)
            if (r[perm[j]].rank < r[best].rank) best = perm[j];
        const double w = sorted_sum(s, e, [&](uint32_t j) { return r[perm[j]].w; });
        rw[i] = w;
        rep[i] = best;
        keep[i] = w != 0.0;
    }
}

__global__ void k_ee_scatter(const EERec* r, int64_t nu, const double* rw, const uint32_t* rep,
                             const uint32_t* keep, const uint32_t* pos, int2* ids, double* w,
                             double* eps, int32_t* dtype)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < nu && keep[i]) {
        const EERec e = r[rep[i]];
        const uint32_t k = pos[i];
        ids[k] = make_int2(e.ea, e.eb);
        w[k] = rw[i];
        eps[k] = e.eps_x;
        dtype[k] = e.dtype;
    }
}

// ---- IMPROVED_MAX_APPROX -------------------------------------------------

// conversion slot keys: 2-bit tag above a 62-bit pair key (ids < 2^30)
constexpr uint64_t CONV_NONE = ~0ull;
enum { CV_EV_VV = 0, CV_EE_EV = 1, CV_FV_EV = 2, CV_FV_VV = 3 };

__device__ __forceinline__ uint64_t conv_key(int tag, uint64_t k)
{
    return ((uint64_t)tag << 62) | k;
}

// Sub-element pairs: 2 per codim edge-vertex candidate (its edge's
// vertices, candidates.cpp:583-620), 4 per edge-edge candidate (each edge
// with the other's vertices, candidates.cpp:665-692), 6 per face-vertex
// candidate (the face's edges, candidates.cpp:640-663, then its vertices);
// a pair whose sub-element distance is active is appended to `out` (one
// atomic per wave), so the sort sees only the active pairs.  Every lane
// reaches the append (its count may be 0).
__global__ __launch_bounds__(NP_BLOCK) void k_conv_slots(
    const double* X, int nV, const int32_t* E, const int32_t* F, const int32_t* F2E, const int2* ev,
    int64_t n_ev, const int2* ee, int64_t n_ee, const int2* fv, int64_t n_fv, double active2,
    unsigned long long* count, uint64_t* out)
{
    const int64_t i = (int64_t)xcd_block() * NP_BLOCK + threadIdx.x;
    uint64_t k[6];
    unsigned nk = 0;
    if (i < n_ev) {
        const int2 c = ev[i];
        const d3 p = ldx(X, nV, c.y);
        for (int j = 0; j < 2; j++) {
            const int vj = E[2 * (int64_t)c.x + j];
            if (pp_dist2(p, ldx(X, nV, vj)) < active2) k[nk++] = conv_key(CV_EV_VV, vv_key(c.y, vj));
        }
    } else if (i < n_ev + n_ee) {
        const int2 c = ee[i - n_ev];
        for (int s = 0; s < 2; s++) {
            const int ei = s ? c.y : c.x, ej = s ? c.x : c.y;
            const d3 q0 = ldx(X, nV, E[2 * (int64_t)ei]), q1 = ldx(X, nV, E[2 * (int64_t)ei + 1]);
            for (int j = 0; j < 2; j++) {
                const int vj = E[2 * (int64_t)ej + j];
                const d3 p = ldx(X, nV, vj);
                if (pe_dist2(p, q0, q1, point_edge_type(p, q0, q1)) < active2)
                    k[nk++] = conv_key(CV_EE_EV, pair_key(ei, vj));
            }
        }
    } else if (i < n_ev + n_ee + n_fv) {
        const int2 c = fv[i - n_ev - n_ee];
        const d3 p = ldx(X, nV, c.y);
        for (int j = 0; j < 3; j++) {
            const int ei = F2E[3 * (int64_t)c.x + j];
            const d3 q0 = ldx(X, nV, E[2 * (int64_t)ei]), q1 = ldx(X, nV, E[2 * (int64_t)ei + 1]);
            if (pe_dist2(p, q0, q1, point_edge_type(p, q0, q1)) < active2)
                k[nk++] = conv_key(CV_FV_EV, pair_key(ei, c.y));
        }
        for (int j = 0; j < 3; j++) {
            const int vj = F[3 * (int64_t)c.x + j];
            if (pp_dist2(p, ldx(X, nV, vj)) < active2) k[nk++] = conv_key(CV_FV_VV, vv_key(c.y, vj));
        }
    }
    const unsigned long long at = wave_append(count, nk);
    for (unsigned j = 0; j < nk; j++) out[at + j] = k[j];
}

__global__ void k_iota_u32(uint32_t* v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) v[i] = (uint32_t)i;
}

struct ImOut {
    uint64_t* vv;
    double* vvw;
    uint64_t* ev;
    double* evw;
    EERec* ee;
    const uint32_t* off; // 3 x cap exclusive offsets (vv, ev, ee)
};

// One lane per unique converted pair: its signed contributions.  COUNT
// writes the numbers of vertex-vertex (0/1), edge-vertex (0/1) and
// edge-edge records to cnt; !COUNT writes the records at the scanned
// offsets, at most as many as counted.
template <bool COUNT>
__global__ __launch_bounds__(NP_BLOCK) void k_conv_eval(
    const uint64_t* ukeys, int64_t nu, const double* X, const double* R, int nV, const int32_t* E,
    Adj J, Areas A, uint32_t* cnt, int64_t cap, ImOut O)
{
    const int64_t i = (int64_t)xcd_block() * NP_BLOCK + threadIdx.x;
    if (i >= nu) return;
    const uint64_t key = ukeys[i];
    uint32_t nvv = 0, nev = 0, nee = 0;
    // writes stay within what the counting pass reserved
    const uint32_t lim_vv = COUNT ? 0 : cnt[i], lim_ev = COUNT ? 0 : cnt[cap + i],
                   lim_ee = COUNT ? 0 : cnt[2 * cap + i];
    // an edge-vertex pair reduced by its point-edge type
    // (normal_collisions_builder.cpp:116-138)
    auto typed = [&](int ei, int vi, int dt, double w) {
        if (dt == PE_P_E) {
            if (!COUNT && lim_ev) {
                O.ev[O.off[cap + i]] = pair_key(ei, vi);
                O.evw[O.off[cap + i]] = w;
            }
            nev = 1;
        } else {
            if (!COUNT && lim_vv) {
                O.vv[O.off[i]] = vv_key(vi, E[2 * (int64_t)ei + (dt == PE_P_E1)]);
                O.vvw[O.off[i]] = w;
            }
            nvv = 1;
        }
    };
    if (key != CONV_NONE) {
        const int tag = (int)(key >> 62);
        const int a = (int)((key >> 32) & 0x3FFFFFFFu), b = (int)(key & 0xFFFFFFFFu);
        if (tag == CV_EV_VV || tag == CV_FV_VV) {
            // normal_collisions_builder.cpp:340-383 (negative) and 385-425
            // (positive); a < b, both orientations summed
            double w = 0.0;
            for (int s = 0; s < 2; s++) {
                const int vi = s ? b : a, vj = s ? a : b;
                const bool inc_i = row_has(J.vv_off, J.vv, vj, vi);
                if (tag == CV_EV_VV) {
                    const int inc = (J.vv_off[vj + 1] - J.vv_off[vj]) - (int)inc_i;
                    if (inc > 1) w += (double)(1 - inc) * (A.va ? 0.5 * A.va[vi] : 1.0);
                } else if (!J.bnd[vj] && !inc_i) {
                    w += A.va ? 0.25 * A.va[vi] : 1.0;
                }
            }
            if (w != 0) {
                if (!COUNT && lim_vv) {
                    O.vv[O.off[i]] = vv_key(a, b);
                    O.vvw[O.off[i]] = w;
                }
                nvv = 1;
            }
        } else if (tag == CV_FV_EV) { // normal_collisions_builder.cpp:427-461
            const int ei = a, vi = b;
            const int inc = (J.ev_off[ei + 1] - J.ev_off[ei]) - (int)row_has(J.ev_off, J.ev, ei, vi);
            if (inc > 1) {
                const double w = (double)(1 - inc) * (A.va ? 0.25 * A.va[vi] : 1.0);
                typed(ei, vi, point_edge_type(ldx(X, nV, vi), ldx(X, nV, E[2 * (int64_t)ei]),
                                              ldx(X, nV, E[2 * (int64_t)ei + 1])),
                      w);
            }
        } else { // CV_EE_EV: normal_collisions_builder.cpp:463-543
            const int ea = a, p = b;
            const int a0 = E[2 * (int64_t)ea], a1 = E[2 * (int64_t)ea + 1];
            const double w = A.va ? -0.25 * A.ea[ea] : -1.0;
            const d3 pa0 = ldx(X, nV, a0), pa1 = ldx(X, nV, a1);
            const int dt = point_edge_type(ldx(X, nV, p), pa0, pa1);
            const d3 ra0 = ldx(R, nV, a0), ra1 = ldx(R, nV, a1);
            int nonmollified = 0;
            for (int t = J.ve_off[p]; t < J.ve_off[p + 1]; t++) {
                const int eb = J.ve[t];
                const int b0 = E[2 * (int64_t)eb], b1 = E[2 * (int64_t)eb + 1];
                const int q = p == b0 ? b1 : b0;
                if (q == a0 || q == a1) continue;
                const double eps_x = ee_mollifier_threshold(ra0, ra1, ldx(R, nV, b0), ldx(R, nV, b1));
                const double x = ee_cross2(pa0, pa1, ldx(X, nV, b0), ldx(X, nV, b1));
                if (x >= eps_x) {
                    nonmollified++;
                    continue;
                }
                const int edt = dt == PE_P_E0 ? (p == b0 ? EE_EA0_EB0 : EE_EA0_EB1)
                    : dt == PE_P_E1           ? (p == b0 ? EE_EA1_EB0 : EE_EA1_EB1)
                                              : (p == b0 ? EE_EA_EB0 : EE_EA_EB1);
                if (!COUNT && nee < lim_ee)
                    O.ee[O.off[2 * cap + i] + nee] = { ea, eb, edt, ea < eb ? 1 : 2, eps_x, w };
                nee++;
            }
            if (nonmollified != 1) typed(ea, p, dt, (double)(nonmollified - 1) * w);
        }
    }
    if (COUNT) {
        cnt[i] = nvv;
        cnt[cap + i] = nev;
        cnt[2 * cap + i] = nee;
    }
}

int key_bits_for(int64_t max_id)
{
    int b = 1;
    while ((1ll << b) <= max_id) b++;
    return 32 + b;
}

} // namespace

static Adj adjacency(const Ctx& c)
{
    return { c.adj_vv_off.as<int32_t>(), c.adj_vv.as<int32_t>(), c.adj_ve_off.as<int32_t>(),
             c.adj_ve.as<int32_t>(),     c.adj_ev_off.as<int32_t>(), c.adj_ev.as<int32_t>(),
             c.adj_vf_off.as<int32_t>(), c.adj_bnd.as<uint8_t>() };
}

// read a device u32 through the pinned staging buffer
static uint32_t read_u32(Ctx& c, const uint32_t* dev)
{
    uint32_t* h = static_cast<uint32_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, dev, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    IPCG_HIP(hipStreamSynchronize(c.stream));
    return h[0];
}

// IMPROVED_MAX_APPROX contributions (normal_collisions.cpp:84-128) into
// c.im_vv / im_vvw, c.im_ev / im_evw and c.im_ee; x = their counts
static void improved_records(Ctx& c, double active2, Areas A, int64_t x[3])
{
    hipStream_t s = c.stream;
    x[0] = x[1] = x[2] = 0;
    const int64_t nc = c.n_ev + c.n_ee + c.n_fv;
    const int64_t nslot = 2 * c.n_ev + 4 * c.n_ee + 6 * c.n_fv;
    if (nslot == 0) return;
    if (nslot >= (int64_t)UINT32_MAX || c.nV >= (1 << 30) || c.nE >= (1 << 30))
        throw Error(IPCG_ERR_ARG, "mesh or candidate set too large for the improved collision set");
    // worst case: every sub-element pair active; the sort buffers are
    // sized to the pairs that are
    c.im_slot.reserve(sizeof(uint64_t) * (size_t)nslot);
    c.counters.reserve(sizeof(unsigned long long) * 16);
    unsigned long long* n_act_dev = c.counters.as<unsigned long long>() + 6;
    uint32_t* nu_dev = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 8);
    uint32_t* tot_dev = nu_dev + 2; // three totals
    uint64_t* slot = c.im_slot.as<uint64_t>();
    IPCG_HIP(hipMemsetAsync(n_act_dev, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_conv_slots, dim3(grid_for(nc, NP_BLOCK, INT32_MAX)), dim3(NP_BLOCK), 0, s,
                       c.X[0].as<double>(), c.nV, c.E.as<int32_t>(), c.F.as<int32_t>(),
                       c.F2E.as<int32_t>(), c.cand_ev.as<int2>(), c.n_ev, c.cand_ee.as<int2>(),
                       c.n_ee, c.cand_fv.as<int2>(), c.n_fv, active2, n_act_dev, slot);
    IPCG_CHECK_LAUNCH();
    unsigned long long* hc = static_cast<unsigned long long*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(hc, n_act_dev, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    const int64_t nact = (int64_t)hc[0];
    if (nact > nslot) throw Error(IPCG_ERR_NUMERIC, "sub-element pair count exceeds its bound");
    if (nact == 0) return;
    c.im_slot2.reserve(sizeof(uint64_t) * (size_t)nact);
    c.im_vals.reserve(sizeof(uint32_t) * (size_t)nact);
    c.im_vals2.reserve(sizeof(uint32_t) * (size_t)nact);
    c.im_ukeys.reserve(sizeof(uint64_t) * (size_t)nact);
    c.im_start.reserve(sizeof(uint32_t) * ((size_t)nact + 1));
    c.np_tmp.reserve(std::max({ radix_tmp_bytes(nact), rle_tmp_bytes(nact), scan_tmp_bytes(nact) }));
    hipLaunchKernelGGL(k_iota_u32, dim3(grid_for(nact, 256, INT32_MAX)), dim3(256), 0, s,
                       c.im_vals.as<uint32_t>(), nact);
    IPCG_CHECK_LAUNCH();
    radix_sort_pairs(slot, c.im_vals.as<uint32_t>(), c.im_slot2.as<uint64_t>(),
                     c.im_vals2.as<uint32_t>(), nact, 64, c.np_tmp.p, s);
    run_length_encode(slot, nact, c.im_ukeys.as<uint64_t>(), c.im_start.as<uint32_t>(), nu_dev,
                      c.np_tmp.p, s);
    const int64_t nu = read_u32(c, nu_dev);
    c.im_cnt.reserve(sizeof(uint32_t) * 3 * (size_t)nu);
    c.im_off.reserve(sizeof(uint32_t) * 3 * (size_t)nu);
    uint32_t* cnt = c.im_cnt.as<uint32_t>();
    const Adj J = adjacency(c);
    ImOut O { nullptr, nullptr, nullptr, nullptr, nullptr, c.im_off.as<uint32_t>() };
    const unsigned g = grid_for(nu, NP_BLOCK, INT32_MAX);
    hipLaunchKernelGGL(k_conv_eval<true>, dim3(g), dim3(NP_BLOCK), 0, s, c.im_ukeys.as<uint64_t>(),
                       nu, c.X[0].as<double>(), c.rest.as<double>(), c.nV, c.E.as<int32_t>(), J, A,
                       cnt, nu, O);
    IPCG_CHECK_LAUNCH();
    for (int k = 0; k < 3; k++)
        exclusive_scan_u32(cnt + k * nu, c.im_off.as<uint32_t>() + k * nu, nu, tot_dev + k,
                           c.np_tmp.p, s);
    uint32_t* h = static_cast<uint32_t*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(h, tot_dev, 3 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < 3; k++) x[k] = h[k];
    c.im_vv.reserve(sizeof(uint64_t) * (size_t)std::max<int64_t>(x[0], 1));
    c.im_vvw.reserve(sizeof(double) * (size_t)std::max<int64_t>(x[0], 1));
    c.im_ev.reserve(sizeof(uint64_t) * (size_t)std::max<int64_t>(x[1], 1));
    c.im_evw.reserve(sizeof(double) * (size_t)std::max<int64_t>(x[1], 1));
    c.im_ee.reserve(sizeof(EERec) * (size_t)std::max<int64_t>(x[2], 1));
    O = { c.im_vv.as<uint64_t>(), c.im_vvw.as<double>(), c.im_ev.as<uint64_t>(),
          c.im_evw.as<double>(), c.im_ee.as<EERec>(), c.im_off.as<uint32_t>() };
    hipLaunchKernelGGL(k_conv_eval<false>, dim3(g), dim3(NP_BLOCK), 0, s,
                       c.im_ukeys.as<uint64_t>(), nu, c.X[0].as<double>(), c.rest.as<double>(),
                       c.nV, c.E.as<int32_t>(), J, A, cnt, nu, O);
    IPCG_CHECK_LAUNCH();
}

void np_build(Ctx& c, double dhat, double dmin)
{
    c.require_mesh();
    c.require_pos(0);
    if (!c.cand_valid || c.cand_swept)
        throw Error(IPCG_ERR_STATE, "no static candidate set (run the static broad phase)");
    hipStream_t s = c.stream;
    c.coll_valid = false;
    c.coff_valid = false;
    const int64_t n = c.n_ee + c.n_fv + c.n_vv + c.n_ev;
    const int64_t waves = (n + WAVE - 1) / WAVE;
    const int64_t scap = std::max<int64_t>(WAVE * ((waves + SHARDS - 1) / SHARDS), WAVE);
    const int64_t cap = SHARDS * scap; // entries per stream
    const bool aw = c.area_weighting;
    const bool improved = c.coll_type == IPCG_COLLISIONS_IMPROVED_MAX_APPROX;
    const bool ogc = c.coll_type == IPCG_COLLISIONS_OGC;
    const bool real_w = aw || improved || ogc; // weights are not plain counts
    if (real_w && c.nE >= (1 << 30))
        throw Error(IPCG_ERR_ARG, "edge ids beyond 2^30 do not fit the edge-edge merge key");
    c.np_raw.reserve((3 * sizeof(uint64_t) + sizeof(EERec) + (real_w ? 2 * sizeof(double) : 0))
                         * (size_t)cap + 256);
    char* base = c.np_raw.as<char>();
    Streams S;
    S.vv = reinterpret_cast<uint64_t*>(base);
    S.ev = S.vv + cap;
    S.fv = S.ev + cap;
    S.ee = reinterpret_cast<EERec*>(S.fv + cap);
    S.vvw = real_w ? reinterpret_cast<double*>(S.ee + cap) : nullptr;
    S.evw = real_w ? S.vvw + cap : nullptr;
    const Areas A { aw ? c.v_area.as<double>() : nullptr, aw ? c.e_area.as<double>() : nullptr };
    c.counters.reserve(sizeof(unsigned long long) * 16);
    c.shard_cnt.reserve(4 * SHARD_CNT_BYTES);
    c.shard_off.reserve(sizeof(int64_t) * 4 * (SHARDS + 1));
    S.cnt = c.shard_cnt.as<unsigned long long>();
    S.scap = scap;
    IPCG_HIP(hipMemsetAsync(S.cnt, 0, 4 * SHARD_CNT_BYTES, s));
    const double off = dmin + dhat;
    c.stage_begin("np_classify");
    if (n > 0)
        hipLaunchKernelGGL(ogc ? k_classify<true> : k_classify<false>,
                           dim3(grid_for(n, NP_BLOCK, INT32_MAX)), dim3(NP_BLOCK), 0, s,
                           c.X[0].as<double>(), c.rest.as<double>(), c.nV, c.E.as<int32_t>(),
                           c.F.as<int32_t>(), c.F2E.as<int32_t>(), c.cand_ee.as<int2>(), c.n_ee,
                           c.cand_fv.as<int2>(), c.n_fv, c.cand_vv.as<int2>(), c.n_vv,
                           c.cand_ev.as<int2>(), c.n_ev, off * off, S, A, adjacency(c));
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    ShardCounts kc[4];
    int64_t nraw[4];
    c.shard_sum.reserve(256);
    shard_counts_n(S.cnt, 4, c.shard_sum.as<unsigned long long>(), c.pinned, s, kc);
    for (int k = 0; k < 4; k++) {
        if (kc[k].max > scap) throw Error(IPCG_ERR_NUMERIC, "collision stream outgrew its shard");
        nraw[k] = kc[k].total;
    }
    int64_t x[3] = { 0, 0, 0 }; // improved vv, ev, ee records
    if (improved) {
        c.stage_begin("np_improved");
        improved_records(c, off * off, A, x);
        c.stage_end();
    }

    c.stage_begin("np_merge");
    int64_t mtot[4];
    int64_t mx = 1;
    for (int k = 0; k < 4; k++) {
        mtot[k] = nraw[k] + (k < 3 ? x[k] : 0);
        mx = std::max(mx, mtot[k]);
    }
    c.np_keys.reserve(sizeof(uint64_t) * mx);
    c.np_keys2.reserve(sizeof(uint64_t) * mx);
    c.np_vals.reserve(sizeof(uint32_t) * mx);
    c.np_vals2.reserve(sizeof(uint32_t) * mx);
    c.np_ukeys.reserve(sizeof(uint64_t) * mx);
    c.np_start.reserve(sizeof(uint32_t) * (mx + 1));
    c.np_tmp.reserve(std::max({ radix_tmp_bytes(mx), rle_tmp_bytes(mx), scan_tmp_bytes(mx) }));
    if (real_w) {
        c.np_rw.reserve(sizeof(double) * mx);
        c.np_flag.reserve(sizeof(uint32_t) * mx);
        c.np_pos.reserve(sizeof(uint32_t) * (mx + 1));
        c.np_rep.reserve(sizeof(uint32_t) * mx);
    }
    uint64_t* keys = c.np_keys.as<uint64_t>();
    uint32_t* vals = c.np_vals.as<uint32_t>();
    uint32_t* nu_dev = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 3);
    uint32_t* kept_dev = nu_dev + 1;
    // (id << 32 | id) keys: only the digits that hold id bits can vary, so
    // the sorts below run those passes without reading the keys back
    const int ibits = key_bits_for(std::max<int64_t>(std::max<int64_t>(c.nV, c.nE), c.nF)) - 32;
    const unsigned long long id_mask = (1ull << ibits) - 1ull;
    const unsigned long long pair_mask = id_mask | (id_mask << 32);
    auto sort_pairs = [&](int64_t m) {
        if (radix_sort_pairs_masked(keys, vals, c.np_keys2.as<uint64_t>(),
                                    c.np_vals2.as<uint32_t>(), m, pair_mask, c.np_tmp.p, s)) {
            std::swap(c.np_keys, c.np_keys2); // sorted pairs in the second buffers
            std::swap(c.np_vals, c.np_vals2);
            keys = c.np_keys.as<uint64_t>();
            vals = c.np_vals.as<uint32_t>();
        }
    };
    for (int k = 0; k < 4; k++) {
        const int64_t m0 = nraw[k], m = mtot[k];
        c.c_ids[k].reserve(sizeof(int2) * (size_t)std::max<int64_t>(m, 1));
        c.c_w[k].reserve(sizeof(double) * (size_t)std::max<int64_t>(m, 1));
        if (k == 2) {
            c.c_eps.reserve(sizeof(double) * (size_t)std::max<int64_t>(m, 1));
            c.c_dtype.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
        }
        c.ncoll[k] = 0;
        if (m == 0) continue;
        const unsigned g = grid_for(m, 256, INT32_MAX);
        const unsigned long long* kcnt = S.cnt + (int64_t)k * SHARDS * SHARD_PAD;
        int64_t* koff = c.shard_off.as<int64_t>() + k * (SHARDS + 1);
        if (k == 2) {
            c.np_eec.reserve(sizeof(EERec) * (size_t)m);
            EERec* ee = c.np_eec.as<EERec>();
            if (m0) shard_compact(S.ee, scap, kcnt, kc[k].max, ee, koff, s);
            if (!real_w) { // one collision per edge-edge candidate: no duplicates
                hipLaunchKernelGGL(k_ee_keys, dim3(g), dim3(256), 0, s, ee, m, keys, vals);
                sort_pairs(m);
                hipLaunchKernelGGL(k_ee_gather, dim3(g), dim3(256), 0, s, ee, vals, m,
                                   c.c_ids[2].as<int2>(), c.c_w[2].as<double>(),
                                   c.c_eps.as<double>(), c.c_dtype.as<int32_t>());
                c.ncoll[2] = m;
            } else {
                if (x[2])
                    IPCG_HIP(hipMemcpyAsync(ee + m0, c.im_ee.p, sizeof(EERec) * (size_t)x[2],
                                            hipMemcpyDeviceToDevice, s));
                hipLaunchKernelGGL(k_ee_typed_keys, dim3(g), dim3(256), 0, s, ee, m, keys, vals);
                radix_sort_pairs(keys, vals, c.np_keys2.as<uint64_t>(), c.np_vals2.as<uint32_t>(),
                                 m, 64, c.np_tmp.p, s);
                run_length_encode(keys, m, c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(),
                                  nu_dev, c.np_tmp.p, s);
                const int64_t nu = read_u32(c, nu_dev);
                const unsigned gu = grid_for(nu, 256, INT32_MAX);
                hipLaunchKernelGGL(k_ee_runs, dim3(gu), dim3(256), 0, s, c.np_start.as<uint32_t>(),
                                   nu, vals, ee, c.np_rw.as<double>(), c.np_rep.as<uint32_t>(),
                                   c.np_flag.as<uint32_t>());
                exclusive_scan_u32(c.np_flag.as<uint32_t>(), c.np_pos.as<uint32_t>(), nu, kept_dev,
                                   c.np_tmp.p, s);
                hipLaunchKernelGGL(k_ee_scatter, dim3(gu), dim3(256), 0, s, ee, nu,
                                   c.np_rw.as<double>(), c.np_rep.as<uint32_t>(),
                                   c.np_flag.as<uint32_t>(), c.np_pos.as<uint32_t>(),
                                   c.c_ids[2].as<int2>(), c.c_w[2].as<double>(),
                                   c.c_eps.as<double>(), c.c_dtype.as<int32_t>());
                IPCG_CHECK_LAUNCH();
                c.ncoll[2] = read_u32(c, kept_dev);
            }
        } else {
            const uint64_t* src = k == 0 ? S.vv : (k == 1 ? S.ev : S.fv);
            if (m0) shard_compact(src, scap, kcnt, kc[k].max, keys, koff, s);
            double* cw = nullptr; // compacted contribution weights
            if (real_w && k < 2) {
                c.np_cw.reserve(sizeof(double) * (size_t)m);
                cw = c.np_cw.as<double>();
                if (m0) shard_compact(k == 0 ? S.vvw : S.evw, scap, kcnt, kc[k].max, cw, koff, s);
                if (x[k]) {
                    IPCG_HIP(hipMemcpyAsync(keys + m0, k == 0 ? c.im_vv.p : c.im_ev.p,
                                            sizeof(uint64_t) * (size_t)x[k], hipMemcpyDeviceToDevice, s));
                    IPCG_HIP(hipMemcpyAsync(cw + m0, k == 0 ? c.im_vvw.p : c.im_evw.p,
                                            sizeof(double) * (size_t)x[k], hipMemcpyDeviceToDevice, s));
                }
            }
            hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, s, vals, m);
            sort_pairs(m);
            if (k == 3) { // each (face, vertex) candidate is unique
                hipLaunchKernelGGL(k_keys_to_pairs, dim3(g), dim3(256), 0, s, keys, m, A,
                                   c.c_ids[3].as<int2>(), c.c_w[3].as<double>());
                c.ncoll[3] = m;
            } else {
                run_length_encode(keys, m, c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(),
                                  nu_dev, c.np_tmp.p, s);
                if (cw) {
                    const int64_t nu = read_u32(c, nu_dev);
                    const unsigned gu = grid_for(nu, 256, INT32_MAX);
                    hipLaunchKernelGGL(k_run_sums, dim3(gu), dim3(256), 0, s,
                                       c.np_start.as<uint32_t>(), nu, vals, cw,
                                       c.np_rw.as<double>(), c.np_flag.as<uint32_t>());
                    exclusive_scan_u32(c.np_flag.as<uint32_t>(), c.np_pos.as<uint32_t>(), nu,
                                       kept_dev, c.np_tmp.p, s);
                    hipLaunchKernelGGL(k_run_scatter, dim3(gu), dim3(256), 0, s,
                                       c.np_ukeys.as<uint64_t>(), nu, c.np_rw.as<double>(),
                                       c.np_flag.as<uint32_t>(), c.np_pos.as<uint32_t>(),
                                       c.c_ids[k].as<int2>(), c.c_w[k].as<double>());
                    IPCG_CHECK_LAUNCH();
                    c.ncoll[k] = read_u32(c, kept_dev);
                } else {
                    hipLaunchKernelGGL(k_runs_to_weights, dim3(g), dim3(256), 0, s,
                                       c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(), nu_dev,
                                       c.c_ids[k].as<int2>(), c.c_w[k].as<double>());
                    IPCG_CHECK_LAUNCH();
                    c.ncoll[k] = read_u32(c, nu_dev);
                }
            }
        }
        IPCG_CHECK_LAUNCH();
    }
    c.stage_end();
    c.coll_dmin = dmin;
    c.coll_valid = true;
    c.hess_valid = false;
    c.fric_valid = false; // a friction set belongs to the collision set it was built on
    c.pattern_valid = false;
    c.inc_valid = false;
}

} // namespace ipcg
