// Continuous collision detection over the swept candidate lists:
// Tight-Inclusion per candidate inside the IPC Toolkit strategy
// (tight_inclusion_ccd.cpp:33-336) and the step-size minimum
// (candidates.cpp:252-292).
//
// Three tiers share one device restatement of the algorithm (ti.hpp):
//  * LDS tier: one lane per candidate that survived the root pass, its level
//    lists in LDS (LDS_CAP packed boxes per list).  Almost every candidate
//    resolves here within a few levels.
//  * slab tier, for searches that outgrew the LDS lists: persistent lanes
//    pull candidates from an atomic counter and run the whole breadth-first
//    search in a private 3 x LANE_CAP box slab in HBM.
//  * wavefront path, for candidates whose levels outgrow the slab: all of
//    them advance one level per round of short kernels.  Every box of the
//    level is evaluated by its own lane (the inclusion test is the work),
//    the sequential per-level rules are resolved from each candidate's first
//    deciding boxes (found by per-box atomic minimums over the t-sorted
//    level), children are emitted through a scan and each candidate's next
//    level is ordered (a merge of the lower / upper t halves, a lane per
//    box).
// The running minimum lives in device memory as the bit pattern of a
// non-negative double (unsigned atomic min); every search hides outcomes at
// or above it, which leaves the final minimum independent of scheduling.
#include "context.hpp"
#include "shard.hpp"
#include "sort.hpp"
#include "ti.hpp"

#include <algorithm>
#include <mutex>

namespace ipcg {

namespace {

constexpr int CB = 64;
constexpr int LDS_CAP = 6;
constexpr int LANE_CAP = 256;
constexpr int LANES = 131072; // 8 waves per CU at this kernel's register use
// Wavefront batches: searches advanced together (one round of launches and
// one host check per level for the whole batch), from WF_BATCH, doubled
// after every batch that fits the box budget up to WF_BATCH_MAX.  Workloads
// with edges parallel and on top of each other send hundreds of thousands
// of searches here; small batches would pay the per-level round trip for
// every few thousand of them.
constexpr int WF_BATCH = 65536;
constexpr int WF_BATCH_MAX = 1 << 18;
constexpr int64_t WF_MAX_BOXES = 1ll << 26;
constexpr double SMALL_TOI = 1e-6; // tight_inclusion_ccd.hpp:14

struct CCDArgs {
    const double* X0;
    const double* X1;
    int nV;
    const int32_t* E;
    const int32_t* F;
    const int2* ee;
    int64_t n_ee;
    const int2* fv;
    int64_t n_fv;
    const int2* vv; // codim vertex pairs
    int64_t n_vv;
    const int2* ev; // (codim edge, codim vertex)
    int64_t n_ev;
    double min_distance, tolerance, rescale;
    long max_iter;
    bool bounded;          // hide outcomes >= running minimum
    double* bound;         // running minimum (double bits), one per scene
    const int32_t* vscene; // batched contexts: scene per vertex (null: one scene)
    double* toi_out;       // per-candidate toi (null: not recorded)
    int* any_hit;          // early exit flag (null: no early exit)
};

__device__ __forceinline__ void atomic_min_nonneg(double* p, double v)
{
    atomicMin(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ void report(const CCDArgs& A, int64_t ci, bool hit, double toi,
                                       int scene)
{
    if (hit) {
        if (A.bounded) atomic_min_nonneg(A.bound + scene, toi);
        if (A.any_hit) atomicOr(A.any_hit, 1);
    }
    if (A.toi_out) A.toi_out[ci] = hit ? toi : INFINITY;
}

// Strategy prologue (ccd_strategy + *_ccd of tight_inclusion_ccd.cpp):
// gathers the stencil, returns 1 (hit at toi = 0), 0 (miss) or -1 (search
// needed, q / eff / adj set); `scene` receives the candidate's scene.
__device__ int candidate_setup(const CCDArgs& A, int64_t ci, TIQuery& q, double& eff, double& adj,
                               int& scene)
{
    // candidates [0, n_ee) edge-edge, then face-vertex, codim vertex-vertex
    // and codim edge-vertex; point-point and point-edge run as degenerate
    // edge-edge queries (p0, p0, p1, p1) and (p, p, e0, e1)
    // (tight_inclusion_ccd.cpp:77-200)
    int ids[4];
    int kind; // 0 ee, 1 fv, 2 vv, 3 ev
    if (ci < A.n_ee) {
        kind = 0;
        const int2 c = A.ee[ci];
        ids[0] = A.E[2 * (int64_t)c.x];
        ids[1] = A.E[2 * (int64_t)c.x + 1];
        ids[2] = A.E[2 * (int64_t)c.y];
        ids[3] = A.E[2 * (int64_t)c.y + 1];
    } else if ((ci -= A.n_ee) < A.n_fv) {
        kind = 1;
        const int2 c = A.fv[ci];
        ids[0] = c.y;
        ids[1] = A.F[3 * (int64_t)c.x];
        ids[2] = A.F[3 * (int64_t)c.x + 1];
        ids[3] = A.F[3 * (int64_t)c.x + 2];
    } else if ((ci -= A.n_fv) < A.n_vv) {
        kind = 2;
        const int2 c = A.vv[ci];
        ids[0] = ids[1] = c.x;
        ids[2] = ids[3] = c.y;
    } else {
        kind = 3;
        const int2 c = A.ev[ci - A.n_vv];
        ids[0] = ids[1] = c.y;
        ids[2] = A.E[2 * (int64_t)c.x];
        ids[3] = A.E[2 * (int64_t)c.x + 1];
    }
    const bool is_ee = kind != 1;
    scene = A.vscene ? A.vscene[ids[0]] : 0;
    d3 p0[4];
    bool moving = false;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        p0[j] = ldx(A.X0, A.nV, ids[j]);
        const d3 p1 = ldx(A.X1, A.nV, ids[j]);
        q.s[3 * j] = p0[j].x;
        q.s[3 * j + 1] = p0[j].y;
        q.s[3 * j + 2] = p0[j].z;
        q.e[3 * j] = p1.x;
        q.e[3 * j + 1] = p1.y;
        q.e[3 * j + 2] = p1.z;
        moving |= p1.x != p0[j].x || p1.y != p0[j].y || p1.z != p0[j].z;
    }
    q.vf = !is_ee;
    double d2;
    if (kind == 0)
        d2 = ee_dist2(p0[0], p0[1], p0[2], p0[3], edge_edge_type(p0[0], p0[1], p0[2], p0[3]));
    else if (kind == 1)
        d2 = pt_dist2(p0[0], p0[1], p0[2], p0[3], point_triangle_type(p0[0], p0[1], p0[2], p0[3]));
    else if (kind == 2)
        d2 = sqn(sub(p0[0], p0[2]));
    else
        d2 = pe_dist2(p0[0], p0[2], p0[3], point_edge_type(p0[0], p0[2], p0[3]));
    const double d0 = sqrt(d2);
    const double dmin = A.min_distance;
    if (!(d0 > dmin)) return 1; // check_initial_distance.hpp:7-21 (also without motion)
    if (!moving) return 0;
    adj = 0.5 * d0 < A.tolerance ? 0.5 * d0 : A.tolerance;
    eff = (1.0 - A.rescale) * (d0 - dmin);
    eff = eff < 1e-4 ? eff : 1e-4;
    eff += dmin;
    return -1;
}

// ---------------------------------------------------------------------------
// root pass: one lane per candidate settles the strategy prologue and the
// first level of the search (the inclusion test of the whole box [0,1]^3);
// only candidates whose root box includes the origin go on to a search.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ccd_root(CCDArgs A, Shards S, uint32_t* stage)
{
    const int64_t ci = (int64_t)xcd_block() * 256 + threadIdx.x;
    bool search = false;
    if (ci < A.n_ee + A.n_fv + A.n_vv + A.n_ev) {
        TIQuery q;
        double eff = 0, adj = 0;
        int sc = 0;
        const int pre = candidate_setup(A, ci, q, eff, adj, sc);
        if (pre >= 0) {
            report(A, ci, pre == 1, 0.0, sc);
        } else {
            ti_filter(q, eff > 0);
            bool box_in;
            double w[3];
            search = ti_inclusion(q, TIBox { 0.0, 0.0, 0.0 }, 1.0, 1.0, 1.0, eff, box_in, w);
            if (!search) report(A, ci, false, INFINITY, sc);
        }
    }
    const int sh = wave_shard();
    unsigned pre;
    const int64_t slot = (int64_t)shard_append_bits(S, sh, search, false, pre) + pre;
    if (search && slot < S.scap) stage[sh * S.scap + slot] = (uint32_t)ci;
}

// ---------------------------------------------------------------------------
// LDS tier
// ---------------------------------------------------------------------------
// ci_lo, ci_hi: only the candidates of this index range are searched (a
// phase of the search; the others return at once)
__global__ __launch_bounds__(CB) void k_ccd_lds(CCDArgs A, const uint32_t* todo, int64_t n_todo,
                                                uint32_t* spill, unsigned long long* n_spill,
                                                int64_t ci_lo, int64_t ci_hi)
{
    __shared__ unsigned long long lists[3 * LDS_CAP * CB];
    const int64_t k = (int64_t)xcd_block() * CB + threadIdx.x;
    if (k >= n_todo) return;
    if (A.any_hit && __hip_atomic_load(A.any_hit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        return;
    const TILdsLists L { lists + threadIdx.x, CB, LDS_CAP };
    const int64_t ci = todo[k];
    if (ci < ci_lo || ci >= ci_hi) return;
    TIQuery q;
    double eff = 0, adj = 0, toi = INFINITY;
    int sc = 0;
    const int pre = candidate_setup(A, ci, q, eff, adj, sc);
    if (pre >= 0) {
        report(A, ci, pre == 1, 0.0, sc);
        return;
    }
    TIBound B1 { A.bounded ? A.bound + sc : nullptr, 1.0, SMALL_TOI };
    int r = ti_query_lds(q, eff, adj, A.max_iter, false, B1, L, toi);
    if (r == TI_HIT && toi < SMALL_TOI) {
        TIBound B2 { A.bounded ? A.bound + sc : nullptr, (1.0 / A.rescale) * (1 + 1e-12), 0.0 };
        r = ti_query_lds(q, A.min_distance, adj, -1, true, B2, L, toi);
        if (r == TI_HIT) toi *= A.rescale;
    }
    if (r >= TI_OVERFLOW) {
        spill[atomicAdd(n_spill, 1ull)] = (uint32_t)ci;
        // searches spilled for depth (the rest: for list length), counted
        // per wave
        const unsigned long long deep = __ballot(r == TI_DEEP);
        if (deep && (unsigned)lane_id() == (unsigned)(__ffsll((long long)deep) - 1))
            atomicAdd(n_spill + 1, (unsigned long long)__popcll(deep));
        return;
    }
    report(A, ci, r == TI_HIT, toi, sc);
}

// ---------------------------------------------------------------------------
// slab tier
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(CB) void k_ccd_lanes(CCDArgs A, const uint32_t* todo, int64_t n_todo,
                                                  TIBox* slab, unsigned long long* work,
                                                  uint32_t* heavy, unsigned long long* n_heavy)
{
    const int64_t lane = (int64_t)blockIdx.x * CB + threadIdx.x;
    TISlab S;
    S.cap = LANE_CAP;
    S.cur = slab + lane * 3 * (int64_t)LANE_CAP;
    S.lo = S.cur + LANE_CAP;
    S.hi = S.lo + LANE_CAP;
    for (;;) {
        if (A.any_hit && __hip_atomic_load(A.any_hit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            break;
        const int64_t k = (int64_t)atomicAdd(work, 1ull);
        if (k >= n_todo) break;
        const int64_t ci = todo[k];
        TIQuery q;
        double eff = 0, adj = 0, toi = INFINITY;
        int sc = 0;
        const int pre = candidate_setup(A, ci, q, eff, adj, sc);
        if (pre >= 0) {
            report(A, ci, pre == 1, 0.0, sc);
            continue;
        }
        TIBound B1 { A.bounded ? A.bound + sc : nullptr, 1.0, SMALL_TOI };
        int r = ti_query(q, eff, adj, A.max_iter, false, B1, S, toi);
        if (r == TI_HIT && toi < SMALL_TOI) {
            TIBound B2 { A.bounded ? A.bound + sc : nullptr, (1.0 / A.rescale) * (1 + 1e-12), 0.0 };
            r = ti_query(q, A.min_distance, adj, -1, true, B2, S, toi);
            if (r == TI_HIT) toi *= A.rescale;
        }
        if (r >= TI_OVERFLOW) {
            heavy[atomicAdd(n_heavy, 1ull)] = (uint32_t)ci;
            continue;
        }
        report(A, ci, r == TI_HIT, toi, sc);
    }
}

// ---------------------------------------------------------------------------
// wavefront path
// ---------------------------------------------------------------------------
struct WFBox {
    double t, u, v;
    int32_t owner, pad;
};

struct WFState {
    TIQuery q;
    int64_t ci;
    double ms, delta, adj, eff;
    double wt, wu, wv;
    double toi_skip, temp_toi, temp_out_tol, result;
    int64_t refine;
    long max_itr;
    int stage, nz_iter, levels, axis;
    int use_skip, done, hit, restart;
    int scene, pad_;
    int64_t cut;                    // boxes of the level before this one may split
    int64_t seg_start, seg_count;   // boxes of the current level
    int64_t next_start, next_count; // boxes of the next level
};

enum : uint8_t { F_ZERO = 1, F_BOXIN = 2, F_TOL = 4, F_SKIP = 8 };

__device__ void wf_begin_query(WFState& S, double ms, double delta, long max_itr)
{
    S.ms = ms;
    S.delta = delta;
    S.max_itr = max_itr;
    ti_param_tol(S.q, delta);
    ti_filter(S.q, ms > 0);
    S.wt = S.wu = S.wv = 1.0;
    S.toi_skip = 4;
    S.temp_toi = INFINITY;
    S.temp_out_tol = delta;
    S.use_skip = 0;
    S.refine = 0;
    S.levels = 0;
    S.restart = 1;
}

// per candidate and level: first box (level index) that stops the scan,
// first zero box out of tolerance, first zero box, first zero box past the
// iteration limit
constexpr int WF_EV = 4;
constexpr unsigned long long WF_NONE = ~0ull;

__global__ void k_wf_init(CCDArgs A, const uint32_t* heavy, int64_t first, int64_t nb,
                          WFState* st, WFBox* boxes, unsigned long long* ev)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    for (int k = 0; k < WF_EV; k++) ev[WF_EV * i + k] = WF_NONE;
    WFState S;
    S.ci = heavy[first + i];
    double eff = 0, adj = 0;
    const int pre = candidate_setup(A, S.ci, S.q, eff, adj, S.scene);
    S.done = 0;
    S.hit = 0;
    S.result = INFINITY;
    S.eff = eff;
    S.adj = adj;
    S.stage = 1;
    S.nz_iter = 0;
    S.axis = 0;
    S.seg_start = i;
    S.seg_count = 1;
    S.cut = i;
    S.next_start = S.next_count = 0;
    if (pre >= 0) { // resolved by the prologue (cannot happen for lane overflows)
        S.done = 1;
        S.hit = pre == 1;
        S.result = 0.0;
        S.restart = 0;
    } else {
        wf_begin_query(S, eff, adj, A.max_iter);
        S.restart = 0;
    }
    boxes[i] = { 0.0, 0.0, 0.0, (int32_t)i, 0 };
    st[i] = S;
}

__device__ __forceinline__ double wf_bound(const CCDArgs& A, const WFState& S)
{
    if (!A.bounded) return INFINITY;
    const double g = __hip_atomic_load(A.bound + S.scene, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
    if (S.stage == 1) return g > SMALL_TOI ? g : SMALL_TOI;
    return g * ((1.0 / A.rescale) * (1 + 1e-12));
}

// Every box of the level by its own lane: skipped (its candidate is done,
// or the box starts at or after the running minimum or the candidate's
// earliest tolerance box), else the inclusion test.  live[j] = not skipped
// (scanned for the iteration counts).  A level is sorted by t (the first
// level is one box; a t split merges its halves in order, other splits keep
// it); the per-level rules below rely on that and a level out of order is
// reported as an error.
// also clears the level's counters (lvl[0..1] for k_wf_nochild, nroot for
// k_wf_segments), which run after it on the same stream
__global__ void k_wf_eval(CCDArgs A, const WFState* st, const WFBox* boxes, int64_t n,
                          uint8_t* flags, uint32_t* live, int* error, unsigned long long* lvl,
                          unsigned long long* nroot)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j == 0) {
        lvl[0] = 0;
        lvl[1] = 0;
        *nroot = 0;
    }
    if (j >= n) return;
    const WFBox b = boxes[j];
    const WFState& S = st[b.owner];
    uint8_t f = 0;
    if (S.done || b.t >= wf_bound(A, S) || b.t >= S.toi_skip) {
        f = F_SKIP;
    } else {
        bool box_in;
        double tw[3];
        const TIBox tb { b.t, b.u, b.v };
        if (ti_inclusion(S.q, tb, S.wt, S.wu, S.wv, S.ms, box_in, tw)) {
            f = F_ZERO;
            if (box_in) f |= F_BOXIN;
            if (tw[0] <= S.delta && tw[1] <= S.delta && tw[2] <= S.delta) f |= F_TOL;
        }
    }
    flags[j] = f;
    live[j] = (f & F_SKIP) ? 0u : 1u;
    if (j > S.seg_start && boxes[j - 1].t > b.t) atomicOr(error, 2);
}

// lowest level index of the lanes whose `hit` is set, per run of one
// candidate inside the wave, into e (one atomic per run); called by every
// lane of the wave
__device__ __forceinline__ void wf_first(bool hit, int owner, int64_t j, unsigned long long* e)
{
    const unsigned long long m = __ballot(hit);
    const unsigned long long below = m & ((1ull << lane_id()) - 1ull);
    const int l = below ? 63 - __clzll((long long)below) : lane_id();
    const int prev_owner = __shfl(owner, l, WAVE);
    if (hit && (!below || prev_owner != owner)) atomicMin(e, (unsigned long long)j);
}

// The sequential per-level rules of ti_bfs, in parallel: the scan of a
// level in order ends at the first zero box that is within tolerance or
// inside the ε box (or any zero box once the level's widths are within
// tolerance: c1); it finishes the search there if no zero box before it was
// out of tolerance, else it sets the earliest tolerance box and everything
// after it is skipped.  The iteration limit counts the scanned boxes.
__global__ void k_wf_mark(const WFState* st, const WFBox* boxes, int64_t n, const uint8_t* flags,
                          const uint32_t* lpre, unsigned long long* ev)
{
    const int64_t j0 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool in_range = j0 < n;
    const int64_t j = in_range ? j0 : n - 1;
    const uint8_t f = in_range ? flags[j] : F_SKIP;
    const bool zero = !(f & F_SKIP) && (f & F_ZERO);
    const int owner = boxes[j].owner;
    const WFState& S = st[owner];
    const TIQuery& q = S.q;
    const bool c1 = S.wt <= q.tol[0] && S.wu <= q.tol[1] && S.wv <= q.tol[2];
    const bool tol = f & F_TOL, in = f & F_BOXIN;
    unsigned long long* e = ev + (int64_t)WF_EV * owner;
    bool past = false;
    if (zero && S.max_itr > 0) {
        const int64_t refine = S.refine + (int64_t)(lpre[j] - lpre[S.seg_start]) + 1;
        past = refine > S.max_itr;
    }
    wf_first(zero && (c1 || tol || in), owner, j, e + 0);
    wf_first(zero && !tol, owner, j, e + 1);
    wf_first(zero, owner, j, e + 2);
    wf_first(past, owner, j, e + 3);
}

// Ends the current search of a candidate with (hit, toi) and applies the
// strategy: retry with no_zero_toi, its halvings, or completion.
__device__ void wf_search_done(const CCDArgs& A, WFState& S, bool hit, double toi, double out_tol)
{
    if (S.stage == 1) {
        if (hit && toi < SMALL_TOI) {
            S.stage = 2;
            S.nz_iter = 0;
            wf_begin_query(S, A.min_distance, S.adj, -1);
            return;
        }
        S.done = 1;
        S.hit = hit;
        S.result = toi;
        return;
    }
    S.nz_iter++;
    if (hit && toi == 0 && S.nz_iter < TI_MAX_NO_ZERO_TOI_ITER && !(out_tol > S.delta)) {
        double ms = S.ms, delta = S.delta;
        if (10 * delta < ms) ms *= 0.5;
        else delta *= 0.5;
        wf_begin_query(S, ms, delta, -1);
        return;
    }
    S.done = 1;
    S.hit = hit;
    S.result = hit ? toi * A.rescale : toi;
}

// one lane per candidate: the outcome of the level from its first events
// (ti_bfs's rules); children may come only from boxes before S.cut
__global__ void k_wf_decide(CCDArgs A, WFState* st, int64_t nb, const WFBox* boxes,
                            const uint32_t* lpre, unsigned long long* ev, int* error)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    WFState& S = st[i]; // updated in place (one lane per search)
    S.restart = 0;
    const int64_t s0 = S.seg_start, s1 = S.seg_start + S.seg_count;
    unsigned long long* e = ev + (int64_t)WF_EV * i;
    const unsigned long long e_stop = e[0], e_bad = e[1], e_zero = e[2], e_max = e[3];
    for (int k = 0; k < WF_EV; k++) e[k] = WF_NONE;
    S.cut = s0;
    if (S.done) return;
    if (++S.levels > TI_MAX_LEVELS) {
        atomicOr(error, 1);
        S.done = 1;
        return;
    }
    const TIQuery& q = S.q;
    const bool c1 = S.wt <= q.tol[0] && S.wu <= q.tol[1] && S.wv <= q.tol[2];
    {
        int axis = 0;
        double best = -INFINITY;
        const double w3[3] = { S.wt, S.wu, S.wv };
        for (int k = 0; k < 3; k++) {
            const double r = w3[k] > q.tol[k] ? w3[k] / q.tol[k] : -INFINITY;
            if (r > best) {
                best = r;
                axis = k;
            }
        }
        S.axis = axis;
    }
    // scanned boxes up to and including level index j
    auto scanned = [&](int64_t j) { return (int64_t)(lpre[j] - lpre[s0]) + 1; };
    if (e_stop == WF_NONE && e_max == WF_NONE) { // every live box scanned
        S.refine += (int64_t)(lpre[s1] - lpre[s0]);
        S.cut = s1;
    } else if (e_stop < e_max || (e_stop == e_max && (c1 || e_bad >= e_stop))) {
        const int64_t j = (int64_t)e_stop;
        S.refine += scanned(j);
        const double t = boxes[j].t;
        if (c1 || e_bad >= e_stop) { // no zero box before j out of tolerance
            wf_search_done(A, S, true, t, S.delta);
        } else { // earliest tolerance box; the rest of the level is skipped
            S.toi_skip = t < S.toi_skip ? t : S.toi_skip;
            S.use_skip = 1;
            S.cut = j;
        }
    } else { // iteration limit: the level's first zero box
        S.refine += scanned((int64_t)e_max);
        const WFBox b = boxes[e_zero];
        bool bi;
        double tw[3];
        const TIBox tb { b.t, b.u, b.v };
        ti_inclusion(q, tb, S.wt, S.wu, S.wv, S.ms, bi, tw);
        double m = S.delta;
        for (int k = 0; k < 3; k++) m = tw[k] > m ? tw[k] : m;
        S.temp_toi = b.t;
        S.temp_out_tol = m;
        wf_search_done(A, S, true, S.temp_toi, S.temp_out_tol);
    }
}

// children of every box (split_and_push: the upper half, then the lower)
__global__ void k_wf_counts(const WFState* st, const WFBox* boxes, const uint8_t* flags,
                            int64_t n, uint8_t* cmask, uint32_t* cnt)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint8_t f = flags[j];
    uint8_t m = 0;
    if (!(f & F_SKIP) && (f & F_ZERO)) {
        const WFBox b = boxes[j];
        const WFState& S = st[b.owner];
        if (j < S.cut && !S.done && !S.restart) {
            if (S.axis == 0) {
                m = 3;
            } else {
                const double hu = S.axis == 1 ? b.u + 0.5 * S.wu : b.u;
                const double hv = S.axis == 2 ? b.v + 0.5 * S.wv : b.v;
                if (!S.q.vf || hu + hv <= 1) m |= 1; // upper half
                if (!S.q.vf || b.u + b.v <= 1) m |= 2; // lower half
            }
        }
    }
    cmask[j] = m;
    cnt[j] = (uint32_t)__popc(m);
}

// a search whose level left no children ends (at its earliest tolerance
// box, if any); then the level's counts for the host, in one read:
// lvl[0] searches still pending, lvl[1] of those restarting at a root box,
// lvl[2] the children's total
// off: the scan of the children counts, off[nbox] their total
__global__ void k_wf_nochild(CCDArgs A, WFState* st, int64_t nb, const uint32_t* off,
                             int64_t nbox, unsigned long long* lvl)
{
    const int64_t total_children = (int64_t)off[nbox];
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i == 0) lvl[2] = (unsigned long long)total_children;
    if (i >= nb) return;
    WFState& S = st[i];
    if (!S.done && !S.restart) {
        const int64_t s0 = S.seg_start, s1 = S.seg_start + S.seg_count;
        const int64_t a = s0 < nbox ? (int64_t)off[s0] : total_children;
        const int64_t b = s1 < nbox ? (int64_t)off[s1] : total_children;
        if (b == a)
            wf_search_done(A, S, S.use_skip != 0, S.use_skip ? S.toi_skip : INFINITY, S.delta);
    }
    if (!S.done) {
        atomicAdd(lvl, 1ull);
        if (S.restart) atomicAdd(lvl + 1, 1ull);
    }
}

__global__ void k_wf_emit(const WFState* st, const WFBox* boxes, const uint8_t* cmask,
                          const uint32_t* off, int64_t n, WFBox* out)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const uint8_t m = cmask[j];
    if (!m) return;
    const WFBox b = boxes[j];
    const WFState& S = st[b.owner];
    WFBox h = b;
    if (S.axis == 0) h.t = b.t + 0.5 * S.wt;
    else if (S.axis == 1) h.u = b.u + 0.5 * S.wu;
    else h.v = b.v + 0.5 * S.wv;
    uint32_t p = off[j];
    if (m & 1) out[p++] = h; // split_and_push pushes the upper half first
    if (m & 2) out[p] = b;
}

// next-level segment of every candidate; restarts get one root box after
// all children
__global__ void k_wf_segments(WFState* st, int64_t nb, const uint32_t* off, int64_t nbox,
                              unsigned long long* nroot, int64_t total_children, WFBox* out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    WFState& S = st[i];
    if (S.done) {
        S.next_start = 0;
        S.next_count = 0;
        return;
    }
    if (S.restart) {
        const int64_t p = total_children + (int64_t)atomicAdd(nroot, 1ull);
        out[p] = { 0.0, 0.0, 0.0, (int32_t)i, 0 };
        S.next_start = p;
        S.next_count = 1;
        return;
    }
    const int64_t s0 = S.seg_start, s1 = S.seg_start + S.seg_count;
    const int64_t a = off[s0], b = s1 < nbox ? (int64_t)off[s1] : total_children;
    S.next_start = a;
    S.next_count = b - a;
}

// Orders each candidate's next level, a lane per box: a t split
// interleaves (upper, lower) pairs of two t-sorted halves, merged with ties
// going to the upper half (it was pushed first): an upper box lands after
// the lower boxes strictly below it, a lower box after the upper boxes at
// or below it.  Other splits and root boxes are already in order.
__global__ void k_wf_order(const WFState* st, int64_t n, const WFBox* in, WFBox* out)
{
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const WFBox b = in[k];
    const WFState& S = st[b.owner];
    const bool was_split = !S.done && !S.restart;
    if (!was_split || S.axis != 0) {
        out[k] = b;
        return;
    }
    const int64_t s = S.next_start, half = S.next_count / 2, i = k - s;
    const bool upper = (i & 1) == 0;
    const int64_t own = i >> 1;
    // other half: element r at s + 2 r + (upper ? 1 : 0)
    const int64_t base = s + (upper ? 1 : 0);
    int64_t lo = 0, hi = half; // count of the other half before b
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        const double t = in[base + 2 * mid].t;
        const bool before = upper ? t < b.t : t <= b.t;
        if (before) lo = mid + 1;
        else hi = mid;
    }
    out[s + own + lo] = b;
}

// after the order: halve the split axis and make the next level current
__global__ void k_wf_advance(WFState* st, int64_t nb)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    WFState& S = st[i];
    if (!S.done && !S.restart) {
        if (S.axis == 0) S.wt *= 0.5;
        else if (S.axis == 1) S.wu *= 0.5;
        else S.wv *= 0.5;
    }
    S.seg_start = S.next_start;
    S.seg_count = S.next_count;
}

__global__ void k_wf_finish(CCDArgs A, const WFState* st, int64_t nb)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nb) return;
    const WFState& S = st[i];
    report(A, S.ci, S.hit != 0, S.result, S.scene);
}

// Runs one batch of heavy candidates to completion.  Returns false, with
// nothing reported, when a level exceeds the box budget.
bool wf_batch(Ctx& c, const CCDArgs& A, const uint32_t* heavy, int64_t first, int64_t nb,
              WFState* states, bool early_exit, int* any_hit, int* error,
              unsigned long long* nroot, unsigned long long* npend, unsigned long long* h)
{
    hipStream_t s = c.stream;
    int64_t nbox = nb;
    c.blk_keys.reserve(sizeof(WFBox) * (size_t)std::max<int64_t>(nbox, 1024));
    c.wf_ev.reserve(sizeof(unsigned long long) * WF_EV * (size_t)nb);
    unsigned long long* ev = c.wf_ev.as<unsigned long long>();
    WFBox* cur = reinterpret_cast<WFBox*>(c.blk_keys.p);
    hipLaunchKernelGGL(k_wf_init, dim3(grid_for(nb, 256)), dim3(256), 0, s, A, heavy, first, nb,
                       states, cur, ev);
    IPCG_CHECK_LAUNCH();
    for (;;) {
        if (early_exit) {
            IPCG_HIP(hipMemcpyAsync(h, any_hit, sizeof(int), hipMemcpyDeviceToHost, s));
            IPCG_HIP(hipStreamSynchronize(s));
            if (*reinterpret_cast<int*>(h)) break;
        }
        c.blk_vals.reserve(sizeof(uint8_t) * 2 * (size_t)std::max<int64_t>(nbox, 1) + 64);
        uint8_t* flags = c.blk_vals.as<uint8_t>();
        uint8_t* cmask = flags + ((nbox + 63) / 64) * 64;
        c.blk_start.reserve(sizeof(uint32_t) * (size_t)(nbox + 1));
        c.blk_vals2.reserve(sizeof(uint32_t) * (size_t)(nbox + 1));
        c.wf_live.reserve(sizeof(uint32_t) * (size_t)(nbox + 1));
        c.wf_lpre.reserve(sizeof(uint32_t) * (size_t)(nbox + 1));
        uint32_t* cnt = c.blk_start.as<uint32_t>();
        uint32_t* off = c.blk_vals2.as<uint32_t>();
        uint32_t* live = c.wf_live.as<uint32_t>();
        uint32_t* lpre = c.wf_lpre.as<uint32_t>();
        c.blk_tmp.reserve(scan_tmp_bytes(nbox + 1));
        const unsigned gb = grid_for(nbox, 256, INT32_MAX);
        hipLaunchKernelGGL(k_wf_eval, dim3(gb), dim3(256), 0, s, A, states, cur, nbox, flags, live,
                           error, npend, nroot);
        exclusive_scan_u32(live, lpre, nbox, lpre + nbox, c.blk_tmp.p, s);
        hipLaunchKernelGGL(k_wf_mark, dim3(gb), dim3(256), 0, s, states, cur, nbox, flags, lpre,
                           ev);
        hipLaunchKernelGGL(k_wf_decide, dim3(grid_for(nb, 256)), dim3(256), 0, s, A, states, nb,
                           cur, lpre, ev, error);
        hipLaunchKernelGGL(k_wf_counts, dim3(gb), dim3(256), 0, s, states, cur, flags, nbox, cmask,
                           cnt);
        exclusive_scan_u32(cnt, off, nbox, off + nbox, c.blk_tmp.p, s);
        hipLaunchKernelGGL(k_wf_nochild, dim3(grid_for(nb, 256)), dim3(256), 0, s, A, states, nb,
                           off, nbox, npend);
        IPCG_CHECK_LAUNCH();
        // one round trip per level: the searches still pending, those
        // restarting and the children's total
        IPCG_HIP(hipMemcpyAsync(h, npend, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                s));
        IPCG_HIP(hipStreamSynchronize(s));
        const int64_t pending = (int64_t)h[0], restarts = (int64_t)h[1];
        const int64_t children = (int64_t)h[2];
        if (pending == 0) break;
        const int64_t nnext = children + restarts; // a restart is one root box
        if (nnext > WF_MAX_BOXES) return false;
        c.blk_keys2.reserve(sizeof(WFBox) * (size_t)std::max<int64_t>(nnext, 1));
        c.blk_ukeys.reserve(sizeof(WFBox) * (size_t)std::max<int64_t>(nnext, 1));
        WFBox* nxt = reinterpret_cast<WFBox*>(c.blk_keys2.p);
        WFBox* ord = reinterpret_cast<WFBox*>(c.blk_ukeys.p);
        hipLaunchKernelGGL(k_wf_emit, dim3(gb), dim3(256), 0, s, states, cur, cmask, off, nbox, nxt);
        hipLaunchKernelGGL(k_wf_segments, dim3(grid_for(nb, 256)), dim3(256), 0, s, states, nb, off,
                           nbox, nroot, children, nxt);
        hipLaunchKernelGGL(k_wf_order, dim3(grid_for(nnext, 256, INT32_MAX)), dim3(256), 0, s,
                           states, nnext, nxt, ord);
        hipLaunchKernelGGL(k_wf_advance, dim3(grid_for(nb, 256)), dim3(256), 0, s, states, nb);
        IPCG_CHECK_LAUNCH();
        std::swap(c.blk_keys, c.blk_ukeys);
        cur = reinterpret_cast<WFBox*>(c.blk_keys.p);
        nbox = nnext;
    }
    hipLaunchKernelGGL(k_wf_finish, dim3(grid_for(nb, 256)), dim3(256), 0, s, A, states, nb);
    IPCG_CHECK_LAUNCH();
    return true;
}

// The slab tier's box lists (3 x LANE_CAP boxes per lane, up to LANES
// lanes: 2.4 GB) are one buffer per device shared by every context on it:
// a context holds the lock while it enqueues its slab launch behind the
// previous user's (stream event), so the memory does not grow with the
// number of contexts.
struct SharedSlab {
    std::mutex mu;
    DevBuf buf;
    hipEvent_t last = nullptr;
};
constexpr int SLAB_DEVICES = 64;
SharedSlab g_slab[SLAB_DEVICES];

SharedSlab& shared_slab(int device)
{
    if (device < 0 || device >= SLAB_DEVICES) throw Error(IPCG_ERR_ARG, "device index out of range");
    return g_slab[device];
}

} // namespace

__global__ void k_fill_f64(double* p, int n, double v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

double ccd_stepsize(Ctx& c, double min_distance, double tol, int64_t max_iter, double rescale,
                    bool early_exit, double* const* toi_out, double* scene_steps)
{
    c.require_mesh();
    c.require_pos(0);
    c.require_pos(1);
    // any candidate set (Candidates::compute_collision_free_stepsize takes
    // the lists as built, static or swept)
    if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set (run a broad phase)");
    hipStream_t s = c.stream;
    const int64_t n = c.n_ee + c.n_fv + c.n_vv + c.n_ev;
    const bool per_cand = toi_out != nullptr;
    const int nsc = c.batched() ? c.n_scenes : 1;
    if (n == 0) {
        if (scene_steps) std::fill(scene_steps, scene_steps + nsc, 1.0);
        return 1.0;
    }
    if (n >= (1ll << 32)) throw Error(IPCG_ERR_ARG, "too many CCD candidates");
    // device scalars: bound, work counter, heavy count, hit flag, error, roots
    c.ccd_status.reserve(512);
    char* st = c.ccd_status.as<char>();
    double* bound = reinterpret_cast<double*>(st);
    unsigned long long* work = reinterpret_cast<unsigned long long*>(st + 64);
    unsigned long long* n_heavy = reinterpret_cast<unsigned long long*>(st + 128);
    int* any_hit = reinterpret_cast<int*>(st + 192);
    int* error = reinterpret_cast<int*>(st + 256);
    unsigned long long* nroot = reinterpret_cast<unsigned long long*>(st + 320);
    unsigned long long* npend = reinterpret_cast<unsigned long long*>(st + 384); // [3]
    unsigned long long* n_spill = reinterpret_cast<unsigned long long*>(st + 448); // [2]
    const double one = 1.0;
    IPCG_HIP(hipMemsetAsync(st, 0, 512, s));
    if (c.batched()) { // one running minimum per scene
        c.ccd_sbound.reserve(sizeof(double) * (size_t)nsc);
        bound = c.ccd_sbound.as<double>();
        hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(nsc, 256)), dim3(256), 0, s, bound, nsc, 1.0);
        IPCG_CHECK_LAUNCH();
    } else {
        IPCG_HIP(hipMemcpyAsync(bound, &one, sizeof(double), hipMemcpyHostToDevice, s));
    }
    c.ccd_overflow.reserve(sizeof(uint32_t) * (size_t)n);
    uint32_t* heavy = c.ccd_overflow.as<uint32_t>();
    if (per_cand) c.ccd_toi.reserve(sizeof(double) * (size_t)n);

    CCDArgs A;
    A.X0 = c.X[0].as<double>();
    A.X1 = c.X[1].as<double>();
    A.nV = c.nV;
    A.E = c.E.as<int32_t>();
    A.F = c.F.as<int32_t>();
    A.ee = c.cand_ee.as<int2>();
    A.n_ee = c.n_ee;
    A.fv = c.cand_fv.as<int2>();
    A.n_fv = c.n_fv;
    A.vv = c.cand_vv.as<int2>();
    A.n_vv = c.n_vv;
    A.ev = c.cand_ev.as<int2>();
    A.n_ev = c.n_ev;
    A.min_distance = min_distance;
    A.tolerance = tol;
    A.rescale = rescale;
    A.max_iter = (long)max_iter;
    A.bounded = !per_cand;
    A.bound = bound;
    A.vscene = c.batched() ? c.vscene.as<int32_t>() : nullptr;
    A.toi_out = per_cand ? c.ccd_toi.as<double>() : nullptr;
    A.any_hit = early_exit ? any_hit : nullptr;
    unsigned long long* h = static_cast<unsigned long long*>(c.pinned);

    // root pass -> candidates that need a search (sharded, then packed)
    c.stage_begin("ccd_root");
    const int64_t waves = (n + WAVE - 1) / WAVE;
    const int64_t scap = std::max<int64_t>(WAVE * ((waves + SHARDS - 1) / SHARDS), WAVE);
    c.ccd_stage.reserve(sizeof(uint32_t) * (size_t)(SHARDS * scap));
    c.shard_cnt.reserve(SHARD_CNT_BYTES);
    c.shard_off.reserve(sizeof(int64_t) * (SHARDS + 1));
    IPCG_HIP(hipMemsetAsync(c.shard_cnt.p, 0, SHARD_CNT_BYTES, s));
    const Shards RS { c.shard_cnt.as<unsigned long long>(), scap };
    hipLaunchKernelGGL(k_ccd_root, dim3(grid_for(n, 256, INT32_MAX)), dim3(256), 0, s, A, RS,
                       c.ccd_stage.as<uint32_t>());
    IPCG_CHECK_LAUNCH();
    c.shard_sum.reserve(256);
    const ShardCounts rc = shard_counts(RS.cnt, c.shard_sum.as<unsigned long long>(), c.pinned, s);
    if (rc.max > scap) throw Error(IPCG_ERR_NUMERIC, "CCD root pass outgrew its shard");
    c.ccd_todo.reserve(sizeof(uint32_t) * (size_t)std::max<int64_t>(rc.total, 1));
    shard_compact(c.ccd_stage.as<uint32_t>(), scap, RS.cnt, rc.max, c.ccd_todo.as<uint32_t>(),
                  c.shard_off.as<int64_t>(), s);
    c.stage_end();
    const int64_t n_todo = rc.total;
    c.stat["ccd.candidates"] = n;
    c.stat["ccd.searched"] = n_todo;

    // One search of every surviving candidate of an index range: the LDS
    // tier, the slab tier for its overflow, then the wavefront for searches
    // that outgrew a slab.  The root pass's staging array is free again and
    // holds the spill list.
    auto search = [&](int64_t ci_lo, int64_t ci_hi) {
        uint32_t* spill = c.ccd_stage.as<uint32_t>();
        IPCG_HIP(hipMemsetAsync(n_spill, 0, 2 * sizeof(unsigned long long), s));
        IPCG_HIP(hipMemsetAsync(work, 0, sizeof(unsigned long long), s));
        IPCG_HIP(hipMemsetAsync(n_heavy, 0, sizeof(unsigned long long), s));
        c.stage_begin("ccd_lds");
        if (n_todo > 0) {
            hipLaunchKernelGGL(k_ccd_lds, dim3((unsigned)((n_todo + CB - 1) / CB)), dim3(CB), 0, s,
                               A, c.ccd_todo.as<uint32_t>(), n_todo, spill, n_spill, ci_lo,
                               ci_hi);
            IPCG_CHECK_LAUNCH();
        }
        c.stage_end();
        IPCG_HIP(hipMemcpyAsync(h, n_spill, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                s));
        IPCG_HIP(hipStreamSynchronize(s));
        const int64_t n2 = (int64_t)h[0];
        c.stat["ccd.spilled"] += n2;
        c.stat["ccd.spilled_deep"] += (int64_t)h[1];
        c.stage_begin("ccd_lanes");
        if (n2 > 0) {
            const int64_t ln = std::min<int64_t>(LANES, ((n2 + CB - 1) / CB) * CB);
            SharedSlab& P = shared_slab(c.device);
            std::lock_guard<std::mutex> lock(P.mu);
            // slabs for the lanes launched only; growing frees the old
            // buffer, which waits for the device to finish with it
            P.buf.reserve(sizeof(TIBox) * 3 * (size_t)LANE_CAP * (size_t)ln);
            if (!P.last) IPCG_HIP(hipEventCreateWithFlags(&P.last, hipEventDisableTiming));
            else IPCG_HIP(hipStreamWaitEvent(s, P.last, 0)); // the previous user's launch
            hipLaunchKernelGGL(k_ccd_lanes, dim3((unsigned)(ln / CB)), dim3(CB), 0, s, A, spill,
                               n2, P.buf.as<TIBox>(), work, heavy, n_heavy);
            IPCG_CHECK_LAUNCH();
            IPCG_HIP(hipEventRecord(P.last, s));
        }
        c.stage_end();
        IPCG_HIP(hipMemcpyAsync(h, n_heavy, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        IPCG_HIP(hipStreamSynchronize(s));
        const int64_t nh = (int64_t)h[0];
        c.stat["ccd.wavefront"] += nh;
        if (nh > 0) {
            c.stage_begin("ccd_wavefront");
            DevBuf& wst = c.ccd_bound; // per-candidate wavefront state
            wst.reserve(sizeof(WFState) * (size_t)std::min<int64_t>(nh, WF_BATCH_MAX));
            WFState* states = wst.as<WFState>();
            // Batches of heavy candidates advance together.  A batch whose
            // level outgrows the box budget is abandoned (results are only
            // reported at the end of a batch, so this has no side effect)
            // and retried with a quarter of the candidates.
            int64_t batch = std::min<int64_t>(WF_BATCH, nh);
            for (int64_t first = 0; first < nh;) {
                const int64_t nb = std::min<int64_t>(batch, nh - first);
                const bool ok = wf_batch(c, A, heavy, first, nb, states, early_exit, any_hit,
                                         error, nroot, npend, h);
                if (!ok) {
                    if (nb == 1)
                        throw Error(IPCG_ERR_NUMERIC, "CCD wavefront level exceeds its box budget");
                    batch = std::max<int64_t>(1, nb / 4);
                    continue;
                }
                first += nb;
                batch = std::min<int64_t>(WF_BATCH_MAX, 2 * batch);
            }
            c.stage_end();
            IPCG_HIP(hipMemcpyAsync(h, error, sizeof(int), hipMemcpyDeviceToHost, s));
            IPCG_HIP(hipStreamSynchronize(s));
            const int err = *reinterpret_cast<int*>(h);
            if (err & 2) throw Error(IPCG_ERR_NUMERIC, "CCD wavefront level out of t order");
            if (err) throw Error(IPCG_ERR_NUMERIC, "CCD root search exceeded the level limit");
        }
    };
    c.stat["ccd.wavefront"] = 0;
    c.stat["ccd.spilled"] = 0;
    c.stat["ccd.spilled_deep"] = 0;
    if (A.bounded) {
        // Face-vertex and codim candidates first: their roots are isolated
        // points, so they resolve fast and lower every scene's running
        // minimum before the edge-edge searches start, which parallel
        // overlapping edges can make long.  Outcomes below the minimum do not
        // depend on the order, so neither does the result.
        search(c.n_ee, n);
        search(0, c.n_ee);
    } else {
        search(0, n);
    }
    if (per_cand) { // toi_out: ee, fv, vv, ev (any null)
        const int64_t cnt[4] = { c.n_ee, c.n_fv, c.n_vv, c.n_ev };
        int64_t off = 0;
        for (int k = 0; k < 4; k++) {
            if (toi_out[k] && cnt[k])
                IPCG_HIP(hipMemcpyAsync(toi_out[k], c.ccd_toi.as<double>() + off,
                                        sizeof(double) * cnt[k], hipMemcpyDeviceToHost, s));
            off += cnt[k];
        }
        IPCG_HIP(hipStreamSynchronize(s));
        return 1.0;
    }
    double* hd = static_cast<double*>(c.pinned);
    IPCG_HIP(hipMemcpyAsync(hd, any_hit, sizeof(int), hipMemcpyDeviceToHost, s));
    std::vector<double> per_scene((size_t)nsc);
    IPCG_HIP(hipMemcpyAsync(per_scene.data(), bound, sizeof(double) * (size_t)nsc,
                            hipMemcpyDeviceToHost, s));
    IPCG_HIP(hipStreamSynchronize(s));
    if (early_exit) return reinterpret_cast<int*>(hd)[0] ? 0.0 : 1.0;
    if (scene_steps) std::copy(per_scene.begin(), per_scene.end(), scene_steps);
    return *std::min_element(per_scene.begin(), per_scene.end());
}

} // namespace ipcg
