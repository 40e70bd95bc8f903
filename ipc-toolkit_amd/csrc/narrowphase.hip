// Narrow phase: classify every candidate's distance type, test activity
// d^2 < (dhat + dmin)^2 and reduce it to a VV / EV / EE / FV collision
// (normal_collisions_builder.cpp:140-336), then merge duplicates summing
// their unit weights (normal_collisions_builder.cpp:547-600) by sorting the
// collision keys and run-length encoding them.  Output order is canonical:
// vv by (min, max), ev by (edge, vertex), ee by (min edge, max edge), fv by
// (face, vertex).
#include "context.hpp"
#include "geometry.hpp"
#include "shard.hpp"
#include "sort.hpp"

namespace ipcg {

namespace {

constexpr int NP_BLOCK = 256;

struct EERec {
    int32_t ea, eb, dtype, pad;
    double eps_x;
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
    double* evw; // (area weighting only, else null)
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

// candidates: [0, n_ee) edge-edge, then face-vertex, codim vertex-vertex,
// codim edge-vertex
__global__ __launch_bounds__(NP_BLOCK) void k_classify(
    const double* X, const double* R, int nV, const int32_t* E, const int32_t* F,
    const int32_t* F2E, const int2* ee, int64_t n_ee, const int2* fv, int64_t n_fv,
    const int2* vv, int64_t n_vv, const int2* ev, int64_t n_ev, double active2, Streams S,
    Areas A)
{
    const int64_t i = (int64_t)xcd_block() * NP_BLOCK + threadIdx.x;
    int kind = -1; // 0 vv, 1 ev, 2 ee, 3 fv
    uint64_t key = 0;
    double w = 1.0; // weight of this candidate's contribution
    EERec er { 0, 0, 0, 0, 0.0 };
    if (i < n_ee) {
        const int2 c = ee[i];
        if (A.va) w = 0.25 * (A.ea[c.x] + A.ea[c.y]);
        const int a0 = E[2 * (int64_t)c.x], a1 = E[2 * (int64_t)c.x + 1];
        const int b0 = E[2 * (int64_t)c.y], b1 = E[2 * (int64_t)c.y + 1];
        const d3 pa0 = ldx(X, nV, a0), pa1 = ldx(X, nV, a1), pb0 = ldx(X, nV, b0),
                 pb1 = ldx(X, nV, b1);
        const int actual = edge_edge_type(pa0, pa1, pb0, pb1);
        const double d = ee_dist2(pa0, pa1, pb0, pb1, actual);
        if (d < active2) {
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
                er = { c.x, c.y, actual, 0, eps_x };
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
        if (d < active2) {
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
        if (pe_dist2(p, q0, q1, dt) < active2) {
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

__global__ void k_ee_gather(const EERec* r, const uint32_t* perm, int64_t n, Areas A, int2* ids,
                            double* w, double* eps, int32_t* dtype)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const EERec e = r[perm[i]];
        ids[i] = make_int2(e.ea, e.eb);
        w[i] = A.va ? 0.25 * (A.ea[e.ea] + A.ea[e.eb]) : 1.0;
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

// area weighting: a run's weight is the sum of its contributions' weights
// in sorted order (the sort is stable over the compacted stream)
__global__ void k_runs_to_weight_sums(const uint64_t* ukeys, const uint32_t* start,
                                      const uint32_t* nu, const uint32_t* perm, const double* cw,
                                      int2* ids, double* w)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < (int64_t)*nu) {
        ids[i] = make_int2((int)(ukeys[i] >> 32), (int)(ukeys[i] & 0xFFFFFFFFu));
        double sum = 0.0;
        for (uint32_t j = start[i]; j < start[i + 1]; j++) sum += cw[perm[j]];
        w[i] = sum;
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

int key_bits_for(int64_t max_id)
{
    int b = 1;
    while ((1ll << b) <= max_id) b++;
    return 32 + b;
}

} // namespace

void np_build(Ctx& c, double dhat, double dmin)
{
    c.require_mesh();
    c.require_pos(0);
    if (!c.cand_valid || c.cand_swept)
        throw Error(IPCG_ERR_STATE, "no static candidate set (run the static broad phase)");
    hipStream_t s = c.stream;
    c.coll_valid = false;
    const int64_t n = c.n_ee + c.n_fv + c.n_vv + c.n_ev;
    const int64_t waves = (n + WAVE - 1) / WAVE;
    const int64_t scap = std::max<int64_t>(WAVE * ((waves + SHARDS - 1) / SHARDS), WAVE);
    const int64_t cap = SHARDS * scap; // entries per stream
    const bool aw = c.area_weighting;
    c.np_raw.reserve((3 * sizeof(uint64_t) + sizeof(EERec) + (aw ? 2 * sizeof(double) : 0))
                         * (size_t)cap + 256);
    char* base = c.np_raw.as<char>();
    Streams S;
    S.vv = reinterpret_cast<uint64_t*>(base);
    S.ev = S.vv + cap;
    S.fv = S.ev + cap;
    S.ee = reinterpret_cast<EERec*>(S.fv + cap);
    S.vvw = aw ? reinterpret_cast<double*>(S.ee + cap) : nullptr;
    S.evw = aw ? S.vvw + cap : nullptr;
    const Areas A { aw ? c.v_area.as<double>() : nullptr, aw ? c.e_area.as<double>() : nullptr };
    c.counters.reserve(sizeof(unsigned long long) * 8);
    c.shard_cnt.reserve(4 * SHARD_CNT_BYTES);
    c.shard_off.reserve(sizeof(int64_t) * 4 * (SHARDS + 1));
    S.cnt = c.shard_cnt.as<unsigned long long>();
    S.scap = scap;
    IPCG_HIP(hipMemsetAsync(S.cnt, 0, 4 * SHARD_CNT_BYTES, s));
    const double off = dmin + dhat;
    c.stage_begin("np_classify");
    if (n > 0)
        hipLaunchKernelGGL(k_classify, dim3(grid_for(n, NP_BLOCK, INT32_MAX)), dim3(NP_BLOCK), 0, s,
                           c.X[0].as<double>(), c.rest.as<double>(), c.nV, c.E.as<int32_t>(),
                           c.F.as<int32_t>(), c.F2E.as<int32_t>(), c.cand_ee.as<int2>(), c.n_ee,
                           c.cand_fv.as<int2>(), c.n_fv, c.cand_vv.as<int2>(), c.n_vv,
                           c.cand_ev.as<int2>(), c.n_ev, off * off, S, A);
    IPCG_CHECK_LAUNCH();
    c.stage_end();
    unsigned long long* h = static_cast<unsigned long long*>(c.pinned);
    ShardCounts kc[4];
    int64_t nraw[4];
    for (int k = 0; k < 4; k++) {
        kc[k] = shard_counts(S.cnt + (int64_t)k * SHARDS * SHARD_PAD, c.pinned, s);
        if (kc[k].max > scap) throw Error(IPCG_ERR_NUMERIC, "collision stream outgrew its shard");
        nraw[k] = kc[k].total;
    }

    c.stage_begin("np_merge");
    int64_t mx = 1;
    for (int k = 0; k < 4; k++) mx = std::max(mx, nraw[k]);
    c.np_keys.reserve(sizeof(uint64_t) * mx);
    c.np_keys2.reserve(sizeof(uint64_t) * mx);
    c.np_vals.reserve(sizeof(uint32_t) * mx);
    c.np_vals2.reserve(sizeof(uint32_t) * mx);
    c.np_ukeys.reserve(sizeof(uint64_t) * mx);
    c.np_start.reserve(sizeof(uint32_t) * (mx + 1));
    c.np_tmp.reserve(std::max(radix_tmp_bytes(mx), rle_tmp_bytes(mx)));
    uint64_t* keys = c.np_keys.as<uint64_t>();
    uint32_t* vals = c.np_vals.as<uint32_t>();
    uint32_t* nu_dev = reinterpret_cast<uint32_t*>(c.counters.as<unsigned long long>() + 3);
    const int vbits = key_bits_for(std::max<int64_t>(c.nV, c.nE));
    const int fbits = key_bits_for(std::max<int64_t>(c.nF, 1));
    for (int k = 0; k < 4; k++) {
        const int64_t m = nraw[k];
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
            shard_compact(S.ee, scap, kcnt, kc[k].max, ee, koff, s);
            hipLaunchKernelGGL(k_ee_keys, dim3(g), dim3(256), 0, s, ee, m, keys, vals);
            radix_sort_pairs(keys, vals, c.np_keys2.as<uint64_t>(), c.np_vals2.as<uint32_t>(), m,
                             vbits, c.np_tmp.p, s);
            hipLaunchKernelGGL(k_ee_gather, dim3(g), dim3(256), 0, s, ee, vals, m, A,
                               c.c_ids[2].as<int2>(), c.c_w[2].as<double>(), c.c_eps.as<double>(),
                               c.c_dtype.as<int32_t>());
            c.ncoll[2] = m; // each unordered edge pair is one candidate: no duplicates
        } else {
            const uint64_t* src = k == 0 ? S.vv : (k == 1 ? S.ev : S.fv);
            shard_compact(src, scap, kcnt, kc[k].max, keys, koff, s);
            double* cw = nullptr; // compacted contribution weights (area weighting)
            if (aw && k < 2) {
                c.np_cw.reserve(sizeof(double) * (size_t)m);
                cw = c.np_cw.as<double>();
                shard_compact(k == 0 ? S.vvw : S.evw, scap, kcnt, kc[k].max, cw, koff, s);
            }
            hipLaunchKernelGGL(k_iota, dim3(g), dim3(256), 0, s, vals, m);
            radix_sort_pairs(keys, vals, c.np_keys2.as<uint64_t>(), c.np_vals2.as<uint32_t>(), m,
                             k == 3 ? fbits : vbits, c.np_tmp.p, s);
            if (k == 3) { // each (face, vertex) candidate is unique
                hipLaunchKernelGGL(k_keys_to_pairs, dim3(g), dim3(256), 0, s, keys, m, A,
                                   c.c_ids[3].as<int2>(), c.c_w[3].as<double>());
                c.ncoll[3] = m;
            } else {
                run_length_encode(keys, m, c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(),
                                  nu_dev, c.np_tmp.p, s);
                if (cw)
                    hipLaunchKernelGGL(k_runs_to_weight_sums, dim3(g), dim3(256), 0, s,
                                       c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(),
                                       nu_dev, vals, cw, c.c_ids[k].as<int2>(),
                                       c.c_w[k].as<double>());
                else
                    hipLaunchKernelGGL(k_runs_to_weights, dim3(g), dim3(256), 0, s,
                                       c.np_ukeys.as<uint64_t>(), c.np_start.as<uint32_t>(), nu_dev,
                                       c.c_ids[k].as<int2>(), c.c_w[k].as<double>());
                uint32_t* hu = reinterpret_cast<uint32_t*>(h);
                IPCG_HIP(hipMemcpyAsync(hu, nu_dev, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                IPCG_HIP(hipStreamSynchronize(s));
                c.ncoll[k] = hu[0];
            }
        }
        IPCG_CHECK_LAUNCH();
    }
    c.stage_end();
    c.coll_dmin = dmin;
    c.coll_valid = true;
    c.hess_valid = false;
    c.pattern_valid = false;
    c.inc_valid = false;
}

} // namespace ipcg
