// C ABI (include/ipcg.h): context lifetime, uploads, and the status-code
// wrappers around the stage entry points.  Nothing here throws across the
// ABI: every entry point catches and records the message in the context.
#include "context.hpp"
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

using namespace ipcg;


namespace ipcg {

void Ctx::require_mesh() const
{
    if (!has_mesh) throw Error(IPCG_ERR_STATE, "no collision mesh uploaded");
}

void Ctx::require_pos(int slot) const
{
    if (!has_pos[slot])
        throw Error(IPCG_ERR_STATE, "positions slot " + std::to_string(slot) + " not uploaded");
}

void Ctx::stage_begin(const char* name)
{
    if (!timer.enabled) return;
    hipEvent_t e;
    IPCG_HIP(hipEventCreate(&e));
    IPCG_HIP(hipEventRecord(e, stream));
    open_stage.push_back({ e, std::string(name) + stage_tag });
}

void Ctx::stage_end()
{
    if (!timer.enabled || open_stage.empty()) return;
    hipEvent_t e;
    IPCG_HIP(hipEventCreate(&e));
    IPCG_HIP(hipEventRecord(e, stream));
    auto st = open_stage.back();
    open_stage.pop_back();
    timer.pending.push_back({ st.first, e, st.second });
}

} // namespace ipcg

namespace {

template <class F> int guard(ipcg_ctx* ctx, F&& f)
{
    if (!ctx) return IPCG_ERR_ARG;
    try {
        f(ctx->c);
        ctx->c.err.clear();
        return IPCG_OK;
    } catch (const Error& e) {
        ctx->c.err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->c.err = e.what();
        return IPCG_ERR_HIP;
    }
}

// caller layout (SoA: x[n], y[n], z[n]; AoS: xyz per vertex) -> 32-byte
// records of the device layout
__global__ void k_to_records(const double* in, int layout, double* out, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double x, y, z;
    if (layout == IPCG_LAYOUT_SOA) {
        x = in[i];
        y = in[n + i];
        z = in[2 * (int64_t)n + i];
    } else {
        x = in[3 * (int64_t)i];
        y = in[3 * (int64_t)i + 1];
        z = in[3 * (int64_t)i + 2];
    }
    double* o = out + PSTRIDE * (int64_t)i;
    *reinterpret_cast<double2*>(o) = make_double2(x, y);
    *reinterpret_cast<double2*>(o + 2) = make_double2(z, 0.0);
}

__global__ void k_gather_pairs(const int2* in, int64_t n, int32_t* out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        out[2 * i] = in[i].x;
        out[2 * i + 1] = in[i].y;
    }
}

void upload_positions(Ctx& c, DevBuf& dst, const double* src, int layout, bool device_src)
{
    if (layout != IPCG_LAYOUT_SOA && layout != IPCG_LAYOUT_AOS)
        throw Error(IPCG_ERR_ARG, "unknown position layout");
    dst.reserve(sizeof(double) * PSTRIDE * (size_t)std::max(c.nV, 1));
    if (c.nV == 0) return;
    const double* in = src;
    if (!device_src) { // stage the caller's array on the device first
        const size_t bytes = sizeof(double) * 3 * (size_t)c.nV;
        c.np_tmp.reserve(bytes);
        IPCG_HIP(hipMemcpyAsync(c.np_tmp.p, src, bytes, hipMemcpyHostToDevice, c.stream));
        in = c.np_tmp.as<double>();
    }
    hipLaunchKernelGGL(k_to_records, dim3(grid_for(c.nV, 256)), dim3(256), 0, c.stream, in, layout,
                       dst.as<double>(), c.nV);
    IPCG_CHECK_LAUNCH();
    if (!device_src) IPCG_HIP(hipStreamSynchronize(c.stream));
}

} // namespace

extern "C" {

int ipcg_ctx_create(int device, ipcg_ctx** out)
{
    if (!out) return IPCG_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return IPCG_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) return IPCG_ERR_ARG;
    ipcg_ctx* ctx = new ipcg_ctx();
    const int rc = guard(ctx, [&](Ctx& c) {
        c.device = device;
        IPCG_HIP(hipSetDevice(device));
        IPCG_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        noalloc_init(device); // device address of the allocation-free energy's scratch
        IPCG_HIP(hipHostMalloc(&c.pinned, PINNED_BYTES, hipHostMallocDefault));
        c.counters.reserve(sizeof(unsigned long long) * 16);
    });
    if (rc != IPCG_OK) {
        delete ctx;
        return rc;
    }
    *out = ctx;
    return IPCG_OK;
}

// every device buffer a context owns
static std::vector<DevBuf*> all_bufs(Ctx& c)
{
    std::vector<DevBuf*> bufs = { &c.E, &c.F, &c.F2E, &c.patch, &c.rest, &c.X[0], &c.X[1],
                       &c.tv.box, &c.tv.keys, &c.tv.keys2, &c.tv.ids, &c.tv.ids2, &c.tv.nodes, &c.tv.info, &c.tv.root,
                       &c.te.box, &c.te.keys, &c.te.keys2, &c.te.ids, &c.te.ids2, &c.te.nodes, &c.te.info, &c.te.root,
                       &c.tf.box, &c.tf.keys, &c.tf.keys2, &c.tf.ids, &c.tf.ids2, &c.tf.nodes, &c.tf.info, &c.tf.root,
                       &c.aabb_partial, &c.aabb, &c.sort_tmp, &c.cand_ee, &c.cand_fv, &c.counters,
                       &c.c_ids[0], &c.c_ids[1], &c.c_ids[2], &c.c_ids[3], &c.c_w[0], &c.c_w[1], &c.c_w[2], &c.c_w[3],
                       &c.c_eps, &c.c_dtype, &c.np_keys, &c.np_keys2, &c.np_vals, &c.np_vals2, &c.np_ukeys,
                       &c.np_start, &c.np_tmp, &c.np_raw, &c.red_partial, &c.red_out, &c.grad, &c.hq,
                       &c.blk_keys, &c.blk_keys2, &c.blk_vals, &c.blk_vals2, &c.blk_ukeys, &c.blk_start,
                       &c.blk_tmp, &c.blk_mask, &c.row_cnt, &c.rowptr, &c.col, &c.val, &c.ccd_bound,
                       &c.ccd_overflow, &c.ccd_toi, &c.ccd_status,
                       &c.hp_mark, &c.hp_bptr, &c.hp_bcol, &c.cand_stage[0], &c.cand_stage[1],
                       &c.shard_cnt, &c.shard_off, &c.np_eec, &c.ccd_stage, &c.ccd_todo,
                       &c.tv.leaf, &c.te.leaf, &c.tf.leaf, &c.inc_keys, &c.inc_keys2,
                       &c.inc_vals, &c.inc_vals2, &c.inc_ptr, &c.inc_vtx, &c.ee_other, &c.hrec, &c.tr_deep, &c.shard_sum,
                       &c.tv.sbox, &c.te.sbox, &c.tf.sbox, &c.cv, &c.ce, &c.cev,
                       &c.cand_vv, &c.cand_ev, &c.cand_stage[2], &c.cand_stage[3],
                       &c.cand_ef, &c.isect, &c.gn_out, &c.gn_p, &c.v_area, &c.e_area,
                       &c.adj_vv_off, &c.adj_vv, &c.adj_ve_off, &c.adj_ve, &c.adj_ev_off,
                       &c.adj_ev, &c.adj_vf_off, &c.adj_bnd, &c.np_rw, &c.np_flag, &c.np_pos, &c.np_rep,
                       &c.im_slot, &c.im_slot2, &c.im_vals, &c.im_vals2, &c.im_ukeys,
                       &c.im_start, &c.im_cnt, &c.im_off, &c.im_vv, &c.im_vvw, &c.im_ev,
                       &c.im_evw, &c.im_ee, &c.fric, &c.vel, &c.fric_mu,
                       &c.np_cw, &c.wf_ev, &c.wf_live, &c.wf_lpre, &c.ro_keys, &c.ro_keys2, &c.ro_ids, &c.ro_ids2 };
    for (Tree* T : { &c.tcv, &c.tce })
        for (DevBuf* b : { &T->box, &T->keys, &T->keys2, &T->ids, &T->ids2, &T->nodes, &T->info,
                           &T->root, &T->sbox, &T->leaf })
            b->release();
    for (Tree* T : { &c.tcv, &c.tce })
        for (DevBuf* b : { &T->box, &T->keys, &T->keys2, &T->ids, &T->ids2, &T->nodes, &T->info,
                           &T->root, &T->sbox, &T->leaf })
            bufs.push_back(b);
    for (Tree* T : { &c.tv, &c.te, &c.tf, &c.tcv, &c.tce })
        for (DevBuf* b : { &T->mort, &T->lscene, &T->loff }) bufs.push_back(b);
    for (DevBuf* b : { &c.vscene, &c.voff_d, &c.scene_active, &c.coff_d, &c.scene_part,
                       &c.ccd_sbound })
        bufs.push_back(b);
    return bufs;
}

extern "C++" {
namespace ipcg {
int64_t device_bytes(Ctx& c)
{
    int64_t n = 0;
    for (DevBuf* b : all_bufs(c)) n += (int64_t)b->bytes;
    return n;
}
} // namespace ipcg
}

int ipcg_ctx_destroy(ipcg_ctx* ctx)
{
    if (!ctx) return IPCG_ERR_ARG;
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (DevBuf* b : all_bufs(c)) b->release();
    for (auto& p : c.timer.pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (c.pinned) (void)hipHostFree(c.pinned);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return IPCG_OK;
}

const char* ipcg_last_error(const ipcg_ctx* ctx)
{
    return ctx ? ctx->c.err.c_str() : "null context";
}

int ipcg_ctx_synchronize(ipcg_ctx* ctx)
{
    return guard(ctx, [&](Ctx& c) { IPCG_HIP(hipStreamSynchronize(c.stream)); });
}

void* ipcg_ctx_stream(ipcg_ctx* ctx) { return ctx ? (void*)ctx->c.stream : nullptr; }

// CSR of rows[k] -> vals[k], each row sorted and duplicate-free
// (collision_mesh.cpp:235-243 remove_duplicates)
static void csr_rows(int32_t n, const std::vector<int32_t>& rows, const std::vector<int32_t>& vals,
                     std::vector<int32_t>& off, std::vector<int32_t>& ids)
{
    std::vector<int32_t> cnt((size_t)n + 1, 0);
    for (int32_t r : rows) cnt[(size_t)r + 1]++;
    for (int32_t i = 0; i < n; i++) cnt[(size_t)i + 1] += cnt[i];
    std::vector<int32_t> tmp(rows.size()), fill(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < rows.size(); k++) tmp[(size_t)fill[rows[k]]++] = vals[k];
    off.assign((size_t)n + 1, 0);
    ids.clear();
    ids.reserve(tmp.size());
    for (int32_t i = 0; i < n; i++) {
        off[i] = (int32_t)ids.size();
        std::sort(tmp.begin() + cnt[i], tmp.begin() + cnt[(size_t)i + 1]);
        for (int32_t k = cnt[i]; k < cnt[(size_t)i + 1]; k++)
            if (k == cnt[i] || tmp[k] != tmp[(size_t)k - 1]) ids.push_back(tmp[k]);
    }
    off[n] = (int32_t)ids.size();
}

static void upload_i32(Ctx& c, DevBuf& b, const std::vector<int32_t>& v)
{
    b.reserve(sizeof(int32_t) * std::max<size_t>(v.size(), 1));
    if (!v.empty())
        IPCG_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice,
                                c.stream));
}

// CollisionMesh::init_adjacencies (collision_mesh.cpp:247-291, 3D rule for
// the boundary flags): used by the IMPROVED_MAX_APPROX collision set
static void upload_adjacencies(Ctx& c, int32_t nV, const int32_t* edges, int32_t nE,
                               const int32_t* faces, int32_t nF, const std::vector<int32_t>& f2e)
{
    if (2 * (int64_t)nE >= INT32_MAX || 3 * (int64_t)nF >= INT32_MAX)
        throw Error(IPCG_ERR_ARG, "mesh too large for 32-bit adjacency offsets");
    std::vector<int32_t> rows(2 * (size_t)nE), vals(2 * (size_t)nE), off, ids;
    for (int32_t e = 0; e < nE; e++) {
        rows[2 * (size_t)e] = edges[2 * e];
        vals[2 * (size_t)e] = edges[2 * e + 1];
        rows[2 * (size_t)e + 1] = edges[2 * e + 1];
        vals[2 * (size_t)e + 1] = edges[2 * e];
    }
    csr_rows(nV, rows, vals, off, ids);
    upload_i32(c, c.adj_vv_off, off);
    upload_i32(c, c.adj_vv, ids);
    for (int32_t e = 0; e < nE; e++) vals[2 * (size_t)e] = vals[2 * (size_t)e + 1] = e;
    csr_rows(nV, rows, vals, off, ids);
    upload_i32(c, c.adj_ve_off, off);
    upload_i32(c, c.adj_ve, ids);
    rows.assign(3 * (size_t)nF, 0);
    vals.assign(3 * (size_t)nF, 0);
    for (int32_t f = 0; f < nF; f++)
        for (int j = 0; j < 3; j++) {
            rows[3 * (size_t)f + j] = f2e[3 * (size_t)f + j];
            vals[3 * (size_t)f + j] = faces[3 * f + (j + 2) % 3];
        }
    csr_rows(nE, rows, vals, off, ids);
    {
        std::vector<int32_t> fr(3 * (size_t)nF), fv(3 * (size_t)nF), foff, fids;
        for (int32_t f = 0; f < nF; f++)
            for (int j = 0; j < 3; j++) {
                fr[3 * (size_t)f + j] = faces[3 * f + j];
                fv[3 * (size_t)f + j] = f;
            }
        csr_rows(nV, fr, fv, foff, fids);
        upload_i32(c, c.adj_vf_off, foff);
    }
    std::vector<uint8_t> bnd((size_t)std::max(nV, 1), 1);
    for (int32_t e = 0; e < nE; e++)
        if (off[(size_t)e + 1] - off[e] >= 2) bnd[edges[2 * e]] = bnd[edges[2 * e + 1]] = 0;
    upload_i32(c, c.adj_ev_off, off);
    upload_i32(c, c.adj_ev, ids);
    c.adj_bnd.reserve(bnd.size());
    IPCG_HIP(hipMemcpyAsync(c.adj_bnd.p, bnd.data(), bnd.size(), hipMemcpyHostToDevice, c.stream));
    IPCG_HIP(hipStreamSynchronize(c.stream)); // the host vectors go out of scope
}

// The collision mesh (any number of scenes: their topology holds global
// vertex ids); cv_out / ce_out receive the codim vertices and edges.
static void mesh_upload(Ctx& c, int32_t nV, const double* rest, int layout, int32_t nE,
                        const int32_t* edges, int32_t nF, const int32_t* faces,
                        const int32_t* vertex_patch, std::vector<int32_t>* cv_out = nullptr,
                        std::vector<int32_t>* ce_out = nullptr)
{
    {
        if (nV < 0 || nE < 0 || nF < 0 || (nV > 0 && !rest) || (nE > 0 && !edges)
            || (nF > 0 && !faces))
            throw Error(IPCG_ERR_ARG, "invalid mesh sizes or null arrays");
        IPCG_HIP(hipSetDevice(c.device));
        for (int64_t i = 0; i < 2 * (int64_t)nE; i++)
            if (edges[i] < 0 || edges[i] >= nV) throw Error(IPCG_ERR_ARG, "edge vertex id out of range");
        for (int64_t i = 0; i < 3 * (int64_t)nF; i++)
            if (faces[i] < 0 || faces[i] >= nV) throw Error(IPCG_ERR_ARG, "face vertex id out of range");
        // faces_to_edges (collision_mesh.cpp:510-541)
        std::vector<std::pair<uint64_t, int32_t>> ek((size_t)nE);
        for (int32_t e = 0; e < nE; e++) {
            uint32_t a = (uint32_t)edges[2 * e], b = (uint32_t)edges[2 * e + 1];
            if (a > b) std::swap(a, b);
            ek[e] = { ((uint64_t)a << 32) | b, e };
        }
        std::sort(ek.begin(), ek.end());
        std::vector<int32_t> f2e(3 * (size_t)nF);
        for (int32_t f = 0; f < nF; f++)
            for (int j = 0; j < 3; j++) {
                uint32_t a = (uint32_t)faces[3 * f + j], b = (uint32_t)faces[3 * f + (j + 1) % 3];
                if (a > b) std::swap(a, b);
                const uint64_t k = ((uint64_t)a << 32) | b;
                auto it = std::lower_bound(ek.begin(), ek.end(), std::make_pair(k, INT32_MIN));
                if (it == ek.end() || it->first != k)
                    throw Error(IPCG_ERR_ARG, "Unable to find edge!");
                f2e[3 * (size_t)f + j] = it->second;
            }
        c.nV = nV;
        c.nE = nE;
        c.nF = nF;
        c.E.reserve(sizeof(int32_t) * 2 * (size_t)std::max(nE, 1));
        c.F.reserve(sizeof(int32_t) * 3 * (size_t)std::max(nF, 1));
        c.F2E.reserve(sizeof(int32_t) * 3 * (size_t)std::max(nF, 1));
        if (nE) IPCG_HIP(hipMemcpyAsync(c.E.p, edges, sizeof(int32_t) * 2 * nE, hipMemcpyHostToDevice, c.stream));
        if (nF) IPCG_HIP(hipMemcpyAsync(c.F.p, faces, sizeof(int32_t) * 3 * nF, hipMemcpyHostToDevice, c.stream));
        if (nF) IPCG_HIP(hipMemcpyAsync(c.F2E.p, f2e.data(), sizeof(int32_t) * 3 * nF, hipMemcpyHostToDevice, c.stream));
        // codim vertices (in no edge) and codim edges (in no face)
        // (collision_mesh.cpp:145-182)
        std::vector<char> in_edge((size_t)std::max(nV, 1), 0), in_face((size_t)std::max(nE, 1), 0);
        for (int64_t i = 0; i < 2 * (int64_t)nE; i++) in_edge[edges[i]] = 1;
        for (int32_t x : f2e) in_face[x] = 1;
        std::vector<int32_t> cv, ce, cev;
        for (int32_t v = 0; v < nV; v++)
            if (!in_edge[v]) cv.push_back(v);
        for (int32_t e = 0; e < nE; e++)
            if (!in_face[e]) {
                ce.push_back(e);
                cev.push_back(edges[2 * e]);
                cev.push_back(edges[2 * e + 1]);
            }
        c.n_cv = (int32_t)cv.size();
        c.n_ce = (int32_t)ce.size();
        c.cv.reserve(sizeof(int32_t) * std::max<size_t>(cv.size(), 1));
        c.ce.reserve(sizeof(int32_t) * std::max<size_t>(ce.size(), 1));
        c.cev.reserve(sizeof(int32_t) * std::max<size_t>(cev.size(), 1));
        if (!cv.empty()) IPCG_HIP(hipMemcpyAsync(c.cv.p, cv.data(), sizeof(int32_t) * cv.size(), hipMemcpyHostToDevice, c.stream));
        if (!ce.empty()) IPCG_HIP(hipMemcpyAsync(c.ce.p, ce.data(), sizeof(int32_t) * ce.size(), hipMemcpyHostToDevice, c.stream));
        if (!cev.empty()) IPCG_HIP(hipMemcpyAsync(c.cev.p, cev.data(), sizeof(int32_t) * cev.size(), hipMemcpyHostToDevice, c.stream));
        c.has_patch = vertex_patch != nullptr;
        if (c.has_patch) {
            c.patch.reserve(sizeof(int32_t) * (size_t)std::max(nV, 1));
            if (nV) IPCG_HIP(hipMemcpyAsync(c.patch.p, vertex_patch, sizeof(int32_t) * nV, hipMemcpyHostToDevice, c.stream));
        }
        c.n_dynamic = c.n_codim_filter = -1;
        // vertex and edge areas over the rest positions (collision_mesh.cpp:
        // 309-376): a third of the incident faces' areas, else half the
        // incident edges' lengths (vertices), else 1; a third of the
        // incident faces' areas, else the edge length (edges)
        {
            auto P = [&](int i, int k) {
                return layout == IPCG_LAYOUT_SOA ? rest[(int64_t)k * nV + i] : rest[3 * (int64_t)i + k];
            };
            auto elen = [&](int e) {
                const int a = edges[2 * e], b = edges[2 * e + 1];
                const double d0 = P(b, 0) - P(a, 0), d1 = P(b, 1) - P(a, 1), d2 = P(b, 2) - P(a, 2);
                return std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
            };
            std::vector<double> ve((size_t)nV, -1.0), vf((size_t)nV, -1.0), ea((size_t)nE, -1.0);
            for (int32_t e = 0; e < nE; e++) {
                const double l = elen(e);
                for (int j = 0; j < 2; j++) {
                    double& x = ve[edges[2 * e + j]];
                    x = x > 0.0 ? x : 0.0;
                    x += 0.5 * l;
                }
            }
            for (int32_t f = 0; f < nF; f++) {
                const int t0 = faces[3 * f], t1 = faces[3 * f + 1], t2 = faces[3 * f + 2];
                const double a0 = P(t1, 0) - P(t0, 0), a1 = P(t1, 1) - P(t0, 1), a2 = P(t1, 2) - P(t0, 2);
                const double b0 = P(t2, 0) - P(t0, 0), b1 = P(t2, 1) - P(t0, 1), b2 = P(t2, 2) - P(t0, 2);
                const double c0 = a1 * b2 - a2 * b1, c1 = a2 * b0 - a0 * b2, c2 = a0 * b1 - a1 * b0;
                const double area = 0.5 * std::sqrt(c0 * c0 + (c1 * c1 + c2 * c2));
                for (int j = 0; j < 3; j++) {
                    double& x = vf[faces[3 * f + j]];
                    x = x > 0.0 ? x : 0.0;
                    x += area / 3.0;
                    double& y = ea[f2e[3 * (size_t)f + j]];
                    y = y > 0.0 ? y : 0.0;
                    y += area / 3.0;
                }
            }
            std::vector<double> va((size_t)nV);
            for (int32_t i = 0; i < nV; i++) va[i] = vf[i] < 0 ? (ve[i] < 0 ? 1.0 : ve[i]) : vf[i];
            for (int32_t e = 0; e < nE; e++)
                if (ea[e] < 0) ea[e] = elen(e);
            c.v_area.reserve(sizeof(double) * (size_t)std::max(nV, 1));
            c.e_area.reserve(sizeof(double) * (size_t)std::max(nE, 1));
            if (nV) IPCG_HIP(hipMemcpyAsync(c.v_area.p, va.data(), sizeof(double) * nV, hipMemcpyHostToDevice, c.stream));
            if (nE) IPCG_HIP(hipMemcpyAsync(c.e_area.p, ea.data(), sizeof(double) * nE, hipMemcpyHostToDevice, c.stream));
            IPCG_HIP(hipStreamSynchronize(c.stream)); // the host vectors go out of scope
        }
        upload_adjacencies(c, nV, edges, nE, faces, nF, f2e);
        upload_positions(c, c.rest, rest, layout, false);
        IPCG_HIP(hipStreamSynchronize(c.stream));
        c.has_mesh = true;
        c.has_pos[0] = c.has_pos[1] = false;
        c.cand_valid = c.coll_valid = c.hess_valid = false;
        c.coff_valid = false;
        c.row_order_valid = c.ro_ready = false;
        c.fric_valid = c.has_vel = false;
        c.scap_ee = c.scap_fv = c.scap_vv = c.scap_ev = 0;
        if (cv_out) *cv_out = cv;
        if (ce_out) *ce_out = ce;
    }
}

int ipcg_mesh_upload(ipcg_ctx* ctx, int32_t nV, const double* rest, int layout, int32_t nE,
                     const int32_t* edges, int32_t nF, const int32_t* faces,
                     const int32_t* vertex_patch)
{
    return guard(ctx, [&](Ctx& c) {
        c.n_scenes = 0;
        mesh_upload(c, nV, rest, layout, nE, edges, nF, faces, vertex_patch);
    });
}

// first element of every scene among the sorted ids, n_scenes + 1 entries
static std::vector<int64_t> scene_starts(const std::vector<int32_t>& ids,
                                         const std::vector<int64_t>& voff)
{
    std::vector<int64_t> out(voff.size());
    for (size_t k = 0; k < voff.size(); k++)
        out[k] = std::lower_bound(ids.begin(), ids.end(), (int32_t)voff[k]) - ids.begin();
    return out;
}

extern "C++" {
namespace ipcg {

void mesh_upload_batched(Ctx& c, int32_t n_scenes, const int64_t* vertex_offsets,
                         const double* rest, int layout, const int64_t* edge_offsets,
                         const int32_t* edges, const int64_t* face_offsets, const int32_t* faces)
{
        if (n_scenes <= 0 || !vertex_offsets || !edge_offsets || !face_offsets)
            throw Error(IPCG_ERR_ARG, "a batch needs at least one scene and its offsets");
        const int64_t* off[3] = { vertex_offsets, edge_offsets, face_offsets };
        for (int a = 0; a < 3; a++) {
            if (off[a][0] != 0) throw Error(IPCG_ERR_ARG, "scene offsets must start at 0");
            for (int32_t k = 0; k < n_scenes; k++)
                if (off[a][k + 1] < off[a][k])
                    throw Error(IPCG_ERR_ARG, "scene offsets must not decrease");
            if (off[a][n_scenes] >= INT32_MAX)
                throw Error(IPCG_ERR_ARG, "too many vertices, edges or faces for one batch");
        }
        const int32_t nV = (int32_t)vertex_offsets[n_scenes];
        const int32_t nE = (int32_t)edge_offsets[n_scenes];
        const int32_t nF = (int32_t)face_offsets[n_scenes];
        if ((nE > 0 && !edges) || (nF > 0 && !faces) || (nV > 0 && !rest))
            throw Error(IPCG_ERR_ARG, "null mesh arrays");
        // scene-local vertex ids -> global ids, each within its own scene
        std::vector<int32_t> ge(2 * (size_t)nE), gf(3 * (size_t)nF);
        for (int32_t k = 0; k < n_scenes; k++) {
            const int64_t v0 = vertex_offsets[k], nv = vertex_offsets[k + 1] - v0;
            for (int64_t i = 2 * edge_offsets[k]; i < 2 * edge_offsets[k + 1]; i++) {
                if (edges[i] < 0 || edges[i] >= nv)
                    throw Error(IPCG_ERR_ARG, "edge vertex id out of its scene's range");
                ge[i] = (int32_t)(edges[i] + v0);
            }
            for (int64_t i = 3 * face_offsets[k]; i < 3 * face_offsets[k + 1]; i++) {
                if (faces[i] < 0 || faces[i] >= nv)
                    throw Error(IPCG_ERR_ARG, "face vertex id out of its scene's range");
                gf[i] = (int32_t)(faces[i] + v0);
            }
        }
        std::vector<int32_t> cv, ce;
        c.n_scenes = 0;
        mesh_upload(c, nV, rest, layout, nE, ge.data(), nF, gf.data(), nullptr, &cv, &ce);
        const size_t S1 = (size_t)n_scenes + 1;
        c.h_voff.assign(vertex_offsets, vertex_offsets + S1);
        c.h_eoff.assign(edge_offsets, edge_offsets + S1);
        c.h_foff.assign(face_offsets, face_offsets + S1);
        // codim vertices are ascending vertex ids, codim edges ascending edge
        // ids: each scene's are contiguous
        c.h_cvoff = scene_starts(cv, c.h_voff);
        c.h_ceoff = scene_starts(ce, c.h_eoff);
        std::vector<int32_t> vs((size_t)std::max(nV, 1), 0);
        for (int32_t k = 0; k < n_scenes; k++)
            for (int64_t v = vertex_offsets[k]; v < vertex_offsets[k + 1]; v++) vs[v] = k;
        std::vector<int64_t> so(3 * S1);
        std::copy(c.h_voff.begin(), c.h_voff.end(), so.begin());
        std::copy(c.h_eoff.begin(), c.h_eoff.end(), so.begin() + S1);
        std::copy(c.h_foff.begin(), c.h_foff.end(), so.begin() + 2 * S1);
        c.vscene.reserve(sizeof(int32_t) * vs.size());
        c.voff_d.reserve(sizeof(int64_t) * so.size());
        c.scene_active.reserve(S1);
        IPCG_HIP(hipMemcpyAsync(c.vscene.p, vs.data(), sizeof(int32_t) * vs.size(),
                                hipMemcpyHostToDevice, c.stream));
        IPCG_HIP(hipMemcpyAsync(c.voff_d.p, so.data(), sizeof(int64_t) * so.size(),
                                hipMemcpyHostToDevice, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream)); // the host vectors go out of scope
        c.all_active = true;
        c.h_active.assign((size_t)n_scenes, 1);
        c.n_scenes = n_scenes;
}

void batch_set_active(Ctx& c, const uint8_t* active)
{
        if (!c.batched()) throw Error(IPCG_ERR_STATE, "not a batched context");
        if (active) c.h_active.assign(active, active + c.n_scenes);
        else c.h_active.assign((size_t)c.n_scenes, 1);
        c.all_active = std::all_of(c.h_active.begin(), c.h_active.end(), [](uint8_t a) { return a != 0; });
        if (!c.all_active) {
            IPCG_HIP(hipMemcpyAsync(c.scene_active.p, c.h_active.data(), c.h_active.size(),
                                    hipMemcpyHostToDevice, c.stream));
            IPCG_HIP(hipStreamSynchronize(c.stream));
        }
        c.cand_valid = c.coll_valid = c.hess_valid = false;
        c.coff_valid = c.fric_valid = false;
}

} // namespace ipcg
}

int ipcg_mesh_upload_batched(ipcg_ctx* ctx, int32_t n_scenes, const int64_t* vertex_offsets,
                             const double* rest, int layout, const int64_t* edge_offsets,
                             const int32_t* edges, const int64_t* face_offsets,
                             const int32_t* faces)
{
    return guard(ctx, [&](Ctx& c) {
        mesh_upload_batched(c, n_scenes, vertex_offsets, rest, layout, edge_offsets, edges,
                            face_offsets, faces);
    });
}

int ipcg_ctx_device_bytes(ipcg_ctx* ctx, int64_t* bytes)
{
    return guard(ctx, [&](Ctx& c) {
        if (!bytes) throw Error(IPCG_ERR_ARG, "null output");
        *bytes = device_bytes(c);
    });
}

int ipcg_ctx_set_autotune(ipcg_ctx* ctx, int enable)
{
    return guard(ctx, [&](Ctx& c) {
        c.autotune = enable != 0;
        if (!c.autotune) c.launch_block.clear();
    });
}

int ipcg_launch_shape(ipcg_ctx* ctx, const char* stage, int32_t* block, double* ms)
{
    return guard(ctx, [&](Ctx& c) {
        if (!stage || !block) throw Error(IPCG_ERR_ARG, "null argument");
        auto it = c.launch_block.find(stage);
        *block = it == c.launch_block.end() ? 0 : it->second;
        if (ms) {
            auto t = c.tune_ms.find(std::string(stage) + "@" + std::to_string(*block));
            *ms = t == c.tune_ms.end() ? 0.0 : (double)t->second;
        }
    });
}

int ipcg_batch_size(ipcg_ctx* ctx, int32_t* n_scenes)
{
    return guard(ctx, [&](Ctx& c) {
        if (!n_scenes) throw Error(IPCG_ERR_ARG, "null output");
        *n_scenes = c.n_scenes;
    });
}

int ipcg_batch_set_active(ipcg_ctx* ctx, const uint8_t* active)
{
    return guard(ctx, [&](Ctx& c) { batch_set_active(c, active); });
}

int ipcg_collisions_count_batched(ipcg_ctx* ctx, int64_t* counts)
{
    return guard(ctx, [&](Ctx& c) {
        if (!counts) throw Error(IPCG_ERR_ARG, "null output");
        scene_collision_offsets(c);
        const size_t S1 = (size_t)c.n_scenes + 1;
        for (int32_t sc = 0; sc < c.n_scenes; sc++)
            for (int k = 0; k < 4; k++)
                counts[4 * (size_t)sc + k] = c.h_coff[k * S1 + sc + 1] - c.h_coff[k * S1 + sc];
    });
}

int ipcg_barrier_energy_batched(ipcg_ctx* ctx, double dhat, double kappa, double* energies)
{
    return guard(ctx, [&](Ctx& c) {
        if (!energies) throw Error(IPCG_ERR_ARG, "null output");
        pot_energy_scenes(c, dhat, kappa, energies);
    });
}

int ipcg_mesh_filter(ipcg_ctx* ctx, int32_t n_dynamic, int32_t n_codim_vertices)
{
    return guard(ctx, [&](Ctx& c) {
        c.require_mesh();
        if (n_dynamic > c.nV || n_codim_vertices > c.nV)
            throw Error(IPCG_ERR_ARG, "filter bound beyond the vertex count");
        c.n_dynamic = n_dynamic < 0 ? -1 : n_dynamic;
        c.n_codim_filter = n_codim_vertices < 0 ? -1 : n_codim_vertices;
        c.cand_valid = c.coll_valid = c.hess_valid = false;
    });
}

int ipcg_positions_upload(ipcg_ctx* ctx, int slot, const double* V, int layout)
{
    return guard(ctx, [&](Ctx& c) {
        c.require_mesh();
        if (slot < 0 || slot > 1 || (c.nV > 0 && !V)) throw Error(IPCG_ERR_ARG, "invalid slot or array");
        upload_positions(c, c.X[slot], V, layout, false);
        c.has_pos[slot] = true;
    });
}

int ipcg_positions_upload_device(ipcg_ctx* ctx, int slot, const double* dV, int layout)
{
    return guard(ctx, [&](Ctx& c) {
        c.require_mesh();
        if (slot < 0 || slot > 1 || (c.nV > 0 && !dV)) throw Error(IPCG_ERR_ARG, "invalid slot or array");
        upload_positions(c, c.X[slot], dV, layout, true);
        c.has_pos[slot] = true;
    });
}

int ipcg_broad_phase(ipcg_ctx* ctx, int swept, double inflation, int method)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(inflation >= 0)) throw Error(IPCG_ERR_ARG, "inflation radius must be >= 0");
        bp_build(c, swept != 0, inflation, method);
    });
}

int ipcg_candidates_count(ipcg_ctx* ctx, int64_t* n_ee, int64_t* n_fv)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set");
        if (n_ee) *n_ee = c.n_ee;
        if (n_fv) *n_fv = c.n_fv;
    });
}

int ipcg_candidates_download(ipcg_ctx* ctx, int32_t* ee, int32_t* fv)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set");
        if (ee && c.n_ee)
            IPCG_HIP(hipMemcpyAsync(ee, c.cand_ee.p, sizeof(int2) * c.n_ee, hipMemcpyDeviceToHost, c.stream));
        if (fv && c.n_fv)
            IPCG_HIP(hipMemcpyAsync(fv, c.cand_fv.p, sizeof(int2) * c.n_fv, hipMemcpyDeviceToHost, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_mesh_codim(ipcg_ctx* ctx, int32_t* n_codim_vertices, int32_t* n_codim_edges)
{
    return guard(ctx, [&](Ctx& c) {
        c.require_mesh();
        if (n_codim_vertices) *n_codim_vertices = c.n_cv;
        if (n_codim_edges) *n_codim_edges = c.n_ce;
    });
}

int ipcg_codim_candidates_count(ipcg_ctx* ctx, int64_t* n_vv, int64_t* n_ev)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set");
        if (n_vv) *n_vv = c.n_vv;
        if (n_ev) *n_ev = c.n_ev;
    });
}

int ipcg_codim_candidates_download(ipcg_ctx* ctx, int32_t* vv, int32_t* ev)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.cand_valid) throw Error(IPCG_ERR_STATE, "no candidate set");
        if (vv && c.n_vv)
            IPCG_HIP(hipMemcpyAsync(vv, c.cand_vv.p, sizeof(int2) * c.n_vv, hipMemcpyDeviceToHost, c.stream));
        if (ev && c.n_ev)
            IPCG_HIP(hipMemcpyAsync(ev, c.cand_ev.p, sizeof(int2) * c.n_ev, hipMemcpyDeviceToHost, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_edge_face_candidates(ipcg_ctx* ctx, double inflation, int64_t* n_ef)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(inflation >= 0)) throw Error(IPCG_ERR_ARG, "inflation must be >= 0");
        bp_edge_face(c, inflation);
        if (n_ef) *n_ef = c.n_ef;
    });
}

int ipcg_edge_face_candidates_download(ipcg_ctx* ctx, int32_t* ef)
{
    return guard(ctx, [&](Ctx& c) {
        if (ef && c.n_ef)
            IPCG_HIP(hipMemcpyAsync(ef, c.cand_ef.p, sizeof(int2) * c.n_ef, hipMemcpyDeviceToHost,
                                    c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_barrier_gauss_newton_diagonal(ipcg_ctx* ctx, double dhat, double kappa, double* diag)
{
    return guard(ctx, [&](Ctx& c) {
        if (!diag) throw Error(IPCG_ERR_ARG, "null output");
        if (!(dhat > 0) || !(kappa > 0)) throw Error(IPCG_ERR_ARG, "dhat and kappa must be > 0");
        pot_gn_diagonal(c, dhat, kappa);
        if (c.nV)
            IPCG_HIP(hipMemcpyAsync(diag, c.gn_out.p, sizeof(double) * 3 * (size_t)c.nV,
                                    hipMemcpyDeviceToHost, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_barrier_gauss_newton_quadratic_form(ipcg_ctx* ctx, double dhat, double kappa,
                                             const double* p, double* value)
{
    return guard(ctx, [&](Ctx& c) {
        if (!value || (!p && c.nV)) throw Error(IPCG_ERR_ARG, "null argument");
        if (!(dhat > 0) || !(kappa > 0)) throw Error(IPCG_ERR_ARG, "dhat and kappa must be > 0");
        *value = pot_gn_quadratic_form(c, dhat, kappa, p);
    });
}

int ipcg_has_intersections(ipcg_ctx* ctx, int* intersecting)
{
    return guard(ctx, [&](Ctx& c) {
        if (!intersecting) throw Error(IPCG_ERR_ARG, "null output");
        *intersecting = has_intersections(c) ? 1 : 0;
    });
}

int ipcg_collisions_build(ipcg_ctx* ctx, double dhat, double dmin, int method)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(dhat > 0) || !(dmin >= 0)) throw Error(IPCG_ERR_ARG, "dhat must be > 0 and dmin >= 0");
        bp_build(c, false, 0.5 * (dhat + dmin), method);
        np_build(c, dhat, dmin);
    });
}

int ipcg_collisions_set_area_weighting(ipcg_ctx* ctx, int use_area_weighting)
{
    return guard(ctx, [&](Ctx& c) { c.area_weighting = use_area_weighting != 0; });
}

int ipcg_collisions_set_type(ipcg_ctx* ctx, int collision_set_type)
{
    return guard(ctx, [&](Ctx& c) {
        if (collision_set_type != IPCG_COLLISIONS_IPC
            && collision_set_type != IPCG_COLLISIONS_IMPROVED_MAX_APPROX
            && collision_set_type != IPCG_COLLISIONS_OGC)
            throw Error(IPCG_ERR_ARG, "unknown collision set type");
        c.coll_type = collision_set_type;
    });
}

int ipcg_barrier_set_physical(ipcg_ctx* ctx, int use_physical_barrier)
{
    return guard(ctx, [&](Ctx& c) { c.physical_barrier = use_physical_barrier != 0; });
}

int ipcg_collisions_build_from_candidates(ipcg_ctx* ctx, double dhat, double dmin)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(dhat > 0) || !(dmin >= 0)) throw Error(IPCG_ERR_ARG, "dhat must be > 0 and dmin >= 0");
        np_build(c, dhat, dmin);
    });
}

int ipcg_collisions_count(ipcg_ctx* ctx, int64_t counts[4])
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.coll_valid) throw Error(IPCG_ERR_STATE, "no collision set");
        for (int k = 0; k < 4; k++) counts[k] = c.ncoll[k];
    });
}

int ipcg_collisions_download(ipcg_ctx* ctx, int32_t* vv, double* vv_w, int32_t* ev, double* ev_w,
                             int32_t* ee, double* ee_w, double* ee_eps_x, int32_t* ee_dtype,
                             int32_t* fv, double* fv_w)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.coll_valid) throw Error(IPCG_ERR_STATE, "no collision set");
        int32_t* ids[4] = { vv, ev, ee, fv };
        double* w[4] = { vv_w, ev_w, ee_w, fv_w };
        for (int k = 0; k < 4; k++) {
            const int64_t n = c.ncoll[k];
            if (!n) continue;
            if (ids[k]) IPCG_HIP(hipMemcpyAsync(ids[k], c.c_ids[k].p, sizeof(int2) * n, hipMemcpyDeviceToHost, c.stream));
            if (w[k]) IPCG_HIP(hipMemcpyAsync(w[k], c.c_w[k].p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream));
        }
        if (c.ncoll[2]) {
            if (ee_eps_x) IPCG_HIP(hipMemcpyAsync(ee_eps_x, c.c_eps.p, sizeof(double) * c.ncoll[2], hipMemcpyDeviceToHost, c.stream));
            if (ee_dtype) IPCG_HIP(hipMemcpyAsync(ee_dtype, c.c_dtype.p, sizeof(int32_t) * c.ncoll[2], hipMemcpyDeviceToHost, c.stream));
        }
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_barrier_energy(ipcg_ctx* ctx, double dhat, double kappa, double* out)
{
    return guard(ctx, [&](Ctx& c) {
        if (!out) throw Error(IPCG_ERR_ARG, "null output");
        *out = pot_energy(c, dhat, kappa);
    });
}

int ipcg_barrier_energy_noalloc(ipcg_ctx* ctx, double dhat, double kappa, double* energy_device)
{
    return guard(ctx, [&](Ctx& c) { pot_energy_noalloc(c, dhat, kappa, energy_device); });
}

int ipcg_barrier_gradient(ipcg_ctx* ctx, double dhat, double kappa, double* grad)
{
    return guard(ctx, [&](Ctx& c) {
        pot_gradient(c, dhat, kappa);
        if (grad && c.nV) {
            IPCG_HIP(hipMemcpyAsync(grad, c.grad.p, sizeof(double) * 3 * c.nV, hipMemcpyDeviceToHost, c.stream));
            IPCG_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int ipcg_barrier_hessian(ipcg_ctx* ctx, double dhat, double kappa, int psd, int64_t* nnz)
{
    return guard(ctx, [&](Ctx& c) {
        pot_hessian(c, dhat, kappa, psd);
        if (nnz) *nnz = c.nnz;
    });
}

int ipcg_velocities_upload(ipcg_ctx* ctx, const double* V, int layout)
{
    return guard(ctx, [&](Ctx& c) {
        c.require_mesh();
        if (c.nV > 0 && !V) throw Error(IPCG_ERR_ARG, "null velocities");
        upload_positions(c, c.vel, V, layout, false);
        c.has_vel = true;
    });
}

int ipcg_friction_build(ipcg_ctx* ctx, double dhat, double kappa, double mu_s, double mu_k)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(dhat > 0) || !(kappa > 0) || !(mu_s >= 0) || !(mu_k >= 0))
            throw Error(IPCG_ERR_ARG, "dhat and kappa must be > 0, mu_s and mu_k >= 0");
        fric_build(c, dhat, kappa, mu_s, mu_k);
    });
}

int ipcg_friction_build_per_vertex(ipcg_ctx* ctx, double dhat, double kappa, const double* mu_s,
                                   const double* mu_k)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(dhat > 0) || !(kappa > 0) || (c.nV > 0 && (!mu_s || !mu_k)))
            throw Error(IPCG_ERR_ARG, "dhat and kappa must be > 0 and the coefficient arrays set");
        for (int32_t i = 0; i < c.nV; i++)
            if (!(mu_s[i] >= 0) || !(mu_k[i] >= 0))
                throw Error(IPCG_ERR_ARG, "friction coefficients must be >= 0");
        fric_build(c, dhat, kappa, 0.0, 0.0, mu_s, mu_k);
    });
}

int ipcg_friction_set_anisotropy(ipcg_ctx* ctx, const double* mu_aniso, const double* mu_s_aniso,
                                 const double* mu_k_aniso)
{
    return guard(ctx, [&](Ctx& c) { fric_set_anisotropy(c, mu_aniso, mu_s_aniso, mu_k_aniso); });
}

int ipcg_friction_update_lagged_mu(ipcg_ctx* ctx)
{
    return guard(ctx, [&](Ctx& c) { fric_update_lagged_mu(c); });
}

int ipcg_friction_download(ipcg_ctx* ctx, int32_t* keep, double* closest_point,
                           double* tangent_basis, double* normal_force)
{
    return guard(ctx, [&](Ctx& c) { fric_download(c, keep, closest_point, tangent_basis, normal_force); });
}

int ipcg_friction_energy(ipcg_ctx* ctx, double eps_v, double* out)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(eps_v > 0)) throw Error(IPCG_ERR_ARG, "eps_v must be > 0");
        const double e = fric_energy(c, eps_v);
        if (out) *out = e;
    });
}

int ipcg_friction_gradient(ipcg_ctx* ctx, double eps_v, double* grad)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(eps_v > 0)) throw Error(IPCG_ERR_ARG, "eps_v must be > 0");
        fric_gradient(c, eps_v);
        if (grad && c.nV) {
            IPCG_HIP(hipMemcpyAsync(grad, c.grad.p, sizeof(double) * 3 * c.nV, hipMemcpyDeviceToHost, c.stream));
            IPCG_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int ipcg_friction_hessian(ipcg_ctx* ctx, double eps_v, int psd, int64_t* nnz)
{
    return guard(ctx, [&](Ctx& c) {
        if (!(eps_v > 0)) throw Error(IPCG_ERR_ARG, "eps_v must be > 0");
        fric_hessian(c, eps_v, psd);
        if (nnz) *nnz = c.nnz;
    });
}

int ipcg_hessian_download(ipcg_ctx* ctx, int64_t* rowptr, int32_t* col, double* val)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.hess_valid) throw Error(IPCG_ERR_STATE, "no Hessian assembled");
        if (rowptr) IPCG_HIP(hipMemcpyAsync(rowptr, c.rowptr.p, sizeof(int64_t) * (3 * (size_t)c.nV + 1), hipMemcpyDeviceToHost, c.stream));
        if (col && c.nnz) IPCG_HIP(hipMemcpyAsync(col, c.col.p, sizeof(int32_t) * c.nnz, hipMemcpyDeviceToHost, c.stream));
        if (val && c.nnz) IPCG_HIP(hipMemcpyAsync(val, c.val.p, sizeof(double) * c.nnz, hipMemcpyDeviceToHost, c.stream));
        IPCG_HIP(hipStreamSynchronize(c.stream));
    });
}

int ipcg_device_results(ipcg_ctx* ctx, const double** grad, const int64_t** rowptr,
                        const int32_t** col, const double** val)
{
    return guard(ctx, [&](Ctx& c) {
        if (grad) *grad = c.grad.as<double>();
        if (rowptr) *rowptr = c.rowptr.as<int64_t>();
        if (col) *col = c.col.as<int32_t>();
        if (val) *val = c.val.as<double>();
    });
}

static void check_ccd_args(double min_distance, double tol, double rescale)
{
    if (!(min_distance >= 0) || !(tol > 0) || !(rescale > 0 && rescale <= 1))
        throw Error(IPCG_ERR_ARG, "invalid CCD parameters");
}

int ipcg_compute_collision_free_stepsize(ipcg_ctx* ctx, double min_distance, double tolerance,
                                         int64_t max_iterations, double rescale, int method,
                                         double* step)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!step) throw Error(IPCG_ERR_ARG, "null output");
        bp_build(c, true, 0.5 * min_distance, method);
        *step = ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, nullptr);
    });
}

int ipcg_compute_collision_free_stepsize_batched(ipcg_ctx* ctx, double min_distance,
                                                 double tolerance, int64_t max_iterations,
                                                 double rescale, double* steps)
{
    return guard(ctx, [&](Ctx& c) {
        if (!c.batched()) throw Error(IPCG_ERR_STATE, "not a batched context");
        check_ccd_args(min_distance, tolerance, rescale);
        if (!steps) throw Error(IPCG_ERR_ARG, "null output");
        bp_build(c, true, 0.5 * min_distance, IPCG_BP_LBVH);
        ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, nullptr, steps);
    });
}

int ipcg_candidates_stepsize(ipcg_ctx* ctx, double min_distance, double tolerance,
                             int64_t max_iterations, double rescale, double* step)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!step) throw Error(IPCG_ERR_ARG, "null output");
        *step = ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, nullptr);
    });
}

int ipcg_is_step_collision_free(ipcg_ctx* ctx, double min_distance, double tolerance,
                                int64_t max_iterations, double rescale, int method,
                                int* collision_free)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!collision_free) throw Error(IPCG_ERR_ARG, "null output");
        bp_build(c, true, 0.5 * min_distance, method);
        const double r = ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, true, nullptr);
        *collision_free = r == 1.0 ? 1 : 0;
    });
}

int ipcg_candidates_is_step_collision_free(ipcg_ctx* ctx, double min_distance, double tolerance,
                                           int64_t max_iterations, double rescale,
                                           int* collision_free)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!collision_free) throw Error(IPCG_ERR_ARG, "null output");
        const double r = ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, true, nullptr);
        *collision_free = r >= 1.0 ? 1 : 0;
    });
}

int ipcg_candidates_per_vertex_safe_distances(ipcg_ctx* ctx, double inflation_radius,
                                              double min_distance, double* out)
{
    return guard(ctx, [&](Ctx& c) {
        if (!out) throw Error(IPCG_ERR_ARG, "null output");
        if (!(min_distance >= 0) || !(inflation_radius >= min_distance))
            throw Error(IPCG_ERR_ARG, "need inflation_radius >= min_distance >= 0");
        cand_safe_distances(c, inflation_radius, min_distance, out);
    });
}

int ipcg_candidates_noncandidate_conservative_stepsize(ipcg_ctx* ctx, double dhat, double* step)
{
    return guard(ctx, [&](Ctx& c) {
        if (!step) throw Error(IPCG_ERR_ARG, "null output");
        if (!(dhat > 0)) throw Error(IPCG_ERR_ARG, "dhat must be > 0");
        *step = cand_noncandidate_stepsize(c, dhat);
    });
}

int ipcg_candidates_cfl_stepsize(ipcg_ctx* ctx, double dhat, double min_distance,
                                 double tolerance, int64_t max_iterations, double rescale,
                                 int method, double* step)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!step) throw Error(IPCG_ERR_ARG, "null output");
        if (!(dhat > 0)) throw Error(IPCG_ERR_ARG, "dhat must be > 0");
        if (method != IPCG_BP_LBVH && method != IPCG_BP_SAP)
            throw Error(IPCG_ERR_ARG, "unsupported broad phase method");
        const double alpha_c =
            ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, nullptr);
        const double alpha_f = cand_noncandidate_stepsize(c, dhat);
        if (alpha_f < 0.5 * alpha_c) { // full CCD (candidates.cpp:352-357)
            bp_build(c, true, 0.5 * min_distance, method);
            *step = ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false,
                                 nullptr);
        } else {
            *step = alpha_c < alpha_f ? alpha_c : alpha_f;
        }
    });
}

int ipcg_candidates_toi(ipcg_ctx* ctx, double min_distance, double tolerance,
                        int64_t max_iterations, double rescale, double* toi_ee, double* toi_fv)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!toi_ee && !toi_fv) throw Error(IPCG_ERR_ARG, "null outputs");
        double* const out[4] = { toi_ee, toi_fv, nullptr, nullptr };
        ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, out);
    });
}

int ipcg_codim_candidates_toi(ipcg_ctx* ctx, double min_distance, double tolerance,
                              int64_t max_iterations, double rescale, double* toi_vv,
                              double* toi_ev)
{
    return guard(ctx, [&](Ctx& c) {
        check_ccd_args(min_distance, tolerance, rescale);
        if (!toi_vv && !toi_ev) throw Error(IPCG_ERR_ARG, "null outputs");
        double* const out[4] = { nullptr, nullptr, toi_vv, toi_ev };
        ccd_stepsize(c, min_distance, tolerance, max_iterations, rescale, false, out);
    });
}

int ipcg_stage_times(ipcg_ctx* ctx, int max, const char** names, double* ms, int64_t* calls)
{
    if (!ctx) return -IPCG_ERR_ARG;
    Ctx& c = ctx->c;
    try {
        IPCG_HIP(hipStreamSynchronize(c.stream));
        for (auto& p : c.timer.pending) {
            float t = 0;
            IPCG_HIP(hipEventElapsedTime(&t, p.a, p.b));
            auto& a = c.timer.acc[p.name];
            a.first += t;
            a.second += 1;
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        c.timer.pending.clear();
    } catch (const Error& e) {
        c.err = e.what();
        return -e.code;
    }
    int i = 0;
    for (auto& kv : c.timer.acc) {
        if (i >= max) break;
        if (names) names[i] = kv.first.c_str();
        if (ms) ms[i] = kv.second.first;
        if (calls) calls[i] = kv.second.second;
        i++;
    }
    return i;
}

int ipcg_counter(ipcg_ctx* ctx, const char* name, int64_t* value)
{
    return guard(ctx, [&](Ctx& c) {
        if (!name || !value) throw Error(IPCG_ERR_ARG, "null argument");
        auto it = c.stat.find(name);
        if (it == c.stat.end()) throw Error(IPCG_ERR_ARG, std::string("no counter ") + name);
        *value = it->second;
    });
}

int ipcg_stage_times_reset(ipcg_ctx* ctx, int enable)
{
    return guard(ctx, [&](Ctx& c) {
        IPCG_HIP(hipStreamSynchronize(c.stream));
        for (auto& p : c.timer.pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        c.timer.pending.clear();
        c.timer.acc.clear();
        c.timer.enabled = enable != 0;
    });
}

} // extern "C"
