// Device-resident state of one collision mesh and the host-side stage entry
// points that enqueue its kernels on the context stream.
#pragma once

#include "common.hpp"

#include <map>
#include <string>
#include <vector>

namespace ipcg {

constexpr size_t PINNED_BYTES = 64 * 1024;

struct StageTimer {
    bool enabled = false;
    struct Rec {
        hipEvent_t a, b;
        std::string name;
    };
    std::vector<Rec> pending;
    std::map<std::string, std::pair<double, int64_t>> acc;
};

// LBVH of one primitive set: internal node j splits sorted keys j | j+1.
struct Node {
    float lo0[3], hi0[3], lo1[3], hi1[3];
    int32_t c0, c1; // child: >= 0 internal node, < 0 leaf ~position
    int32_t r0, r1; // rightmost leaf position of each child
};
static_assert(sizeof(Node) == 64, "Node is one 64-byte line");

struct Tree {
    int32_t n = 0;       // leaves
    DevBuf box;          // float[6] (LBVH) or double[6] (sweep) per primitive
    DevBuf keys, keys2;  // uint64 Morton codes (sorted in place)
    DevBuf ids, ids2;    // uint32 primitive id per sorted position
    DevBuf nodes;        // Node[n-1]
    DevBuf info;         // construction counters / ranges
    DevBuf root;         // int32 root node (-1: single leaf)
    DevBuf sbox;         // sweep and prune: FP64 boxes in sorted order
    DevBuf leaf;         // int4 per sorted leaf: (mesh id, its vertices, -1 padded)
    // forest of a batched context: Morton code per primitive, scene per
    // sorted leaf; `root` then holds one entry per scene
    DevBuf mort, lscene, loff;
};

// Root entries of a batched context's forest: an internal node (>= 0), a
// single leaf (~position), no primitive of this kind in the scene, or not
// yet completed by the refit.
constexpr int32_t ROOT_NONE = INT32_MIN + 1;
constexpr int32_t ROOT_UNFINISHED = INT32_MIN;

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    // mesh
    int32_t nV = 0, nE = 0, nF = 0;
    bool has_mesh = false, has_patch = false;
    int32_t n_dynamic = -1, n_codim_filter = -1; // optional vertex-pair filters
    DevBuf E, F, F2E, patch, rest;
    DevBuf v_area, e_area;     // CollisionMesh vertex / edge areas (rest positions)
    bool area_weighting = false; // NormalCollisions::use_area_weighting
    int coll_type = 0; // NormalCollisions::CollisionSetType: 0 IPC, 1 IMPROVED_MAX_APPROX, 2 OGC
    bool physical_barrier = false; // BarrierPotential use_physical_barrier
    // CollisionMesh adjacencies (collision_mesh.cpp:247-291), CSR with int32
    // offsets, rows sorted and duplicate-free: vertex -> vertices sharing an
    // edge, vertex -> incident edges, edge -> vertices opposite it in its
    // faces, vertex -> incident faces (offsets only); is_vertex_on_boundary
    // as bytes
    DevBuf adj_vv_off, adj_vv, adj_ve_off, adj_ve, adj_ev_off, adj_ev, adj_vf_off, adj_bnd;
    DevBuf X[2];
    bool has_pos[2] = { false, false };

    // Batched context: independent scenes concatenated; scene s owns
    // vertices [voff[s], voff[s+1]) and likewise edges, faces, codim
    // vertices and codim edges.  n_scenes = 0: one scene, no scene arrays.
    int32_t n_scenes = 0;
    std::vector<int64_t> h_voff, h_eoff, h_foff, h_cvoff, h_ceoff;
    DevBuf vscene;              // int32 scene per vertex
    DevBuf voff_d;              // int64 vertex offsets (n_scenes + 1)
    DevBuf scene_active;        // uint8 per scene: queries of inactive scenes are skipped
    bool all_active = true;
    std::vector<uint8_t> h_active;
    bool batched() const { return n_scenes > 0; }
    // per-scene collision offsets per kind (4 x (n_scenes + 1), kind-major),
    // valid with the collision set
    bool coff_valid = false;
    std::vector<int64_t> h_coff;
    DevBuf coff_d, scene_part;

    // codim vertices (in no edge) and codim edges (in no face), their
    // endpoint pairs (collision_mesh.cpp:145-182)
    int32_t n_cv = 0, n_ce = 0;
    DevBuf cv, ce, cev;

    // broad phase
    Tree tv, te, tf, tcv, tce;
    DevBuf aabb_partial, aabb; // mesh bounds
    DevBuf sort_tmp;
    bool cand_valid = false, cand_swept = false;
    DevBuf cand_ee, cand_fv;   // int2 pairs
    DevBuf counters;           // uint64 counters (device)
    int64_t n_ee = 0, n_fv = 0, n_vv = 0, n_ev = 0;
    DevBuf cand_vv, cand_ev;           // codim lists: (v, v) and (edge, vertex)
    DevBuf cand_ef;                    // (edge, face) pairs of the last bp_edge_face
    int64_t n_ef = 0, scap_ef = 0;
    DevBuf isect;                      // intersection test: flag, deferred count, list
    int64_t scap_ee = 0, scap_fv = 0, scap_vv = 0, scap_ev = 0; // shard regions
    DevBuf cand_stage[4];              // sharded traversal output (ee, fv, vv, ev)
    DevBuf tr_deep;                    // traversal stack entries beyond LDS
    DevBuf shard_sum;                  // per counter set: total, largest shard (u64)
    DevBuf shard_cnt, shard_off;       // shard counters / offsets of the append lists

    // collision set
    bool coll_valid = false;
    double coll_dmin = 0;
    int64_t ncoll[4] = { 0, 0, 0, 0 };
    DevBuf c_ids[4], c_w[4], c_eps, c_dtype; // ids: int2 per collision
    DevBuf np_cw; // contribution weights of the merge (area weighting)
    DevBuf np_keys, np_keys2, np_vals, np_vals2, np_ukeys, np_start, np_tmp, np_raw, np_eec;
    DevBuf np_rw, np_flag, np_pos, np_rep; // signed merge: run sums, kept flags, positions
    // IMPROVED_MAX_APPROX: conversion slots, their unique keys, per-key
    // output counts / offsets, and the signed vertex-vertex, edge-vertex and
    // edge-edge records they add
    DevBuf im_slot, im_slot2, im_vals, im_vals2, im_ukeys, im_start, im_cnt, im_off;
    DevBuf im_vv, im_vvw, im_ev, im_evw, im_ee;

    // potential
    DevBuf red_partial, red_out, grad, hq;
    DevBuf blk_keys, blk_keys2, blk_vals, blk_vals2, blk_ukeys, blk_start, blk_tmp;
    DevBuf blk_mask, row_cnt, rowptr, col, val;
    DevBuf hp_mark, hp_bptr, hp_bcol; // block pattern of the collision set
    DevBuf inc_keys, inc_keys2, inc_vals, inc_vals2, inc_ptr; // vertex -> (collision, point)
    DevBuf ee_other;                  // edge-edge collisions of non-interior types (list)
    DevBuf inc_vtx;                   // stencil vertices per sorted incidence (int4)
    DevBuf hrec;                      // projected Hessian per collision (Q basis)
    DevBuf gn_out, gn_p;              // Gauss-Newton diagonal (3 nV) and direction (3 nV)
    DevBuf fric;                      // friction record per collision (global order)
    bool fric_valid = false;
    DevBuf fric_mu;                   // per-vertex static, then kinetic coefficients
    DevBuf vel;                       // velocities (3 nV, vertex major)
    bool has_vel = false;
    bool inc_valid = false;
    // aabb holds the box of every scene (set by a broad phase), so the
    // per-vertex row kernels can visit the vertices in spatial order:
    // ro_ids, built with the incidences (ro_ready)
    bool row_order_valid = false, ro_ready = false;
    DevBuf ro_keys, ro_keys2, ro_ids, ro_ids2;
    int64_t nblocks = 0;
    bool pattern_valid = false;
    int64_t nnz = 0;
    bool hess_valid = false;

    // ccd
    DevBuf ccd_bound, ccd_overflow, ccd_toi, ccd_status;
    DevBuf ccd_stage, ccd_todo; // root-pass survivors (sharded, packed)
    DevBuf ccd_sbound;          // batched contexts: running minimum per scene
    DevBuf wf_ev, wf_live, wf_lpre; // CCD wavefront: per-candidate events, scanned boxes

    // host pinned staging (PINNED_BYTES)
    void* pinned = nullptr;

    // launch shapes: workgroup size per tuned stage (absent: the default),
    // filled on first use when autotune is on; tune_ms: each size's time
    bool autotune = false;
    std::map<std::string, int> launch_block;
    std::map<std::string, float> tune_ms;

    StageTimer timer;
    void stage_begin(const char* name);
    void stage_end();
    std::vector<std::pair<hipEvent_t, std::string>> open_stage;
    std::string stage_tag; // appended to stage names (e.g. ".static")
    std::map<std::string, int64_t> stat; // work counts of the last calls

    void require_mesh() const;
    void require_pos(int slot) const;
};

// ---- stage entry points (one translation unit each) -------------------------
void bp_build(Ctx& c, bool swept, double inflation, int method);
void np_build(Ctx& c, double dhat, double dmin);
double pot_energy(Ctx& c, double dhat, double kappa);
// allocation-free energy: block partials in a per-device static array of
// NOALLOC_PARTIALS doubles; enqueues on the context stream and writes the
// energy to device memory (no synchronization)
constexpr int64_t NOALLOC_PARTIALS = 1 << 20;
void noalloc_init(int device);
// batched contexts: upload (scene-local topology, ragged offsets), the
// active-scene mask, the bytes of device memory a context holds
void mesh_upload_batched(Ctx& c, int32_t n_scenes, const int64_t* vertex_offsets,
                         const double* rest, int layout, const int64_t* edge_offsets,
                         const int32_t* edges, const int64_t* face_offsets, const int32_t* faces);
void batch_set_active(Ctx& c, const uint8_t* active);
int64_t device_bytes(Ctx& c);
// batched contexts: per-scene collision offsets and energies (n_scenes
// values); ccd_stepsize's scene_steps receives the per-scene step sizes
void scene_collision_offsets(Ctx& c);
void pot_energy_scenes(Ctx& c, double dhat, double kappa, double* out);

void pot_energy_noalloc(Ctx& c, double dhat, double kappa, double* d_out);
void pot_gradient(Ctx& c, double dhat, double kappa);
void pot_hessian(Ctx& c, double dhat, double kappa, int psd);
// toi_out: null (step size) or 4 host arrays (ee, fv, vv, ev; any null)
// receiving per-candidate times of impact
double ccd_stepsize(Ctx& c, double min_distance, double tol, int64_t max_iter,
                    double rescale, bool early_exit, double* const* toi_out,
                    double* scene_steps = nullptr);
void bp_edge_face(Ctx& c, double inflation);
void fric_build(Ctx& c, double dhat, double kappa, double mu_s, double mu_k,
                const double* mu_sv = nullptr, const double* mu_kv = nullptr);
void fric_set_anisotropy(Ctx& c, const double* mu_aniso, const double* mu_s_aniso,
                         const double* mu_k_aniso);
void fric_update_lagged_mu(Ctx& c);
void fric_download(Ctx& c, int32_t* keep, double* beta, double* basis, double* normal_force);
double fric_energy(Ctx& c, double eps_v);
void fric_gradient(Ctx& c, double eps_v);
void fric_hessian(Ctx& c, double eps_v, int psd);
void pot_gn_diagonal(Ctx& c, double dhat, double kappa);
double pot_gn_quadratic_form(Ctx& c, double dhat, double kappa, const double* p_host);
// candidate-set helpers (candidates.cpp:294-408) over the current lists
void cand_safe_distances(Ctx& c, double inflation_radius, double min_distance, double* out);
double cand_noncandidate_stepsize(Ctx& c, double dhat);
bool has_intersections(Ctx& c);

} // namespace ipcg

// the C ABI's opaque context
struct ipcg_ctx {
    ipcg::Ctx c;
};
