// Reference-side shim: plugs libipcg (include/ipcg.h) into IPC Toolkit
// v1.6.0 behind its own interfaces.  A maintainer adds this header to the
// reference build (it includes the reference's headers) and links
// libipcg.so; nothing in the reference changes except the one branch in
// compute_collision_free_stepsize shown at the end (ipc.cpp:63-85 holds the
// same kind of branch for its CUDA broad phase).
//
//   ipc::gpu::Context ctx(mesh);             // topology resident on the GPU
//   ipc::gpu::GPULBVH bp(ctx);               // an ipc::BroadPhase
//   ipc::NormalCollisions C;
//   C.build(mesh, V, dhat, 0, &bp);          // reference CPU narrow phase, GPU candidates
//   ipc::gpu::build(C, ctx, V, dhat);        // or the whole build on the GPU
//   double t = ipc::compute_collision_free_stepsize(mesh, V0, V1, 0, &bp);
#pragma once

#include <ipcg.h>

#include <ipc/broad_phase/broad_phase.hpp>
#include <ipc/candidates/candidates.hpp>
#include <ipc/ccd/tight_inclusion_ccd.hpp>
#include <ipc/collision_mesh.hpp>
#include <ipc/collisions/normal/normal_collisions.hpp>
#include <ipc/utils/eigen_ext.hpp>

#include <Eigen/Core>
#include <Eigen/Sparse>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace ipc::gpu {

// status -> exception, as log_and_throw_error does (logger.cpp:41-45)
inline void check(ipcg_ctx* c, int rc)
{
    if (rc != IPCG_OK) throw std::runtime_error(std::string("ipcg: ") + ipcg_last_error(c));
}

// an Eigen index matrix (any storage order) as row-major int32
inline std::vector<int32_t> rows_of(const Eigen::MatrixXi& M)
{
    std::vector<int32_t> out(static_cast<size_t>(M.rows() * M.cols()));
    for (Eigen::Index i = 0; i < M.rows(); i++)
        for (Eigen::Index j = 0; j < M.cols(); j++)
            out[static_cast<size_t>(i * M.cols() + j)] = M(i, j);
    return out;
}

/// One device context per collision mesh: topology, rest positions and the
/// mesh's vertex-pair filter stay on the GPU until the context is destroyed.
class Context {
public:
    explicit Context(const CollisionMesh& mesh, int device = 0)
    {
        if (ipcg_ctx_create(device, &c) != IPCG_OK)
            throw std::runtime_error("ipcg: no HIP device");
        const std::vector<int32_t> E = rows_of(mesh.edges()), F = rows_of(mesh.faces());
        // Eigen::MatrixXd is column-major: x[], y[], z[] = IPCG_LAYOUT_SOA
        const int rc = ipcg_mesh_upload(c, static_cast<int32_t>(mesh.num_vertices()),
                                        mesh.rest_positions().data(), IPCG_LAYOUT_SOA,
                                        static_cast<int32_t>(mesh.edges().rows()), E.data(),
                                        static_cast<int32_t>(mesh.faces().rows()), F.data(),
                                        nullptr);
        if (rc != IPCG_OK) {
            const std::string msg = ipcg_last_error(c);
            ipcg_ctx_destroy(c);
            throw std::runtime_error("ipcg: " + msg);
        }
        nV = static_cast<int>(mesh.num_vertices());
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    ~Context() { ipcg_ctx_destroy(c); }

    /// slot 0: the configuration x; slot 1: the end of a CCD step
    void positions(int slot, Eigen::ConstRef<Eigen::MatrixXd> V)
    {
        if (V.rows() != nV || V.cols() != 3)
            throw std::invalid_argument("ipcg: positions must be num_vertices x 3");
        check(c, ipcg_positions_upload(c, slot, V.data(), IPCG_LAYOUT_SOA));
    }

    ipcg_ctx* c = nullptr;
    int nV = 0;
};

/// ipc::BroadPhase on the GPU (broad_phase.hpp:19-133): LBVH or sweep and
/// prune over FP32 / FP64 boxes, the same candidate sets as the reference's
/// LBVH / SweepAndPrune.  The filter is the context mesh's (the patch,
/// static-obstacle and codim-cross filters of ipcg_mesh_upload /
/// ipcg_mesh_filter); a can_vertices_collide set on this object is not
/// consulted.
class GPULBVH : public BroadPhase {
public:
    explicit GPULBVH(Context& ctx, int method = IPCG_BP_LBVH) : ctx(ctx), method(method) { }

    std::string name() const override
    {
        return method == IPCG_BP_SAP ? "GPUSweepAndPrune" : "GPULBVH";
    }

    void build(Eigen::ConstRef<Eigen::MatrixXd> vertices, Eigen::ConstRef<Eigen::MatrixXi>,
               Eigen::ConstRef<Eigen::MatrixXi>, const double inflation_radius = 0) override
    {
        ctx.positions(0, vertices);
        radius = inflation_radius;
        check(ctx.c, ipcg_broad_phase(ctx.c, 0, inflation_radius, method));
        built = true;
    }

    void build(Eigen::ConstRef<Eigen::MatrixXd> vertices_t0,
               Eigen::ConstRef<Eigen::MatrixXd> vertices_t1, Eigen::ConstRef<Eigen::MatrixXi>,
               Eigen::ConstRef<Eigen::MatrixXi>, const double inflation_radius = 0) override
    {
        ctx.positions(0, vertices_t0);
        ctx.positions(1, vertices_t1);
        radius = inflation_radius;
        check(ctx.c, ipcg_broad_phase(ctx.c, 1, inflation_radius, method));
        built = true;
    }

    /// Boxes built elsewhere cannot be uploaded: the device builds its own
    /// boxes from positions (the two overloads above).
    void build(const AABBs&, Eigen::ConstRef<Eigen::MatrixXi>, Eigen::ConstRef<Eigen::MatrixXi>,
               const uint8_t) override
    {
        throw std::runtime_error("GPULBVH builds its boxes from vertex positions");
    }

    void clear() override { built = false; }

    /// codim vertex pairs (candidates.cpp:65-118); empty for surface meshes
    void detect_vertex_vertex_candidates(std::vector<VertexVertexCandidate>& out) const override
    {
        require_built();
        int64_t nvv = 0, nev = 0;
        check(ctx.c, ipcg_codim_candidates_count(ctx.c, &nvv, &nev));
        std::vector<int32_t> vv(2 * static_cast<size_t>(nvv));
        check(ctx.c, ipcg_codim_candidates_download(ctx.c, vv.data(), nullptr));
        for (int64_t i = 0; i < nvv; i++) out.emplace_back(vv[2 * i], vv[2 * i + 1]);
    }

    /// (codim edge, codim vertex) pairs
    void detect_edge_vertex_candidates(std::vector<EdgeVertexCandidate>& out) const override
    {
        require_built();
        int64_t nvv = 0, nev = 0;
        check(ctx.c, ipcg_codim_candidates_count(ctx.c, &nvv, &nev));
        std::vector<int32_t> ev(2 * static_cast<size_t>(nev));
        check(ctx.c, ipcg_codim_candidates_download(ctx.c, nullptr, ev.data()));
        for (int64_t i = 0; i < nev; i++) out.emplace_back(ev[2 * i], ev[2 * i + 1]);
    }

    void detect_edge_edge_candidates(std::vector<EdgeEdgeCandidate>& out) const override
    {
        require_built();
        int64_t ne = 0, nf = 0;
        check(ctx.c, ipcg_candidates_count(ctx.c, &ne, &nf));
        std::vector<int32_t> ee(2 * static_cast<size_t>(ne));
        check(ctx.c, ipcg_candidates_download(ctx.c, ee.data(), nullptr));
        for (int64_t i = 0; i < ne; i++) out.emplace_back(ee[2 * i], ee[2 * i + 1]);
    }

    void detect_face_vertex_candidates(std::vector<FaceVertexCandidate>& out) const override
    {
        require_built();
        int64_t ne = 0, nf = 0;
        check(ctx.c, ipcg_candidates_count(ctx.c, &ne, &nf));
        std::vector<int32_t> fv(2 * static_cast<size_t>(nf));
        check(ctx.c, ipcg_candidates_download(ctx.c, nullptr, fv.data()));
        for (int64_t i = 0; i < nf; i++) out.emplace_back(fv[2 * i], fv[2 * i + 1]);
    }

    /// edge-face pairs over the static boxes of the last build (lbvh.cpp:
    /// 771-784); rebuilds the edge and face trees, the other lists stay
    void detect_edge_face_candidates(std::vector<EdgeFaceCandidate>& out) const override
    {
        require_built();
        int64_t n = 0;
        check(ctx.c, ipcg_edge_face_candidates(ctx.c, radius, &n));
        std::vector<int32_t> ef(2 * static_cast<size_t>(n));
        check(ctx.c, ipcg_edge_face_candidates_download(ctx.c, ef.data()));
        for (int64_t i = 0; i < n; i++) out.emplace_back(ef[2 * i], ef[2 * i + 1]);
    }

    /// face-face pairs are used by 2D-free intersection checks of other
    /// broad phases only; the device does not produce them
    void detect_face_face_candidates(std::vector<FaceFaceCandidate>&) const override
    {
        throw std::runtime_error("GPULBVH does not produce face-face candidates");
    }

    Context& context() const { return ctx; }
    int broad_phase_method() const { return method; }

private:
    void require_built() const
    {
        if (!built) throw std::runtime_error("GPULBVH: build() first");
    }

    Context& ctx;
    int method;
    double radius = 0;
    bool built = false;
};

/// SweepAndTiniestQueue (sweep_and_tiniest_queue.hpp, .cu:9-280) on this
/// library, for builds without CUDA.  The overrides are GPULBVH's over the
/// device sweep and prune (FP64 boxes): box overlap plus the mesh's filter,
/// the sets the reference's STQ detect_* return.  name() is the one that
/// compute_collision_free_stepsize dispatches on (ipc.cpp:63-85).
class SweepAndTiniestQueue : public GPULBVH {
public:
    explicit SweepAndTiniestQueue(Context& ctx) : GPULBVH(ctx, IPCG_BP_SAP) { }

    std::string name() const override { return "SweepAndTiniestQueue"; }
};

/// NormalCollisions::build (normal_collisions.cpp:20-158) entirely on the
/// GPU: broad phase, narrow phase and merge; the collision vectors are
/// filled in the reference's layout (canonical order, merged weights).
inline void build(NormalCollisions& C, Context& ctx, Eigen::ConstRef<Eigen::MatrixXd> V,
                  double dhat, double dmin = 0, int method = IPCG_BP_LBVH)
{
    ctx.positions(0, V);
    check(ctx.c, ipcg_collisions_set_area_weighting(ctx.c, C.use_area_weighting() ? 1 : 0));
    check(ctx.c, ipcg_collisions_build(ctx.c, dhat, dmin, method));
    int64_t n[4];
    check(ctx.c, ipcg_collisions_count(ctx.c, n));
    std::vector<int32_t> vv(2 * n[0]), ev(2 * n[1]), ee(2 * n[2]), fv(2 * n[3]), dt(n[2]);
    std::vector<double> wvv(n[0]), wev(n[1]), wee(n[2]), eps(n[2]), wfv(n[3]);
    check(ctx.c, ipcg_collisions_download(ctx.c, vv.data(), wvv.data(), ev.data(), wev.data(),
                                          ee.data(), wee.data(), eps.data(), dt.data(),
                                          fv.data(), wfv.data()));
    C.clear();
    const Eigen::SparseVector<double> no_gradient; // no shape derivatives
    for (int64_t i = 0; i < n[0]; i++)
        C.vv_collisions.emplace_back(vv[2 * i], vv[2 * i + 1], wvv[i], no_gradient);
    for (int64_t i = 0; i < n[1]; i++)
        C.ev_collisions.emplace_back(ev[2 * i], ev[2 * i + 1], wev[i], no_gradient);
    for (int64_t i = 0; i < n[2]; i++)
        C.ee_collisions.emplace_back(ee[2 * i], ee[2 * i + 1], eps[i], wee[i], no_gradient,
                                     static_cast<EdgeEdgeDistanceType>(dt[i]));
    for (int64_t i = 0; i < n[3]; i++)
        C.fv_collisions.emplace_back(fv[2 * i], fv[2 * i + 1], wfv[i], no_gradient);
    for (size_t ci = 0; ci < C.size(); ci++) C[ci].dmin = dmin; // normal_collisions.cpp:154-157
}

/// BarrierPotential::operator() / gradient / hessian (potential.cpp:36-222)
/// over the context's last GPU-built collision set.
inline double energy(Context& ctx, double dhat, double kappa)
{
    double e = 0;
    check(ctx.c, ipcg_barrier_energy(ctx.c, dhat, kappa, &e));
    return e;
}

inline Eigen::VectorXd gradient(Context& ctx, double dhat, double kappa)
{
    Eigen::VectorXd g(3 * ctx.nV); // RowMajor vertex derivatives (config.hpp.in:45)
    check(ctx.c, ipcg_barrier_gradient(ctx.c, dhat, kappa, g.data()));
    return g;
}

inline Eigen::SparseMatrix<double> hessian(Context& ctx, double dhat, double kappa,
                                           PSDProjectionMethod psd)
{
    int64_t nnz = 0;
    check(ctx.c, ipcg_barrier_hessian(ctx.c, dhat, kappa, static_cast<int>(psd), &nnz));
    const int n = 3 * ctx.nV;
    std::vector<int64_t> rp(static_cast<size_t>(n) + 1);
    std::vector<int32_t> col(static_cast<size_t>(nnz));
    std::vector<double> val(static_cast<size_t>(nnz));
    check(ctx.c, ipcg_hessian_download(ctx.c, rp.data(), col.data(), val.data()));
    // symmetric: the CSR arrays are the CSC arrays of the same matrix
    Eigen::SparseMatrix<double> H(n, n);
    H.resizeNonZeros(static_cast<Eigen::Index>(nnz));
    std::transform(rp.begin(), rp.end(), H.outerIndexPtr(),
                   [](int64_t x) { return static_cast<int>(x); });
    std::copy(col.begin(), col.end(), H.innerIndexPtr());
    std::copy(val.begin(), val.end(), H.valuePtr());
    return H;
}

/// compute_collision_free_stepsize (ipc.cpp:45-101) with a GPU broad phase:
/// the swept broad phase and Tight-Inclusion CCD on the device, with the
/// TightInclusionCCD's tolerance and iteration limit (the defaults for any
/// other narrow phase) and the narrow phase's conservative rescaling.
inline double compute_collision_free_stepsize(GPULBVH& bp, Eigen::ConstRef<Eigen::MatrixXd> V0,
                                              Eigen::ConstRef<Eigen::MatrixXd> V1,
                                              double min_distance,
                                              const NarrowPhaseCCD& narrow_phase_ccd)
{
    double tol = TightInclusionCCD::DEFAULT_TOLERANCE;
    long max_iter = TightInclusionCCD::DEFAULT_MAX_ITERATIONS;
    const double rescale = narrow_phase_ccd.conservative_rescaling;
    if (const auto* ti = dynamic_cast<const TightInclusionCCD*>(&narrow_phase_ccd)) {
        tol = ti->tolerance;
        max_iter = ti->max_iterations;
    }
    Context& ctx = bp.context();
    ctx.positions(0, V0);
    ctx.positions(1, V1);
    double step = 1;
    check(ctx.c, ipcg_compute_collision_free_stepsize(ctx.c, min_distance, tol, max_iter, rescale,
                                                      bp.broad_phase_method(), &step));
    return step;
}

/// The SweepAndTiniestQueue branch of compute_collision_free_stepsize
/// (ipc.cpp:63-85): 3D only, no minimum distance whatever the caller asks,
/// TightInclusionCCD's tolerance and iteration limit (the defaults for any
/// other narrow phase), and the earliest time of impact scaled by 0.8 when
/// it is below 1.  The device search runs with conservative rescaling 1 (no
/// minimum separation inside the search, no rescaling of small times): the
/// 0.8 of that branch is the only margin, as in the reference.
inline double compute_collision_free_stepsize(SweepAndTiniestQueue& bp,
                                              Eigen::ConstRef<Eigen::MatrixXd> V0,
                                              Eigen::ConstRef<Eigen::MatrixXd> V1,
                                              const NarrowPhaseCCD& narrow_phase_ccd)
{
    if (V0.cols() != 3)
        throw std::runtime_error("Sweep and Tiniest Queue is only supported in 3D!");
    double tol = TightInclusionCCD::DEFAULT_TOLERANCE;
    long max_iter = TightInclusionCCD::DEFAULT_MAX_ITERATIONS;
    if (const auto* ti = dynamic_cast<const TightInclusionCCD*>(&narrow_phase_ccd)) {
        tol = ti->tolerance;
        max_iter = ti->max_iterations;
    }
    Context& ctx = bp.context();
    ctx.positions(0, V0);
    ctx.positions(1, V1);
    double toi = 1;
    check(ctx.c, ipcg_compute_collision_free_stepsize(ctx.c, 0.0, tol, max_iter, 1.0,
                                                      IPCG_BP_SAP, &toi));
    return toi < 1.0 ? 0.8 * toi : 1.0;
}

} // namespace ipc::gpu

// The branch a maintainer adds at the top of ipc::compute_collision_free_
// stepsize (ipc.cpp:45-101), beside the reference's own dispatch on
// broad_phase->name() for its CUDA broad phase (ipc.cpp:63-85).  The STQ
// test comes first and by type: a CUDA build has its own class of that name.
//
//   if (auto* stq = dynamic_cast<gpu::SweepAndTiniestQueue*>(broad_phase)) {
//       return gpu::compute_collision_free_stepsize(
//           *stq, vertices_t0, vertices_t1, narrow_phase_ccd);
//   }
//   if (broad_phase != nullptr
//       && (broad_phase->name() == "GPULBVH"
//           || broad_phase->name() == "GPUSweepAndPrune")) {
//       return gpu::compute_collision_free_stepsize(
//           static_cast<gpu::GPULBVH&>(*broad_phase), vertices_t0,
//           vertices_t1, min_distance, narrow_phase_ccd);
//   }
