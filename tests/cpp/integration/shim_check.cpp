// The reference-side shim (ipc-toolkit_amd/integration/ipc_gpu.hpp) compiled
// against a stand-in of the reference interface (stand_in/) and linked with
// libipcg.so.  Without a HIP device it reports that and exits with 3; with
// one it drives every override of the GPU broad phase through the
// reference's BroadPhase interface, the GPU collision build into the
// reference's collision vectors, the barrier and the step size through the
// name dispatch a maintainer adds to compute_collision_free_stepsize.
#include <ipc_gpu.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <utility>

using namespace ipc;

static int failures = 0;
#define EXPECT(cond)                                                                         \
    do {                                                                                     \
        if (!(cond)) {                                                                       \
            std::printf("FAILED: %s (line %d)\n", #cond, __LINE__);                          \
            failures++;                                                                      \
        }                                                                                    \
    } while (0)

// two stacked n x n unit sheets, the top one `gap` above the bottom one
static void two_sheets(int n, double gap, Eigen::MatrixXd& V, Eigen::MatrixXi& E,
                       Eigen::MatrixXi& F)
{
    const int nv = n * n, nf = 2 * (n - 1) * (n - 1);
    V = Eigen::MatrixXd(2 * nv, 3);
    F = Eigen::MatrixXi(2 * nf, 3);
    std::map<std::pair<int, int>, int> edges;
    for (int s = 0; s < 2; s++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                const int v = s * nv + i * n + j;
                V(v, 0) = j / double(n - 1);
                V(v, 1) = i / double(n - 1);
                V(v, 2) = s * gap;
            }
        int f = s * nf;
        for (int i = 0; i + 1 < n; i++)
            for (int j = 0; j + 1 < n; j++) {
                const int a = s * nv + i * n + j, b = a + 1, c = a + n + 1, d = a + n;
                const int tri[2][3] = { { a, b, c }, { a, c, d } };
                for (auto& t : tri) {
                    for (int k = 0; k < 3; k++) {
                        F(f, k) = t[k];
                        const int x = t[k], y = t[(k + 1) % 3];
                        edges.emplace(std::make_pair(std::min(x, y), std::max(x, y)), 0);
                    }
                    f++;
                }
            }
    }
    E = Eigen::MatrixXi(static_cast<Eigen::Index>(edges.size()), 2);
    int e = 0;
    for (auto& kv : edges) {
        E(e, 0) = kv.first.first;
        E(e, 1) = kv.first.second;
        e++;
    }
}

// the branch of compute_collision_free_stepsize (ipc.cpp:63-85 dispatches
// its CUDA broad phase the same way)
static double stepsize_by_name(BroadPhase* bp, const Eigen::MatrixXd& V0,
                               const Eigen::MatrixXd& V1, const NarrowPhaseCCD& ccd)
{
    if (auto* stq = dynamic_cast<gpu::SweepAndTiniestQueue*>(bp))
        return gpu::compute_collision_free_stepsize(*stq, V0, V1, ccd);
    if (bp != nullptr && (bp->name() == "GPULBVH" || bp->name() == "GPUSweepAndPrune"))
        return gpu::compute_collision_free_stepsize(static_cast<gpu::GPULBVH&>(*bp), V0, V1, 0,
                                                    ccd);
    return -1; // the reference's CPU path would run here
}

int main()
{
    const double dhat = 1e-3, kappa = 1e3;
    Eigen::MatrixXd V, V1;
    Eigen::MatrixXi E, F;
    two_sheets(12, 0.5 * dhat, V, E, F);
    V1 = V;
    for (Eigen::Index v = V.rows() / 2; v < V.rows(); v++) V1(v, 2) -= 1e-3;
    CollisionMesh mesh(V, E, F);
    try {
        gpu::Context probe(mesh);
    } catch (const std::runtime_error& e) {
        std::printf("no device: %s\n", e.what());
        return 3;
    }
    gpu::Context ctx(mesh);
    gpu::GPULBVH gpu_bp(ctx);
    BroadPhase& bp = gpu_bp; // every call below goes through the reference interface
    EXPECT(bp.name() == "GPULBVH");
    bp.build(V, E, F, 0.5 * dhat);
    std::vector<EdgeEdgeCandidate> ee;
    std::vector<FaceVertexCandidate> fv;
    std::vector<VertexVertexCandidate> vv;
    std::vector<EdgeVertexCandidate> ev;
    std::vector<EdgeFaceCandidate> ef;
    std::vector<FaceFaceCandidate> ff;
    bp.detect_edge_edge_candidates(ee);
    bp.detect_face_vertex_candidates(fv);
    bp.detect_vertex_vertex_candidates(vv);
    bp.detect_edge_vertex_candidates(ev);
    bp.detect_edge_face_candidates(ef);
    EXPECT(!ee.empty() && !fv.empty() && !ef.empty());
    EXPECT(vv.empty() && ev.empty()); // a surface mesh has no codim primitives
    bool threw = false;
    try {
        bp.detect_face_face_candidates(ff);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    threw = false;
    try {
        bp.build(AABBs(), E, F, 3);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    bp.clear();
    threw = false;
    try {
        bp.detect_edge_edge_candidates(ee);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);

    NormalCollisions C;
    gpu::build(C, ctx, V, dhat);
    EXPECT(C.size() > 0);
    int64_t n[4];
    gpu::check(ctx.c, ipcg_collisions_count(ctx.c, n));
    EXPECT((int64_t)C.vv_collisions.size() == n[0] && (int64_t)C.ev_collisions.size() == n[1]
           && (int64_t)C.ee_collisions.size() == n[2] && (int64_t)C.fv_collisions.size() == n[3]);
    const double e = gpu::energy(ctx, dhat, kappa);
    const Eigen::VectorXd g = gpu::gradient(ctx, dhat, kappa);
    const Eigen::SparseMatrix<double> H = gpu::hessian(ctx, dhat, kappa, PSDProjectionMethod::CLAMP);
    double gn = 0;
    for (Eigen::Index i = 0; i < g.size(); i++) gn += g.data()[i] * g.data()[i];
    EXPECT(e > 0 && gn > 0 && H.nonZeros() > 0);

    TightInclusionCCD ccd;
    const double t = stepsize_by_name(&bp, V, V1, ccd);
    double t_c = 0;
    gpu::check(ctx.c, ipcg_compute_collision_free_stepsize(ctx.c, 0, ccd.tolerance,
                                                           ccd.max_iterations,
                                                           ccd.conservative_rescaling,
                                                           IPCG_BP_LBVH, &t_c));
    EXPECT(t > 0 && t < 1 && t == t_c);

    // SweepAndTiniestQueue: the sets of the device sweep and prune (FP64
    // boxes), inside those of the LBVH (FP32 outward-rounded boxes), and the
    // branch's rule on the step (ipc.cpp:63-85): no minimum distance, the
    // earliest time of impact times 0.8 below 1, else 1
    gpu::SweepAndTiniestQueue gpu_stq(ctx);
    BroadPhase& stq = gpu_stq;
    EXPECT(stq.name() == "SweepAndTiniestQueue");
    using Pairs = std::vector<std::pair<int, int>>;
    auto sorted = [](Pairs p) {
        std::sort(p.begin(), p.end());
        return p;
    };
    auto ee_pairs = [&](const std::vector<EdgeEdgeCandidate>& v) {
        Pairs p;
        for (const auto& c : v) p.emplace_back(c.edge0_id, c.edge1_id);
        return sorted(p);
    };
    auto fv_pairs = [&](const std::vector<FaceVertexCandidate>& v) {
        Pairs p;
        for (const auto& c : v) p.emplace_back(c.face_id, c.vertex_id);
        return sorted(p);
    };
    auto subset = [&](Pairs a, Pairs b) { // as unordered pairs: each broad phase
        for (Pairs* p : { &a, &b }) {     // orders its edge-edge pairs its own way
            for (auto& x : *p)
                if (x.second < x.first) std::swap(x.first, x.second);
            *p = sorted(*p);
        }
        return std::includes(b.begin(), b.end(), a.begin(), a.end());
    };
    gpu::GPULBVH gpu_sap(ctx, IPCG_BP_SAP);
    BroadPhase& sap = gpu_sap;
    EXPECT(sap.name() == "GPUSweepAndPrune");
    std::vector<EdgeEdgeCandidate> ee_l, ee_p, ee_s;
    std::vector<FaceVertexCandidate> fv_l, fv_p, fv_s;
    bp.build(V, E, F, 0.5 * dhat);
    bp.detect_edge_edge_candidates(ee_l);
    bp.detect_face_vertex_candidates(fv_l);
    sap.build(V, E, F, 0.5 * dhat);
    sap.detect_edge_edge_candidates(ee_p);
    sap.detect_face_vertex_candidates(fv_p);
    stq.build(V, E, F, 0.5 * dhat);
    stq.detect_edge_edge_candidates(ee_s);
    stq.detect_face_vertex_candidates(fv_s);
    EXPECT(!ee_s.empty() && ee_pairs(ee_s) == ee_pairs(ee_p) && fv_pairs(fv_s) == fv_pairs(fv_p));
    EXPECT(subset(ee_pairs(ee_s), ee_pairs(ee_l)));
    EXPECT(subset(fv_pairs(fv_s), fv_pairs(fv_l))); // (face, vertex): one order
    const double t_stq = stepsize_by_name(&stq, V, V1, ccd);
    double toi = 0;
    gpu::check(ctx.c, ipcg_compute_collision_free_stepsize(ctx.c, 0, ccd.tolerance,
                                                           ccd.max_iterations, 1.0,
                                                           IPCG_BP_SAP, &toi));
    EXPECT(toi > 0 && toi < 1 && t_stq == 0.8 * toi);
    EXPECT(stepsize_by_name(&stq, V, V, ccd) == 1.0); // no motion, no impact
    Eigen::MatrixXd V2(V.rows(), 2);
    for (Eigen::Index i = 0; i < V.rows(); i++) {
        V2(i, 0) = V(i, 0);
        V2(i, 1) = V(i, 1);
    }
    threw = false;
    try {
        stepsize_by_name(&stq, V2, V2, ccd);
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    std::printf("candidates ee %zu fv %zu ef %zu; collisions %zu; energy %.6e; nnz %ld; "
                "step %.9g; STQ step %.9g\n",
                ee.size(), fv.size(), ef.size(), C.size(), e, (long)H.nonZeros(), t, t_stq);
    std::printf("%d failure(s)\n", failures);
    return failures ? 1 : 0;
}
