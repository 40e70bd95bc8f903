// ipc_amd.hpp -- C++ host API of the MI355X contact path, mirroring the
// reference's interface for it (IPC Toolkit v1.6.0):
//
//   ipc::CollisionMesh          collision_mesh.hpp:14-126
//   ipc::Candidates             candidates/candidates.hpp (build static /
//                               swept, compute_collision_free_stepsize,
//                               is_step_collision_free)
//   ipc::NormalCollisions       collisions/normal/normal_collisions.hpp:45-63
//   ipc::BarrierPotential       potentials/barrier_potential.hpp,
//                               potentials/potential.hpp:34-60
//   ipc::TangentialCollisions   collisions/tangential/tangential_collisions.hpp
//                               (scalar or per-vertex friction coefficients)
//   ipc::FrictionPotential      potentials/friction_potential.hpp
//   ipc::has_intersections      ipc.hpp (ipc.cpp:105-166)
//   ipc::compute_collision_free_stepsize, ipc::is_step_collision_free
//                               ipc.hpp (ipc.cpp:20-101)
//   ipc::TightInclusionCCD      ccd/tight_inclusion_ccd.hpp:11-28 (parameters)
//   ipc::PSDProjectionMethod    utils/eigen_ext.hpp:202-206
//
// Same names, argument meanings and defaults; errors are thrown as
// std::runtime_error with the library's message, as the reference's
// log_and_throw_error does (logger.cpp:41-45).  Everything runs through the
// C ABI of ipcg.h on one HIP device; a CollisionMesh owns one device
// context and keeps topology and positions resident.
//
// Matrices are small dense row-major containers (no Eigen dependency):
// MatrixXd V{{x0, y0, z0}, {x1, y1, z1}, ...}, MatrixXi E{{0, 1}, ...}.
// Gradients are RowMajor [x0, y0, z0, x1, ...] (config.hpp.in:45) and
// Hessians are CSR (the matrix is symmetric, so CSR equals the reference's
// column-major CSC).
#pragma once

#include "ipcg.h"

#include <algorithm>
#include <cstdint>
#include <initializer_list>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace ipc {

template <class T> class Matrix {
public:
    Matrix() = default;
    Matrix(int rows, int cols) : r(rows), c(cols), v((size_t)rows * cols) {}
    Matrix(std::initializer_list<std::initializer_list<T>> rows_)
    {
        r = (int)rows_.size();
        c = r ? (int)rows_.begin()->size() : 0;
        v.reserve((size_t)r * c);
        for (const auto& row : rows_) {
            if ((int)row.size() != c) throw std::invalid_argument("ragged matrix literal");
            v.insert(v.end(), row.begin(), row.end());
        }
    }
    int rows() const { return r; }
    int cols() const { return c; }
    T& operator()(int i, int j) { return v[(size_t)i * c + j]; }
    const T& operator()(int i, int j) const { return v[(size_t)i * c + j]; }
    T* data() { return v.data(); }
    const T* data() const { return v.data(); }

private:
    int r = 0, c = 0;
    std::vector<T> v;
};
using MatrixXd = Matrix<double>;
using MatrixXi = Matrix<int32_t>;

enum class PSDProjectionMethod : int { NONE = IPCG_PSD_NONE, CLAMP = IPCG_PSD_CLAMP,
                                       ABS = IPCG_PSD_ABS };

// broad_phase/broad_phase.hpp: the device broad phases
enum class BroadPhaseMethod : int { LBVH = IPCG_BP_LBVH, SWEEP_AND_PRUNE = IPCG_BP_SAP };

namespace detail {
inline void check(ipcg_ctx* c, int rc)
{
    if (rc != IPCG_OK) throw std::runtime_error(std::string(c ? ipcg_last_error(c) : "ipcg"));
}
} // namespace detail

// ccd/tight_inclusion_ccd.hpp:11-28
struct TightInclusionCCD {
    static constexpr double DEFAULT_TOLERANCE = 1e-6;
    static constexpr long DEFAULT_MAX_ITERATIONS = 10'000'000;
    static constexpr double DEFAULT_CONSERVATIVE_RESCALING = 0.8;
    double tolerance = DEFAULT_TOLERANCE;
    long max_iterations = DEFAULT_MAX_ITERATIONS;
    double conservative_rescaling = DEFAULT_CONSERVATIVE_RESCALING;
    TightInclusionCCD(double tol = DEFAULT_TOLERANCE, long max_iter = DEFAULT_MAX_ITERATIONS,
                      double rescaling = DEFAULT_CONSERVATIVE_RESCALING)
        : tolerance(tol), max_iterations(max_iter), conservative_rescaling(rescaling)
    {
    }
};

// CollisionFilter (collision_filter.hpp:30-152) as data: the intersection
// (operator&) of an optional patch filter, static obstacle filter and codim
// cross filter; the device evaluates the same predicate.
struct CollisionFilter {
    std::vector<int32_t> patches; // make_vertex_patches_filter (empty: none)
    int32_t n_dynamic = -1;       // make_static_obstacle_filter (< 0: none)
    int32_t n_codim_vertices = -1; // make_codim_cross_filter (< 0: none)

    bool operator()(size_t vi, size_t vj) const
    {
        return (patches.empty() || patches[vi] != patches[vj])
            && (n_dynamic < 0 || vi < (size_t)n_dynamic || vj < (size_t)n_dynamic)
            && (n_codim_vertices < 0
                || ((vi < (size_t)n_codim_vertices) != (vj < (size_t)n_codim_vertices)));
    }
    // collision_filter.hpp:74-81; only intersections of distinct filter kinds
    // are representable here
    friend CollisionFilter operator&(CollisionFilter a, const CollisionFilter& b)
    {
        if ((!a.patches.empty() && !b.patches.empty())
            || (a.n_dynamic >= 0 && b.n_dynamic >= 0 && a.n_dynamic != b.n_dynamic)
            || (a.n_codim_vertices >= 0 && b.n_codim_vertices >= 0
                && a.n_codim_vertices != b.n_codim_vertices))
            throw std::invalid_argument("unsupported collision filter intersection");
        if (a.patches.empty()) a.patches = b.patches;
        if (a.n_dynamic < 0) a.n_dynamic = b.n_dynamic;
        if (a.n_codim_vertices < 0) a.n_codim_vertices = b.n_codim_vertices;
        return a;
    }
};

// collision_filter.hpp:113-118
inline CollisionFilter make_vertex_patches_filter(std::vector<int32_t> patch_ids)
{
    CollisionFilter f;
    f.patches = std::move(patch_ids);
    return f;
}

// collision_filter.hpp:126-131
inline CollisionFilter make_static_obstacle_filter(size_t n_dynamic)
{
    CollisionFilter f;
    f.n_dynamic = (int32_t)n_dynamic;
    return f;
}

// collision_filter.hpp:138-143
inline CollisionFilter make_codim_cross_filter(size_t n_codim_vertices)
{
    CollisionFilter f;
    f.n_codim_vertices = (int32_t)n_codim_vertices;
    return f;
}

// collision_filter.hpp:152: vertices joined by faces share a label (union
// find); a vertex in no face is its own component
inline CollisionFilter make_connected_components_filter(const MatrixXi& faces, int num_vertices)
{
    std::vector<int32_t> root((size_t)num_vertices);
    for (int i = 0; i < num_vertices; i++) root[(size_t)i] = i;
    auto find = [&](int32_t x) {
        while (root[(size_t)x] != x) x = root[(size_t)x] = root[(size_t)root[(size_t)x]];
        return x;
    };
    for (int f = 0; f < faces.rows(); f++)
        for (int k = 1; k < 3; k++) {
            const int32_t a = find(faces(f, 0)), b = find(faces(f, k));
            if (a != b) root[(size_t)std::max(a, b)] = std::min(a, b);
        }
    for (int i = 0; i < num_vertices; i++) root[(size_t)i] = find(i);
    return make_vertex_patches_filter(std::move(root));
}

// collision_mesh.hpp:14-126 (full-mesh form; vertex_patches gives the
// patch filter of collision_filter.hpp:113-118)
class CollisionMesh {
public:
    CollisionMesh(const MatrixXd& rest_positions, const MatrixXi& edges = MatrixXi(0, 2),
                  const MatrixXi& faces = MatrixXi(0, 3),
                  const std::vector<int32_t>& vertex_patches = {}, int device = 0)
        : nV(rest_positions.rows()), nE(edges.rows()), nF(faces.rows()), E(edges), F(faces),
          patches(vertex_patches)
    {
        if (rest_positions.rows() && rest_positions.cols() != 3)
            throw std::invalid_argument("positions must be #V x 3");
        if ((edges.rows() && edges.cols() != 2) || (faces.rows() && faces.cols() != 3))
            throw std::invalid_argument("edges must be #E x 2 and faces #F x 3");
        ipcg_ctx* c = nullptr;
        const int rc = ipcg_ctx_create(device, &c);
        if (rc != IPCG_OK) throw std::runtime_error("no HIP device for the collision mesh");
        ctx.reset(c);
        if (!vertex_patches.empty() && (int)vertex_patches.size() != nV)
            throw std::invalid_argument("one patch id per vertex");
        detail::check(c, ipcg_mesh_upload(c, nV, rest_positions.data(), IPCG_LAYOUT_AOS, nE,
                                          edges.data(), nF, faces.data(),
                                          vertex_patches.empty() ? nullptr
                                                                 : vertex_patches.data()));
        detail::check(c, ipcg_mesh_codim(c, &ncv, &nce));
    }
    int num_vertices() const { return nV; }
    int num_edges() const { return nE; }
    int num_faces() const { return nF; }
    int num_codim_vertices() const { return ncv; }
    int num_codim_edges() const { return nce; }
    const MatrixXi& edges() const { return E; }
    const MatrixXi& faces() const { return F; }
    ipcg_ctx* context() const { return ctx.get(); }
    // counts the device broad-phase calls, so a Candidates object can tell
    // whether the device still holds its lists
    uint64_t broad_phase_generation() const { return generation; }
    void touch_broad_phase() const { generation++; }
    // counts the device collision-set builds
    uint64_t collisions_generation() const { return cgeneration; }
    void touch_collisions() const { cgeneration++; }

    // CollisionMesh::can_collide (collision_mesh.hpp:338).  The patch labels
    // are part of the uploaded mesh: a filter's patches must be those given
    // at construction (or none).
    void set_can_collide(const CollisionFilter& f)
    {
        if (!f.patches.empty() && f.patches != patches)
            throw std::invalid_argument("patch labels are fixed when the mesh is built");
        if (f.patches.empty() && !patches.empty())
            throw std::invalid_argument("the mesh was built with patch labels");
        detail::check(ctx.get(), ipcg_mesh_filter(ctx.get(), f.n_dynamic, f.n_codim_vertices));
        touch_broad_phase();
        touch_collisions();
    }

    // slot 0: x / x_t0, slot 1: x_t1
    void set_positions(int slot, const MatrixXd& V) const
    {
        if (V.rows() != nV || (nV && V.cols() != 3))
            throw std::invalid_argument("positions must be #V x 3");
        detail::check(ctx.get(), ipcg_positions_upload(ctx.get(), slot, V.data(), IPCG_LAYOUT_AOS));
    }

private:
    struct Del {
        void operator()(ipcg_ctx* c) const { ipcg_ctx_destroy(c); }
    };
    int nV, nE, nF;
    int32_t ncv = 0, nce = 0;
    mutable uint64_t generation = 0, cgeneration = 0;
    MatrixXi E, F;
    std::vector<int32_t> patches;
    std::unique_ptr<ipcg_ctx, Del> ctx;
};

struct VertexVertexCandidate { int32_t vertex0_id, vertex1_id; };
struct EdgeVertexCandidate { int32_t edge_id, vertex_id; };
struct EdgeEdgeCandidate { int32_t edge0_id, edge1_id; };
struct FaceVertexCandidate { int32_t face_id, vertex_id; };

// candidates/candidates.hpp
class Candidates {
public:
    std::vector<VertexVertexCandidate> vv_candidates;
    std::vector<EdgeVertexCandidate> ev_candidates;
    std::vector<EdgeEdgeCandidate> ee_candidates;
    std::vector<FaceVertexCandidate> fv_candidates;

    // candidates.cpp:43-118 (static)
    void build(const CollisionMesh& mesh, const MatrixXd& vertices, double inflation_radius = 0,
               BroadPhaseMethod bp = BroadPhaseMethod::LBVH)
    {
        mesh.set_positions(0, vertices);
        mesh.touch_broad_phase();
        detail::check(mesh.context(), ipcg_broad_phase(mesh.context(), 0, inflation_radius, int(bp)));
        download(mesh);
        remember(mesh, false, vertices, vertices, inflation_radius, bp);
    }
    // candidates.cpp:120-222 (swept)
    void build(const CollisionMesh& mesh, const MatrixXd& vertices_t0, const MatrixXd& vertices_t1,
               double inflation_radius = 0, BroadPhaseMethod bp = BroadPhaseMethod::LBVH)
    {
        mesh.set_positions(0, vertices_t0);
        mesh.set_positions(1, vertices_t1);
        mesh.touch_broad_phase();
        detail::check(mesh.context(), ipcg_broad_phase(mesh.context(), 1, inflation_radius, int(bp)));
        download(mesh);
        remember(mesh, true, vertices_t0, vertices_t1, inflation_radius, bp);
    }
    size_t size() const
    {
        return vv_candidates.size() + ev_candidates.size() + ee_candidates.size()
            + fv_candidates.size();
    }
    bool empty() const { return size() == 0; }

    // candidates.cpp:252-292 over these candidates (static or swept)
    double compute_collision_free_stepsize(const CollisionMesh& mesh, const MatrixXd& vertices_t0,
                                           const MatrixXd& vertices_t1, double min_distance = 0,
                                           const TightInclusionCCD& ccd = {}) const
    {
        require_lists(mesh, vertices_t0, vertices_t1);
        double step = 1;
        detail::check(mesh.context(),
                      ipcg_candidates_stepsize(mesh.context(), min_distance, ccd.tolerance,
                                               ccd.max_iterations, ccd.conservative_rescaling,
                                               &step));
        return step;
    }
    // candidates.cpp:224-250
    bool is_step_collision_free(const CollisionMesh& mesh, const MatrixXd& vertices_t0,
                                const MatrixXd& vertices_t1, double min_distance = 0,
                                const TightInclusionCCD& ccd = {}) const
    {
        require_lists(mesh, vertices_t0, vertices_t1);
        int free_ = 0;
        detail::check(mesh.context(), ipcg_candidates_is_step_collision_free(
                                          mesh.context(), min_distance, ccd.tolerance,
                                          ccd.max_iterations, ccd.conservative_rescaling, &free_));
        return free_ != 0;
    }

    // candidates.cpp:372-408
    std::vector<double> compute_per_vertex_safe_distances(const CollisionMesh& mesh,
                                                          const MatrixXd& vertices,
                                                          double inflation_radius,
                                                          double min_distance = 0) const
    {
        require_lists(mesh);
        mesh.set_positions(0, vertices);
        std::vector<double> out((size_t)mesh.num_vertices());
        detail::check(mesh.context(), ipcg_candidates_per_vertex_safe_distances(
                                          mesh.context(), inflation_radius, min_distance,
                                          out.data()));
        return out;
    }
    // candidates.cpp:294-333
    double compute_noncandidate_conservative_stepsize(const CollisionMesh& mesh,
                                                      const MatrixXd& displacements,
                                                      double dhat) const
    {
        require_lists(mesh);
        // slot 1 - slot 0 is the displacement itself
        mesh.set_positions(0, MatrixXd(displacements.rows(), displacements.cols()));
        mesh.set_positions(1, displacements);
        double step = 1;
        detail::check(mesh.context(), ipcg_candidates_noncandidate_conservative_stepsize(
                                          mesh.context(), dhat, &step));
        return step;
    }
    // candidates.cpp:335-360; the full CCD fallback replaces the device lists
    double compute_cfl_stepsize(const CollisionMesh& mesh, const MatrixXd& vertices_t0,
                                const MatrixXd& vertices_t1, double dhat, double min_distance = 0,
                                BroadPhaseMethod bp = BroadPhaseMethod::LBVH,
                                const TightInclusionCCD& ccd = {}) const
    {
        require_lists(mesh, vertices_t0, vertices_t1);
        double step = 1;
        const int rc = ipcg_candidates_cfl_stepsize(mesh.context(), dhat, min_distance,
                                                    ccd.tolerance, ccd.max_iterations,
                                                    ccd.conservative_rescaling, int(bp), &step);
        mesh.touch_broad_phase();
        detail::check(mesh.context(), rc);
        return step;
    }

private:
    const CollisionMesh* built_for = nullptr;
    bool swept = false;
    MatrixXd build_v0, build_v1; // inputs of the build, to rebuild the device lists
    uint64_t generation = 0;
    double inflation = 0;
    BroadPhaseMethod method = BroadPhaseMethod::LBVH;
    void remember(const CollisionMesh& mesh, bool swept_, const MatrixXd& V0, const MatrixXd& V1,
                  double r, BroadPhaseMethod bp)
    {
        built_for = &mesh;
        swept = swept_;
        build_v0 = V0;
        build_v1 = swept_ ? V1 : MatrixXd();
        generation = mesh.broad_phase_generation();
        inflation = r;
        method = bp;
    }
    // The device evaluates its own copy of the lists; when another call
    // replaced them since build(), the same broad phase runs again on the
    // build's inputs, which reproduces the same lists.
    void require_lists(const CollisionMesh& mesh) const
    {
        if (built_for != &mesh) throw std::runtime_error("candidates were not built for this mesh");
        if (mesh.broad_phase_generation() != generation) {
            mesh.set_positions(0, build_v0);
            if (swept) mesh.set_positions(1, build_v1);
            mesh.touch_broad_phase();
            detail::check(mesh.context(),
                          ipcg_broad_phase(mesh.context(), swept ? 1 : 0, inflation, int(method)));
            const_cast<Candidates*>(this)->generation = mesh.broad_phase_generation();
        }
    }
    void require_lists(const CollisionMesh& mesh, const MatrixXd& V0, const MatrixXd& V1) const
    {
        require_lists(mesh);
        mesh.set_positions(0, V0);
        mesh.set_positions(1, V1);
    }
    void download(const CollisionMesh& mesh)
    {
        ipcg_ctx* c = mesh.context();
        int64_t ne = 0, nf = 0, nv = 0, nev = 0;
        detail::check(c, ipcg_candidates_count(c, &ne, &nf));
        detail::check(c, ipcg_codim_candidates_count(c, &nv, &nev));
        std::vector<int32_t> ee(2 * ne), fv(2 * nf), vv(2 * nv), ev(2 * nev);
        detail::check(c, ipcg_candidates_download(c, ee.data(), fv.data()));
        detail::check(c, ipcg_codim_candidates_download(c, vv.data(), ev.data()));
        ee_candidates.resize(ne);
        fv_candidates.resize(nf);
        vv_candidates.resize(nv);
        ev_candidates.resize(nev);
        for (int64_t i = 0; i < ne; i++) ee_candidates[i] = { ee[2 * i], ee[2 * i + 1] };
        for (int64_t i = 0; i < nf; i++) fv_candidates[i] = { fv[2 * i], fv[2 * i + 1] };
        for (int64_t i = 0; i < nv; i++) vv_candidates[i] = { vv[2 * i], vv[2 * i + 1] };
        for (int64_t i = 0; i < nev; i++) ev_candidates[i] = { ev[2 * i], ev[2 * i + 1] };
    }
};

struct VertexVertexNormalCollision { int32_t vertex0_id, vertex1_id; double weight; };
struct EdgeVertexNormalCollision { int32_t edge_id, vertex_id; double weight; };
struct EdgeEdgeNormalCollision {
    int32_t edge0_id, edge1_id;
    double weight, eps_x;
    int32_t dtype; // EdgeEdgeDistanceType stored at build (edge_edge.hpp:96)
};
struct FaceVertexNormalCollision { int32_t face_id, vertex_id; double weight; };

// collisions/normal/normal_collisions.hpp:45-63
class NormalCollisions {
public:
    std::vector<VertexVertexNormalCollision> vv_collisions;
    std::vector<EdgeVertexNormalCollision> ev_collisions;
    std::vector<EdgeEdgeNormalCollision> ee_collisions;
    std::vector<FaceVertexNormalCollision> fv_collisions;

    // normal_collisions.hpp set_use_area_weighting / use_area_weighting
    void set_use_area_weighting(bool use) { m_area_weighting = use; }
    bool use_area_weighting() const { return m_area_weighting; }

    // normal_collisions.hpp CollisionSetType
    enum class CollisionSetType { IPC = IPCG_COLLISIONS_IPC,
                                  IMPROVED_MAX_APPROX = IPCG_COLLISIONS_IMPROVED_MAX_APPROX,
                                  OGC = IPCG_COLLISIONS_OGC };
    void set_collision_set_type(CollisionSetType type) { m_type = type; }
    CollisionSetType collision_set_type() const { return m_type; }

    // normal_collisions.cpp:20-36
    void build(const CollisionMesh& mesh, const MatrixXd& vertices, double dhat, double dmin = 0,
               BroadPhaseMethod bp = BroadPhaseMethod::LBVH)
    {
        mesh.set_positions(0, vertices);
        mesh.touch_broad_phase();
        mesh.touch_collisions();
        detail::check(mesh.context(),
                      ipcg_collisions_set_area_weighting(mesh.context(), m_area_weighting ? 1 : 0));
        detail::check(mesh.context(), ipcg_collisions_set_type(mesh.context(), int(m_type)));
        detail::check(mesh.context(), ipcg_collisions_build(mesh.context(), dhat, dmin, int(bp)));
        download(mesh);
        built_for = &mesh;
        m_generation = mesh.collisions_generation();
        m_dmin = dmin;
    }
    size_t size() const
    {
        return vv_collisions.size() + ev_collisions.size() + ee_collisions.size()
            + fv_collisions.size();
    }
    bool empty() const { return size() == 0; }
    double dmin() const { return m_dmin; }
    const CollisionMesh* mesh() const { return built_for; }
    // the device holds this set (no later build on the same mesh)
    bool resident() const
    {
        return built_for && built_for->collisions_generation() == m_generation;
    }

private:
    const CollisionMesh* built_for = nullptr;
    uint64_t m_generation = 0;
    double m_dmin = 0;
    bool m_area_weighting = false;
    CollisionSetType m_type = CollisionSetType::IPC;
    void download(const CollisionMesh& mesh)
    {
        ipcg_ctx* c = mesh.context();
        int64_t n[4];
        detail::check(c, ipcg_collisions_count(c, n));
        std::vector<int32_t> vv(2 * n[0]), ev(2 * n[1]), ee(2 * n[2]), fv(2 * n[3]), dt(n[2]);
        std::vector<double> wvv(n[0]), wev(n[1]), wee(n[2]), eps(n[2]), wfv(n[3]);
        detail::check(c, ipcg_collisions_download(c, vv.data(), wvv.data(), ev.data(), wev.data(),
                                                  ee.data(), wee.data(), eps.data(), dt.data(),
                                                  fv.data(), wfv.data()));
        vv_collisions.resize(n[0]);
        ev_collisions.resize(n[1]);
        ee_collisions.resize(n[2]);
        fv_collisions.resize(n[3]);
        for (int64_t i = 0; i < n[0]; i++) vv_collisions[i] = { vv[2 * i], vv[2 * i + 1], wvv[i] };
        for (int64_t i = 0; i < n[1]; i++) ev_collisions[i] = { ev[2 * i], ev[2 * i + 1], wev[i] };
        for (int64_t i = 0; i < n[2]; i++)
            ee_collisions[i] = { ee[2 * i], ee[2 * i + 1], wee[i], eps[i], dt[i] };
        for (int64_t i = 0; i < n[3]; i++) fv_collisions[i] = { fv[2 * i], fv[2 * i + 1], wfv[i] };
    }
};

// CSR of a symmetric 3V x 3V matrix (int64 row pointers)
struct SparseMatrix {
    int64_t rows = 0, cols = 0;
    std::vector<int64_t> outer;
    std::vector<int32_t> inner;
    std::vector<double> values;
    int64_t nonZeros() const { return (int64_t)values.size(); }
    double coeff(int64_t r, int64_t c) const
    {
        for (int64_t k = outer[r]; k < outer[r + 1]; k++)
            if (inner[k] == c) return values[k];
        return 0;
    }
};

// potentials/barrier_potential.hpp (use_physical_barrier = false)
class BarrierPotential {
public:
    BarrierPotential(double dhat, double stiffness = 1.0, bool use_physical_barrier = false)
        : m_dhat(dhat), m_stiffness(stiffness), m_physical(use_physical_barrier)
    {
        if (!(dhat > 0)) throw std::invalid_argument("dhat must be positive");
    }
    double dhat() const { return m_dhat; }
    double stiffness() const { return m_stiffness; }
    bool use_physical_barrier() const { return m_physical; }
    // the device flag follows this potential for the calls it makes
    void apply(const CollisionMesh& mesh) const
    {
        detail::check(mesh.context(), ipcg_barrier_set_physical(mesh.context(), m_physical ? 1 : 0));
    }

    // potential.cpp:36-57 (over the collision set at `vertices`)
    double operator()(const NormalCollisions& collisions, const CollisionMesh& mesh,
                      const MatrixXd& vertices) const
    {
        prepare(collisions, mesh, vertices);
        double e = 0;
        detail::check(mesh.context(), ipcg_barrier_energy(mesh.context(), m_dhat, m_stiffness, &e));
        return e;
    }
    // potential.cpp:59-95
    std::vector<double> gradient(const NormalCollisions& collisions, const CollisionMesh& mesh,
                                 const MatrixXd& vertices) const
    {
        prepare(collisions, mesh, vertices);
        std::vector<double> g(3 * (size_t)mesh.num_vertices());
        detail::check(mesh.context(),
                      ipcg_barrier_gradient(mesh.context(), m_dhat, m_stiffness, g.data()));
        return g;
    }
    // potential.cpp:97-222
    SparseMatrix hessian(const NormalCollisions& collisions, const CollisionMesh& mesh,
                         const MatrixXd& vertices,
                         PSDProjectionMethod project_hessian_to_psd = PSDProjectionMethod::NONE) const
    {
        prepare(collisions, mesh, vertices);
        ipcg_ctx* c = mesh.context();
        int64_t nnz = 0;
        detail::check(c, ipcg_barrier_hessian(c, m_dhat, m_stiffness, int(project_hessian_to_psd),
                                              &nnz));
        SparseMatrix H;
        H.rows = H.cols = 3 * (int64_t)mesh.num_vertices();
        H.outer.resize(H.rows + 1);
        H.inner.resize(nnz);
        H.values.resize(nnz);
        detail::check(c, ipcg_hessian_download(c, H.outer.data(), H.inner.data(), H.values.data()));
        return H;
    }

    // normal_potential.cpp:15-46
    std::vector<double> gauss_newton_hessian_diagonal(const NormalCollisions& collisions,
                                                      const CollisionMesh& mesh,
                                                      const MatrixXd& vertices) const
    {
        prepare(collisions, mesh, vertices);
        std::vector<double> d(3 * (size_t)mesh.num_vertices());
        detail::check(mesh.context(), ipcg_barrier_gauss_newton_diagonal(
                                          mesh.context(), m_dhat, m_stiffness, d.data()));
        return d;
    }
    // normal_potential.cpp:48-79; p: 3 #V values, RowMajor
    double gauss_newton_hessian_quadratic_form(const NormalCollisions& collisions,
                                               const CollisionMesh& mesh,
                                               const MatrixXd& vertices,
                                               const std::vector<double>& p) const
    {
        if (p.size() != 3 * (size_t)mesh.num_vertices())
            throw std::invalid_argument("p must hold 3 #V values");
        prepare(collisions, mesh, vertices);
        double v = 0;
        detail::check(mesh.context(), ipcg_barrier_gauss_newton_quadratic_form(
                                          mesh.context(), m_dhat, m_stiffness, p.data(), &v));
        return v;
    }

private:
    double m_dhat, m_stiffness;
    bool m_physical;
    // the device evaluates the collision set it holds at slot-0 positions
    void prepare(const NormalCollisions& collisions, const CollisionMesh& mesh,
                 const MatrixXd& vertices) const
    {
        if (collisions.mesh() != &mesh)
            throw std::runtime_error("collisions were not built for this mesh");
        if (!collisions.resident())
            throw std::runtime_error("a later collision build replaced this set on the device");
        mesh.set_positions(0, vertices);
        apply(mesh);
    }
};

// collisions/tangential/tangential_collisions.hpp: friction records over a
// normal collision set (tangential_collisions.cpp:61-186), with scalar or
// per-vertex static and kinetic coefficients
class TangentialCollisions {
public:
    void build(const CollisionMesh& mesh, const MatrixXd& vertices,
               const NormalCollisions& collisions, const BarrierPotential& normal_potential,
               double mu_s, double mu_k)
    {
        begin(mesh, vertices, collisions);
        normal_potential.apply(mesh);
        detail::check(mesh.context(), ipcg_friction_build(mesh.context(), normal_potential.dhat(),
                                                          normal_potential.stiffness(), mu_s, mu_k));
        finish(mesh, collisions);
    }
    void build(const CollisionMesh& mesh, const MatrixXd& vertices,
               const NormalCollisions& collisions, const BarrierPotential& normal_potential,
               double mu)
    {
        build(mesh, vertices, collisions, normal_potential, mu, mu);
    }
    // per-vertex coefficients, averaged per collision (default_blend_mu)
    void build(const CollisionMesh& mesh, const MatrixXd& vertices,
               const NormalCollisions& collisions, const BarrierPotential& normal_potential,
               const std::vector<double>& mu_s, const std::vector<double>& mu_k)
    {
        if (mu_s.size() != (size_t)mesh.num_vertices() || mu_k.size() != mu_s.size())
            throw std::invalid_argument("one friction coefficient per vertex");
        begin(mesh, vertices, collisions);
        normal_potential.apply(mesh);
        detail::check(mesh.context(),
                      ipcg_friction_build_per_vertex(mesh.context(), normal_potential.dhat(),
                                                     normal_potential.stiffness(), mu_s.data(),
                                                     mu_k.data()));
        finish(mesh, collisions);
    }
    size_t size() const { return m_size; }
    bool empty() const { return m_size == 0; }

    // tangential_collision.hpp:108-139: per collision of the normal set (in
    // its order), 2 values each; an empty vector keeps the current values.
    // Resets the effective coefficients (reset_lagged_anisotropic_friction_
    // coefficients).
    void set_anisotropy(const CollisionMesh& mesh, const std::vector<double>& mu_aniso,
                        const std::vector<double>& mu_s_aniso = {},
                        const std::vector<double>& mu_k_aniso = {}) const
    {
        if (!resident(mesh)) throw std::runtime_error("the friction set is not on the device");
        auto ptr = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
        detail::check(mesh.context(), ipcg_friction_set_anisotropy(mesh.context(), ptr(mu_aniso),
                                                                   ptr(mu_s_aniso), ptr(mu_k_aniso)));
    }
    // tangential_collisions.cpp:19-57 at the lagged positions of the build
    void update_lagged_anisotropic_friction_coefficients(const CollisionMesh& mesh,
                                                         const MatrixXd& velocities) const
    {
        if (!resident(mesh)) throw std::runtime_error("the friction set is not on the device");
        if (velocities.rows() != mesh.num_vertices() || (velocities.rows() && velocities.cols() != 3))
            throw std::invalid_argument("velocities must be #V x 3");
        detail::check(mesh.context(), ipcg_velocities_upload(mesh.context(), velocities.data(),
                                                             IPCG_LAYOUT_AOS));
        detail::check(mesh.context(), ipcg_friction_update_lagged_mu(mesh.context()));
    }
    // the device holds this set (no later collision build on the mesh)
    bool resident(const CollisionMesh& mesh) const
    {
        return built_for == &mesh && mesh.collisions_generation() == m_generation;
    }

private:
    const CollisionMesh* built_for = nullptr;
    uint64_t m_generation = 0;
    size_t m_size = 0;
    static void begin(const CollisionMesh& mesh, const MatrixXd& vertices,
                      const NormalCollisions& collisions)
    {
        if (collisions.mesh() != &mesh)
            throw std::runtime_error("collisions were not built for this mesh");
        if (!collisions.resident())
            throw std::runtime_error("a later collision build replaced this set on the device");
        mesh.set_positions(0, vertices);
    }
    void finish(const CollisionMesh& mesh, const NormalCollisions& collisions)
    {
        std::vector<int32_t> keep(collisions.size());
        detail::check(mesh.context(),
                      ipcg_friction_download(mesh.context(), keep.data(), nullptr, nullptr, nullptr));
        m_size = 0;
        for (int32_t k : keep) m_size += k != 0;
        built_for = &mesh;
        m_generation = mesh.collisions_generation();
    }
};

// potentials/friction_potential.hpp (tangential_potential.cpp:156-331) at
// velocities #V x 3, with the set's anisotropy and effective coefficients
class FrictionPotential {
public:
    explicit FrictionPotential(double eps_v) : m_eps_v(eps_v)
    {
        if (!(eps_v > 0)) throw std::invalid_argument("eps_v must be positive");
    }
    double eps_v() const { return m_eps_v; }

    double operator()(const TangentialCollisions& collisions, const CollisionMesh& mesh,
                      const MatrixXd& velocities) const
    {
        prepare(collisions, mesh, velocities);
        double e = 0;
        detail::check(mesh.context(), ipcg_friction_energy(mesh.context(), m_eps_v, &e));
        return e;
    }
    std::vector<double> gradient(const TangentialCollisions& collisions,
                                 const CollisionMesh& mesh, const MatrixXd& velocities) const
    {
        prepare(collisions, mesh, velocities);
        std::vector<double> g(3 * (size_t)mesh.num_vertices());
        detail::check(mesh.context(), ipcg_friction_gradient(mesh.context(), m_eps_v, g.data()));
        return g;
    }
    SparseMatrix hessian(const TangentialCollisions& collisions, const CollisionMesh& mesh,
                         const MatrixXd& velocities,
                         PSDProjectionMethod project_hessian_to_psd = PSDProjectionMethod::NONE) const
    {
        prepare(collisions, mesh, velocities);
        ipcg_ctx* c = mesh.context();
        int64_t nnz = 0;
        detail::check(c, ipcg_friction_hessian(c, m_eps_v, int(project_hessian_to_psd), &nnz));
        SparseMatrix H;
        H.rows = H.cols = 3 * (int64_t)mesh.num_vertices();
        H.outer.resize(H.rows + 1);
        H.inner.resize(nnz);
        H.values.resize(nnz);
        detail::check(c, ipcg_hessian_download(c, H.outer.data(), H.inner.data(), H.values.data()));
        return H;
    }

private:
    double m_eps_v;
    static void prepare(const TangentialCollisions& collisions, const CollisionMesh& mesh,
                        const MatrixXd& velocities)
    {
        if (!collisions.resident(mesh))
            throw std::runtime_error("the friction set is not the one on the device for this mesh");
        if (velocities.rows() != mesh.num_vertices() || (velocities.rows() && velocities.cols() != 3))
            throw std::invalid_argument("velocities must be #V x 3");
        detail::check(mesh.context(), ipcg_velocities_upload(mesh.context(), velocities.data(),
                                                             IPCG_LAYOUT_AOS));
    }
};

// ipc.cpp:45-101
inline double compute_collision_free_stepsize(const CollisionMesh& mesh,
                                              const MatrixXd& vertices_t0,
                                              const MatrixXd& vertices_t1,
                                              double min_distance = 0,
                                              BroadPhaseMethod bp = BroadPhaseMethod::LBVH,
                                              const TightInclusionCCD& ccd = {})
{
    mesh.set_positions(0, vertices_t0);
    mesh.set_positions(1, vertices_t1);
    mesh.touch_broad_phase();
    double step = 1;
    detail::check(mesh.context(),
                  ipcg_compute_collision_free_stepsize(mesh.context(), min_distance, ccd.tolerance,
                                                       ccd.max_iterations,
                                                       ccd.conservative_rescaling, int(bp), &step));
    return step;
}

// ipc.cpp:20-41
inline bool is_step_collision_free(const CollisionMesh& mesh, const MatrixXd& vertices_t0,
                                   const MatrixXd& vertices_t1, double min_distance = 0,
                                   BroadPhaseMethod bp = BroadPhaseMethod::LBVH,
                                   const TightInclusionCCD& ccd = {})
{
    mesh.set_positions(0, vertices_t0);
    mesh.set_positions(1, vertices_t1);
    mesh.touch_broad_phase();
    int free_ = 0;
    detail::check(mesh.context(),
                  ipcg_is_step_collision_free(mesh.context(), min_distance, ccd.tolerance,
                                              ccd.max_iterations, ccd.conservative_rescaling,
                                              int(bp), &free_));
    return free_ != 0;
}

// ipc.cpp:105-166 (3D): some edge crosses some triangle.  The broad phase
// method does not change the answer, so none is taken.
inline bool has_intersections(const CollisionMesh& mesh, const MatrixXd& vertices)
{
    mesh.set_positions(0, vertices);
    mesh.touch_broad_phase();
    int hit = 0;
    detail::check(mesh.context(), ipcg_has_intersections(mesh.context(), &hit));
    return hit != 0;
}

// ---- Eigen interoperability ------------------------------------------------
// Where Eigen is available (the reference's own types), matrices convert to
// and from it: the mirror's dense matrices are row-major copies of Eigen's
// column-major ones, gradients become VectorXd and the Hessian (symmetric,
// so CSR equals CSC) an Eigen::SparseMatrix<double> with int indices, as the
// reference returns it.
#if __has_include(<Eigen/Core>) && __has_include(<Eigen/Sparse>)
} // namespace ipc
#include <Eigen/Core>
#include <Eigen/Sparse>
namespace ipc {

template <class T, class E> inline Matrix<T> from_eigen(const E& M)
{
    Matrix<T> out((int)M.rows(), (int)M.cols());
    for (int i = 0; i < out.rows(); i++)
        for (int j = 0; j < out.cols(); j++) out(i, j) = static_cast<T>(M(i, j));
    return out;
}

inline MatrixXd from_eigen(const Eigen::MatrixXd& M) { return from_eigen<double>(M); }
inline MatrixXi from_eigen(const Eigen::MatrixXi& M) { return from_eigen<int32_t>(M); }

inline Eigen::MatrixXd to_eigen(const MatrixXd& M)
{
    Eigen::MatrixXd out(M.rows(), M.cols());
    for (int i = 0; i < M.rows(); i++)
        for (int j = 0; j < M.cols(); j++) out(i, j) = M(i, j);
    return out;
}

inline Eigen::VectorXd to_eigen(const std::vector<double>& v)
{
    Eigen::VectorXd out((Eigen::Index)v.size());
    std::copy(v.begin(), v.end(), out.data());
    return out;
}

inline Eigen::SparseMatrix<double> to_eigen(const SparseMatrix& H)
{
    if (H.nonZeros() > INT32_MAX) throw std::runtime_error("Hessian too large for int indices");
    Eigen::SparseMatrix<double> out((Eigen::Index)H.rows, (Eigen::Index)H.cols);
    out.resizeNonZeros((Eigen::Index)H.nonZeros());
    std::transform(H.outer.begin(), H.outer.end(), out.outerIndexPtr(),
                   [](int64_t x) { return static_cast<int>(x); });
    std::copy(H.inner.begin(), H.inner.end(), out.innerIndexPtr());
    std::copy(H.values.begin(), H.values.end(), out.valuePtr());
    return out;
}
#endif

} // namespace ipc
