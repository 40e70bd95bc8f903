// C++ host API (include/ipc_amd.hpp) on a device: the reference's codim
// known-answer tests (tests/src/tests/collisions/test_normal_collisions.cpp:
// 14-206) and the squash-tet / vertex-on-plane cases, written against the
// mirrored interface.  Prints one line per check; exit status = failures.
#include "ipc_amd.hpp"

#include <cmath>
#include <cstdio>

static int failures = 0;
#define CHECK(cond)                                                                     \
    do {                                                                                \
        const bool ok_ = (cond);                                                        \
        std::printf("%s %s:%d %s\n", ok_ ? "ok  " : "FAIL", __FILE__, __LINE__, #cond); \
        if (!ok_) failures++;                                                           \
    } while (0)

static bool approx(double a, double b, double rel = 1.1920928955078125e-05)
{
    return std::fabs(a - b) <= rel * std::fmax(std::fabs(a), std::fabs(b));
}

static void codim_vertex_vertex()
{
    using namespace ipc;
    constexpr double thickness = 0.4, min_distance = 2 * thickness;
    MatrixXd V { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 1, 0 }, { 0, 1, 1 },
                 { 1, 0, 0 }, { 1, 0, 1 }, { 1, 1, 0 }, { 1, 1, 1 } };
    for (int i = 0; i < 8; i++)
        for (int k = 0; k < 3; k++) V(i, k) -= 0.5;
    CollisionMesh mesh(V);
    CHECK(mesh.num_vertices() == 8);
    CHECK(mesh.num_codim_vertices() == 8);
    CHECK(mesh.num_edges() == 0 && mesh.num_faces() == 0);

    MatrixXd V1 = V;
    for (int i = 0; i < 8; i++) V1(i, 1) *= 0.5;
    Candidates candidates;
    candidates.build(mesh, V, V1, thickness);
    CHECK(!candidates.empty());
    CHECK(candidates.vv_candidates.size() == candidates.size());
    CHECK(!candidates.is_step_collision_free(mesh, V, V1, min_distance));
    const double expected_toi = (1 - (min_distance + 1e-4)) / 2.0 / 0.25;
    CHECK(approx(candidates.compute_collision_free_stepsize(mesh, V, V1, min_distance),
                 expected_toi));

    const double dhat = 0.25;
    NormalCollisions collisions;
    collisions.build(mesh, V, dhat, min_distance);
    CHECK(collisions.size() == 12);
    CHECK(collisions.vv_collisions.size() == 12);
    BarrierPotential barrier_potential(dhat, 1.0);
    CHECK(barrier_potential(collisions, mesh, V) > 0.0);
    const std::vector<double> grad = barrier_potential.gradient(collisions, mesh, V);
    bool radial = true;
    for (int i = 0; i < 8; i++) {
        double f[3], n = 0, m = 0, d = 0;
        for (int k = 0; k < 3; k++) {
            f[k] = -grad[3 * i + k];
            n += f[k] * f[k];
            m += V(i, k) * V(i, k);
        }
        for (int k = 0; k < 3; k++) d += f[k] * V(i, k);
        radial &= std::fabs(d / std::sqrt(n * m) - 1) < 1e-12;
    }
    CHECK(radial);
}

static void codim_edge_vertex()
{
    using namespace ipc;
    constexpr double thickness = 1e-3, min_distance = 2 * thickness;
    const MatrixXd V { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 0, -1 }, { -1, 0, 0 },
                       { 0, 0, 1 }, { 0, 1, 0 }, { 0, 2, 0 }, { 0, 3, 0 } };
    const MatrixXi E { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 0, 4 } };
    CollisionMesh mesh(V, E);
    CHECK(mesh.num_vertices() == 8);
    CHECK(mesh.num_codim_vertices() == 3);
    CHECK(mesh.num_codim_edges() == 4);

    MatrixXd V1 = V;
    for (int i = 5; i < 8; i++) V1(i, 1) -= 4;
    Candidates candidates;
    candidates.build(mesh, V, V1, thickness);
    CHECK(candidates.size() == 15);
    CHECK(candidates.vv_candidates.size() == 3);
    CHECK(candidates.ev_candidates.size() == 12);
    CHECK(candidates.ee_candidates.size() == 0);
    CHECK(candidates.fv_candidates.size() == 0);
    CHECK(!candidates.is_step_collision_free(mesh, V, V1, min_distance));
    const double expected_toi = (1 - (min_distance + 1e-4)) / 4;
    CHECK(approx(candidates.compute_collision_free_stepsize(mesh, V, V1, min_distance),
                 expected_toi));

    NormalCollisions collisions;
    const double dhat = 0.25;
    collisions.build(mesh, V, dhat, /*min_distance=*/0.8);
    CHECK(collisions.size() == 6);
    CHECK(collisions.vv_collisions.size() == 2);
    CHECK(collisions.ev_collisions.size() == 4);
    CHECK(collisions.ee_collisions.size() == 0);
    CHECK(collisions.fv_collisions.size() == 0);
    CHECK(BarrierPotential(dhat, 1.0)(collisions, mesh, V) > 0.0);
}

static void vertex_on_plane()
{
    using namespace ipc;
    const double h = 1e-2;
    const MatrixXd V0 { { -10, -10, 0 }, { 10, -10, 0 }, { 0, 10, 0 }, { 0.1, 0.2, h } };
    MatrixXd V1 = V0;
    V1(3, 2) = -h;
    const MatrixXi F { { 0, 1, 2 } };
    const MatrixXi E { { 0, 1 }, { 0, 2 }, { 1, 2 } };
    CollisionMesh mesh(V0, E, F);
    const double t = compute_collision_free_stepsize(mesh, V0, V1);
    const double t_c = 0.5 - 1e-4 / (2 * h);
    CHECK(t <= t_c && t >= t_c - 2e-6 / (2 * h) - 1e-9);
    MatrixXd Vt = V0;
    for (int k = 0; k < 3; k++) Vt(3, k) = V0(3, k) + t * (V1(3, k) - V0(3, k));
    CHECK(is_step_collision_free(mesh, V0, Vt));
    CHECK(!is_step_collision_free(mesh, V0, V1));
    // the same through the sweep
    CHECK(compute_collision_free_stepsize(mesh, V0, V1, 0, BroadPhaseMethod::SWEEP_AND_PRUNE) == t);
    // a vertex hovering at 0.5 dhat over the face: one face-vertex collision
    NormalCollisions collisions;
    MatrixXd Vh = V0;
    Vh(3, 2) = 5e-4;
    collisions.build(mesh, Vh, 1e-3);
    CHECK(collisions.fv_collisions.size() == 1 && collisions.size() == 1);
    BarrierPotential B(1e-3, 1e3);
    const SparseMatrix H = B.hessian(collisions, mesh, Vh, PSDProjectionMethod::CLAMP);
    CHECK(H.rows == 12 && H.nonZeros() > 0);
    CHECK(H.coeff(11, 11) > 0);
}

static void errors_throw()
{
    using namespace ipc;
    bool threw = false;
    try {
        CollisionMesh mesh(MatrixXd { { 0, 0, 0 }, { 1, 0, 0 } }, MatrixXi { { 0, 5 } });
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        CollisionMesh mesh(MatrixXd { { 0, 0, 0 }, { 1, 0, 0 } });
        BarrierPotential(1e-3)(NormalCollisions(), mesh, MatrixXd { { 0, 0, 0 }, { 1, 0, 0 } });
    } catch (const std::runtime_error&) {
        threw = true;
    }
    CHECK(threw);
}

// Vertex-pair filters (collision_filter.hpp:113-152) on the codim cube:
// its 12 vertex-vertex collisions are the cube's edges; vertex i has x = 1
// iff i >= 4.  Static obstacles {4..7} drop the 4 edges of the x = 1 face;
// the codim cross filter with the first 4 vertices keeps the 4 edges along
// x; without faces every vertex is its own connected component.
static void collision_filters()
{
    using namespace ipc;
    MatrixXd V { { 0, 0, 0 }, { 0, 0, 1 }, { 0, 1, 0 }, { 0, 1, 1 },
                 { 1, 0, 0 }, { 1, 0, 1 }, { 1, 1, 0 }, { 1, 1, 1 } };
    const double dhat = 0.25, dmin = 0.8;
    CollisionMesh mesh(V);
    auto count = [&](const CollisionFilter& f) {
        mesh.set_can_collide(f);
        NormalCollisions c;
        c.build(mesh, V, dhat, dmin);
        return c.vv_collisions.size();
    };
    CHECK(count(CollisionFilter()) == 12);
    CHECK(count(make_static_obstacle_filter(4)) == 8);
    CHECK(count(make_codim_cross_filter(4)) == 4);
    CHECK(count(make_static_obstacle_filter(4) & make_codim_cross_filter(4)) == 4);
    CHECK(count(make_static_obstacle_filter(0)) == 0);
    CHECK(make_connected_components_filter(MatrixXi(0, 3), 8)(0, 1));
    CHECK(count(CollisionFilter()) == 12);
    MatrixXi F { { 0, 1, 2 }, { 4, 5, 6 } };
    const CollisionFilter cc = make_connected_components_filter(F, 8);
    CHECK(!cc(0, 2) && cc(0, 4) && !cc(4, 6) && cc(3, 7));
    bool threw = false;
    try {
        mesh.set_can_collide(cc); // patch labels belong to the uploaded mesh
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
}

// A triangle and an edge through it (ipc.cpp:105-166): the edge's
// endpoints on either side of the plane and its crossing point inside the
// triangle; moved aside, the same edge misses.
static void intersections()
{
    using namespace ipc;
    const MatrixXi F { { 0, 1, 2 } };
    const MatrixXi E { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 3, 4 } };
    MatrixXd V { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0.25, 0.25, -1 }, { 0.25, 0.25, 1 } };
    CollisionMesh mesh(V, E, F);
    CHECK(has_intersections(mesh, V));
    V(3, 0) = V(4, 0) = 2.0;
    CHECK(!has_intersections(mesh, V));
    V(3, 0) = V(4, 0) = 0.25;
    V(3, 2) = 0.0; // touching the triangle from above counts (t = 0)
    CHECK(has_intersections(mesh, V));
}

// test_per_vertex_safe_distances.cpp: three points, inflation 10, minimum
// distance 0.5; then the noncandidate step for a uniform displacement of
// norm 13/1024 and the CFL step against the full CCD (test_cfl.cpp).
static void candidate_helpers()
{
    using namespace ipc;
    const MatrixXd V { { -1, 0, 0 }, { 1, 0, 0 }, { 10, 0, 0 } };
    CollisionMesh mesh(V);
    Candidates C;
    C.build(mesh, V, 10.0);
    CHECK(C.vv_candidates.size() == 3);
    const std::vector<double> d = C.compute_per_vertex_safe_distances(mesh, V, 10.0, 0.5);
    CHECK(d.size() == 3 && d[0] == 1.5 && d[1] == 1.5 && d[2] == 8.5);
    MatrixXd D(3, 3);
    for (int i = 0; i < 3; i++) {
        D(i, 0) = 3.0 / 1024;
        D(i, 1) = 4.0 / 1024;
        D(i, 2) = 12.0 / 1024;
    }
    CHECK(C.compute_noncandidate_conservative_stepsize(mesh, D, 1e-3) == 0.5 * 1e-3 / (13.0 / 1024));
    MatrixXd V1 = V;
    V1(0, 0) = 0.5;
    V1(1, 0) = -0.5; // points 0 and 1 pass through each other
    const double cfl = C.compute_cfl_stepsize(mesh, V, V1, 1e-3);
    const double ccd = compute_collision_free_stepsize(mesh, V, V1);
    CHECK(cfl <= ccd && ccd < 1.0 && cfl > 0.0);
}

// Area weighting (normal_collisions_builder.cpp:26-54): isolated points
// weigh 1 each, so a point pair weighs 0.5 (1 + 1) = 1; a vertex hovering over
// a unit right triangle contributes 0.25 of its area (1 for a point with no
// edge or face) to the face-vertex collision.
static void area_weighting()
{
    using namespace ipc;
    const MatrixXd V { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 1, 0 }, { 0.2, 0.2, 5e-4 } };
    const MatrixXi F { { 0, 1, 2 } };
    const MatrixXi E { { 0, 1 }, { 0, 2 }, { 1, 2 } };
    CollisionMesh mesh(V, E, F);
    NormalCollisions C;
    C.set_use_area_weighting(true);
    C.build(mesh, V, 1e-3);
    CHECK(C.fv_collisions.size() == 1 && C.size() == 1);
    CHECK(C.fv_collisions[0].weight == 0.25);
    C.set_use_area_weighting(false);
    C.build(mesh, V, 1e-3);
    CHECK(C.fv_collisions.size() == 1 && C.fv_collisions[0].weight == 1.0);
}

// test_normal_collisions.cpp:120-200: the codim star gains one
// vertex-vertex collision under the improved max approximation and keeps
// its six under OGC
static void improved_max_approx()
{
    using namespace ipc;
    const MatrixXd V { { 0, 0, 0 }, { 1, 0, 0 }, { 0, 0, -1 }, { -1, 0, 0 },
                       { 0, 0, 1 }, { 0, 1, 0 }, { 0, 2, 0 }, { 0, 3, 0 } };
    const MatrixXi E { { 0, 1 }, { 0, 2 }, { 0, 3 }, { 0, 4 } };
    CollisionMesh mesh(V, E);
    NormalCollisions C;
    C.build(mesh, V, 0.25, 0.8);
    CHECK(C.size() == 6 && C.vv_collisions.size() == 2 && C.ev_collisions.size() == 4);
    C.set_collision_set_type(NormalCollisions::CollisionSetType::IMPROVED_MAX_APPROX);
    CHECK(C.collision_set_type() == NormalCollisions::CollisionSetType::IMPROVED_MAX_APPROX);
    C.build(mesh, V, 0.25, 0.8);
    CHECK(C.size() == 7 && C.vv_collisions.size() == 3 && C.ev_collisions.size() == 4);
    C.set_collision_set_type(NormalCollisions::CollisionSetType::OGC);
    C.build(mesh, V, 0.25, 0.8);
    CHECK(C.size() == 6 && C.vv_collisions.size() == 2 && C.ev_collisions.size() == 4);
}

// a codim point pair along z: the friction energy at velocity (vx, vy, *)
// of the first point is mu N f0(|(vx, vy)|), N = -kappa b'(d^2) 2 d
static void friction_point_pair()
{
    using namespace ipc;
    const double d = 0.004, dhat = 0.01, kappa = 2.0, mu = 0.4, eps_v = 1e-3;
    const MatrixXd V { { 0, 0, 0 }, { 0, 0, d } };
    CollisionMesh mesh(V);
    NormalCollisions C;
    C.build(mesh, V, dhat);
    CHECK(C.vv_collisions.size() == 1);
    BarrierPotential B(dhat, kappa);
    TangentialCollisions T;
    T.build(mesh, V, C, B, mu);
    CHECK(T.size() == 1);
    const double x = d * d, t = dhat * dhat;
    const double db = -2 * (x - t) * std::log(x / t) - (x - t) * (x - t) / x;
    const double N = -kappa * db * 2 * d;
    const MatrixXd Vel { { 3e-3, -4e-3, 0.5 }, { 0, 0, -0.1 } };
    const double e = FrictionPotential(eps_v)(T, mesh, Vel);
    CHECK(std::abs(e - mu * N * 5e-3) <= 1e-12 * mu * N * 5e-3);
    const std::vector<double> g = FrictionPotential(eps_v).gradient(T, mesh, Vel);
    // sliding: force mu N along u / |u| on the first point, opposite on the second
    CHECK(std::abs(g[0] - mu * N * 0.6) <= 1e-12 * mu * N && std::abs(g[1] + mu * N * 0.8) <= 1e-12 * mu * N);
    CHECK(std::abs(g[3] + g[0]) <= 1e-15 * mu * N && g[2] == 0.0);
}

int main()
{
    try {
        collision_filters();
        codim_vertex_vertex();
        codim_edge_vertex();
        vertex_on_plane();
        errors_throw();
        intersections();
        candidate_helpers();
        area_weighting();
        improved_max_approx();
        friction_point_pair();
    } catch (const std::exception& e) {
        std::printf("FAIL exception: %s\n", e.what());
        failures++;
    }
    std::printf("%d failure(s)\n", failures);
    return failures;
}
