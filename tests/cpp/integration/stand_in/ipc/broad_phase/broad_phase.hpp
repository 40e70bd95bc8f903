// Stand-in: the reference's abstract broad phase (broad_phase.hpp:19-133):
// the virtual interface a GPU broad phase overrides.
#pragma once

#include <ipc/candidates/candidates.hpp>
#include <ipc/utils/eigen_ext.hpp>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace ipc {

struct AABB {
    double min[3], max[3];
};
using AABBs = std::vector<AABB>;
using CollisionFilter = std::function<bool(size_t, size_t)>;

class BroadPhase {
public:
    virtual ~BroadPhase() = default;
    virtual std::string name() const = 0;
    virtual void build(Eigen::ConstRef<Eigen::MatrixXd> vertices,
                       Eigen::ConstRef<Eigen::MatrixXi> edges,
                       Eigen::ConstRef<Eigen::MatrixXi> faces, const double inflation_radius = 0)
        = 0;
    virtual void build(Eigen::ConstRef<Eigen::MatrixXd> vertices_t0,
                       Eigen::ConstRef<Eigen::MatrixXd> vertices_t1,
                       Eigen::ConstRef<Eigen::MatrixXi> edges,
                       Eigen::ConstRef<Eigen::MatrixXi> faces, const double inflation_radius = 0)
        = 0;
    virtual void build(const AABBs& vertex_boxes, Eigen::ConstRef<Eigen::MatrixXi> edges,
                       Eigen::ConstRef<Eigen::MatrixXi> faces, const uint8_t dim)
        = 0;
    virtual void clear() = 0;
    virtual void detect_vertex_vertex_candidates(std::vector<VertexVertexCandidate>&) const = 0;
    virtual void detect_edge_vertex_candidates(std::vector<EdgeVertexCandidate>&) const = 0;
    virtual void detect_edge_edge_candidates(std::vector<EdgeEdgeCandidate>&) const = 0;
    virtual void detect_face_vertex_candidates(std::vector<FaceVertexCandidate>&) const = 0;
    virtual void detect_edge_face_candidates(std::vector<EdgeFaceCandidate>&) const = 0;
    virtual void detect_face_face_candidates(std::vector<FaceFaceCandidate>&) const = 0;
    CollisionFilter can_vertices_collide;
};

} // namespace ipc
