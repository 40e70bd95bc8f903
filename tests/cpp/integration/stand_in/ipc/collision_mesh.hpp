// Stand-in: the reference's collision mesh accessors (collision_mesh.hpp).
#pragma once

#include <Eigen/Core>

#include <cstddef>

namespace ipc {

class CollisionMesh {
public:
    CollisionMesh(const Eigen::MatrixXd& rest, const Eigen::MatrixXi& edges,
                  const Eigen::MatrixXi& faces)
        : rest_(rest), edges_(edges), faces_(faces)
    {
    }
    size_t num_vertices() const { return static_cast<size_t>(rest_.rows()); }
    const Eigen::MatrixXd& rest_positions() const { return rest_; }
    const Eigen::MatrixXi& edges() const { return edges_; }
    const Eigen::MatrixXi& faces() const { return faces_; }

private:
    Eigen::MatrixXd rest_;
    Eigen::MatrixXi edges_, faces_;
};

} // namespace ipc
