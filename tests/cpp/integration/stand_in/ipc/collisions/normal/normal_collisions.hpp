// Stand-in: the reference's normal collision set (normal_collisions.hpp,
// normal/*.hpp): the four collision vectors, their constructors with weight
// and weight gradient, dmin per collision, the area-weighting flag.
#pragma once

#include <ipc/candidates/candidates.hpp>

#include <Eigen/Sparse>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace ipc {

enum class EdgeEdgeDistanceType : uint8_t {
    EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB, AUTO
};

struct NormalCollision {
    NormalCollision(double w, const Eigen::SparseVector<double>& wg) : weight(w), weight_gradient(wg) { }
    virtual ~NormalCollision() = default;
    double dmin = 0;
    double weight = 1;
    Eigen::SparseVector<double> weight_gradient;
};

struct VertexVertexNormalCollision : VertexVertexCandidate, NormalCollision {
    VertexVertexNormalCollision(index_t a, index_t b, double w, const Eigen::SparseVector<double>& g)
        : VertexVertexCandidate(a, b), NormalCollision(w, g)
    {
    }
};
struct EdgeVertexNormalCollision : EdgeVertexCandidate, NormalCollision {
    EdgeVertexNormalCollision(index_t e, index_t v, double w, const Eigen::SparseVector<double>& g)
        : EdgeVertexCandidate(e, v), NormalCollision(w, g)
    {
    }
};
struct EdgeEdgeNormalCollision : EdgeEdgeCandidate, NormalCollision {
    EdgeEdgeNormalCollision(index_t a, index_t b, double eps, double w,
                            const Eigen::SparseVector<double>& g, EdgeEdgeDistanceType t)
        : EdgeEdgeCandidate(a, b), NormalCollision(w, g), eps_x(eps), dtype(t)
    {
    }
    double eps_x;
    EdgeEdgeDistanceType dtype;
};
struct FaceVertexNormalCollision : FaceVertexCandidate, NormalCollision {
    FaceVertexNormalCollision(index_t f, index_t v, double w, const Eigen::SparseVector<double>& g)
        : FaceVertexCandidate(f, v), NormalCollision(w, g)
    {
    }
};

class NormalCollisions {
public:
    size_t size() const
    {
        return vv_collisions.size() + ev_collisions.size() + ee_collisions.size()
            + fv_collisions.size();
    }
    NormalCollision& operator[](size_t i)
    {
        if (i < vv_collisions.size()) return vv_collisions[i];
        i -= vv_collisions.size();
        if (i < ev_collisions.size()) return ev_collisions[i];
        i -= ev_collisions.size();
        if (i < ee_collisions.size()) return ee_collisions[i];
        return fv_collisions[i - ee_collisions.size()];
    }
    void clear()
    {
        vv_collisions.clear();
        ev_collisions.clear();
        ee_collisions.clear();
        fv_collisions.clear();
    }
    bool use_area_weighting() const { return area_weighting; }
    void set_use_area_weighting(bool w) { area_weighting = w; }

    std::vector<VertexVertexNormalCollision> vv_collisions;
    std::vector<EdgeVertexNormalCollision> ev_collisions;
    std::vector<EdgeEdgeNormalCollision> ee_collisions;
    std::vector<FaceVertexNormalCollision> fv_collisions;

private:
    bool area_weighting = false;
};

} // namespace ipc
