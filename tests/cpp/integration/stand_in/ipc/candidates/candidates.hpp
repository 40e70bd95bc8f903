// Stand-in: the reference's candidate pair types (candidates/*.hpp), as far
// as the integration shim constructs them.
#pragma once

#include <cstdint>
#include <vector>

namespace ipc {

using index_t = int32_t;

struct VertexVertexCandidate {
    VertexVertexCandidate(index_t a, index_t b) : vertex0_id(a), vertex1_id(b) { }
    index_t vertex0_id, vertex1_id;
};
struct EdgeVertexCandidate {
    EdgeVertexCandidate(index_t e, index_t v) : edge_id(e), vertex_id(v) { }
    index_t edge_id, vertex_id;
};
struct EdgeEdgeCandidate {
    EdgeEdgeCandidate(index_t a, index_t b) : edge0_id(a), edge1_id(b) { }
    index_t edge0_id, edge1_id;
};
struct FaceVertexCandidate {
    FaceVertexCandidate(index_t f, index_t v) : face_id(f), vertex_id(v) { }
    index_t face_id, vertex_id;
};
struct EdgeFaceCandidate {
    EdgeFaceCandidate(index_t e, index_t f) : edge_id(e), face_id(f) { }
    index_t edge_id, face_id;
};
struct FaceFaceCandidate {
    FaceFaceCandidate(index_t a, index_t b) : face0_id(a), face1_id(b) { }
    index_t face0_id, face1_id;
};

} // namespace ipc
