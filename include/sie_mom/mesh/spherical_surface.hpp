// Triangle meshes: MeshData and SphericalSurface at the reference's path and
// with its interface (include/sie_mom/mesh/spherical_surface.hpp:10-38), and
// the edge numbering shared by every mesh source.
#pragma once

#include <string>
#include <vector>

#include "sie_mom/core.hpp"

namespace sie_mom {
namespace mesh {

struct MeshData {
    std::vector<vec3_t> vertices;  // (nV, 3)
    std::vector<int3> faces;       // (nF, 3) vertex indices
    std::vector<int2> edges;       // (nE, 2) unique edges, first-encounter order
    std::vector<int3> faces_edges; // (nF, 3) edge indices, local edge j = (v_j, v_{j+1})
};

// Icosphere of 20 * 4^level outward CCW faces (SphericalSurface(radius, level)).
class SphericalSurface {
public:
    SphericalSurface(float radius, int refinement_level);
    const MeshData& getMeshData() const { return data_; }

private:
    MeshData data_;
};

// Edges as (min, max) vertex pairs in first-encounter order over faces (local
// edge j = (v_j, v_{j+1})) and the face-to-edge map of a face list; throws
// std::invalid_argument unless every edge is shared by exactly two faces.
void build_edges(MeshData& mesh);

}  // namespace mesh
}  // namespace sie_mom
