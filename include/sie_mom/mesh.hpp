// Triangle meshes: MeshData and SphericalSurface as in the reference
// (include/sie_mom/mesh/spherical_surface.hpp:10-38), the circular-arc
// spindle of its gmsh CircularOgive (src/sie_mom/mesh/circular_ogive.py:13-80),
// and a strict OBJ loader (src/sie_mom/mesh/obj_mesh_loader.cpp:83-386).
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

// Revolution of a circular arc of radius arc_radius about its chord of length chord.
MeshData make_spindle(double chord, double arc_radius, int n_rings);

// Edges as (min, max) vertex pairs in first-encounter order over faces (local
// edge j = (v_j, v_{j+1})) and the face-to-edge map of a face list; throws
// std::invalid_argument unless every edge is shared by exactly two faces.
void build_edges(MeshData& mesh);

// Strict Wavefront OBJ reader with the reference loader's contract: `v x y z`
// and triangular `f a b c` records (1-based, negative relative indices and
// `a/t/n` forms accepted; face and vertex order preserved; other records
// ignored), then the closed-surface checks (no unreferenced vertices,
// degenerate or duplicate faces, boundary or non-manifold edges; one connected
// component; consistent orientation; positive signed volume). Throws
// std::runtime_error naming the file, line or problems.
MeshData load_obj(const std::string& path);

}  // namespace mesh
}  // namespace sie_mom
