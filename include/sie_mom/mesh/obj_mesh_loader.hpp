// Strict OBJ loader at the reference's path and with its interface
// (include/sie_mom/mesh/obj_mesh_loader.hpp: class ObjMeshLoader;
// src/sie_mom/mesh/obj_mesh_loader.cpp:83-386): vertex and face order and
// winding preserved, invalid closed surfaces rejected rather than repaired,
// RWG edges numbered in first-encounter order.
#pragma once

#include <string>

#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
namespace mesh {

// Strict Wavefront OBJ reader: `v x y z` and triangular `f a b c` records
// (1-based, negative relative indices and `a/t/n` forms accepted; face and
// vertex order preserved; other records ignored), then the closed-surface
// checks (no unreferenced vertices, degenerate or duplicate faces, boundary or
// non-manifold edges; one connected component; consistent orientation;
// positive signed volume). Throws std::runtime_error naming the file, line or
// problems.
MeshData load_obj(const std::string& path);

class ObjMeshLoader {
public:
    explicit ObjMeshLoader(const std::string& filepath) : mesh_data_(load_obj(filepath)) {}
    const MeshData& getMeshData() const { return mesh_data_; }

private:
    MeshData mesh_data_;
};

}  // namespace mesh
}  // namespace sie_mom
