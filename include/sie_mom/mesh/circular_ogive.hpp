// The circular-arc spindle of the reference's gmsh CircularOgive
// (src/sie_mom/mesh/circular_ogive.py:13-80) as a deterministic mesher: the
// surface of revolution of a circular arc of radius arc_radius about its chord
// of length chord, n_rings rings of triangles, outward CCW, RWG edges in
// first-encounter order.
#pragma once

#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
namespace mesh {

MeshData make_spindle(double chord, double arc_radius, int n_rings);

}  // namespace mesh
}  // namespace sie_mom
