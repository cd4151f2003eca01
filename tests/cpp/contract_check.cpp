// Compile-time drop-in check: a caller written against the reference's
// headers (the includes, namespaces, global host types and signatures that
// tests/cpp/gstc_sie_contract_test.cpp:15-33 and the reference's apps use)
// compiles against include/ unchanged. Built with a plain C++17 compiler
// (-fsyntax-only) by tests/test_cpp_interface.py; nothing here runs.
#include "sie_mom/excitation/excitation.hpp"
#include "sie_mom/fields.hpp"
#include "sie_mom/gstc_utils.hpp"
#include "sie_mom/mesh/obj_mesh_loader.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"
#include "sie_mom/solvers/dielectric/dielectric_solver.hpp"
#include "sie_mom/solvers/gstc_sie.hpp"
#include "sie_mom/solvers/pec_solver.hpp"
#include "sie_mom/solvers/surface_current_fields.hpp"

#include <complex>
#include <string>
#include <type_traits>
#include <vector>

// the two GSTC entry points, pinned by their exact function types
using GstcSingle = sie_mom::solvers::GstcSieResult (*)(const sie_mom::mesh::MeshData&,
                                                        const std::vector<std::complex<real_t>>&,
                                                        const sie_mom::solvers::GstcSieOptions&,
                                                        const sie_mom::excitation::PreparedExcitation&);
using GstcBatch = sie_mom::solvers::GstcSieBatchResult (*)(const sie_mom::mesh::MeshData&,
                                                            const std::vector<std::complex<real_t>>&,
                                                            const sie_mom::solvers::GstcSieOptions&,
                                                            const std::vector<sie_mom::solvers::GstcExcitationCase>&);
static_assert(std::is_same_v<decltype(&sie_mom::solvers::solve_gstc_dense), GstcSingle>);
static_assert(std::is_same_v<decltype(&sie_mom::solvers::solve_gstc_dense_batch), GstcBatch>);

// the PEC and dielectric entry points
using PecSolve = sie_mom::solvers::PecResult (*)(const sie_mom::mesh::MeshData&, const sie_mom::solvers::PecOptions&,
                                                  const sie_mom::excitation::PreparedExcitation&);
static_assert(std::is_same_v<decltype(&sie_mom::solvers::solve_pec_dense), PecSolve>);
using DielectricSolve = sie_mom::solvers::DielectricResult (*)(const sie_mom::mesh::MeshData&,
                                                                const sie_mom::solvers::DielectricOptions&,
                                                                const sie_mom::excitation::PreparedExcitation&);
static_assert(std::is_same_v<decltype(&sie_mom::solvers::solve_dielectric_dense), DielectricSolve>);

// global host types of the reference's platform header, FP64 build
static_assert(std::is_same_v<real_t, double>);
static_assert(std::is_same_v<decltype(std::declval<vec3_t>().x), double>);
static_assert(std::is_same_v<decltype(sie_mom::mesh::MeshData{}.faces), std::vector<int3>>);
static_assert(std::is_same_v<decltype(sie_mom::mesh::MeshData{}.edges), std::vector<int2>>);
static_assert(std::is_same_v<decltype(sie_mom::mesh::MeshData{}.vertices), std::vector<vec3_t>>);

// a one-triangle mesh built the way the reference's contract test builds it
inline sie_mom::mesh::MeshData one_triangle() {
    sie_mom::mesh::MeshData m;
    m.vertices = {make_vec3(real_t(0), real_t(0), real_t(0)), make_vec3(real_t(1), real_t(0), real_t(0)),
                  make_vec3(real_t(0), real_t(1), real_t(0))};
    m.faces = {make_int3(0, 1, 2)};
    m.edges = {make_int2(0, 1), make_int2(1, 2), make_int2(2, 0)};
    m.faces_edges = {make_int3(0, 1, 2)};
    return m;
}

// options defaults a caller relies on
inline bool face_basis_default() {
    sie_mom::solvers::GstcSieOptions o;
    return o.basis == sie_mom::GstcBasisType::face && o.quadrature_rule == 17 &&
           o.singularity_threshold_lambda0 == real_t(0.1);
}

// the app-side loaders: OBJ class and the tensor loader with .x / .y values
inline std::vector<std::complex<real_t>> load_tensors(const std::string& path, std::size_t faces) {
    const auto raw = sie_mom::gstc::load_face_tensors_from_file(path, faces);
    std::vector<std::complex<real_t>> t;
    for (const auto& v : raw) t.emplace_back(v.x, v.y);
    return t;
}
inline std::size_t obj_faces(const std::string& path) {
    const sie_mom::mesh::ObjMeshLoader loader(path);
    return loader.getMeshData().faces.size();
}
inline std::size_t sphere_edges() { return sie_mom::mesh::SphericalSurface(1.0f, 2).getMeshData().edges.size(); }

// sources and media
inline sie_mom::excitation::PreparedExcitation exterior_wave() {
    sie_mom::excitation::PlaneWaveExcitation pw;
    pw.electric_amplitude = make_vec3(1, 0, 0);
    pw.propagation_direction = make_vec3(0, 0, 1);
    sie_mom::IsotropicMedium vac;
    return sie_mom::excitation::prepare_source(pw, 1e9, vac, vac);
}
