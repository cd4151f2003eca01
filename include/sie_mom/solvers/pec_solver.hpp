// Dense PEC T-EFIE solver with the reference's interface
// (include/sie_mom/solvers/pec_solver.hpp:10-61). Only the T-EFIE formulation
// is on this library's path; N-EFIE / N-MFIE are rejected with
// std::invalid_argument.
#pragma once

#include <cstddef>
#include <vector>

#include "sie_mom/excitation/excitation.hpp"
#include "sie_mom/fields.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
namespace solvers {

enum class PecFormulation { TEfie, NEfie, NMfie };
enum class PecNtlForm { Line, Surface };

struct PecOptions {
    real_t frequency_hz = real_t(0);
    IsotropicMedium exterior;
    int quadrature_rule = 7;
    real_t singularity_threshold_lambda0 = real_t(0.1);
    bool compute_residual = true;
    bool capture_assembled_matrix = false;
    PecFormulation formulation = PecFormulation::TEfie;
    PecNtlForm ntl_form = PecNtlForm::Surface;
    int line_quadrature_points = 20;
    real_t vacuum_permittivity = kCodata2018Eps0;
    real_t vacuum_permeability = kCodata2018Mu0;
};

struct PecMetrics {
    std::size_t vertex_count = 0;
    std::size_t face_count = 0;
    std::size_t edge_count = 0;
    std::size_t matrix_dimension = 0;
    std::size_t dense_matrix_bytes = 0;
    double assembly_seconds = 0.0;
    double solve_seconds = 0.0;
    double residual_seconds = 0.0;
    bool residual_computed = false;
    double absolute_residual = 0.0;
    double relative_residual = 0.0;
    double maximum_j_coefficient = 0.0;
};

struct PecResult {
    std::vector<std::complex<real_t>> exterior_current;   // J2 in mesh edge order
    std::vector<std::complex<real_t>> assembled_matrix;   // column-major, when captured
    PecMetrics metrics;
};

PecResult solve_pec_dense(const mesh::MeshData& mesh, const PecOptions& options,
                          const excitation::PreparedExcitation& source);

}  // namespace solvers
}  // namespace sie_mom
