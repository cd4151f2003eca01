// Dense SIE-GSTC solver with the reference's interface
// (include/sie_mom/solvers/gstc_sie.hpp:32-239): options, metrics and results
// with the same fields and meaning, solve_gstc_dense / solve_gstc_dense_batch
// with the same validation and exception types. One extension: `precision`
// selects the plain FP64 LU (the reference's path) or the FP32-accurate
// matrix-core LU with FP64 iterative refinement to `refinement_tolerance`.
#pragma once

#include <cstddef>
#include <optional>
#include <vector>

#include "sie_mom/excitation/excitation.hpp"
#include "sie_mom/fields.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
namespace solvers {

struct GstcSieOptions {
    real_t frequency_hz = real_t(0);
    IsotropicMedium interior;
    IsotropicMedium exterior;
    int quadrature_rule = 17;
    real_t singularity_threshold_lambda0 = real_t(0.1);
    bool tensors_are_uniform = false;
    bool compute_residual = true;
    GstcBasisType basis = GstcBasisType::face;
    real_t vacuum_permittivity = kCodata2022Eps0;
    real_t vacuum_permeability = kCodata2022Mu0;
    SolvePrecision precision = SolvePrecision::fp64;
    // normwise backward error |b - Z x| / (|Z|_F |x| + |b|) the refinement reaches
    double refinement_tolerance = 1e-15;
    int refinement_max_sweeps = 40;
};

struct PhysicalParameters {
    real_t omega;
    real_t k_interior;
    real_t k_exterior;
    real_t eta0;
    real_t singularity_threshold;
};

PhysicalParameters make_physical_parameters(const GstcSieOptions& options);

struct GstcSieMetrics {
    std::size_t vertex_count = 0;
    std::size_t face_count = 0;
    std::size_t edge_count = 0;
    std::size_t matrix_dimension = 0;
    std::size_t dense_matrix_bytes = 0;
    double device_preparation_seconds = 0.0;
    double assembly_seconds = 0.0;
    double solve_seconds = 0.0;
    double residual_seconds = 0.0;
    double rhs_norm = 0.0;
    double absolute_residual = 0.0;
    double relative_residual = 0.0;
    bool residual_computed = false;
    double maximum_j1_coefficient = 0.0;
    double maximum_m1_coefficient = 0.0;
    double maximum_j2_coefficient = 0.0;
    double maximum_m2_coefficient = 0.0;
};

// physical_currents: [J1, M1, J2, M2], one coefficient per RWG edge per block,
// magnetic blocks multiplied by eta0 (no scaling left).
struct GstcSieResult {
    std::vector<std::complex<real_t>> physical_currents;
    GstcSieMetrics metrics;
};

struct GstcExcitationCase {
    std::optional<excitation::PreparedExcitation> interior;
    std::optional<excitation::PreparedExcitation> exterior;
};

struct GstcSieBatchMetrics {
    std::size_t vertex_count = 0;
    std::size_t face_count = 0;
    std::size_t edge_count = 0;
    std::size_t matrix_dimension = 0;
    std::size_t dense_matrix_bytes = 0;
    std::size_t case_count = 0;
    double device_preparation_seconds = 0.0;
    double matrix_assembly_seconds = 0.0;
    double factorization_seconds = 0.0;
    double solve_seconds = 0.0;
    double residual_matrix_assembly_seconds = 0.0;
};

struct GstcSieCaseMetrics {
    double rhs_assembly_seconds = 0.0;
    double residual_seconds = 0.0;
    double rhs_norm = 0.0;
    double absolute_residual = 0.0;
    double relative_residual = 0.0;
    bool residual_computed = false;
    double maximum_j1_coefficient = 0.0;
    double maximum_m1_coefficient = 0.0;
    double maximum_j2_coefficient = 0.0;
    double maximum_m2_coefficient = 0.0;
};

struct GstcSieCaseResult {
    std::vector<std::complex<real_t>> physical_currents;
    GstcSieCaseMetrics metrics;
};

struct GstcSieBatchResult {
    std::vector<GstcSieCaseResult> cases;
    GstcSieBatchMetrics metrics;
};

// gstc_tensors: 32 complex per face (or 32 for a uniform sheet): the 4x4
// row-major fields-to-currents map followed by its inverse, in the frame of
// options.basis; phasors exp(+j omega t).
GstcSieResult solve_gstc_dense(const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& gstc_tensors,
                               const GstcSieOptions& options, const excitation::PreparedExcitation& source);

// One assembly and one factorisation shared by every case; one RHS per case
// (the sum of its interior and exterior incident contributions).
GstcSieBatchResult solve_gstc_dense_batch(const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& gstc_tensors,
                                          const GstcSieOptions& options,
                                          const std::vector<GstcExcitationCase>& cases);

}  // namespace solvers
}  // namespace sie_mom
