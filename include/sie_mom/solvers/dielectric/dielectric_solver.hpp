// Dense T-PMCHWT dielectric solver with the reference's interface
// (include/sie_mom/solvers/dielectric/dielectric_solver.hpp:15-98). Only
// T-PMCHWT is on this library's path; N-PMCHWT / N-Muller are rejected with
// std::invalid_argument.
#pragma once

#include <cstddef>
#include <vector>

#include "sie_mom/excitation/excitation.hpp"
#include "sie_mom/fields.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
namespace solvers {

enum class DielectricFormulation { T_PMCHWT, N_PMCHWT, N_Muller };
enum class NtlIntegralForm { Line, Surface };

struct DielectricOptions {
    real_t frequency_hz = 0;
    IsotropicMedium interior;
    IsotropicMedium exterior;
    int quadrature_rule = 17;
    real_t singularity_threshold_lambda0 = 0.1;
    bool compute_residual = true;
    real_t vacuum_permittivity = kCodata2018Eps0;
    real_t vacuum_permeability = kCodata2018Mu0;
    DielectricFormulation formulation = DielectricFormulation::T_PMCHWT;
    NtlIntegralForm ntl_form = NtlIntegralForm::Surface;
    int line_quadrature_points = 20;
    // extension: FP32 matrix-core LU + FP64 refinement (GMRES-IR for one
    // right-hand side, blocked refinement sweeps for several)
    SolvePrecision precision = SolvePrecision::fp64;
    // normwise backward error |b - Z x| / (|Z|_F |x| + |b|) the refinement reaches
    double refinement_tolerance = 1e-15;
    int refinement_max_sweeps = 40;
};

struct DielectricMetrics {
    std::size_t vertex_count = 0;
    std::size_t face_count = 0;
    std::size_t edge_count = 0;
    std::size_t matrix_dimension = 0;
    std::size_t dense_matrix_bytes = 0;
    double device_preparation_seconds = 0.0;
    double assembly_seconds = 0.0;
    double solve_seconds = 0.0;
    double residual_seconds = 0.0;
    bool residual_computed = false;
    double absolute_residual = 0.0;
    double relative_residual = 0.0;
    double maximum_j_coefficient = 0.0;
    double maximum_m_coefficient = 0.0;
};

// physical_currents: interior-oriented [J1, M1] (M multiplied by eta0); the
// exterior pair is [-J1, -M1].
struct DielectricResult {
    std::vector<std::complex<real_t>> physical_currents;
    DielectricMetrics metrics;
};

DielectricResult solve_dielectric_dense(const mesh::MeshData& mesh, const DielectricOptions& options,
                                        const excitation::PreparedExcitation& source);

// Many plane waves on one operator (the multi-RHS entry the reference's
// GSTC driver has, solve_gstc_dense_batch gstc_sie.hpp:214-239, for T-PMCHWT):
// one assembly, one factorisation, one RHS column per source; each case
// carries its own residual and coefficient maxima, `metrics` the shared
// timings of the batch.
struct DielectricBatchResult {
    std::vector<DielectricResult> cases;
    DielectricMetrics metrics;
};

DielectricBatchResult solve_dielectric_dense_batch(const mesh::MeshData& mesh, const DielectricOptions& options,
                                                   const std::vector<excitation::PreparedExcitation>& sources);

}  // namespace solvers
}  // namespace sie_mom
