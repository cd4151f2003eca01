// Field samples and the surface-current field evaluator of the C++ interface:
// ComplexCartesianVector / ElectromagneticFieldSample (include/sie_mom/fields.hpp:
// 18-45 of the reference) and solvers::evaluate_surface_integral_fields
// (include/sie_mom/solvers/surface_current_fields.hpp; implementation
// src/sie_mom/solvers/surface_current_fields.cu:170-299), on libsie_amd.
#pragma once

#include <array>
#include <complex>
#include <cstddef>
#include <vector>

#include "sie_mom/core.hpp"
#include "sie_mom/mesh.hpp"

namespace sie_mom {

using ComplexCartesianVector = std::array<std::complex<real_t>, 3>;

struct ElectromagneticFieldSample {
    ComplexCartesianVector electric;
    ComplexCartesianVector magnetic;
};

inline ElectromagneticFieldSample subtract_fields(const ElectromagneticFieldSample& left,
                                                  const ElectromagneticFieldSample& right) {
    ElectromagneticFieldSample diff;
    for (std::size_t c = 0; c < 3; ++c) {
        diff.electric[c] = left.electric[c] - right.electric[c];
        diff.magnetic[c] = left.magnetic[c] - right.magnetic[c];
    }
    return diff;
}

inline ElectromagneticFieldSample operator-(const ElectromagneticFieldSample& left,
                                            const ElectromagneticFieldSample& right) {
    return subtract_fields(left, right);
}

namespace solvers {

// E and H radiated at the observation points by physical per-edge currents
// (E = jw mu L(J) + K(M), H = -K(J) + jw eps L(M) in the medium). Same
// argument checks and messages as the reference (std::invalid_argument,
// std::overflow_error, std::runtime_error without a device).
std::vector<ElectromagneticFieldSample> evaluate_surface_integral_fields(
    const mesh::MeshData& mesh, const std::vector<std::complex<real_t>>& electric_currents,
    const std::vector<std::complex<real_t>>& magnetic_currents, const std::vector<vec3_t>& observation_points,
    real_t frequency_hz, const IsotropicMedium& medium, int quadrature_rule = 17,
    real_t vacuum_permittivity = kCodata2022Eps0, real_t vacuum_permeability = kCodata2022Mu0);

}  // namespace solvers
}  // namespace sie_mom
