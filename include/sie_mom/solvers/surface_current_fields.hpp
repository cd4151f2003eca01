// The surface-current field evaluator of the C++ interface:
// solvers::evaluate_surface_integral_fields at the reference's path and with
// its signature (include/sie_mom/solvers/surface_current_fields.hpp:34-44;
// implementation src/sie_mom/solvers/surface_current_fields.cu:170-299), on
// libsie_amd.
#pragma once

#include <complex>
#include <vector>

#include "sie_mom/fields.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"

namespace sie_mom {
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
