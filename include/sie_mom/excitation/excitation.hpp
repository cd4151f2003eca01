// Incident sources bound to their medium: PreparedExcitation / prepare_source
// (include/sie_mom/excitation/excitation.hpp:14-90, src/sie_mom/excitation/excitation.cu:110-137).
// Plane waves are supported; dipoles are rejected at preparation time.
#pragma once

#include <variant>

#include "sie_mom/fields.hpp"

namespace sie_mom {
namespace excitation {

enum class SourceRegion : int { interior = 1, exterior = 2 };

struct DipoleExcitation {
    vec3_t effective_moment = make_vec3(0, 0, 1);
    vec3_t position = make_vec3(0, 0, 0);
    bool is_magnetic = false;
    SourceRegion region = SourceRegion::interior;
};

struct PlaneWaveExcitation {
    vec3_t electric_amplitude = make_vec3(1, 0, 0);
    vec3_t propagation_direction = make_vec3(0, 0, 1);
    SourceRegion region = SourceRegion::exterior;
};

using SourceSpec = std::variant<DipoleExcitation, PlaneWaveExcitation>;

class PreparedExcitation {
public:
    SourceRegion region() const { return region_; }
    // The plane wave with k and eta of its medium and a unit direction.
    const sie_plane_wave& plane_wave() const { return wave_; }
    // Rejects use with a solver whose frequency or media differ (std::invalid_argument).
    void validate_environment(real_t frequency_hz, const IsotropicMedium& interior,
                              const IsotropicMedium& exterior, real_t vacuum_permittivity,
                              real_t vacuum_permeability) const;

private:
    sie_plane_wave wave_{};
    SourceRegion region_ = SourceRegion::exterior;
    real_t frequency_hz_ = 0;
    IsotropicMedium medium_;
    real_t vacuum_permittivity_ = 0, vacuum_permeability_ = 0;
    friend PreparedExcitation prepare_source(const SourceSpec&, real_t, const IsotropicMedium&,
                                             const IsotropicMedium&, real_t, real_t);
};

PreparedExcitation prepare_source(const SourceSpec& specification, real_t frequency_hz,
                                  const IsotropicMedium& interior, const IsotropicMedium& exterior,
                                  real_t vacuum_permittivity = kCodata2022Eps0,
                                  real_t vacuum_permeability = kCodata2022Mu0);

}  // namespace excitation
}  // namespace sie_mom
