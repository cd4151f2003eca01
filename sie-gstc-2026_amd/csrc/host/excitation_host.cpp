// prepare_source / validate_environment (src/sie_mom/excitation/excitation.cu:110-160).
#include <cmath>
#include <limits>

#include "host_common.hpp"
#include "sie_mom/excitation/excitation.hpp"

namespace sie_mom {
namespace excitation {
namespace {

void require_positive(double v, const char* name) {
    if (!(v > 0) || !std::isfinite(v)) throw std::invalid_argument(std::string(name) + " must be positive and finite");
}

const IsotropicMedium& source_medium(SourceRegion r, const IsotropicMedium& interior,
                                     const IsotropicMedium& exterior) {
    if (r == SourceRegion::interior) return interior;
    if (r == SourceRegion::exterior) return exterior;
    throw std::invalid_argument("source region must be interior or exterior");
}

}  // namespace

PreparedExcitation prepare_source(const SourceSpec& specification, real_t frequency_hz,
                                  const IsotropicMedium& interior, const IsotropicMedium& exterior,
                                  real_t vacuum_permittivity, real_t vacuum_permeability) {
    require_positive(frequency_hz, "source frequency");
    require_positive(vacuum_permittivity, "vacuum permittivity");
    require_positive(vacuum_permeability, "vacuum permeability");
    if (std::holds_alternative<DipoleExcitation>(specification))
        throw std::invalid_argument("dipole sources are not supported by this library (plane waves only)");
    const PlaneWaveExcitation& spec = std::get<PlaneWaveExcitation>(specification);
    const IsotropicMedium& medium = source_medium(spec.region, interior, exterior);
    require_positive(medium.relative_permittivity, "source relative permittivity");
    require_positive(medium.relative_permeability, "source relative permeability");
    const double eps = medium.relative_permittivity * vacuum_permittivity;
    const double mu = medium.relative_permeability * vacuum_permeability;
    const double omega = 2.0 * std::acos(-1.0) * frequency_hz;
    const double k = omega * std::sqrt(mu * eps), eta = std::sqrt(mu / eps);
    require_positive(k, "source wavenumber");
    require_positive(eta, "source impedance");
    const vec3_t d = spec.propagation_direction, a = spec.electric_amplitude;
    const double dn = std::hypot(d.x, d.y, d.z), an = std::hypot(a.x, a.y, a.z);
    require_positive(dn, "plane-wave direction magnitude");
    require_positive(an, "plane-wave amplitude magnitude");
    const double u[3] = {d.x / dn, d.y / dn, d.z / dn};
    const double lf = (u[0] * a.x + u[1] * a.y + u[2] * a.z) / an;
    if (std::abs(lf) > 64 * std::numeric_limits<double>::epsilon())
        throw std::invalid_argument("plane-wave electric amplitude must be transverse to propagation");
    PreparedExcitation p;
    p.wave_.E0[0] = a.x;
    p.wave_.E0[1] = a.y;
    p.wave_.E0[2] = a.z;
    for (int i = 0; i < 3; ++i) p.wave_.khat[i] = u[i];
    p.wave_.k = k;
    p.wave_.eta = eta;
    p.wave_.region = (int)spec.region;
    p.region_ = spec.region;
    p.frequency_hz_ = frequency_hz;
    p.medium_ = medium;
    p.vacuum_permittivity_ = vacuum_permittivity;
    p.vacuum_permeability_ = vacuum_permeability;
    return p;
}

void PreparedExcitation::validate_environment(real_t frequency_hz, const IsotropicMedium& interior,
                                              const IsotropicMedium& exterior, real_t vacuum_permittivity,
                                              real_t vacuum_permeability) const {
    const IsotropicMedium& medium = source_medium(region_, interior, exterior);
    if (frequency_hz != frequency_hz_ || medium.relative_permittivity != medium_.relative_permittivity ||
        medium.relative_permeability != medium_.relative_permeability ||
        vacuum_permittivity != vacuum_permittivity_ || vacuum_permeability != vacuum_permeability_)
        throw std::invalid_argument(
            "prepared excitation does not match the solver frequency, source medium, or vacuum constants");
}

}  // namespace excitation
}  // namespace sie_mom
