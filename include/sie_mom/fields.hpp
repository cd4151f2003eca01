// Material and field-sample types of the C++ interface: IsotropicMedium,
// ComplexCartesianVector and ElectromagneticFieldSample as the reference's
// include/sie_mom/fields.hpp:13-45 declares them.
#pragma once

#include <array>
#include <complex>
#include <cstddef>

#include "sie_mom/core.hpp"

namespace sie_mom {

struct IsotropicMedium {
    real_t relative_permittivity = real_t(1);
    real_t relative_permeability = real_t(1);
};

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

}  // namespace sie_mom
