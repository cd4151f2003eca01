// Core types of the C++ interface over libsie_amd (MI355X). They mirror the
// reference's host-side types (real_t, vec3_t, int2/int3 of
// include/sie_mom/cuda/platform.cuh; IsotropicMedium, include/sie_mom/fields.hpp:13-16;
// GstcBasisType, include/sie_mom/cuda/gstc_tensor_utils.cuh:30-33) without any
// device-language dependency: a caller compiles against them with a plain C++17
// compiler and links libsie_amd.so.
#pragma once

#include <complex>
#include <stdexcept>
#include <string>

#include "sie_mom/sie_capi.h"

namespace sie_mom {

using real_t = double;
using complex_t = std::complex<double>;

struct vec3_t {
    double x = 0, y = 0, z = 0;
};
struct int2 {
    int x = 0, y = 0;
};
struct int3 {
    int x = 0, y = 0, z = 0;
};
inline vec3_t make_vec3(double x, double y, double z) { return vec3_t{x, y, z}; }

struct IsotropicMedium {
    real_t relative_permittivity = real_t(1);
    real_t relative_permeability = real_t(1);
};

enum class GstcBasisType : int { face = 0, spherical = 1 };

namespace solvers {
// Arithmetic of the dense solve: plain FP64 LU, or FP32 LU on the matrix cores
// with FP64 refinement to refinement_tolerance (an extension of the reference's
// option structs; fp64 is the reference's behaviour).
enum class SolvePrecision { fp64, mixed };
}  // namespace solvers

// CODATA vacuum constants used by the reference's option defaults.
constexpr real_t kCodata2022Eps0 = 8.8541878188e-12;
constexpr real_t kCodata2022Mu0 = 1.25663706127e-6;
constexpr real_t kCodata2018Eps0 = 8.8541878128e-12;
constexpr real_t kCodata2018Mu0 = 1.25663706212e-6;

namespace detail {
// Maps a libsie_amd status to the reference's exception classes:
// invalid_argument, overflow_error, runtime_error (singular, non-finite, device).
inline void check(int code) {
    if (code == SIE_OK) return;
    const std::string msg = sie_last_error();
    if (code == SIE_ERR_INVALID) throw std::invalid_argument(msg);
    if (code == SIE_ERR_OVERFLOW) throw std::overflow_error(msg);
    throw std::runtime_error(msg);
}
}  // namespace detail

}  // namespace sie_mom
