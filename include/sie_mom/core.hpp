// Core types of the C++ interface over libsie_amd (MI355X). The global host
// types are those the reference's callers use unqualified (real_t, vec3_t,
// int2, int3, complex__float and make_vec3 / make_int2 / make_int3 /
// make_complex: the host definitions of include/sie_mom/cuda/platform.cuh in
// its FP64 build), so a caller of the reference compiles unchanged. The
// vector types are the HIP runtime's own (hip/hip_vector_types.h, usable from
// any host C++ compiler with -D__HIP_PLATFORM_AMD__ and ROCm's include path),
// so these headers and the HIP runtime's headers can meet in one translation
// unit. The solver interfaces take std::complex<real_t>, as the reference's do.
#pragma once

#include <hip/hip_vector_types.h>

#include <complex>
#include <stdexcept>
#include <string>

#include "sie_mom/sie_capi.h"

using real_t = double;
using vec3_t = double3;
// complex value with the .x (real) / .y (imaginary) members of the
// reference's host complex type
struct complex_t {
    real_t x, y;
};
using complex__float = complex_t;

inline vec3_t make_vec3(real_t x, real_t y, real_t z) { return make_double3(x, y, z); }
inline complex__float make_complex(real_t re, real_t im) { return complex__float{re, im}; }

namespace sie_mom {

// include/sie_mom/cuda/gstc_tensor_utils.cuh:30-33 of the reference
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
