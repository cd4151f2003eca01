// The reference's on-disk exchange formats, byte compatible:
//  - tensors.bin: 32 little-endian complex128 per face (the record layout of
//    solve_gstc_dense), or complex64 detected by file size
//    (src/sie_mom/gstc/gstc_utils.hpp:165-210; writer tensor_io.py:118-120);
//  - currents.bin: complex128 [J1, M1, J2, M2] in physical units
//    (applications/anisotropic_gstc_sphere/anisotropic_sphere_app.cpp:519-540).
#pragma once

#include <string>
#include <vector>

#include "sie_mom/core.hpp"

namespace sie_mom {
namespace io {

// Reads face_count * 32 values (or 32 when uniform); std::runtime_error if the
// file is missing or its size matches neither complex128 nor complex64 records.
std::vector<complex_t> read_tensors(const std::string& path, std::size_t face_count, bool uniform = false);
void write_tensors(const std::string& path, const std::vector<complex_t>& tensors);
void write_currents(const std::string& path, const std::vector<complex_t>& currents);
std::vector<complex_t> read_currents(const std::string& path);

}  // namespace io
}  // namespace sie_mom
