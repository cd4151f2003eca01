// The reference's on-disk exchange formats, byte compatible, at the path of
// its tensor loader (include/sie_mom/gstc_utils.hpp):
//  - tensors.bin: 32 little-endian complex128 per face (the record layout of
//    solve_gstc_dense), or complex64 detected by file size
//    (src/sie_mom/gstc/gstc_utils.hpp:165-210; writer tensor_io.py:118-120);
//  - currents.bin: complex128 [J1, M1, J2, M2] in physical units
//    (applications/anisotropic_gstc_sphere/anisotropic_sphere_app.cpp:519-540).
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "sie_mom/core.hpp"

namespace sie_mom {
namespace io {

// Reads face_count * 32 values (or 32 when uniform); std::runtime_error if the
// file is missing or its size matches neither complex128 nor complex64 records.
std::vector<std::complex<real_t>> read_tensors(const std::string& path, std::size_t face_count, bool uniform = false);
void write_tensors(const std::string& path, const std::vector<std::complex<real_t>>& tensors);
void write_currents(const std::string& path, const std::vector<std::complex<real_t>>& currents);
std::vector<std::complex<real_t>> read_currents(const std::string& path);

}  // namespace io

namespace gstc {
// The reference's loader (gstc_utils.hpp:165-210): 32 complex values per face
// from a complex128 or complex64 file, as the host complex type with .x / .y
// members; the same size check and messages as io::read_tensors.
std::vector<complex__float> load_face_tensors_from_file(const std::string& filename, std::size_t num_faces);
}  // namespace gstc
}  // namespace sie_mom
