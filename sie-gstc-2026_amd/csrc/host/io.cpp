// tensors.bin / currents.bin (the reference's include/sie_mom/gstc_utils.hpp:165-210;
// applications/anisotropic_gstc_sphere/anisotropic_sphere_app.cpp:519-540).
#include <fstream>

#include "host_common.hpp"
#include "sie_mom/gstc_utils.hpp"

namespace sie_mom {
namespace io {

std::vector<std::complex<real_t>> read_tensors(const std::string& path, std::size_t face_count, bool uniform) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Could not open tensor file: " + path);
    in.seekg(0, std::ios::end);
    const std::size_t size = (std::size_t)in.tellg();
    in.seekg(0, std::ios::beg);
    const std::size_t values = (uniform ? 1 : face_count) * 32 * 2;
    const std::size_t b32 = values * sizeof(float), b64 = values * sizeof(double);
    if (size != b32 && size != b64)
        throw std::runtime_error("Tensor file size mismatch. Expected " + std::to_string(b32) + " or " +
                                 std::to_string(b64) + " bytes, got " + std::to_string(size));
    std::vector<std::complex<real_t>> t(values / 2);
    if (size == b32) {
        std::vector<float> raw(values);
        in.read(reinterpret_cast<char*>(raw.data()), b32);
        if (!in) throw std::runtime_error("Error reading tensor data");
        for (std::size_t i = 0; i < t.size(); ++i) t[i] = std::complex<real_t>(raw[2 * i], raw[2 * i + 1]);
    } else {
        in.read(reinterpret_cast<char*>(t.data()), b64);
        if (!in) throw std::runtime_error("Error reading tensor data");
    }
    return t;
}

void write_tensors(const std::string& path, const std::vector<std::complex<real_t>>& tensors) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Could not open tensor file for writing: " + path);
    out.write(reinterpret_cast<const char*>(tensors.data()), tensors.size() * sizeof(std::complex<real_t>));
    if (!out) throw std::runtime_error("Error writing tensor data");
}

void write_currents(const std::string& path, const std::vector<std::complex<real_t>>& currents) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Could not open currents file for writing: " + path);
    out.write(reinterpret_cast<const char*>(currents.data()), currents.size() * sizeof(std::complex<real_t>));
    if (!out) throw std::runtime_error("Error writing currents");
}

std::vector<std::complex<real_t>> read_currents(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) throw std::runtime_error("Could not open currents file: " + path);
    in.seekg(0, std::ios::end);
    const std::size_t size = (std::size_t)in.tellg();
    in.seekg(0, std::ios::beg);
    if (size % sizeof(std::complex<real_t>) != 0) throw std::runtime_error("currents file size is not a multiple of 16 bytes");
    std::vector<std::complex<real_t>> c(size / sizeof(std::complex<real_t>));
    in.read(reinterpret_cast<char*>(c.data()), size);
    if (!in) throw std::runtime_error("Error reading currents");
    return c;
}

}  // namespace io
namespace gstc {

std::vector<complex__float> load_face_tensors_from_file(const std::string& filename, std::size_t num_faces) {
    const std::vector<std::complex<real_t>> t = io::read_tensors(filename, num_faces, false);
    std::vector<complex__float> out(t.size());
    for (std::size_t i = 0; i < t.size(); ++i) out[i] = make_complex(t[i].real(), t[i].imag());
    return out;
}

}  // namespace gstc
}  // namespace sie_mom
