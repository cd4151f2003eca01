// Shared helpers of the C++ interface layer (device buffers, timing, the
// float-rounded kernel parameters). The layer uses only the C ABI of
// libsie_amd plus the HIP runtime for device memory.
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <complex>
#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "sie_mom/core.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"

// the C ABI takes vertices and index triples as packed arrays
static_assert(sizeof(vec3_t) == 3 * sizeof(double) && sizeof(int3) == 3 * sizeof(int) &&
                  sizeof(int2) == 2 * sizeof(int),
              "host vector types must be packed");

namespace sie_mom {
namespace host {

inline void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
class DeviceArray {
public:
    explicit DeviceArray(std::size_t n = 0) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(std::size_t n) {
        release();
        if (n) hip_ok(hipMalloc(&p_, n * sizeof(T)), "device allocation");
        n_ = n;
    }
    void release() {
        if (p_) {
            (void)hipDeviceSynchronize();
            (void)hipFree(p_);
        }
        p_ = nullptr;
        n_ = 0;
    }
    T* data() const { return p_; }
    std::size_t size() const { return n_; }
    void upload(const T* h, std::size_t n) {
        if (n) hip_ok(hipMemcpy(p_, h, n * sizeof(T), hipMemcpyHostToDevice), "host to device copy");
    }
    void download(T* h, std::size_t n, std::size_t offset = 0) const {
        if (n) hip_ok(hipMemcpy(h, p_ + offset, n * sizeof(T), hipMemcpyDeviceToHost), "device to host copy");
    }

private:
    T* p_ = nullptr;
    std::size_t n_ = 0;
};

// Owns a sie_mesh_t uploaded from MeshData.
class MeshHandle {
public:
    explicit MeshHandle(const mesh::MeshData& m);
    ~MeshHandle() {
        if (h_) sie_mesh_destroy(h_);
    }
    MeshHandle(const MeshHandle&) = delete;
    MeshHandle& operator=(const MeshHandle&) = delete;
    sie_mesh_t get() const { return h_; }

private:
    sie_mesh_t h_ = nullptr;
};

using Clock = std::chrono::steady_clock;
inline double seconds_since(Clock::time_point t0) {
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

// Norm of a host vector accumulated in long double (the reference's vector_norm,
// src/sie_mom/solvers/gstc_sie.cu:61-69).
inline double host_norm(const std::vector<std::complex<real_t>>& v) {
    long double s = 0.0L;
    for (const auto& z : v) {
        const long double a = z.real(), b = z.imag();
        s += a * a + b * b;
    }
    return std::sqrt((double)s);
}

inline void require_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw std::runtime_error("no GPU device is available for the SIE solve");
}

}  // namespace host
}  // namespace sie_mom
