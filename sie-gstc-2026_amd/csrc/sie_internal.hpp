// Internal declarations shared by the HIP translation units of libsie_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/sie_mom/sie_capi.h"

namespace sie {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(SIE_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define SIE_HIP(x) ::sie::hip_check((x), #x)

inline void require(bool ok, const std::string& msg) {
    if (!ok) throw Error(SIE_ERR_INVALID, msg);
}

void set_last_error(const std::string& m);

void lu_solve_f32_f64(int op, int n, const float* LU, long long lda, const int* ipiv, double* b, hipStream_t s);
// Throws SIE_ERR_DEVICE when a device-side wait gave up since the last check
// (synchronises the stream).
void check_device_waits(hipStream_t s);
int gmres_run(int n, int m, int maxnit, double tol, sie_matvec_fn matvec, void* user, const double* b, double* x,
              double* out_res, int* out_iters, double* history, int history_cap, int* history_len,
              hipStream_t s);

// ------------------------------------------------------------ device memory
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count) SIE_HIP(hipMalloc(&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        // kernels still queued may reference the buffer: drain the device first
        if (p) {
            (void)hipDeviceSynchronize();
            (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
    }
    void upload(const T* h, size_t count, hipStream_t s) {
        if (count) SIE_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

// ------------------------------------------------------------------ mesh
// Per-face geometry, precomputed on the host in the same arithmetic the
// reference uses on the device for every face pair (mesh_geometry.cuh:7-80,
// sie_kernels_assembly.cuh:193-197), so the near/far classification and the
// signed RWG lengths are bit-identical to it.
struct FaceSoA {
    // centroid (v0+v1+v2)/3
    std::vector<double> cx, cy, cz;
    // vertices relative to the centroid: o_a = v_a - c  (a = 0,1,2)
    std::vector<double> o[3][3];  // [vertex][xyz]
    // signed RWG length of local edge j (vertex j -> j+1, free vertex j+2)
    std::vector<double> len[3];
    std::vector<int> edge[3];
    std::vector<int> vid[3];
    std::vector<double> area;
};

struct Mesh {
    int nv = 0, nf = 0, ne = 0;
    std::vector<double> V;    // nv x 3
    std::vector<int> F;       // nf x 3
    std::vector<int> FE;      // nf x 3
    FaceSoA h;
    // edge -> its two faces and local indices (plus, minus as first/second encounter)
    std::vector<int> ef[2], el[2];
    // device copies
    DevBuf<double> d_c;       // 3 x nf  (SoA)
    DevBuf<double> d_o;       // 9 x nf  (SoA: vertex a, component x -> row 3a+x)
    DevBuf<double> d_len;     // 3 x nf
    DevBuf<int> d_edge;       // 3 x nf
    DevBuf<int> d_vid;        // 3 x nf
    DevBuf<double> d_V;       // nv x 3
    DevBuf<int> d_F;          // nf x 3
    DevBuf<int> d_ef;         // 2 x ne
    DevBuf<int> d_el;         // 2 x ne
};

void build_mesh(Mesh& m, const double* V, int nv, const int* F, int nf, const int* FE, int ne,
                hipStream_t s);

// Quadrature table in device memory: rule r -> (xi, eta, w) as float32.
struct QuadDev {
    int rule = 0, n = 0;
    DevBuf<float> xy, w;
};
void quad_upload(QuadDev& q, int rule, hipStream_t s);
void quad_host(int rule, std::vector<float>& xy, std::vector<float>& w);

}  // namespace sie
