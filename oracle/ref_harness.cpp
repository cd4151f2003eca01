// Test-infrastructure harness: exposes the reference's own host-compilable
// FP64 numerics (the __host__ __device__ headers of sie-mom, built with
// -DUSE_DOUBLE_PRECISION) through a flat C ABI, so that the oracle restatement
// in sie_oracle.cpp can be pinned against the reference itself.
//
// This file contains no reference source. It includes the reference headers
// where they lie (the recipe in oracle/Makefile points -I at the reference
// tree) and calls their functions. The only logic here is the per-face-pair
// loop that the reference runs inside its __global__ kernels (which cannot be
// compiled for the host); those loops call the reference's own per-pair
// helpers and epilogue factors.
//
// Only tests/ and the bench's CPU baseline load the library built from this.

#include "sie_mom/cuda/dunavant_quadrature.cuh"
#include "sie_mom/cuda/gstc_tensor_utils.cuh"
#include "sie_mom/cuda/sie_scaling.cuh"
#include "sie_mom/quadrature/triangle_dunavant_rule.hpp"
#include "sie_mom/mesh/spherical_surface.hpp"
#include "sie_mom/mesh/obj_mesh_loader.hpp"

#include <cstring>
#include <exception>
#include <string>
#include <vector>

namespace {

inline vec3_t v3(const double* p) { return make_vec3(p[0], p[1], p[2]); }

void load_tri(const double* t, vec3_t tri[3]) {
    for (int a = 0; a < 3; ++a) tri[a] = v3(t + 3 * a);
}

void put_block(const complex__float m[3][3], double* out) {
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            out[2 * (3 * j + i) + 0] = m[j][i].x;
            out[2 * (3 * j + i) + 1] = m[j][i].y;
        }
}

struct Quad {
    std::vector<float> pts, w;
    int n = 0;
};

Quad quad(int rule) {
    auto q = sie_mom::quadrature::make_quadrature(rule);
    return {q.points, q.weights, q.order};
}

}  // namespace

extern "C" {

int ref_dunavant_order_num(int rule) { return dunavant_order_num(rule); }

void ref_make_quadrature(int rule, float* pts, float* w) {
    Quad q = quad(rule);
    std::memcpy(pts, q.pts.data(), sizeof(float) * 2 * q.n);
    std::memcpy(w, q.w.data(), sizeof(float) * q.n);
}

// Icosphere of SphericalSurface(radius, level): counts, then arrays.
void ref_icosphere_counts(int level, int* nv, int* nf, int* ne) {
    sie_mom::mesh::SphericalSurface s(1.0f, level);
    const auto& m = s.getMeshData();
    *nv = (int)m.vertices.size();
    *nf = (int)m.faces.size();
    *ne = (int)m.edges.size();
}

void ref_icosphere(float radius, int level, double* V, int* F, int* E, int* FE) {
    sie_mom::mesh::SphericalSurface s(radius, level);
    const auto& m = s.getMeshData();
    for (size_t i = 0; i < m.vertices.size(); ++i) {
        V[3 * i] = m.vertices[i].x; V[3 * i + 1] = m.vertices[i].y; V[3 * i + 2] = m.vertices[i].z;
    }
    for (size_t i = 0; i < m.faces.size(); ++i) {
        F[3 * i] = m.faces[i].x; F[3 * i + 1] = m.faces[i].y; F[3 * i + 2] = m.faces[i].z;
        FE[3 * i] = m.faces_edges[i].x; FE[3 * i + 1] = m.faces_edges[i].y; FE[3 * i + 2] = m.faces_edges[i].z;
    }
    for (size_t i = 0; i < m.edges.size(); ++i) {
        E[2 * i] = m.edges[i].x; E[2 * i + 1] = m.edges[i].y;
    }
}

// calculateDoubleQuadrature (TL/TK at k1, k2). out: 4 blocks x 9 complex
// in the order TL(k1), TL(k2), TK(k1), TK(k2).
void ref_pair_regular(const double* tt, const int* ti, const double* st, const int* si,
                      int rule, double k1, double k2, int is_singular, double* out) {
    Quad q = quad(rule);
    vec3_t a[3], b[3];
    load_tri(tt, a);
    load_tri(st, b);
    complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
    calculateDoubleQuadrature(a, ti, b, si, q.pts.data(), q.w.data(), q.n, k1, k2,
                              is_singular != 0, tl1, tl2, tk1, tk2);
    put_block(tl1, out); put_block(tl2, out + 18); put_block(tk1, out + 36); put_block(tk2, out + 54);
}

// calculateSingularPartQuadrature (Wilton inner integrals, outer Dunavant).
void ref_pair_singular(const double* tt, const int* ti, const double* st, const int* si,
                       int rule, double k1, double k2, double* out) {
    Quad q = quad(rule);
    vec3_t a[3], b[3];
    load_tri(tt, a);
    load_tri(st, b);
    complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
    calculateSingularPartQuadrature(a, ti, b, si, q.pts.data(), q.w.data(), q.n, k1, k2,
                                    tl1, tl2, tk1, tk2);
    put_block(tl1, out); put_block(tl2, out + 18); put_block(tk1, out + 36); put_block(tk2, out + 54);
}

// calculateAllWiltonPotentials: phi0, phi1_base[3], phi2, phi3_base[3], r_pi[3], d
void ref_wilton(const double* r, const double* tri, double* out) {
    vec3_t t[3];
    load_tri(tri, t);
    WiltonPotentials p = calculateAllWiltonPotentials(v3(r), t);
    out[0] = p.phi_0;
    out[1] = p.phi_1_base.x; out[2] = p.phi_1_base.y; out[3] = p.phi_1_base.z;
    out[4] = p.phi_2;
    out[5] = p.phi_3_base.x; out[6] = p.phi_3_base.y; out[7] = p.phi_3_base.z;
    out[8] = p.r_pi.x; out[9] = p.r_pi.y; out[10] = p.r_pi.z;
    out[11] = p.d;
}

// Batched 8-tensor mass quadrature: T and out are 8 x 3 x 3 complex.
void ref_mass_batch(const double* tri, const int* idx, int rule, const double* T, double* out) {
    Quad q = quad(rule);
    vec3_t t[3];
    load_tri(tri, t);
    complex__float tensors[8][3][3], res[8][3][3];
    for (int k = 0; k < 8; ++k)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c)
                tensors[k][r][c] = make_complex(T[2 * (9 * k + 3 * r + c)], T[2 * (9 * k + 3 * r + c) + 1]);
    calculateMassMatrixQuadrature(t, idx, q.pts.data(), q.w.data(), q.n, tensors, res);
    for (int k = 0; k < 8; ++k) put_block(res[k], out + 18 * k);
}

void ref_excitation_planewave(const double* tri, const int* idx, int rule, const double* E0,
                              const double* khat, double k, double eta, double* VE, double* VH) {
    Quad q = quad(rule);
    vec3_t t[3];
    load_tri(tri, t);
    sie_mom::cuda::PlaneWave pw{v3(E0), v3(khat), k, eta};
    complex__float ve[3], vh[3];
    calculateExcitationQuadrature(t, idx, q.pts.data(), q.w.data(), q.n, pw, ve, vh);
    for (int i = 0; i < 3; ++i) {
        VE[2 * i] = ve[i].x; VE[2 * i + 1] = ve[i].y;
        VH[2 * i] = vh[i].x; VH[2 * i + 1] = vh[i].y;
    }
}

void ref_face_basis(const double* tri, int basis, double* u, double* v) {
    vec3_t t[3], uu, vv;
    load_tri(tri, t);
    get_local_basis_vectors(t[0], t[1], t[2], static_cast<sie_mom::GstcBasisType>(basis), uu, vv);
    u[0] = uu.x; u[1] = uu.y; u[2] = uu.z;
    v[0] = vv.x; v[1] = vv.y; v[2] = vv.z;
}

// ---------------------------------------------------------------------------
// Host restatements of the reference __global__ kernel loops, built from the
// reference's own per-pair helpers. Z is column-major, dim x dim complex.
// ---------------------------------------------------------------------------

static void load_face(const double* V, const int* F, int f, vec3_t tri[3], int idx[3]) {
    for (int a = 0; a < 3; ++a) {
        idx[a] = F[3 * f + a];
        tri[a] = v3(V + 3 * idx[a]);
    }
}

// Shared Calderon quadrature of one face pair (sie_kernels_assembly.cuh:187-220 pattern).
static void pair_blocks(const vec3_t tt[3], const int ti[3], const vec3_t st[3], const int si[3],
                        const Quad& q, float k1f, float k2f, float thr,
                        complex__float tl1[3][3], complex__float tl2[3][3],
                        complex__float tk1[3][3], complex__float tk2[3][3]) {
    vec3_t ct = (tt[0] + tt[1] + tt[2]) / 3.f;
    vec3_t cs = (st[0] + st[1] + st[2]) / 3.f;
    float dist = length(ct - cs);
    if (dist < thr) {
        calculateDoubleQuadrature(tt, ti, st, si, q.pts.data(), q.w.data(), q.n, k1f, k2f, true,
                                  tl1, tl2, tk1, tk2);
        complex__float s1[3][3], s2[3][3], s3[3][3], s4[3][3];
        calculateSingularPartQuadrature(tt, ti, st, si, q.pts.data(), q.w.data(), q.n, k1f, k2f,
                                        s1, s2, s3, s4);
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                complexAddInPlace(tl1[j][i], s1[j][i]);
                complexAddInPlace(tl2[j][i], s2[j][i]);
                complexAddInPlace(tk1[j][i], s3[j][i]);
                complexAddInPlace(tk2[j][i], s4[j][i]);
            }
    } else {
        calculateDoubleQuadrature(tt, ti, st, si, q.pts.data(), q.w.data(), q.n, k1f, k2f, false,
                                  tl1, tl2, tk1, tk2);
    }
}

static inline void addz(double* Z, size_t dim, size_t row, size_t col, complex__float v) {
    Z[2 * (col * dim + row)] += v.x;
    Z[2 * (col * dim + row) + 1] += v.y;
}

// assemble_Z_calderon_eta0_scaled_kernel + assemble_Z_mass_eta0_scaled_kernel.
// All float parameters are the float-cast values the reference passes.
void ref_assemble_gstc(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                       float k1, float k2, float mu_r1, float mu_r2, float eps_r1, float eps_r2,
                       float mu0, float eps0, float omega, float thr,
                       const double* tensors, int uniform, int basis, int with_mass, double* Z) {
    Quad q = quad(rule);
    const size_t dim = 4 * (size_t)ne;
    std::memset(Z, 0, sizeof(double) * 2 * dim * dim);
    const real_t eta0 = sqrt(mu0 / eps0);
    const auto f1 = sie_mom::cuda::make_eta0_scaled_calderon_factors(omega, mu_r1 * mu0, eps_r1 * eps0, eta0);
    const auto f2 = sie_mom::cuda::make_eta0_scaled_calderon_factors(omega, mu_r2 * mu0, eps_r2 * eps0, eta0);
    for (int m = 0; m < nf; ++m) {
        vec3_t tt[3]; int ti[3];
        load_face(V, F, m, tt, ti);
        for (int n = 0; n < nf; ++n) {
            vec3_t st[3]; int si[3];
            load_face(V, F, n, st, si);
            complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
            pair_blocks(tt, ti, st, si, q, k1, k2, thr, tl1, tl2, tk1, tk2);
            for (int j = 0; j < 3; ++j) {
                size_t te = FE[3 * m + j];
                for (int i = 0; i < 3; ++i) {
                    size_t se = FE[3 * n + i];
                    auto b1 = sie_mom::cuda::make_eta0_scaled_calderon_block(tl1[j][i], tk1[j][i], f1);
                    auto b2 = sie_mom::cuda::make_eta0_scaled_calderon_block(tl2[j][i], tk2[j][i], f2);
                    size_t E = ne;
                    addz(Z, dim, te, se, b1.jj);
                    addz(Z, dim, te, se + E, b1.jm);
                    addz(Z, dim, te + E, se, b1.mj);
                    addz(Z, dim, te + E, se + E, b1.mm);
                    addz(Z, dim, te + 2 * E, se + 2 * E, b2.jj);
                    addz(Z, dim, te + 2 * E, se + 3 * E, b2.jm);
                    addz(Z, dim, te + 3 * E, se + 2 * E, b2.mj);
                    addz(Z, dim, te + 3 * E, se + 3 * E, b2.mm);
                }
            }
        }
    }
    if (!with_mass) return;
    const float eta0f = sqrtf(mu0 / eps0);
    const float eta0sq = eta0f * eta0f;
    for (int m = 0; m < nf; ++m) {
        vec3_t tri[3]; int idx[3];
        load_face(V, F, m, tri, idx);
        vec3_t u, v;
        get_local_basis_vectors(tri[0], tri[1], tri[2], static_cast<sie_mom::GstcBasisType>(basis), u, v);
        const double* tp = uniform ? tensors : tensors + 2 * 32 * (size_t)m;
        auto c = [&](int k) { return make_complex(tp[2 * k], tp[2 * k + 1]); };
        auto rot = [&](int base, int rb, int cb, complex__float T[3][3]) {
            int r0 = rb * 2, c0 = cb * 2;
            rotate_tensor_2x2_to_3x3(c(base + r0 * 4 + c0), c(base + r0 * 4 + c0 + 1),
                                     c(base + (r0 + 1) * 4 + c0), c(base + (r0 + 1) * 4 + c0 + 1), u, v, T);
        };
        complex__float Y[4][3][3], Zt[4][3][3];
        rot(0, 0, 0, Y[0]); rot(0, 0, 1, Y[1]); rot(0, 1, 0, Y[2]); rot(0, 1, 1, Y[3]);
        rot(16, 0, 0, Zt[0]); rot(16, 0, 1, Zt[1]); rot(16, 1, 0, Zt[2]); rot(16, 1, 1, Zt[3]);
        vec3_t nv = normalize(cross(tri[1] - tri[0], tri[2] - tri[0]));
        complex__float batch[8][3][3], res[8][3][3];
        for (int k = 0; k < 4; ++k) apply_perp_transform(Y[k], nv, batch[k]);
        for (int k = 0; k < 4; ++k) copy_tensor_3x3(Zt[k], batch[4 + k]);
        calculateMassMatrixQuadrature(tri, idx, q.pts.data(), q.w.data(), q.n, batch, res);
        size_t E = ne;
        for (int j = 0; j < 3; ++j) {
            size_t te = FE[3 * m + j];
            for (int i = 0; i < 3; ++i) {
                size_t se = FE[3 * m + i];
                const float f = 0.5f, fq = 0.5f * 0.25f;
                auto Zee = res[4][j][i], Zem = res[5][j][i], Zme = res[6][j][i], Zmm = res[7][j][i];
                auto Yee = res[0][j][i], Yem = res[1][j][i], Yme = res[2][j][i], Ymm = res[3][j][i];
                auto add = [&](size_t ro, size_t co, complex__float val, float s) {
                    addz(Z, dim, te + ro, se + co, complexScale(val, s));
                };
                add(0, 0, complexAdd(complexScale(Zee, f), complexScale(Ymm, fq)), 1.f);
                add(0, E, complexSub(complexScale(Zem, f), complexScale(Yme, fq)), eta0f);
                add(0, 2 * E, complexSub(complexScale(Zee, f), complexScale(Ymm, fq)), 1.f);
                add(0, 3 * E, complexAdd(complexScale(Zem, f), complexScale(Yme, fq)), eta0f);
                add(E, 0, complexSub(complexScale(Zme, f), complexScale(Yem, fq)), eta0f);
                add(E, E, complexAdd(complexScale(Zmm, f), complexScale(Yee, fq)), eta0sq);
                add(E, 2 * E, complexAdd(complexScale(Zme, f), complexScale(Yem, fq)), eta0f);
                add(E, 3 * E, complexSub(complexScale(Zmm, f), complexScale(Yee, fq)), eta0sq);
                add(2 * E, 0, complexSub(complexScale(Zee, f), complexScale(Ymm, fq)), 1.f);
                add(2 * E, E, complexAdd(complexScale(Zem, f), complexScale(Yme, fq)), eta0f);
                add(2 * E, 2 * E, complexAdd(complexScale(Zee, f), complexScale(Ymm, fq)), 1.f);
                add(2 * E, 3 * E, complexSub(complexScale(Zem, f), complexScale(Yme, fq)), eta0f);
                add(3 * E, 0, complexAdd(complexScale(Zme, f), complexScale(Yem, fq)), eta0f);
                add(3 * E, E, complexSub(complexScale(Zmm, f), complexScale(Yee, fq)), eta0sq);
                add(3 * E, 2 * E, complexSub(complexScale(Zme, f), complexScale(Yem, fq)), eta0f);
                add(3 * E, 3 * E, complexAdd(complexScale(Zmm, f), complexScale(Yee, fq)), eta0sq);
            }
        }
    }
}

// assemble_V_vector_eta0_scaled_kernel with a plane wave (region 1 or 2).
void ref_rhs_gstc_planewave(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                            const double* E0, const double* khat, double k, double eta, int region,
                            float mu0, float eps0, double* rhs) {
    Quad q = quad(rule);
    sie_mom::cuda::PlaneWave pw{v3(E0), v3(khat), k, eta};
    real_t eta0 = sqrt(static_cast<real_t>(mu0) / static_cast<real_t>(eps0));
    for (int m = 0; m < nf; ++m) {
        vec3_t tri[3]; int idx[3];
        load_face(V, F, m, tri, idx);
        complex__float ve[3], vh[3];
        calculateExcitationQuadrature(tri, idx, q.pts.data(), q.w.data(), q.n, pw, ve, vh);
        for (int i = 0; i < 3; ++i) {
            size_t e = FE[3 * m + i];
            size_t o = region == 1 ? 0 : 2 * (size_t)ne;
            complex__float h = complexScale(vh[i], eta0);
            rhs[2 * (e + o)] += ve[i].x; rhs[2 * (e + o) + 1] += ve[i].y;
            rhs[2 * (e + o + ne)] += h.x; rhs[2 * (e + o + ne) + 1] += h.y;
        }
    }
}

// assemble_dielectric_rhs_kernel with the T-PMCHWT scaling of assemble_rhs
// (src/sie_mom/solvers/dielectric/dielectric_solver.cu:400-430, 515-531):
// sign +1 for an exterior source, -1 for an interior one, and sign * eta0 on
// the magnetic rows with eta0 = sqrt(vac_mu / vac_eps) in double.
void ref_rhs_pmchwt_planewave(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                              const double* E0, const double* khat, double k, double eta, int region,
                              double vac_mu, double vac_eps, double* rhs) {
    Quad q = quad(rule);
    sie_mom::cuda::PlaneWave pw{v3(E0), v3(khat), k, eta};
    const real_t eta0 = std::sqrt(static_cast<real_t>(vac_mu) / static_cast<real_t>(vac_eps));
    const real_t sign = region == 2 ? real_t(1) : real_t(-1);
    for (int m = 0; m < nf; ++m) {
        vec3_t tri[3]; int idx[3];
        load_face(V, F, m, tri, idx);
        complex__float ve[3], vh[3];
        calculateExcitationQuadrature(tri, idx, q.pts.data(), q.w.data(), q.n, pw, ve, vh);
        for (int i = 0; i < 3; ++i) {
            size_t e = FE[3 * m + i];
            complex__float a = complexScale(ve[i], sign), h = complexScale(vh[i], sign * eta0);
            rhs[2 * e] += a.x; rhs[2 * e + 1] += a.y;
            rhs[2 * (e + ne)] += h.x; rhs[2 * (e + ne) + 1] += h.y;
        }
    }
}

// assemble_pec_efie_operator_kernel (k2 passed twice, as pec_solver.cu:352-359 does).
void ref_assemble_pec(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                      float k2, float mu_r2, float mu0, float omega, float thr, double* Z) {
    Quad q = quad(rule);
    const size_t dim = ne;
    std::memset(Z, 0, sizeof(double) * 2 * dim * dim);
    const complex__float c = make_complex(0.f, omega * mu_r2 * mu0);
    for (int m = 0; m < nf; ++m) {
        vec3_t tt[3]; int ti[3];
        load_face(V, F, m, tt, ti);
        for (int n = 0; n < nf; ++n) {
            vec3_t st[3]; int si[3];
            load_face(V, F, n, st, si);
            complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
            pair_blocks(tt, ti, st, si, q, k2, k2, thr, tl1, tl2, tk1, tk2);
            for (int j = 0; j < 3; ++j)
                for (int i = 0; i < 3; ++i)
                    addz(Z, dim, FE[3 * m + j], FE[3 * n + i], c * tl2[j][i]);
        }
    }
}

// t_pmchwt_kernel (dielectric_kernels.cu:23-150).
void ref_assemble_pmchwt(const double* V, const int* F, const int* FE, int nf, int ne, int rule,
                         float k1, float k2, float mu_r1, float mu_r2, float eps_r1, float eps_r2,
                         float mu0, float eps0, float omega, float thr, double* Z) {
    Quad q = quad(rule);
    const size_t dim = 2 * (size_t)ne;
    std::memset(Z, 0, sizeof(double) * 2 * dim * dim);
    const complex__float jw = make_complex(0.f, omega);
    const float mu1 = mu_r1 * mu0, mu2 = mu_r2 * mu0, eps1 = eps_r1 * eps0, eps2 = eps_r2 * eps0;
    const float eta0 = sqrtf(mu0 / eps0);
    const complex__float cm1 = jw * mu1, cm2 = jw * mu2, ce1 = jw * eps1, ce2 = jw * eps2;
    for (int m = 0; m < nf; ++m) {
        vec3_t tt[3]; int ti[3];
        load_face(V, F, m, tt, ti);
        for (int n = 0; n < nf; ++n) {
            vec3_t st[3]; int si[3];
            load_face(V, F, n, st, si);
            complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
            pair_blocks(tt, ti, st, si, q, k1, k2, thr, tl1, tl2, tk1, tk2);
            for (int j = 0; j < 3; ++j) {
                size_t te = FE[3 * m + j];
                for (int i = 0; i < 3; ++i) {
                    size_t se = FE[3 * n + i];
                    complex__float tk = tk1[j][i] + tk2[j][i];
                    complex__float zjj = cm1 * tl1[j][i] + cm2 * tl2[j][i];
                    complex__float zmm = ce1 * tl1[j][i] + ce2 * tl2[j][i];
                    complex__float zmj = -1.f * tk;
                    complex__float zjm = 1.f * tk;
                    addz(Z, dim, te, se, zjj);
                    addz(Z, dim, te, se + ne, complexScale(zjm, eta0));
                    addz(Z, dim, te + ne, se, complexScale(zmj, eta0));
                    addz(Z, dim, te + ne, se + ne, complexScale(zmm, eta0 * eta0));
                }
            }
        }
    }
}

}  // extern "C"

extern "C" double ref_eta0_calderon(float mu0, float eps0) {
    // Same expression as the Calderon kernel's eta0 (double result of a float ratio).
    const real_t eta0 = sqrt(mu0 / eps0);
    return eta0;
}

// Reference host numerics on a sample of the Calderon assembly: test faces
// [m0, m1) against every source face, with the kernel's near/far branching
// (sie_kernels_assembly.cuh:187-220). OpenMP over test faces. Returns a
// checksum so the work cannot be elided.
extern "C" double ref_sample_pairs(const double* V, const int* F, int nf, int m0, int m1, int rule, float k1,
                                   float k2, float thr) {
    Quad q = quad(rule);
    double sum = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : sum)
    for (int m = m0; m < m1; ++m) {
        vec3_t tt[3]; int ti[3];
        load_face(V, F, m, tt, ti);
        for (int n = 0; n < nf; ++n) {
            vec3_t st[3]; int si[3];
            load_face(V, F, n, st, si);
            complex__float tl1[3][3], tl2[3][3], tk1[3][3], tk2[3][3];
            pair_blocks(tt, ti, st, si, q, k1, k2, thr, tl1, tl2, tk1, tk2);
            sum += tl1[0][0].x + tk2[2][2].y;
        }
    }
    return sum;
}

// The field loop of compute_scattered_fields_kernel (sie_kernels_utils.cuh:21-120)
// run sequentially over faces for each observation point, calling the
// reference's calculateScatterFieldQuadrature / calculateScatterFieldSingular;
// J, M, E, H interleaved complex doubles (E, H: 3 per point).
extern "C" void ref_fields(const double* V, const int* F, const int* FE, int nf, int rule, const double* J,
                           const double* M, const double* obs, int nobs, double omega, double eps, double mu,
                           double k, double* E, double* H) {
    Quad q = quad(rule);
    const complex__float cJ = make_complex(0.0, omega * mu), cM = make_complex(1.0, 0.0);
    const complex__float cJh = make_complex(-1.0, 0.0), cMh = make_complex(0.0, omega * eps);
#pragma omp parallel for schedule(dynamic, 4)
    for (int o = 0; o < nobs; ++o) {
        const vec3_t r = v3(obs + 3 * o);
        const complex__float z = make_complex(0.0, 0.0);
        cvec3_t e = make_cvec3(z, z, z), h = make_cvec3(z, z, z);
        for (int f = 0; f < nf; ++f) {
            vec3_t t[3]; int idx[3];
            load_face(V, F, f, t, idx);
            const vec3_t c = (t[0] + t[1] + t[2]) / 3.f;
            const float dist = length(r - c);
            const float lambda = 2.0f * static_cast<float>(M_PI) / k;
            const float thr = 0.1f * lambda;
            cvec3_t L[3], K[3];
            if (dist < thr) {
                calculateScatterFieldQuadrature(r, t, idx, q.pts.data(), q.w.data(), q.n, k, true, L, K);
                cvec3_t Ls[3], Ks[3];
                calculateScatterFieldSingular(r, t, idx, k, Ls, Ks);
                for (int i = 0; i < 3; ++i) { L[i] += Ls[i]; K[i] += Ks[i]; }
            } else {
                calculateScatterFieldQuadrature(r, t, idx, q.pts.data(), q.w.data(), q.n, k, false, L, K);
            }
            for (int i = 0; i < 3; ++i) {
                const int ed = FE[3 * f + i];
                const complex__float Je = make_complex(J[2 * ed], J[2 * ed + 1]);
                const complex__float Me = make_complex(M[2 * ed], M[2 * ed + 1]);
                e += complexMul(cJ, Je) * L[i];
                e += complexMul(cM, Me) * K[i];
                h += complexMul(cJh, Je) * K[i];
                h += complexMul(cMh, Me) * L[i];
            }
        }
        const complex__float* pe = &e.x;
        const complex__float* ph = &h.x;
        for (int x = 0; x < 3; ++x) {
            E[6 * o + 2 * x] = pe[x].x; E[6 * o + 2 * x + 1] = pe[x].y;
            H[6 * o + 2 * x] = ph[x].x; H[6 * o + 2 * x + 1] = ph[x].y;
        }
    }
}

// The reference's ObjMeshLoader (src/sie_mom/mesh/obj_mesh_loader.cpp:83-386):
// counts = {nv, nf, ne}; the arrays are filled when non-null. Returns 0, or 1
// with the loader's exception message in err (errlen bytes).
extern "C" int ref_load_obj(const char* path, int* counts, double* V, int* F, int* E, int* FE, char* err,
                            int errlen) {
    try {
        sie_mom::mesh::ObjMeshLoader loader(path);
        const auto& m = loader.getMeshData();
        counts[0] = (int)m.vertices.size();
        counts[1] = (int)m.faces.size();
        counts[2] = (int)m.edges.size();
        if (V && F && E && FE) {
            for (size_t i = 0; i < m.vertices.size(); ++i) {
                V[3 * i] = m.vertices[i].x; V[3 * i + 1] = m.vertices[i].y; V[3 * i + 2] = m.vertices[i].z;
            }
            for (size_t i = 0; i < m.faces.size(); ++i) {
                F[3 * i] = m.faces[i].x; F[3 * i + 1] = m.faces[i].y; F[3 * i + 2] = m.faces[i].z;
                FE[3 * i] = m.faces_edges[i].x; FE[3 * i + 1] = m.faces_edges[i].y; FE[3 * i + 2] = m.faces_edges[i].z;
            }
            for (size_t i = 0; i < m.edges.size(); ++i) { E[2 * i] = m.edges[i].x; E[2 * i + 1] = m.edges[i].y; }
        }
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(err, e.what(), errlen - 1);
        err[errlen - 1] = 0;
        return 1;
    }
}
