// Adjoint gradient of a far-field objective with respect to per-face GSTC
// surface parameters. Not present in the reference (its solves have no
// gradient); the forward pieces follow its conventions:
//  * surface map Lambda (4x4, [J_s; M_s] = Lambda [E_avg; H_avg]) in the
//    reciprocal lossless form [[jB, K], [-K^T, jX]] (src/sie_mom/gstc/
//    surface_parameters.py:280-306, omega special case omega_biisotropic.py:
//    196-219) and the record layout Lambda (row-major) then Lambda^{-1}
//    (tensor_io.py:17-130, inverted in impedance-balanced form);
//  * far field of the exterior currents by the Schelkunoff radiation vectors
//    F_theta = -jk0/4pi (eta0 N_theta + L_phi), F_phi = -jk0/4pi (eta0 N_phi -
//    L_theta), N = sum_e J_e int f_e e^{+jk0 r.r'} (applications/
//    anisotropic_gstc_sphere/plot_far_field.py:150-219);
//  * the mass-matrix derivative is the exact linearisation of the GSTC mass
//    kernel (sie_kernels_assembly.cuh:460-594) in the 32 record entries.
//
// With Zs xs = bs the eta0-scaled system and J = sum_d w_d |F_d|^2 / (2 eta0),
//   dJ/dtheta = -2 Re( lambda^H (dZs/dtheta) xs ),  Zs^H lambda = dJ/d conj(xs).
#include "sie_internal.hpp"
#include "prof.hpp"

#include <cmath>
#include <vector>

struct sie_mesh;  // defined in assembly.hip

namespace sie {
Mesh& mesh_of(sie_mesh* h);
const double* mesh_area(sie_mesh* h);
QuadDev& mesh_quad(sie_mesh* h, int rule, hipStream_t s);

namespace {

struct cplx { double re, im; };
__device__ __forceinline__ cplx cm(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }

// ---------------------------------------------------------- records
// In-place Gauss-Jordan inverse of a 4x4 complex matrix with partial pivoting.
__device__ bool inv4(cplx a[4][4], cplx out[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) out[i][j] = {i == j ? 1.0 : 0.0, 0.0};
    for (int c = 0; c < 4; ++c) {
        int p = c;
        double best = -1;
        for (int r = c; r < 4; ++r) {
            const double v = fabs(a[r][c].re) + fabs(a[r][c].im);
            if (v > best) { best = v; p = r; }
        }
        if (!(best > 0)) return false;
        if (p != c)
            for (int j = 0; j < 4; ++j) {
                cplx t = a[c][j]; a[c][j] = a[p][j]; a[p][j] = t;
                t = out[c][j]; out[c][j] = out[p][j]; out[p][j] = t;
            }
        const double d = a[c][c].re * a[c][c].re + a[c][c].im * a[c][c].im;
        const cplx iv = {a[c][c].re / d, -a[c][c].im / d};
        for (int j = 0; j < 4; ++j) { a[c][j] = cm(a[c][j], iv); out[c][j] = cm(out[c][j], iv); }
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const cplx f = a[r][c];
            for (int j = 0; j < 4; ++j) {
                const cplx u = cm(f, a[c][j]), v = cm(f, out[c][j]);
                a[r][j].re -= u.re; a[r][j].im -= u.im;
                out[r][j].re -= v.re; out[r][j].im -= v.im;
            }
        }
    }
    return true;
}

__device__ void build_lambda(int kind, const double* p, double bs, double xs, double ks, cplx L[4][4]) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) L[i][j] = {0, 0};
    if (kind == 3) {
        L[0][0] = L[1][1] = {0, bs * p[0]};
        L[2][2] = L[3][3] = {0, xs * p[1]};
        const double k = ks * p[2];
        L[0][3] = {k, 0}; L[1][2] = {-k, 0};
        L[2][1] = {k, 0}; L[3][0] = {-k, 0};
    } else {
        L[0][0] = {0, bs * p[0]}; L[1][1] = {0, bs * p[1]}; L[0][1] = L[1][0] = {0, bs * p[2]};
        L[2][2] = {0, xs * p[3]}; L[3][3] = {0, xs * p[4]}; L[2][3] = L[3][2] = {0, xs * p[5]};
        const double K00 = ks * p[6], K01 = ks * p[7], K10 = ks * p[8], K11 = ks * p[9];
        L[0][2] = {K00, 0}; L[0][3] = {K01, 0}; L[1][2] = {K10, 0}; L[1][3] = {K11, 0};
        // -K^T
        L[2][0] = {-K00, 0}; L[2][1] = {-K10, 0}; L[3][0] = {-K01, 0}; L[3][1] = {-K11, 0};
    }
}

__global__ void records_kernel(int nf, int kind, const double* __restrict__ params, double bs, double xs, double ks,
                               double eta, cplx* __restrict__ rec, int* __restrict__ nsing) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    cplx L[4][4], Bal[4][4], Bi[4][4];
    build_lambda(kind, params + (size_t)f * kind, bs, xs, ks, L);
    // impedance-balanced inverse: T = diag(sqrt(eta) x2, 1/sqrt(eta) x2), Lambda^-1 = T (T L T)^-1 T
    const double se = sqrt(eta);
    const double t[4] = {se, se, 1.0 / se, 1.0 / se};
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Bal[i][j] = {L[i][j].re * t[i] * t[j], L[i][j].im * t[i] * t[j]};
    const bool ok = inv4(Bal, Bi);
    if (!ok) atomicAdd(nsing, 1);
    cplx* o = rec + (size_t)f * 32;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            o[4 * i + j] = L[i][j];
            o[16 + 4 * i + j] = ok ? cplx{Bi[i][j].re * t[i] * t[j], Bi[i][j].im * t[i] * t[j]} : cplx{0, 0};
        }
}

// ---------------------------------------------------------- far field
struct Dir { double r[3], th[3], ph[3]; };

// One workgroup per direction: N and L vectors summed over (face, point).
__global__ __launch_bounds__(256) void farfield_kernel(int nf, int ne, int Q, const double* __restrict__ V,
                                                       const int* __restrict__ F, const int* __restrict__ edge,
                                                       const double* __restrict__ len, const float* __restrict__ qxy,
                                                       const float* __restrict__ qw, double k0, double eta0,
                                                       const cplx* __restrict__ xs, const Dir* __restrict__ dirs,
                                                       cplx* __restrict__ out) {
    const Dir d = dirs[blockIdx.x];
    double acc[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};  // N (re,im)x3, L (re,im)x3
    const long long tot = (long long)nf * Q;
    for (long long t = threadIdx.x; t < tot; t += blockDim.x) {
        const int f = (int)(t / Q), q = (int)(t - (long long)f * Q);
        double v[3][3];
        for (int a = 0; a < 3; ++a) {
            const int vi = F[3 * f + a];
            v[a][0] = V[3 * vi]; v[a][1] = V[3 * vi + 1]; v[a][2] = V[3 * vi + 2];
        }
        const double xi = qxy[2 * q], eta = qxy[2 * q + 1], ze = 1.0 - xi - eta, w = qw[q];
        double r[3];
        for (int x = 0; x < 3; ++x) r[x] = v[0][x] * ze + v[1][x] * xi + v[2][x] * eta;
        double s, c;
        sincos(k0 * (d.r[0] * r[0] + d.r[1] * r[1] + d.r[2] * r[2]), &s, &c);
        // e^{+j k0 rhat.r} * w
        const cplx ph = {c * w, s * w};
        for (int le = 0; le < 3; ++le) {
            const int e = edge[le * (size_t)nf + f];
            const double h = 0.5 * len[le * (size_t)nf + f];
            const cplx J = xs[2 * (size_t)ne + e];
            const cplx M = {xs[3 * (size_t)ne + e].re * eta0, xs[3 * (size_t)ne + e].im * eta0};
            const double* fv = v[(le + 2) % 3];
            const cplx pj = cm(ph, {J.re * h, J.im * h}), pm = cm(ph, {M.re * h, M.im * h});
            for (int x = 0; x < 3; ++x) {
                const double rho = r[x] - fv[x];
                acc[2 * x] += pj.re * rho; acc[2 * x + 1] += pj.im * rho;
                acc[6 + 2 * x] += pm.re * rho; acc[6 + 2 * x + 1] += pm.im * rho;
            }
        }
    }
    __shared__ double red[12][256];
    for (int k = 0; k < 12; ++k) red[k][threadIdx.x] = acc[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 12; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cplx N[3], L[3];
        for (int x = 0; x < 3; ++x) { N[x] = {red[2 * x][0], red[2 * x + 1][0]}; L[x] = {red[6 + 2 * x][0], red[7 + 2 * x][0]}; }
        auto dotc = [](const cplx* a, const double* u) {
            return cplx{a[0].re * u[0] + a[1].re * u[1] + a[2].re * u[2], a[0].im * u[0] + a[1].im * u[1] + a[2].im * u[2]};
        };
        const cplx Nt = dotc(N, d.th), Np = dotc(N, d.ph), Lt = dotc(L, d.th), Lp = dotc(L, d.ph);
        const cplx coef = {0.0, -k0 / (4.0 * 3.14159265358979323846)};
        const cplx ft = cm(coef, {eta0 * Nt.re + Lp.re, eta0 * Nt.im + Lp.im});
        const cplx fp = cm(coef, {eta0 * Np.re - Lt.re, eta0 * Np.im - Lt.im});
        out[2 * blockIdx.x] = ft;
        out[2 * blockIdx.x + 1] = fp;
    }
}

// g(xs) = A^H u for the far-field map F = A xs, u[d] = (u_theta, u_phi).
// One thread per (edge, J/M component).
__global__ void farfield_adjoint_kernel(int nf, int ne, int Q, int ndir, const double* __restrict__ V,
                                        const int* __restrict__ F, const int* __restrict__ ef,
                                        const int* __restrict__ el, const double* __restrict__ len,
                                        const float* __restrict__ qxy, const float* __restrict__ qw, double k0,
                                        double eta0, const Dir* __restrict__ dirs, const cplx* __restrict__ u,
                                        cplx* __restrict__ g) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const cplx coef = {0.0, -k0 / (4.0 * 3.14159265358979323846)};
    cplx gJ = {0, 0}, gM = {0, 0};
    for (int d = 0; d < ndir; ++d) {
        const Dir D = dirs[d];
        // n_de = sum over the edge's faces and points of w e^{jk0 r.r'} (l/2) rho
        cplx n[3] = {{0, 0}, {0, 0}, {0, 0}};
        for (int b = 0; b < 2; ++b) {
            const int f = ef[b * ne + e], le = el[b * ne + e];
            double v[3][3];
            for (int a = 0; a < 3; ++a) {
                const int vi = F[3 * f + a];
                v[a][0] = V[3 * vi]; v[a][1] = V[3 * vi + 1]; v[a][2] = V[3 * vi + 2];
            }
            const double h = 0.5 * len[le * (size_t)nf + f];
            const double* fv = v[(le + 2) % 3];
            for (int q = 0; q < Q; ++q) {
                const double xi = qxy[2 * q], et = qxy[2 * q + 1], ze = 1.0 - xi - et, w = qw[q];
                double r[3];
                for (int x = 0; x < 3; ++x) r[x] = v[0][x] * ze + v[1][x] * xi + v[2][x] * et;
                double s, c;
                sincos(k0 * (D.r[0] * r[0] + D.r[1] * r[1] + D.r[2] * r[2]), &s, &c);
                for (int x = 0; x < 3; ++x) {
                    const double rho = (r[x] - fv[x]) * h * w;
                    n[x].re += c * rho;
                    n[x].im += s * rho;
                }
            }
        }
        const cplx nt = {n[0].re * D.th[0] + n[1].re * D.th[1] + n[2].re * D.th[2],
                         n[0].im * D.th[0] + n[1].im * D.th[1] + n[2].im * D.th[2]};
        const cplx np = {n[0].re * D.ph[0] + n[1].re * D.ph[1] + n[2].re * D.ph[2],
                         n[0].im * D.ph[0] + n[1].im * D.ph[1] + n[2].im * D.ph[2]};
        // dF_theta/dJ = c eta0 nt, dF_phi/dJ = c eta0 np ; dF_theta/dxsM = c eta0 np, dF_phi/dxsM = -c eta0 nt
        const cplx a_t = cm(coef, {eta0 * nt.re, eta0 * nt.im});
        const cplx a_p = cm(coef, {eta0 * np.re, eta0 * np.im});
        const cplx ut = u[2 * d], up = u[2 * d + 1];
        const cplx t1 = cm(cconj(a_t), ut), t2 = cm(cconj(a_p), up);
        gJ.re += t1.re + t2.re; gJ.im += t1.im + t2.im;
        const cplx t3 = cm(cconj(a_p), ut), t4 = cm(cconj({-a_t.re, -a_t.im}), up);
        gM.re += t3.re + t4.re; gM.im += t3.im + t4.im;
    }
    g[e] = {0, 0};
    g[(size_t)ne + e] = {0, 0};
    g[2 * (size_t)ne + e] = gJ;
    g[3 * (size_t)ne + e] = gM;
}

// ---------------------------------------------------------- mass sensitivity
// dS/drec_k, S = lambda^H Zs_mass(rec) xs, per face (one lane per face).
// Block pattern of the eta0-scaled GSTC mass block (sie_kernels_assembly.cuh:
// 555-592): entry (r, c) = s_rc (0.5 I_Z[zb] + sg * 0.125 I_Y[yb]).
__constant__ int kZb[4][4] = {{0, 1, 0, 1}, {2, 3, 2, 3}, {0, 1, 0, 1}, {2, 3, 2, 3}};
__constant__ int kYb[4][4] = {{3, 2, 3, 2}, {1, 0, 1, 0}, {3, 2, 3, 2}, {1, 0, 1, 0}};
__constant__ int kSg[4][4] = {{1, -1, -1, 1}, {-1, 1, 1, -1}, {-1, 1, 1, -1}, {1, -1, -1, 1}};

__global__ __launch_bounds__(64) void mass_sensitivity_kernel(
    int nf, int ne, int Q, const double* __restrict__ V, const int* __restrict__ F, const int* __restrict__ edge,
    const double* __restrict__ len, const double* __restrict__ area, const float* __restrict__ qxy,
    const float* __restrict__ qw, int basis, float mu0, float eps0, const cplx* __restrict__ lam,
    const cplx* __restrict__ xs, cplx* __restrict__ G) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double t[3][3];
    for (int a = 0; a < 3; ++a) {
        const int v = F[3 * f + a];
        t[a][0] = V[3 * v]; t[a][1] = V[3 * v + 1]; t[a][2] = V[3 * v + 2];
    }
    // frame (gstc_tensor_utils.cuh:72-147)
    double u[3], w[3];
    {
        const double e1[3] = {t[1][0] - t[0][0], t[1][1] - t[0][1], t[1][2] - t[0][2]};
        const double e2[3] = {t[2][0] - t[0][0], t[2][1] - t[0][1], t[2][2] - t[0][2]};
        const double nu[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        if (basis == 0) {
            const double l1 = sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
            const double ln = sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
            const double nh[3] = {nu[0] / ln, nu[1] / ln, nu[2] / ln};
            for (int x = 0; x < 3; ++x) u[x] = e1[x] / l1;
            w[0] = nh[1] * u[2] - nh[2] * u[1]; w[1] = nh[2] * u[0] - nh[0] * u[2]; w[2] = nh[0] * u[1] - nh[1] * u[0];
        } else {
            const double c[3] = {(t[0][0] + t[1][0] + t[2][0]) / 3.0, (t[0][1] + t[1][1] + t[2][1]) / 3.0,
                                 (t[0][2] + t[1][2] + t[2][2]) / 3.0};
            const double r = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            double ct = c[2] / r;
            ct = ct > 1 ? 1 : (ct < -1 ? -1 : ct);
            const double th = acos(ct), ph = atan2(c[1], c[0]);
            u[0] = cos(th) * cos(ph); u[1] = cos(th) * sin(ph); u[2] = -sin(th);
            w[0] = -sin(ph); w[1] = cos(ph); w[2] = 0.0;
        }
    }
    double nvec[3];
    {
        const double e1[3] = {t[1][0] - t[0][0], t[1][1] - t[0][1], t[1][2] - t[0][2]};
        const double e2[3] = {t[2][0] - t[0][0], t[2][1] - t[0][1], t[2][2] - t[0][2]};
        const double nu[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double ln = 1.0 / sqrt(nu[0] * nu[0] + nu[1] * nu[1] + nu[2] * nu[2]);
        for (int x = 0; x < 3; ++x) nvec[x] = nu[x] * ln;
    }
    double Nf[3][3] = {{0.0, -nvec[2], nvec[1]}, {nvec[2], 0.0, -nvec[0]}, {-nvec[1], nvec[0], 0.0}};
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Nf[a][b] = (double)(float)Nf[a][b];
    // moments Mo[j][i][a][b] = fs(j,i) sum_p w_p rho_j,a rho_i,b
    float el[3];
    for (int i = 0; i < 3; ++i) el[i] = (float)len[i * (size_t)nf + f];
    const float sc = 1.f / (4.f * (float)area[f]);
    double Mo[3][3][3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) Mo[j][i][a][b] = 0;
    for (int p = 0; p < Q; ++p) {
        const double xi = qxy[2 * p], et = qxy[2 * p + 1], ze = 1.0 - xi - et, wp = qw[p];
        double rho[3][3];
        for (int x = 0; x < 3; ++x) {
            const double r = t[0][x] * ze + t[1][x] * xi + t[2][x] * et;
            for (int i = 0; i < 3; ++i) rho[i][x] = r - t[(i + 2) % 3][x];
        }
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) Mo[j][i][a][b] += wp * rho[j][a] * rho[i][b];
    }
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            const double fs = (double)(sc * el[j] * el[i]);
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) Mo[j][i][a][b] *= fs;
        }
    // P[r][c][j][i] = conj(lambda_{r,e_j}) xs_{c,e_i} s_rc
    const float e0 = sqrtf(mu0 / eps0);
    const double s1 = e0, s2 = (double)(e0 * e0);
    int eidx[3];
    for (int i = 0; i < 3; ++i) eidx[i] = edge[i * (size_t)nf + f];
    cplx lv[4][3], xv[4][3];
    for (int r = 0; r < 4; ++r)
        for (int i = 0; i < 3; ++i) {
            lv[r][i] = cconj(lam[(size_t)r * ne + eidx[i]]);
            xv[r][i] = xs[(size_t)r * ne + eidx[i]];
        }
    const double* R0 = u;
    const double* R1 = w;
    cplx* Gf = G + (size_t)f * 32;
    // record entry k: matrix (0 = Lambda -> Y tensors, 1 = Lambda^-1 -> Z tensors), block, local (a', b')
    for (int src = 0; src < 2; ++src)
        for (int A = 0; A < 4; ++A)
            for (int B = 0; B < 4; ++B) {
                const int blk = (A >> 1) * 2 + (B >> 1);
                const double* ua = (A & 1) ? R1 : R0;
                const double* ub = (B & 1) ? R1 : R0;
                // unit local tensor -> global 3x3 (real)
                double Tg[3][3];
                for (int x = 0; x < 3; ++x)
                    for (int y = 0; y < 3; ++y) Tg[x][y] = ua[x] * ub[y];
                if (src == 0) {  // perp: -N T N
                    double tmp[3][3], o2[3][3];
                    for (int x = 0; x < 3; ++x)
                        for (int y = 0; y < 3; ++y) {
                            double s = 0;
                            for (int k = 0; k < 3; ++k) s += Tg[x][k] * Nf[k][y];
                            tmp[x][y] = s;
                        }
                    for (int x = 0; x < 3; ++x)
                        for (int y = 0; y < 3; ++y) {
                            double s = 0;
                            for (int k = 0; k < 3; ++k) s += Nf[x][k] * tmp[k][y];
                            o2[x][y] = -s;
                        }
                    for (int x = 0; x < 3; ++x)
                        for (int y = 0; y < 3; ++y) Tg[x][y] = o2[x][y];
                }
                // I[j][i] = sum_ab Tg[a][b] Mo[j][i][a][b]
                double I[3][3];
                for (int j = 0; j < 3; ++j)
                    for (int i = 0; i < 3; ++i) {
                        double s = 0;
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b) s += Tg[a][b] * Mo[j][i][a][b];
                        I[j][i] = s;
                    }
                cplx acc = {0, 0};
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 4; ++c) {
                        double coef;
                        if (src == 1) coef = (kZb[r][c] == blk) ? 0.5 : 0.0;
                        else coef = (kYb[r][c] == blk) ? 0.125 * kSg[r][c] : 0.0;
                        if (coef == 0.0) continue;
                        const bool mr = r & 1, mc = c & 1;
                        const double srs = (mr && mc) ? s2 : (mr || mc) ? s1 : 1.0;
                        coef *= srs;
                        for (int j = 0; j < 3; ++j)
                            for (int i = 0; i < 3; ++i) {
                                const cplx pr = cm(lv[r][j], xv[c][i]);
                                acc.re += coef * I[j][i] * pr.re;
                                acc.im += coef * I[j][i] * pr.im;
                            }
                    }
                Gf[src * 16 + 4 * A + B] = acc;
            }
}

// Chain rule from record sensitivities to real parameters:
// G_tot = G_Lambda - Lambda^-T G_inv Lambda^-T ; dJ/dtheta = -2 Re sum G_tot dLambda/dtheta.
__global__ void param_grad_kernel(int nf, int kind, const cplx* __restrict__ rec, const cplx* __restrict__ G,
                                  double bs, double xs, double ks, double* __restrict__ grad) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const cplx* R = rec + (size_t)f * 32;
    const cplx* Gf = G + (size_t)f * 32;
    cplx T[4][4], Gt[4][4];
    // T = Lambda^-T G_inv  ->  T[a][b] = sum_c inv[c][a] Ginv[c][b]
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            cplx s = {0, 0};
            for (int c = 0; c < 4; ++c) {
                const cplx p = cm(R[16 + 4 * c + a], Gf[16 + 4 * c + b]);
                s.re += p.re; s.im += p.im;
            }
            T[a][b] = s;
        }
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) {
            cplx s = {0, 0};
            for (int c = 0; c < 4; ++c) {  // (T Lambda^-T)[a][b] = sum_c T[a][c] inv[b][c]
                const cplx p = cm(T[a][c], R[16 + 4 * b + c]);
                s.re += p.re; s.im += p.im;
            }
            Gt[a][b] = {Gf[4 * a + b].re - s.re, Gf[4 * a + b].im - s.im};
        }
    double* g = grad + (size_t)f * kind;
    auto dre = [&](cplx z) { return -2.0 * z.re; };         // dLambda real
    auto dim = [&](cplx z) { return -2.0 * (-z.im); };      // dLambda = j * (.) : Re(G j) = -Im G
    if (kind == 3) {
        g[0] = dim({(Gt[0][0].re + Gt[1][1].re) * bs, (Gt[0][0].im + Gt[1][1].im) * bs});
        g[1] = dim({(Gt[2][2].re + Gt[3][3].re) * xs, (Gt[2][2].im + Gt[3][3].im) * xs});
        g[2] = dre({(Gt[0][3].re - Gt[1][2].re + Gt[2][1].re - Gt[3][0].re) * ks,
                    (Gt[0][3].im - Gt[1][2].im + Gt[2][1].im - Gt[3][0].im) * ks});
    } else {
        g[0] = dim({Gt[0][0].re * bs, Gt[0][0].im * bs});
        g[1] = dim({Gt[1][1].re * bs, Gt[1][1].im * bs});
        g[2] = dim({(Gt[0][1].re + Gt[1][0].re) * bs, (Gt[0][1].im + Gt[1][0].im) * bs});
        g[3] = dim({Gt[2][2].re * xs, Gt[2][2].im * xs});
        g[4] = dim({Gt[3][3].re * xs, Gt[3][3].im * xs});
        g[5] = dim({(Gt[2][3].re + Gt[3][2].re) * xs, (Gt[2][3].im + Gt[3][2].im) * xs});
        g[6] = dre({(Gt[0][2].re - Gt[2][0].re) * ks, (Gt[0][2].im - Gt[2][0].im) * ks});
        g[7] = dre({(Gt[0][3].re - Gt[3][0].re) * ks, (Gt[0][3].im - Gt[3][0].im) * ks});
        g[8] = dre({(Gt[1][2].re - Gt[2][1].re) * ks, (Gt[1][2].im - Gt[2][1].im) * ks});
        g[9] = dre({(Gt[1][3].re - Gt[3][1].re) * ks, (Gt[1][3].im - Gt[3][1].im) * ks});
    }
}

}  // namespace
}  // namespace sie

using namespace sie;

#define AGUARD_BEGIN try {
#define AGUARD_END                    \
    set_last_error("");               \
    return SIE_OK;                    \
    }                                 \
    catch (const Error& e) {          \
        set_last_error(e.what());     \
        return e.code;                \
    }                                 \
    catch (const std::exception& e) { \
        set_last_error(e.what());     \
        return SIE_ERR_DEVICE;        \
    }

extern "C" {

int sie_gstc_records(int nf, int kind, const double* params_dev, double b_scale, double x_scale, double k_scale,
                     double eta_ref, double* records_dev, int* n_singular, void* stream) {
    AGUARD_BEGIN
    require(nf > 0 && (kind == 3 || kind == 10), "kind must be 3 (omega) or 10 (general lossless)");
    require(params_dev && records_dev && n_singular, "null argument");
    hipStream_t s = (hipStream_t)stream;
    DevBuf<int> cnt(1);
    SIE_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int), s));
    records_kernel<<<(nf + 127) / 128, 128, 0, s>>>(nf, kind, params_dev, b_scale, x_scale, k_scale, eta_ref,
                                                    (cplx*)records_dev, cnt.p);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipMemcpyAsync(n_singular, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SIE_HIP(hipStreamSynchronize(s));
    if (*n_singular)
        throw Error(SIE_ERR_SINGULAR, "surface map is singular at " + std::to_string(*n_singular) + " faces");
    AGUARD_END
}

int sie_farfield(sie_mesh_t mesh, int rule, double k0, double eta0, const double* x_dev, int ndir,
                 const double* dirs, double* F_dev, void* stream) {
    AGUARD_BEGIN
    require(mesh && x_dev && dirs && F_dev && ndir > 0, "null argument");
    hipStream_t s = (hipStream_t)stream;
    Mesh& m = mesh_of(mesh);
    QuadDev& q = mesh_quad(mesh, rule, s);
    DevBuf<Dir> d(ndir);
    d.upload((const Dir*)dirs, ndir, s);
    farfield_kernel<<<ndir, 256, 0, s>>>(m.nf, m.ne, q.n, m.d_V.p, m.d_F.p, m.d_edge.p, m.d_len.p,
                                         q.xy.p, q.w.p, k0, eta0, (const cplx*)x_dev, d.p, (cplx*)F_dev);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    AGUARD_END
}

int sie_farfield_adjoint(sie_mesh_t mesh, int rule, double k0, double eta0, int ndir, const double* dirs,
                         const double* u_dev, double* g_dev, void* stream) {
    AGUARD_BEGIN
    require(mesh && u_dev && dirs && g_dev && ndir > 0, "null argument");
    hipStream_t s = (hipStream_t)stream;
    Mesh& m = mesh_of(mesh);
    QuadDev& q = mesh_quad(mesh, rule, s);
    DevBuf<Dir> d(ndir);
    d.upload((const Dir*)dirs, ndir, s);
    farfield_adjoint_kernel<<<(m.ne + 127) / 128, 128, 0, s>>>(m.nf, m.ne, q.n, ndir, m.d_V.p, m.d_F.p, m.d_ef.p,
                                                               m.d_el.p, m.d_len.p, q.xy.p, q.w.p, k0, eta0, d.p,
                                                               (const cplx*)u_dev, (cplx*)g_dev);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    AGUARD_END
}

int sie_gstc_mass_sensitivity(sie_mesh_t mesh, int rule, int basis, float mu0, float eps0, const double* lambda_dev,
                              const double* x_dev, double* G_dev, void* stream) {
    AGUARD_BEGIN
    require(mesh && lambda_dev && x_dev && G_dev, "null argument");
    require(basis == 0 || basis == 1, "unsupported GSTC basis type");
    hipStream_t s = (hipStream_t)stream;
    Mesh& m = mesh_of(mesh);
    QuadDev& q = mesh_quad(mesh, rule, s);
    prof().begin(PROF_CONTRACT, s);
    mass_sensitivity_kernel<<<(m.nf + 63) / 64, 64, 0, s>>>(m.nf, m.ne, q.n, m.d_V.p, m.d_F.p, m.d_edge.p,
                                                            m.d_len.p, mesh_area(mesh), q.xy.p, q.w.p, basis, mu0,
                                                            eps0, (const cplx*)lambda_dev, (const cplx*)x_dev,
                                                            (cplx*)G_dev);
    // algorithmic bytes: lambda and x (4E complex each) read once, the 32
    // complex sensitivities per face written, the face geometry (vertices,
    // indices, lengths, area) read once
    prof().end(PROF_CONTRACT, s, 0.0,
               2.0 * 16.0 * 4 * m.ne + 32.0 * 16.0 * m.nf + (double)m.nf * (9 * 8 + 6 * 4 + 4 * 8));
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    AGUARD_END
}

int sie_gstc_param_grad(int nf, int kind, const double* records_dev, const double* G_dev, double b_scale,
                        double x_scale, double k_scale, double* grad_dev, void* stream) {
    AGUARD_BEGIN
    require(nf > 0 && (kind == 3 || kind == 10), "kind must be 3 or 10");
    hipStream_t s = (hipStream_t)stream;
    param_grad_kernel<<<(nf + 127) / 128, 128, 0, s>>>(nf, kind, (const cplx*)records_dev, (const cplx*)G_dev, b_scale,
                                                       x_scale, k_scale, grad_dev);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    AGUARD_END
}

}  // extern "C"
