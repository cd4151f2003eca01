// SIE-MoM impedance assembly on MI355X (gfx950): T-EFIE, T-PMCHWT, SIE-GSTC.
//
// Pipeline per chunk of source faces [n0, n0+nc):
//   1. near_list_kernel  - pairs whose float centroid distance is below the
//                          threshold (same predicate as the reference kernels,
//                          sie_kernels_assembly.cuh:193-197) are compacted.
//   2. far_pairs_kernel  - one lane per test face, one wave per source face:
//                          the Q x Q double quadrature of every pair but the
//                          self pair in a moment-factored form (below), with
//                          the regularised kernels on near pairs, writing the
//                          raw TL/TK 3x3 blocks of each pair to the scratch W.
//   3. self_pairs_kernel - one workgroup per self pair: its regularised
//                          quadrature (dunavant_quadrature.cuh:207-324 with
//                          is_singular) into its slots of W;
//      near_pairs_kernel - one lane per near pair: the Wilton singular part
//                          (dunavant_quadrature.cuh:769-875) added to the near
//                          slots of W.
//   4. gather kernels    - one lane per (test edge, source edge): sum the four
//                          face-pair contributions of the RWG pair in a fixed
//                          order, apply the formulation's epilogue (and the
//                          GSTC mass blocks of shared faces) and update Z with
//                          coalesced column-major stores. No atomics: the
//                          matrix is bitwise reproducible run to run.
//
// Moment factoring of the regular double quadrature. With rho_j(p) = a_p + d_j
// (a_p = r_p - c_m, d_j = c_m - v_j) and rho_i(q) = b_q + e_i (b_q = r_q - c_n,
// e_i = c_n - v_i):
//   sum_pq w w G (rho_j . rho_i) = S_ab + S_a . e_i + d_j . S_b + (d_j . e_i) S_0
// with S_0 = sum w w G, S_a = sum w w G a_p, S_b = sum w w G b_q,
// S_ab = sum w w G a_p . b_q; and since grad'G = g(R) (r_p - r_q) and
// (rho_j x rho_i) . (r_p - r_q) = (rho_j x rho_i) . (v_j - v_i),
//   sum_pq w w g (rho_j x rho_i) . R = [U_ab + U_a x e_i + d_j x U_b
//                                       + U_0 (d_j x e_i)] . (v_j - v_i).
// The inner q-loop therefore accumulates 4 complex moments per operator and
// wavenumber (G, G b_q and g, g b_q) instead of 9 per operator.
#include "sie_internal.hpp"
#include "prof.hpp"
#define SIE_HD __host__ __device__
#include "trig.hpp"

#include <algorithm>
#include <cstdlib>
#include <cmath>

namespace sie {

namespace {

constexpr double kPi = 3.14159265358979323846;
// FP64 operations per point pair in far_pairs_kernel, counted on the source:
// geometry (|R|^2 = |P|^2 + |b|^2 - 2 P.b, |R|, kR, weight scale) and, per
// wavenumber, G, g and the 8 complex moment accumulations. sqrt/rsqrt and
// sincos are counted as one operation each (their instruction sequences are not).
constexpr double kFarGeomFlops = 7 + 1 + 2;
constexpr double kFarPairFlops = 1 + 1 + 2 + 6 + 4 + 28;
constexpr double kInv4Pi = 1.0 / (4.0 * kPi);

struct cplx { double re, im; };

__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// 1/sqrt(x) for finite x > 0: the hardware estimate and one third-order
// correction, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2 (the device library's
// rsqrt without its zero/infinity handling).
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// Largest quadrature rule of the far-pair kernel (its weights are staged in LDS).
constexpr int kMaxFarQ = 128;

// Float centroid distance < float threshold, without FMA contraction so the
// predicate is bit-identical to the reference (float dist = length(c_m - c_n)).
__device__ __forceinline__ bool is_near(const double* __restrict__ c, int nf, int m, int n, float thr) {
    const double dx = __dsub_rn(c[m], c[n]);
    const double dy = __dsub_rn(c[nf + m], c[nf + n]);
    const double dz = __dsub_rn(c[2 * nf + m], c[2 * nf + n]);
    const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
    const float dist = __double2float_rn(__dsqrt_rn(r2));
    return dist < thr;
}

// ------------------------------------------------------------------ step 0
// Source points relative to their centroid, b_q = r_q - c (mapTriQuadPoint),
// in units of 1/k (scale = k) with their squared length:
// bq[4 t .. 4 t + 3] = (k b_x, k b_y, k b_z, k^2 |b|^2).
__global__ void source_points_kernel(int nf, int Q, const double* __restrict__ o,
                                     const float* __restrict__ qxy, double scale, double* __restrict__ bq) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nf * Q) return;
    const int f = t / Q, q = t - f * Q;
    const double xi = qxy[2 * q], eta = qxy[2 * q + 1], ze = 1.0 - xi - eta;
    double b[3];
    for (int x = 0; x < 3; ++x) {
        b[x] = scale * (o[(0 + x) * (size_t)nf + f] * ze + o[(3 + x) * (size_t)nf + f] * xi +
                        o[(6 + x) * (size_t)nf + f] * eta);
        bq[4 * (size_t)t + x] = b[x];
    }
    bq[4 * (size_t)t + 3] = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
}

// ------------------------------------------------------------------ step 1
__global__ void near_list_kernel(int nf, int n0, int nc, const double* __restrict__ c, float thr,
                                 int2* __restrict__ list, unsigned int* __restrict__ count,
                                 unsigned int cap) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nf * nc) return;
    const int nl = (int)(t / nf), m = (int)(t - (long long)nl * nf);
    const int n = n0 + nl;
    if (is_near(c, nf, m, n, thr)) {
        const unsigned int k = atomicAdd(count, 1u);
        if (k < cap) list[k] = make_int2(m, n);
    }
}

// ------------------------------------------------------------------ step 2
// W layout: W[((nl * NS) + s) * nf + m], s = (kq*2 + op)*9 + j*3 + i,
// op 0 = TL, 1 = TK; j = test local edge, i = source local edge.
// SYM (one chunk holding every source face, n0 = 0): only pairs with test
// face m > source face n are integrated, and each also fills the slot of the
// swapped pair (test n, source m) with the transposed blocks, which are equal:
// TL(n, m)[i][j] = TL(m, n)[j][i] because the same rule integrates both faces,
// and TK likewise since swapping the faces flips both rho_j x rho_i and r - r'.
// Lengths in the point loops are in units of 1/k1 (bq from source_points_kernel
// with scale k1): kR is then |R|^2 / |R|, the weights carry k1, and g / k1^2 =
// G (1/R' + j k/k1) / R' (R' = k1 R) takes two operations fewer; the moments
// are scaled back before the epilogue.
// One wavenumber: at most 168 registers, three waves per SIMD (the compiler
// keeps a few of the per-test-point accumulators outside the point loop in
// scratch: 2.6% less time for the far pairs at 72k than at two waves).
template <int NK, bool SYM>
__global__ __launch_bounds__(256, NK == 1 ? 3 : 1) void far_pairs_kernel(
    int nf, int n0, int nc, int Q, const double* __restrict__ c, const double* __restrict__ o,
    const double* __restrict__ len, const double* __restrict__ bq, const float* __restrict__ qxy,
    const float* __restrict__ qw, const double* __restrict__ sctab, double k1, double k2, float thr,
    cplx* __restrict__ W) {
    constexpr int NS = NK * 18;
    // the e^{j 2 pi i / 2048} table and the source weights k1 w_q / (4 pi), shared by the workgroup
    __shared__ double stab[2 * kSinCosTabN];
    __shared__ double swq[kMaxFarQ];
    for (int t = threadIdx.x; t < 2 * kSinCosTabN; t += blockDim.x) stab[t] = sctab[t];
    for (int t = threadIdx.x; t < Q; t += blockDim.x) swq[t] = (double)qw[t] * kInv4Pi * k1;
    __syncthreads();
    const int m = blockIdx.x * 64 + (threadIdx.x & 63);
    const int nl = uniform(blockIdx.y * 4 + (threadIdx.x >> 6));
    if (nl >= nc) return;
    const int n = n0 + nl;
    if (SYM && (int)blockIdx.x * 64 + 63 <= n) return;  // no test face of this wave lies above n
    const int mm = m < nf ? m : nf - 1;
    // every pair but the self pair; near pairs (the reference's float centroid
    // predicate) take the regularised kernels G - 1/(4 pi R) and
    // K - (1/R^3 + k^2/(2R))/(4 pi) (integrandKernelsRegular,
    // green_functions_utils.cuh:295-361), whose singular parts
    // near_pairs_kernel adds
    const bool active = (m < nf) && (!SYM || m > n) && m != n;
    const bool nearp = active && is_near(c, nf, mm, n, thr);
    const double rg = nearp ? 1.0 : 0.0;
    const bool anyreg = __any(nearp);

    // test-face geometry (per lane)
    double om[3][3];
    for (int a = 0; a < 3; ++a)
        for (int x = 0; x < 3; ++x) om[a][x] = o[(3 * a + x) * (size_t)nf + mm];
    const double dlt[3] = {c[mm] - c[n], c[nf + mm] - c[nf + n], c[2 * nf + mm] - c[2 * nf + n]};
    const double kk[2] = {k1, k2};
    const double kap[2] = {1.0, k2 / k1};  // wavenumbers in units of k1

    // outer accumulators: per wavenumber, TL: S0, Sa[3], Sb[3], Sab; TK: U0, Ua[3], Ub[3], Uab[3]
    cplx S0[NK], Sa[NK][3], Sb[NK][3], Sab[NK], U0[NK], Ua[NK][3], Ub[NK][3], Uab[NK][3];
#pragma unroll
    for (int w = 0; w < NK; ++w) {
        S0[w] = {0, 0}; Sab[w] = {0, 0}; U0[w] = {0, 0};
#pragma unroll
        for (int x = 0; x < 3; ++x) { Sa[w][x] = Sb[w][x] = Ua[w][x] = Ub[w][x] = Uab[w][x] = {0, 0}; }
    }
    const double* bqn = bq + (size_t)n * Q * 4;

    for (int p = 0; p < Q; ++p) {
        const double xi = qxy[2 * p], eta = qxy[2 * p + 1], wp = qw[p];
        const double ze = 1.0 - xi - eta;
        double ap[3], P[3];  // P in units of 1/k1
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            ap[x] = om[0][x] * ze + om[1][x] * xi + om[2][x] * eta;
            P[x] = k1 * (dlt[x] + ap[x]);
        }
        // |R|^2 = |P - b_q|^2 = (|P|^2 + |b_q|^2) - 2 P.b_q: four operations per
        // point pair instead of six (the rounding error, ~4 eps (|P|^2 + |b|^2),
        // is relative to |R|^2 within 20 eps on far pairs; near pairs take
        // the same R in the pass that removes the static terms below)
        const double m2P[3] = {-2.0 * P[0], -2.0 * P[1], -2.0 * P[2]};
        const double PP = P[0] * P[0] + P[1] * P[1] + P[2] * P[2];
        auto rsq = [&](int q) {
            const double* b = bqn + 4 * q;
            return fma(m2P[0], b[0], fma(m2P[1], b[1], fma(m2P[2], b[2], PP + b[3])));
        };
        cplx T0[NK], Tb[NK][3], V0[NK], Vb[NK][3];
#pragma unroll
        for (int w = 0; w < NK; ++w) {
            T0[w] = {0, 0}; V0[w] = {0, 0};
#pragma unroll
            for (int x = 0; x < 3; ++x) { Tb[w][x] = {0, 0}; Vb[w][x] = {0, 0}; }
        }
        for (int q = 0; q < Q; ++q) {
            const double bx = bqn[4 * q], by = bqn[4 * q + 1], bz = bqn[4 * q + 2];
            const double R2 = rsq(q);  // (k1 R)^2
            // R2 > 0 on far pairs; lanes of other pairs compute values that are not stored
            const double iR = rsqrt_pos(R2);
            const double kR = R2 * iR;  // k1 R
            const double gsc = swq[q] * iR;
#pragma unroll
            for (int w = 0; w < NK; ++w) {
                double s, co;
                sincos_tab(w == 0 ? kR : kap[w] * kR, stab, &s, &co);
                // w_q G = w_q e^{-jkR} / (4 pi R)
                const double Gr = co * gsc, Gi = -s * gsc;
                // w_q g / k1^2 = w_q G (1/R' + j k/k1) / R'   (grad'G = g R_vec, R' = k1 R)
                const double gr = fma(Gr, iR, w == 0 ? -Gi : -Gi * kap[w]) * iR;
                const double gi = fma(Gi, iR, w == 0 ? Gr : Gr * kap[w]) * iR;
                T0[w].re += Gr; T0[w].im += Gi;
                V0[w].re += gr; V0[w].im += gi;
                Tb[w][0].re += Gr * bx; Tb[w][0].im += Gi * bx;
                Tb[w][1].re += Gr * by; Tb[w][1].im += Gi * by;
                Tb[w][2].re += Gr * bz; Tb[w][2].im += Gi * bz;
                Vb[w][0].re += gr * bx; Vb[w][0].im += gi * bx;
                Vb[w][1].re += gr * by; Vb[w][1].im += gi * by;
                Vb[w][2].re += gr * bz; Vb[w][2].im += gi * bz;
            }
        }
        if (anyreg) {
            // near lanes: the static parts of the kernels, w_q/(4 pi R) in G and
            // w_q (1/R^3 + k^2/(2R))/(4 pi) in g, taken out of the real moments
            // (a separate pass, so that the loop above stays that of far pairs)
            for (int q = 0; q < Q; ++q) {
                const double bx = bqn[4 * q], by = bqn[4 * q + 1], bz = bqn[4 * q + 2];
                const double iR = rsqrt_pos(rsq(q));
                const double gs = rg * swq[q] * iR;
#pragma unroll
                for (int w = 0; w < NK; ++w) {
                    const double gd = gs * (iR * iR + 0.5 * kap[w] * kap[w]);  // / k1^2, as g above
                    T0[w].re -= gs;
                    V0[w].re -= gd;
                    Tb[w][0].re -= gs * bx; Tb[w][1].re -= gs * by; Tb[w][2].re -= gs * bz;
                    Vb[w][0].re -= gd * bx; Vb[w][1].re -= gd * by; Vb[w][2].re -= gd * bz;
                }
            }
        }
#pragma unroll
        for (int w = 0; w < NK; ++w) {
            S0[w].re += wp * T0[w].re; S0[w].im += wp * T0[w].im;
            U0[w].re += wp * V0[w].re; U0[w].im += wp * V0[w].im;
            double adr = 0, adi = 0;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                Sa[w][x].re += wp * ap[x] * T0[w].re; Sa[w][x].im += wp * ap[x] * T0[w].im;
                Ua[w][x].re += wp * ap[x] * V0[w].re; Ua[w][x].im += wp * ap[x] * V0[w].im;
                Sb[w][x].re += wp * Tb[w][x].re; Sb[w][x].im += wp * Tb[w][x].im;
                Ub[w][x].re += wp * Vb[w][x].re; Ub[w][x].im += wp * Vb[w][x].im;
                adr += ap[x] * Tb[w][x].re; adi += ap[x] * Tb[w][x].im;
            }
            Sab[w].re += wp * adr; Sab[w].im += wp * adi;
            // a_p x Vb
            Uab[w][0].re += wp * (ap[1] * Vb[w][2].re - ap[2] * Vb[w][1].re);
            Uab[w][0].im += wp * (ap[1] * Vb[w][2].im - ap[2] * Vb[w][1].im);
            Uab[w][1].re += wp * (ap[2] * Vb[w][0].re - ap[0] * Vb[w][2].re);
            Uab[w][1].im += wp * (ap[2] * Vb[w][0].im - ap[0] * Vb[w][2].im);
            Uab[w][2].re += wp * (ap[0] * Vb[w][1].re - ap[1] * Vb[w][0].re);
            Uab[w][2].im += wp * (ap[0] * Vb[w][1].im - ap[1] * Vb[w][0].im);
        }
    }
    if (!active) return;
    // back to lengths in units of 1: the b moments were taken over k1 b_q, the
    // g moments over g / k1^2
    {
        const double ik1 = 1.0 / k1, k1s = k1 * k1;
#pragma unroll
        for (int w = 0; w < NK; ++w) {
            Sab[w].re *= ik1; Sab[w].im *= ik1;
            U0[w].re *= k1s; U0[w].im *= k1s;
#pragma unroll
            for (int x = 0; x < 3; ++x) {
                Sb[w][x].re *= ik1; Sb[w][x].im *= ik1;
                Ua[w][x].re *= k1s; Ua[w][x].im *= k1s;
                Ub[w][x].re *= k1; Ub[w][x].im *= k1;
                Uab[w][x].re *= k1; Uab[w][x].im *= k1;
            }
        }
    }

    // epilogue: the 3x3 TL / TK blocks
    double lm[3], ln[3], dj[3][3], ei[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lm[a] = len[a * (size_t)nf + mm];
        ln[a] = len[a * (size_t)nf + n];
        const int fv = (a + 2) % 3;  // free vertex of local edge a
#pragma unroll
        for (int x = 0; x < 3; ++x) {
            dj[a][x] = -om[fv][x];
            ei[a][x] = -o[(3 * fv + x) * (size_t)nf + n];
        }
    }
    cplx* Wn = W + (size_t)nl * NS * nf + m;
#pragma unroll
    for (int w = 0; w < NK; ++w) {
        const double ik2 = 1.0 / (kk[w] * kk[w]);
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double de = dj[j][0] * ei[i][0] + dj[j][1] * ei[i][1] + dj[j][2] * ei[i][2];
                cplx tl;
                tl.re = Sab[w].re + Sa[w][0].re * ei[i][0] + Sa[w][1].re * ei[i][1] + Sa[w][2].re * ei[i][2] +
                        dj[j][0] * Sb[w][0].re + dj[j][1] * Sb[w][1].re + dj[j][2] * Sb[w][2].re + de * S0[w].re;
                tl.im = Sab[w].im + Sa[w][0].im * ei[i][0] + Sa[w][1].im * ei[i][1] + Sa[w][2].im * ei[i][2] +
                        dj[j][0] * Sb[w][0].im + dj[j][1] * Sb[w][1].im + dj[j][2] * Sb[w][2].im + de * S0[w].im;
                const double sc = lm[j] * ln[i];
                tl.re = (0.25 * tl.re - ik2 * S0[w].re) * sc;
                tl.im = (0.25 * tl.im - ik2 * S0[w].im) * sc;
                // D = v_j - v_i = dlt - d_j + e_i
                const double D[3] = {dlt[0] - dj[j][0] + ei[i][0], dlt[1] - dj[j][1] + ei[i][1],
                                     dlt[2] - dj[j][2] + ei[i][2]};
                // X = Uab + Ua x e_i + d_j x Ub + U0 (d_j x e_i)
                const double dxe[3] = {dj[j][1] * ei[i][2] - dj[j][2] * ei[i][1],
                                       dj[j][2] * ei[i][0] - dj[j][0] * ei[i][2],
                                       dj[j][0] * ei[i][1] - dj[j][1] * ei[i][0]};
                cplx tk = {0, 0};
#pragma unroll
                for (int x = 0; x < 3; ++x) {
                    const int y = (x + 1) % 3, z = (x + 2) % 3;
                    const double Xr = Uab[w][x].re + (Ua[w][y].re * ei[i][z] - Ua[w][z].re * ei[i][y]) +
                                      (dj[j][y] * Ub[w][z].re - dj[j][z] * Ub[w][y].re) + U0[w].re * dxe[x];
                    const double Xi = Uab[w][x].im + (Ua[w][y].im * ei[i][z] - Ua[w][z].im * ei[i][y]) +
                                      (dj[j][y] * Ub[w][z].im - dj[j][z] * Ub[w][y].im) + U0[w].im * dxe[x];
                    tk.re += Xr * D[x];
                    tk.im += Xi * D[x];
                }
                tk.re *= 0.25 * sc;
                tk.im *= 0.25 * sc;
                Wn[(size_t)((w * 2 + 0) * 9 + j * 3 + i) * nf] = tl;
                Wn[(size_t)((w * 2 + 1) * 9 + j * 3 + i) * nf] = tk;
                if (SYM) {
                    cplx* Wt = W + (size_t)m * NS * nf + n;
                    Wt[(size_t)((w * 2 + 0) * 9 + i * 3 + j) * nf] = tl;
                    Wt[(size_t)((w * 2 + 1) * 9 + i * 3 + j) * nf] = tk;
                }
            }
    }
}

// ------------------------------------------------------------------ step 3
// Device restatement of the reference near-pair path (direct form).
struct V3 { double x, y, z; };
__device__ __forceinline__ V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 vscl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 vcross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 vunit(V3 a) { return vscl(a, 1.0 / sqrt(vdot(a, a))); }

struct Wil { double phi0; V3 phi1b; double phi2; V3 phi3b; V3 rpi; };

// calculateWiltonGeometricTerms + calculateAllWiltonPotentials
// (semi_analytic_integrals.cuh:35-92, 563-667); Wilton et al. 1984.
__device__ Wil wilton(V3 r, const V3 t[3], V3 n) {
    Wil o;
    const double d = vdot(vsub(r, t[0]), n);
    const V3 rpi = vsub(r, vscl(n, d));
    o.rpi = rpi;
    double phi0 = 0, phi2 = 0;
    V3 phi1 = {0, 0, 0}, phi3 = {0, 0, 0};
    const double ad = fabs(d);
    const double tiny = (double)1e-9f;
    for (int e = 0; e < 3; ++e) {
        const V3 vs = t[e], ve = t[(e + 1) % 3];
        const V3 th = vunit(vsub(ve, vs));
        const V3 u = vcross(th, n);
        const V3 rs = vsub(vs, vscl(n, vdot(n, vs)));
        const V3 re = vsub(ve, vscl(n, vdot(n, ve)));
        const V3 ts = vsub(rs, rpi);
        const double tu = vdot(ts, u);
        const double p0 = fabs(tu);
        const double orient = tu >= 0.0 ? 1.0 : -1.0;
        const double lp = vdot(vsub(re, rpi), th), lmn = vdot(vsub(rs, rpi), th);
        const double Pp = hypot(p0, lp), Pm = hypot(p0, lmn);
        const double R0 = hypot(p0, ad), Rp = hypot(Pp, ad), Rmn = hypot(Pm, ad);
        const double lt = (log(Rp) + log1p(lp / Rp)) - (log(Rmn) + log1p(lmn / Rmn));
        // shared arctangent difference of phi_0 (|d| > 1e-9f) and phi_2 (|d| > 1e-6)
        double atn = 0.0;
        if (ad > tiny) atn = atan2(p0 * lp, R0 * R0 + ad * Rp) - atan2(p0 * lmn, R0 * R0 + ad * Rmn);
        if (p0 > tiny) phi0 += orient * (p0 * lt - ad * atn);
        const double st = (R0 * R0) * lt + (lp * Rp) - (lmn * Rmn);
        phi1 = vadd(phi1, vscl(u, st));
        double term = 0.0;
        if (ad > 1e-6) term = (1.0 / ad) * atn;
        else if (p0 > 1e-9) term = -((lp / (p0 * Pp)) - (lmn / (p0 * Pm)));
        phi2 += orient * term;
        phi3 = vsub(phi3, vscl(u, lt));
    }
    o.phi0 = phi0;
    o.phi1b = vscl(phi1, 0.5);
    o.phi2 = phi2;
    o.phi3b = phi3;
    return o;
}

// One wavenumber: held to two waves per SIMD (256 registers, a few spilled;
// it needs 300 otherwise, one wave): 22 against 29 ms at 72k.
// Regular part of the self pairs (m, m) of the chunk's source faces: the
// regularised double quadrature (integrandKernelsRegular) in the reference's
// direct form, with its limit -jk/(4 pi) where the two points coincide; one
// workgroup per face, its threads over the Q x Q point pairs, then a
// fixed-order reduction (wave shuffles, then the four waves in order). It
// stores TL (TK vanishes on a flat face); near_pairs_kernel then adds the
// singular part. (One lane per self pair, as the reference loops, left 63 of
// 64 lanes idle through Q^2 point pairs.)
template <int NK>
__global__ __launch_bounds__(256) void self_pairs_kernel(int nf, int n0, int Q, const double* __restrict__ V,
                                                         const int* __restrict__ F, const double* __restrict__ len,
                                                         const float* __restrict__ qxy, const float* __restrict__ qw,
                                                         double k1, double k2, cplx* __restrict__ W) {
    constexpr int NS = NK * 18, NV = NK * 18;
    __shared__ double part[4][NV];
    const int m = n0 + blockIdx.x, tid = threadIdx.x;
    V3 tt[3], tf[3];
    for (int a = 0; a < 3; ++a) {
        const int v = F[3 * m + a];
        tt[a] = {V[3 * v], V[3 * v + 1], V[3 * v + 2]};
    }
    for (int a = 0; a < 3; ++a) tf[a] = tt[(a + 2) % 3];
    const double kk[2] = {k1, k2};
    double acc[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] = 0.0;
    for (int pq = tid; pq < Q * Q; pq += 256) {
        const int p = pq / Q, q = pq - p * Q;
        const double xp = qxy[2 * p], yp = qxy[2 * p + 1], wp = qw[p];
        const double zp = 1.0 - xp - yp;
        const V3 rp = {tt[0].x * zp + tt[1].x * xp + tt[2].x * yp, tt[0].y * zp + tt[1].y * xp + tt[2].y * yp,
                       tt[0].z * zp + tt[1].z * xp + tt[2].z * yp};
        const double xq = qxy[2 * q], yq = qxy[2 * q + 1], wq = qw[q];
        const double zq = 1.0 - xq - yq;
        const V3 rq = {tt[0].x * zq + tt[1].x * xq + tt[2].x * yq, tt[0].y * zq + tt[1].y * xq + tt[2].y * yq,
                       tt[0].z * zq + tt[1].z * xq + tt[2].z * yq};
        const V3 R = vsub(rp, rq);
        const double Rm = sqrt(vdot(R, R));
        const double cw = wp * wq;
#pragma unroll
        for (int w = 0; w < NK; ++w) {
            cplx G;
            if (Rm < 1e-9) {
                G = {0.0, -kk[w] * kInv4Pi};
            } else {
                const double iR = 1.0 / Rm;
                double sn, co;
                sincos_moderate(kk[w] * Rm, &sn, &co);
                G = {(co - 1.0) * kInv4Pi * iR, -sn * kInv4Pi * iR};
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const V3 rm = vsub(rp, tf[j]);
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const V3 rn = vsub(rq, tf[i]);
                    const double md = 0.25 * vdot(rm, rn) - 1.0 / (kk[w] * kk[w]);
                    acc[w * 18 + (j * 3 + i) * 2] += G.re * md * cw;
                    acc[w * 18 + (j * 3 + i) * 2 + 1] += G.im * md * cw;
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < NV; ++v)
        for (int off = 32; off > 0; off >>= 1) acc[v] += __shfl_xor(acc[v], off);
    if ((tid & 63) == 0)
        for (int v = 0; v < NV; ++v) part[tid >> 6][v] = acc[v];
    __syncthreads();
    if (tid < NK * 9) {
        const int w = tid / 9, ji = tid % 9, j = ji / 3, i = ji % 3;
        const int v = w * 18 + ji * 2;
        const double sc = len[j * (size_t)nf + m] * len[i * (size_t)nf + m];
        const double re = ((part[0][v] + part[1][v]) + (part[2][v] + part[3][v])) * sc;
        const double im = ((part[0][v + 1] + part[1][v + 1]) + (part[2][v + 1] + part[3][v + 1])) * sc;
        cplx* Wm = W + (size_t)(m - n0) * NS * nf + m;
        Wm[(size_t)((w * 2 + 0) * 9 + ji) * nf] = {re, im};
        Wm[(size_t)((w * 2 + 1) * 9 + ji) * nf] = {0.0, 0.0};
    }
}

template <int NK>
__global__ __launch_bounds__(128, NK == 1 ? 2 : 1) void near_pairs_kernel(
    int npairs, const int2* __restrict__ list, int nf, int n0, int Q, const double* __restrict__ V,
    const int* __restrict__ F, const double* __restrict__ len, const double* __restrict__ area,
    const float* __restrict__ qxy, const float* __restrict__ qw, double k1, double k2,
    cplx* __restrict__ W) {
    constexpr int NS = NK * 18;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npairs) return;
    const int m = list[t].x, n = list[t].y;
    V3 tt[3], st[3];
    for (int a = 0; a < 3; ++a) {
        const int vm = F[3 * m + a], vn = F[3 * n + a];
        tt[a] = {V[3 * vm], V[3 * vm + 1], V[3 * vm + 2]};
        st[a] = {V[3 * vn], V[3 * vn + 1], V[3 * vn + 2]};
    }
    const bool self = (m == n);
    double lm[3], ln[3];
    V3 tf[3], sf[3];
    for (int a = 0; a < 3; ++a) {
        lm[a] = len[a * (size_t)nf + m];
        ln[a] = len[a * (size_t)nf + n];
        tf[a] = tt[(a + 2) % 3];
        sf[a] = st[(a + 2) % 3];
    }
    const double kk[2] = {k1, k2};

    // singular part: Wilton inner integrals, Dunavant outer rule
    const V3 nn = vunit(vcross(vsub(st[1], st[0]), vsub(st[2], st[0])));
    double stl[NK][3][3], stk[NK][3][3];
    for (int w = 0; w < NK; ++w)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) stl[w][j][i] = stk[w][j][i] = 0.0;
    for (int p = 0; p < Q; ++p) {
        const double xp = qxy[2 * p], yp = qxy[2 * p + 1], wp = qw[p];
        const double zp = 1.0 - xp - yp;
        const V3 rp = {tt[0].x * zp + tt[1].x * xp + tt[2].x * yp, tt[0].y * zp + tt[1].y * xp + tt[2].y * yp,
                       tt[0].z * zp + tt[1].z * xp + tt[2].z * yp};
        const Wil P = wilton(rp, st, nn);
        for (int i = 0; i < 3; ++i) {
            const V3 fv = sf[i];
            // Psi_1 = phi1_base + phi0 r_pi - phi0 free
            const V3 ps = vsub(vadd(P.phi1b, vscl(P.rpi, P.phi0)), vscl(fv, P.phi0));
            const V3 phi3 = vadd(P.phi3b, vscl(vsub(P.rpi, fv), P.phi2));
            const V3 Ri = vsub(rp, fv);
            for (int j = 0; j < 3; ++j) {
                const V3 rm = vsub(rp, tf[j]);
                const double db = 0.25 * vdot(rm, ps);
                for (int w = 0; w < NK; ++w) {
                    const float kf = (float)kk[w];
                    const float k2o2 = kf * kf * 0.5f;
                    const V3 comb = vadd(phi3, vscl(ps, (double)k2o2));
                    const V3 ups = vscl(vcross(Ri, comb), -1.0);
                    stl[w][j][i] += (db - 1.0 / (kk[w] * kk[w]) * P.phi0) * wp;
                    if (!self) stk[w][j][i] += 0.25 * vdot(rm, ups) * wp;
                }
            }
        }
    }
    const double asc = 1.0 / (4.0 * kPi * area[n]);
    cplx* Wn = W + (size_t)(n - n0) * NS * nf + m;
    for (int w = 0; w < NK; ++w)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                const double s = asc * lm[j] * ln[i];
                cplx* pa = &Wn[(size_t)((w * 2 + 0) * 9 + j * 3 + i) * nf];
                cplx* pb = &Wn[(size_t)((w * 2 + 1) * 9 + j * 3 + i) * nf];
                cplx a = *pa, b = *pb;
                a.re += stl[w][j][i] * s;
                b.re += stk[w][j][i] * s;
                *pa = a;
                *pb = b;
            }
}

// ------------------------------------------------------------------ step 4
struct Epi {
    // GSTC: eta0-scaled Calderon factors (sie_scaling.cuh:48-77)
    cplx jj[2], mm[2];
    double jm;
    // PEC: c = j w mu2 ; PMCHWT: c_mu1, c_mu2, c_eps1, c_eps2, eta0 (float)
    cplx c[4];
    double eta;
};

enum Kind { K_GSTC = 0, K_PEC = 1, K_PMCHWT = 2, K_RAW = 3 };

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ void zacc(cplx* Z, long long ldz, long long r, long long c, cplx v) {
    cplx* p = Z + c * ldz + r;
    cplx z = *p;
    z.re += v.re; z.im += v.im;
    *p = z;
}

// --------------------------------------------------------- surface fields
// E and H at observation points from physical RWG currents J, M (one complex
// per edge): evaluate_surface_integral_fields (surface_current_fields.cu:170-299)
// through compute_scattered_fields_kernel (sie_kernels_utils.cuh:21-120):
//   E = jw mu sum J L + sum M K,   H = -sum J K + jw eps sum M L,
// with L_i = l_i sum_q w_q (rho_i G / 2 - grad'G / k^2) and
// K_i = l_i sum_q w_q rho_i x grad'G / 2 (calculateScatterFieldQuadrature,
// dunavant_quadrature.cuh:1634-1712); faces whose float centroid distance is
// below 0.1 lambda (float) use the regularised Green's function plus the
// Wilton singular part (calculateScatterFieldSingular, :1719-1766). One
// workgroup per observation point, its lanes over faces, then a fixed-order
// tree reduction (deterministic).
constexpr int kFieldThreads = 256;

__global__ __launch_bounds__(kFieldThreads) void surface_fields_kernel(
    int nf, int Q, const double* __restrict__ Vx, const int* __restrict__ Fx, const int* __restrict__ edge,
    const double* __restrict__ c, const double* __restrict__ len, const float* __restrict__ qxy,
    const float* __restrict__ qw, const cplx* __restrict__ J, const cplx* __restrict__ M,
    const double* __restrict__ obs, double omega_mu, double omega_eps, double k, cplx* __restrict__ E,
    cplx* __restrict__ H) {
    const int o = blockIdx.x;
    const V3 r = {obs[3 * o], obs[3 * o + 1], obs[3 * o + 2]};
    const float kf = (float)k;
    const float ksq = 1.f / (kf * kf);
    const float lambda = (float)(2.0f * (float)kPi / k);
    const float thr = 0.1f * lambda;
    const double kd = (double)kf;
    cplx e[3] = {{0, 0}, {0, 0}, {0, 0}}, h[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int f = threadIdx.x; f < nf; f += kFieldThreads) {
        V3 t[3];
        for (int a = 0; a < 3; ++a) {
            const int v = Fx[3 * f + a];
            t[a] = {Vx[3 * v], Vx[3 * v + 1], Vx[3 * v + 2]};
        }
        const V3 d = {r.x - c[f], r.y - c[nf + f], r.z - c[2 * nf + f]};
        const bool near = (float)sqrt(vdot(d, d)) < thr;
        V3 fv[3];
        float el[3];
        for (int i = 0; i < 3; ++i) {
            fv[i] = t[(i + 2) % 3];
            el[i] = (float)len[i * (size_t)nf + f];
        }
        cplx L[3][3], K[3][3];
        for (int i = 0; i < 3; ++i)
            for (int x = 0; x < 3; ++x) L[i][x] = K[i][x] = {0, 0};
        for (int q = 0; q < Q; ++q) {
            const double xi = qxy[2 * q], eta = qxy[2 * q + 1], wq = qw[q];
            const double ze = 1.0 - xi - eta;
            const V3 rq = {t[0].x * ze + t[1].x * xi + t[2].x * eta, t[0].y * ze + t[1].y * xi + t[2].y * eta,
                           t[0].z * ze + t[1].z * xi + t[2].z * eta};
            const V3 R = vsub(r, rq);
            const double Rm = sqrt(vdot(R, R));
            cplx G = {0, 0}, g[3] = {{0, 0}, {0, 0}, {0, 0}};
            double sn, cs;
            sincos_moderate(kd * Rm, &sn, &cs);
            if (!near) {  // greenAndGradSourceGreenFunction (green_functions_utils.cuh:142-175)
                if (Rm >= 1e-12) {
                    const double iR = 1.0 / Rm, c4 = 1.0 / (4.0 * kPi * Rm);
                    G = {cs * c4, -sn * c4};
                    const cplx wg = cmul({iR, kd}, G);
                    const double Rh[3] = {R.x * iR, R.y * iR, R.z * iR};
                    for (int x = 0; x < 3; ++x) g[x] = {wg.re * Rh[x], wg.im * Rh[x]};
                }
            } else if (Rm < 1e-6) {  // ...Regular (:177-222), R -> 0
                G = {0.0, -kd / (4.0 * kPi)};
            } else {
                const double i4 = 1.0 / (4.0 * kPi), iR = 1.0 / Rm, iR2 = iR * iR, iR3 = iR2 * iR, i4R = i4 * iR;
                G = {(cs - 1.0) * i4R, -sn * i4R};
                cplx Kr = cmul({iR3, kd * iR2}, {cs, -sn});
                Kr.re += -(iR3 + kd * kd * 0.5 * iR);
                const cplx sc = {Kr.re * i4, Kr.im * i4};
                const double Rv[3] = {R.x, R.y, R.z};
                for (int x = 0; x < 3; ++x) g[x] = {sc.re * Rv[x], sc.im * Rv[x]};
            }
            for (int i = 0; i < 3; ++i) {
                const double rv[3] = {rq.x - fv[i].x, rq.y - fv[i].y, rq.z - fv[i].z};
                for (int x = 0; x < 3; ++x) {
                    const cplx t1 = {G.re * rv[x] * 0.5, G.im * rv[x] * 0.5};
                    const cplx t2 = {g[x].re * -(double)ksq, g[x].im * -(double)ksq};
                    L[i][x].re += (t1.re + t2.re) * wq;
                    L[i][x].im += (t1.im + t2.im) * wq;
                }
                for (int x = 0; x < 3; ++x) {
                    const int y = (x + 1) % 3, z = (x + 2) % 3;
                    const cplx cr = {g[z].re * rv[y] - g[y].re * rv[z], g[z].im * rv[y] - g[y].im * rv[z]};
                    K[i][x].re += cr.re * 0.5 * wq;
                    K[i][x].im += cr.im * 0.5 * wq;
                }
            }
        }
        for (int i = 0; i < 3; ++i)
            for (int x = 0; x < 3; ++x) {
                L[i][x].re *= el[i]; L[i][x].im *= el[i];
                K[i][x].re *= el[i]; K[i][x].im *= el[i];
            }
        if (near) {
            const V3 cr = vcross(vsub(t[1], t[0]), vsub(t[2], t[0]));
            const float A = (float)(0.5f * sqrt(vdot(cr, cr)));
            const float base = 1.0f / (8.0f * (float)kPi * A);
            const Wil P = wilton(r, t, vunit(cr));
            const float k2o2 = kf * kf * 0.5f;
            for (int i = 0; i < 3; ++i) {
                const V3 Rf = vsub(r, fv[i]);
                // Psi_1 = phi1_base + phi0 r_pi - phi0 free; phi_3 = phi3_base + (r_pi - free) phi2
                const V3 ps = vsub(vadd(P.phi1b, vscl(P.rpi, P.phi0)), vscl(fv[i], P.phi0));
                const V3 ph3 = vadd(P.phi3b, vscl(vsub(P.rpi, fv[i]), P.phi2));
                const V3 Li = vadd(vsub(ps, vscl(Rf, 0.5 * P.phi0)), vscl(vsub(ph3, vscl(Rf, P.phi2)), (double)ksq));
                // calculate_Upsilon_coul: -(r - free) x (phi_3 + k^2/2 Psi_1)
                const V3 Ki = vscl(vcross(Rf, vadd(ph3, vscl(ps, (double)k2o2))), -1.0);
                const double ls = (double)(el[i] * base);
                L[i][0].re += 2.0 * Li.x * ls; L[i][1].re += 2.0 * Li.y * ls; L[i][2].re += 2.0 * Li.z * ls;
                K[i][0].re += Ki.x * ls; K[i][1].re += Ki.y * ls; K[i][2].re += Ki.z * ls;
            }
        }
        for (int i = 0; i < 3; ++i) {
            const int ed = edge[i * (size_t)nf + f];
            const cplx Je = J[ed], Me = M[ed];
            const cplx wEJ = {-omega_mu * Je.im, omega_mu * Je.re};    // j w mu J
            const cplx wHM = {-omega_eps * Me.im, omega_eps * Me.re};  // j w eps M
            for (int x = 0; x < 3; ++x) {
                const cplx a = cmul(wEJ, L[i][x]), b = cmul(Me, K[i][x]);
                const cplx hc = cmul(Je, K[i][x]), hd = cmul(wHM, L[i][x]);
                e[x].re += a.re + b.re; e[x].im += a.im + b.im;
                h[x].re += hd.re - hc.re; h[x].im += hd.im - hc.im;
            }
        }
    }
    __shared__ double red[12][kFieldThreads];
    for (int x = 0; x < 3; ++x) {
        red[2 * x][threadIdx.x] = e[x].re; red[2 * x + 1][threadIdx.x] = e[x].im;
        red[6 + 2 * x][threadIdx.x] = h[x].re; red[7 + 2 * x][threadIdx.x] = h[x].im;
    }
    __syncthreads();
    for (int st = kFieldThreads / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int v = 0; v < 12; ++v) red[v][threadIdx.x] += red[v][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        E[3 * o + threadIdx.x] = {red[2 * threadIdx.x][0], red[2 * threadIdx.x + 1][0]};
        H[3 * o + threadIdx.x] = {red[6 + 2 * threadIdx.x][0], red[7 + 2 * threadIdx.x][0]};
    }
}

// One lane per (test edge em, listed source edge en).
template <int NK, int KIND>
__global__ __launch_bounds__(256) void gather_kernel(
    int nf, int ne, int n0, int nc, const int* __restrict__ elist, int nlist,
    const int* __restrict__ ef, const int* __restrict__ el, const cplx* __restrict__ W,
    const cplx* __restrict__ M /* nf x 144, GSTC only */, Epi ep, cplx* __restrict__ Z,
    long long ldz, int col_stride, int overwrite, int wave_pass = -1) {
    // col_stride 0: columns of the full matrix (edge id + block E); > 0: a slab
    // whose block c holds the chunk's edges in list order at c * col_stride.
    // overwrite: the only chunk (every face pair present), so every entry of
    // the lane's blocks is stored, zeros included, and Z need not be cleared
    // first; otherwise the chunk's contributions are added.
    // wave_pass (T-PMCHWT, NK = 1 blocks): the blocks of wavenumber k_(pass+1)
    // only, so that Z is the sum of one pass per region's wavenumber.
    constexpr int NS = NK * 18;
    const int em = blockIdx.x * blockDim.x + threadIdx.x;
    const int li = uniform(blockIdx.y);
    if (em >= ne || li >= nlist) return;
    const int en = elist[li];
    cplx acc[NK * 2];
#pragma unroll
    for (int s = 0; s < NK * 2; ++s) acc[s] = {0, 0};
    const int fa0 = ef[em], fa1 = ef[ne + em], ja0 = el[em], ja1 = el[ne + em];
    int shared_g = -1, shared_j = 0, shared_i = 0;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int g = ef[b * ne + en], i = el[b * ne + en];
        if (g < n0 || g >= n0 + nc) continue;
        const cplx* Wg = W + (size_t)(g - n0) * NS * nf;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int f = a ? fa1 : fa0, j = a ? ja1 : ja0;
#pragma unroll
            for (int s = 0; s < NK * 2; ++s) {
                const cplx v = Wg[(size_t)(s * 9 + j * 3 + i) * nf + f];
                acc[s].re += v.re; acc[s].im += v.im;
            }
            if (f == g) { shared_g = g; shared_j = j; shared_i = i; }
        }
    }
    const long long E = ne;
    auto col = [&](int c) -> long long { return col_stride > 0 ? (long long)li + (long long)c * col_stride : en + c * E; };
    auto zput = [&](long long r, long long c, cplx v) {
        if (overwrite) Z[c * ldz + r] = v;
        else zacc(Z, ldz, r, c, v);
    };
    if (KIND == K_GSTC) {
        const cplx TL1 = acc[0], TK1 = acc[1];
        const cplx TL2 = NK == 2 ? acc[2] : acc[0], TK2 = NK == 2 ? acc[3] : acc[1];
        cplx blk[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) blk[r][c] = {0, 0};
        blk[0][0] = cmul(ep.jj[0], TL1);
        blk[0][1] = {TK1.re * ep.jm, TK1.im * ep.jm};
        blk[1][0] = {-blk[0][1].re, -blk[0][1].im};
        blk[1][1] = cmul(ep.mm[0], TL1);
        blk[2][2] = cmul(ep.jj[1], TL2);
        blk[2][3] = {TK2.re * ep.jm, TK2.im * ep.jm};
        blk[3][2] = {-blk[2][3].re, -blk[2][3].im};
        blk[3][3] = cmul(ep.mm[1], TL2);
        if (shared_g >= 0 && M) {
            const cplx* Mg = M + (size_t)shared_g * 144;
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const cplx v = Mg[(r * 3 + shared_j) * 12 + c * 3 + shared_i];
                    blk[r][c].re += v.re; blk[r][c].im += v.im;
                }
            // an edge paired with itself shares both of its faces
            if (em == en) {
                const int g2 = (shared_g == fa0) ? fa1 : fa0;
                const int j2 = (shared_g == fa0) ? ja1 : ja0;
                if (g2 >= n0 && g2 < n0 + nc && g2 != shared_g) {
                    const cplx* M2 = M + (size_t)g2 * 144;
#pragma unroll
                    for (int r = 0; r < 4; ++r)
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const cplx v = M2[(r * 3 + j2) * 12 + c * 3 + j2];
                            blk[r][c].re += v.re; blk[r][c].im += v.im;
                        }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const bool cal = (r >> 1) == (c >> 1);
                if (cal || shared_g >= 0 || overwrite) zput(em + r * E, col(c), blk[r][c]);
            }
    } else if (KIND == K_PEC) {
        zput(em, col(0), cmul(ep.c[0], NK == 2 ? acc[2] : acc[0]));
    } else if (KIND == K_PMCHWT && NK == 1 && wave_pass >= 0) {
        // one region's terms: jw mu_p TL_p, eta0 TK_p, jw eps_p TL_p (eta0^2 scaled)
        const cplx TL = acc[0], TK = acc[1];
        const cplx zjj = cmul(ep.c[wave_pass], TL), zmm = cmul(ep.c[2 + wave_pass], TL);
        const double et = ep.eta, et2 = ep.jm;
        zput(em, col(0), zjj);
        zput(em, col(1), {TK.re * et, TK.im * et});
        zput(em + E, col(0), {-TK.re * et, -TK.im * et});
        zput(em + E, col(1), {zmm.re * et2, zmm.im * et2});
    } else if (KIND == K_PMCHWT) {
        const cplx TL1 = acc[0], TK1 = acc[1];
        const cplx TL2 = NK == 2 ? acc[2] : acc[0], TK2 = NK == 2 ? acc[3] : acc[1];
        const cplx tk = {TK1.re + TK2.re, TK1.im + TK2.im};
        const cplx a = cmul(ep.c[0], TL1), b = cmul(ep.c[1], TL2);
        const cplx zjj = {a.re + b.re, a.im + b.im};
        const cplx d = cmul(ep.c[2], TL1), e = cmul(ep.c[3], TL2);
        const cplx zmm = {d.re + e.re, d.im + e.im};
        const double et = ep.eta, et2 = ep.jm;  // eta0 and (float) eta0^2
        zput(em, col(0), zjj);
        zput(em, col(1), {tk.re * et, tk.im * et});
        zput(em + E, col(0), {-tk.re * et, -tk.im * et});
        zput(em + E, col(1), {zmm.re * et2, zmm.im * et2});
    }
}

// Raw per-pair blocks (validation): W chunk -> out[(m*nf+n)*36 complex].
template <int NK>
__global__ void raw_blocks_kernel(int nf, int n0, int nc, const cplx* __restrict__ W, cplx* __restrict__ out) {
    constexpr int NS = NK * 18;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nf * nc) return;
    const int nl = (int)(t / nf), m = (int)(t - (long long)nl * nf);
    const int n = n0 + nl;
    cplx* o = out + ((size_t)m * nf + n) * 36;
    // order TL(k1), TL(k2), TK(k1), TK(k2)
    for (int op = 0; op < 2; ++op)
        for (int w = 0; w < 2; ++w) {
            const int ws = NK == 2 ? w : 0;
            for (int ji = 0; ji < 9; ++ji)
                o[(op * 2 + w) * 9 + ji] = W[((size_t)nl * NS + (ws * 2 + op) * 9 + ji) * nf + m];
        }
}

// ------------------------------------------------------------- GSTC mass
// One lane per face: frames, 2x2 -> 3x3 lift, perp transform and the batched
// 8-tensor Gram quadrature with the eta0-scaled 4x4 block pattern
// (sie_kernels_assembly.cuh:460-594, dunavant_quadrature.cuh:1422-1508,
// gstc_tensor_utils.cuh:72-309). Output M[f][(r*3+j)*12 + c*3+i].
__global__ __launch_bounds__(64) void mass_kernel(
    int nf, int Q, const double* __restrict__ V, const int* __restrict__ F,
    const double* __restrict__ len, const double* __restrict__ area, const float* __restrict__ qxy,
    const float* __restrict__ qw, const cplx* __restrict__ tensors, int uniform_t, int basis,
    float mu0, float eps0, int scaled, cplx* __restrict__ M) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    V3 t[3];
    for (int a = 0; a < 3; ++a) {
        const int v = F[3 * f + a];
        t[a] = {V[3 * v], V[3 * v + 1], V[3 * v + 2]};
    }
    // local frame
    V3 u, w;
    if (basis == 0) {
        const V3 e1 = vsub(t[1], t[0]), e2 = vsub(t[2], t[0]);
        const V3 nu = vcross(e1, e2);
        const double l1 = sqrt(vdot(e1, e1)), lnu = sqrt(vdot(nu, nu));
        if (l1 < 1e-20 || lnu < 1e-20) { u = {1, 0, 0}; w = {0, 1, 0}; }
        else {
            const V3 nh = vscl(nu, 1.0 / lnu);
            u = vscl(e1, 1.0 / l1);
            w = vcross(nh, u);
        }
    } else {
        const V3 c = {(t[0].x + t[1].x + t[2].x) / 3.0, (t[0].y + t[1].y + t[2].y) / 3.0,
                      (t[0].z + t[1].z + t[2].z) / 3.0};
        const double r = sqrt(vdot(c, c));
        if (r < 1e-12) { u = {1, 0, 0}; w = {0, 1, 0}; }
        else {
            double ct = c.z / r;
            ct = ct > 1.0 ? 1.0 : (ct < -1.0 ? -1.0 : ct);
            const double th = acos(ct), ph = atan2(c.y, c.x);
            const double cth = cos(th), sth = sin(th), cph = cos(ph), sph = sin(ph);
            u = {cth * cph, cth * sph, -sth};
            w = {-sph, cph, 0.0};
        }
    }
    const cplx* tp = tensors + (uniform_t ? 0 : (size_t)f * 32);
    const double U[3] = {u.x, u.y, u.z}, Wv[3] = {w.x, w.y, w.z};
    const V3 nv = vunit(vcross(vsub(t[1], t[0]), vsub(t[2], t[0])));
    const double Nx[3][3] = {{0.0, -nv.z, nv.y}, {nv.z, 0.0, -nv.x}, {-nv.y, nv.x, 0.0}};
    double Nf[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Nf[a][b] = (double)(float)Nx[a][b];
    // the 8 global tensors: 0..3 Y_perp (ee, em, me, mm), 4..7 Z (ee, em, me, mm)
    cplx T[8][3][3];
    for (int blk = 0; blk < 4; ++blk) {
        const int r0 = (blk >> 1) * 2, c0 = (blk & 1) * 2;
        for (int src = 0; src < 2; ++src) {
            const cplx* b = tp + src * 16;
            const cplx t00 = b[r0 * 4 + c0], t01 = b[r0 * 4 + c0 + 1], t10 = b[(r0 + 1) * 4 + c0],
                       t11 = b[(r0 + 1) * 4 + c0 + 1];
            cplx G[3][3];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) {
                    cplx s = {0, 0};
                    s.re += t00.re * (U[i] * U[j]); s.im += t00.im * (U[i] * U[j]);
                    s.re += t01.re * (U[i] * Wv[j]); s.im += t01.im * (U[i] * Wv[j]);
                    s.re += t10.re * (Wv[i] * U[j]); s.im += t10.im * (Wv[i] * U[j]);
                    s.re += t11.re * (Wv[i] * Wv[j]); s.im += t11.im * (Wv[i] * Wv[j]);
                    G[i][j] = s;
                }
            if (src == 1) {
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) T[4 + blk][i][j] = G[i][j];
            } else {
                cplx tmp[3][3];
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        cplx s = {0, 0};
                        for (int k = 0; k < 3; ++k) { s.re += G[r][k].re * Nf[k][c]; s.im += G[r][k].im * Nf[k][c]; }
                        tmp[r][c] = s;
                    }
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c) {
                        cplx s = {0, 0};
                        for (int k = 0; k < 3; ++k) { s.re += tmp[k][c].re * Nf[r][k]; s.im += tmp[k][c].im * Nf[r][k]; }
                        T[blk][r][c] = {-s.re, -s.im};
                    }
            }
        }
    }
    // Gram quadrature with float area / lengths / weights
    V3 fv[3];
    float el[3];
    for (int i = 0; i < 3; ++i) {
        fv[i] = t[(i + 2) % 3];
        el[i] = (float)len[i * (size_t)nf + f];
    }
    cplx R[8][3][3];
    for (int k = 0; k < 8; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) R[k][j][i] = {0, 0};
    for (int p = 0; p < Q; ++p) {
        const double xi = qxy[2 * p], eta = qxy[2 * p + 1], ze = 1.0 - xi - eta;
        const double wp = qw[p];
        const V3 rp = {t[0].x * ze + t[1].x * xi + t[2].x * eta, t[0].y * ze + t[1].y * xi + t[2].y * eta,
                       t[0].z * ze + t[1].z * xi + t[2].z * eta};
        V3 rho[3];
        for (int i = 0; i < 3; ++i) rho[i] = vsub(rp, fv[i]);
        for (int k = 0; k < 8; ++k)
            for (int i = 0; i < 3; ++i) {
                // T rho_i
                cplx tr[3];
                for (int r = 0; r < 3; ++r) {
                    tr[r].re = T[k][r][0].re * rho[i].x + T[k][r][1].re * rho[i].y + T[k][r][2].re * rho[i].z;
                    tr[r].im = T[k][r][0].im * rho[i].x + T[k][r][1].im * rho[i].y + T[k][r][2].im * rho[i].z;
                }
                for (int j = 0; j < 3; ++j) {
                    R[k][j][i].re += (tr[0].re * rho[j].x + tr[1].re * rho[j].y + tr[2].re * rho[j].z) * wp;
                    R[k][j][i].im += (tr[0].im * rho[j].x + tr[1].im * rho[j].y + tr[2].im * rho[j].z) * wp;
                }
            }
    }
    const float ar = (float)area[f];
    const float sc = 1.f / (4.f * ar);
    // unscaled operator (assemble_Z_mass_kernel, sie_kernels_assembly.cuh:282-452): eta0 -> 1
    const float e0 = scaled ? sqrtf(mu0 / eps0) : 1.f;
    const float e0sq = e0 * e0;
    // row/col scaling D = diag(1, eta0, 1, eta0)
    const double rs[4] = {1.0, (double)e0, 1.0, (double)e0};
    cplx* Mf = M + (size_t)f * 144;
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            const double fs = (double)(sc * el[j] * el[i]);
            cplx I[8];
            for (int k = 0; k < 8; ++k) I[k] = {R[k][j][i].re * fs, R[k][j][i].im * fs};
            // I: 0 Yee 1 Yem 2 Yme 3 Ymm 4 Zee 5 Zem 6 Zme 7 Zmm
            auto comb = [&](int zk, int yk, double sgn) {
                return cplx{0.5 * I[zk].re + sgn * 0.125 * I[yk].re, 0.5 * I[zk].im + sgn * 0.125 * I[yk].im};
            };
            cplx b[4][4];
            b[0][0] = comb(4, 3, +1); b[0][1] = comb(5, 2, -1); b[0][2] = comb(4, 3, -1); b[0][3] = comb(5, 2, +1);
            b[1][0] = comb(6, 1, -1); b[1][1] = comb(7, 0, +1); b[1][2] = comb(6, 1, +1); b[1][3] = comb(7, 0, -1);
            b[2][0] = comb(4, 3, -1); b[2][1] = comb(5, 2, +1); b[2][2] = comb(4, 3, +1); b[2][3] = comb(5, 2, -1);
            b[3][0] = comb(6, 1, +1); b[3][1] = comb(7, 0, -1); b[3][2] = comb(6, 1, -1); b[3][3] = comb(7, 0, +1);
            for (int r = 0; r < 4; ++r)
                for (int c = 0; c < 4; ++c) {
                    const bool mr = (r & 1), mc = (c & 1);
                    const double s = (mr && mc) ? (double)e0sq : (mr || mc) ? rs[1] : 1.0;
                    Mf[(r * 3 + j) * 12 + c * 3 + i] = {b[r][c].re * s, b[r][c].im * s};
                }
        }
}

// Grid-stride fill: the dispatch grid (blocks x threads) must stay below 2^32
// work-items, so large matrices are covered by a capped grid.
// xs[c L + li] = x[c E + elist[li]] for the slab matvec.
__global__ void slab_x_kernel(int L, int blocks, long long E, const int* __restrict__ elist,
                              const cplx* __restrict__ x, cplx* __restrict__ xs) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L * blocks) return;
    const int c = t / L, li = t - c * L;
    xs[t] = x[c * E + elist[li]];
}

__global__ void fill_kernel(cplx* __restrict__ Z, long long ldz, int rows, int cols) {
    const long long tot = (long long)rows * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / rows, r = t - c * rows;
        Z[c * ldz + r] = {0, 0};
    }
}

// ------------------------------------------------------------------ RHS
// Per (face, wave): RWG testing of one plane wave (calculateExcitationQuadrature,
// dunavant_quadrature.cuh:1543-1611; PlaneWave, source_functors.cuh:14-42).
struct PW { double E0[3], kh[3], k, ieta; int region; };

__global__ void rhs_face_kernel(int nf, int nw, int Q, const double* __restrict__ V,
                                const int* __restrict__ F, const double* __restrict__ len,
                                const float* __restrict__ qxy, const float* __restrict__ qw,
                                const PW* __restrict__ waves, cplx* __restrict__ out /* [nw][nf][6] */) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nf * nw) return;
    const int w = (int)(t / nf), f = (int)(t - (long long)w * nf);
    const PW pw = waves[w];
    V3 tv[3];
    for (int a = 0; a < 3; ++a) {
        const int v = F[3 * f + a];
        tv[a] = {V[3 * v], V[3 * v + 1], V[3 * v + 2]};
    }
    cplx VE[3] = {{0, 0}, {0, 0}, {0, 0}}, VH[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int p = 0; p < Q; ++p) {
        const double xi = qxy[2 * p], eta = qxy[2 * p + 1], ze = 1.0 - xi - eta;
        const double wp = qw[p];
        const V3 rp = {tv[0].x * ze + tv[1].x * xi + tv[2].x * eta, tv[0].y * ze + tv[1].y * xi + tv[2].y * eta,
                       tv[0].z * ze + tv[1].z * xi + tv[2].z * eta};
        const double kr = (pw.kh[0] * rp.x + pw.kh[1] * rp.y + pw.kh[2] * rp.z) * pw.k;
        double s, c;
        sincos(-kr, &s, &c);
        const cplx E[3] = {{c * pw.E0[0], s * pw.E0[0]}, {c * pw.E0[1], s * pw.E0[1]}, {c * pw.E0[2], s * pw.E0[2]}};
        const cplx H[3] = {{(pw.kh[1] * E[2].re - pw.kh[2] * E[1].re) * pw.ieta, (pw.kh[1] * E[2].im - pw.kh[2] * E[1].im) * pw.ieta},
                           {(pw.kh[2] * E[0].re - pw.kh[0] * E[2].re) * pw.ieta, (pw.kh[2] * E[0].im - pw.kh[0] * E[2].im) * pw.ieta},
                           {(pw.kh[0] * E[1].re - pw.kh[1] * E[0].re) * pw.ieta, (pw.kh[0] * E[1].im - pw.kh[1] * E[0].im) * pw.ieta}};
        for (int i = 0; i < 3; ++i) {
            const V3 r = vsub(rp, tv[(i + 2) % 3]);
            VE[i].re += (r.x * E[0].re + r.y * E[1].re + r.z * E[2].re) * wp;
            VE[i].im += (r.x * E[0].im + r.y * E[1].im + r.z * E[2].im) * wp;
            VH[i].re += (r.x * H[0].re + r.y * H[1].re + r.z * H[2].re) * wp;
            VH[i].im += (r.x * H[0].im + r.y * H[1].im + r.z * H[2].im) * wp;
        }
    }
    cplx* o = out + ((size_t)w * nf + f) * 6;
    for (int i = 0; i < 3; ++i) {
        const double sc = (double)(0.5f * (float)len[i * (size_t)nf + f]);
        o[i] = {VE[i].re * sc, VE[i].im * sc};
        o[3 + i] = {VH[i].re * sc, VH[i].im * sc};
    }
}

// Per (edge, wave): sum the two faces in a fixed order and place the rows.
__global__ void rhs_gather_kernel(int nf, int ne, int nw, int kind, const int* __restrict__ ef,
                                  const int* __restrict__ el, const PW* __restrict__ waves,
                                  const cplx* __restrict__ face, double eta_g, double eta_d,
                                  cplx* __restrict__ rhs, long long ldb) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)ne * nw) return;
    const int w = (int)(t / ne), e = (int)(t - (long long)w * ne);
    cplx ve = {0, 0}, vh = {0, 0};
    for (int b = 0; b < 2; ++b) {
        const int f = ef[b * ne + e], j = el[b * ne + e];
        const cplx* o = face + ((size_t)w * nf + f) * 6;
        ve.re += o[j].re; ve.im += o[j].im;
        vh.re += o[3 + j].re; vh.im += o[3 + j].im;
    }
    cplx* col = rhs + (long long)w * ldb;
    const int region = waves[w].region;
    if (kind == SIE_RHS_GSTC) {
        const long long base = region == 1 ? 0 : 2LL * ne;
        for (int r = 0; r < 4; ++r) col[(long long)r * ne + e] = {0, 0};
        col[base + e] = ve;
        col[base + ne + e] = {vh.re * eta_g, vh.im * eta_g};
    } else if (kind == SIE_RHS_PMCHWT) {
        const double sg = region == 2 ? 1.0 : -1.0;
        col[e] = {ve.re * sg, ve.im * sg};
        col[ne + e] = {vh.re * sg * eta_d, vh.im * sg * eta_d};
    } else {
        col[e] = ve;
    }
}

}  // namespace

// =================================================================== host
struct Assembler {
    Mesh* mesh;
    QuadDev quad;
    DevBuf<double> bq;  // source points in units of 1/bq_scale (source_points_kernel)
    int bq_rule = 0;
    double bq_scale = 0.0;
    DevBuf<double> sctab;  // e^{j 2 pi i / 2048}: (cos, sin) pairs for sincos_tab
    DevBuf<cplx> W;
    DevBuf<int2> near;
    DevBuf<unsigned int> count;
    DevBuf<int> elist;
    DevBuf<cplx> slab, xs;  // matrix-free products: the column slab of a chunk and its x entries
    unsigned int* h_count = nullptr;
    size_t scratch_bytes = size_t(4) << 30;  // W chunk budget

    explicit Assembler(Mesh* m) : mesh(m) {}
    ~Assembler() {
        if (h_count) (void)hipHostFree(h_count);
    }

    void prepare(int rule, hipStream_t s) {
        quad_upload(quad, rule, s);
        if (!sctab.p) {
            double h[2 * kSinCosTabN];
            sincos_table_fill(h);
            sctab.alloc(2 * kSinCosTabN);
            SIE_HIP(hipMemcpy(sctab.p, h, sizeof(h), hipMemcpyHostToDevice));
        }
        if (bq_rule != rule) {
            bq.alloc((size_t)mesh->nf * quad.n * 4);
            bq_rule = rule;
            bq_scale = 0.0;
        }
        if (!h_count) SIE_HIP(hipHostMalloc(&h_count, sizeof(unsigned int)));
        if (!count.p) count.alloc(1);
    }

    // Runs steps 1-3 for source faces [n0, n0+nc) into W.
    // count_geom = false: a second pass over the same pairs for another
    // wavenumber (its geometric work is not algorithmic work again)
    template <int NK>
    void chunk_blocks(int n0, int nc, double k1, double k2, float thr, DevBuf<double>& area, hipStream_t s,
                      bool sym = false, bool count_geom = true) {
        const int nf = mesh->nf;
        const int Q = quad.n;
        // near list
        SIE_HIP(hipMemsetAsync(count.p, 0, sizeof(unsigned int), s));
        const long long tot = (long long)nf * nc;
        unsigned int cap = (unsigned int)std::min<long long>(tot, std::max<long long>(1 << 16, tot / 16));
        if (near.n < cap) near.alloc(cap);
        near_list_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, n0, nc, mesh->d_c.p, thr, near.p,
                                                                        count.p, cap);
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipMemcpyAsync(h_count, count.p, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
        // far pairs (source points in units of 1/k1)
        if (bq_scale != k1) {
            const int tot = nf * Q;
            source_points_kernel<<<(tot + 255) / 256, 256, 0, s>>>(nf, Q, mesh->d_o.p, quad.xy.p, k1, bq.p);
            SIE_HIP(hipGetLastError());
            bq_scale = k1;
        }
        dim3 grid((nf + 63) / 64, (nc + 3) / 4);
        require(Q <= kMaxFarQ, "quadrature rule has more points than the far-pair kernel stages");
        prof().begin(PROF_FAR_PAIRS, s);
        if (sym) {
            require(n0 == 0 && nc == nf, "symmetric far pairs need every source face in one chunk");
            far_pairs_kernel<NK, true><<<grid, 256, 0, s>>>(nf, n0, nc, Q, mesh->d_c.p, mesh->d_o.p, mesh->d_len.p,
                                                            bq.p, quad.xy.p, quad.w.p, sctab.p, k1, k2, thr, W.p);
        } else {
            far_pairs_kernel<NK, false><<<grid, 256, 0, s>>>(nf, n0, nc, Q, mesh->d_c.p, mesh->d_o.p, mesh->d_len.p,
                                                             bq.p, quad.xy.p, quad.w.p, sctab.p, k1, k2, thr, W.p);
        }
        // flops of the moment-factored point-pair update (kFarPairFlops per pair and wavenumber),
        // over the pairs integrated (m > n only when symmetric)
        const double pairs = sym ? 0.5 * (double)nf * (nf - 1) : (double)nf * nc;
        prof().end(PROF_FAR_PAIRS, s, pairs * Q * Q * ((count_geom ? kFarGeomFlops : 0) + NK * kFarPairFlops),
                   (double)nf * nc * NK * 18 * 16.0);
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipStreamSynchronize(s));
        unsigned int nn = *h_count;
        if (nn > cap) {  // rare: more near pairs than the first guess
            near.alloc(nn);
            SIE_HIP(hipMemsetAsync(count.p, 0, sizeof(unsigned int), s));
            near_list_kernel<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, n0, nc, mesh->d_c.p, thr, near.p,
                                                                            count.p, nn);
            SIE_HIP(hipGetLastError());
        }
        if (nn) {
            prof().begin(PROF_NEAR_PAIRS, s);
            self_pairs_kernel<NK><<<nc, 256, 0, s>>>(nf, n0, Q, mesh->d_V.p, mesh->d_F.p, mesh->d_len.p, quad.xy.p,
                                                     quad.w.p, k1, k2, W.p);
            near_pairs_kernel<NK><<<(nn + 127) / 128, 128, 0, s>>>((int)nn, near.p, nf, n0, Q, mesh->d_V.p,
                                                                   mesh->d_F.p, mesh->d_len.p, area.p, quad.xy.p,
                                                                   quad.w.p, k1, k2, W.p);
            // no flop model (transcendental-bound); bytes: the blocks it adds to
            prof().end(PROF_NEAR_PAIRS, s, 0.0, (double)nn * NK * 18 * 16.0 * 2);
            SIE_HIP(hipGetLastError());
        }
    }

    // One chunk of every source face, integrated with the pair symmetry, when
    // its scratch (F^2 NK 18 complex) stays under the cap and fits in free
    // device memory with a margin; otherwise scratch_bytes-sized chunks. The
    // cap is kSymBytes for the scratch of every wavenumber together and
    // kSymWaveBytes for one wavenumber at a time (T-PMCHWT with two regions:
    // one symmetric pass per wavenumber); SIE_SYM_SCRATCH_BYTES, when set,
    // replaces both (tests reach the chunked form at small meshes).
    static constexpr size_t kSymBytes = size_t(64) << 30;
    static constexpr size_t kSymWaveBytes = size_t(160) << 30;
    static size_t sym_cap(bool per_wave) {
        const char* e = std::getenv("SIE_SYM_SCRATCH_BYTES");
        if (e != nullptr && std::atoll(e) > 0) return (size_t)std::atoll(e);
        return per_wave ? kSymWaveBytes : kSymBytes;
    }
    bool symmetric_fits(int NK, bool per_wave = false) const {
        const size_t need = (size_t)mesh->nf * mesh->nf * NK * 18 * sizeof(cplx);
        if (need > sym_cap(per_wave)) return false;
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
        return need + (size_t(8) << 30) <= fr + W.n * sizeof(cplx);
    }

    int chunk_faces(int NK) const {
        const size_t per_face = (size_t)mesh->nf * NK * 18 * sizeof(cplx);
        size_t c = scratch_bytes / per_face;
        if (c < 4) c = 4;
        return (int)std::min<size_t>(c, (size_t)mesh->nf);
    }
};

}  // namespace sie

// ------------------------------------------------------------ C ABI glue
struct sie_mesh {
    sie::Mesh m;
    sie::DevBuf<double> area;
    sie::Assembler* as = nullptr;
    ~sie_mesh() { delete as; }
};

namespace sie {

Mesh& mesh_of(sie_mesh* h) { return h->m; }
const double* mesh_area(sie_mesh* h) { return h->area.p; }
QuadDev& mesh_quad(sie_mesh* h, int rule, hipStream_t s) {
    quad_upload(h->as->quad, rule, s);
    return h->as->quad;
}

static int NKof(const sie_calderon_params* p) { return (p->k1 == p->k2) ? 1 : 2; }

template <int NK, int KIND>
static void assemble_generic(sie_mesh* h, const sie_calderon_params* p, const Epi& ep, const cplx* Mblk,
                             cplx* Z, long long ldz, int dim, hipStream_t s, const cplx* x = nullptr,
                             cplx* y = nullptr) {
    // x != nullptr: y = Z x without storing Z. Each chunk of source faces is
    // gathered into a slab of the columns of its edges, then y += slab x(those edges).
    Mesh& m = h->m;
    Assembler& A = *h->as;
    A.prepare(p->rule, s);
    // assembly into Z integrates each far pair once (both orientations);
    // products keep the chunked form so that the slab stays a few columns wide
    const bool sym = x == nullptr && A.symmetric_fits(NK);
    if constexpr (NK == 2 && KIND == K_PMCHWT) {
        if (x == nullptr && A.symmetric_fits(1, true)) {
            // one symmetric pass per region's wavenumber, the first storing every
            // entry of Z, the second adding its terms: half the scratch of both
            // wavenumbers together (at N = 61,440 the joint one does not fit
            // beside Z) and the kernel for one wavenumber holds 168 registers
            // (three waves per SIMD) where the joint one needs 360 (one wave):
            // at N = 15,360 the far pairs take 211 against 250 ms, at 61,440
            // 3.1 s against 5.6 s for the non-symmetric chunks of both
            const size_t w1 = (size_t)m.nf * 18 * m.nf;
            if (A.W.n < w1) A.W.alloc(w1);
            std::vector<int> all(m.ne);
            for (int e = 0; e < m.ne; ++e) all[e] = e;
            if (A.elist.n < all.size()) A.elist.alloc(m.ne);
            SIE_HIP(hipMemcpyAsync(A.elist.p, all.data(), all.size() * sizeof(int), hipMemcpyHostToDevice, s));
            dim3 grid((m.ne + 255) / 256, (unsigned)m.ne);
            for (int pass = 0; pass < 2; ++pass) {
                const double k = pass == 0 ? (double)p->k1 : (double)p->k2;
                A.chunk_blocks<1>(0, m.nf, k, k, p->threshold, h->area, s, true, pass == 0);
                prof().begin(PROF_GATHER, s);
                gather_kernel<1, KIND><<<grid, 256, 0, s>>>(m.nf, m.ne, 0, m.nf, A.elist.p, m.ne, m.d_ef.p, m.d_el.p,
                                                            A.W.p, Mblk, ep, Z, ldz, 0, pass == 0 ? 1 : 0, pass);
                prof().end(PROF_GATHER, s, 0.0,
                           (double)m.nf * m.nf * 18 * 16.0 + (double)dim * 2 * m.ne * 16.0 * (pass == 0 ? 1 : 2));
                SIE_HIP(hipGetLastError());
                SIE_HIP(hipStreamSynchronize(s));
            }
            if (w1 * sizeof(cplx) > A.scratch_bytes) A.W.release();
            return;
        }
        if (x != nullptr) {
            // products: per chunk of source faces one pass per region's
            // wavenumber with the one-wavenumber kernel, both gathered into the
            // chunk's slab before its product (L5 sphere: 6.97 s per product
            // with the joint kernel at one wave per SIMD)
            const int nc1 = A.chunk_faces(1);
            const size_t wsz1 = (size_t)nc1 * 18 * m.nf;
            if (A.W.n < wsz1) A.W.alloc(wsz1);
            prof().begin(PROF_MATVEC, s);
            fill_kernel<<<(unsigned)std::min<long long>((dim + 255) / 256, 1 << 16), 256, 0, s>>>(y, dim, dim, 1);
            SIE_HIP(hipGetLastError());
            std::vector<int> list;
            list.reserve(m.ne);
            for (int n0 = 0; n0 < m.nf; n0 += nc1) {
                const int nc = std::min(nc1, m.nf - n0);
                list.clear();
                for (int e = 0; e < m.ne; ++e) {
                    const int g0 = m.ef[0][e], g1 = m.ef[1][e];
                    if ((g0 >= n0 && g0 < n0 + nc) || (g1 >= n0 && g1 < n0 + nc)) list.push_back(e);
                }
                const int L = (int)list.size();
                if (A.elist.n < list.size()) A.elist.alloc(m.ne);
                SIE_HIP(hipMemcpyAsync(A.elist.p, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, s));
                const size_t cols = (size_t)2 * L;
                if (A.slab.n < (size_t)dim * cols) A.slab.alloc((size_t)dim * cols);
                if (A.xs.n < cols) A.xs.alloc((size_t)2 * m.ne);
                const long long tot = (long long)dim * (long long)cols;
                fill_kernel<<<(unsigned)std::min<long long>((tot + 255) / 256, 1 << 16), 256, 0, s>>>(A.slab.p, dim,
                                                                                                      dim, (int)cols);
                dim3 grid((m.ne + 255) / 256, (unsigned)L);
                for (int pass = 0; pass < 2; ++pass) {
                    const double k = pass == 0 ? (double)p->k1 : (double)p->k2;
                    A.chunk_blocks<1>(n0, nc, k, k, p->threshold, h->area, s, false, pass == 0);
                    gather_kernel<1, KIND><<<grid, 256, 0, s>>>(m.nf, m.ne, n0, nc, A.elist.p, L, m.d_ef.p, m.d_el.p,
                                                                A.W.p, Mblk, ep, A.slab.p, dim, L, 0, pass);
                    SIE_HIP(hipGetLastError());
                }
                slab_x_kernel<<<(unsigned)((cols + 255) / 256), 256, 0, s>>>(L, 2, m.ne, A.elist.p, x, A.xs.p);
                SIE_HIP(hipGetLastError());
                const int rc = sie_zgemv(0, dim, (int)cols, 1.0, (const double*)A.slab.p, dim, (const double*)A.xs.p,
                                         1.0, (double*)y, s);
                if (rc != SIE_OK) throw Error(rc, sie_last_error());
                SIE_HIP(hipStreamSynchronize(s));  // list buffer reused by the next chunk
            }
            const double Q = A.quad.n;
            prof().end(PROF_MATVEC, s, (double)m.nf * m.nf * Q * Q * (kFarGeomFlops + 2 * kFarPairFlops),
                       2.0 * 16.0 * dim);
            if (wsz1 * sizeof(cplx) > A.scratch_bytes) A.W.release();
            return;
        }
    }
    const int nc_max = sym ? m.nf : A.chunk_faces(NK);
    const size_t wsz = (size_t)nc_max * NK * 18 * m.nf;
    if (A.W.n < wsz) A.W.alloc(wsz);
    const int blocks = KIND == K_GSTC ? 4 : (KIND == K_PMCHWT ? 2 : 1);
    // one chunk holding every source face: the gather stores every entry of Z
    const bool single = nc_max >= m.nf;
    // a matrix-free product is timed as a whole (flops of its far-pair
    // quadrature over every ordered face pair; bytes: x and y)
    if (x) prof().begin(PROF_MATVEC, s);
    if (x) {
        fill_kernel<<<(unsigned)std::min<long long>((dim + 255) / 256, 1 << 16), 256, 0, s>>>(y, dim, dim, 1);
    } else if (!single) {
        const long long tot = (long long)dim * dim;
        fill_kernel<<<(unsigned)std::min<long long>((tot + 255) / 256, 1 << 16), 256, 0, s>>>(Z, ldz, dim, dim);
    }
    SIE_HIP(hipGetLastError());
    std::vector<int> list;
    list.reserve(m.ne);
    for (int n0 = 0; n0 < m.nf; n0 += nc_max) {
        const int nc = std::min(nc_max, m.nf - n0);
        A.chunk_blocks<NK>(n0, nc, (double)p->k1, (double)p->k2, p->threshold, h->area, s, sym);
        // edges with a face in the chunk
        list.clear();
        for (int e = 0; e < m.ne; ++e) {
            const int g0 = m.ef[0][e], g1 = m.ef[1][e];
            if ((g0 >= n0 && g0 < n0 + nc) || (g1 >= n0 && g1 < n0 + nc)) list.push_back(e);
        }
        const int L = (int)list.size();
        if (A.elist.n < list.size()) A.elist.alloc(m.ne);
        SIE_HIP(hipMemcpyAsync(A.elist.p, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice, s));
        dim3 grid((m.ne + 255) / 256, (unsigned)L);
        if (x) {
            const size_t cols = (size_t)blocks * L;
            if (A.slab.n < (size_t)dim * cols) A.slab.alloc((size_t)dim * cols);
            if (A.xs.n < cols) A.xs.alloc((size_t)blocks * m.ne);
            const long long tot = (long long)dim * (long long)cols;
            fill_kernel<<<(unsigned)std::min<long long>((tot + 255) / 256, 1 << 16), 256, 0, s>>>(A.slab.p, dim, dim,
                                                                                                  (int)cols);
            gather_kernel<NK, KIND><<<grid, 256, 0, s>>>(m.nf, m.ne, n0, nc, A.elist.p, L, m.d_ef.p, m.d_el.p, A.W.p,
                                                         Mblk, ep, A.slab.p, dim, L, 0);
            slab_x_kernel<<<(unsigned)((cols + 255) / 256), 256, 0, s>>>(L, blocks, m.ne, A.elist.p, x, A.xs.p);
            SIE_HIP(hipGetLastError());
            const int rc = sie_zgemv(0, dim, (int)cols, 1.0, (const double*)A.slab.p, dim, (const double*)A.xs.p,
                                     1.0, (double*)y, s);
            if (rc != SIE_OK) throw Error(rc, sie_last_error());
        } else {
            prof().begin(PROF_GATHER, s);
            gather_kernel<NK, KIND><<<grid, 256, 0, s>>>(m.nf, m.ne, n0, nc, A.elist.p, L, m.d_ef.p, m.d_el.p, A.W.p,
                                                         Mblk, ep, Z, ldz, 0, single ? 1 : 0);
            // bytes: the chunk's face-pair blocks read once, its columns of Z written
            prof().end(PROF_GATHER, s, 0.0,
                       (double)m.nf * nc * NK * 18 * 16.0 + (double)dim * blocks * L * 16.0 * (single ? 1 : 2));
        }
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipStreamSynchronize(s));  // list buffer reused by the next chunk
    }
    if (x) {
        const double Q = A.quad.n;
        prof().end(PROF_MATVEC, s, (double)m.nf * m.nf * Q * Q * (kFarGeomFlops + NK * kFarPairFlops),
                   2.0 * 16.0 * dim);
    }
    if (wsz * sizeof(cplx) > A.scratch_bytes) A.W.release();  // the whole-matrix scratch is not kept
}

}  // namespace sie

using namespace sie;

extern "C" {

int sie_mesh_create(const double* vertices, int nv, const int* faces, int nf, const int* face_edges, int ne,
                    sie_mesh_t* out) {
    try {
        require(out != nullptr, "output handle pointer is null");
        auto* h = new sie_mesh();
        try {
            build_mesh(h->m, vertices, nv, faces, nf, face_edges, ne, 0);
            h->area.alloc(nf);
            h->area.upload(h->m.h.area.data(), nf, 0);
            SIE_HIP(hipDeviceSynchronize());
            h->as = new Assembler(&h->m);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        set_last_error("");
        return SIE_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SIE_ERR_DEVICE;
    }
}

int sie_mesh_destroy(sie_mesh_t mesh) {
    delete mesh;
    return SIE_OK;
}

#define SIE_GUARD_BEGIN try {
#define SIE_GUARD_END                         \
    set_last_error("");                       \
    return SIE_OK;                            \
    }                                         \
    catch (const Error& e) {                  \
        set_last_error(e.what());             \
        return e.code;                        \
    }                                         \
    catch (const std::exception& e) {         \
        set_last_error(e.what());             \
        return SIE_ERR_DEVICE;                \
    }

static void check_params(sie_mesh_t mesh, const sie_calderon_params* p) {
    require(mesh && p, "null mesh or parameter pointer");
    require(p->rule >= 1 && p->rule <= 20, "Dunavant quadrature rule must be in [1, 20]");
    require(p->k1 > 0 && p->k2 > 0 && std::isfinite(p->k1) && std::isfinite(p->k2), "wavenumbers must be positive");
    require(p->threshold > 0, "singularity threshold must be positive");
}

int sie_calderon_pair_blocks(sie_mesh_t mesh, const sie_calderon_params* p, double* out_dev, void* stream) {
    SIE_GUARD_BEGIN
    check_params(mesh, p);
    require(out_dev != nullptr, "null output");
    hipStream_t s = (hipStream_t)stream;
    Assembler& A = *mesh->as;
    A.prepare(p->rule, s);
    const int nf = mesh->m.nf;
    const int NK = NKof(p);
    const int ncm = A.chunk_faces(NK);
    const size_t wsz = (size_t)ncm * NK * 18 * nf;
    if (A.W.n < wsz) A.W.alloc(wsz);
    for (int n0 = 0; n0 < nf; n0 += ncm) {
        const int nc = std::min(ncm, nf - n0);
        const long long tot = (long long)nf * nc;
        if (NK == 1) {
            A.chunk_blocks<1>(n0, nc, p->k1, p->k2, p->threshold, mesh->area, s);
            raw_blocks_kernel<1><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, n0, nc, A.W.p, (cplx*)out_dev);
        } else {
            A.chunk_blocks<2>(n0, nc, p->k1, p->k2, p->threshold, mesh->area, s);
            raw_blocks_kernel<2><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(nf, n0, nc, A.W.p, (cplx*)out_dev);
        }
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipStreamSynchronize(s));
    }
    SIE_GUARD_END
}

}  // extern "C"

// The three operators, assembled into Z (x == nullptr) or applied as y = Z x.
static void run_gstc(sie_mesh_t mesh, const sie_calderon_params* p, const double* tensors_dev, int uniform_t,
                     int basis, int flags, cplx* Z, long long ldz, const cplx* x, cplx* y, hipStream_t s) {
    check_params(mesh, p);
    require(basis == 0 || basis == 1, "unsupported GSTC basis type");
    const int E = mesh->m.ne;
    const long long dim = 4LL * E;
    require(dim <= 2147483647LL, "SIE matrix dimension exceeds the int limit");
    if (!x) require(ldz >= dim, "leading dimension smaller than the matrix dimension");
    const bool unscaled = (flags & SIE_ASM_UNSCALED) != 0;
    DevBuf<cplx> M;
    if (flags & SIE_ASM_MASS) {
        require(tensors_dev != nullptr, "GSTC tensors are required for the mass blocks");
        Assembler& A = *mesh->as;
        A.prepare(p->rule, s);
        M.alloc((size_t)mesh->m.nf * 144);
        mass_kernel<<<(mesh->m.nf + 63) / 64, 64, 0, s>>>(mesh->m.nf, A.quad.n, mesh->m.d_V.p, mesh->m.d_F.p,
                                                          mesh->m.d_len.p, mesh->area.p, A.quad.xy.p, A.quad.w.p,
                                                          (const cplx*)tensors_dev, uniform_t, basis, p->mu0,
                                                          p->eps0, unscaled ? 0 : 1, M.p);
        SIE_HIP(hipGetLastError());
    }
    Epi ep{};
    // make_eta0_scaled_calderon_factors: eta0 = sqrt(double(float(mu0/eps0)));
    // the unscaled operator (assemble_Z_calderon_kernel, sie_kernels_assembly.cuh:24-148) has eta0 = 1
    const double eta0 = unscaled ? 1.0 : std::sqrt((double)(p->mu0 / p->eps0));
    const double om = p->omega;
    const double mu[2] = {(double)(p->mu_r1 * p->mu0), (double)(p->mu_r2 * p->mu0)};
    const double ep_[2] = {(double)(p->eps_r1 * p->eps0), (double)(p->eps_r2 * p->eps0)};
    for (int b = 0; b < 2; ++b) {
        ep.jj[b] = {0.0 * mu[b], om * mu[b]};
        ep.mm[b] = {0.0 * (eta0 * eta0 * ep_[b]), om * (eta0 * eta0 * ep_[b])};
    }
    ep.jm = eta0;
    if (!(flags & SIE_ASM_CALDERON)) {
        ep.jj[0] = ep.jj[1] = ep.mm[0] = ep.mm[1] = {0, 0};
        ep.jm = 0;
    }
    if (NKof(p) == 1)
        assemble_generic<1, K_GSTC>(mesh, p, ep, M.p, Z, ldz, (int)dim, s, x, y);
    else
        assemble_generic<2, K_GSTC>(mesh, p, ep, M.p, Z, ldz, (int)dim, s, x, y);
    SIE_HIP(hipStreamSynchronize(s));
}

static void run_pec(sie_mesh_t mesh, const sie_calderon_params* p, cplx* Z, long long ldz, const cplx* x, cplx* y,
                    hipStream_t s) {
    check_params(mesh, p);
    const int E = mesh->m.ne;
    if (!x) require(ldz >= E, "leading dimension smaller than the matrix dimension");
    Epi ep{};
    ep.c[0] = {0.0, (double)(p->omega * p->mu_r2 * p->mu0)};
    sie_calderon_params q = *p;
    q.k1 = p->k2;  // the reference passes k2 for both wavenumbers
    assemble_generic<1, K_PEC>(mesh, &q, ep, nullptr, Z, ldz, E, s, x, y);
    SIE_HIP(hipStreamSynchronize(s));
}

static void run_pmchwt(sie_mesh_t mesh, const sie_calderon_params* p, cplx* Z, long long ldz, const cplx* x,
                       cplx* y, hipStream_t s) {
    check_params(mesh, p);
    const long long dim = 2LL * mesh->m.ne;
    if (!x) require(ldz >= dim, "leading dimension smaller than the matrix dimension");
    Epi ep{};
    const float m1 = p->mu_r1 * p->mu0, m2 = p->mu_r2 * p->mu0, e1 = p->eps_r1 * p->eps0, e2 = p->eps_r2 * p->eps0;
    const double om = p->omega;
    ep.c[0] = {0.0 * (double)m1, om * (double)m1};
    ep.c[1] = {0.0 * (double)m2, om * (double)m2};
    ep.c[2] = {0.0 * (double)e1, om * (double)e1};
    ep.c[3] = {0.0 * (double)e2, om * (double)e2};
    const float eta0 = sqrtf(p->mu0 / p->eps0);
    ep.eta = eta0;
    ep.jm = (double)(eta0 * eta0);
    if (NKof(p) == 1)
        assemble_generic<1, K_PMCHWT>(mesh, p, ep, nullptr, Z, ldz, (int)dim, s, x, y);
    else
        assemble_generic<2, K_PMCHWT>(mesh, p, ep, nullptr, Z, ldz, (int)dim, s, x, y);
    SIE_HIP(hipStreamSynchronize(s));
}

extern "C" {

int sie_assemble_gstc(sie_mesh_t mesh, const sie_calderon_params* p, const double* tensors_dev, int uniform_t,
                      int basis, int flags, double* Z_dev, long long ldz, void* stream) {
    SIE_GUARD_BEGIN
    require(Z_dev != nullptr, "null matrix");
    run_gstc(mesh, p, tensors_dev, uniform_t, basis, flags, (cplx*)Z_dev, ldz, nullptr, nullptr,
             (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_assemble_pec(sie_mesh_t mesh, const sie_calderon_params* p, double* Z_dev, long long ldz, void* stream) {
    SIE_GUARD_BEGIN
    require(Z_dev != nullptr, "null matrix");
    run_pec(mesh, p, (cplx*)Z_dev, ldz, nullptr, nullptr, (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_assemble_pmchwt(sie_mesh_t mesh, const sie_calderon_params* p, double* Z_dev, long long ldz,
                        void* stream) {
    SIE_GUARD_BEGIN
    require(Z_dev != nullptr, "null matrix");
    run_pmchwt(mesh, p, (cplx*)Z_dev, ldz, nullptr, nullptr, (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_gstc_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* tensors_dev, int uniform_t,
                    int basis, int flags, const double* x_dev, double* y_dev, void* stream) {
    SIE_GUARD_BEGIN
    require(x_dev != nullptr && y_dev != nullptr && x_dev != y_dev, "matvec needs distinct x and y vectors");
    run_gstc(mesh, p, tensors_dev, uniform_t, basis, flags, nullptr, 0, (const cplx*)x_dev, (cplx*)y_dev,
             (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_pec_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* x_dev, double* y_dev,
                   void* stream) {
    SIE_GUARD_BEGIN
    require(x_dev != nullptr && y_dev != nullptr && x_dev != y_dev, "matvec needs distinct x and y vectors");
    run_pec(mesh, p, nullptr, 0, (const cplx*)x_dev, (cplx*)y_dev, (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_pmchwt_matvec(sie_mesh_t mesh, const sie_calderon_params* p, const double* x_dev, double* y_dev,
                      void* stream) {
    SIE_GUARD_BEGIN
    require(x_dev != nullptr && y_dev != nullptr && x_dev != y_dev, "matvec needs distinct x and y vectors");
    run_pmchwt(mesh, p, nullptr, 0, (const cplx*)x_dev, (cplx*)y_dev, (hipStream_t)stream);
    SIE_GUARD_END
}

int sie_surface_fields(sie_mesh_t mesh, int rule, double omega, double epsilon, double mu, const double* J_dev,
                       const double* M_dev, const double* obs_dev, int nobs, double* E_dev, double* H_dev,
                       void* stream) {
    SIE_GUARD_BEGIN
    require(mesh != nullptr, "null mesh");
    require(rule >= 1 && rule <= 20, "Dunavant quadrature rule must be in [1, 20]");
    require(omega > 0 && epsilon > 0 && mu > 0 && std::isfinite(omega) && std::isfinite(epsilon) &&
                std::isfinite(mu),
            "field-evaluation frequency and medium must be positive and finite");
    require(nobs >= 0, "observation point count must be non-negative");
    if (nobs == 0) return SIE_OK;
    require(J_dev && M_dev && obs_dev && E_dev && H_dev, "null field-evaluation argument");
    hipStream_t s = (hipStream_t)stream;
    Assembler& A = *mesh->as;
    quad_upload(A.quad, rule, s);
    const Mesh& m = mesh->m;
    const double k = omega * std::sqrt(mu * epsilon);
    surface_fields_kernel<<<(unsigned)nobs, kFieldThreads, 0, s>>>(
        m.nf, A.quad.n, m.d_V.p, m.d_F.p, m.d_edge.p, m.d_c.p, m.d_len.p, A.quad.xy.p, A.quad.w.p,
        (const cplx*)J_dev, (const cplx*)M_dev, obs_dev, omega * mu, omega * epsilon, k, (cplx*)E_dev, (cplx*)H_dev);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    SIE_GUARD_END
}

int sie_rhs_plane_waves(sie_mesh_t mesh, int kind, int rule, const sie_plane_wave* waves, int nwaves, float mu0,
                        float eps0, double vac_mu, double vac_eps, double* rhs_dev, long long ldb, void* stream) {
    SIE_GUARD_BEGIN
    require(mesh && waves && rhs_dev, "null argument");
    require(nwaves > 0, "at least one excitation is required");
    require(kind >= 0 && kind <= 2, "unknown RHS kind");
    const int nf = mesh->m.nf, ne = mesh->m.ne;
    const long long dim = kind == SIE_RHS_GSTC ? 4LL * ne : kind == SIE_RHS_PMCHWT ? 2LL * ne : ne;
    require(ldb >= dim, "leading dimension smaller than the RHS length");
    hipStream_t s = (hipStream_t)stream;
    Assembler& A = *mesh->as;
    quad_upload(A.quad, rule, s);
    std::vector<PW> hw(nwaves);
    for (int w = 0; w < nwaves; ++w) {
        const sie_plane_wave& q = waves[w];
        require(q.region == 1 || q.region == 2, "source region must be interior (1) or exterior (2)");
        require(q.k > 0 && q.eta > 0, "plane-wave k and eta must be positive");
        for (int x = 0; x < 3; ++x) { hw[w].E0[x] = q.E0[x]; hw[w].kh[x] = q.khat[x]; }
        hw[w].k = q.k;
        hw[w].ieta = 1.0 / q.eta;
        hw[w].region = q.region;
    }
    DevBuf<PW> dw(nwaves);
    dw.upload(hw.data(), nwaves, s);
    DevBuf<cplx> face((size_t)nwaves * nf * 6);
    const long long t1 = (long long)nf * nwaves;
    rhs_face_kernel<<<(unsigned)((t1 + 127) / 128), 128, 0, s>>>(nf, nwaves, A.quad.n, mesh->m.d_V.p, mesh->m.d_F.p,
                                                                 mesh->m.d_len.p, A.quad.xy.p, A.quad.w.p, dw.p,
                                                                 face.p);
    SIE_HIP(hipGetLastError());
    const double eta_g = std::sqrt((double)mu0 / (double)eps0);
    const double eta_d = std::sqrt(vac_mu / vac_eps);
    const long long t2 = (long long)ne * nwaves;
    rhs_gather_kernel<<<(unsigned)((t2 + 255) / 256), 256, 0, s>>>(nf, ne, nwaves, kind, mesh->m.d_ef.p,
                                                                   mesh->m.d_el.p, dw.p, face.p, eta_g, eta_d,
                                                                   (cplx*)rhs_dev, ldb);
    SIE_HIP(hipGetLastError());
    SIE_HIP(hipStreamSynchronize(s));
    SIE_GUARD_END
}

}  // extern "C"
