// Dense complex linear algebra on MI355X (gfx950): the solver half of the
// SIE-MoM hot path. Replaces the reference's cuSOLVER getrf/getrs and cuBLAS
// gemv (include/sie_mom/cuda/dense_lu_solver.cuh:42-166, gstc_sie.cu:398-457)
// with own kernels; algorithms follow the textbook right-looking blocked LU
// with partial pivoting (Golub & Van Loan, Matrix Computations, 3.4).
//
//  * cgemm_mfma<T>: C = beta C + alpha A B for complex column-major blocks on
//    the matrix cores. FP32: v_mfma_f32_32x32x2f32 with a 128x128 workgroup
//    tile, real 4M (four real products) or 3M (Gauss) decomposition; FP64:
//    v_mfma_f64_16x16x4f64 with a 64x128 tile, 4M (its D layout: lane l,
//    register v holds D[4v + l/16][l%16]). The roles of A and B are
//    swapped in the MFMA so that each lane owns rows of C: the epilogue's
//    column-major stores are contiguous across lanes.
//  * panel_getf2<T>: the partial-pivoting factorisation of a tall panel of
//    width W as one cooperative kernel: one row per lane held in registers,
//    one grid barrier per column (local argmax -> publish -> barrier -> every
//    workgroup reduces the candidates and swaps/scales/updates its rows).
//  * getrf: outer blocks of NB = 256 columns (the trailing GEMM's K), panels
//    of W columns inside; row swaps, unit-lower TRSM and the trailing GEMM.
//  * getrs (op N / C), GEMV (op N / C) with deterministic two-stage
//    reductions, and FP64 iterative refinement around an FP32 factorisation.
#include "sie_internal.hpp"
#include "prof.hpp"

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace sie {

template <typename T>
struct cx {
    T re, im;
};
using c32 = cx<float>;
using c64 = cx<double>;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// Bound of the device-side waits: the panels' grid-wide exchanges and the
// sweeps' published entries. ticks counts the 100 MHz constant clock
// (s_memrealtime). A wait that reaches the bound writes a nonzero code into
// *err and its kernel returns; every exchange also gives up as soon as *err
// is set, so one stalled workgroup ends the whole grid. The host reads *err
// at its next synchronisation and reports SIE_ERR_DEVICE. force (testing):
// the panel exchanges wait as if a workgroup never arrived, so they give up
// at the bound.
struct WaitArg {
    unsigned long long ticks;
    unsigned int* err;
    int force;
};
// codes in *err
constexpr unsigned kWaitPanel = 1, kWaitSweep = 2, kWaitRow = 3;

__device__ __forceinline__ void flag_wait_error(unsigned int* err, unsigned code) {
    unsigned zero = 0;
    __hip_atomic_compare_exchange_strong(err, &zero, code, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool wait_error_set(const unsigned int* err) {
    return __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

namespace {

// =================================================================== GEMM
// FP32: 4 waves as 2x2, wave tile (32 TI) x (32 TI) complex = TI x TI MFMA
// 32x32 tiles; block tile (64 TI)^2. TI = 2 (128 x 128) for long K; TI = 1
// (64 x 64) for the LU's K = 32 updates, whose 128 x 128 tiles are too few
// and too long (16k cycles of MFMA per wave) to fill the chip.
constexpr int G32_KB = 16;

template <int MODE, int TI>  // MODE 0 = 4M, 1 = 3M
__global__ __launch_bounds__(256, 2) void cgemm_f32_kernel(int M, int N, int K, float alpha, float beta,
                                                           const c32* __restrict__ A, long long lda,
                                                           const c32* __restrict__ B, long long ldb,
                                                           c32* __restrict__ C, long long ldc) {
    constexpr int BM = 64 * TI, BN = 64 * TI;
    constexpr int LA = G32_KB * BM / 256, LB = BN * G32_KB / 256;  // elements per thread and K tile
    __shared__ c32 sA[2][G32_KB][BM];
    __shared__ c32 sB[2][BN][G32_KB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const long long i0 = (long long)blockIdx.x * BM, j0 = (long long)blockIdx.y * BN;

    constexpr int NACC = MODE == 0 ? 2 : 3;
    floatx16 acc[NACC][TI][TI];
#pragma unroll
    for (int a = 0; a < NACC; ++a)
#pragma unroll
        for (int ti = 0; ti < TI; ++ti)
#pragma unroll
            for (int tj = 0; tj < TI; ++tj)
#pragma unroll
                for (int v = 0; v < 16; ++v) acc[a][ti][tj][v] = 0.f;

    c32 ra[LA], rb[LB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < LA; ++r) {
            // A tile: [k][i], i fastest
            const int L = r * 256 + tid;
            const int k = L / BM, i = L % BM;
            const long long gi = i0 + i, gk = k0 + k;
            ra[r] = (gi < M && gk < K) ? A[gk * lda + gi] : c32{0.f, 0.f};
        }
#pragma unroll
        for (int r = 0; r < LB; ++r) {
            // B tile: [j][k], k fastest
            const int L = r * 256 + tid;
            const int j = L / G32_KB, kk = L % G32_KB;
            const long long gj = j0 + j, gkk = k0 + kk;
            rb[r] = (gj < N && gkk < K) ? B[gj * ldb + gkk] : c32{0.f, 0.f};
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int r = 0; r < LA; ++r) {
            const int L = r * 256 + tid;
            sA[buf][L / BM][L % BM] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < LB; ++r) {
            const int L = r * 256 + tid;
            sB[buf][L / G32_KB][L % G32_KB] = rb[r];
        }
    };

    const int nk = (K + G32_KB - 1) / G32_KB;
    gload(0);
    // short K (TI = 1): C is loaded together with the first operand tiles, so
    // that the epilogue waits for no second memory round trip (the products
    // are still summed from zero and added to beta C once, as for long K)
    c32 cpre[TI == 1 ? 16 : 1];
    if constexpr (TI == 1) {
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const long long i = i0 + wm * 32 + (lane & 31);
            const long long j = j0 + wn * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            cpre[v] = beta != 0.f && i < M && j < N ? C[j * ldc + i] : c32{0.f, 0.f};
        }
    }
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * G32_KB);
#pragma unroll
        for (int kk = 0; kk < G32_KB; kk += 2) {
            c32 af[TI], bf[TI];
#pragma unroll
            for (int ti = 0; ti < TI; ++ti)
                af[ti] = sA[buf][kk + (lane >> 5)][wm * 32 * TI + ti * 32 + (lane & 31)];
#pragma unroll
            for (int tj = 0; tj < TI; ++tj)
                bf[tj] = sB[buf][wn * 32 * TI + tj * 32 + (lane & 31)][kk + (lane >> 5)];
#pragma unroll
            for (int ti = 0; ti < TI; ++ti)
#pragma unroll
                for (int tj = 0; tj < TI; ++tj) {
                    if (MODE == 0) {
                        acc[0][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].re, af[ti].re, acc[0][ti][tj], 0, 0, 0);
                        acc[0][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(-bf[tj].im, af[ti].im, acc[0][ti][tj], 0, 0, 0);
                        acc[1][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].re, af[ti].im, acc[1][ti][tj], 0, 0, 0);
                        acc[1][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].im, af[ti].re, acc[1][ti][tj], 0, 0, 0);
                    } else {
                        acc[0][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].re, af[ti].re, acc[0][ti][tj], 0, 0, 0);
                        acc[1][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].im, af[ti].im, acc[1][ti][tj], 0, 0, 0);
                        acc[2][ti][tj] = __builtin_amdgcn_mfma_f32_32x32x2f32(bf[tj].re + bf[tj].im, af[ti].re + af[ti].im,
                                                                              acc[2][ti][tj], 0, 0, 0);
                    }
                }
        }
        if (kt + 1 < nk) {
            lstore(buf ^ 1);
        }
        __syncthreads();
    }
    // epilogue: lane owns row i = ibase + lane%32, column j = jbase + 8(v/4) + 4(lane/32) + v%4
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < TI; ++tj)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long i = i0 + wm * 32 * TI + ti * 32 + (lane & 31);
                const long long j = j0 + wn * 32 * TI + tj * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                if (i < M && j < N) {
                    float re, im;
                    if (MODE == 0) {
                        re = acc[0][ti][tj][v];
                        im = acc[1][ti][tj][v];
                    } else {
                        re = acc[0][ti][tj][v] - acc[1][ti][tj][v];
                        im = acc[2][ti][tj][v] - acc[0][ti][tj][v] - acc[1][ti][tj][v];
                    }
                    c32* p = C + j * ldc + i;
                    c32 c;
                    if constexpr (TI == 1) c = cpre[v];
                    else c = beta != 0.f ? *p : c32{0.f, 0.f};
                    c.re = beta * c.re + alpha * re;
                    c.im = beta * c.im + alpha * im;
                    *p = c;
                }
            }
}

// FP64: 4 waves as 2x2, wave tile 32x64 complex = 2x4 MFMA 16x16 tiles.
constexpr int G64_BM = 64, G64_BN = 128, G64_KB = 8;

__global__ __launch_bounds__(256, 2) void zgemm_f64_kernel(int M, int N, int K, double alpha, double beta,
                                                           const c64* __restrict__ A, long long lda,
                                                           const c64* __restrict__ B, long long ldb,
                                                           c64* __restrict__ C, long long ldc) {
    __shared__ c64 sA[2][G64_KB][G64_BM];
    __shared__ c64 sB[2][G64_BN][G64_KB + 1];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const long long i0 = (long long)blockIdx.x * G64_BM, j0 = (long long)blockIdx.y * G64_BN;
    doublex4 acc[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[a][ti][tj][v] = 0.0;
    c64 ra[2], rb[4];
    auto gload = [&](int k0) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {  // A: 8 x 64 = 512 elements
            const int L = r * 256 + tid;
            const int k = L >> 6, i = L & 63;
            const long long gi = i0 + i, gk = k0 + k;
            ra[r] = (gi < M && gk < K) ? A[gk * lda + gi] : c64{0.0, 0.0};
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // B: 128 x 8 = 1024 elements
            const int L = r * 256 + tid;
            const int j = L >> 3, kk = L & 7;
            const long long gj = j0 + j, gkk = k0 + kk;
            rb[r] = (gj < N && gkk < K) ? B[gj * ldb + gkk] : c64{0.0, 0.0};
        }
    };
    auto lstore = [&](int buf) {
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int L = r * 256 + tid;
            sA[buf][L >> 6][L & 63] = ra[r];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int L = r * 256 + tid;
            sB[buf][L >> 3][L & 7] = rb[r];
        }
    };
    const int nk = (K + G64_KB - 1) / G64_KB;
    gload(0);
    lstore(0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
        const int buf = kt & 1;
        if (kt + 1 < nk) gload((kt + 1) * G64_KB);
#pragma unroll
        for (int kk = 0; kk < G64_KB; kk += 4) {
            c64 af[2], bf[4];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) af[ti] = sA[buf][kk + (lane >> 4)][wm * 32 + ti * 16 + (lane & 15)];
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) bf[tj] = sB[buf][wn * 64 + tj * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 4; ++tj) {
                    acc[0][ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[tj].re, af[ti].re, acc[0][ti][tj], 0, 0, 0);
                    acc[0][ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bf[tj].im, af[ti].im, acc[0][ti][tj], 0, 0, 0);
                    acc[1][ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[tj].re, af[ti].im, acc[1][ti][tj], 0, 0, 0);
                    acc[1][ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[tj].im, af[ti].re, acc[1][ti][tj], 0, 0, 0);
                }
        }
        if (kt + 1 < nk) lstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const long long i = i0 + wm * 32 + ti * 16 + (lane & 15);
                // v_mfma_f64_16x16x4f64: lane l, register v holds D[4v + l/16][l%16]
                const long long j = j0 + wn * 64 + tj * 16 + 4 * v + (lane >> 4);
                if (i < M && j < N) {
                    c64* p = C + j * ldc + i;
                    c64 c = beta != 0.0 ? *p : c64{0.0, 0.0};
                    c.re = beta * c.re + alpha * acc[0][ti][tj][v];
                    c.im = beta * c.im + alpha * acc[1][ti][tj][v];
                    *p = c;
                }
            }
}

// ===================================================== split-FP16 GEMM
// C = beta C + alpha A B for complex FP32 operands on the FP16 matrix cores
// (v_mfma_f32_32x32x16_f16, FP32 accumulation). After power-of-two scaling of
// each operand to max |re|, |im| < 1, every real entry x is split as
//   h = fp16(x),  l' = fp16((x - h) 2^11)
// (22 significant bits; the 2^11 offset keeps l' out of the FP16 subnormal
// range down to |x| ~ 2^-13). A real product x y is then formed, scaled by
// 2^11, from three FP16 products
//   x y 2^11 ~ h_x (h_y 2^11) + h_x l'_y + l'_x h_y
// (the dropped l l' term is ~2^-22 relative). B is stored as (h 2^11, l');
// the kernel recovers h_y = (h_y 2^11) 2^-11 exactly in FP16. Complex products
// use the 4M form. Lane layout of the FP16 MFMA: lane l holds X[l%32][8(l/32)+e]
// and Y[8(l/32)+e][l%32], e = 0..7; the output follows the 32x32x2f32 layout.
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
constexpr float kLoShift = 2048.f;  // 2^11

// max(|re|, |im|) of a column-major rows x cols block, as the bits of a
// non-negative float (which order like unsigned integers) in *out.
__global__ __launch_bounds__(256) void maxabs_kernel(long long rows, long long cols, const c32* __restrict__ X,
                                                     long long ld, unsigned int* __restrict__ out) {
    __shared__ float wmax[4];
    float m = 0.f;
    const long long tot = rows * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / rows, r = t - c * rows;
        const c32 v = X[c * ld + r];
        m = fmaxf(m, fmaxf(fabsf(v.re), fabsf(v.im)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    // one atomic per block: non-negative floats order like their bit patterns
    if (threadIdx.x == 0) atomicMax(out, __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
}

// mx[0] = a bound given by the caller (float bits), mx[1] = 0 for maxabs_kernel.
__global__ void set_mx_kernel(unsigned int* __restrict__ mx, unsigned int a_bits) {
    mx[0] = a_bits;
    mx[1] = 0u;
}

// 2^-e with max = f 2^e, f in [0.5, 1): scaled entries have modulus < 1.
__device__ __forceinline__ float pow2_scale(unsigned int bits) {
    const float m = __uint_as_float(bits);
    if (!(m > 0.f) || !(m < INFINITY)) return 1.f;
    int e;
    frexpf(m, &e);
    return ldexpf(1.f, -e);
}

// Split planes of an operand with "row" index r (i for A, j for B) and depth k:
// P[plane][k/16][(k/8)%2][r][8], planes {re hi, re lo', im hi, im lo'}, depth
// padded with zeros to a multiple of 16; the 8 halves at (k/16, (k/8)%2, r) are
// one MFMA lane's operand. trans = 0: X is A (element (r, k) at X[k ld + r]);
// trans = 1: X is B (element (k, r) at X[r ld + k]) and its hi planes carry the
// 2^11 offset.
__global__ void split_f16_kernel(int R, int K, const c32* __restrict__ X, long long ld, int trans,
                                 const unsigned int* __restrict__ mx, _Float16* __restrict__ P) {
    const int KB = (K + 15) / 16;
    const long long total = (long long)KB * R * 16;
    const size_t plane = (size_t)total;
    const float sc = pow2_scale(*mx);
    const float hs = trans ? kLoShift : 1.f;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(t & 7);
        const long long rest = t >> 3;
        const int r = (int)(rest % R);
        const long long kp = rest / R;  // kb * 2 + part
        const int k = (int)(kp * 8) + e;
        c32 v = {0.f, 0.f};
        if (k < K) v = trans ? X[(long long)r * ld + k] : X[(long long)k * ld + r];
        const float xr = v.re * sc, xi = v.im * sc;
        const _Float16 rh = (_Float16)xr, ih = (_Float16)xi;
        P[t] = (_Float16)((float)rh * hs);
        P[plane + t] = (_Float16)((xr - (float)rh) * kLoShift);
        P[2 * plane + t] = (_Float16)((float)ih * hs);
        P[3 * plane + t] = (_Float16)((xi - (float)ih) * kLoShift);
    }
}

// The same split into the tile-blocked layout of cgemm_f16x3_r3_kernel:
// P[blk][kb][plane][part][rr][8] with rows r = blk RB + rr padded with zeros to
// whole blocks, so that one tile's K slice (all four planes) is one contiguous
// RB x 128-byte piece and its whole K panel one contiguous range. One thread
// per (row, 8-wide k group) writes a 16-byte vector to each plane.
__global__ void split_f16_blk_kernel(int R, int K, int RB, const c32* __restrict__ X, long long ld, int trans,
                                     const unsigned int* __restrict__ mx, _Float16* __restrict__ P) {
    const int KB = (K + 15) / 16, nblk = (R + RB - 1) / RB;
    const long long total = (long long)nblk * RB * KB * 2;
    const float sc = pow2_scale(*mx);
    const float hs = trans ? kLoShift : 1.f;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        // rows fastest: consecutive lanes read consecutive rows of a column of X
        const int rr = (int)(t % RB);
        long long rest = t / RB;
        const int part = (int)(rest & 1);
        rest >>= 1;
        const int kb = (int)(rest % KB);
        const int blk = (int)(rest / KB);
        const int r = blk * RB + rr;
        half8 h[4];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int k = kb * 16 + part * 8 + e;
            c32 v = {0.f, 0.f};
            if (r < R && k < K) v = trans ? X[(long long)r * ld + k] : X[(long long)k * ld + r];
            const float xr = v.re * sc, xi = v.im * sc;
            const _Float16 rh = (_Float16)xr, ih = (_Float16)xi;
            h[0][e] = (_Float16)((float)rh * hs);
            h[1][e] = (_Float16)((xr - (float)rh) * kLoShift);
            h[2][e] = (_Float16)((float)ih * hs);
            h[3][e] = (_Float16)((xi - (float)ih) * kLoShift);
        }
        half8* out = (half8*)P + (((long long)blk * KB + kb) * 8) * RB + (long long)part * RB + rr;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) out[(long long)pl * 2 * RB] = h[pl];
    }
}

// Tile of a linear block id. Work-groups are dispatched round-robin over the
// 8 XCDs, each with its own L2: give every XCD a contiguous range of tiles and
// walk that range in groups of GM tile rows, column by column, so that the
// blocks resident on one XCD share their A and B panels in its L2.
// Tile rows per group of the 256 x 128 kernel: its 32 co-resident workgroups
// per XCD then cover 4 x 8 tiles, a 1024 x 1024 block of C whose A and B
// panels are equally long (the 128 x 128 kernel's 64 cover 8 x 8 with GM = 8).
constexpr int kWideGroupRows = 4;

template <bool SWZ, int GM>
__device__ __forceinline__ void f16x3_tile(int id, int nbm, int nbn, int& tm, int& tn) {
    if (!SWZ) {
        tm = id % nbm;
        tn = id / nbm;
        return;
    }
    const int total = nbm * nbn;
    const int per = total >> 3, rem = total & 7, x = id & 7;
    const int lin = x * per + min(x, rem) + (id >> 3);
    const int gsz = GM * nbn;
    const int g = lin / gsz, first = g * GM, rows = min(nbm - first, GM);
    const int r = lin - g * gsz;
    tm = first + r % rows;
    tn = r / rows;
}

typedef unsigned int uint4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ half8 neg8(half8 x) {
    // flip the eight sign bits, one 32-bit xor per pair of halves
    return __builtin_bit_cast(half8, __builtin_bit_cast(uint4v, x) ^ 0x80008000u);
}
__device__ __forceinline__ half8 unshift8(half8 x) { return x * (_Float16)(1.f / kLoShift); }

// The accumulators start from beta C in accumulator units (C g with
// g = beta 2^11 sA sB / alpha, exact for alpha, beta = +-1), loaded before the
// first K slice so that the read of C overlaps that slice's loads and does
// not stall the end of the tile; the epilogue then only writes C = f acc with
// f = alpha / (2^11 sA sB). Requires |beta C g| < FLT_MAX.
template <int TI>
__device__ __forceinline__ void f16x3_acc_init(floatx16 (&accR)[TI][2], floatx16 (&accI)[TI][2],
                                               const c32* __restrict__ C, long long ldc, int M, int N, int ib,
                                               int jb, int r32, int part, float g, bool load) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long i = ib + ti * 32 + r32;
                const long long j = jb + tj * 32 + 8 * (v >> 2) + 4 * part + (v & 3);
                c32 c = {0.f, 0.f};
                if (load && i < M && j < N) c = C[j * ldc + i];
                accR[ti][tj][v] = c.re * g;
                accI[ti][tj][v] = c.im * g;
            }
}

template <int TI>
__device__ __forceinline__ void f16x3_acc_store(const floatx16 (&accR)[TI][2], const floatx16 (&accI)[TI][2],
                                                c32* __restrict__ C, long long ldc, int M, int N, int ib, int jb,
                                                int r32, int part, float f) {
#pragma unroll
    for (int ti = 0; ti < TI; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const long long i = ib + ti * 32 + r32;
                const long long j = jb + tj * 32 + 8 * (v >> 2) + 4 * part + (v & 3);
                if (i < M && j < N) C[j * ldc + i] = c32{f * accR[ti][tj][v], f * accI[ti][tj][v]};
            }
}

// One K slice of 16: the 4M complex product of the wave tile, 3 MFMAs per
// real product (hi*hi, hi*lo, lo*hi) with the small terms first.
template <int TI>
__device__ __forceinline__ void f16x3_mfma(const half8 (&a)[TI][4], const half8 (&b)[2][4], floatx16 (&accR)[TI][2],
                                           floatx16 (&accI)[TI][2]) {
#pragma unroll
    for (int tj = 0; tj < 2; ++tj) {
        // B hi planes carry 2^11: b?2 = h 2^11, b?l = l 2^11, b? = h
        const half8 br2 = b[tj][0], brl = b[tj][1], bi2 = b[tj][2], bil = b[tj][3];
        const half8 br = unshift8(br2), bi = unshift8(bi2);
#pragma unroll
        for (int ti = 0; ti < TI; ++ti) {
            const half8 arh = a[ti][0], arl = a[ti][1], aih = a[ti][2], ail = a[ti][3];
            const half8 naih = neg8(aih), nail = neg8(ail);
            floatx16 cr = accR[ti][tj], ci = accI[ti][tj];
            // re: ar br - ai bi, small terms first
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(brl, arh, cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(br, arl, cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(bil, naih, cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi, nail, cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(br2, arh, cr, 0, 0, 0);
            cr = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi2, naih, cr, 0, 0, 0);
            // im: ar bi + ai br
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(bil, arh, ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi, arl, ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(brl, aih, ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(br, ail, ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(bi2, arh, ci, 0, 0, 0);
            ci = __builtin_amdgcn_mfma_f32_32x32x16_f16(br2, aih, ci, 0, 0, 0);
            accR[ti][tj] = cr;
            accI[ti][tj] = ci;
        }
    }
}

#define SIE_GPTR(p) ((const void __attribute__((address_space(1)))*)(p))
#define SIE_LPTR(p) ((void __attribute__((address_space(3)))*)(p))

// Block tile (64 TI) x 128, four waves in a 2 x 2 grid, wave tile (32 TI) x 64.
// K is consumed 16 at a time through two LDS buffers filled by direct
// global->LDS loads (global_load_lds_dwordx4: lane q of a wave writes 16 bytes
// at the wave's LDS base + 16 q). Each iteration waits for its own buffer,
// reads its fragments, then starts the next slice's loads into the other
// buffer so that they overlap this slice's MFMAs. Rows past M (columns past
// N) load a clamped in-range row; they only feed outputs that are not stored.
template <int TI, int MINB, bool SWZ>
__global__ __launch_bounds__(256, MINB) void cgemm_f16x3_kernel(int M, int N, int K, float alpha, float beta,
                                                                const _Float16* __restrict__ PA,
                                                                const _Float16* __restrict__ PB,
                                                                const unsigned int* __restrict__ mx,
                                                                c32* __restrict__ C, long long ldc) {
    constexpr int BM = 64 * TI, BN = 128;
    constexpr int LA = BM * 8 / 256, LB = BN * 8 / 256;
    // [buffer][plane][k half][row]: a lane's 16 bytes; 32 lanes read 512 contiguous bytes
    __shared__ half8 sA[2][4][2][BM];
    __shared__ half8 sB[2][4][2][BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    int tm, tn;
    f16x3_tile<SWZ, 8>(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int i0 = tm * BM, j0 = tn * BN;
    const int KB = (K + 15) / 16;
    const size_t planeA = (size_t)KB * M * 16, planeB = (size_t)KB * N * 16;
    // source offsets (halves) of this thread's 16-byte pieces in slice 0
    size_t offA[LA], offB[LB];
#pragma unroll
    for (int q = 0; q < LA; ++q) {
        const int L = q * 256 + tid;  // LDS piece (plane, k half, row), row fastest
        const int plane = L / (2 * BM), part = (L / BM) & 1, row = L % BM;
        offA[q] = plane * planeA + ((size_t)part * M + min(i0 + row, M - 1)) * 8;
    }
#pragma unroll
    for (int q = 0; q < LB; ++q) {
        const int L = q * 256 + tid;
        const int plane = L / (2 * BN), part = (L / BN) & 1, row = L % BN;
        offB[q] = plane * planeB + ((size_t)part * N + min(j0 + row, N - 1)) * 8;
    }
    auto dma = [&](int kb, int buf) {
        const size_t ka = (size_t)kb * M * 16, kbb = (size_t)kb * N * 16;
        half8* la = &sA[buf][0][0][0] + wave * 64;
        half8* lb = &sB[buf][0][0][0] + wave * 64;  // lane q lands at base + 16 q bytes
#pragma unroll
        for (int q = 0; q < LA; ++q)
            __builtin_amdgcn_global_load_lds(SIE_GPTR(PA + offA[q] + ka), SIE_LPTR(la + q * 256), 16, 0, 0);
#pragma unroll
        for (int q = 0; q < LB; ++q)
            __builtin_amdgcn_global_load_lds(SIE_GPTR(PB + offB[q] + kbb), SIE_LPTR(lb + q * 256), 16, 0, 0);
    };
    floatx16 accR[TI][2], accI[TI][2];
    const float unit = pow2_scale(mx[0]) * pow2_scale(mx[1]) * kLoShift;
    f16x3_acc_init<TI>(accR, accI, C, ldc, M, N, i0 + wm * 32 * TI, j0 + wn * 64, lane & 31, lane >> 5,
                       beta * unit / alpha, beta != 0.f);
    dma(0, 0);
    const int part = lane >> 5, r32 = lane & 31;
    for (int kb = 0; kb < KB; ++kb) {
        const int buf = kb & 1;
        // this slice's loads (the only ones in flight) have landed for every wave,
        // and every wave is done reading the other buffer
        asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        half8 a[TI][4], b[2][4];
#pragma unroll
        for (int t = 0; t < TI; ++t)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) a[t][pl] = sA[buf][pl][part][wm * 32 * TI + t * 32 + r32];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) b[t][pl] = sB[buf][pl][part][wn * 64 + t * 32 + r32];
        if (kb + 1 < KB) dma(kb + 1, buf ^ 1);
        f16x3_mfma<TI>(a, b, accR, accI);
    }
    f16x3_acc_store<TI>(accR, accI, C, ldc, M, N, i0 + wm * 32 * TI, j0 + wn * 64, lane & 31, lane >> 5, alpha / unit);
}

// Block tile 256 x 128, eight waves in a 4 x 2 grid, wave tile 64 x 64, one
// workgroup per CU. Three LDS stages (48 KB each, 144 KB) let the loads of
// slice k + 2 start while slice k is multiplied, two slices of MFMAs ahead of
// their use; each stage is its own LDS object so that the compiler's wait
// for outstanding direct-to-LDS loads before an LDS read can be limited to
// the loads into that stage. The wave tile and the per-slice product are
// those of cgemm_f16x3_kernel; a 256-row tile halves the B-panel traffic per
// MFMA of the 128-row tile.
struct F16Stage {
    half8 a[4][2][256];
    half8 b[4][2][128];
};

template <bool SWZ>
__global__ __launch_bounds__(512, 1) void cgemm_f16x3_r3_kernel(int M, int N, int K, float alpha, float beta,
                                                               const _Float16* __restrict__ PA,
                                                               const _Float16* __restrict__ PB,
                                                               const unsigned int* __restrict__ mx,
                                                               c32* __restrict__ C, long long ldc) {
    constexpr int BM = 256, BN = 128, TI = 2;
    constexpr int LA = BM * 8 / 512, LB = BN * 8 / 512;  // 16-byte pieces per thread and slice
    __shared__ F16Stage st0, st1, st2;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 3, wn = wave >> 2;
    int tm, tn;
    f16x3_tile<SWZ, kWideGroupRows>(blockIdx.x, (M + BM - 1) / BM, (N + BN - 1) / BN, tm, tn);
    const int i0 = tm * BM, j0 = tn * BN;
    const int KB = (K + 15) / 16;
    // Operands in the tile-blocked layout of split_f16_blk_kernel: the K slice
    // of this tile is 32 KB (A) and 16 KB (B) of contiguous bytes whose order
    // is the LDS order, so thread t copies bytes 16 (q 512 + t). 32-bit byte
    // offsets: the caller keeps each split operand below 4 GiB.
    const unsigned sliceA = BM * 128u, sliceB = BN * 128u;
    const unsigned baseA = (unsigned)tm * KB * sliceA + tid * 16u, baseB = (unsigned)tn * KB * sliceB + tid * 16u;
    const char* pa8 = (const char*)PA;
    const char* pb8 = (const char*)PB;
    auto dma = [&](F16Stage& d, int kb) {
        const unsigned oa = baseA + (unsigned)kb * sliceA, ob = baseB + (unsigned)kb * sliceB;
        half8* la = &d.a[0][0][0] + wave * 64;
        half8* lb = &d.b[0][0][0] + wave * 64;  // lane q lands at base + 16 q bytes
#pragma unroll
        for (int q = 0; q < LA; ++q)
            __builtin_amdgcn_global_load_lds(SIE_GPTR(pa8 + (oa + q * 8192u)), SIE_LPTR(la + q * 512), 16, 0, 0);
#pragma unroll
        for (int q = 0; q < LB; ++q)
            __builtin_amdgcn_global_load_lds(SIE_GPTR(pb8 + (ob + q * 8192u)), SIE_LPTR(lb + q * 512), 16, 0, 0);
    };
    floatx16 accR[TI][2], accI[TI][2];
    const float unit = pow2_scale(mx[0]) * pow2_scale(mx[1]) * kLoShift;
    f16x3_acc_init<TI>(accR, accI, C, ldc, M, N, i0 + wm * 64, j0 + wn * 64, lane & 31, lane >> 5,
                       beta * unit / alpha, beta != 0.f);
    const int part = lane >> 5, r32 = lane & 31;
    // slice kb from stage `cur`; the loads of slice kb + 2 go to `nxt2` (the
    // stage read by slice kb - 1, which every wave finished before this barrier)
    auto step = [&](F16Stage& cur, F16Stage& nxt2, int kb) {
        if (kb + 1 < KB)  // slice kb + 1's loads may stay in flight
            asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(LA + LB) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory");
        half8 a[TI][4], b[2][4];
#pragma unroll
        for (int t = 0; t < TI; ++t)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) a[t][pl] = cur.a[pl][part][wm * 64 + t * 32 + r32];
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int pl = 0; pl < 4; ++pl) b[t][pl] = cur.b[pl][part][wn * 64 + t * 32 + r32];
        if (kb + 2 < KB) dma(nxt2, kb + 2);
        f16x3_mfma<TI>(a, b, accR, accI);
    };
    dma(st0, 0);
    if (KB > 1) dma(st1, 1);
    for (int kb = 0; kb < KB; kb += 3) {
        step(st0, st2, kb);
        if (kb + 1 < KB) step(st1, st0, kb + 1);
        if (kb + 2 < KB) step(st2, st1, kb + 2);
    }
    f16x3_acc_store<TI>(accR, accI, C, ldc, M, N, i0 + wm * 64, j0 + wn * 64, lane & 31, lane >> 5, alpha / unit);
}

// Sustained rate of the FP16 matrix cores under the split-FP16 GEMM's
// instruction mix: every wave runs the per-slice product of that GEMM's wave
// tile (48 independent-accumulator 32x32x16 FP16 MFMAs with their operand
// negations and unshifts) on register operands, with no memory traffic, so
// the rate is what the matrix cores deliver at the clock the device holds
// under this load. Operands are pseudo-random FP16 of modulus < 1 (the bit
// activity of real split planes). Wave 0 of each workgroup records its shader
// clock cycles (s_memtime) and 100 MHz constant-clock ticks (s_memrealtime).
__global__ __launch_bounds__(512, 1) void mfma_f16_rate_kernel(int iters, float* __restrict__ sink,
                                                               unsigned long long* __restrict__ clocks) {
    const int tid = threadIdx.x;
    unsigned h = 2654435761u * (unsigned)(blockIdx.x * 512 + tid + 1);
    auto next = [&]() {
        h ^= h << 13;
        h ^= h >> 17;
        h ^= h << 5;
        return (_Float16)((float)(h & 0xFFFF) * (1.f / 65536.f) - 0.5f);
    };
    half8 a[2][4], b[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                a[t][q][e] = next();
                b[t][q][e] = next();
            }
    floatx16 accR[2][2], accI[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) accR[i][j][v] = accI[i][j][v] = 0.f;
    const unsigned long long c0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
    for (int it = 0; it < iters; ++it) {
        f16x3_mfma<2>(a, b, accR, accI);
        asm volatile("" ::: "memory");
    }
    const unsigned long long c1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
    float sum = 0.f;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) sum += accR[i][j][v] + accI[i][j][v];
    sink[blockIdx.x * 512 + tid] = sum;
    if (tid == 0) {
        clocks[2 * blockIdx.x] = c1 - c0;
        clocks[2 * blockIdx.x + 1] = r1 - r0;
    }
}

// =================================================================== panel
// Grid barrier over `nblocks` co-resident workgroups: monotone counter.
// Bounded (WaitArg): returns true, uniformly over the workgroup, when the
// barrier gave up; the caller then returns.
__device__ __forceinline__ bool grid_barrier(unsigned int* counter, unsigned int target, const WaitArg& wa) {
    __shared__ int s_stalled;
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        // relaxed polling, then one acquire fence
        const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
        int stalled = 0;
        while (wa.force || __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__builtin_amdgcn_s_memrealtime() - t0 > wa.ticks || wait_error_set(wa.err)) {
                stalled = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (stalled) flag_wait_error(wa.err, kWaitPanel);
        else __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_stalled = stalled;
    }
    __syncthreads();
    return s_stalled != 0;
}

template <typename T>
struct Cand {
    T val;
    int idx;
    int pad;
};

template <typename T>
__device__ __forceinline__ T cabs1(cx<T> a) {
    return fabs(a.re) + fabs(a.im);
}

__device__ __forceinline__ bool better(double v, int i, double bv, int bi) {
    return v > bv || (v == bv && i < bi);
}

// P points at A(k1, k1) (column-major, lda); the panel is M rows x W columns.
// ipiv[j] receives the 0-based panel row swapped with row j.
// pub: [W][G] candidate rows (W values each) ; cand: [W][G] ; drow: [W][W].
// Per column: local argmax, publication of the workgroup's candidate (value,
// row index, row values) and of row j, one grid barrier; then every workgroup
// loads the candidates together with row j, picks the global winner (largest
// |re|+|im|, lowest index on ties), loads its row, swaps and updates its rows.
template <typename T, int W, int BS>
__global__ __launch_bounds__(BS) void panel_getf2_kernel(int M, int w, cx<T>* __restrict__ P, long long lda,
                                                         int* __restrict__ ipiv, int ipiv_off,
                                                         Cand<T>* __restrict__ cand, cx<T>* __restrict__ pub,
                                                         cx<T>* __restrict__ drow, unsigned int* __restrict__ counter,
                                                         unsigned int counter_base, int* __restrict__ info,
                                                         int info_base, WaitArg wa) {
    const int G = gridDim.x;
    const int r = blockIdx.x * BS + threadIdx.x;  // owned row
    const bool own = r < M;
    // identity pivots until each column's choice is written (by thread 0
    // after this workgroup's barriers): a launch that gives up leaves valid ones
    if (blockIdx.x == 0 && threadIdx.x < w) ipiv[threadIdx.x] = threadIdx.x + ipiv_off;
    cx<T> a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = (own && k < w) ? P[(long long)k * lda + r] : cx<T>{0, 0};

    __shared__ T sv[BS / 64];
    __shared__ int si[BS / 64];
    __shared__ int swin;
    __shared__ cx<T> sprow[W];
    __shared__ cx<T> sdrow[W];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

    for (int j = 0; j < w; ++j) {
        // 1. local argmax over rows >= j of column j (static register indexing)
        cx<T> aj = {0, 0};
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k == j) aj = a[k];
        double v = -1.0;
        int idx = 0x7fffffff;
        if (own && r >= j) {
            v = (double)cabs1(aj);
            idx = r;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(v, off);
            const int oi = __shfl_xor(idx, off);
            if (better(ov, oi, v, idx)) { v = ov; idx = oi; }
        }
        if (lane == 0) { sv[wv] = (T)v; si[wv] = idx; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double bv = sv[0];
            int bi = si[0];
            for (int w = 1; w < BS / 64; ++w)
                if (better((double)sv[w], si[w], bv, bi)) { bv = sv[w]; bi = si[w]; }
            swin = bi;
            cand[(size_t)j * G + blockIdx.x] = Cand<T>{(T)bv, bi, 0};
        }
        __syncthreads();
        // publish the local winner's row and, by its owner, row j
        if (own && r == swin) {
#pragma unroll
            for (int k = 0; k < W; ++k) pub[((size_t)j * G + blockIdx.x) * W + k] = a[k];
        }
        if (own && r == j) {
#pragma unroll
            for (int k = 0; k < W; ++k) drow[(size_t)j * W + k] = a[k];
        }
        // 2. grid barrier (a stalled exchange ends the launch; the host reports it)
        if (grid_barrier(counter, counter_base + (unsigned)(G * (j + 1)), wa)) return;
        // 3. global winner (every workgroup computes the same); row j is
        // loaded in the same round trip as the candidates
        if (threadIdx.x < W) sdrow[threadIdx.x] = drow[(size_t)j * W + threadIdx.x];
        if (threadIdx.x < 64) {
            double bv = -1.0;
            int bi = 0x7fffffff, bg = 0;
            for (int g = threadIdx.x; g < G; g += 64) {
                const Cand<T> c = cand[(size_t)j * G + g];
                if (better((double)c.val, c.idx, bv, bi)) { bv = c.val; bi = c.idx; bg = g; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(bi, off);
                const int og = __shfl_xor(bg, off);
                if (better(ov, oi, bv, bi)) { bv = ov; bi = oi; bg = og; }
            }
            if (threadIdx.x == 0) {
                swin = (bv > 0.0) ? (bi | (bg << 20)) : -1;
            }
        }
        __syncthreads();
        const int code = swin;
        if (code < 0) {
            // exactly singular column: record, keep going without a swap (zgetf2 semantics)
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ipiv[j] = j + ipiv_off;
                if (*info == 0) *info = info_base + j + 1;
            }
            continue;
        }
        const int p = code & 0xFFFFF, pg = code >> 20;
        // the pivot row, shared by the workgroup
        if (threadIdx.x < W) sprow[threadIdx.x] = pub[((size_t)j * G + pg) * W + threadIdx.x];
        __syncthreads();
        const cx<T>* prow = sprow;
        if (blockIdx.x == 0 && threadIdx.x == 0) ipiv[j] = p + ipiv_off;
        // swap: position p takes old row j, position j takes the pivot row
        if (own && r == p && p != j) {
#pragma unroll
            for (int k = 0; k < W; ++k) a[k] = sdrow[k];
        }
        if (own && r == j) {
#pragma unroll
            for (int k = 0; k < W; ++k) a[k] = prow[k];
        }
        // 4. scale and rank-1 update of rows below j
        const cx<T> pj = prow[j];
        if (own && r > j) {
            cx<T> ajj = {0, 0};
#pragma unroll
            for (int k = 0; k < W; ++k)
                if (k == j) ajj = a[k];
            const T d = pj.re * pj.re + pj.im * pj.im;
            const cx<T> inv = {pj.re / d, -pj.im / d};
            const cx<T> l = {ajj.re * inv.re - ajj.im * inv.im, ajj.re * inv.im + ajj.im * inv.re};
#pragma unroll
            for (int k = 0; k < W; ++k) {
                if (k == j) a[k] = l;
                if (k > j) {
                    a[k].re -= l.re * prow[k].re - l.im * prow[k].im;
                    a[k].im -= l.re * prow[k].im + l.im * prow[k].re;
                }
            }
        }
        // sprow and sdrow are rewritten after the next grid barrier, which
        // every thread of the workgroup reaches after this column's reads
    }
    if (own) {
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < w) P[(long long)k * lda + r] = a[k];
    }
}

// 64-bit pivot key of a candidate of the FP32 panel: |re| + |im| (a
// non-negative float, whose bit pattern orders as an unsigned integer) above
// the complement of the row index, so that the largest key is the largest
// modulus with the lowest row index (izamax's choice, as better() above).
// Zero means "not yet published": the keys are cleared before each launch and
// every published key has a nonzero low word (row indices are below 2^31).
__device__ __forceinline__ unsigned long long pivot_key(float v, int idx) {
    return ((unsigned long long)__float_as_uint(v) << 32) | (0xFFFFFFFFu - (unsigned)idx);
}

// Largest grid of the keyed panel (one 512-thread workgroup per CU fits its
// registers, so a co-resident grid has at most one workgroup per CU anyway).
constexpr int kPanelKeyedMaxG = 256;

// The block triangular sweeps publish their solution in place: the output
// vector is first filled with kUnsolved words, each solved entry is written
// with agent-scope (device-coherent) 64-bit stores, which workgroups on other
// XCDs read without an L2 writeback or invalidate, and a reader polls the
// entries themselves (an entry is there once none of its 64-bit words is
// kUnsolved), so the values are their own ready signal and no flag or store
// fence is needed. kUnsolved is a NaN pattern (sign and every payload bit
// set) that arithmetic on other operands does not produce; a computed word
// equal to it is published as the default quiet NaN.
constexpr unsigned long long kUnsolved = ~0ull;

template <typename T>
__device__ __forceinline__ unsigned long long published_word(unsigned long long w) {
    constexpr unsigned long long qnan = sizeof(T) == 8 ? 0x7FF8000000000000ull : 0x7FC000007FC00000ull;
    return w == kUnsolved ? qnan : w;
}

template <typename T>
__device__ __forceinline__ void publish_cx(cx<T>* p, cx<T> v) {
    if constexpr (sizeof(T) == 8) {
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(&p->re),
                           published_word<T>(__builtin_bit_cast(unsigned long long, v.re)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(&p->im),
                           published_word<T>(__builtin_bit_cast(unsigned long long, v.im)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __hip_atomic_store(reinterpret_cast<unsigned long long*>(p),
                           published_word<T>(__builtin_bit_cast(unsigned long long, v)), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Bounded wait for one published entry. The sweeps wait only for blocks
// solved earlier in the sweep or by a workgroup of lower index in the same
// launch (dispatched earlier and waiting only on workgroups of even lower
// index), so the wait ends; the bound (seconds) only keeps a broken invariant
// from hanging the device.
template <typename T>
__device__ __forceinline__ cx<T> wait_published(const cx<T>* p, unsigned int* err) {
    unsigned long long w0 = kUnsolved, w1 = kUnsolved;
    bool seen = false;
    for (unsigned it = 0; it < (1u << 22); ++it) {
        if constexpr (sizeof(T) == 8) {
            w0 = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(&p->re), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            w1 = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(&p->im), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        } else {
            w0 = w1 = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(p), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        }
        if (w0 != kUnsolved && w1 != kUnsolved) {
            seen = true;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    if (!seen) flag_wait_error(err, kWaitSweep);
    if constexpr (sizeof(T) == 8) return cx<T>{__builtin_bit_cast(T, w0), __builtin_bit_cast(T, w1)};
    else return __builtin_bit_cast(cx<T>, w0);
}

__global__ void fill_unsolved_kernel(long long nwords, unsigned long long* __restrict__ p) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nwords) p[i] = kUnsolved;
}

// A published row of the FP32 panel: each complex entry as two 64-bit words,
// (re bits | tag << 32) and (im bits | tag << 32), with a tag unique to the
// launch and column. A reader that sees every word carry its tag has the
// whole row, so the writer need not wait for its stores to complete before
// publishing its key: the reader re-reads the (rare) row not yet complete.
// One wave publishes a row that one of its lanes holds (mine): the row goes
// through the wave's LDS stage and each lane stores one word.
template <int W>
__device__ __forceinline__ void publish_row_tagged(float* __restrict__ stage, const cx<float>* a, bool mine,
                                                   unsigned long long* __restrict__ dst, unsigned tag) {
    if (mine) {
#pragma unroll
        for (int k = 0; k < W; ++k) {
            stage[2 * k] = a[k].re;
            stage[2 * k + 1] = a[k].im;
        }
    }
    // the same wave reads what its lane wrote
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int lane = threadIdx.x & 63;
    if (lane < 2 * W)
        __hip_atomic_store(dst + lane, ((unsigned long long)tag << 32) | __float_as_uint(stage[lane]),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Maximum of a 32-bit unsigned value over the wave, returned to every lane:
// DPP steps within rows of 16 lanes (quad permutations, half-row and row
// mirrors), then the broadcasts of lane 15 and lane 31 into the rows above;
// lane 63 ends with the maximum.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    unsigned t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    v = v > t ? v : t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    v = v > t ? v : t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    v = v > t ? v : t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x140, 0xF, 0xF, false);  // row_mirror
    v = v > t ? v : t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v = v > t ? v : t;
    t = (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    v = v > t ? v : t;
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// Wave 0 reads two tagged rows of W <= 32 entries (2W words each, two per
// lane: one round trip for the pivot row and row j), both carrying the same
// tag, into LDS, re-reading until every word carries the tag (bounded: the
// words were stored before the key that led here was, so they arrive; rows
// still incomplete at the bound are reported through *err).
__device__ __forceinline__ void read_rows_tagged(const unsigned long long* __restrict__ src1,
                                                 const unsigned long long* __restrict__ src2, int W, unsigned tag,
                                                 float* __restrict__ out1, float* __restrict__ out2,
                                                 unsigned int* err) {
    const int lane = threadIdx.x & 63;
    unsigned long long v1 = 0, v2 = 0;
    bool complete = false;
    for (unsigned it = 0; it < (1u << 20); ++it) {
        const unsigned long long mark = (unsigned long long)tag << 32;
        v1 = lane < 2 * W ? __hip_atomic_load(src1 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : mark;
        v2 = lane < 2 * W ? __hip_atomic_load(src2 + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : mark;
        if (__all((unsigned)(v1 >> 32) == tag && (unsigned)(v2 >> 32) == tag)) {
            complete = true;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    if (!complete && lane == 0) flag_wait_error(err, kWaitRow);
    if (lane < 2 * W) {
        out1[lane] = __uint_as_float((unsigned)v1);
        out2[lane] = __uint_as_float((unsigned)v2);
    }
}

// The interchanges of rows kb..ke-1 (applied in order; step j interchanges
// rows kb + j and sp[j] > kb + j, sp[j] < 0: none), at most L of them, as one
// permutation composed by the NT threads of a workgroup in LDS:
//   last(j): the last earlier step whose pivot row is kb + j;
//   prev(j): the last earlier step with the same pivot row as step j;
//   A(j) = the original row found at kb + j just before step j
//        = A(last(j)) if last(j) exists, else kb + j (pointer jumping);
// row kb + j then ends with A(prev(j)) if prev(j) exists, else p_j (step j
// interchanges) or with A(j) (no interchange at j), and a pivot row p >= ke
// ends with A(j) of the last step that chose it. Writes the moves (row
// mpos[i] receives the entry of row msrc[i], i < *nmv; *nmv cleared by the
// caller before the first barrier here). last, prv, par, fin: L words of LDS.
template <int L, int NT>
__device__ void compose_interchanges(const int* sp, int nL, int kb, int ke, int* last, int* prv, int* par, int* fin,
                                     int* mpos, int* msrc, int* nmv) {
    const int tid = threadIdx.x;
    for (int j = tid; j < L; j += NT) {
        last[j] = -1;
        fin[j] = 1;
    }
    __syncthreads();
    for (int j = tid; j < nL; j += NT) {
        const int p = sp[j];
        int pr = -1;
        if (p >= 0) {
            if (p < ke) atomicMax(&last[p - kb], j);
            for (int i = 0; i < j; ++i)
                if (sp[i] == p) pr = i;
        }
        prv[j] = pr;
    }
    __syncthreads();
    for (int j = tid; j < nL; j += NT) {
        par[j] = last[j] >= 0 ? last[j] : j;
        if (prv[j] >= 0) fin[prv[j]] = 0;  // a later step chose the same pivot row
    }
    __syncthreads();
    for (int span = 1; span < L; span *= 2) {
        int np[(L + NT - 1) / NT];
#pragma unroll
        for (int u = 0; u < (L + NT - 1) / NT; ++u) {
            const int j = tid + NT * u;
            np[u] = j < nL ? par[par[j]] : 0;
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < (L + NT - 1) / NT; ++u) {
            const int j = tid + NT * u;
            if (j < nL) par[j] = np[u];
        }
        __syncthreads();
    }
    for (int j = tid; j < nL; j += NT) {
        const int p = sp[j];
        const int src = p < 0 ? kb + par[j] : prv[j] >= 0 ? kb + par[prv[j]] : p;
        if (src != kb + j) {
            const int t = atomicAdd(nmv, 1);
            mpos[t] = kb + j;
            msrc[t] = src;
        }
        if (p >= ke && fin[j]) {
            const int t = atomicAdd(nmv, 1);
            mpos[t] = p;
            msrc[t] = kb + par[j];
        }
    }
}

// The FP32 panel with the candidates as keys: per column, each wave finds its
// candidate (DPP maximum), the workgroup's candidate is the best of its waves;
// each workgroup publishes its candidate row (tagged words), then its key into
// the column's own slot, and the owner of row j publishes row j. Wave 0 polls
// the G keys of the column until all are published (this wait is the
// grid-wide synchronisation: no counter), takes the largest and reads that
// row and row j in one round trip; the tags prove them complete. Rows, keys and
// their reads are agent-scope (device-coherent) accesses, so no L2 writeback
// or invalidate is needed per column. Slots are per column and tags per
// launch, so no stale word is taken for a current one. Three workgroup
// barriers per column. Requires all G workgroups co-resident (the launch
// forms: g_cooperative_panels); the waits are bounded either way.
template <int W, int BS>
__global__ __launch_bounds__(BS) void panel_getf2_keyed_kernel(int M, int w, cx<float>* __restrict__ P,
                                                               long long lda, int* __restrict__ ipiv, int ipiv_off,
                                                               unsigned long long* __restrict__ keys,
                                                               unsigned long long* __restrict__ keys_next, int nkeys,
                                                               unsigned long long* __restrict__ pub,
                                                               unsigned long long* __restrict__ drow,
                                                               unsigned tag_base, int* __restrict__ info,
                                                               int info_base, WaitArg wa, int sc0, int snc, int sc1,
                                                               int snc1) {
    static_assert(W % 8 == 0 && W <= 32, "columns are handled in groups of 8, a row read by one wave");
    constexpr int kStalled = -2;  // swin code: the key exchange gave up (-1: singular column)
    constexpr int NG = W / 8;
    const int G = gridDim.x;
    // the key slots of the next launch (those of the previous one, which has
    // completed) start cleared
    for (int t = blockIdx.x * BS + threadIdx.x; t < nkeys; t += G * BS) keys_next[t] = 0ull;
    const int r = blockIdx.x * BS + threadIdx.x;  // owned row
    const bool own = r < M;
    // identity pivots until each column's choice is written (by thread 0
    // after this workgroup's barriers): a launch that gives up leaves valid ones
    if (blockIdx.x == 0 && threadIdx.x < w) ipiv[threadIdx.x] = threadIdx.x + ipiv_off;
    cx<float> a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = (own && k < w) ? P[(long long)k * lda + r] : cx<float>{0, 0};

    __shared__ unsigned long long skey[BS / 64];
    __shared__ int swin;
    __shared__ cx<float> sprow[W];
    __shared__ cx<float> sdrow[W];
    __shared__ float stage_c[2 * W], stage_j[2 * W];  // rows on their way out
    __shared__ int spv[W];  // the pivot row of each column (-1: no interchange)
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;

    for (int j = 0; j < w; ++j) {
        const unsigned tag = tag_base | (unsigned)(j + 1);
        const int jg = j >> 3;  // group of 8 columns holding column j (uniform)
        // 1. each wave's candidate over rows >= j of column j (static register
        // indexing within the group of column j): the largest |re| + |im| as
        // its bits + 1 (0: no row >= j in the wave), the lowest row among equals
        cx<float> aj = {0, 0};
#pragma unroll
        for (int g = 0; g < NG; ++g)
            if (g == jg) {
#pragma unroll
                for (int kk = 0; kk < 8; ++kk)
                    if (8 * g + kk == j) aj = a[8 * g + kk];
            }
        const unsigned enc = (own && r >= j) ? __float_as_uint(cabs1(aj)) + 1u : 0u;
        const unsigned m = wave_max_u32(enc);
        const int L = __ffsll((long long)__ballot(enc == m)) - 1;
        if (lane == 0)
            skey[wv] = m != 0u ? pivot_key(__uint_as_float(m - 1u), blockIdx.x * BS + wv * 64 + L)
                               : pivot_key(0.f, 0x7fffffff);
        __syncthreads();
        // 2. every wave takes the workgroup's candidate (the largest key);
        // thread 0 publishes the key, the holders of the candidate row and of
        // row j publish those rows
        unsigned long long b = skey[0];
#pragma unroll
        for (int q = 1; q < BS / 64; ++q) b = skey[q] > b ? skey[q] : b;
        if (threadIdx.x == 0)
            __hip_atomic_store(&keys[(size_t)j * G + blockIdx.x], b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // (measured: polling that starts while the other waves publish rows
        // makes the column 0.25 us slower)
        __syncthreads();
        const int lw = (int)(0xFFFFFFFFu - (unsigned)(b & 0xFFFFFFFFu));
        const bool mc = own && r == lw, mj = own && r == j;
        if (__any(mc)) publish_row_tagged<W>(stage_c, a, mc, pub + ((size_t)j * G + blockIdx.x) * 2 * W, tag);
        if (__any(mj)) publish_row_tagged<W>(stage_j, a, mj, drow + (size_t)j * 2 * W, tag);
        // 3. wave 0 waits for every workgroup's key of column j, takes the
        // largest and reads that row together with row j.
        // G <= kPanelKeyedMaxG: each lane polls its (up to) 4 keys with the
        // loads of one round in flight together
        if (wv == 0) {
            unsigned long long best = 0;
            int bg = 0;
            bool stalled = false;
            const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
            for (;;) {
                unsigned long long kv[kPanelKeyedMaxG / 64];
#pragma unroll
                for (int u = 0; u < kPanelKeyedMaxG / 64; ++u) {
                    const int g = lane + 64 * u;
                    // 1: a ready key that loses to every published one
                    kv[u] = g < G ? __hip_atomic_load(&keys[(size_t)j * G + g], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 1ull;
                }
                bool ready = true;
                best = 0;
#pragma unroll
                for (int u = 0; u < kPanelKeyedMaxG / 64; ++u) {
                    ready = ready && kv[u] != 0;
                    if (kv[u] > best) { best = kv[u]; bg = lane + 64 * u; }
                }
                if (__all(ready) && !wa.force) break;
                // the clock read is scalar and the error word is read by the
                // whole wave: the exit is uniform
                if (__builtin_amdgcn_s_memrealtime() - t0 > wa.ticks || __any(wait_error_set(wa.err))) {
                    stalled = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            // the largest key of the wave: its high word, then the low word
            // among the lanes holding that high word (keys with a nonzero
            // high word are distinct)
            const unsigned hi = (unsigned)(best >> 32), lo = (unsigned)best;
            const unsigned mh = wave_max_u32(hi);
            const unsigned ml = wave_max_u32(hi == mh ? lo : 0u);
            const int lb = __ffsll((long long)__ballot(hi == mh && lo == ml)) - 1;
            bg = __builtin_amdgcn_readlane(bg, lb);
            const int bi = (int)(0xFFFFFFFFu - ml);
            int code = mh != 0u ? (bi | (bg << 20)) : -1;
            if (stalled) {
                if (lane == 0) flag_wait_error(wa.err, kWaitPanel);
                code = kStalled;
            }
            if (code >= 0)
                read_rows_tagged(pub + ((size_t)j * G + bg) * 2 * W, drow + (size_t)j * 2 * W, W, tag,
                                 (float*)sprow, (float*)sdrow, wa.err);
            if (lane == 0) swin = code;
        }
        __syncthreads();
        const int code = swin;
        if (code == kStalled) return;  // uniform: the host reports the stalled exchange
        if (code < 0) {
            // exactly singular column: record, keep going without a swap (zgetf2 semantics)
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                ipiv[j] = j + ipiv_off;
                if (*info == 0) *info = info_base + j + 1;
            }
            if (threadIdx.x == 0) spv[j] = -1;
            continue;
        }
        const int p = code & 0xFFFFF;
        if (blockIdx.x == 0 && threadIdx.x == 0) ipiv[j] = p + ipiv_off;
        if (threadIdx.x == 0) spv[j] = p > j ? p : -1;
        // swap: position p takes old row j, position j takes the pivot row
        if (own && r == p && p != j) {
#pragma unroll
            for (int k = 0; k < W; ++k) a[k] = sdrow[k];
        }
        if (own && r == j) {
#pragma unroll
            for (int k = 0; k < W; ++k) a[k] = sprow[k];
        }
        // 4. scale and rank-1 update of rows below j, over the groups of
        // columns >= j only
        const cx<float> pj = sprow[j];
        if (own && r > j) {
            cx<float> ajj = {0, 0};
#pragma unroll
            for (int g = 0; g < NG; ++g)
                if (g == jg) {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk)
                        if (8 * g + kk == j) ajj = a[8 * g + kk];
                }
            const float d = pj.re * pj.re + pj.im * pj.im;
            const cx<float> inv = {pj.re / d, -pj.im / d};
            const cx<float> l = {ajj.re * inv.re - ajj.im * inv.im, ajj.re * inv.im + ajj.im * inv.re};
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                if (g > jg) {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) {
                        const int k = 8 * g + kk;
                        a[k].re -= l.re * sprow[k].re - l.im * sprow[k].im;
                        a[k].im -= l.re * sprow[k].im + l.im * sprow[k].re;
                    }
                } else if (g == jg) {
#pragma unroll
                    for (int kk = 0; kk < 8; ++kk) {
                        const int k = 8 * g + kk;
                        if (k == j) a[k] = l;
                        if (k > j) {
                            a[k].re -= l.re * sprow[k].re - l.im * sprow[k].im;
                            a[k].im -= l.re * sprow[k].im + l.im * sprow[k].re;
                        }
                    }
                }
            }
        }
        // skey, swin, sprow and sdrow are rewritten in the next column only
        // after __syncthreads that every thread reaches after these reads
    }
    if (own) {
#pragma unroll
        for (int k = 0; k < W; ++k)
            if (k < w) P[(long long)k * lda + r] = a[k];
    }
    // The panel's interchanges on the other columns of its NB block: column
    // offsets (from P) sc0 .. sc0 + snc - 1 and sc1 .. sc1 + snc1 - 1, rows of
    // the panel (local 0 .. M-1). Every workgroup composes them (spv is the
    // same in all) and takes every G-th group of BS/64 columns, one column
    // per wave, one move per lane: all of a column's moved entries are loaded
    // before any is stored. The panel's own columns are not among them.
    const int ns = snc + snc1;
    if (ns > 0) {
        __shared__ int last[W], prv[W], par[W], fin[W], mpos[2 * W], msrc[2 * W];
        __shared__ int nmv;
        if (threadIdx.x == 0) nmv = 0;
        compose_interchanges<W, BS>(spv, w, 0, w, last, prv, par, fin, mpos, msrc, &nmv);
        __syncthreads();
        const int nm = nmv;  // <= 2W = 64: one lane each
        for (int t0 = blockIdx.x * (BS / 64); t0 < ns; t0 += G * (BS / 64)) {
            const int t = t0 + wv;
            const bool live = t < ns && lane < nm;
            cx<float>* col = P + (long long)(t < snc ? sc0 + t : sc1 + (t - snc)) * lda;
            cx<float> v = {0, 0};
            if (live) v = col[msrc[lane]];
            __builtin_amdgcn_wave_barrier();  // one wave owns the column: its loads precede its stores
            if (live) col[mpos[lane]] = v;
        }
    }
}

// ============================================================== row swaps
// Apply ipiv[k] (0-based, relative to row offset r0) for k in [k_begin, k_end)
// in order to columns [c0, c0+nc) and [c1, c1+nc1) of A (the second range may
// be empty). One lane per column keeps the sequence of swaps in order.
// Pivots outside [k, rows) (never produced by a completed factorisation) are
// skipped, so that incomplete factors cannot move memory outside the matrix.
template <typename T>
__global__ void laswp_kernel(cx<T>* __restrict__ A, long long lda, int rows, int c0, int nc, int c1, int nc1,
                             const int* __restrict__ ipiv, int k_begin, int k_end, int r0, int reverse) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc + nc1) return;
    cx<T>* col = A + (long long)(c < nc ? c0 + c : c1 + (c - nc)) * lda + r0;
    if (!reverse) {
        for (int k = k_begin; k < k_end; ++k) {
            const int p = ipiv[k];
            if (p != k && p > k && p < rows) { const cx<T> t = col[k]; col[k] = col[p]; col[p] = t; }
        }
    } else {
        for (int k = k_end - 1; k >= k_begin; --k) {
            const int p = ipiv[k];
            if (p != k && p > k && p < rows) { const cx<T> t = col[k]; col[k] = col[p]; col[p] = t; }
        }
    }
}

// The interchanges ipiv[kb..ke) (p > k), at most L, applied as one permutation
// (compose_interchanges, in every workgroup) to the columns [c0, c0+nc) and
// [c1, c1+nc1): each column's moved entries are all loaded before any is
// stored. cols: columns per workgroup. reverse: the interchanges in the
// opposite order (the inverse permutation: every move taken backwards).
template <typename T, int L>
__global__ __launch_bounds__(256) void laswp_perm_kernel(cx<T>* __restrict__ A, long long lda, int rows, int c0,
                                                         int nc, int c1, int nc1, const int* __restrict__ ipiv,
                                                         int kb, int ke, int cols, int reverse = 0) {
    constexpr int MV = 2 * L;              // moved rows at most
    constexpr int PER = (MV + 255) / 256;  // moves per thread
    // columns per batch: a thread keeps CB x PER complex entries in flight,
    // 32 registers (an FP32 entry takes 2, an FP64 entry 4)
    constexpr int REGS = PER * (int)sizeof(T) / 2;  // registers per column
    constexpr int CB = 32 / REGS > 0 ? 32 / REGS : 1;
    __shared__ int sp[L], last[L], prv[L], par[L], fin[L];
    __shared__ int mpos[MV], msrc[MV];
    __shared__ int nmv;
    const int tid = threadIdx.x, nL = ke - kb;
    if (tid == 0) nmv = 0;
    for (int j = tid; j < L; j += 256) {
        int p = -1;
        if (j < nL) {
            p = ipiv[kb + j];
            if (!(p > kb + j && p < rows)) p = -1;
        }
        sp[j] = p;
    }
    compose_interchanges<L, 256>(sp, nL, kb, ke, last, prv, par, fin, mpos, msrc, &nmv);
    __syncthreads();
    const int nm = nmv;
    const int* from = reverse ? mpos : msrc;
    const int* to = reverse ? msrc : mpos;
    const int cbeg = blockIdx.x * cols, cend = min(cbeg + cols, nc + nc1);
    for (int cb = cbeg; cb < cend; cb += CB) {
        cx<T> v[CB][PER];
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = cb + b;
            cx<T>* col = A + (long long)(c < nc ? c0 + c : c1 + (c - nc)) * lda;
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = tid + 256 * u;
                v[b][u] = (c < cend && i < nm) ? col[from[i]] : cx<T>{0, 0};
            }
        }
        __syncthreads();  // every entry of these columns read before any is written
#pragma unroll
        for (int b = 0; b < CB; ++b) {
            const int c = cb + b;
            cx<T>* col = A + (long long)(c < nc ? c0 + c : c1 + (c - nc)) * lda;
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int i = tid + 256 * u;
                if (c < cend && i < nm) col[to[i]] = v[b][u];
            }
        }
    }
}

// ============================================== small triangular solves
// In-place solve of a w x w (w <= WM) triangular system on ncols right-hand columns:
//   uplo 0: unit lower L (forward), 1: non-unit upper U (backward),
//   2: conj-transposed non-unit upper U^H (forward), 3: conj-transposed unit lower L^H (backward).
// The triangle T lives at A (column-major, lda); B at B (ldb). One wave per
// 64 columns, one lane per column: the columns' leading WM rows are loaded
// and stored coalesced (32 lanes per 256-byte column segment) through LDS, and
// each lane solves its column in registers (fully unrolled, static indices).
template <typename T, int WM, int UPLO>
__global__ __launch_bounds__(64) void trsm_small_kernel(int w, const cx<T>* __restrict__ Tm, long long lda,
                                                        cx<T>* __restrict__ B, long long ldb, int ncols) {
    static_assert(WM == 32, "two columns per coalesced load of a wave");
    constexpr int uplo = UPLO;
    constexpr int TL = WM * WM / 64;  // triangle entries per lane
    __shared__ cx<T> sT[WM * WM];     // column-major w x w
    __shared__ cx<T> sB[64][WM + 1];  // the wave's columns, padded
    const int lane = threadIdx.x;
    const long long c0 = (long long)blockIdx.x * 64;
    {
        // the triangle in one round trip: every load issued before any LDS store
        cx<T> v[TL];
#pragma unroll
        for (int q = 0; q < TL; ++q) {
            const int t = q * 64 + lane, col = t / WM, row = t % WM;
            v[q] = (col < w && row < w) ? Tm[(long long)col * lda + row] : cx<T>{0, 0};
        }
#pragma unroll
        for (int q = 0; q < TL; ++q) sT[q * 64 + lane] = v[q];
    }
    {
        // the 64 columns: load q covers columns 2q and 2q + 1
        cx<T> u[WM];
#pragma unroll
        for (int q = 0; q < WM; ++q) {
            const int col = 2 * q + (lane >> 5), row = lane & 31;
            u[q] = (c0 + col < ncols && row < w) ? B[(c0 + col) * ldb + row] : cx<T>{0, 0};
        }
#pragma unroll
        for (int q = 0; q < WM; ++q) sB[2 * q + (lane >> 5)][lane & 31] = u[q];
    }
    __syncthreads();
    auto at = [&](int i, int j) { return sT[j * WM + i]; };
    cx<T> x[WM];
#pragma unroll
    for (int i = 0; i < WM; ++i) x[i] = sB[lane][i];
    auto msub = [](cx<T>& s, cx<T> a, cx<T> v) {  // s -= a v
        s.re -= a.re * v.re - a.im * v.im;
        s.im -= a.re * v.im + a.im * v.re;
    };
    auto msubc = [](cx<T>& s, cx<T> a, cx<T> v) {  // s -= conj(a) v
        s.re -= a.re * v.re + a.im * v.im;
        s.im -= a.re * v.im - a.im * v.re;
    };
    auto divd = [](cx<T> v, cx<T> d) {
        const T q = d.re * d.re + d.im * d.im;
        return cx<T>{(v.re * d.re + v.im * d.im) / q, (v.im * d.re - v.re * d.im) / q};
    };
    if constexpr (uplo == 0) {
#pragma unroll
        for (int i = 1; i < WM; ++i)
#pragma unroll
            for (int k = 0; k < i; ++k)
                if (i < w) msub(x[i], at(i, k), x[k]);
    } else if constexpr (uplo == 1) {
#pragma unroll
        for (int i = WM - 1; i >= 0; --i) {
            if (i < w) {
#pragma unroll
                for (int k = i + 1; k < WM; ++k)
                    if (k < w) msub(x[i], at(i, k), x[k]);
                x[i] = divd(x[i], at(i, i));
            }
        }
    } else if constexpr (uplo == 2) {  // U^H lower: (U^H)(i,k) = conj(U(k,i))
#pragma unroll
        for (int i = 0; i < WM; ++i) {
            if (i < w) {
#pragma unroll
                for (int k = 0; k < i; ++k) msubc(x[i], at(k, i), x[k]);
                const cx<T> d = {at(i, i).re, -at(i, i).im};
                x[i] = divd(x[i], d);
            }
        }
    } else {  // L^H upper unit: (L^H)(i,k) = conj(L(k,i))
#pragma unroll
        for (int i = WM - 1; i >= 0; --i)
#pragma unroll
            for (int k = i + 1; k < WM; ++k)
                if (i < w && k < w) msubc(x[i], at(k, i), x[k]);
    }
#pragma unroll
    for (int i = 0; i < WM; ++i) sB[lane][i] = x[i];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < WM; ++q) {
        const int col = 2 * q + (lane >> 5), row = lane & 31;
        if (c0 + col < ncols && row < w) B[(c0 + col) * ldb + row] = sB[2 * q + (lane >> 5)][lane & 31];
    }
}

// In-place solve of a w x w (w <= WM) triangular system on ncols right-hand columns:
//   uplo 0: unit lower L (forward), 1: non-unit upper U (backward),
//   2: conj-transposed non-unit upper U^H (forward), 3: conj-transposed unit lower L^H (backward).
// The triangle T lives at A (column-major, lda); B at B (ldb). This form,
// for few columns: WM lanes per column, lane i owning row i; at step k the
// solved x_k is broadcast within the lane group and every lane below
// (forward) or above (backward) subtracts its multiple, so a column costs WM
// dependent steps of one complex FMA (latency) instead of WM^2 / 2 serial
// FMAs in one lane. trsm_small_kernel (one lane per column) does less LDS and
// shuffle work per column and wins when the columns fill the device.
template <typename T, int WM, int UPLO>
__global__ __launch_bounds__(256) void trsm_lanes_kernel(int w, const cx<T>* __restrict__ Tm, long long lda,
                                                         cx<T>* __restrict__ B, long long ldb, int ncols) {
    constexpr int TL = WM * WM / 256;  // triangle entries per thread for the LDS load
    constexpr int CPB = 256 / WM;      // columns per workgroup
    __shared__ cx<T> sT[WM * WM];      // column-major w x w, zero-padded
    {
        cx<T> v[TL];
#pragma unroll
        for (int q = 0; q < TL; ++q) {
            const int t = q * 256 + threadIdx.x, col = t / WM, row = t % WM;
            v[q] = (col < w && row < w) ? Tm[(long long)col * lda + row] : cx<T>{0, 0};
        }
#pragma unroll
        for (int q = 0; q < TL; ++q) sT[q * 256 + threadIdx.x] = v[q];
    }
    __syncthreads();
    const int i = threadIdx.x % WM;
    const int c = blockIdx.x * CPB + threadIdx.x / WM;
    // the multipliers this lane applies at step k: row i of the operator
    //   0: L(i, k)   1: U(i, k)   2: conj(U(k, i))   3: conj(L(k, i))
    cx<T> m[WM];
#pragma unroll
    for (int k = 0; k < WM; ++k) {
        const cx<T> t = (UPLO < 2) ? sT[k * WM + i] : sT[i * WM + k];
        m[k] = (UPLO < 2) ? t : cx<T>{t.re, -t.im};
    }
    cx<T>* b = B + (long long)c * ldb;
    const bool live = c < ncols && i < w;
    cx<T> x = live ? b[i] : cx<T>{0, 0};
    auto divd = [](cx<T> v, cx<T> d) {
        const T q = d.re * d.re + d.im * d.im;
        return cx<T>{(v.re * d.re + v.im * d.im) / q, (v.im * d.re - v.re * d.im) / q};
    };
    // (steps k >= w are skipped by a uniform test, not a break, so that the
    // loop unrolls and m[] stays in registers)
    if constexpr (UPLO == 0 || UPLO == 2) {  // forward
#pragma unroll
        for (int k = 0; k < WM; ++k) {
            if (k < w) {
                if (UPLO == 2 && i == k) x = divd(x, m[k]);  // diagonal conj(U(k, k))
                const T xr = __shfl(x.re, k, WM), xi = __shfl(x.im, k, WM);
                if (i > k) {
                    x.re -= m[k].re * xr - m[k].im * xi;
                    x.im -= m[k].re * xi + m[k].im * xr;
                }
            }
        }
    } else {  // backward
#pragma unroll
        for (int k = WM - 1; k >= 0; --k) {
            if (k < w) {
                if (UPLO == 1 && i == k) x = divd(x, m[k]);  // diagonal U(k, k)
                const T xr = __shfl(x.re, k, WM), xi = __shfl(x.im, k, WM);
                if (i < k) {
                    x.re -= m[k].re * xr - m[k].im * xi;
                    x.im -= m[k].re * xi + m[k].im * xr;
                }
            }
        }
    }
    if (live) b[i] = x;
}

// ================================================================= GEMV
// Streaming (nontemporal) load of one complex entry of an operator read once.
typedef double double2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cx<double> ld_stream(const cx<double>* p) {
    const double2v v = __builtin_nontemporal_load(reinterpret_cast<const double2v*>(p));
    return {v[0], v[1]};
}
__device__ __forceinline__ cx<float> ld_stream(const cx<float>* p) {
    const float2v v = __builtin_nontemporal_load(reinterpret_cast<const float2v*>(p));
    return {v[0], v[1]};
}

// Partial y_part[chunk][i] = sum_{j in chunk} A(i, j) x(j), op N.
template <typename T, typename TX>
__global__ void gemv_n_partial(int m, int n, int chunk, const cx<T>* __restrict__ A, long long lda,
                               const cx<TX>* __restrict__ x, cx<double>* __restrict__ part) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int j0 = c * chunk, j1 = min(n, j0 + chunk);
    if (i >= m) return;
    // eight column loads in flight per lane, two independent accumulator
    // pairs (even / odd columns) combined at the end in a fixed order
    double sr0 = 0, si0 = 0, sr1 = 0, si1 = 0;
    const cx<T>* a = A + (long long)j0 * lda + i;
    int j = j0;
    for (; j + 8 <= j1; j += 8, a += 8 * lda) {
        cx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = ld_stream(a + u * lda);
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            const cx<TX> x0 = x[j + u], x1 = x[j + u + 1];
            sr0 += (double)v[u].re * x0.re - (double)v[u].im * x0.im;
            si0 += (double)v[u].re * x0.im + (double)v[u].im * x0.re;
            sr1 += (double)v[u + 1].re * x1.re - (double)v[u + 1].im * x1.im;
            si1 += (double)v[u + 1].re * x1.im + (double)v[u + 1].im * x1.re;
        }
    }
    for (; j < j1; ++j, a += lda) {
        const cx<T> v = *a;
        const cx<TX> xv = x[j];
        sr0 += (double)v.re * xv.re - (double)v.im * xv.im;
        si0 += (double)v.re * xv.im + (double)v.im * xv.re;
    }
    part[(size_t)c * m + i] = {sr0 + sr1, si0 + si1};
}

// op C: y_part[chunk][j] = sum_{i in chunk} conj(A(i, j)) x(i); one wave per column.
template <typename T, typename TX>
__global__ void gemv_c_partial(int m, int n, int chunk, const cx<T>* __restrict__ A, long long lda,
                               const cx<TX>* __restrict__ x, cx<double>* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int c = blockIdx.y;
    if (j >= n) return;
    const int i0 = c * chunk, i1 = min(m, i0 + chunk);
    // four 1 KB row segments in flight per wave, accumulated in order
    double sr = 0, si = 0;
    const cx<T>* a = A + (long long)j * lda;
    int i = i0 + lane;
    for (; i + 192 < i1; i += 256) {
        cx<T> v[4];
        cx<TX> xv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            v[u] = ld_stream(a + i + 64 * u);
            xv[u] = x[i + 64 * u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sr += (double)v[u].re * xv[u].re + (double)v[u].im * xv[u].im;
            si += (double)v[u].re * xv[u].im - (double)v[u].im * xv[u].re;
        }
    }
    for (; i < i1; i += 64) {
        const cx<T> v = a[i];
        const cx<TX> xv = x[i];
        sr += (double)v.re * xv.re + (double)v.im * xv.im;
        si += (double)v.re * xv.im - (double)v.im * xv.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if (lane == 0) part[(size_t)c * n + j] = {sr, si};
}

// y = beta*y + alpha*sum_chunks part (y of type TY)
template <typename TY>
__global__ void gemv_finish(int m, int nchunks, const cx<double>* __restrict__ part, double alpha, double beta,
                            cx<TY>* __restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double sr = 0, si = 0;
    for (int c = 0; c < nchunks; ++c) {
        const cx<double> p = part[(size_t)c * m + i];
        sr += p.re;
        si += p.im;
    }
    cx<TY> o = y[i];
    const double yr = beta != 0.0 ? (double)o.re : 0.0, yi = beta != 0.0 ? (double)o.im : 0.0;
    o.re = (TY)(beta * yr + alpha * sr);
    o.im = (TY)(beta * yi + alpha * si);
    y[i] = o;
}

// ============================================================= utilities
// Grid-stride: the dispatch must stay below 2^32 work-items.
template <typename TI, typename TO>
__global__ void convert_kernel(long long rows, long long cols, const cx<TI>* __restrict__ A, long long lda,
                               cx<TO>* __restrict__ B, long long ldb, double scale) {
    const long long tot = rows * cols;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / rows, r = t - c * rows;
        const cx<TI> v = A[c * lda + r];
        B[c * ldb + r] = {(TO)(v.re * scale), (TO)(v.im * scale)};
    }
}

// B = scale A (FP64 -> TO) with, per workgroup, the partial sum of |A|^2: the
// Frobenius norm of the operator (the matrix norm of the refinement's
// backward-error test) in the same pass over A. Fixed grid, fixed order:
// deterministic.
template <typename TO>
__global__ __launch_bounds__(256) void convert_fro_kernel(long long rows, long long cols, const c64* __restrict__ A,
                                                          long long lda, cx<TO>* __restrict__ B, long long ldb,
                                                          double scale, double* __restrict__ part) {
    __shared__ double s[256];
    const long long tot = rows * cols;
    double a = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / rows, r = t - c * rows;
        const c64 v = A[c * lda + r];
        a += v.re * v.re + v.im * v.im;
        B[c * ldb + r] = {(TO)(v.re * scale), (TO)(v.im * scale)};
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// partial sums of |x|^2 per block (deterministic two-stage)
__global__ void sumsq_partial(long long n, const c64* __restrict__ x, double* __restrict__ part) {
    __shared__ double s[256];
    double a = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        a += x[i].re * x[i].re + x[i].im * x[i].im;
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// x += s * d (d in T), d optionally FP32
template <typename T>
__global__ void axpy_kernel(long long n, double s, const cx<T>* __restrict__ d, c64* __restrict__ x) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i].re += s * (double)d[i].re;
    x[i].im += s * (double)d[i].im;
}

// Per-column sums of squares of an n x nrhs block (one workgroup per column,
// fixed-order tree reduction).
__global__ void colsumsq_kernel(long long n, const c64* __restrict__ R, long long ldr, double* __restrict__ out) {
    __shared__ double s[256];
    const c64* col = R + (long long)blockIdx.x * ldr;
    double a = 0;
    for (long long i = threadIdx.x; i < n; i += 256) a += col[i].re * col[i].re + col[i].im * col[i].im;
    s[threadIdx.x] = a;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

// D32[:, c] = R[:, c] * sc[c]  (FP64 -> FP32, per-column scale)
__global__ void scale_cols_f32_kernel(long long n, int nrhs, const c64* __restrict__ R, long long ldr,
                                      const double* __restrict__ sc, c32* __restrict__ D, long long ldd) {
    const long long tot = n * nrhs;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / n, r = t - c * n;
        const c64 v = R[c * ldr + r];
        D[c * ldd + r] = {(float)(v.re * sc[c]), (float)(v.im * sc[c])};
    }
}

// X[:, c] += D[:, c] * g[c]  (FP32 correction into the FP64 iterate)
__global__ void axpy_cols_kernel(long long n, int nrhs, const c32* __restrict__ D, long long ldd,
                                 const double* __restrict__ g, c64* __restrict__ X, long long ldx) {
    const long long tot = n * nrhs;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / n, r = t - c * n;
        X[c * ldx + r].re += g[c] * (double)D[c * ldd + r].re;
        X[c * ldx + r].im += g[c] * (double)D[c * ldd + r].im;
    }
}

// r = b - r  (r holds Z x on entry)
__global__ void residual_kernel(long long n, const c64* __restrict__ b, c64* __restrict__ r) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = {b[i].re - r[i].re, b[i].im - r[i].im};
}

// ================================================ single-RHS triangular solves
// Blocked sweeps with TB-wide diagonal blocks whose inverses are formed once
// per solve; each step is one launch that applies the block inverse to x_k
// and updates every remaining row (memory-bound on the factor).
constexpr int TB = 64;

// inv[blk] (TB x TB, column-major) = inverse of diagonal block blk of the
// unit-lower L (which = 0) or of the upper U (which = 1). One lane per column.
// Inverse of every B x B diagonal block of the unit lower (which = 0) or upper
// (which = 1) factor, one thread per column; inv[blk][c B + i] = X(i, c).
// Partial last blocks are padded with the identity.
// Factors in TF, inverse computed and stored in T (T = double over FP32
// factors gives the exact inverse of the FP32 blocks to FP64 rounding).
template <typename TF, typename T, int B>
__global__ void diag_inv_kernel(int n, const cx<TF>* __restrict__ LU, long long lda, int which,
                                cx<T>* __restrict__ inv) {
    const int blk = blockIdx.x, c = threadIdx.x;
    const int k0 = blk * B;
    const int w = min(B, n - k0);
    cx<T>* X = inv + (size_t)blk * B * B;
    if (c >= B) return;
    auto A = [&](int i, int j) {
        const cx<TF> a = LU[(long long)(k0 + j) * lda + k0 + i];
        return cx<T>{(T)a.re, (T)a.im};
    };
    for (int i = 0; i < B; ++i) X[c * B + i] = {0, 0};
    if (c >= w) {
        X[c * B + c] = {1, 0};
        return;
    }
    if (which == 0) {
        X[c * B + c] = {1, 0};
        for (int i = c + 1; i < w; ++i) {
            T sr = 0, si = 0;
            for (int j = c; j < i; ++j) {
                const cx<T> a = A(i, j), x = X[c * B + j];
                sr -= a.re * x.re - a.im * x.im;
                si -= a.re * x.im + a.im * x.re;
            }
            X[c * B + i] = {sr, si};
        }
    } else {
        auto divd = [](cx<T> x, cx<T> d) {
            const T q = d.re * d.re + d.im * d.im;
            return cx<T>{(x.re * d.re + x.im * d.im) / q, (x.im * d.re - x.re * d.im) / q};
        };
        X[c * B + c] = divd({1, 0}, A(c, c));
        for (int i = c - 1; i >= 0; --i) {
            T sr = 0, si = 0;
            for (int j = i + 1; j <= c; ++j) {
                const cx<T> a = A(i, j), x = X[c * B + j];
                sr -= a.re * x.re - a.im * x.im;
                si -= a.re * x.im + a.im * x.re;
            }
            X[c * B + i] = divd({sr, si}, A(i, i));
        }
    }
}

// Sweep step of a single-RHS block triangular solve with B-row diagonal
// blocks, with the diagonal solve of the next block fused in. Y holds the
// solution of block k (columns [k0, k0 + w)). Every workgroup updates B rows
// rb.. of vec in [r0, r1):
//   op N: vec[r] -= sum_c A(r, k0 + c) y_c        (A(r, c) = LU[c lda + r])
//   op C: vec[r] -= sum_c conj(A(k0 + c, r)) y_c   (a contiguous run of column r)
// and the workgroup whose rows start at nxt0 (the next diagonal block, nw
// rows) then solves it: out[nxt0 + i] = sum_j M(i, j) vec[nxt0 + j] with M the
// block's precomputed inverse (op N) or its conjugate transpose (op C).
// Threads per workgroup of the sweep kernels: the diagonal-block solve and
// the update of the next block are each done by one workgroup on the sweep's
// critical path, so a workgroup keeps B^2 / 1024 independent loads per
// thread in flight at once.
constexpr int kSweepThreads = 1024;

// Loads of trsv_block_apply: thread t holds M[j B + t % B] for its B / P
// columns j (P = kSweepThreads / B column groups). The block is a full B x B
// array, so every load is in bounds whatever the block's width.
template <typename T, int B>
__device__ __forceinline__ void trsv_block_load(const cx<T>* __restrict__ M, cx<T>* m) {
    constexpr int P = kSweepThreads / B, J = B / P;
    const int t = threadIdx.x, i = t % B, part = t / B;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) m[jj] = M[(part * J + jj) * B + i];
}

// out[i] = sum_j M[j B + i] v_j with M's entries from trsv_block_load: M is
// the inverse (op N) or its conjugate transpose (op C), column-major; entries
// past nw contribute zero.
template <typename T, int B, bool AGENT = false>
__device__ __forceinline__ void trsv_block_apply(const cx<T>* m, const cx<T>* v, int nw, T (*spr)[B], T (*spi)[B],
                                                 cx<T>* __restrict__ out) {
    constexpr int P = kSweepThreads / B, J = B / P;
    const int t = threadIdx.x, i = t % B, part = t / B;
    T ar = 0, ai = 0;
#pragma unroll
    for (int jj = 0; jj < J; ++jj) {
        const int j = part * J + jj;
        const cx<T> x = j < nw ? v[j] : cx<T>{0, 0};
        ar += m[jj].re * x.re - m[jj].im * x.im;
        ai += m[jj].re * x.im + m[jj].im * x.re;
    }
    spr[part][i] = ar;
    spi[part][i] = ai;
    __syncthreads();
    if (t < nw) {
        T re = 0, im = 0;
#pragma unroll
        for (int q = 0; q < P; ++q) { re += spr[q][t]; im += spi[q][t]; }
        if constexpr (AGENT) publish_cx(out + t, cx<T>{re, im});
        else out[t] = cx<T>{re, im};
    }
}

template <typename T, int B, bool AGENT = false>
__device__ __forceinline__ void trsv_block_solve(const cx<T>* __restrict__ M, const cx<T>* v, int nw, T (*spr)[B],
                                                 T (*spi)[B], cx<T>* __restrict__ out) {
    // all loads are issued before the first use
    cx<T> m[B * B / kSweepThreads];
    trsv_block_load<T, B>(M, m);
    trsv_block_apply<T, B, AGENT>(m, v, nw, spr, spi, out);
}

// Mh = M^H for every B x B block: Mh[blk][j B + i] = conj(M[blk][i B + j]).
template <typename T, int B>
__global__ void conj_transpose_blocks_kernel(long long total, const cx<T>* __restrict__ M, cx<T>* __restrict__ Mh) {
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long long)gridDim.x * blockDim.x) {
        const long long blk = e / (B * B);
        const int r = (int)(e % (B * B)), j = r / B, i = r % B;
        const cx<T> m = M[blk * B * B + (long long)i * B + j];
        Mh[e] = cx<T>{m.re, -m.im};
    }
}

// Factor-tile entries per thread of sweep_tile_load / sweep_tile_apply.
template <int B>
constexpr int kSweepTileRegs = B * B / kSweepThreads;

// Loads of sweep_tile_apply: the factor tile of rows [rb, rb + B) and source
// columns [k0, k0 + w), op 0 by columns (thread t: row t % B, columns of
// group t / B), op 2 by rows of the stored factor (B / 4 lanes per row, 4
// entries each). Columns past w read column w - 1 (in bounds; multiplied by
// sy = 0); rows past r1 are not read.
template <typename TF, int B>
__device__ __forceinline__ void sweep_tile_load(int op, int w, int k0, const cx<TF>* __restrict__ LU, long long lda,
                                                int rb, int r1, cx<TF>* f) {
    constexpr int P = kSweepThreads / B;   // op N: column groups
    constexpr int CG = B / P;              // columns per group
    constexpr int L = B / 4;               // op C: lanes per row, 4 entries each
    constexpr int RP = kSweepThreads / L;  // op C: rows per pass
    const int t = threadIdx.x;
    if (op == 0) {
        const int row = t % B, grp = t / B, r = rb + row;
        const cx<TF>* a = LU + (long long)k0 * lda + r;
#pragma unroll
        for (int cc = 0; cc < CG; ++cc)
            f[cc] = r < r1 ? a[(long long)min(grp * CG + cc, w - 1) * lda] : cx<TF>{0, 0};
    } else {
        const int sub = t % L, rr = t / L;
#pragma unroll
        for (int pass = 0; pass < B / RP; ++pass) {
            const int r = rb + pass * RP + rr;
            const cx<TF>* a = LU + (long long)r * lda + k0;
#pragma unroll
            for (int e = 0; e < 4; ++e) f[pass * 4 + e] = r < r1 ? a[min(sub * 4 + e, w - 1)] : cx<TF>{0, 0};
        }
    }
}

// sres (the workgroup's B rows of the right-hand side, rows [rb, rb + B), in
// LDS) -= the factor tile f (from sweep_tile_load) times sy: op 0 the tile of
// L or U itself, op 2 the tile of U^H or L^H (rows of the stored factor,
// conjugated). Entry e of the tile is f[e * FS] (FS = 1: registers; the
// thread's column of an LDS array otherwise). Each row's products are summed
// in a fixed order, then subtracted once.
template <typename TF, typename T, int B, int FS = 1>
__device__ __forceinline__ void sweep_tile_apply(int op, int rb, int r1, const cx<TF>* f, const cx<T>* sy,
                                                 cx<T>* sres, T (*spr)[B], T (*spi)[B]) {
    constexpr int P = kSweepThreads / B;
    constexpr int CG = B / P;
    constexpr int L = B / 4;
    constexpr int RP = kSweepThreads / L;
    const int t = threadIdx.x;
    if (op == 0) {
        const int row = t % B, grp = t / B;
        T ar = 0, ai = 0;
#pragma unroll
        for (int cc = 0; cc < CG; ++cc) {
            const int c = grp * CG + cc;
            const cx<T> v = {(T)f[cc * FS].re, (T)f[cc * FS].im};
            ar += v.re * sy[c].re - v.im * sy[c].im;
            ai += v.re * sy[c].im + v.im * sy[c].re;
        }
        spr[grp][row] = ar;
        spi[grp][row] = ai;
        __syncthreads();
        if (t < B && rb + t < r1) {
            T re = 0, im = 0;
#pragma unroll
            for (int q = 0; q < P; ++q) { re += spr[q][t]; im += spi[q][t]; }
            sres[t].re -= re;
            sres[t].im -= im;
        }
    } else {
        const int sub = t % L, rr = t / L;
#pragma unroll
        for (int pass = 0; pass < B / RP; ++pass) {
            const int row = pass * RP + rr;
            T ar = 0, ai = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = sub * 4 + e;
                const cx<TF> fe = f[(pass * 4 + e) * FS];
                const cx<T> v = {(T)fe.re, (T)fe.im};
                ar += v.re * sy[c].re + v.im * sy[c].im;
                ai += v.re * sy[c].im - v.im * sy[c].re;
            }
            for (int off = L / 2; off > 0; off >>= 1) {
                ar += __shfl_xor(ar, off, L);
                ai += __shfl_xor(ai, off, L);
            }
            if (sub == 0 && rb + row < r1) {
                sres[row].re -= ar;
                sres[row].im -= ai;
            }
        }
    }
}

// sres (the workgroup's B rows of the right-hand side, rows [rb, rb + B), in
// LDS) -= the factor tile of those rows and source columns [k0, k0 + w) times
// sy: op 0 the tile of L or U itself (column-major reads), op 2 the tile of
// U^H or L^H (row reads of the stored factor, conjugated). Each row's
// products are summed in a fixed order, then subtracted once.
template <typename TF, typename T, int B>
__device__ __forceinline__ void sweep_tile(int op, int w, int k0, const cx<TF>* __restrict__ LU, long long lda, int rb,
                                           int r1, const cx<T>* sy, cx<T>* sres, T (*spr)[B], T (*spi)[B]) {
    constexpr int P = kSweepThreads / B;   // op N: column groups
    constexpr int CG = B / P;              // columns per group
    constexpr int L = B / 4;               // op C: lanes per row, 4 entries each
    constexpr int RP = kSweepThreads / L;  // op C: rows per pass
    const int t = threadIdx.x;
    if (op == 0) {
        const int row = t % B, grp = t / B, r = rb + row;
        T ar = 0, ai = 0;
        if (r < r1) {
            // every load of the row issued before the first use: columns past
            // w read column w - 1 (in bounds) and are multiplied by sy = 0
            const cx<TF>* a = LU + (long long)k0 * lda + r;
            cx<TF> f[CG];
#pragma unroll
            for (int cc = 0; cc < CG; ++cc) f[cc] = a[(long long)min(grp * CG + cc, w - 1) * lda];
#pragma unroll
            for (int cc = 0; cc < CG; ++cc) {
                const int c = grp * CG + cc;
                const cx<T> v = {(T)f[cc].re, (T)f[cc].im};
                ar += v.re * sy[c].re - v.im * sy[c].im;
                ai += v.re * sy[c].im + v.im * sy[c].re;
            }
        }
        spr[grp][row] = ar;
        spi[grp][row] = ai;
        __syncthreads();
        if (t < B && rb + t < r1) {
            T re = 0, im = 0;
#pragma unroll
            for (int q = 0; q < P; ++q) { re += spr[q][t]; im += spi[q][t]; }
            sres[t].re -= re;
            sres[t].im -= im;
        }
    } else {
        const int sub = t % L, rr = t / L;
#pragma unroll
        for (int pass = 0; pass < B / RP; ++pass) {
            const int row = pass * RP + rr, r = rb + row;
            T ar = 0, ai = 0;
            if (r < r1) {
                // loads first; columns past w read column w - 1, times sy = 0
                const cx<TF>* a = LU + (long long)r * lda + k0;
                cx<TF> f[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) f[e] = a[min(sub * 4 + e, w - 1)];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = sub * 4 + e;
                    const cx<T> v = {(T)f[e].re, (T)f[e].im};
                    ar += v.re * sy[c].re + v.im * sy[c].im;
                    ai += v.re * sy[c].im - v.im * sy[c].re;
                }
            }
            for (int off = L / 2; off > 0; off >>= 1) {
                ar += __shfl_xor(ar, off, L);
                ai += __shfl_xor(ai, off, L);
            }
            if (sub == 0) {
                sres[row].re -= ar;
                sres[row].im -= ai;
            }
        }
    }
}

// Diagonal blocks solved per launch of the sweep kernel.
constexpr int kSweepChain = 16;
// Row blocks a bulk workgroup keeps in LDS across a launch (one workgroup per
// CU shares ~n / (64 * 240) blocks: 5 at n = 72k).
constexpr int kSweepBulkRows = 12;

// One launch of a block triangular sweep: with the solution block kb known
// (in out), solves the next D blocks (kb+1 .. kb+D forward, kb-1 .. kb-D
// backward) and applies source blocks kb .. kb+-(D-1) to every row block
// beyond them. Workgroup d < D owns block kb+-(1+d): it applies source blocks
// kb .. kb+-d in order (waiting for each block solved in this launch), then
// solves its block with the diagonal-block inverse and publishes it in place
// in out (publish_cx; readers poll the entries). The nbulk workgroups >= D
// share the row blocks beyond (block i to workgroup D + i mod nbulk) and apply
// the D source blocks one after the other, each as soon as it is published;
// nbulk is at most the number of resident workgroups minus D. Every row
// receives the same products in the same order as with one source block per
// launch.
template <typename TF, typename T, int B>
__global__ __launch_bounds__(kSweepThreads, sizeof(TF) == 4 ? 8 : 4) void trsv_chain_kernel(int n, int dir, int kb, int D, int nbulk, int op,
                                                                   const cx<TF>* __restrict__ LU, long long lda,
                                                                   cx<T>* __restrict__ vec, const cx<T>* __restrict__ inv,
                                                                   cx<T>* __restrict__ out, unsigned int* err) {
    constexpr int P = kSweepThreads / B;
    __shared__ cx<T> sy[B];
    __shared__ cx<T> sres[B];
    __shared__ T spr[P][B], spi[P][B];
    __shared__ cx<TF> stile[kSweepTileRegs<B>][kSweepThreads];
    const int t = threadIdx.x;
    const int wg = blockIdx.x;
    const int nb = (n + B - 1) / B;
    if (wg < D) {
        const int rbk = dir > 0 ? kb + 1 + wg : kb - 1 - wg;
        const int rb = rbk * B;
        // the last source block's factor tile and this block's inverse do not
        // depend on the solution: they are loaded before the wait for the last
        // source block, so that once it is published only products remain.
        // The tile is loaded first and parked in LDS (each thread reads back
        // its own entries); the inverse is loaded into registers after the
        // other source blocks, whose tiles need the registers.
        const int last = dir > 0 ? kb + wg : kb - wg;
        const int wl = min(B, n - last * B);
        {
            cx<TF> f[kSweepTileRegs<B>];
            sweep_tile_load<TF, B>(op, wl, last * B, LU, lda, rb, n, f);
#pragma unroll
            for (int e = 0; e < kSweepTileRegs<B>; ++e) stile[e][t] = f[e];
        }
        if (t < B) sres[t] = rb + t < n ? vec[rb + t] : cx<T>{0, 0};
        for (int q = 0; q < wg; ++q) {
            const int src = dir > 0 ? kb + q : kb - q;
            __syncthreads();
            const int w = min(B, n - src * B);
            if (t < B) sy[t] = t < w ? wait_published(&out[(size_t)src * B + t], err) : cx<T>{0, 0};
            __syncthreads();
            sweep_tile<TF, T, B>(op, w, src * B, LU, lda, rb, n, sy, sres, spr, spi);
        }
        cx<T> m[B * B / kSweepThreads];
        trsv_block_load<T, B>(inv + (size_t)rbk * B * B, m);
        __syncthreads();
        if (t < B) sy[t] = t < wl ? wait_published(&out[(size_t)last * B + t], err) : cx<T>{0, 0};
        __syncthreads();
        sweep_tile_apply<TF, T, B, kSweepThreads>(op, rb, n, &stile[0][t], sy, sres, spr, spi);
        __syncthreads();
        const int nw = min(B, n - rb);
        trsv_block_apply<T, B, true>(m, sres, nw, spr, spi, out + rb);
        return;
    }
    // bulk: the workgroup's row blocks stay in LDS across the D source blocks
    // (up to kSweepBulkRows blocks; a larger share streams through global)
    const int nbeyond = dir > 0 ? nb - (kb + 1 + D) : kb - D;
    __shared__ cx<T> sown[kSweepBulkRows][B];
    int nown = 0;
    for (int i = wg - D; i < nbeyond && nown < kSweepBulkRows; i += nbulk, ++nown) {
        const int rb = (dir > 0 ? kb + 1 + D + i : i) * B;
        if (t < B) sown[nown][t] = rb + t < n ? vec[rb + t] : cx<T>{0, 0};
    }
    for (int q = 0; q < D; ++q) {
        const int src = dir > 0 ? kb + q : kb - q;
        __syncthreads();
        const int w = min(B, n - src * B);
        if (t < B) sy[t] = t < w ? wait_published(&out[(size_t)src * B + t], err) : cx<T>{0, 0};
        int o = 0;
        for (int i = wg - D; i < nbeyond; i += nbulk, ++o) {
            const int rb = (dir > 0 ? kb + 1 + D + i : i) * B;
            cx<T>* res = o < kSweepBulkRows ? sown[o] : sres;
            if (o >= kSweepBulkRows && t < B) sres[t] = rb + t < n ? vec[rb + t] : cx<T>{0, 0};
            __syncthreads();
            sweep_tile<TF, T, B>(op, w, src * B, LU, lda, rb, n, sy, res, spr, spi);
            __syncthreads();
            if (o >= kSweepBulkRows && t < B && rb + t < n) vec[rb + t] = sres[t];
        }
    }
    int o = 0;
    for (int i = wg - D; i < nbeyond && o < kSweepBulkRows; i += nbulk, ++o) {
        const int rb = (dir > 0 ? kb + 1 + D + i : i) * B;
        if (t < B && rb + t < n) vec[rb + t] = sown[o][t];
    }
}

// out[k0 + i] = M vec[k0 + i] for one diagonal block (the first of a sweep),
// published as the sweep kernels publish theirs.
template <typename T, int B>
__global__ __launch_bounds__(kSweepThreads) void trsv_diag_kernel(int k0, int w, const cx<T>* __restrict__ inv, int op,
                                                                  const cx<T>* __restrict__ vec,
                                                                  cx<T>* __restrict__ out) {
    __shared__ T spr[kSweepThreads / B][B], spi[kSweepThreads / B][B];
    trsv_block_solve<T, B, true>(inv, vec + k0, w, spr, spi, out + k0);
}

// out[i] = in[perm[i]] (gather) or out[perm[i]] = in[i] (scatter)
template <typename T>
__global__ void permute_kernel(int n, const int* __restrict__ perm, const cx<T>* __restrict__ in,
                               cx<T>* __restrict__ out, int scatter) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (scatter) out[perm[i]] = in[i];
    else out[i] = in[perm[i]];
}

}  // namespace

// ==================================================================== host
struct Workspace {
    DevBuf<unsigned char> buf;
    void* get(size_t bytes) {
        if (buf.n < bytes) buf.alloc(bytes);
        return buf.p;
    }
};
// The library's long-lived device state (workspaces, captured sweep graphs,
// the wait-error word) is heap-allocated and never destroyed: at process exit
// the HIP runtime may already be torn down, so no device call may run from a
// static destructor. The memory goes with the process.
static Workspace& ws(int slot) {
    static Workspace* w = new Workspace[8];
    return w[slot];
}

// Bound of the device-side waits and its testing switch (sie_set_wait_bound).
struct WaitCtl {
    double seconds = 5.0;
    int force = 0;
};
static WaitCtl& wait_ctl() {
    static WaitCtl c;
    return c;
}
// Launch form of the panel kernels (sie_set_panel_launch): plain dispatches by
// default. A panel's workgroups wait for one another, so they must all be
// resident at once; a plain dispatch of at most one workgroup per CU on a
// stream whose previous kernel has finished gets that from the dispatcher
// unless kernels of other streams hold the CUs, and then the bounded waits
// give up (SIE_ERR_DEVICE). The cooperative form has the runtime guarantee it,
// at ~21 us more per launch (178 against 157 us per 72k-row panel and two
// small kernels, measured), 47 ms per factorisation at N = 72k.
static bool g_cooperative_panels = false;
static unsigned int* g_wait_word = nullptr;
static unsigned int* wait_word() {
    if (!g_wait_word) {
        unsigned int* q = nullptr;
        SIE_HIP(hipMalloc(&q, sizeof(unsigned int)));
        SIE_HIP(hipMemset(q, 0, sizeof(unsigned int)));
        g_wait_word = q;
    }
    return g_wait_word;
}
static WaitArg wait_arg() {
    const WaitCtl& c = wait_ctl();
    return WaitArg{(unsigned long long)(c.seconds * 1e8), wait_word(), c.force};
}
// Reads (and clears) the wait-error word after the stream's work: a device
// wait that gave up becomes SIE_ERR_DEVICE. Synchronises the stream.
static void check_waits(hipStream_t s) {
    unsigned int* w = wait_word();
    unsigned int h = 0;
    SIE_HIP(hipMemcpyAsync(&h, w, sizeof(h), hipMemcpyDeviceToHost, s));
    SIE_HIP(hipStreamSynchronize(s));
    if (h == 0) return;
    SIE_HIP(hipMemsetAsync(w, 0, sizeof(unsigned int), s));
    SIE_HIP(hipStreamSynchronize(s));
    const char* what = h == kWaitPanel ? "a panel's grid-wide pivot exchange"
                       : h == kWaitSweep ? "a triangular sweep's published block"
                                         : "a panel's published pivot row";
    throw Error(SIE_ERR_DEVICE, std::string("device wait gave up: ") + what + " did not arrive within " +
                                    std::to_string(wait_ctl().seconds) +
                                    " s (workgroups not co-resident, or the device is shared)");
}

template <typename T>
static void gemm(int M, int N, int K, double alpha, const cx<T>* A, long long lda, const cx<T>* B, long long ldb,
                 double beta, cx<T>* C, long long ldc, hipStream_t s, int mode3m = 0);

template <>
void gemm<float>(int M, int N, int K, double alpha, const c32* A, long long lda, const c32* B, long long ldb,
                 double beta, c32* C, long long ldc, hipStream_t s, int mode3m) {
    if (M <= 0 || N <= 0) return;
    prof().begin(PROF_GEMM_F32, s);
    if (K <= 64) {
        dim3 grid((M + 63) / 64, (N + 63) / 64);
        if (mode3m)
            cgemm_f32_kernel<1, 1><<<grid, 256, 0, s>>>(M, N, K, (float)alpha, (float)beta, A, lda, B, ldb, C, ldc);
        else
            cgemm_f32_kernel<0, 1><<<grid, 256, 0, s>>>(M, N, K, (float)alpha, (float)beta, A, lda, B, ldb, C, ldc);
    } else {
        dim3 grid((M + 127) / 128, (N + 127) / 128);
        if (mode3m)
            cgemm_f32_kernel<1, 2><<<grid, 256, 0, s>>>(M, N, K, (float)alpha, (float)beta, A, lda, B, ldb, C, ldc);
        else
            cgemm_f32_kernel<0, 2><<<grid, 256, 0, s>>>(M, N, K, (float)alpha, (float)beta, A, lda, B, ldb, C, ldc);
    }
    // complex-equivalent flops 8MNK; bytes: A and B once, C read + written
    prof().end(PROF_GEMM_F32, s, 8.0 * M * N * K, 8.0 * ((double)M * K + (double)K * N + 2.0 * M * N));
    SIE_HIP(hipGetLastError());
}

template <>
void gemm<double>(int M, int N, int K, double alpha, const c64* A, long long lda, const c64* B, long long ldb,
                  double beta, c64* C, long long ldc, hipStream_t s, int) {
    if (M <= 0 || N <= 0) return;
    dim3 grid((M + G64_BM - 1) / G64_BM, (N + G64_BN - 1) / G64_BN);
    prof().begin(PROF_GEMM_F64, s);
    zgemm_f64_kernel<<<grid, 256, 0, s>>>(M, N, K, alpha, beta, A, lda, B, ldb, C, ldc);
    prof().end(PROF_GEMM_F64, s, 8.0 * M * N * K, 16.0 * ((double)M * K + (double)K * N + 2.0 * M * N));
    SIE_HIP(hipGetLastError());
}

// FP32-complex GEMM on the FP16 matrix cores (split-FP16, see the kernel).
// a_bound > 0: a bound on max(|re|, |im|) of A known to the caller (the LU's
// L blocks: |l| <= sqrt 2 under |re| + |im| pivoting), used for A's scale in
// place of a pass over A.
static void gemm_f16x3(int M, int N, int K, double alpha, const c32* A, long long lda, const c32* B, long long ldb,
                       double beta, c32* C, long long ldc, hipStream_t s, float a_bound = 0.f) {
    if (M <= 0 || N <= 0) return;
    if (K <= 0 || alpha == 0.0) {  // C = beta C
        gemm<float>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, s, 0);
        return;
    }
    const int KB = (K + 15) / 16;
    // wide products (the LU's super-panel updates) use the 256 x 128 three-stage
    // kernel; narrow ones (in-panel updates, N <= 1024) keep more tiles with 128 x 128
    const bool r3 = K >= 512 && N > 1024;
    if (r3) {
        // the kernel addresses each split operand with 32-bit byte offsets:
        // split larger products into row / column chunks below 4 GiB each
        const long long per_row = (long long)KB * 128;  // bytes per padded row of a split operand
        const long long cap = ((1LL << 32) - 1) / per_row / 256 * 256;
        require(cap >= 256, "split-FP16 GEMM: K too large");
        if (M > cap) {
            for (long long r0 = 0; r0 < M; r0 += cap)
                gemm_f16x3((int)std::min<long long>(cap, M - r0), N, K, alpha, A + r0, lda, B, ldb, beta, C + r0, ldc, s,
                           a_bound);
            return;
        }
        if (N > cap) {
            for (long long c0 = 0; c0 < N; c0 += cap)
                gemm_f16x3(M, (int)std::min<long long>(cap, N - c0), K, alpha, A, lda, B + c0 * ldb, ldb, beta,
                           C + c0 * ldc, ldc, s, a_bound);
            return;
        }
    }
    // planes padded to whole 256-row (A) and 128-column (B) blocks
    const size_t pa = (size_t)KB * ((M + 255) / 256 * 256) * 16 * 4, pb = (size_t)KB * ((N + 127) / 128 * 128) * 16 * 4;
    unsigned char* w = (unsigned char*)ws(6).get((pa + pb) * sizeof(_Float16) + 256);
    _Float16* PA = (_Float16*)w;
    _Float16* PB = PA + pa;
    unsigned int* mx = (unsigned int*)(w + (pa + pb) * sizeof(_Float16));
    const long long ta = (long long)M * K, tb = (long long)K * N;
    if (a_bound > 0.f) {
        set_mx_kernel<<<1, 1, 0, s>>>(mx, __builtin_bit_cast(unsigned int, a_bound));
    } else {
        SIE_HIP(hipMemsetAsync(mx, 0, 2 * sizeof(unsigned int), s));
        maxabs_kernel<<<(unsigned)std::min<long long>((ta + 255) / 256, 1024), 256, 0, s>>>(M, K, A, lda, mx);
    }
    maxabs_kernel<<<(unsigned)std::min<long long>((tb + 255) / 256, 1024), 256, 0, s>>>(K, N, B, ldb, mx + 1);
    if (r3) {
        const long long ra = (long long)((M + 255) / 256) * 256 * KB * 2, rb = (long long)((N + 127) / 128) * 128 * KB * 2;
        split_f16_blk_kernel<<<(unsigned)std::min<long long>((ra + 255) / 256, 1 << 16), 256, 0, s>>>(M, K, 256, A, lda,
                                                                                                   0, mx, PA);
        split_f16_blk_kernel<<<(unsigned)std::min<long long>((rb + 255) / 256, 1 << 16), 256, 0, s>>>(N, K, 128, B, ldb,
                                                                                                   1, mx + 1, PB);
    } else {
        const long long sa = (long long)KB * M * 16, sb = (long long)KB * N * 16;
        split_f16_kernel<<<(unsigned)std::min<long long>((sa + 255) / 256, 1 << 16), 256, 0, s>>>(M, K, A, lda, 0,
                                                                                                 mx, PA);
        split_f16_kernel<<<(unsigned)std::min<long long>((sb + 255) / 256, 1 << 16), 256, 0, s>>>(N, K, B, ldb, 1,
                                                                                                 mx + 1, PB);
    }
    const int slot = r3 ? PROF_GEMM_F16X3_WIDE : PROF_GEMM_F16X3;
    prof().begin(slot, s);
    if (r3) {
        const long long tiles = (long long)((M + 255) / 256) * ((N + 127) / 128);
        require(tiles * 512 < (1LL << 32), "split-FP16 GEMM: too many tiles for one dispatch");
        cgemm_f16x3_r3_kernel<true><<<(unsigned)tiles, 512, 0, s>>>(M, N, K, (float)alpha, (float)beta, PA, PB, mx,
                                                                    C, ldc);
    } else {
        const long long tiles = (long long)((M + 127) / 128) * ((N + 127) / 128);
        require(tiles * 256 < (1LL << 32), "split-FP16 GEMM: too many tiles for one dispatch");
        cgemm_f16x3_kernel<2, 2, true><<<(unsigned)tiles, 256, 0, s>>>(M, N, K, (float)alpha, (float)beta, PA, PB, mx,
                                                                       C, ldc);
    }
    // algorithmic (complex-equivalent) flops 8MNK; the kernel issues 3x that on
    // the FP16 matrix cores. bytes: split A and B planes once, C read + written
    prof().end(slot, s, 8.0 * M * N * K, 8.0 * ((double)M * K + (double)K * N) + 16.0 * M * N);
    SIE_HIP(hipGetLastError());
}

template <typename T>
static void laswp(cx<T>* A, long long lda, int rows, int c0, int nc, const int* ipiv, int kb, int ke, int r0,
                  bool rev, hipStream_t s) {
    if (nc <= 0 || ke <= kb) return;
    laswp_kernel<T><<<(nc + 127) / 128, 128, 0, s>>>(A, lda, rows, c0, nc, 0, 0, ipiv, kb, ke, r0, rev ? 1 : 0);
    SIE_HIP(hipGetLastError());
}

// The forward swaps of rows [kb, ke) on two column ranges in one launch (the
// LU's columns left and right of the block just factored).
template <typename T>
static void laswp2(cx<T>* A, long long lda, int rows, int c0, int nc, int c1, int nc1, const int* ipiv, int kb,
                   int ke, hipStream_t s) {
    nc = std::max(nc, 0);
    nc1 = std::max(nc1, 0);
    if (nc + nc1 <= 0 || ke <= kb) return;
    // up to 256 interchanges (a panel's, an NB block's): composed per
    // workgroup and moved with the loads in flight together; longer
    // sequences (a level's 4096, on the wide trailing columns) one column per
    // lane, in order (bound by the scattered pivot rows' traffic)
    const int tot = nc + nc1, len = ke - kb;
    if (len <= 32) {
        constexpr int cols = 16;
        laswp_perm_kernel<T, 32><<<(tot + cols - 1) / cols, 256, 0, s>>>(A, lda, rows, c0, nc, c1, nc1, ipiv, kb, ke,
                                                                         cols);
    } else if (len <= 256) {
        constexpr int cols = 16;
        laswp_perm_kernel<T, 256><<<(tot + cols - 1) / cols, 256, 0, s>>>(A, lda, rows, c0, nc, c1, nc1, ipiv, kb,
                                                                          ke, cols);
    } else {
        laswp_kernel<T><<<(tot + 127) / 128, 128, 0, s>>>(A, lda, rows, c0, nc, c1, nc1, ipiv, kb, ke, 0, 0);
    }
    SIE_HIP(hipGetLastError());
}

template <typename T>
static void trsm_small(int w, const cx<T>* Tm, long long lda, cx<T>* B, long long ldb, int ncols, int uplo,
                       hipStream_t s) {
    if (ncols <= 0 || w <= 0) return;
    constexpr int WM = 32;  // panel width (FP32) and the multi-RHS solve block
    require(w <= WM, "internal: triangular block wider than the register tile");
    if (ncols <= 65536) {  // WM lanes per column (measured faster up to ~64k columns at w = 32)
        const unsigned g = (unsigned)((ncols + 256 / WM - 1) / (256 / WM));
        switch (uplo) {
            case 0: trsm_lanes_kernel<T, WM, 0><<<g, 256, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            case 1: trsm_lanes_kernel<T, WM, 1><<<g, 256, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            case 2: trsm_lanes_kernel<T, WM, 2><<<g, 256, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            default: trsm_lanes_kernel<T, WM, 3><<<g, 256, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
        }
    } else {  // many columns: one lane per column, one wave per 64 columns
        const unsigned g = (unsigned)((ncols + 63) / 64);
        switch (uplo) {
            case 0: trsm_small_kernel<T, WM, 0><<<g, 64, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            case 1: trsm_small_kernel<T, WM, 1><<<g, 64, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            case 2: trsm_small_kernel<T, WM, 2><<<g, 64, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
            default: trsm_small_kernel<T, WM, 3><<<g, 64, 0, s>>>(w, Tm, lda, B, ldb, ncols); break;
        }
    }
    SIE_HIP(hipGetLastError());
}

template <typename T>
struct PanelCfg;
template <>
struct PanelCfg<float> {
    static constexpr int W = 32;
};
template <>
struct PanelCfg<double> {
    static constexpr int W = 16;
};
constexpr int kPanelBS = 512;

// Panel workspace of one factorisation (sized for its first, tallest panel):
// two arrays of W x Gmax pivot keys (FP32: each launch uses one and clears the
// other for the next launch), the candidates, published rows, row j and the
// barrier counter (FP64: monotone over the factorisation, so never reset
// between launches). Both are cleared once per factorisation.
struct PanelState {
    int gmax = 0;
    int parity = 0;
    unsigned int counter_base = 0;
};
static PanelState& panel_state() {
    static PanelState st;
    return st;
}
template <typename T>
static unsigned char* panel_workspace(int gmax, size_t* keys_b, size_t* cand_b, size_t* pub_b, size_t* drow_b) {
    constexpr int W = PanelCfg<T>::W;
    // rows: W cx<T> each, or 2W tagged 64-bit words in the FP32 keyed form
    const size_t row_b = std::max(sizeof(cx<T>), 2 * sizeof(unsigned long long)) * W;
    *keys_b = 2 * (size_t)W * gmax * sizeof(unsigned long long);
    *cand_b = (size_t)W * gmax * sizeof(Cand<T>);
    *pub_b = (size_t)W * gmax * row_b;
    *drow_b = (size_t)W * row_b;
    return (unsigned char*)ws(1).get(*keys_b + *cand_b + *pub_b + *drow_b + 256);
}
template <typename T>
static void panel_begin(int n, hipStream_t s) {
    PanelState& st = panel_state();
    st.gmax = (n + kPanelBS - 1) / kPanelBS;
    st.parity = 0;
    st.counter_base = 0;
    size_t keys_b, cand_b, pub_b, drow_b;
    unsigned char* wsp = panel_workspace<T>(st.gmax, &keys_b, &cand_b, &pub_b, &drow_b);
    SIE_HIP(hipMemsetAsync(wsp, 0, keys_b, s));
    SIE_HIP(hipMemsetAsync(wsp + keys_b + cand_b + pub_b + drow_b, 0, sizeof(unsigned int), s));
}

// Factors the M x w panel at P; ipiv[j] receives the absolute row (panel row
// + ipiv_off) swapped with row j. The FP32 (keyed) panel also applies its
// interchanges to the columns at offsets [sc0, sc0 + snc) and [w, w + snc1)
// from P (the rest of its NB block) and returns true; the FP64 panel leaves
// them to the caller (false).
template <typename T>
static bool panel(int M, int w, cx<T>* P, long long lda, int* ipiv, int ipiv_off, int* d_info, int info_base,
                  hipStream_t s, int sc0 = 0, int snc = 0, int snc1 = 0) {
    constexpr int W = PanelCfg<T>::W;
    constexpr bool keyed = std::is_same<T, float>::value;
    const void* kern = keyed ? (const void*)panel_getf2_keyed_kernel<W, kPanelBS>
                             : (const void*)panel_getf2_kernel<T, W, kPanelBS>;
    int G = (M + kPanelBS - 1) / kPanelBS;
    static int max_coresident = -1;
    if (max_coresident < 0) {
        int dev = 0, ncu = 0, per = 0;
        SIE_HIP(hipGetDevice(&dev));
        SIE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        SIE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, kPanelBS, 0));
        max_coresident = ncu * per;
        if (keyed) max_coresident = std::min(max_coresident, kPanelKeyedMaxG);
    }
    if (G > max_coresident)
        throw Error(SIE_ERR_OVERFLOW, "panel height " + std::to_string(M) + " exceeds the co-resident capacity (" +
                                          std::to_string(max_coresident * kPanelBS) + " rows)");
    require(w >= 1 && w <= W, "internal: panel width out of range");
    PanelState& st = panel_state();
    require(G <= st.gmax, "internal: panel taller than its factorisation");
    size_t keys_b, cand_b, pub_b, drow_b;
    unsigned char* wsp = panel_workspace<T>(st.gmax, &keys_b, &cand_b, &pub_b, &drow_b);
    unsigned long long* keys = (unsigned long long*)wsp;
    Cand<T>* cand = (Cand<T>*)(wsp + keys_b);
    cx<T>* pub = (cx<T>*)(wsp + keys_b + cand_b);
    cx<T>* drow = (cx<T>*)(wsp + keys_b + cand_b + pub_b);
    unsigned int* counter = (unsigned int*)(wsp + keys_b + cand_b + pub_b + drow_b);
    prof().begin(PROF_PANEL, s);
    if (keyed) {
        int nkeys = W * st.gmax;
        unsigned long long* kcur = keys + (size_t)st.parity * nkeys;
        unsigned long long* knext = keys + (size_t)(st.parity ^ 1) * nkeys;
        st.parity ^= 1;
        // row tags: a launch number above the 6-bit column number (launch
        // numbers wrap after 2^26 launches; never 0, the cleared state)
        static unsigned launch_no = 0;
        launch_no = (launch_no + 1) & ((1u << 26) - 1);
        if (launch_no == 0) launch_no = 1;
        unsigned tag_base = launch_no << 6;
        unsigned long long* pubw = (unsigned long long*)pub;
        unsigned long long* droww = (unsigned long long*)drow;
        WaitArg wa = wait_arg();
        int sc1 = w;
        void* args[] = {&M,       &w,         &P,      &lda,       &ipiv, &ipiv_off, &kcur, &knext, &nkeys, &pubw,
                        &droww,   &tag_base,  &d_info, &info_base, &wa,   &sc0,      &snc,  &sc1,   &snc1};
        if (g_cooperative_panels) SIE_HIP(hipLaunchCooperativeKernel(kern, dim3(G), dim3(kPanelBS), args, 0, s));
        else SIE_HIP(hipLaunchKernel(kern, dim3(G), dim3(kPanelBS), args, 0, s));
    } else {
        unsigned int base = st.counter_base;
        st.counter_base += (unsigned)(G * w);  // one arrival per workgroup and column
        WaitArg wa = wait_arg();
        void* args[] = {&M, &w, &P, &lda, &ipiv, &ipiv_off, &cand, &pub, &drow, &counter, &base, &d_info,
                        &info_base, &wa};
        if (g_cooperative_panels) SIE_HIP(hipLaunchCooperativeKernel(kern, dim3(G), dim3(kPanelBS), args, 0, s));
        else SIE_HIP(hipLaunchKernel(kern, dim3(G), dim3(kPanelBS), args, 0, s));
    }
    prof().end(PROF_PANEL, s, 8.0 * M * w * w / 2, 2.0 * sizeof(cx<T>) * M * w);
    return keyed;
}

// block of the right-looking steps of the innermost level (NB = 512 measured
// slower at 72k: 2.96 s against 2.91 s per factorisation)
constexpr int kNB = 256;
// Column-block widths of the blocking hierarchy, outermost first: each level
// factors its block by blocks of the next level and updates the rest of its
// block once per sub-block with K = that width. Measured at N = 72k (FP32
// split-FP16 updates): one level of 4096 columns over the NB = 256 steps
// factors in 3.69 s, 1024 in 3.84 s, 8192 in 3.90 s and 4096 over 1024 in 3.80 s
// (an earlier state of the kernels); with the present ones, on one device in
// turn, 4096 alone 2.90 s against 4096 over 2048 2.92 s, and on other devices
// 4096 over 1024 2.92 s, 8192 over 2048 2.89 s, 4096 over 2048 over 1024 2.94 s
// (one level 2.86-2.91 s there).
constexpr int kLevels[] = {4096};
constexpr int kNumLevels = 1;

// C -= L U (M x N, K) on the matrix cores: the split-FP16 GEMM for variant 2
// in FP32, else the FP32 (4M / 3M) or FP64 MFMA kernels.
template <typename T>
static void update(int M, int N, int K, const cx<T>* L, const cx<T>* U, cx<T>* C, long long lda, hipStream_t s,
                   int variant) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    if constexpr (std::is_same<T, float>::value) {
        if (variant == 2) {
            gemm_f16x3(M, N, K, -1.0, L, lda, U, lda, 1.0, C, lda, s, 1.5f);
            return;
        }
    }
    gemm<T>(M, N, K, -1.0, L, lda, U, lda, 1.0, C, lda, s, variant == 1 ? 1 : 0);
}

// U12 = L11^{-1} A12 for the unit lower triangle L11 at rows/columns
// [r0, r0 + h) and the columns [c0, c0 + nc), in W-row blocks with
// NB-blocked updates of the rows below.
template <typename T>
static void trsm_lower(cx<T>* A, long long lda, int r0, int h, int c0, int nc, hipStream_t s, int variant,
                       int level = 0) {
    constexpr int W = PanelCfg<T>::W;
    if (nc <= 0) return;
    if (level < kNumLevels && h > kLevels[level]) {
        // blocks of this level's width, the rows below each updated with K = width
        const int bw = kLevels[level];
        for (int kb = r0; kb < r0 + h; kb += bw) {
            const int nb = std::min(bw, r0 + h - kb);
            trsm_lower<T>(A, lda, kb, nb, c0, nc, s, variant, level + 1);
            update<T>(r0 + h - (kb + nb), nc, nb, A + (long long)kb * lda + kb + nb, A + (long long)c0 * lda + kb,
                      A + (long long)c0 * lda + kb + nb, lda, s, variant);
        }
        return;
    }
    if (level < kNumLevels) {
        trsm_lower<T>(A, lda, r0, h, c0, nc, s, variant, level + 1);
        return;
    }
    for (int kb = r0; kb < r0 + h; kb += kNB) {
        const int nb = std::min(kNB, r0 + h - kb);
        for (int k1 = kb; k1 < kb + nb; k1 += W) {
            const int w = std::min(W, kb + nb - k1);
            cx<T>* B12 = A + (long long)c0 * lda + k1;
            trsm_small<T>(w, A + (long long)k1 * lda + k1, lda, B12, lda, nc, 0, s);
            const int below = kb + nb - (k1 + w);
            if (below > 0)
                gemm<T>(below, nc, w, -1.0, A + (long long)k1 * lda + k1 + w, lda, B12, lda, 1.0,
                        A + (long long)c0 * lda + k1 + w, lda, s, 0);
        }
        // rows of the triangle below this NB block
        update<T>(r0 + h - (kb + nb), nc, nb, A + (long long)kb * lda + kb + nb, A + (long long)c0 * lda + kb,
                  A + (long long)c0 * lda + kb + nb, lda, s, variant);
    }
}

// Factors the column block [c0, c1) over rows [c0, n) by right-looking NB
// steps of W-wide cooperative panels; row swaps are applied inside the block only.
template <typename T>
static void getrf_block(int n, cx<T>* A, long long lda, int c0, int c1, int* ipiv, int* d_info, hipStream_t s,
                        int variant) {
    constexpr int W = PanelCfg<T>::W;
    for (int k0 = c0; k0 < c1; k0 += kNB) {
        const int nb = std::min(kNB, c1 - k0);
        // the NB panel A[k0:n, k0:k0+nb) in W-wide sub-panels
        for (int k1 = k0; k1 < k0 + nb; k1 += W) {
            const int w = std::min(W, k0 + nb - k1);
            cx<T>* P = A + (long long)k1 * lda + k1;
            if (!panel<T>(n - k1, w, P, lda, ipiv + k1, k1, d_info, k1, s, k0 - k1, k1 - k0, k0 + nb - (k1 + w)))
                laswp2<T>(A, lda, n, k0, k1 - k0, k1 + w, k0 + nb - (k1 + w), ipiv, k1, k1 + w, s);
            const int rest = k0 + nb - (k1 + w);
            if (rest > 0) {
                trsm_small<T>(w, P, lda, A + (long long)(k1 + w) * lda + k1, lda, rest, 0, s);
                gemm<T>(n - k1 - w, rest, w, -1.0, A + (long long)k1 * lda + k1 + w, lda,
                        A + (long long)(k1 + w) * lda + k1, lda, 1.0, A + (long long)(k1 + w) * lda + k1 + w, lda, s,
                        0);
            }
        }
        // this NB panel's swaps on the other columns of the block
        const int right = c1 - k0 - nb;
        laswp2<T>(A, lda, n, c0, k0 - c0, k0 + nb, right, ipiv, k0, k0 + nb, s);
        if (right > 0) {
            trsm_lower<T>(A, lda, k0, nb, k0 + nb, right, s, variant);
            update<T>(n - k0 - nb, right, nb, A + (long long)k0 * lda + k0 + nb, A + (long long)(k0 + nb) * lda + k0,
                      A + (long long)(k0 + nb) * lda + k0 + nb, lda, s, variant);
        }
    }
}

// Factors the column block [c0, c1) over rows [c0, n) at a level of the
// blocking hierarchy: sub-blocks of the level's width are factored one after
// the other (recursively, NB steps at the innermost level); each then passes
// its swaps to the block's other columns and, for the columns to its right
// within the block, its U12 and one update with K = its width.
template <typename T>
static void getrf_level(int n, cx<T>* A, long long lda, int c0, int c1, int* ipiv, int* d_info, hipStream_t s,
                        int variant, int level) {
    if (level >= kNumLevels) {
        getrf_block<T>(n, A, lda, c0, c1, ipiv, d_info, s, variant);
        return;
    }
    const int bw = kLevels[level];
    for (int K0 = c0; K0 < c1; K0 += bw) {
        const int sp = std::min(bw, c1 - K0);
        getrf_level<T>(n, A, lda, K0, K0 + sp, ipiv, d_info, s, variant, level + 1);
        const int right = c1 - K0 - sp;
        // the outermost level's swaps reach the columns to its left only after
        // the factorisation (apply_deferred_swaps)
        laswp2<T>(A, lda, n, c0, level == 0 ? 0 : K0 - c0, K0 + sp, right, ipiv, K0, K0 + sp, s);
        if (right > 0) {
            trsm_lower<T>(A, lda, K0, sp, K0 + sp, right, s, variant, level + 1);
            update<T>(n - K0 - sp, right, sp, A + (long long)K0 * lda + K0 + sp, A + (long long)(K0 + sp) * lda + K0,
                      A + (long long)(K0 + sp) * lda + K0 + sp, lda, s, variant);
        }
    }
}

// tmp[col][r] = A[c0 + col][src[r]] for r < m (a row gather of nc columns;
// blockIdx.x runs over rows, so one column's gathers are in flight together).
template <typename T>
__global__ void gather_rows_kernel(const cx<T>* __restrict__ A, long long lda, int c0, int m,
                                   const int* __restrict__ src, cx<T>* __restrict__ tmp) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const long long col = blockIdx.y;
    if (r < m) tmp[col * m + r] = A[(c0 + col) * lda + src[r]];
}

// A[c0 + col][a + r] = tmp[col][r] for r < m.
template <typename T>
__global__ void put_rows_kernel(cx<T>* __restrict__ A, long long lda, int c0, int a, int m,
                                const cx<T>* __restrict__ tmp) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const long long col = blockIdx.y;
    if (r < m) A[(c0 + col) * lda + a + r] = tmp[col * m + r];
}

// The outermost level's row swaps on the columns to its left, applied once
// after the factorisation: the columns of block L (kLevels[0] wide) receive
// the swaps of every later block as one row permutation of rows [end of
// block L, n), composed on the host from ipiv, gathered into a temporary (at
// most deferred_tmp_bytes() of columns at a time) and copied back. Rows are
// only moved, so the factors are those of applying each block's swaps as it is
// factored; the 4096 scattered swaps per column and block become one
// coalesced pass per column.
constexpr size_t kDeferredTmpBytes = size_t(1) << 30;

// The temporary's size: kDeferredTmpBytes, or SIE_DEFERRED_TMP_BYTES (at least
// 4096) when set, so that tests reach the chunked loop at small n.
static size_t deferred_tmp_bytes() {
    const char* e = std::getenv("SIE_DEFERRED_TMP_BYTES");
    if (e == nullptr) return kDeferredTmpBytes;
    const long long v = std::atoll(e);
    return v >= 4096 ? (size_t)v : kDeferredTmpBytes;
}

template <typename T>
static void apply_deferred_swaps(int n, cx<T>* A, long long lda, const int* ipiv, hipStream_t s) {
    const int bw = kLevels[0];
    if (n <= bw) return;
    std::vector<int> hpiv(n);
    SIE_HIP(hipMemcpyAsync(hpiv.data(), ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    check_waits(s);  // a panel that gave up leaves ipiv incomplete: report that first
    // src rows of every block, concatenated: block L's run covers rows [a_L, n)
    std::vector<int> src;
    std::vector<size_t> off;
    for (int c = 0; c + bw < n; c += bw) {
        const int a = c + bw, m = n - a;
        off.push_back(src.size());
        src.resize(src.size() + m);
        int* p = src.data() + off.back();
        for (int r = 0; r < m; ++r) p[r] = a + r;
        for (int k = a; k < n; ++k) {
            const int q = hpiv[k];
            require(q >= k && q < n, "internal: pivot row outside the trailing rows");
            if (q != k) std::swap(p[k - a], p[q - a]);
        }
    }
    const size_t src_bytes = (src.size() * sizeof(int) + 255) / 256 * 256;
    const size_t tmp_bytes = deferred_tmp_bytes();
    unsigned char* w = (unsigned char*)ws(7).get(src_bytes + tmp_bytes);
    int* dsrc = (int*)w;
    cx<T>* tmp = (cx<T>*)(w + src_bytes);
    SIE_HIP(hipMemcpyAsync(dsrc, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice, s));
    SIE_HIP(hipStreamSynchronize(s));
    for (size_t L = 0; L < off.size(); ++L) {
        const int c = (int)L * bw, a = c + bw, m = n - a;
        const int chunk = (int)std::max<size_t>(1, std::min<size_t>(bw, tmp_bytes / ((size_t)m * sizeof(cx<T>))));
        for (int c0 = c; c0 < a; c0 += chunk) {
            const int nc = std::min(chunk, a - c0);
            const dim3 g((unsigned)((m + 255) / 256), (unsigned)nc);
            gather_rows_kernel<T><<<g, 256, 0, s>>>(A, lda, c0, m, dsrc + off[L], tmp);
            put_rows_kernel<T><<<g, 256, 0, s>>>(A, lda, c0, a, m, tmp);
        }
    }
    SIE_HIP(hipGetLastError());
}

// Blocked right-looking LU with partial pivoting; ipiv is 0-based absolute.
// Hierarchical blocking (kLevels): the trailing matrix receives one update per
// 4096 columns and the inner updates K = 256, so the trailing matrix of
// N = 72k is read and written 18 times instead of 282 times.
template <typename T>
void getrf(int n, cx<T>* A, long long lda, int* ipiv, int* d_info, hipStream_t s, int variant) {
    note_factorisation(A);
    SIE_HIP(hipMemsetAsync(d_info, 0, sizeof(int), s));
    panel_begin<T>(n, s);
    getrf_level<T>(n, A, lda, 0, n, ipiv, d_info, s, variant, 0);
    apply_deferred_swaps<T>(n, A, lda, ipiv, s);
}

template <typename T>
__global__ void copy_vec_kernel(int n, const cx<T>* __restrict__ in, cx<T>* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// Diagonal block size of the single-RHS sweeps. Each dependent step needs a
// B x B factor tile and a B x B inverse in one workgroup, whose CU streams
// ~45 GB/s: the data per sweep through that path grows as n B while the fixed
// cost per step (waiting for the block, reading it, publishing: ~2 us) grows
// as n / B; B = 64 balances them (measured against 128 at N = 72k).
template <typename T>
struct SweepCfg {
    static constexpr int B = 64;
};

// Generation of the last factorisation of each array, so that the sweeps can
// keep the diagonal-block inverses of a factorisation across solves.
static std::map<const void*, unsigned long long>& factor_generation() {
    static std::map<const void*, unsigned long long> g;
    return g;
}
static void note_factorisation(const void* A) {
    static unsigned long long counter = 0;
    factor_generation()[A] = ++counter;
}

// Single right-hand-side solve by B-blocked sweeps (op 0 = N, 2 = C): the
// diagonal-block inverses (computed once per factorisation), then one launch
// per block that updates the rows still to be solved and solves the next
// diagonal block. The launch sequence is fixed by (op, n, factors, vector,
// workspace), so it is captured once into a HIP graph and replayed:
// iterative refinement calls it with the same pointers every sweep.
// invL / invU: diagonal-block inverses for op N, their conjugate transposes for op C.
template <typename TF, typename T>
static void record_sweeps(int op, int n, const cx<TF>* LU, long long lda, cx<T>* b, const cx<T>* invL,
                          const cx<T>* invU, cx<T>* y, cx<T>* t, const int* perm, hipStream_t s) {
    constexpr int B = SweepCfg<TF>::B;
    const int nb = (n + B - 1) / B;
    const unsigned g = (unsigned)((n + 255) / 256);
    auto width = [&](int k) { return std::min(B, n - k); };
    // bulk workgroups per launch: as many as are resident with the
    // kSweepChain chain workgroups (FP32 factors: two 1024-thread workgroups
    // per CU, the kernel being held to 64 registers; FP64 factors: one)
    int dev = 0, ncu = 0;
    SIE_HIP(hipGetDevice(&dev));
    SIE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int per_cu = sizeof(TF) == 4 ? 2 : 1;
    const int bulk_max = std::max(1, per_cu * ncu - kSweepChain);
    unsigned int* werr = wait_word();
    // each sweep's output starts as kUnsolved words (a kernel, not a memset
    // node, so that the sequence can be captured and replayed)
    const long long nwords = (long long)n * (long long)(sizeof(cx<T>) / 8);
    auto unsolved = [&](cx<T>* out) {
        fill_unsolved_kernel<<<(unsigned)((nwords + 255) / 256), 256, 0, s>>>(nwords, (unsigned long long*)out);
    };
    // forward sweep over blocks 0, 1, ...: solve block 0, then kSweepChain blocks per launch
    auto forward = [&](const cx<T>* inv, int sop, cx<T>* vec, cx<T>* out) {
        unsolved(out);
        trsv_diag_kernel<T, B><<<1, kSweepThreads, 0, s>>>(0, width(0), inv, sop, vec, out);
        for (int kb = 0; kb < nb - 1; kb += kSweepChain) {
            const int D = std::min(kSweepChain, nb - 1 - kb);
            const int nbulk = std::min(std::max(0, nb - (kb + 1 + D)), bulk_max);
            trsv_chain_kernel<TF, T, B><<<D + nbulk, kSweepThreads, 0, s>>>(n, 1, kb, D, nbulk, sop, LU, lda, vec,
                                                                           inv, out, werr);
        }
    };
    // backward sweep over blocks nb-1, ..., 0
    auto backward = [&](const cx<T>* inv, int sop, cx<T>* vec, cx<T>* out) {
        const int kl = (nb - 1) * B;
        unsolved(out);
        trsv_diag_kernel<T, B><<<1, kSweepThreads, 0, s>>>(kl, width(kl), inv + (size_t)(nb - 1) * B * B, sop, vec,
                                                            out);
        for (int kb = nb - 1; kb > 0; kb -= kSweepChain) {
            const int D = std::min(kSweepChain, kb);
            const int nbulk = std::min(std::max(0, kb - D), bulk_max);
            trsv_chain_kernel<TF, T, B><<<D + nbulk, kSweepThreads, 0, s>>>(n, -1, kb, D, nbulk, sop, LU, lda, vec,
                                                                           inv, out, werr);
        }
    };
    if (op == 0) {
        permute_kernel<T><<<g, 256, 0, s>>>(n, perm, b, t, 0);  // t = P^T b
        forward(invL, 0, t, y);   // L y = t
        backward(invU, 0, y, b);  // U b = y
    } else {
        copy_vec_kernel<T><<<g, 256, 0, s>>>(n, b, t);
        forward(invU, 2, t, y);   // U^H y = b
        backward(invL, 2, y, t);  // L^H t = y
        permute_kernel<T><<<g, 256, 0, s>>>(n, perm, t, b, 1);  // b = P t
    }
    SIE_HIP(hipGetLastError());
}

struct SweepGraph {
    int tsize, op, n;
    const void* LU;
    long long lda;
    const void* b;
    const void* w;
    hipGraphExec_t exec;
};

struct SweepGraphCache {
    std::vector<SweepGraph> entries;
    hipStream_t capture = nullptr;
    // diagonal-block inverses held in the workspace
    const void* inv_LU = nullptr;
    int inv_n = 0, inv_tsize = 0;
    long long inv_lda = 0;
    unsigned long long inv_gen = 0;
    const void* inv_w = nullptr;
    // the row permutation held in the workspace
    const void* perm_LU = nullptr;
    const int* perm_ipiv = nullptr;
    int perm_n = 0;
    unsigned long long perm_gen = 0;
    const void* perm_w = nullptr;
    int perm_tsize = 0;
};
// Never destroyed (see ws()): graph executables and the capture stream are
// released with the process.
static SweepGraphCache& sweep_graphs() {
    static SweepGraphCache* c = new SweepGraphCache;
    return *c;
}

// Factors in TF, right-hand side, inverses and arithmetic in T.
template <typename TF, typename T>
void getrs_vec(int op, int n, const cx<TF>* LU, long long lda, const int* ipiv, cx<T>* b, hipStream_t s) {
    constexpr int B = SweepCfg<TF>::B;
    constexpr int tkey = (int)(sizeof(TF) * 100 + sizeof(T));  // cache key of the type pair
    const int nb = (n + B - 1) / B;
    const size_t invsz = (size_t)nb * B * B * sizeof(cx<T>);
    unsigned char* w = (unsigned char*)ws(5).get(4 * invsz + 2 * (size_t)n * sizeof(cx<T>) + (size_t)n * sizeof(int) + 16);
    cx<T>* invL = (cx<T>*)w;
    cx<T>* invU = (cx<T>*)(w + invsz);
    cx<T>* invLh = (cx<T>*)(w + 2 * invsz);
    cx<T>* invUh = (cx<T>*)(w + 3 * invsz);
    cx<T>* y = (cx<T>*)(w + 4 * invsz);
    cx<T>* t = y + n;
    int* perm = (int*)(t + n);
    SweepGraphCache& cache = sweep_graphs();
    const auto it = factor_generation().find(LU);
    const unsigned long long gen = it == factor_generation().end() ? 0 : it->second;
    // the row permutation of P^T, composed on the host from the swap sequence;
    // reused while the factors and pivots are those of the last factorisation
    // of this array (a copy of the pivots to the host and a pass over them per
    // solve otherwise)
    // (the permutation's place in the workspace depends on n and on the
    // vector type, so both are part of the key)
    if (!(gen != 0 && cache.perm_LU == LU && cache.perm_ipiv == ipiv && cache.perm_n == n && cache.perm_gen == gen &&
          cache.perm_w == w && cache.perm_tsize == tkey)) {
        std::vector<int> hp(n), hpiv(n);
        SIE_HIP(hipMemcpyAsync(hpiv.data(), ipiv, sizeof(int) * n, hipMemcpyDeviceToHost, s));
        SIE_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < n; ++i) hp[i] = i;
        for (int k = 0; k < n; ++k) {
            require(hpiv[k] >= k && hpiv[k] < n,
                    "pivot index outside the trailing rows (factors not from sie_lu_factor?)");
            std::swap(hp[k], hp[hpiv[k]]);
        }
        SIE_HIP(hipMemcpyAsync(perm, hp.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
        SIE_HIP(hipStreamSynchronize(s));  // hp goes out of scope
        cache.perm_LU = LU;
        cache.perm_ipiv = ipiv;
        cache.perm_n = n;
        cache.perm_gen = gen;
        cache.perm_w = w;
        cache.perm_tsize = tkey;
    }
    // inverses of the diagonal blocks: reused while the factors are those of
    // the last factorisation of this array
    if (!(gen != 0 && cache.inv_LU == LU && cache.inv_n == n && cache.inv_lda == lda &&
          cache.inv_tsize == tkey && cache.inv_gen == gen && cache.inv_w == w)) {
        diag_inv_kernel<TF, T, B><<<nb, B, 0, s>>>(n, LU, lda, 0, invL);
        diag_inv_kernel<TF, T, B><<<nb, B, 0, s>>>(n, LU, lda, 1, invU);
        const long long tot = (long long)nb * B * B;
        const unsigned gt = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
        conj_transpose_blocks_kernel<T, B><<<gt, 256, 0, s>>>(tot, invL, invLh);
        conj_transpose_blocks_kernel<T, B><<<gt, 256, 0, s>>>(tot, invU, invUh);
        SIE_HIP(hipGetLastError());
        cache.inv_LU = LU;
        cache.inv_n = n;
        cache.inv_lda = lda;
        cache.inv_tsize = tkey;
        cache.inv_gen = gen;
        cache.inv_w = w;
    }
    const cx<T>* iL = op == 0 ? invL : invLh;
    const cx<T>* iU = op == 0 ? invU : invUh;
    // timed as one unit (both sweeps); algorithmic bytes: every factor entry
    // once (L strictly below and U on and above the diagonal), the two
    // diagonal-block inverse sets, and the vector passes
    const double sweep_bytes = (double)n * n * sizeof(cx<TF>) + 2.0 * nb * B * B * sizeof(cx<T>) +
                               8.0 * n * sizeof(cx<T>);
    if (nb < 8) {  // short sequences: launch directly
        prof().begin(PROF_SWEEPS, s);
        record_sweeps<TF, T>(op, n, LU, lda, b, iL, iU, y, t, perm, s);
        prof().end(PROF_SWEEPS, s, 8.0 * n * n, sweep_bytes);
        return;
    }
    hipGraphExec_t exec = nullptr;
    for (auto& e : cache.entries)
        if (e.tsize == tkey && e.op == op && e.n == n && e.LU == LU && e.lda == lda && e.b == b && e.w == w)
            exec = e.exec;
    if (!exec) {
        // drop graphs that point into an older workspace, keep at most 16
        for (size_t i = 0; i < cache.entries.size();) {
            if (cache.entries[i].w != w || cache.entries.size() >= 16) {
                SIE_HIP(hipStreamSynchronize(s));
                (void)hipGraphExecDestroy(cache.entries[i].exec);
                cache.entries.erase(cache.entries.begin() + i);
            } else {
                ++i;
            }
        }
        if (!cache.capture) SIE_HIP(hipStreamCreateWithFlags(&cache.capture, hipStreamNonBlocking));
        hipGraph_t graph = nullptr;
        SIE_HIP(hipStreamBeginCapture(cache.capture, hipStreamCaptureModeThreadLocal));
        record_sweeps<TF, T>(op, n, LU, lda, b, iL, iU, y, t, perm, cache.capture);
        SIE_HIP(hipStreamEndCapture(cache.capture, &graph));
        SIE_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        SIE_HIP(hipGraphDestroy(graph));
        cache.entries.push_back(SweepGraph{tkey, op, n, LU, lda, b, w, exec});
    }
    prof().begin(PROF_SWEEPS, s);
    SIE_HIP(hipGraphLaunch(exec, s));
    prof().end(PROF_SWEEPS, s, 8.0 * n * n, sweep_bytes);
}

// One right-hand side in complex double through complex float LU factors
// (op 0 = N, 2 = C): the FP32 factors applied in FP64 arithmetic, a fixed
// linear operator whatever the vector.
void lu_solve_f32_f64(int op, int n, const float* LU, long long lda, const int* ipiv, double* b, hipStream_t s) {
    require(op == 0 || op == 2, "op must be 0 (N) or 2 (C)");
    getrs_vec<float, double>(op, n, (const c32*)LU, lda, ipiv, (c64*)b, s);
}

// B = conj(B) on an n x nrhs block.
template <typename T>
__global__ void conj_kernel(int n, int nrhs, cx<T>* __restrict__ B, long long ldb) {
    const long long tot = (long long)n * nrhs;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / n, r = t - c * n;
        B[c * ldb + r].im = -B[c * ldb + r].im;
    }
}

// D = S^H for an r x c block S (column-major, lds) into D (c x r, ldd):
// 32 x 32 tiles through LDS, loads and stores coalesced.
template <typename T>
__global__ __launch_bounds__(256) void conj_transpose_kernel(int r, int c, const cx<T>* __restrict__ S, long long lds,
                                                             cx<T>* __restrict__ D, long long ldd) {
    __shared__ cx<T> t[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
    const long long r0 = (long long)blockIdx.x * 32, c0 = (long long)blockIdx.y * 32;
    for (int k = ty; k < 32; k += 8) {
        const long long i = r0 + tx, j = c0 + k;  // S(i, j)
        if (i < r && j < c) t[k][tx] = S[j * lds + i];
    }
    __syncthreads();
    for (int k = ty; k < 32; k += 8) {
        const long long a = c0 + tx, b = r0 + k;  // D(a, b) = conj(S(b, a))
        if (a < c && b < r) {
            const cx<T> v = t[tx][k];
            D[b * ldd + a] = cx<T>{v.re, -v.im};
        }
    }
}

template <typename T>
void getrs(int op, int n, int nrhs, const cx<T>* LU, long long lda, const int* ipiv, cx<T>* B, long long ldb,
           hipStream_t s, bool split_f16 = false);

// Solve op(A) X = B with the factors of getrf (op 0 = N, 1 = T, 2 = C).
// A^T X = B is solved as A^H conj(X) = conj(B). split_f16 (FP32 factors, op
// N, more than 4 columns): the products beyond each 256-row block on the
// split-FP16 matrix cores, whose one power-of-two scale per operand makes
// them accurate relative to the largest entry of all columns; for callers
// whose columns are of one scale (the blocked refinement's normalised
// corrections). Otherwise those products are FP32.
template <typename T>
void getrs(int op, int n, int nrhs, const cx<T>* LU, long long lda, const int* ipiv, cx<T>* B, long long ldb,
           hipStream_t s, bool split_f16) {
    constexpr int BS = 32;
    if (op == 1) {
        const unsigned g = (unsigned)std::min<long long>(((long long)n * nrhs + 255) / 256, 1 << 16);
        conj_kernel<T><<<g, 256, 0, s>>>(n, nrhs, B, ldb);
        getrs<T>(2, n, nrhs, LU, lda, ipiv, B, ldb, s, false);
        conj_kernel<T><<<g, 256, 0, s>>>(n, nrhs, B, ldb);
        SIE_HIP(hipGetLastError());
        return;
    }
    if (nrhs <= 4) {
        for (int c = 0; c < nrhs; ++c) getrs_vec<T, T>(op, n, LU, lda, ipiv, B + (long long)c * ldb, s);
        return;
    }
    if (op == 0) {
        // the interchanges, composed per NB of them
        for (int k = 0; k < n; k += kNB) laswp2<T>(B, ldb, n, 0, nrhs, 0, 0, ipiv, k, std::min(k + kNB, n), s);
        // NB-row blocks: inside a block, BS-row triangles and their K = BS
        // updates of the block's other rows; the rows beyond the block take one
        // K = NB product
        auto block_update = [&](int M, const cx<T>* Lk, cx<T>* Bk, cx<T>* Ck, int K, float a_bound) {
            if (M <= 0) return;
            if constexpr (std::is_same<T, float>::value) {
                if (split_f16) {
                    gemm_f16x3(M, nrhs, K, -1.0, Lk, lda, Bk, ldb, 1.0, Ck, ldb, s, a_bound);
                    return;
                }
            }
            (void)a_bound;
            gemm<T>(M, nrhs, K, -1.0, Lk, lda, Bk, ldb, 1.0, Ck, ldb, s, 0);
        };
        // L y = b (unit lower, multipliers bounded by sqrt 2), forward
        for (int kb = 0; kb < n; kb += kNB) {
            const int nb = std::min(kNB, n - kb);
            for (int k = kb; k < kb + nb; k += BS) {
                const int w = std::min(BS, kb + nb - k);
                trsm_small<T>(w, LU + (long long)k * lda + k, lda, B + k, ldb, nrhs, 0, s);
                const int below = kb + nb - k - w;
                if (below > 0)
                    gemm<T>(below, nrhs, w, -1.0, LU + (long long)k * lda + k + w, lda, B + k, ldb, 1.0, B + k + w,
                            ldb, s, 0);
            }
            block_update(n - kb - nb, LU + (long long)kb * lda + kb + nb, B + kb, B + kb + nb, nb, 1.5f);
        }
        // U x = y, backward
        for (int kb = ((n - 1) / kNB) * kNB; kb >= 0; kb -= kNB) {
            const int nb = std::min(kNB, n - kb);
            for (int k = kb + ((nb - 1) / BS) * BS; k >= kb; k -= BS) {
                const int w = std::min(BS, kb + nb - k);
                trsm_small<T>(w, LU + (long long)k * lda + k, lda, B + k, ldb, nrhs, 1, s);
                if (k > kb)
                    gemm<T>(k - kb, nrhs, w, -1.0, LU + (long long)k * lda + kb, lda, B + k, ldb, 1.0, B + kb, ldb,
                            s, 0);
            }
            block_update(kb, LU + (long long)kb * lda, B + kb, B, nb, 0.f);
        }
    } else {
        // A^H = U^H L^H P^T: U^H z = b (forward), L^H w = z (backward), x = P w.
        // Right-looking in NB-row blocks as for op N; the products take the
        // needed rows of U (columns of L) conjugate-transposed into a
        // temporary: C -= S^H X with S an r x c block of the factors.
        cx<T>* tmp = (cx<T>*)ws(2).get((size_t)n * kNB * sizeof(cx<T>));
        auto update_h = [&](int r, int c, const cx<T>* S, const cx<T>* X, cx<T>* C) {
            if (r <= 0 || c <= 0) return;
            conj_transpose_kernel<T><<<dim3((unsigned)((r + 31) / 32), (unsigned)((c + 31) / 32)), 256, 0, s>>>(
                r, c, S, lda, tmp, c);
            gemm<T>(c, nrhs, r, -1.0, tmp, c, X, ldb, 1.0, C, ldb, s, 0);
        };
        // U^H z = b: U^H is lower; the rows of U of a solved block reach the
        // entries below it
        for (int kb = 0; kb < n; kb += kNB) {
            const int nb = std::min(kNB, n - kb);
            for (int k = kb; k < kb + nb; k += BS) {
                const int w = std::min(BS, kb + nb - k);
                trsm_small<T>(w, LU + (long long)k * lda + k, lda, B + k, ldb, nrhs, 2, s);
                update_h(w, kb + nb - k - w, LU + (long long)(k + w) * lda + k, B + k, B + k + w);
            }
            update_h(nb, n - kb - nb, LU + (long long)(kb + nb) * lda + kb, B + kb, B + kb + nb);
        }
        // L^H w = z: L^H is unit upper; the columns of L of a solved block
        // reach the entries above it
        for (int kb = ((n - 1) / kNB) * kNB; kb >= 0; kb -= kNB) {
            const int nb = std::min(kNB, n - kb);
            for (int k = kb + ((nb - 1) / BS) * BS; k >= kb; k -= BS) {
                const int w = std::min(BS, kb + nb - k);
                trsm_small<T>(w, LU + (long long)k * lda + k, lda, B + k, ldb, nrhs, 3, s);
                update_h(w, k - kb, LU + (long long)kb * lda + k, B + k, B + kb);
            }
            update_h(nb, kb, LU + kb, B + kb, B);
        }
        // x = P w: the interchanges backwards, NB of them at a time, last first
        for (int k = ((n - 1) / kNB) * kNB; k >= 0; k -= kNB) {
            const int ke = std::min(k + kNB, n);
            laswp_perm_kernel<T, kNB><<<(nrhs + 15) / 16, 256, 0, s>>>(B, ldb, n, 0, nrhs, 0, 0, ipiv, k, ke, 16, 1);
        }
    }
    SIE_HIP(hipGetLastError());
}

// y = beta y + alpha op(A) x for an m x n FP64 matrix (op 0 = N, 2 = C):
// per-chunk partial sums (chunks of 2048 columns for op N, 8192 rows for op C)
// into part, then a fixed-order reduction, so the result is deterministic.
constexpr int kGemvChunkN = 2048, kGemvChunkC = 8192;
static size_t zgemv_partial_bytes(int op, int m, int n) {
    return op == 0 ? (size_t)((n + kGemvChunkN - 1) / kGemvChunkN) * m * sizeof(cx<double>)
                   : (size_t)((m + kGemvChunkC - 1) / kGemvChunkC) * n * sizeof(cx<double>);
}
static void zgemv_launch(int op, int m, int n, double alpha, const c64* A, long long lda, const c64* x, double beta,
                         c64* y, cx<double>* part, hipStream_t s) {
    if (op == 0) {
        const int nch = (n + kGemvChunkN - 1) / kGemvChunkN;
        gemv_n_partial<double, double><<<dim3((m + 255) / 256, nch), 256, 0, s>>>(m, n, kGemvChunkN, A, lda, x, part);
        gemv_finish<double><<<(m + 255) / 256, 256, 0, s>>>(m, nch, part, alpha, beta, y);
    } else {
        const int nch = (m + kGemvChunkC - 1) / kGemvChunkC;
        gemv_c_partial<double, double><<<dim3((n + 3) / 4, nch), 256, 0, s>>>(m, n, kGemvChunkC, A, lda, x, part);
        gemv_finish<double><<<(n + 255) / 256, 256, 0, s>>>(n, nch, part, alpha, beta, y);
    }
}
static void zgemv(int op, int m, int n, double alpha, const c64* A, long long lda, const c64* x, double beta, c64* y,
                  hipStream_t s) {
    // timed: products of at least 2^24 entries (the operator's, not the Krylov basis')
    const bool timed = (double)m * n >= 16777216.0;
    cx<double>* part = (cx<double>*)ws(3).get(zgemv_partial_bytes(op, m, n));
    if (timed) prof().begin(PROF_GEMV, s);
    zgemv_launch(op, m, n, alpha, A, lda, x, beta, y, part, s);
    // algorithmic bytes: A once, x and y (read + write)
    if (timed) prof().end(PROF_GEMV, s, 8.0 * m * n, 16.0 * ((double)m * n + (op == 0 ? n + 2.0 * m : m + 2.0 * n)));
    SIE_HIP(hipGetLastError());
}

// ------------------------------------------------- GEMV without allocation
// The partial sums of sie_zgemv_noalloc live in this statically sized device
// array ("scratch"): each device the library runs on holds its own copy until
// the process ends. Host code never uses the symbol itself as an address (it
// is a host-side stand-in): the device address is looked up once per device
// with hipGetSymbolAddress, on the first call with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mutex;

static cx<double>* scratch_of_current_device() {
    int dev = 0;
    SIE_HIP(hipGetDevice(&dev));
    require(dev >= 0 && dev < kMaxDevices, "device index beyond the library's device table");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(SIE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    return (cx<double>*)g_scratch_ptr[dev];
}

// Blocked FP64 refinement of Z X = B around FP32 LU factors (op N): per sweep
// R = B - Z X (FP64 GEMM), column norms on the host, D = LU32^{-1} (R / |r_c|)
// for all columns at once, X += |r_c| D. Stops when every column reaches tol.
static double norm2(long long n, const c64* x, hipStream_t s);
static void ir_blocked(int n, int nrhs, const c64* Z, long long ldz, const c32* LU, long long ldl, const int* ipiv,
                       const c64* B, long long ldb, c64* X, long long ldx, double tol, double znorm, int maxit,
                       double* rel_out, double* be_out, int* it_out, hipStream_t s) {
    DevBuf<c64> R((size_t)n * nrhs);
    DevBuf<c32> D((size_t)n * nrhs);
    DevBuf<double> nrm(nrhs), sc(nrhs), gain(nrhs);
    std::vector<double> bn(nrhs), rn(nrhs), xn(nrhs, 0.0), hs(nrhs), hg(nrhs);
    const long long tot = (long long)n * nrhs;
    const unsigned g = (unsigned)std::min<long long>((tot + 255) / 256, 1 << 16);
    auto column_norms = [&](const c64* M, long long ld, std::vector<double>& out) {
        colsumsq_kernel<<<nrhs, 256, 0, s>>>(n, M, ld, nrm.p);
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipMemcpyAsync(out.data(), nrm.p, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
        SIE_HIP(hipStreamSynchronize(s));
        for (double& v : out) v = std::sqrt(v);
    };
    column_norms(B, ldb, bn);
    for (int c = 0; c < nrhs; ++c)
        if (!(bn[c] > 0) || !std::isfinite(bn[c]))
            throw Error(SIE_ERR_NONFINITE, "right-hand side norm must be finite and nonzero");
    for (int c = 0; c < nrhs; ++c) SIE_HIP(hipMemsetAsync(X + (long long)c * ldx, 0, sizeof(c64) * n, s));
    for (int c = 0; c < nrhs; ++c)
        SIE_HIP(hipMemcpyAsync(R.p + (size_t)c * n, B + (long long)c * ldb, sizeof(c64) * n, hipMemcpyDeviceToDevice, s));
    rn = bn;
    std::vector<int> its(nrhs, 0);
    std::vector<char> done(nrhs, 0);
    for (int it = 0; it < maxit; ++it) {
        int active = 0;
        for (int c = 0; c < nrhs; ++c) {
            if (!done[c] && rn[c] <= tol * (znorm * xn[c] + bn[c])) done[c] = 1;
            // converged columns get a zero correction (their iterate is kept)
            hs[c] = done[c] ? 0.0 : 1.0 / rn[c];
            hg[c] = done[c] ? 0.0 : rn[c];
            if (!done[c]) { ++active; ++its[c]; }
        }
        if (!active) break;
        SIE_HIP(hipMemcpyAsync(sc.p, hs.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
        SIE_HIP(hipMemcpyAsync(gain.p, hg.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, s));
        scale_cols_f32_kernel<<<g, 256, 0, s>>>(n, nrhs, R.p, n, sc.p, D.p, n);
        SIE_HIP(hipGetLastError());
        getrs<float>(0, n, nrhs, LU, ldl, ipiv, D.p, n, s, true);  // columns normalised above
        axpy_cols_kernel<<<g, 256, 0, s>>>(n, nrhs, D.p, n, gain.p, X, ldx);
        SIE_HIP(hipGetLastError());
        // R = B - Z X
        for (int c = 0; c < nrhs; ++c)
            SIE_HIP(hipMemcpyAsync(R.p + (size_t)c * n, B + (long long)c * ldb, sizeof(c64) * n,
                                   hipMemcpyDeviceToDevice, s));
        gemm<double>(n, nrhs, n, -1.0, Z, ldz, X, ldx, 1.0, R.p, n, s, 0);
        column_norms(R.p, n, rn);
        for (int c = 0; c < nrhs; ++c)
            if (!std::isfinite(rn[c])) throw Error(SIE_ERR_NONFINITE, "iterative refinement diverged (non-finite residual)");
        if (znorm > 0) column_norms(X, ldx, xn);
    }
    for (int c = 0; c < nrhs; ++c) {
        if (rel_out) rel_out[c] = rn[c] / bn[c];
        if (be_out) be_out[c] = rn[c] / (znorm * xn[c] + bn[c]);
        if (it_out) it_out[c] = its[c];
    }
}

void check_device_waits(hipStream_t s) { check_waits(s); }

static double norm2(long long n, const c64* x, hipStream_t s) {
    const int nb = 512;
    double* part = (double*)ws(0).get(nb * sizeof(double));
    sumsq_partial<<<nb, 256, 0, s>>>(n, x, part);
    SIE_HIP(hipGetLastError());
    double h[nb];
    SIE_HIP(hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s));
    SIE_HIP(hipStreamSynchronize(s));
    long double a = 0;
    for (int i = 0; i < nb; ++i) a += h[i];
    return std::sqrt((double)a);
}

}  // namespace sie

using namespace sie;

#define GUARD_BEGIN try {
#define GUARD_END                     \
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

int sie_gemm(int precision, int m, int n, int k, double alpha, const void* A, long long lda, const void* B,
             long long ldb, double beta, void* C, long long ldc, int variant, void* stream) {
    GUARD_BEGIN
    require(precision == 32 || precision == 64, "precision must be 32 or 64");
    require(m >= 0 && n >= 0 && k >= 0 && lda >= std::max(1, m) && ldb >= std::max(1, k) && ldc >= std::max(1, m),
            "invalid GEMM dimensions");
    if (precision == 32 && variant == 2)
        gemm_f16x3(m, n, k, alpha, (const c32*)A, lda, (const c32*)B, ldb, beta, (c32*)C, ldc, (hipStream_t)stream);
    else if (precision == 32)
        gemm<float>(m, n, k, alpha, (const c32*)A, lda, (const c32*)B, ldb, beta, (c32*)C, ldc, (hipStream_t)stream,
                    variant);
    else
        gemm<double>(m, n, k, alpha, (const c64*)A, lda, (const c64*)B, ldb, beta, (c64*)C, ldc, (hipStream_t)stream,
                     0);
    GUARD_END
}

int sie_lu_factor(int precision, int n, void* A, long long lda, int* ipiv, int* info, int variant, void* stream) {
    GUARD_BEGIN
    require(precision == 32 || precision == 64, "precision must be 32 or 64");
    require(n > 0, "DenseLuSolver dimension must be positive");
    require(lda >= n, "leading dimension smaller than the matrix dimension");
    require(A && ipiv && info, "null argument");
    hipStream_t s = (hipStream_t)stream;
    int* d_info = (int*)ws(4).get(sizeof(int) * 2);
    if (precision == 32)
        getrf<float>(n, (c32*)A, lda, ipiv, d_info, s, variant);
    else
        getrf<double>(n, (c64*)A, lda, ipiv, d_info, s, 0);
    SIE_HIP(hipMemcpyAsync(info, d_info, sizeof(int), hipMemcpyDeviceToHost, s));
    check_waits(s);
    if (*info > 0)
        throw Error(SIE_ERR_SINGULAR, "dense matrix is singular at U diagonal " + std::to_string(*info));
    GUARD_END
}

int sie_lu_solve(int precision, int op, int n, int nrhs, const void* LU, long long lda, const int* ipiv, void* B,
                 long long ldb, void* stream) {
    GUARD_BEGIN
    require(precision == 32 || precision == 64, "precision must be 32 or 64");
    require(op == 0 || op == 1 || op == 2, "op must be 0 (N), 1 (T) or 2 (C)");
    require(n > 0 && nrhs > 0 && lda >= n && ldb >= n, "invalid solve dimensions");
    hipStream_t s = (hipStream_t)stream;
    if (precision == 32)
        getrs<float>(op, n, nrhs, (const c32*)LU, lda, ipiv, (c32*)B, ldb, s);
    else
        getrs<double>(op, n, nrhs, (const c64*)LU, lda, ipiv, (c64*)B, ldb, s);
    if (nrhs <= 4) check_waits(s);  // the single-RHS sweeps wait on published blocks
    GUARD_END
}

int sie_release(void) {
    GUARD_BEGIN
    SweepGraphCache& c = sweep_graphs();
    bool held = g_wait_word != nullptr || c.capture != nullptr || !c.entries.empty();
    for (int i = 0; i < 8; ++i) held = held || ws(i).buf.p != nullptr;
    if (!held) {
        set_last_error("");
        return SIE_OK;
    }
    SIE_HIP(hipDeviceSynchronize());
    for (auto& e : c.entries) SIE_HIP(hipGraphExecDestroy(e.exec));
    c.entries.clear();
    if (c.capture) SIE_HIP(hipStreamDestroy(c.capture));
    c.capture = nullptr;
    c.inv_LU = nullptr;
    c.inv_w = nullptr;
    c.inv_gen = 0;
    c.perm_LU = nullptr;
    c.perm_w = nullptr;
    c.perm_gen = 0;
    for (int i = 0; i < 8; ++i) ws(i).buf.release();
    if (g_wait_word) SIE_HIP(hipFree(g_wait_word));
    g_wait_word = nullptr;
    for (int k = 0; k < 64; ++k)
        if (sie_profile_enable(k, 0) != SIE_OK) break;  // destroys recorded events
    GUARD_END
}

int sie_set_wait_bound(double seconds, int force_timeout) {
    GUARD_BEGIN
    require(seconds > 0 && seconds <= 3600, "wait bound must be in (0, 3600] seconds");
    wait_ctl().seconds = seconds;
    wait_ctl().force = force_timeout ? 1 : 0;
    GUARD_END
}

int sie_set_panel_launch(int cooperative) {
    GUARD_BEGIN
    g_cooperative_panels = cooperative != 0;
    GUARD_END
}

int sie_zgemv(int op, int m, int n, double alpha, const double* A, long long lda, const double* x, double beta,
              double* y, void* stream) {
    GUARD_BEGIN
    require(op == 0 || op == 2, "op must be 0 (N) or 2 (C)");
    require(m > 0 && n > 0 && lda >= m, "invalid GEMV dimensions");
    zgemv(op, m, n, alpha, (const c64*)A, lda, (const c64*)x, beta, (c64*)y, (hipStream_t)stream);
    GUARD_END
}

int sie_zgemv_noalloc(int op, int m, int n, double alpha, const double* A, long long lda, const double* x,
                      double beta, double* y, void* stream) {
    GUARD_BEGIN
    require(op == 0 || op == 2, "op must be 0 (N) or 2 (C)");
    require(m > 0 && n > 0 && lda >= m, "invalid GEMV dimensions");
    require(A && x && y, "null GEMV argument");
    const size_t need = zgemv_partial_bytes(op, m, n);
    if (need > kNoAllocScratchBytes)
        throw Error(SIE_ERR_INVALID, "sie_zgemv_noalloc: " + std::to_string(m) + " x " + std::to_string(n) +
                                         " needs " + std::to_string(need) + " bytes of partial sums, above the " +
                                         std::to_string(kNoAllocScratchBytes) + "-byte static scratch");
    hipStream_t s = (hipStream_t)stream;
    zgemv_launch(op, m, n, alpha, (const c64*)A, lda, (const c64*)x, beta, (c64*)y, scratch_of_current_device(), s);
    SIE_HIP(hipGetLastError());
    GUARD_END
}

int sie_convert(int from_precision, int to_precision, long long rows, long long cols, const void* A, long long lda,
                void* B, long long ldb, double scale, double* fro_norm, void* stream) {
    GUARD_BEGIN
    require(rows > 0 && cols > 0 && lda >= rows && ldb >= rows, "invalid conversion dimensions");
    hipStream_t s = (hipStream_t)stream;
    const long long t = rows * cols;
    const unsigned g = (unsigned)std::min<long long>((t + 255) / 256, 1 << 16);
    if (fro_norm) {
        require(from_precision == 64 && (to_precision == 32 || to_precision == 64),
                "the Frobenius norm is formed for complex double sources only");
        double* part = (double*)ws(0).get(sizeof(double) * (1 << 16));
        if (to_precision == 32)
            convert_fro_kernel<float><<<g, 256, 0, s>>>(rows, cols, (const c64*)A, lda, (c32*)B, ldb, scale, part);
        else
            convert_fro_kernel<double><<<g, 256, 0, s>>>(rows, cols, (const c64*)A, lda, (c64*)B, ldb, scale, part);
        SIE_HIP(hipGetLastError());
        std::vector<double> h(g);
        SIE_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * g, hipMemcpyDeviceToHost, s));
        SIE_HIP(hipStreamSynchronize(s));
        long double acc = 0;
        for (unsigned i = 0; i < g; ++i) acc += h[i];
        *fro_norm = std::sqrt((double)acc);
        if (!std::isfinite(*fro_norm)) throw Error(SIE_ERR_NONFINITE, "matrix has non-finite entries");
        set_last_error("");
        return SIE_OK;
    }
    if (from_precision == 64 && to_precision == 32)
        convert_kernel<double, float><<<g, 256, 0, s>>>(rows, cols, (const c64*)A, lda, (c32*)B, ldb, scale);
    else if (from_precision == 32 && to_precision == 64)
        convert_kernel<float, double><<<g, 256, 0, s>>>(rows, cols, (const c32*)A, lda, (c64*)B, ldb, scale);
    else if (from_precision == 64 && to_precision == 64)
        convert_kernel<double, double><<<g, 256, 0, s>>>(rows, cols, (const c64*)A, lda, (c64*)B, ldb, scale);
    else
        throw Error(SIE_ERR_INVALID, "unsupported conversion");
    SIE_HIP(hipGetLastError());
    GUARD_END
}

// FP64 iterative refinement of op(Z) X = B around an FP32 LU of Z.
int sie_ir_solve(int op, int n, int nrhs, const double* Z, long long ldz, const float* LU, long long ldl,
                 const int* ipiv, const double* B, long long ldb, double* X, long long ldx, double tol, double znorm,
                 int maxit, double* rel_residual, double* backward_error, int* iterations, void* stream) {
    GUARD_BEGIN
    require(op == 0 || op == 2, "op must be 0 (N) or 2 (C)");
    require(n > 0 && nrhs > 0 && ldz >= n && ldl >= n && ldb >= n && ldx >= n, "invalid IR dimensions");
    require(tol > 0 && std::isfinite(tol), "refinement tolerance must be positive");
    require(znorm >= 0 && std::isfinite(znorm), "refinement matrix norm must be finite and non-negative");
    hipStream_t s = (hipStream_t)stream;
    if (op == 0 && nrhs > 4) {
        // many right-hand sides: every sweep is one FP64 GEMM residual and one
        // multi-RHS FP32 solve on the matrix cores, columns scaled to unit norm
        ir_blocked(n, nrhs, (const c64*)Z, ldz, (const c32*)LU, ldl, ipiv, (const c64*)B, ldb, (c64*)X, ldx, tol,
                   znorm, maxit, rel_residual, backward_error, iterations, s);
        set_last_error("");
        return SIE_OK;
    }
    DevBuf<c64> r(n);
    DevBuf<c32> d32(n);
    const unsigned g = (unsigned)((n + 255) / 256);
    for (int c = 0; c < nrhs; ++c) {
        const c64* b = (const c64*)B + (long long)c * ldb;
        c64* x = (c64*)X + (long long)c * ldx;
        const double bn = norm2(n, b, s);
        if (!(bn > 0) || !std::isfinite(bn)) throw Error(SIE_ERR_NONFINITE, "right-hand side norm must be finite and nonzero");
        SIE_HIP(hipMemsetAsync(x, 0, sizeof(c64) * n, s));
        SIE_HIP(hipMemcpyAsync(r.p, b, sizeof(c64) * n, hipMemcpyDeviceToDevice, s));
        double rn = bn, rel = 1.0, xn = 0.0;
        int it = 0;
        for (; it < maxit; ++it) {
            // d = op(LU32)^{-1} (r / |r|), x += |r| d
            convert_kernel<double, float><<<g, 256, 0, s>>>(n, 1, r.p, n, d32.p, n, 1.0 / rn);
            getrs<float>(op, n, 1, (const c32*)LU, ldl, ipiv, d32.p, n, s);
            axpy_kernel<float><<<g, 256, 0, s>>>(n, rn, d32.p, x);
            // r = b - op(Z) x
            zgemv(op, n, n, 1.0, (const c64*)Z, ldz, x, 0.0, r.p, s);
            residual_kernel<<<g, 256, 0, s>>>(n, b, r.p);
            rn = norm2(n, r.p, s);
            check_waits(s);
            if (!std::isfinite(rn)) throw Error(SIE_ERR_NONFINITE, "iterative refinement diverged (non-finite residual)");
            if (znorm > 0) xn = norm2(n, x, s);
            rel = rn / bn;
            // normwise backward error test |r| <= tol (znorm |x| + |b|)
            if (rn <= tol * (znorm * xn + bn)) { ++it; break; }
        }
        if (rel_residual) rel_residual[c] = rel;
        if (backward_error) backward_error[c] = rn / (znorm * xn + bn);
        if (iterations) iterations[c] = it;
    }
    GUARD_END
}

double sie_norm2(long long n, const double* x, void* stream) {
    try {
        return norm2(n, (const c64*)x, (hipStream_t)stream);
    } catch (...) {
        return -1.0;
    }
}

int sie_mfma_f16_rate(int iters, int reps, double* tflops, double* clock_mhz, void* stream) {
    GUARD_BEGIN
    require(iters > 0 && reps > 0 && tflops && clock_mhz, "invalid calibration arguments");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 0;
    SIE_HIP(hipGetDevice(&dev));
    SIE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = ncu;  // one 512-thread workgroup per CU (its registers allow no second)
    DevBuf<float> sink((size_t)grid * 512);
    DevBuf<unsigned long long> clocks((size_t)grid * 2);
    hipEvent_t e0, e1;
    SIE_HIP(hipEventCreate(&e0));
    SIE_HIP(hipEventCreate(&e1));
    double best = 0.0, mhz = 0.0;
    std::vector<unsigned long long> hc((size_t)grid * 2);
    for (int r = 0; r < reps; ++r) {
        SIE_HIP(hipEventRecord(e0, s));
        mfma_f16_rate_kernel<<<grid, 512, 0, s>>>(iters, sink.p, clocks.p);
        SIE_HIP(hipGetLastError());
        SIE_HIP(hipEventRecord(e1, s));
        SIE_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SIE_HIP(hipEventElapsedTime(&ms, e0, e1));
        // 48 MFMAs of 2 * 32 * 32 * 16 flops per wave and iteration, 8 waves
        const double fl = (double)grid * 8 * iters * 48 * 32768.0;
        const double tf = fl / (ms * 1e-3) / 1e12;
        if (tf > best) {
            best = tf;
            SIE_HIP(hipMemcpy(hc.data(), clocks.p, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
            double cyc = 0.0, ticks = 0.0;
            for (int b = 0; b < grid; ++b) {
                cyc += (double)hc[2 * b];
                ticks += (double)hc[2 * b + 1];
            }
            mhz = ticks > 0.0 ? cyc / ticks * 100.0 : 0.0;
        }
    }
    SIE_HIP(hipEventDestroy(e0));
    SIE_HIP(hipEventDestroy(e1));
    *tflops = best;
    *clock_mhz = mhz;
    GUARD_END
}

}  // extern "C"
