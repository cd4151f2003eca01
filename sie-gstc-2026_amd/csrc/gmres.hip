// Restarted GMRES(m) on the GPU for complex double systems A x = b.
//
// Same contract as the reference's gmres_restarted / GMRESSolver
// (include/sie_mom/cuda/gmres/gmres_solver.cuh:258-537): x holds the initial
// guess on entry and the solution on exit; the tolerance is relative to |b|;
// a zero right-hand side (|b| < 1e-20) returns x = 0; the returned residual is
// the absolute residual estimate |g_{k+1}| of the last Arnoldi step; the
// iteration count is the total number of inner (Arnoldi) steps; the history
// holds the relative residual after every inner step; info is 0 when
// converged and 1 otherwise.
//
// MI355X form: the Krylov basis V (n x (m+1), column-major) stays in HBM and
// each Arnoldi step orthogonalises against all previous vectors at once by
// classical Gram-Schmidt with one reorthogonalisation (CGS2): h = V^H w and
// w -= V h, twice, as two GEMV calls each — four HBM passes over the basis
// instead of the 2(k+1) small dot / axpy launches and k+1 host round trips of
// modified Gram-Schmidt, with the same orthogonality in floating point. The
// (m+1) x m Hessenberg matrix, the Givens rotations and the triangular solve
// are host work of O(m^2).
#include "sie_internal.hpp"

#include <cmath>
#include <complex>
#include <vector>

namespace sie {
namespace {

using zc = std::complex<double>;

__global__ void gm_scale_kernel(long long n, double a, double* __restrict__ x) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += (long long)gridDim.x * blockDim.x)
        x[i] *= a;
}

// r = b - r
__global__ void gm_residual_kernel(long long n, const double* __restrict__ b, double* __restrict__ r) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += (long long)gridDim.x * blockDim.x)
        r[i] = b[i] - r[i];
}

// y = x / a
__global__ void gm_copy_scaled_kernel(long long n, const double* __restrict__ x, double a, double* __restrict__ y) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < 2 * n; i += (long long)gridDim.x * blockDim.x)
        y[i] = x[i] / a;
}

unsigned blocks_for(long long n) { return (unsigned)std::min<long long>((2 * n + 255) / 256, 1 << 14); }

void check(int code) {
    if (code != SIE_OK) throw Error(code, sie_last_error());
}

// Givens rotation (c real, s complex) with [c s; -conj(s) c] [a; b] = [r; 0].
void givens(zc a, zc b, double& c, zc& s) {
    const double aa = std::abs(a), ab = std::abs(b);
    if (ab == 0.0) {
        c = 1.0;
        s = 0.0;
    } else if (aa == 0.0) {
        c = 0.0;
        s = std::conj(b) / ab;
    } else {
        const double rho = std::hypot(aa, ab);
        c = aa / rho;
        s = (a / aa) * std::conj(b) / rho;
    }
}

}  // namespace

int gmres_run(int n, int m, int maxnit, double tol, sie_matvec_fn matvec, void* user, const double* b, double* x,
              double* out_res, int* out_iters, double* history, int history_cap, int* history_len,
              hipStream_t s) {
    require(n > 0, "GMRES dimension must be positive");
    require(m > 0 && maxnit > 0, "GMRES restart length and restart count must be positive");
    require(tol > 0 && std::isfinite(tol), "GMRES tolerance must be positive");
    require(matvec != nullptr && b != nullptr && x != nullptr, "null GMRES argument");
    int hist_n = 0;
    auto push_history = [&](double v) {
        if (history && hist_n < history_cap) history[hist_n] = v;
        ++hist_n;
    };
    const double bn = sie_norm2(n, b, s);
    if (bn < 0) throw Error(SIE_ERR_DEVICE, sie_last_error());
    if (!std::isfinite(bn)) throw Error(SIE_ERR_NONFINITE, "GMRES right-hand side is not finite");
    if (bn < 1e-20) {
        SIE_HIP(hipMemsetAsync(x, 0, sizeof(double) * 2 * (size_t)n, s));
        SIE_HIP(hipStreamSynchronize(s));
        if (out_res) *out_res = 0.0;
        if (out_iters) *out_iters = 0;
        if (history_len) *history_len = 0;
        return 0;
    }
    const double res_tol = tol * bn;
    DevBuf<double> V((size_t)2 * n * (m + 1));
    DevBuf<double> r((size_t)2 * n);
    DevBuf<double> h((size_t)2 * (m + 1)), h2((size_t)2 * (m + 1));
    const int ldH = m + 1;
    std::vector<zc> H((size_t)ldH * m), sn(m), g(ldH), y(m), hv(m + 1), hv2(m + 1);
    std::vector<double> cs(m);
    const unsigned gb = blocks_for(n);
    int info = 1, total = 0;
    double cur = bn;
    for (int outer = 0; outer < maxnit; ++outer) {
        // r = b - A x
        check(matvec(user, x, r.p, s));
        gm_residual_kernel<<<gb, 256, 0, s>>>(n, b, r.p);
        SIE_HIP(hipGetLastError());
        const double beta = sie_norm2(n, r.p, s);
        if (!std::isfinite(beta) || beta < 0) throw Error(SIE_ERR_NONFINITE, "GMRES residual is not finite");
        cur = beta;
        if (cur <= res_tol) {
            info = 0;
            break;
        }
        gm_copy_scaled_kernel<<<gb, 256, 0, s>>>(n, r.p, beta, V.p);
        std::fill(g.begin(), g.end(), zc(0));
        g[0] = beta;
        int k = 0;
        for (; k < m; ++k) {
            ++total;
            double* vk = V.p + (size_t)2 * n * k;
            double* w = V.p + (size_t)2 * n * (k + 1);
            check(matvec(user, vk, w, s));
            // CGS2: h = V^H w, w -= V h (twice)
            check(sie_zgemv(2, n, k + 1, 1.0, V.p, n, w, 0.0, h.p, s));
            check(sie_zgemv(0, n, k + 1, -1.0, V.p, n, h.p, 1.0, w, s));
            check(sie_zgemv(2, n, k + 1, 1.0, V.p, n, w, 0.0, h2.p, s));
            check(sie_zgemv(0, n, k + 1, -1.0, V.p, n, h2.p, 1.0, w, s));
            SIE_HIP(hipMemcpyAsync(hv.data(), h.p, sizeof(zc) * (k + 1), hipMemcpyDeviceToHost, s));
            SIE_HIP(hipMemcpyAsync(hv2.data(), h2.p, sizeof(zc) * (k + 1), hipMemcpyDeviceToHost, s));
            const double hn = sie_norm2(n, w, s);  // synchronises the stream
            if (!std::isfinite(hn) || hn < 0) throw Error(SIE_ERR_NONFINITE, "GMRES Arnoldi vector is not finite");
            zc* col = &H[(size_t)k * ldH];
            for (int j = 0; j <= k; ++j) col[j] = hv[j] + hv2[j];
            col[k + 1] = hn;
            if (hn > 1e-12) {  // else: invariant subspace reached
                gm_scale_kernel<<<gb, 256, 0, s>>>(n, 1.0 / hn, w);
                SIE_HIP(hipGetLastError());
            }
            // previous rotations, then a new one that annihilates H(k+1, k)
            for (int i = 0; i < k; ++i) {
                const zc a = col[i], c1 = col[i + 1];
                col[i] = cs[i] * a + sn[i] * c1;
                col[i + 1] = -std::conj(sn[i]) * a + cs[i] * c1;
            }
            givens(col[k], col[k + 1], cs[k], sn[k]);
            col[k] = cs[k] * col[k] + sn[k] * col[k + 1];
            col[k + 1] = 0.0;
            const zc gk = g[k];
            g[k] = cs[k] * gk;
            g[k + 1] = -std::conj(sn[k]) * gk;
            cur = std::abs(g[k + 1]);
            push_history(cur / bn);
            if (cur <= res_tol) {
                ++k;
                break;
            }
        }
        // x += V(:, 0:k) y with H(0:k, 0:k) y = g(0:k)
        const int meff = k;
        for (int i = meff - 1; i >= 0; --i) {
            zc acc = g[i];
            for (int j = i + 1; j < meff; ++j) acc -= H[(size_t)j * ldH + i] * y[j];
            y[i] = acc / H[(size_t)i * ldH + i];
        }
        if (meff > 0) {
            SIE_HIP(hipMemcpyAsync(h.p, y.data(), sizeof(zc) * meff, hipMemcpyHostToDevice, s));
            check(sie_zgemv(0, n, meff, 1.0, V.p, n, h.p, 1.0, x, s));
        }
        SIE_HIP(hipStreamSynchronize(s));
        if (cur <= res_tol) {
            info = 0;
            break;
        }
        info = 1;
    }
    if (out_res) *out_res = cur;
    if (out_iters) *out_iters = total;
    if (history_len) *history_len = std::min(hist_n, history ? history_cap : 0);
    return info;
}

// GMRES(m) refinement of op(Z) x = b (Z complex double) right-preconditioned
// by an FP32 LU of Z (GMRES-IR, Carson & Higham, SIAM J. Sci. Comput. 40
// (2018) A817-A847): z_k = LU32^{-1} v_k is applied with FP64 arithmetic to the
// FP32 factors, so the preconditioner is one fixed linear operator and the
// Krylov space removes the few slow error components of an FP32
// factorisation in a handful of steps instead of one refinement sweep each.
// w = op(Z) z_k is orthogonalised in FP64 (CGS2) and the z_k are kept
// (flexible form), so x += [z_0 .. z_{k-1}] y satisfies the FP64 Arnoldi
// relation. Every cycle ends with the true residual b - op(Z) x. maxit bounds
// the total number of inner steps (one preconditioner solve and one FP64 GEMV each).
// Stopping test: |r| <= tol (znorm |x| + |b|), the normwise backward error
// (Rigal & Gaches) with znorm a norm of Z (its Frobenius norm: the minimal
// perturbation is of rank one, so its 2- and Frobenius norms agree); znorm = 0
// gives the relative-residual test |r| <= tol |b|. Inside a cycle |x| is
// estimated by beta |z_0| (the preconditioned first direction) in the first
// cycle and by the last true |x| afterwards; each cycle's end tests the true
// residual and |x|. A restart cycle that does not halve the true residual
// ends the solve: the residual has reached the floor of its FP64 evaluation
// (the Arnoldi estimate inside a cycle keeps falling below it, so further
// cycles would only run to maxit); the caller compares the backward error
// reached with its own acceptance bound.
int fgmres_ir_run(int op, int n, const double* Z, long long ldz, const float* LU, long long ldl, const int* ipiv,
                  const double* b, double* x, double tol, double znorm, int m, int maxit, double* out_rel,
                  double* out_be, hipStream_t s) {
    require(op == 0 || op == 2, "op must be 0 (N) or 2 (C)");
    require(n > 0 && m > 0 && maxit > 0, "invalid FGMRES-IR dimensions");
    const double bn = sie_norm2(n, b, s);
    if (bn < 0) throw Error(SIE_ERR_DEVICE, sie_last_error());
    if (!(bn > 0) || !std::isfinite(bn)) throw Error(SIE_ERR_NONFINITE, "right-hand side norm must be finite and nonzero");
    const size_t n2 = (size_t)2 * n;
    DevBuf<double> V(n2 * (m + 1)), Zb(n2 * m), r(n2), zt(n2), h((size_t)2 * (m + 1)), h2((size_t)2 * (m + 1));
    const int ldH = m + 1;
    std::vector<zc> H((size_t)ldH * m), sn(m), g(ldH), y(m), hv(m + 1), hv2(m + 1);
    std::vector<double> cs(m);
    const unsigned gb = blocks_for(n);
    SIE_HIP(hipMemsetAsync(x, 0, sizeof(double) * n2, s));
    SIE_HIP(hipMemcpyAsync(r.p, b, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
    double beta = bn, xn = 0.0;
    bool have_xn = false;
    auto bound = [&](double xnorm) { return tol * (znorm * xnorm + bn); };
    int total = 0, cycles = 0;
    double prev_beta = beta;
    while (beta > bound(xn) && total < maxit) {
        gm_copy_scaled_kernel<<<gb, 256, 0, s>>>(n, r.p, beta, V.p);
        SIE_HIP(hipGetLastError());
        std::fill(g.begin(), g.end(), zc(0));
        g[0] = beta;
        int k = 0;
        for (; k < m && total < maxit; ++k) {
            ++total;
            double* vk = V.p + n2 * k;
            double* zk = Zb.p + n2 * k;
            double* w = V.p + n2 * (k + 1);
            // z_k = LU32^{-1} v_k with FP64 arithmetic, through a fixed buffer
            // so that the captured sweep sequence is reused every step
            SIE_HIP(hipMemcpyAsync(zt.p, vk, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
            lu_solve_f32_f64(op, n, LU, ldl, ipiv, zt.p, s);
            SIE_HIP(hipMemcpyAsync(zk, zt.p, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
            if (!have_xn && znorm > 0) {
                const double zn = sie_norm2(n, zk, s);
                if (!std::isfinite(zn) || zn < 0) throw Error(SIE_ERR_NONFINITE, "FGMRES-IR preconditioned vector is not finite");
                xn = beta * zn;
                have_xn = true;
            }
            check(sie_zgemv(op, n, n, 1.0, Z, ldz, zk, 0.0, w, s));
            check(sie_zgemv(2, n, k + 1, 1.0, V.p, n, w, 0.0, h.p, s));
            check(sie_zgemv(0, n, k + 1, -1.0, V.p, n, h.p, 1.0, w, s));
            check(sie_zgemv(2, n, k + 1, 1.0, V.p, n, w, 0.0, h2.p, s));
            check(sie_zgemv(0, n, k + 1, -1.0, V.p, n, h2.p, 1.0, w, s));
            SIE_HIP(hipMemcpyAsync(hv.data(), h.p, sizeof(zc) * (k + 1), hipMemcpyDeviceToHost, s));
            SIE_HIP(hipMemcpyAsync(hv2.data(), h2.p, sizeof(zc) * (k + 1), hipMemcpyDeviceToHost, s));
            const double hn = sie_norm2(n, w, s);  // synchronises the stream
            check_device_waits(s);
            if (!std::isfinite(hn) || hn < 0) throw Error(SIE_ERR_NONFINITE, "FGMRES-IR Arnoldi vector is not finite");
            zc* col = &H[(size_t)k * ldH];
            for (int j = 0; j <= k; ++j) col[j] = hv[j] + hv2[j];
            col[k + 1] = hn;
            if (hn > 0) {
                gm_scale_kernel<<<gb, 256, 0, s>>>(n, 1.0 / hn, w);
                SIE_HIP(hipGetLastError());
            }
            for (int i = 0; i < k; ++i) {
                const zc a = col[i], c1 = col[i + 1];
                col[i] = cs[i] * a + sn[i] * c1;
                col[i + 1] = -std::conj(sn[i]) * a + cs[i] * c1;
            }
            givens(col[k], col[k + 1], cs[k], sn[k]);
            col[k] = cs[k] * col[k] + sn[k] * col[k + 1];
            col[k + 1] = 0.0;
            const zc gk = g[k];
            g[k] = cs[k] * gk;
            g[k + 1] = -std::conj(sn[k]) * gk;
            if (std::abs(g[k + 1]) <= bound(xn) || hn == 0) {
                ++k;
                break;
            }
        }
        // x += Zk y with H(0:k, 0:k) y = g(0:k)
        for (int i = k - 1; i >= 0; --i) {
            zc acc = g[i];
            for (int j = i + 1; j < k; ++j) acc -= H[(size_t)j * ldH + i] * y[j];
            y[i] = acc / H[(size_t)i * ldH + i];
        }
        if (k > 0) {
            SIE_HIP(hipMemcpyAsync(h.p, y.data(), sizeof(zc) * k, hipMemcpyHostToDevice, s));
            check(sie_zgemv(0, n, k, 1.0, Zb.p, n, h.p, 1.0, x, s));
        }
        // true residual r = b - op(Z) x
        SIE_HIP(hipMemcpyAsync(r.p, b, sizeof(double) * n2, hipMemcpyDeviceToDevice, s));
        check(sie_zgemv(op, n, n, -1.0, Z, ldz, x, 1.0, r.p, s));
        beta = sie_norm2(n, r.p, s);
        if (!std::isfinite(beta) || beta < 0) throw Error(SIE_ERR_NONFINITE, "FGMRES-IR diverged (non-finite residual)");
        if (znorm > 0) {
            xn = sie_norm2(n, x, s);
            have_xn = true;
        }
        if (++cycles >= 2 && beta > 0.5 * prev_beta) break;  // stagnation at the FP64 residual floor
        prev_beta = beta;
    }
    if (out_rel) *out_rel = beta / bn;
    if (out_be) *out_be = beta / (znorm * xn + bn);
    return total;
}

namespace {
struct DenseOp {
    const double* Z;
    long long ldz;
    int n;
};
int dense_matvec(void* user, const double* x, double* y, void* stream) {
    const DenseOp* op = (const DenseOp*)user;
    return sie_zgemv(0, op->n, op->n, 1.0, op->Z, op->ldz, x, 0.0, y, stream);
}
}  // namespace

}  // namespace sie

using namespace sie;

extern "C" {

int sie_gmres(int n, int m, int maxnit, double tol, sie_matvec_fn matvec, void* user, const double* b_dev,
              double* x_dev, double* out_res, int* out_iters, int* info, double* history, int history_cap,
              int* history_len, void* stream) {
    try {
        const int r = gmres_run(n, m, maxnit, tol, matvec, user, b_dev, x_dev, out_res, out_iters, history,
                                history_cap, history_len, (hipStream_t)stream);
        if (info) *info = r;
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

int sie_fgmres_ir_solve(int op, int n, int nrhs, const double* Z, long long ldz, const float* LU, long long ldl,
                        const int* ipiv, const double* B, long long ldb, double* X, long long ldx, double tol,
                        double znorm, int maxit, int restart, double* rel_residual, double* backward_error,
                        int* iterations, void* stream) {
    try {
        require(n > 0 && nrhs > 0 && ldz >= n && ldl >= n && ldb >= n && ldx >= n, "invalid FGMRES-IR dimensions");
        require(Z && LU && ipiv && B && X, "null FGMRES-IR argument");
        require(tol > 0 && std::isfinite(tol), "FGMRES-IR tolerance must be positive");
        require(znorm >= 0 && std::isfinite(znorm), "FGMRES-IR matrix norm must be finite and non-negative");
        for (int c = 0; c < nrhs; ++c) {
            double rel = 0, be = 0;
            const int it = fgmres_ir_run(op, n, Z, ldz, LU, ldl, ipiv, B + 2 * (long long)c * ldb,
                                         X + 2 * (long long)c * ldx, tol, znorm, restart, maxit, &rel, &be,
                                         (hipStream_t)stream);
            if (rel_residual) rel_residual[c] = rel;
            if (backward_error) backward_error[c] = be;
            if (iterations) iterations[c] = it;
        }
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

int sie_gmres_dense(int n, const double* Z_dev, long long ldz, int m, int maxnit, double tol, const double* b_dev,
                    double* x_dev, double* out_res, int* out_iters, int* info, double* history, int history_cap,
                    int* history_len, void* stream) {
    if (!Z_dev || ldz < n) {
        set_last_error("GMRES: null matrix or leading dimension below n");
        return SIE_ERR_INVALID;
    }
    DenseOp op{Z_dev, ldz, n};
    return sie_gmres(n, m, maxnit, tol, dense_matvec, &op, b_dev, x_dev, out_res, out_iters, info, history,
                     history_cap, history_len, stream);
}

}  // extern "C"
