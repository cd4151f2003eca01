// Operator-level dense solvers over libsie_amd: DenseLuSolver and
// solve_dense_lu with the reference's interface (include/sie_mom/cuda/
// dense_lu_solver.cuh:17-197, namespace sie_mom::cuda there, sie_mom::gpu here),
// plus MixedLuSolver: the FP32-accurate matrix-core LU with FP64 refinement
// (GMRES-IR). Matrices are caller-owned complex128 device arrays, column-major.
#pragma once

#include <vector>

#include "sie_mom/core.hpp"

namespace sie_mom {
namespace gpu {

// cublasOperation_t numbering, which the reference's solve() takes.
enum class Op : int { N = 0, T = 1, C = 2 };

class DenseLuSolver {
public:
    DenseLuSolver() = default;
    explicit DenseLuSolver(int dimension) { allocate(dimension); }
    ~DenseLuSolver();
    DenseLuSolver(const DenseLuSolver&) = delete;
    DenseLuSolver& operator=(const DenseLuSolver&) = delete;

    // std::invalid_argument("DenseLuSolver dimension must be positive") if dimension <= 0.
    void allocate(int dimension);
    int dimension() const { return dimension_; }
    // In-place LU with partial pivoting (zgetrf); dimension 0 keeps the allocated one.
    // std::runtime_error("dense matrix is singular at U diagonal k") on a zero pivot.
    void factorize(std::complex<real_t>* matrix, int dimension = 0);
    // op(A) X = B in place for nrhs columns (ldb 0 = dimension).
    void solve(Op op, int nrhs, const std::complex<real_t>* factored_matrix, std::complex<real_t>* rhs, int ldb = 0) const;
    void solve(const std::complex<real_t>* factored_matrix, std::complex<real_t>* rhs) const { solve(Op::N, 1, factored_matrix, rhs); }
    const int* pivots() const { return pivots_; }
    void set_stream(void* stream) { stream_ = stream; }

private:
    int dimension_ = 0;
    int* pivots_ = nullptr;  // device, 0-based
    void* stream_ = nullptr;
};

// Factorize and solve A x = b in place (solve_dense_lu, dense_lu_solver.cuh:189-197).
void solve_dense_lu(int dimension, std::complex<real_t>* matrix, std::complex<real_t>* rhs);

// A = P L U in complex FP32 computed on the FP16 matrix cores with FP32-level
// accuracy (each FP32 product from three FP16 products of a hi/lo split), then
// op(A) X = B solved by FP64 iterative refinement against the complex128
// matrix, which must stay unchanged and resident while the solver is used,
// until the normwise backward error |b - op(A) x| / (|A|_F |x| + |b|) reaches
// tol. Owns the FP32 factors (8 dimension^2 bytes).
class MixedLuSolver {
public:
    MixedLuSolver() = default;
    ~MixedLuSolver();
    MixedLuSolver(const MixedLuSolver&) = delete;
    MixedLuSolver& operator=(const MixedLuSolver&) = delete;

    void factorize(const std::complex<real_t>* matrix, int dimension, long long ld = 0);
    // GMRES-IR: GMRES(20) right-preconditioned by the FP32 factors (applied in
    // FP64 arithmetic) on the FP64 matrix, per column; for more than four
    // columns with op N, blocked refinement sweeps instead. maxit bounds the
    // steps (sweeps) per column. Returns the per-column normwise backward
    // errors reached (<= tol unless maxit ran out); relative_residuals
    // receives |r| / |b| per column.
    std::vector<double> solve(Op op, int nrhs, const std::complex<real_t>* rhs, int ldb, std::complex<real_t>* x, int ldx,
                              double tol = 1e-15, int maxit = 30, std::vector<int>* sweeps = nullptr,
                              std::vector<double>* relative_residuals = nullptr) const;
    int dimension() const { return n_; }
    // Frobenius norm of the FP64 matrix of the last factorisation.
    double matrix_norm() const { return znorm_; }
    void set_stream(void* stream) { stream_ = stream; }

private:
    int n_ = 0;
    long long ld_ = 0;
    const std::complex<real_t>* A_ = nullptr;
    void* lu32_ = nullptr;
    int* pivots_ = nullptr;
    void* stream_ = nullptr;
    double znorm_ = 0.0;
};

}  // namespace gpu
}  // namespace sie_mom
