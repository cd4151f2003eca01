// DenseLuSolver, solve_dense_lu and MixedLuSolver over the C ABI.
#include "host_common.hpp"
#include "sie_mom/gpu/dense_lu_solver.hpp"

namespace sie_mom {
namespace gpu {

DenseLuSolver::~DenseLuSolver() {
    if (pivots_) (void)hipFree(pivots_);
}

void DenseLuSolver::allocate(int dimension) {
    if (dimension <= 0) throw std::invalid_argument("DenseLuSolver dimension must be positive");
    if (pivots_) {
        (void)hipDeviceSynchronize();
        (void)hipFree(pivots_);
        pivots_ = nullptr;
    }
    host::hip_ok(hipMalloc(&pivots_, sizeof(int) * (size_t)dimension), "allocate LU pivots");
    dimension_ = dimension;
}

void DenseLuSolver::factorize(std::complex<real_t>* matrix, int dimension) {
    if (dimension > 0) allocate(dimension);
    if (dimension_ <= 0) throw std::invalid_argument("DenseLuSolver dimension must be initialized before factorize");
    int info = 0;
    detail::check(sie_lu_factor(64, dimension_, matrix, dimension_, pivots_, &info, 0, stream_));
}

void DenseLuSolver::solve(Op op, int nrhs, const std::complex<real_t>* factored_matrix, std::complex<real_t>* rhs, int ldb) const {
    if (dimension_ <= 0) throw std::invalid_argument("DenseLuSolver dimension must be initialized before solve");
    if (nrhs <= 0) throw std::invalid_argument("DenseLuSolver right-hand-side count must be positive");
    detail::check(sie_lu_solve(64, (int)op, dimension_, nrhs, factored_matrix, dimension_, pivots_, rhs,
                               ldb > 0 ? ldb : dimension_, stream_));
    host::hip_ok(hipStreamSynchronize((hipStream_t)stream_), "LU solve");
}

void solve_dense_lu(int dimension, std::complex<real_t>* matrix, std::complex<real_t>* rhs) {
    DenseLuSolver solver(dimension);
    solver.factorize(matrix);
    solver.solve(matrix, rhs);
}

MixedLuSolver::~MixedLuSolver() {
    if (lu32_) (void)hipFree(lu32_);
    if (pivots_) (void)hipFree(pivots_);
}

void MixedLuSolver::factorize(const std::complex<real_t>* matrix, int dimension, long long ld) {
    if (dimension <= 0) throw std::invalid_argument("MixedLuSolver dimension must be positive");
    if (ld == 0) ld = dimension;
    if (ld < dimension) throw std::invalid_argument("leading dimension below the matrix dimension");
    if (n_ != dimension) {
        (void)hipDeviceSynchronize();
        if (lu32_) (void)hipFree(lu32_);
        if (pivots_) (void)hipFree(pivots_);
        lu32_ = nullptr;
        pivots_ = nullptr;
        host::hip_ok(hipMalloc(&lu32_, sizeof(float) * 2 * (size_t)dimension * dimension), "allocate FP32 factors");
        host::hip_ok(hipMalloc(&pivots_, sizeof(int) * (size_t)dimension), "allocate LU pivots");
        n_ = dimension;
    }
    A_ = matrix;
    ld_ = ld;
    // the FP32 copy and the Frobenius norm of the FP64 operator in one pass
    detail::check(sie_convert(64, 32, dimension, dimension, matrix, ld, lu32_, dimension, 1.0, &znorm_, stream_));
    int info = 0;
    detail::check(sie_lu_factor(32, dimension, lu32_, dimension, pivots_, &info, 2, stream_));
}

std::vector<double> MixedLuSolver::solve(Op op, int nrhs, const std::complex<real_t>* rhs, int ldb, std::complex<real_t>* x, int ldx,
                                         double tol, int maxit, std::vector<int>* sweeps,
                                         std::vector<double>* relative_residuals) const {
    if (n_ <= 0 || !A_) throw std::invalid_argument("MixedLuSolver must be factorized before solve");
    if (op == Op::T) throw std::invalid_argument("MixedLuSolver supports op N and C");
    std::vector<double> rel(nrhs), be(nrhs);
    std::vector<int> its(nrhs);
    if (op == Op::N && nrhs > 4) {
        // many columns: blocked refinement sweeps (one FP64 GEMM residual and one
        // multi-RHS FP32 solve per sweep)
        detail::check(sie_ir_solve((int)op, n_, nrhs, reinterpret_cast<const double*>(A_), ld_,
                                   static_cast<const float*>(lu32_), n_, pivots_, reinterpret_cast<const double*>(rhs),
                                   ldb > 0 ? ldb : n_, reinterpret_cast<double*>(x), ldx > 0 ? ldx : n_, tol, znorm_,
                                   maxit, rel.data(), be.data(), its.data(), stream_));
        if (sweeps) *sweeps = its;
        if (relative_residuals) *relative_residuals = rel;
        return be;
    }
    // GMRES-IR (FP32 factors applied in FP64 as the preconditioner), restart 20
    detail::check(sie_fgmres_ir_solve((int)op, n_, nrhs, reinterpret_cast<const double*>(A_), ld_,
                                      static_cast<const float*>(lu32_), n_, pivots_,
                                      reinterpret_cast<const double*>(rhs), ldb > 0 ? ldb : n_,
                                      reinterpret_cast<double*>(x), ldx > 0 ? ldx : n_, tol, znorm_, maxit, 20,
                                      rel.data(), be.data(), its.data(), stream_));
    if (sweeps) *sweeps = its;
    if (relative_residuals) *relative_residuals = rel;
    return be;
}

}  // namespace gpu
}  // namespace sie_mom
