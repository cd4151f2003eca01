// GMRESSolver<MatVecOp> with the reference's interface (include/sie_mom/cuda/
// gmres/gmres_solver.cuh:524-537) over sie_gmres. MatVecOp is callable as
// op(const std::complex<real_t>* x_dev, std::complex<real_t>* y_dev) and writes y = A x on the device
// (enqueued on the solver's stream). Semantics of solve(): as sie_gmres in
// sie_capi.h (initial guess in d_solution, tol relative to |b|, returns 0 when
// converged and 1 otherwise; out_res is the absolute residual estimate).
#pragma once

#include <exception>
#include <vector>

#include "sie_mom/core.hpp"

namespace sie_mom {
namespace gpu {

template <typename MatVecOp>
class GMRESSolver {
public:
    GMRESSolver(int neqns, MatVecOp matvec, void* stream = nullptr)
        : neqns_(neqns), matvec_(matvec), stream_(stream) {}

    int solve(const std::complex<real_t>* d_rhs, std::complex<real_t>* d_solution, int m, int maxnit, real_t tol,
              real_t* out_res = nullptr, int* out_iters = nullptr,
              std::vector<real_t>* out_res_history = nullptr) {
        Ctx ctx{&matvec_, nullptr};
        std::vector<double> hist(out_res_history ? (size_t)m * (size_t)maxnit : 0);
        double res = 0;
        int iters = 0, info = 0, hlen = 0;
        const int code = sie_gmres(neqns_, m, maxnit, tol, &GMRESSolver::trampoline, &ctx,
                                   reinterpret_cast<const double*>(d_rhs), reinterpret_cast<double*>(d_solution),
                                   &res, &iters, &info, hist.empty() ? nullptr : hist.data(), (int)hist.size(),
                                   &hlen, stream_);
        if (ctx.error) std::rethrow_exception(ctx.error);
        detail::check(code);
        if (out_res) *out_res = res;
        if (out_iters) *out_iters = iters;
        if (out_res_history) out_res_history->assign(hist.begin(), hist.begin() + hlen);
        return info;
    }

private:
    struct Ctx {
        MatVecOp* op;
        std::exception_ptr error;
    };
    static int trampoline(void* user, const double* x, double* y, void*) {
        Ctx* c = static_cast<Ctx*>(user);
        try {
            (*c->op)(reinterpret_cast<const std::complex<real_t>*>(x), reinterpret_cast<std::complex<real_t>*>(y));
            return SIE_OK;
        } catch (...) {
            c->error = std::current_exception();
            return SIE_ERR_DEVICE;
        }
    }
    int neqns_;
    MatVecOp matvec_;
    void* stream_;
};

// y = Z x for a stored complex128 column-major matrix (the stored-Z matvec).
struct DenseMatVec {
    const std::complex<real_t>* Z;
    long long ldz;
    int n;
    void* stream = nullptr;
    void operator()(const std::complex<real_t>* x, std::complex<real_t>* y) const {
        detail::check(sie_zgemv(0, n, n, 1.0, reinterpret_cast<const double*>(Z), ldz,
                                reinterpret_cast<const double*>(x), 0.0, reinterpret_cast<double*>(y), stream));
    }
};

}  // namespace gpu
}  // namespace sie_mom
