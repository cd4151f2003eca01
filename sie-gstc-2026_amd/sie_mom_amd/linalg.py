"""Dense complex linear algebra on the GPU: LU (FP64 or FP32 matrix cores),
triangular solves with op N / C, GEMV and FP64 iterative refinement.

Mirrors the reference's operator-level DenseLuSolver
(include/sie_mom/cuda/dense_lu_solver.cuh:17-197): factorize(A) then
solve(op, nrhs, B); op "N" solves A X = B, op "C" solves A^H X = B.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .ops import _dptr, _stream, colmajor

OPS = {"N": 0, "T": 1, "C": 2}
# FP32 GEMM variants: 4M real decomposition, 3M, FP32-accurate split-FP16 (3 FP16 MFMAs per real product)
VARIANTS = {"4m": 0, "3m": 1, "f16x3": 2}


def _prec(t):
    if t.dtype == torch.complex128:
        return 64
    if t.dtype == torch.complex64:
        return 32
    raise TypeError("complex64 or complex128 device tensors expected")


def _ld(A):
    if A.stride(0) != 1:
        raise ValueError("column-major storage (stride(0) == 1) required")
    return A.stride(1) if A.dim() == 2 and A.shape[1] > 1 else max(A.shape[0], 1)


def gemm(A, B, C=None, alpha=1.0, beta=0.0, variant="4m"):
    """C = beta C + alpha A B on the matrix cores (column-major operands)."""
    m, k = A.shape
    k2, n = B.shape
    assert k == k2
    if C is None:
        C = colmajor(m, n, dtype=A.dtype)
        beta = 0.0
    _lib.check(_lib.lib.sie_gemm(_prec(A), m, n, k, float(alpha), _dptr(A), _ld(A), _dptr(B), _ld(B),
                                 float(beta), _dptr(C), _ld(C), VARIANTS[variant], _stream()))
    return C


class DenseLuSolver:
    """In-place LU with partial pivoting of a column-major square matrix."""

    def __init__(self, dimension=0, variant="4m"):
        if dimension < 0:
            raise ValueError("DenseLuSolver dimension must be positive")
        if variant not in VARIANTS:
            raise ValueError(f"unknown GEMM variant {variant!r}")
        self.dimension = dimension
        self.variant = variant
        self.pivots = None
        self.factors = None

    def factorize(self, A):
        n = A.shape[0]
        if A.shape[1] != n:
            raise ValueError("square matrix expected")
        self.dimension = n
        self.pivots = torch.empty(n, dtype=torch.int32, device=A.device)
        info = ctypes.c_int(0)
        _lib.check(_lib.lib.sie_lu_factor(_prec(A), n, _dptr(A), _ld(A), _dptr(self.pivots), ctypes.byref(info),
                                          VARIANTS[self.variant], _stream()))
        self.factors = A
        return A

    def solve(self, B, op="N", factored=None):
        """Solve op(A) X = B in place (B: n or n x nrhs column-major)."""
        LU = self.factors if factored is None else factored
        n = self.dimension
        Bm = B if B.dim() == 2 else B.view(n, 1)
        nrhs = Bm.shape[1]
        ldb = Bm.stride(1) if nrhs > 1 else n
        if Bm.stride(0) != 1:
            raise ValueError("column-major right-hand side required")
        _lib.check(_lib.lib.sie_lu_solve(_prec(LU), OPS[op], n, nrhs, _dptr(LU), _ld(LU), _dptr(self.pivots),
                                         _dptr(Bm), ldb, _stream()))
        return B


def zgemv(A, x, y=None, alpha=1.0, beta=0.0, op="N"):
    m, n = A.shape
    out = (n if op == "C" else m)
    if y is None:
        y = torch.zeros(out, dtype=torch.complex128, device=A.device)
    _lib.check(_lib.lib.sie_zgemv(OPS[op], m, n, float(alpha), _dptr(A), _ld(A), _dptr(x), float(beta),
                                  _dptr(y), _stream()))
    return y


def to_precision(A, precision, fro=False):
    """Column-major copy of A in complex64 (32) or complex128 (64). fro=True
    (complex128 A) also returns the Frobenius norm of A, formed in the same
    pass: (B, norm)."""
    rows, cols = A.shape
    dtype = torch.complex64 if precision == 32 else torch.complex128
    B = colmajor(rows, cols, dtype=dtype)
    nrm = ctypes.c_double(0.0)
    _lib.check(_lib.lib.sie_convert(_prec(A), precision, rows, cols, _dptr(A), _ld(A), _dptr(B), _ld(B), 1.0,
                                    ctypes.byref(nrm) if fro else None, _stream()))
    return (B, nrm.value) if fro else B


IR_METHODS = ("ir", "fgmres")


def ir_solve(Z, LU32, pivots, B, op="N", tol=1e-12, maxit=30, method="ir", restart=20, znorm=0.0):
    """FP64-accurate solve with an FP32 LU: classic iterative refinement
    (method "ir") or flexible GMRES(restart) preconditioned by the FP32 LU
    (method "fgmres"). Stops when |r| <= tol (znorm |x| + |b|): the normwise
    backward error with znorm a norm of Z (its Frobenius norm from
    to_precision(..., fro=True)); znorm = 0 gives |r| <= tol |b|.
    Returns (X, rel_residuals, iterations, backward_errors)."""
    if method not in IR_METHODS:
        raise ValueError(f"unknown refinement method {method!r}")
    n = Z.shape[0]
    Bm = B if B.dim() == 2 else B.view(n, 1)
    nrhs = Bm.shape[1]
    X = colmajor(n, nrhs)
    rr = np.zeros(nrhs)
    be = np.zeros(nrhs)
    its = np.zeros(nrhs, np.int32)
    head = (OPS[op], n, nrhs, _dptr(Z), _ld(Z), _dptr(LU32), _ld(LU32), _dptr(pivots),
            _dptr(Bm), Bm.stride(1) if nrhs > 1 else n, _dptr(X), X.stride(1) if nrhs > 1 else n, float(tol),
            float(znorm), int(maxit))
    tail = (rr.ctypes.data_as(ctypes.c_void_p), be.ctypes.data_as(ctypes.c_void_p),
            its.ctypes.data_as(ctypes.c_void_p), _stream())
    if method == "ir":
        _lib.check(_lib.lib.sie_ir_solve(*head, *tail))
    else:
        _lib.check(_lib.lib.sie_fgmres_ir_solve(*head, int(restart), *tail))
    return (X if B.dim() == 2 else X[:, 0]), rr, its, be


def set_wait_bound(seconds, force_timeout=False):
    """Bound of the library's device-side waits (LU panel exchanges, sweep
    blocks); force_timeout makes the panel exchanges give up (testing)."""
    _lib.check(_lib.lib.sie_set_wait_bound(float(seconds), 1 if force_timeout else 0))


def mfma_f16_rate(iters=4000, reps=3):
    """(TFLOP/s, MHz): the sustained FP16 matrix-core rate of the current
    device under the split-FP16 GEMM's instruction mix and the shader clock
    held meanwhile (register operands, no memory traffic)."""
    tf, mhz = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib.sie_mfma_f16_rate(int(iters), int(reps), ctypes.byref(tf), ctypes.byref(mhz),
                                          _stream()))
    return tf.value, mhz.value


def set_panel_launch(cooperative):
    """Launch form of the LU panel kernels: plain dispatches (default) or
    cooperative launches (the runtime guarantees that every workgroup of a
    panel is resident; ~21 us more per panel)."""
    _lib.check(_lib.lib.sie_set_panel_launch(1 if cooperative else 0))


class _DeviceVector:
    """Zero-copy view of n complex doubles at a device address (for torch.as_tensor)."""

    def __init__(self, ptr, n):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<c16", "data": (ptr, False), "version": 3,
                                         "strides": None}


class GmresResult:
    def __init__(self, x, info, residual, iterations, history):
        self.x, self.info, self.residual, self.iterations, self.history = x, info, residual, iterations, history

    @property
    def converged(self):
        return self.info == 0


def _gmres_call(n, rhs, x, m, maxnit, tol, runner):
    hist = np.zeros(m * maxnit, dtype=np.float64)
    res, iters, info, hlen = ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    runner(res, iters, info, hist, hlen)
    return GmresResult(x, info.value, res.value, iters.value, hist[:hlen.value].copy())


def gmres(Z, b, x0=None, m=50, maxnit=20, tol=1e-8):
    """Restarted GMRES(m) on a stored complex128 column-major matrix Z (the
    stored-Z matvec of gmres_restarted, gmres_solver.cuh:258-520)."""
    n = Z.shape[0]
    if Z.dtype != torch.complex128 or b.dtype != torch.complex128:
        raise TypeError("complex128 operator and right-hand side expected")
    bb = b.contiguous()
    x = torch.zeros_like(bb) if x0 is None else x0.clone().contiguous()

    def run(res, iters, info, hist, hlen):
        _lib.check(_lib.lib.sie_gmres_dense(n, _dptr(Z), _ld(Z), m, maxnit, float(tol), _dptr(bb), _dptr(x),
                                            ctypes.byref(res), ctypes.byref(iters), ctypes.byref(info),
                                            hist.ctypes.data_as(ctypes.c_void_p), hist.size, ctypes.byref(hlen),
                                            _stream()))
    return _gmres_call(n, bb, x, m, maxnit, tol, run)


class GMRESSolver:
    """GMRESSolver<MatVecOp> (gmres_solver.cuh:524-537): matvec(x, y) writes
    y = A x for complex128 device vectors of length neqns."""

    def __init__(self, neqns, matvec):
        if neqns <= 0:
            raise ValueError("GMRES dimension must be positive")
        self.neqns, self.matvec = neqns, matvec

    def solve(self, rhs, x, m=50, maxnit=20, tol=1e-8):
        n = self.neqns
        errors = []

        def cb(user, xp, yp, stream):
            try:
                xv = torch.as_tensor(_DeviceVector(xp, n), device="cuda")
                yv = torch.as_tensor(_DeviceVector(yp, n), device="cuda")
                self.matvec(xv, yv)
                return _lib.SIE_OK
            except Exception as e:  # reported after the C call returns
                errors.append(e)
                return _lib.SIE_ERR_INVALID
        fn = _lib.MATVEC_FN(cb)
        bb = rhs.contiguous()

        def run(res, iters, info, hist, hlen):
            code = _lib.lib.sie_gmres(n, m, maxnit, float(tol), fn, None, _dptr(bb), _dptr(x), ctypes.byref(res),
                                      ctypes.byref(iters), ctypes.byref(info), hist.ctypes.data_as(ctypes.c_void_p),
                                      hist.size, ctypes.byref(hlen), _stream())
            if errors:
                raise errors[0]
            _lib.check(code)
        return _gmres_call(n, bb, x, m, maxnit, tol, run)
