"""GSTC synthesis objective and its hand-written adjoint gradient.

Objective: far-field power radiated into a cone about +z,
    J(theta) = sum_d w_d (|F_theta,d|^2 + |F_phi,d|^2) / (2 eta0),
with directions on a Gauss-Legendre (cos theta) x uniform (phi) grid over the
cone, F the Schelkunoff far field of the exterior currents of the GSTC solve
(applications/anisotropic_gstc_sphere/plot_far_field.py:150-219) and theta the
per-face real surface parameters (libsie_amd: sie_gstc_records).

Gradient (all on the GPU): Zs xs = bs (eta0-scaled system, FP64 LU or FP32 LU
+ FP64 iterative refinement), lambda = Zs^-H dJ/d conj(xs) reusing the
device-resident factors, then dJ/dtheta = -2 Re lambda^H (dZs/dtheta) xs from
the per-face mass-block sensitivities.
"""
import ctypes
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .linalg import DenseLuSolver, ir_solve, to_precision
from .ops import _dptr, _stream, assemble_gstc, rhs_plane_waves
from .physics import physical_parameters


def susceptibility_scales(options):
    """(b, x, k) so that B = w eps0 chi_ee, X = w mu0 chi_mm, K = k0 chi_em."""
    ph = physical_parameters(options)
    k0 = ph.omega * math.sqrt(options.vacuum_permeability * options.vacuum_permittivity)
    return ph.omega * options.vacuum_permittivity, ph.omega * options.vacuum_permeability, k0


def records(params, kind, options):
    """(nf, kind) float64 device parameters -> (nf*32,) complex128 records."""
    nf = params.shape[0]
    bs, xs, ks = susceptibility_scales(options)
    eta = math.sqrt(options.vacuum_permeability / options.vacuum_permittivity)
    p = params.to(device="cuda", dtype=torch.float64).contiguous()
    rec = torch.empty(nf * 32, dtype=torch.complex128, device="cuda")
    ns = ctypes.c_int(0)
    _lib.check(_lib.lib.sie_gstc_records(nf, kind, _dptr(p), bs, xs, ks, eta, _dptr(rec), ctypes.byref(ns),
                                         _stream()))
    return rec


def cone_directions(half_angle_deg=10.0, n_theta=8, n_phi=16):
    """(ndir, 9) host array (r_hat, theta_hat, phi_hat) and solid-angle weights."""
    x, w = np.polynomial.legendre.leggauss(n_theta)
    c0 = math.cos(math.radians(half_angle_deg))
    ct = 0.5 * (1 - c0) * x + 0.5 * (1 + c0)
    wt = 0.5 * (1 - c0) * w
    phis = 2 * np.pi * (np.arange(n_phi) + 0.5) / n_phi
    dirs, weights = [], []
    for c, wc in zip(ct, wt):
        s = math.sqrt(max(0.0, 1 - c * c))
        for ph in phis:
            cp, sp = math.cos(ph), math.sin(ph)
            dirs.append([s * cp, s * sp, c, c * cp, c * sp, -s, -sp, cp, 0.0])
            weights.append(wc * 2 * np.pi / n_phi)
    return np.ascontiguousarray(dirs, np.float64), np.asarray(weights)


def farfield(mesh, options, xs, dirs, rule=None):
    ph = physical_parameters(options)
    k0 = ph.omega * math.sqrt(options.vacuum_permeability * options.vacuum_permittivity)
    nd = len(dirs)
    F = torch.empty(nd * 2, dtype=torch.complex128, device="cuda")
    x = xs.contiguous()
    _lib.check(_lib.lib.sie_farfield(mesh.handle, int(rule or options.quadrature_rule), k0, ph.eta0, _dptr(x), nd,
                                     dirs.ctypes.data_as(ctypes.c_void_p), _dptr(F), _stream()))
    return F.view(nd, 2)


def farfield_adjoint(mesh, options, u, dirs, rule=None):
    ph = physical_parameters(options)
    k0 = ph.omega * math.sqrt(options.vacuum_permeability * options.vacuum_permittivity)
    g = torch.empty(4 * mesh.ne, dtype=torch.complex128, device="cuda")
    uu = u.contiguous()
    _lib.check(_lib.lib.sie_farfield_adjoint(mesh.handle, int(rule or options.quadrature_rule), k0, ph.eta0,
                                             len(dirs), dirs.ctypes.data_as(ctypes.c_void_p), _dptr(uu), _dptr(g),
                                             _stream()))
    return g


def mass_sensitivity(mesh, options, lam, xs):
    G = torch.empty(mesh.nf * 32, dtype=torch.complex128, device="cuda")
    f32 = np.float32
    basis = {"face": 0, "spherical": 1}[options.basis]
    _lib.check(_lib.lib.sie_gstc_mass_sensitivity(mesh.handle, int(options.quadrature_rule), basis,
                                                  float(f32(options.vacuum_permeability)),
                                                  float(f32(options.vacuum_permittivity)),
                                                  _dptr(lam.contiguous()), _dptr(xs.contiguous()), _dptr(G),
                                                  _stream()))
    return G


def param_grad(kind, rec, G, options):
    nf = rec.numel() // 32
    bs, xs, ks = susceptibility_scales(options)
    grad = torch.empty((nf, kind), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib.sie_gstc_param_grad(nf, kind, _dptr(rec), _dptr(G), bs, xs, ks, _dptr(grad), _stream()))
    return grad


@dataclass
class SolveState:
    xs: torch.Tensor
    lu: DenseLuSolver
    Z: torch.Tensor = None      # FP64 operator (mixed mode)
    rec: torch.Tensor = None
    znorm: float = 0.0          # Frobenius norm of Z (mixed mode)
    ir_residual: float = 0.0
    ir_iterations: int = 0
    ir_backward_error: float = 0.0
    adj_residual: float = 0.0
    adj_iterations: int = 0
    adj_backward_error: float = 0.0


class RefinementNotConverged(RuntimeError):
    """The FP64 refinement around the FP32 factors ended above its tolerance."""


class GstcProblem:
    """Mesh, options, excitation and cone objective of one GSTC synthesis step.

    mode "fp64": FP64 LU and solves. mode "mixed": FP32-accurate LU on the
    matrix cores and FP64 refinement of both solves to ir_tol, measured as the
    normwise backward error |r| / (|Z|_F |x| + |b|) (ir_stop="backward") or as
    |r| / |b| (ir_stop="residual"); a solve that ends above ir_accept (default
    ir_tol) raises RefinementNotConverged, as the C++ drivers do. An ir_tol
    below the attainable FP64 floor with a looser ir_accept refines to that
    floor (FGMRES-IR stops at the first restart cycle that does not halve the
    residual) and still accepts the solve."""

    def __init__(self, mesh, options, waves, kind, half_angle_deg=10.0, n_theta=8, n_phi=16, mode="mixed",
                 ir_tol=1e-15, ir_maxit=30, variant="4m", ir_method="fgmres", ir_stop="backward", ir_accept=None):
        if ir_stop not in ("backward", "residual"):
            raise ValueError(f"unknown refinement stop {ir_stop!r}")
        self.mesh, self.options, self.waves, self.kind = mesh, options, waves, kind
        self.dirs, self.weights = cone_directions(half_angle_deg, n_theta, n_phi)
        self.w_dev = torch.from_numpy(self.weights).to("cuda", torch.float64)
        self.mode = mode
        self.ir_tol, self.ir_maxit, self.ir_method, self.ir_stop = ir_tol, ir_maxit, ir_method, ir_stop
        self.ir_accept = ir_tol if ir_accept is None else max(ir_accept, ir_tol)
        self.variant = variant
        self.eta0 = physical_parameters(options).eta0
        self.timings = {}

    def _refine(self, st, b, op):
        znorm = st.znorm if self.ir_stop == "backward" else 0.0
        x, rr, its, be = ir_solve(st.Z, st.lu.factors, st.lu.pivots, b, op=op, tol=self.ir_tol, maxit=self.ir_maxit,
                                  method=self.ir_method, znorm=znorm)
        measure = float(be[0]) if self.ir_stop == "backward" else float(rr[0])
        if not measure <= self.ir_accept:
            what = "backward error" if self.ir_stop == "backward" else "relative residual"
            raise RefinementNotConverged(f"op {op} refinement ended at {what} {measure:.3e} > {self.ir_accept:.1e} "
                                         f"after {int(its[0])} steps")
        return x, float(rr[0]), int(its[0]), float(be[0])

    def _sync_time(self, key, t0):
        torch.cuda.synchronize()
        import time
        self.timings[key] = self.timings.get(key, 0.0) + (time.perf_counter() - t0)

    def solve(self, params):
        import time
        t0 = time.perf_counter()
        rec = records(params, self.kind, self.options)
        Z = assemble_gstc(self.mesh, self.options, rec)
        b = rhs_plane_waves(self.mesh, self.options, self.waves)[:, 0].contiguous()
        self._sync_time("assemble", t0)
        t0 = time.perf_counter()
        if self.mode == "fp64":
            lu = DenseLuSolver()
            lu.factorize(Z)
            self._sync_time("factor", t0)
            t0 = time.perf_counter()
            xs = b.clone()
            lu.solve(xs, op="N")
            st = SolveState(xs=xs, lu=lu, rec=rec)
        else:
            LU32, znorm = to_precision(Z, 32, fro=True)
            lu = DenseLuSolver(variant=self.variant)
            lu.factorize(LU32)
            self._sync_time("factor", t0)
            t0 = time.perf_counter()
            st = SolveState(xs=None, lu=lu, Z=Z, rec=rec, znorm=znorm)
            st.xs, st.ir_residual, st.ir_iterations, st.ir_backward_error = self._refine(st, b, "N")
        self._sync_time("solve", t0)
        return st

    def objective(self, st):
        F = farfield(self.mesh, self.options, st.xs, self.dirs)
        J = ((F.abs() ** 2).sum(dim=1) * self.w_dev).sum() / (2 * self.eta0)
        return J, F

    def gradient(self, st, F):
        import time
        t0 = time.perf_counter()
        u = (F * (self.w_dev / (2 * self.eta0))[:, None]).contiguous()
        g = farfield_adjoint(self.mesh, self.options, u.view(-1), self.dirs)
        if self.mode == "fp64":
            lam = g.clone()
            st.lu.solve(lam, op="C")
        else:
            lam, st.adj_residual, st.adj_iterations, st.adj_backward_error = self._refine(st, g, "C")
        self._sync_time("adjoint_solve", t0)
        t0 = time.perf_counter()
        G = mass_sensitivity(self.mesh, self.options, lam, st.xs)
        grad = param_grad(self.kind, st.rec, G, self.options)
        self._sync_time("contraction", t0)
        return grad


class GstcConeObjective(torch.autograd.Function):
    """torch.autograd.Function: params (nf, kind) float64 -> J (scalar).
    backward uses the device-resident factors of forward's solve."""

    @staticmethod
    def forward(ctx, params, problem):
        st = problem.solve(params)
        J, F = problem.objective(st)
        ctx.problem, ctx.state, ctx.F = problem, st, F
        return J.detach()

    @staticmethod
    def backward(ctx, gout):
        grad = ctx.problem.gradient(ctx.state, ctx.F)
        return grad * gout, None
