"""Device operators over the C ABI, on torch CUDA (HIP) tensors.

Matrices are returned as column-major views: the storage is written by the
library with leading dimension `dim` (column-major, as the reference's cuSOLVER
path stores them) and exposed as ``buf.T`` so that ``Z[r, c]`` indexes row r,
column c.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .physics import GstcOptions, PecOptions, DielectricOptions, calderon_params

_P = ctypes.c_void_p


def _stream():
    return _P(torch.cuda.current_stream().cuda_stream)


def _dptr(t):
    return _P(t.data_ptr())


def colmajor(dim, cols=None, dtype=torch.complex128, device="cuda"):
    """Uninitialised column-major (dim x cols) device matrix."""
    cols = dim if cols is None else cols
    buf = torch.empty((cols, dim), dtype=dtype, device=device)
    return buf.T


def _check_colmajor(Z, dim):
    if Z.shape[0] != dim or Z.stride(0) != 1 or Z.stride(1) < dim:
        raise ValueError("expected a column-major (dim x n) matrix with leading dimension >= dim")
    return Z.stride(1)


def calderon_pair_blocks(mesh, params):
    """Raw per-pair TL(k1), TL(k2), TK(k1), TK(k2) blocks: (nf, nf, 4, 3, 3)."""
    out = torch.empty((mesh.nf, mesh.nf, 4, 3, 3), dtype=torch.complex128, device="cuda")
    _lib.check(_lib.lib.sie_calderon_pair_blocks(mesh.handle, ctypes.byref(params), _dptr(out), _stream()))
    return out


def _gstc_args(mesh, options, tensors, calderon, mass, scaled):
    nrec = 1 if options.tensors_are_uniform else mesh.nf
    if tensors.numel() != 32 * nrec:
        raise ValueError(f"GSTC tensor count mismatch: expected {32 * nrec}, got {tensors.numel()}")
    t = tensors.to(device="cuda", dtype=torch.complex128).contiguous()
    flags = (1 if calderon else 0) | (2 if mass else 0) | (0 if scaled else 4)
    basis = {"face": 0, "spherical": 1}[options.basis]
    return t, int(options.tensors_are_uniform), basis, flags


def assemble_gstc(mesh, options: GstcOptions, tensors, calderon=True, mass=True, out=None, scaled=True,
                  params=None):
    """eta0-scaled SIE-GSTC matrix (dim 4E). tensors: complex128 device tensor of
    32 values per face (or 32 if options.tensors_are_uniform). scaled=False gives
    the unscaled operator (assemble_Z_calderon_kernel + assemble_Z_mass_kernel,
    sie_kernels_assembly.cuh:24-148, 282-452). params: kernel parameters to use
    instead of those derived from options."""
    dim = 4 * mesh.ne
    p = calderon_params(options) if params is None else params
    t, uni, basis, flags = _gstc_args(mesh, options, tensors, calderon, mass, scaled)
    Z = colmajor(dim) if out is None else out
    ld = _check_colmajor(Z, dim)
    _lib.check(_lib.lib.sie_assemble_gstc(mesh.handle, ctypes.byref(p), _dptr(t), uni, basis, flags,
                                          _dptr(Z), ld, _stream()))
    return Z


def _vec(x, dim):
    if x.shape != (dim,):
        raise ValueError(f"expected a vector of length {dim}, got shape {tuple(x.shape)}")
    return x.to(device="cuda", dtype=torch.complex128).contiguous()


def gstc_matvec(mesh, options: GstcOptions, tensors, x, calderon=True, mass=True, scaled=True, out=None,
                params=None):
    """Matrix-free y = Z x with the SIE-GSTC operator of assemble_gstc (Z is not
    stored). scaled=False applies the reference matvec's unscaled operator
    (matvec_Z_calderon_kernel + matvec_Z_mass_kernel, sie_kernels_matvec.cuh:9-327).
    params: kernel parameters to use instead of those derived from options."""
    dim = 4 * mesh.ne
    p = calderon_params(options) if params is None else params
    t, uni, basis, flags = _gstc_args(mesh, options, tensors, calderon, mass, scaled)
    xv = _vec(x, dim)
    y = torch.empty(dim, dtype=torch.complex128, device="cuda") if out is None else out
    _lib.check(_lib.lib.sie_gstc_matvec(mesh.handle, ctypes.byref(p), _dptr(t), uni, basis, flags,
                                        _dptr(xv), _dptr(y), _stream()))
    return y


def pec_matvec(mesh, options: PecOptions, x, out=None):
    """Matrix-free y = Z x with the T-EFIE operator of assemble_pec."""
    dim = mesh.ne
    p = calderon_params(options)
    xv = _vec(x, dim)
    y = torch.empty(dim, dtype=torch.complex128, device="cuda") if out is None else out
    _lib.check(_lib.lib.sie_pec_matvec(mesh.handle, ctypes.byref(p), _dptr(xv), _dptr(y), _stream()))
    return y


def pmchwt_matvec(mesh, options: DielectricOptions, x, out=None):
    """Matrix-free y = Z x with the T-PMCHWT operator of assemble_pmchwt."""
    dim = 2 * mesh.ne
    p = calderon_params(options)
    xv = _vec(x, dim)
    y = torch.empty(dim, dtype=torch.complex128, device="cuda") if out is None else out
    _lib.check(_lib.lib.sie_pmchwt_matvec(mesh.handle, ctypes.byref(p), _dptr(xv), _dptr(y), _stream()))
    return y


def assemble_pec(mesh, options: PecOptions, out=None):
    dim = mesh.ne
    p = calderon_params(options)
    Z = colmajor(dim) if out is None else out
    ld = _check_colmajor(Z, dim)
    _lib.check(_lib.lib.sie_assemble_pec(mesh.handle, ctypes.byref(p), _dptr(Z), ld, _stream()))
    return Z


def assemble_pmchwt(mesh, options: DielectricOptions, out=None):
    dim = 2 * mesh.ne
    p = calderon_params(options)
    Z = colmajor(dim) if out is None else out
    ld = _check_colmajor(Z, dim)
    _lib.check(_lib.lib.sie_assemble_pmchwt(mesh.handle, ctypes.byref(p), _dptr(Z), ld, _stream()))
    return Z


RHS_KIND = {GstcOptions: 0, DielectricOptions: 1, PecOptions: 2}


def rhs_plane_waves(mesh, options, waves):
    """One RHS column per prepared plane wave (physics.plane_wave)."""
    kind = RHS_KIND[type(options)]
    dim = {0: 4, 1: 2, 2: 1}[kind] * mesh.ne
    arr = (_lib.PlaneWaveC * len(waves))(*waves)
    B = colmajor(dim, len(waves))
    f32 = np.float32
    _lib.check(_lib.lib.sie_rhs_plane_waves(
        mesh.handle, kind, int(options.quadrature_rule), arr, len(waves),
        float(f32(options.vacuum_permeability)), float(f32(options.vacuum_permittivity)),
        float(options.vacuum_permeability), float(options.vacuum_permittivity),
        _dptr(B), B.stride(1), _stream()))
    return B


def surface_fields(mesh, J, M, points, frequency_hz, medium=None, quadrature_rule=7,
                   vacuum_permittivity=None, vacuum_permeability=None):
    """E, H (n x 3 complex128 device tensors) radiated at points (n x 3) by the
    physical per-edge currents J, M: evaluate_surface_integral_fields
    (surface_current_fields.cu:170-299), with its argument checks."""
    from .physics import Medium, CODATA2022_EPS0, CODATA2022_MU0
    medium = medium or Medium(1.0, 1.0)
    eps0 = CODATA2022_EPS0 if vacuum_permittivity is None else vacuum_permittivity
    mu0 = CODATA2022_MU0 if vacuum_permeability is None else vacuum_permeability
    fin = np.isfinite
    if not (frequency_hz > 0 and fin(frequency_hz)):
        raise ValueError("field-evaluation frequency must be positive and finite")
    if not (medium.relative_permittivity > 0 and medium.relative_permeability > 0
            and fin(medium.relative_permittivity) and fin(medium.relative_permeability)):
        raise ValueError("field-evaluation relative permittivity and permeability must be positive and finite")
    if not (eps0 > 0 and mu0 > 0 and fin(eps0) and fin(mu0)):
        raise ValueError("field-evaluation vacuum constants must be positive and finite")
    if not 1 <= quadrature_rule <= 20:
        raise ValueError("Dunavant quadrature rule must be in [1, 20]")
    pts = torch.as_tensor(points, dtype=torch.float64)
    if pts.dim() != 2 or pts.shape[1] != 3:
        raise ValueError("observation points must be an n x 3 array")
    if not torch.isfinite(pts).all():
        bad = int((~torch.isfinite(pts).all(dim=1)).nonzero()[0])
        raise ValueError(f"non-finite observation point at index {bad}")
    n = pts.shape[0]
    if J.numel() != mesh.ne or M.numel() != mesh.ne:
        raise ValueError("field evaluation requires one electric and magnetic coefficient per mesh edge")
    Ed = torch.empty((n, 3), dtype=torch.complex128, device="cuda")
    Hd = torch.empty((n, 3), dtype=torch.complex128, device="cuda")
    if n == 0:
        return Ed, Hd
    Jd = J.to(device="cuda", dtype=torch.complex128).contiguous()
    Md = M.to(device="cuda", dtype=torch.complex128).contiguous()
    if not (torch.isfinite(torch.view_as_real(Jd)).all() and torch.isfinite(torch.view_as_real(Md)).all()):
        raise ValueError("non-finite surface-current coefficient")
    Pd = pts.to("cuda").contiguous()
    omega = 2.0 * np.pi * frequency_hz
    eps = medium.relative_permittivity * eps0
    mu = medium.relative_permeability * mu0
    _lib.check(_lib.lib.sie_surface_fields(mesh.handle, int(quadrature_rule), omega, eps, mu, _dptr(Jd), _dptr(Md),
                                           _dptr(Pd), n, _dptr(Ed), _dptr(Hd), _stream()))
    return Ed, Hd
