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
