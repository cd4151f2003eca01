#!/usr/bin/env python3
"""Writes tests/golden/records_vectors.npz from the reference's own Python
(src/sie_mom/gstc in the reference tree, imported here, in the container that
holds it; the fixture then pins the oracle and the GPU records everywhere).

What no closed form fixes, and the reference does:
  * the order of the 10 real parameters of the general reciprocal lossless
    map, surface_parameters._reciprocal_basis("general", lossless=True)
    (src/sie_mom/gstc/surface_parameters.py:110-127): Lambda = sum c_i B_i;
  * the record layout and the impedance-balanced inversion of
    tensor_io.prepare_tensor_records (src/sie_mom/gstc/tensor_io.py:17-130);
  * the omega model J = jB E - chi n x H, M = -chi n x E + jX H in the
    oriented tangent basis (omega_biisotropic.reconstruct_omega_currents,
    omega_biisotropic.py:196-219).

Inputs are seeded; coefficients c_i are the physical entries (B, X in
siemens / ohms, K dimensionless) for parameters p ~ U[-2, 2]/k0 scaled by
(w eps0, w mu0, k0) as the solver's records do.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("SIE_REFERENCE", "/root/reference")


def main():
    sys.path.insert(0, os.path.join(REF, "src"))
    from sie_mom.gstc.omega_biisotropic import BoundaryQuantities, OmegaParameters, reconstruct_omega_currents
    from sie_mom.gstc.surface_parameters import SurfaceParameters, _reciprocal_basis
    from sie_mom.gstc.tensor_io import prepare_tensor_records

    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    k0 = 2 * np.pi
    omega = k0 / np.sqrt(mu0 * eps0)
    scales = np.array([omega * eps0, omega * mu0, k0])
    eta = np.sqrt(mu0 / eps0)
    rng = np.random.default_rng(20260101)
    nf = 24
    out = {"scales": scales, "eta": np.array(eta)}
    # general (10 parameters)
    p10 = rng.uniform(-2, 2, size=(nf, 10)) / k0
    per = np.array([scales[0]] * 3 + [scales[1]] * 3 + [scales[2]] * 4)
    basis = _reciprocal_basis("general", lossless=True)
    lam = np.einsum("fi,irc->frc", p10 * per[None, :], basis)
    blocks = {"Y_ee": lam[:, :2, :2], "Y_em": lam[:, :2, 2:], "Y_me": lam[:, 2:, :2], "Y_mm": lam[:, 2:, 2:]}
    sp = SurfaceParameters(**{k: np.moveaxis(v, 0, -1) for k, v in blocks.items()})
    rec10, _ = prepare_tensor_records(sp, nf, reference_impedance=eta, extraction_method="reciprocal_lossless",
                                      tensor_structure="general")
    out.update(params10=p10, records10=rec10.ravel())
    # omega (3 parameters: chi_ee, chi_mm, chi_em): the map is read off the
    # reference's omega currents for unit fields E_u, E_v, H_u, H_v
    p3 = rng.uniform(-2, 2, size=(nf, 3)) / k0
    B, X, chi = p3[:, 0] * scales[0], p3[:, 1] * scales[1], p3[:, 2] * scales[2]
    lam3 = np.zeros((nf, 4, 4), complex)
    for col in range(4):
        f = np.zeros((4, nf), complex)
        f[col] = 1.0
        q = BoundaryQuantities(electric_average=f[:2], magnetic_average=f[2:], electric_current=np.zeros((2, nf)),
                               magnetic_current=np.zeros((2, nf)))
        J, M = reconstruct_omega_currents(q, OmegaParameters(B=B, X=X, chi=chi))
        lam3[:, :2, col] = J.T
        lam3[:, 2:, col] = M.T
    blocks = {"Y_ee": lam3[:, :2, :2], "Y_em": lam3[:, :2, 2:], "Y_me": lam3[:, 2:, :2], "Y_mm": lam3[:, 2:, 2:]}
    sp3 = SurfaceParameters(**{k: np.moveaxis(v, 0, -1) for k, v in blocks.items()})
    rec3, _ = prepare_tensor_records(sp3, nf, reference_impedance=eta, extraction_method="reciprocal_lossless",
                                     tensor_structure="general")
    out.update(params3=p3, records3=rec3.ravel())
    np.savez_compressed(os.path.join(HERE, "records_vectors.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
