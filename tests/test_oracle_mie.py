"""The Mie series (closed form) and, against it, the oracle's whole PEC T-EFIE
chain (assembly, plane-wave RHS, dense solve, far field): bistatic RCS of a
unit sphere at k a = 1 converges to the series at second order in the mesh
size. This pins the oracle's EFIE, excitation and far-field operators
independently of the reference."""
import math

import numpy as np
import pytest
import scipy.linalg

import mie
import oracle_lib as ol


def test_mie_closed_form_checks():
    # optical theorem: a lossless sphere scatters what it extinguishes
    for ka in (0.05, 1.0, 3.0, 8.0):
        ss, se = mie.cross_sections(ka)
        assert abs(ss - se) / se < 1e-12
    # Rayleigh limit of the backscatter: 9 pi a^2 (k a)^4
    a = 0.01
    sE, sH = mie.bistatic_rcs(a, np.array([math.pi]))
    assert abs(sE[0] / (9 * math.pi * a ** 2 * a ** 4) - 1) < 1e-3
    assert abs(sE[0] - sH[0]) / sE[0] < 1e-12  # backscatter does not depend on the plane
    # forward scattering amplitude: S1(0) = S2(0)
    S1, S2 = mie.amplitudes(2.0, np.array([0.0]))
    assert abs(S1[0] - S2[0]) < 1e-12 * abs(S1[0])
    # optics limit: the backscatter tends to the geometric pi a^2
    assert abs(mie.bistatic_rcs(60.0, np.array([math.pi]))[0][0] / (math.pi * 60.0 ** 2) - 1) < 0.1


def cut_directions(n=19):
    th = np.linspace(0.01, math.pi - 0.01, n)
    e_plane = [[math.sin(t), 0, math.cos(t), math.cos(t), 0, -math.sin(t), 0, 1, 0] for t in th]
    h_plane = [[0, math.sin(t), math.cos(t), 0, math.cos(t), -math.sin(t), -1, 0, 0] for t in th]
    return th, np.array(e_plane + h_plane, float)


def rcs_errors(F, th):
    """Largest deviation of the E- and H-plane RCS from the series, relative to its peak."""
    n = len(th)
    sE = 4 * math.pi * np.abs(F[:n, 0]) ** 2
    sH = 4 * math.pi * np.abs(F[n:, 1]) ** 2
    mE, mH = mie.bistatic_rcs(1.0, th)
    cross = max(np.abs(F[:n, 1]).max(), np.abs(F[n:, 0]).max()) / np.abs(F).max()
    return np.abs(sE - mE).max() / mE.max(), np.abs(sH - mH).max() / mH.max(), cross


def test_oracle_pec_sphere_converges_to_mie():
    import sie_mom_amd as sm
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(1.0), quadrature_rule=5)
    p = sm.calderon_params(o)
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1))
    ph = sm.physical_parameters(o)
    th, dirs = cut_directions()
    errs = []
    for level in (2, 3):
        V, F, E, FE = ol.icosphere(level)

        class Mm:
            pass
        mesh = Mm()
        mesh.vertices, mesh.faces, mesh.edges, mesh.face_edges = V, F, E, FE
        mesh.nf, mesh.ne = len(F), len(E)
        Z = ol.oracle_pec(mesh, p)
        b = ol.oracle_rhs(mesh, 2, o.quadrature_rule, w, o.vacuum_permeability, o.vacuum_permittivity,
                          o.vacuum_permeability, o.vacuum_permittivity)
        J = scipy.linalg.solve(Z, b)
        xs = np.zeros(4 * mesh.ne, complex)
        xs[2 * mesh.ne:3 * mesh.ne] = J  # exterior electric current
        A = ol.oracle_farfield_matrix(mesh, o.quadrature_rule, 1.0, ph.eta0, dirs)
        eE, eH, cross = rcs_errors((A @ xs).reshape(-1, 2), th)
        assert cross < 1e-6  # no cross-polarised field in the principal planes
        errs.append(max(eE, eH))
    assert errs[0] < 0.04 and errs[1] < 0.01, errs
    assert errs[0] / errs[1] > 3.0, errs  # second order in the mesh size
