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


def test_dielectric_mie_closed_form_checks():
    m = 2.0
    for ka in (0.1, 1.0, 5.0):
        ss, se = mie.cross_sections(ka, m=m)
        assert abs(ss - se) / se < 1e-12  # lossless: optical theorem
    # Rayleigh backscatter 4 pi k^4 a^6 |(m^2 - 1) / (m^2 + 2)|^2
    a = 0.01
    sE, sH = mie.bistatic_rcs(a, np.array([math.pi]), m=m)
    ray = 4 * math.pi * a ** 6 * abs((m * m - 1) / (m * m + 2)) ** 2
    assert abs(sE[0] / ray - 1) < 1e-4 and abs(sE[0] - sH[0]) < 1e-12 * sE[0]
    # no contrast, no scattering
    _, a1, b1 = mie.coefficients(1.0, m=1.0)
    assert np.abs(a1).max() < 1e-15 and np.abs(b1).max() < 1e-15


class _Mesh:
    def __init__(self, level):
        V, F, E, FE = ol.icosphere(level)
        self.vertices, self.faces, self.edges, self.face_edges = V, F, E, FE
        self.nf, self.ne = len(F), len(E)


def _pmchwt_rcs(ka, eps_r, level, rule, th, dirs):
    import sie_mom_amd as sm
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(ka, sm.physics.CODATA2018_EPS0,
                                                              sm.physics.CODATA2018_MU0), quadrature_rule=rule)
    o.interior.relative_permittivity = eps_r
    p, ph = sm.calderon_params(o), sm.physical_parameters(o)
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")
    mesh = _Mesh(level)
    Z = ol.oracle_pmchwt(mesh, p)
    b = ol.oracle_rhs(mesh, 1, rule, w, o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability,
                      o.vacuum_permittivity)
    x = scipy.linalg.solve(Z, b)
    xs = np.zeros(4 * mesh.ne, complex)
    xs[2 * mesh.ne:3 * mesh.ne], xs[3 * mesh.ne:] = x[:mesh.ne], x[mesh.ne:]  # exterior J and M / eta0
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    A = ol.oracle_farfield_matrix(mesh, rule, k0, ph.eta0, dirs)
    return (A @ xs).reshape(-1, 2), k0


def test_oracle_pmchwt_sphere_converges_to_dielectric_mie():
    """The oracle's T-PMCHWT chain (assembly, plane-wave RHS with its signs and
    eta0 on the magnetic rows, solve, far field of J and M) against the Mie
    series of a dielectric sphere, eps_r = 4, k a = 1: second order in the
    mesh size, no cross-polarised field in the principal planes."""
    th, dirs = cut_directions()
    errs = []
    for level in (2, 3):
        F, k0 = _pmchwt_rcs(1.0, 4.0, level, 5, th, dirs)
        n = len(th)
        sE, sH = 4 * math.pi * np.abs(F[:n, 0]) ** 2, 4 * math.pi * np.abs(F[n:, 1]) ** 2
        mE, mH = mie.bistatic_rcs(1.0, th, k=k0, m=2.0)
        cross = max(np.abs(F[:n, 1]).max(), np.abs(F[n:, 0]).max()) / np.abs(F).max()
        assert cross < 1e-6
        errs.append(max(np.abs(sE - mE).max() / mE.max(), np.abs(sH - mH).max() / mH.max()))
    assert errs[0] < 0.09 and errs[1] < 0.025, errs
    assert errs[0] / errs[1] > 3.0, errs


def test_oracle_pec_sphere_ka3_181_angles():
    """PEC sphere at k a = 3, icosphere level 3 (1,920 edges), Dunavant rule 5,
    x-polarised +z plane wave, bistatic RCS at 181 angles of the E-plane: within
    1.5% of the series' peak and 0.1 dB rms (level 2: 5% and 0.3 dB)."""
    import sie_mom_amd as sm
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(3.0), quadrature_rule=5)
    p, ph = sm.calderon_params(o), sm.physical_parameters(o)
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1))
    th = np.radians(np.arange(181.0))
    dirs = np.array([[math.sin(t), 0, math.cos(t), math.cos(t), 0, -math.sin(t), 0, 1, 0] for t in th])
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    mE, _ = mie.bistatic_rcs(3.0, th, k=k0)
    out = []
    for level in (2, 3):
        mesh = _Mesh(level)
        Z = ol.oracle_pec(mesh, p)
        b = ol.oracle_rhs(mesh, 2, 5, w, o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability,
                          o.vacuum_permittivity)
        xs = np.zeros(4 * mesh.ne, complex)
        xs[2 * mesh.ne:3 * mesh.ne] = scipy.linalg.solve(Z, b)
        F = (ol.oracle_farfield_matrix(mesh, 5, k0, ph.eta0, dirs) @ xs).reshape(-1, 2)
        sE = 4 * math.pi * np.abs(F[:, 0]) ** 2
        out.append((np.abs(sE - mE).max() / mE.max(), np.sqrt(np.mean((10 * np.log10(sE / mE)) ** 2))))
    assert out[0][0] < 0.05 and out[0][1] < 0.3, out
    assert out[1][0] < 0.015 and out[1][1] < 0.1, out


def test_fibonacci_plane_waves():
    import sie_mom_amd as sm
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(5.0))
    d = sm.fibonacci_directions(256)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0) and abs(d.mean(axis=0)).max() < 1e-2
    waves = sm.fibonacci_plane_waves(o, 256)
    for i, w in enumerate(waves):
        e, k = np.array(w.E0), np.array(w.khat)
        assert np.allclose(k, d[i]) and abs(e @ k) < 1e-12 and abs(np.linalg.norm(e) - 1) < 1e-12
        # theta polarisation on even waves (no azimuthal part), phi on odd ones (horizontal)
        phi_hat = np.array([-d[i][1], d[i][0], 0.0]) / np.hypot(d[i][0], d[i][1])
        assert (abs(e @ phi_hat) < 1e-12) if i % 2 == 0 else (abs(e[2]) < 1e-12)
