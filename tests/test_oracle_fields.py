"""Closed-form pins of the oracle's surface-current field evaluator
(evaluate_surface_integral_fields, surface_current_fields.cu:170-299),
independent of the reference: for the T-EFIE currents of a PEC sphere under a
plane wave, (1) the extinction theorem — inside the closed surface the
evaluator returns the incident field (the sign is the evaluator's: E = jw mu
L(J) + K(M)), for E and H, far from and close to the surface (where the
regularised quadrature plus the Wilton singular part is used); (2) the far
zone — r e^{jkr} E tends to minus the Schelkunoff far field of the same
currents (plot_far_field.py:150-219), to O(1/kr)."""
import numpy as np
import pytest
import scipy.linalg

import oracle_lib as ol


@pytest.fixture(scope="module")
def pec():
    import sie_mom_amd as sm
    V, F, E, FE = ol.icosphere(2)

    class Mm:
        pass
    mesh = Mm()
    mesh.vertices, mesh.faces, mesh.edges, mesh.face_edges = V, F, E, FE
    mesh.nf, mesh.ne = len(F), len(E)
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(1.0), quadrature_rule=5)
    Z = ol.oracle_pec(mesh, sm.calderon_params(o))
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1))
    b = ol.oracle_rhs(mesh, 2, 5, w, o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability,
                      o.vacuum_permittivity)
    J = scipy.linalg.solve(Z, b)
    ph = sm.physical_parameters(o)
    eps, mu = o.vacuum_permittivity, o.vacuum_permeability
    return mesh, (V, F, E, FE), J, ph.omega, eps, mu, ph.omega * np.sqrt(mu * eps), ph.eta0


@pytest.mark.parametrize("radius,tol", [(0.5, 2e-3), (0.9, 2e-2)])
def test_extinction_inside_pec_sphere(pec, radius, tol):
    mesh, arrays, J, om, eps, mu, k, eta = pec
    rng = np.random.default_rng(5)
    u = rng.normal(size=(10, 3))
    pts = radius * u / np.linalg.norm(u, axis=1)[:, None]
    Es, Hs = ol.fields(ol.ORACLE_SO, "oracle_", arrays, 7, J, np.zeros_like(J), pts, om, eps, mu, k)
    ph = np.exp(-1j * k * pts[:, 2])
    Einc = np.zeros((len(pts), 3), complex)
    Einc[:, 0] = ph
    Hinc = np.zeros((len(pts), 3), complex)
    Hinc[:, 1] = ph / np.sqrt(mu / eps)
    assert np.abs(Es - Einc).max() < tol * np.abs(Einc).max()
    assert np.abs(Hs - Hinc).max() < tol * np.abs(Hinc).max()
    # the sign matters: the opposite field is far off
    assert np.abs(Es + Einc).max() > 1.9


def test_far_zone_matches_far_field(pec):
    mesh, arrays, J, om, eps, mu, k, eta = pec
    R = 400.0
    th = np.array([0.3, 1.2, 2.0, 2.8])
    dirs = np.array([[np.sin(t), 0, np.cos(t), np.cos(t), 0, -np.sin(t), 0, 1, 0] for t in th])
    A = ol.oracle_farfield_matrix(mesh, 5, k, eta, dirs)
    xs = np.zeros(4 * mesh.ne, complex)
    xs[2 * mesh.ne:3 * mesh.ne] = J
    Fff = (A @ xs).reshape(-1, 2)
    Ef, _ = ol.fields(ol.ORACLE_SO, "oracle_", arrays, 5, J, np.zeros_like(J), R * dirs[:, :3], om, eps, mu, k)
    Eth = np.einsum("ij,ij->i", Ef, dirs[:, 3:6]) * R * np.exp(1j * k * R)
    assert np.abs(Eth + Fff[:, 0]).max() < 5e-3 * np.abs(Fff[:, 0]).max()


def test_far_zone_electric_and_magnetic_currents(pec):
    """The magnetic-current columns of the far-field matrix: random J and M
    (scaled M, physical eta0 M) radiate in the far zone as minus the
    Schelkunoff far field, F_theta = c (eta0 N_theta + L_phi), F_phi =
    c (eta0 N_phi - L_theta), for both components and directions off the
    x-z plane. A sign or duality error in the L term fails: M alone is
    checked too, and its contribution is as large as J's."""
    mesh, arrays, _, om, eps, mu, k, eta = pec
    rng = np.random.default_rng(17)
    ne = mesh.ne
    J = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    Ms = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    R = 400.0
    th = np.array([0.4, 1.1, 1.9, 2.7])
    ph = np.array([0.3, 2.0, 3.5, 5.2])
    st, ct, sp, cp = np.sin(th), np.cos(th), np.sin(ph), np.cos(ph)
    dirs = np.stack([st * cp, st * sp, ct, ct * cp, ct * sp, -st, -sp, cp, 0 * th], axis=1)
    A = ol.oracle_farfield_matrix(mesh, 5, k, eta, dirs)
    for Jx, Mx in ((J, Ms), (0 * J, Ms)):
        xs = np.zeros(4 * ne, complex)
        xs[2 * ne:3 * ne] = Jx
        xs[3 * ne:] = Mx
        Fff = (A @ xs).reshape(-1, 2)
        Ef, _ = ol.fields(ol.ORACLE_SO, "oracle_", arrays, 5, Jx, eta * Mx, R * dirs[:, :3], om, eps, mu, k)
        g = R * np.exp(1j * k * R)
        Eth = np.einsum("ij,ij->i", Ef, dirs[:, 3:6]) * g
        Eph = np.einsum("ij,ij->i", Ef, dirs[:, 6:9]) * g
        scale = np.abs(Fff).max()
        assert np.abs(Eth + Fff[:, 0]).max() < 5e-3 * scale
        assert np.abs(Eph + Fff[:, 1]).max() < 5e-3 * scale
