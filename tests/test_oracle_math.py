"""Closed forms and invariants that pin the oracle independently of the
reference: quadrature exactness, Wilton potentials against brute force,
singularity extraction consistency, operator symmetries, scaling identity."""
import math

import numpy as np
import pytest

import oracle_lib as ol


@pytest.fixture(scope="module")
def o():
    return ol.oracle()


@pytest.mark.parametrize("rule", list(range(1, 21)))
def test_dunavant_exactness(o, rule):
    xy, w = o.rule(rule)
    xy = xy.astype(float)
    w = w.astype(float)
    assert abs(w.sum() - 1) < 1e-6
    for i in range(rule + 1):
        for j in range(rule + 1 - i):
            exact = 2.0 * math.factorial(i) * math.factorial(j) / math.factorial(i + j + 2)
            got = np.sum(w * xy[:, 0] ** i * xy[:, 1] ** j)
            assert abs(got - exact) < 2e-6, (i, j)


def brute_phi(r, tri, n=400):
    """Midpoint-rule integrals of 1/R and (r'-r)/R over the triangle (subdivided)."""
    a, b, c = tri
    s = (np.arange(n) + 0.5) / n
    u, v = np.meshgrid(s, s, indexing="ij")
    mask = u + v < 1
    u, v = u[mask], v[mask]
    pts = a + np.outer(u, b - a) + np.outer(v, c - a)
    area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
    dA = area / (len(u))
    R = np.linalg.norm(pts - r, axis=1)
    return np.sum(1 / R) * dA, np.sum((pts - r) / R[:, None], axis=0) * dA


def test_wilton_phi0_phi1_bruteforce(o):
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.2, 0.9, 0]])
    for r in ([0.3, 0.3, 0.4], [1.5, -0.2, 0.3], [0.4, 0.2, -0.7]):
        out = o.wilton(np.array(r), tri)
        phi0, phi1b, rpi, d = out[0], out[1:4], out[8:11], out[11]
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n /= np.linalg.norm(n)
        phi1 = phi1b - d * phi0 * n  # phi_1 = phi_1_base - d n phi_0 (semi_analytic_integrals.cuh:437-440)
        b0, b1 = brute_phi(np.array(r), tri)
        assert abs(phi0 - b0) / abs(b0) < 2e-3
        assert np.linalg.norm(phi1 - b1) / np.linalg.norm(b1) < 2e-3


def test_regular_plus_singular_equals_direct(o):
    """For well-separated faces the regularised quadrature plus the analytic
    1/R part equals the plain quadrature (up to the outer rule's error)."""
    V, F, E, FE = ol.icosphere(2)
    m, n = 0, 40
    tt, st = V[F[m]], V[F[n]]
    assert np.linalg.norm(tt.mean(0) - st.mean(0)) > 0.3
    for rule in (7, 13):
        direct = o.pair_regular(tt, F[m], st, F[n], rule, 3.0, 5.0, 0)
        split = o.pair_regular(tt, F[m], st, F[n], rule, 3.0, 5.0, 1) + o.pair_singular(tt, F[m], st, F[n], rule, 3.0, 5.0)
        assert np.linalg.norm(direct - split) / np.linalg.norm(direct) < 1e-4


def test_tl_symmetry_and_dad_identity():
    """TL blocks are symmetric (Z_TL(m,n) = Z_TL(n,m)^T) and the GSTC matrix is
    D A D of the unscaled Calderon operator with D = diag(1, eta0, 1, eta0)."""
    V, F, E, FE = ol.icosphere(1)

    class Mm:
        pass
    mesh = Mm()
    mesh.vertices, mesh.faces, mesh.edges, mesh.face_edges = V, F, E, FE
    mesh.nf, mesh.ne = len(F), len(E)

    class P:
        pass
    p = P()
    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    c0 = 1 / np.sqrt(mu0 * eps0)
    f = 4.0 * c0 / (2 * np.pi)
    om = 2 * np.pi * f
    for k, v in dict(k1=om * np.sqrt(mu0 * eps0), k2=om * np.sqrt(mu0 * eps0), mu_r1=1, mu_r2=1, eps_r1=1, eps_r2=1,
                     mu0=mu0, eps0=eps0, omega=om, threshold=0.1 * c0 / f).items():
        setattr(p, k, float(np.float32(v)))
    p.rule = 7
    Z = ol.oracle_gstc(mesh, p, np.zeros(32 * len(F), complex), False, 0, True, False)
    ne = len(E)
    Jb = Z[:ne, :ne]
    # far-pair quadrature is exactly symmetric; the singular extraction
    # (analytic inner integral over the source face only) is symmetric up to
    # the outer rule's discretisation error
    assert np.linalg.norm(Jb - Jb.T) / np.linalg.norm(Jb) < 1e-4
    eta0 = np.sqrt(np.float64(np.float32(p.mu0) / np.float32(p.eps0)))
    # J-M coupling block = eta0 * TK, M-J = -eta0 * TK, M-M = eta0^2 eps/mu * J-J
    jm, mj, mm = Z[:ne, ne:2 * ne], Z[ne:2 * ne, :ne], Z[ne:2 * ne, ne:2 * ne]
    assert np.linalg.norm(jm + mj) / np.linalg.norm(jm) < 1e-15
    ratio = (eta0 ** 2) * np.float64(np.float32(p.eps0)) / np.float64(np.float32(p.mu0))
    assert np.linalg.norm(mm - ratio * Jb) / np.linalg.norm(mm) < 1e-6


def test_unscaled_operator_is_dad_of_scaled():
    """The unscaled GSTC operator (eta0 = 1, the operator of the reference's
    matrix-free matvec) and the eta0-scaled one satisfy Z_s = D Z_u D with
    D = diag(1, eta0, 1, eta0), for the Calderon and the mass parts alike. The
    mass part rounds eta0 in float, hence the 1e-6 bound."""
    V, F, E, FE = ol.icosphere(1)

    class Mm:
        pass
    mesh = Mm()
    mesh.vertices, mesh.faces, mesh.edges, mesh.face_edges = V, F, E, FE
    mesh.nf, mesh.ne = len(F), len(E)

    class P:
        pass
    p = P()
    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    c0 = 1 / np.sqrt(mu0 * eps0)
    f = 4.0 * c0 / (2 * np.pi)
    om = 2 * np.pi * f
    for k, v in dict(k1=om * np.sqrt(2.0 * mu0 * eps0), k2=om * np.sqrt(mu0 * eps0), mu_r1=1, mu_r2=1, eps_r1=2,
                     eps_r2=1, mu0=mu0, eps0=eps0, omega=om, threshold=0.1 * c0 / f).items():
        setattr(p, k, float(np.float32(v)))
    p.rule = 5
    t = ol.random_gstc_tensors(len(F), seed=11)
    ne = len(E)
    eta0 = np.sqrt(np.float64(np.float32(p.mu0) / np.float32(p.eps0)))
    d = np.tile(np.repeat([1.0, eta0], ne), 2)
    for cal, mass in ((True, False), (False, True), (True, True)):
        Zs = ol.oracle_gstc(mesh, p, t, False, 1, cal, mass, scaled=True)
        Zu = ol.oracle_gstc(mesh, p, t, False, 1, cal, mass, scaled=False)
        assert np.linalg.norm(Zs - Zu) / np.linalg.norm(Zs) > 1e-2  # eta0 ~ 377 is really applied
        dzd = d[:, None] * Zu * d[None, :]
        assert np.linalg.norm(Zs - dzd) / np.linalg.norm(Zs) < 1e-6, (cal, mass)
