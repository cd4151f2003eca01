"""Oracle vs the reference's own host build (oracle/_ref), on random inputs.
Skipped where the reference build is absent."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.skipif(not ol.ref_available(), reason="oracle/_ref not built")


def test_random_pairs_bitexact():
    o, r = ol.oracle(), ol.ref()
    rng = np.random.default_rng(3)
    V, F, E, FE = ol.icosphere(2)
    for t in range(30):
        m, n = rng.integers(0, len(F), 2)
        if t % 5 == 0:
            n = m
        tt, st = V[F[m]], V[F[n]]
        k1, k2 = rng.uniform(1, 12, 2)
        for rule in (3, 6, 11):
            for sing in (0, 1):
                np.testing.assert_array_equal(o.pair_regular(tt, F[m], st, F[n], rule, k1, k2, sing),
                                              r.pair_regular(tt, F[m], st, F[n], rule, k1, k2, sing))
            np.testing.assert_array_equal(o.pair_singular(tt, F[m], st, F[n], rule, k1, k2),
                                          r.pair_singular(tt, F[m], st, F[n], rule, k1, k2))
        p = rng.normal(size=3)
        np.testing.assert_array_equal(o.wilton(p, tt), r.wilton(p, tt))
        for b in (0, 1):
            np.testing.assert_array_equal(np.concatenate(o.face_basis(tt, b)), np.concatenate(r.face_basis(tt, b)))


def test_all_rules_bitexact():
    o, r = ol.oracle(), ol.ref()
    for rule in range(1, 21):
        a, b = o.rule(rule), r.rule(rule)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_full_gstc_level1():
    V, F, E, FE = ol.icosphere(1)
    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    c0 = 1 / np.sqrt(mu0 * eps0)
    f = 4.0 * c0 / (2 * np.pi)
    om = 2 * np.pi * f
    phys = dict(k1=om * np.sqrt(mu0 * eps0 * 3.0), k2=om * np.sqrt(mu0 * eps0), mu_r1=1, mu_r2=1, eps_r1=3.0,
                eps_r2=1, mu0=mu0, eps0=eps0, omega=om, thr=0.1 * c0 / f)
    t = ol.random_gstc_tensors(len(F), seed=2)
    for basis in (0, 1):
        Zo = ol.assemble_gstc(ol.ORACLE_SO, "oracle_", (V, F, E, FE), 5, phys, t, False, basis)
        Zr = ol.assemble_gstc(ol.REF_SO, "ref_", (V, F, E, FE), 5, phys, t, False, basis)
        assert np.linalg.norm(Zo - Zr) / np.linalg.norm(Zr) < 1e-15


def test_fields_vs_reference():
    """Surface-current fields: the oracle against the reference's own
    calculateScatterFieldQuadrature / calculateScatterFieldSingular loop, at
    far, near-surface (singular extraction) and interior points."""
    rng = np.random.default_rng(21)
    mesh = ol.icosphere(1)
    V, F, E, FE = mesh
    ne = len(E)
    J = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    M = rng.normal(size=ne) + 1j * rng.normal(size=ne)
    obs = ol.field_points(V, F, rng)
    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    for k0, er in ((4.0, 1.0), (2.0, 3.0)):
        om = k0 / np.sqrt(mu0 * eps0)
        eps, mu = er * eps0, mu0
        k = om * np.sqrt(mu * eps)
        Eo, Ho = ol.fields(ol.ORACLE_SO, "oracle_", mesh, 7, J, M, obs, om, eps, mu, k)
        Er, Hr = ol.fields(ol.REF_SO, "ref_", mesh, 7, J, M, obs, om, eps, mu, k)
        np.testing.assert_array_equal(Eo, Er)
        np.testing.assert_array_equal(Ho, Hr)
