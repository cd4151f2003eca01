"""Pins the oracle (own CPU restatement) against golden vectors written by the
reference's own host numerics (tests/golden/make_golden.py). Runs without the
reference tree and without a GPU."""
import ctypes
import os

import numpy as np
import pytest

import oracle_lib as ol

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


class M:
    def __init__(self, V, F, E, FE):
        self.vertices, self.faces, self.edges, self.face_edges = V, F, E, FE
        self.nf, self.ne = len(F), len(E)


def P(vals, rule):
    class p:
        pass
    q = p()
    for k, v in zip(("k1", "k2", "mu_r1", "mu_r2", "eps_r1", "eps_r2", "mu0", "eps0", "omega", "threshold"), vals):
        setattr(q, k, float(v))
    q.rule = rule
    return q


@pytest.fixture(scope="module")
def o():
    return ol.oracle()


@pytest.mark.parametrize("rule", [5, 7, 17])
def test_rules_bitexact(o, rule):
    xy, w = o.rule(rule)
    assert np.array_equal(xy.view(np.uint32), G[f"rule{rule}_xy"].view(np.uint32))
    assert np.array_equal(w.view(np.uint32), G[f"rule{rule}_w"].view(np.uint32))


def test_icosphere_bitexact():
    V, F, E, FE = ol.icosphere(1)
    for a, n in zip((V, F, E, FE), ("V", "F", "E", "FE")):
        assert np.array_equal(a, G[f"ico1_{n}"])


@pytest.mark.parametrize("name", ["self", "edge", "vertex", "far"])
@pytest.mark.parametrize("rule", [5, 7])
def test_pair_blocks(o, name, rule):
    V, F = G["ico1_V"], G["ico1_F"]
    m, n = G[f"pairidx_{name}"]
    for sing in (0, 1):
        got = o.pair_regular(V[F[m]], F[m], V[F[n]], F[n], rule, 4.0, 9.0, sing)
        np.testing.assert_array_equal(got, G[f"pair_{name}_r{rule}_s{sing}"])
    got = o.pair_singular(V[F[m]], F[m], V[F[n]], F[n], rule, 4.0, 9.0)
    np.testing.assert_array_equal(got, G[f"sing_{name}_r{rule}"])


def test_wilton(o):
    tri = G["wilton_tri"]
    for p, ref in zip(G["wilton_pts"], G["wilton_out"]):
        np.testing.assert_array_equal(o.wilton(p, tri), ref)


def test_mass_and_excitation(o):
    tri = G["wilton_tri"]
    F = G["ico1_F"]
    np.testing.assert_array_equal(o.mass_batch(tri, F[3], 7, G["mass_T"]), G["mass_out"])
    ve, vh = o.excitation_planewave(tri, F[3], 7, np.array([1.0, 0, 0]), np.array([0, 0, 1.0]), 4.0, 376.73)
    np.testing.assert_array_equal(ve, G["exc_VE"])
    np.testing.assert_array_equal(vh, G["exc_VH"])


@pytest.mark.parametrize("basis", [0, 1])
def test_gstc_matrix(basis):
    mesh = M(G["ico0_V"], G["ico0_F"], G["ico0_E"], G["ico0_FE"])
    Z = ol.oracle_gstc(mesh, P(G["gstc_phys"], 7), G["gstc_tensors"], False, basis)
    ref = G[f"gstc_Z_b{basis}"]
    assert np.linalg.norm(Z - ref) / np.linalg.norm(ref) < 1e-15


def test_pec_and_pmchwt_matrices():
    m1 = M(G["ico1_V"], G["ico1_F"], G["ico1_E"], G["ico1_FE"])
    k2, mur, mu0, om, thr = G["pec_phys"]
    Z = ol.oracle_pec(m1, P([k2, k2, 1, mur, 1, 1, mu0, 1, om, thr], 5))
    assert np.linalg.norm(Z - G["pec_Z"]) / np.linalg.norm(G["pec_Z"]) < 1e-15
    m0 = M(G["ico0_V"], G["ico0_F"], G["ico0_E"], G["ico0_FE"])
    Z = ol.oracle_pmchwt(m0, P(G["pmchwt_phys"], 7))
    assert np.linalg.norm(Z - G["pmchwt_Z"]) / np.linalg.norm(G["pmchwt_Z"]) < 1e-15


def test_rhs():
    m1 = M(G["ico1_V"], G["ico1_F"], G["ico1_E"], G["ico1_FE"])
    k, eta = G["rhs_gstc_k_eta"]

    class W:
        E0 = (1.0, 0.0, 0.0)
        khat = (0.0, 0.0, 1.0)
    w = W()
    w.k, w.eta, w.region = k, eta, 2
    mu0, eps0 = 1.25663706127e-6, 8.8541878188e-12
    rhs = ol.oracle_rhs(m1, 0, 7, w, mu0, eps0, mu0, eps0)
    np.testing.assert_array_equal(rhs, G["rhs_gstc"])


@pytest.mark.parametrize("region", [2, 1])
def test_pmchwt_rhs(region):
    """T-PMCHWT plane-wave RHS: sign +1 (exterior source) / -1 (interior) and
    sign * eta0 on the magnetic rows (dielectric_solver.cu:400-430, 515-531)."""
    m1 = M(G["ico1_V"], G["ico1_F"], G["ico1_E"], G["ico1_FE"])
    k, eta = G[f"rhs_pmchwt_r{region}_k_eta"]
    vmu, veps = G["rhs_pmchwt_vac"]

    class W:
        E0 = (1.0, 0.0, 0.0)
        khat = (0.0, 0.0, 1.0)
    w = W()
    w.k, w.eta, w.region = k, eta, region
    rhs = ol.oracle_rhs(m1, 1, 7, w, vmu, veps, vmu, veps)
    np.testing.assert_array_equal(rhs, G[f"rhs_pmchwt_r{region}"])


R = np.load(os.path.join(os.path.dirname(__file__), "golden", "records_vectors.npz"))


@pytest.mark.parametrize("kind", [10, 3])
def test_records_match_reference(kind):
    """Surface-map records from real parameters against the reference's own
    parameter basis, omega model and record export (tests/golden/
    make_records_golden.py): the maps bit for bit, the inverses (a different
    solve) to 1e-13 of each face's largest entry; and the reference's
    export checks, restated: reciprocity and skew-Hermitian balanced maps."""
    bs, xs, ks = R["scales"]
    eta = float(R["eta"])
    got = ol.oracle_records(R[f"params{kind}"], kind, bs, xs, ks, eta).reshape(-1, 2, 16)
    ref = R[f"records{kind}"].reshape(-1, 2, 16)
    np.testing.assert_array_equal(got[:, 0], ref[:, 0])
    scale = np.abs(ref[:, 1]).max(axis=1, keepdims=True)
    assert (np.abs(got[:, 1] - ref[:, 1]) / scale).max() < 1e-13
    L = got[:, 0].reshape(-1, 4, 4)
    t = np.array([np.sqrt(eta)] * 2 + [1 / np.sqrt(eta)] * 2)
    bal = L * t[None, :, None] * t[None, None, :]
    sc = np.abs(bal).max(axis=(1, 2))
    assert (np.abs(bal + np.conj(np.transpose(bal, (0, 2, 1)))).max(axis=(1, 2)) / sc).max() < 1e-15
    assert (np.abs(bal[:, :2, :2] - bal[:, :2, :2].transpose(0, 2, 1)).max(axis=(1, 2)) / sc).max() < 1e-15
    assert (np.abs(bal[:, 2:, :2] + bal[:, :2, 2:].transpose(0, 2, 1)).max(axis=(1, 2)) / sc).max() < 1e-15
    Li = got[:, 1].reshape(-1, 4, 4)
    assert np.abs(L @ Li - np.eye(4)).max() < 1e-8


def test_fields():
    """E and H of random currents on the level-1 sphere (far, near-surface with
    singular extraction, interior points) against the reference's numerics."""
    mesh = (G["ico1_V"], G["ico1_F"], G["ico1_E"], G["ico1_FE"])
    om, eps, mu, k = G["fields_phys"]
    Eo, Ho = ol.fields(ol.ORACLE_SO, "oracle_", mesh, 7, G["fields_J"], G["fields_M"], G["fields_obs"], om, eps,
                       mu, k)
    np.testing.assert_array_equal(Eo, G["fields_E"])
    np.testing.assert_array_equal(Ho, G["fields_H"])
