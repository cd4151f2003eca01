"""Adjoint gradient of the GSTC cone objective: oracle parity of J, central
finite differences, FP64 vs FP32-MFMA + FP64 refinement."""
import math

import numpy as np
import pytest
import scipy.linalg

import oracle_lib as ol

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sm():
    import sie_mom_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return sie_mom_amd


def make_problem(sm, level=1, kind=3, rule=5, mode="fp64", k0=4.0, basis="face", seed=20260101, ir_tol=1e-13,
                 ir_stop="residual"):
    from sie_mom_amd import adjoint as ad
    mesh = sm.Mesh.icosphere(1.0, level)
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule, basis=basis)
    waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")]
    rng = np.random.default_rng(seed)
    params = rng.uniform(-2, 2, size=(mesh.nf, kind)) / k0
    prob = ad.GstcProblem(mesh, o, waves, kind, mode=mode, ir_tol=ir_tol, ir_stop=ir_stop)
    return mesh, o, waves, prob, params


def J_of(prob, params):
    st = prob.solve(torch.from_numpy(params))
    J, F = prob.objective(st)
    return float(J), st, F


def test_objective_matches_oracle(sm):
    from sie_mom_amd import adjoint as ad
    mesh, o, waves, prob, params = make_problem(sm, level=1)
    J, st, F = J_of(prob, params)
    # oracle: records, matrix, rhs, dense solve, far field in numpy
    bs, xs_, ks = ad.susceptibility_scales(o)
    eta = math.sqrt(o.vacuum_permeability / o.vacuum_permittivity)
    rec = ol.oracle_records(params, 3, bs, xs_, ks, eta)
    rec_gpu = ad.records(torch.from_numpy(params), 3, o).cpu().numpy()
    assert np.abs(rec_gpu - rec).max() / np.abs(rec).max() < 1e-13
    Z = ol.oracle_gstc(mesh, sm.calderon_params(o), rec, False, 0)
    b = ol.oracle_rhs(mesh, 0, o.quadrature_rule, waves[0], o.vacuum_permeability, o.vacuum_permittivity,
                      o.vacuum_permeability, o.vacuum_permittivity)
    x = scipy.linalg.solve(Z, b)
    assert np.linalg.norm(st.xs.cpu().numpy() - x) / np.linalg.norm(x) < 1e-10
    ph = sm.physical_parameters(o)
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    A = ol.oracle_farfield_matrix(mesh, o.quadrature_rule, k0, ph.eta0, prob.dirs)
    Fo = (A @ x).reshape(-1, 2)
    Jo = np.sum(prob.weights * np.sum(np.abs(Fo) ** 2, axis=1)) / (2 * ph.eta0)
    assert np.abs(F.cpu().numpy() - Fo).max() / np.abs(Fo).max() < 1e-10
    assert abs(J - Jo) / abs(Jo) < 1e-10


@pytest.mark.parametrize("kind,basis", [(3, "face"), (10, "face"), (3, "spherical")])
def test_gradient_vs_central_fd(sm, kind, basis):
    from sie_mom_amd import adjoint as ad
    mesh, o, waves, prob, params = make_problem(sm, level=1, kind=kind, basis=basis)
    J, st, F = J_of(prob, params)
    grad = prob.gradient(st, F).cpu().numpy()
    rng = np.random.default_rng(4)
    h = 1e-6
    for t in range(6):
        f = int(rng.integers(0, mesh.nf))
        k = int(rng.integers(0, kind))
        pp, pm = params.copy(), params.copy()
        pp[f, k] += h
        pm[f, k] -= h
        fd = (J_of(prob, pp)[0] - J_of(prob, pm)[0]) / (2 * h)
        scale = np.abs(grad).max()
        assert abs(grad[f, k] - fd) <= 1e-6 * scale + 1e-5 * abs(fd), (f, k, grad[f, k], fd)


def test_mixed_matches_fp64(sm):
    mesh, o, waves, p64, params = make_problem(sm, level=2, mode="fp64")
    _, _, _, pmx, _ = make_problem(sm, level=2, mode="mixed")
    J64, s64, F64 = J_of(p64, params)
    Jmx, smx, Fmx = J_of(pmx, params)
    x64, xmx = s64.xs.cpu().numpy(), smx.xs.cpu().numpy()
    assert np.linalg.norm(xmx - x64) / np.linalg.norm(x64) < 1e-10
    assert abs(Jmx - J64) / abs(J64) < 1e-10
    g64 = p64.gradient(s64, F64).cpu().numpy()
    gmx = pmx.gradient(smx, Fmx).cpu().numpy()
    assert np.linalg.norm(gmx - g64) / np.linalg.norm(g64) < 1e-9


def test_autograd_function(sm):
    from sie_mom_amd import adjoint as ad
    mesh, o, waves, prob, params = make_problem(sm, level=1, mode="mixed")
    p = torch.from_numpy(params).cuda().requires_grad_(True)
    J = ad.GstcConeObjective.apply(p, prob)
    J.backward()
    st = prob.solve(torch.from_numpy(params))
    J2, F = prob.objective(st)
    g = prob.gradient(st, F)
    assert torch.allclose(p.grad, g, rtol=1e-9, atol=0)
    assert abs(float(J) - float(J2)) <= 1e-12 * abs(float(J))


def oracle_gradient_of(sm, mesh, o, waves, prob, params, kind):
    """The CPU FP64 gradient (tests/oracle_lib.oracle_gradient) of the same
    problem: oracle matrix, scipy solves, numpy far field, record
    sensitivities from the oracle's mass assembly."""
    from sie_mom_amd import adjoint as ad
    from threadpoolctl import threadpool_limits
    ph = sm.physical_parameters(o)
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    eta = math.sqrt(o.vacuum_permeability / o.vacuum_permittivity)
    rhs = ol.oracle_rhs(mesh, 0, o.quadrature_rule, waves[0], o.vacuum_permeability, o.vacuum_permittivity,
                        o.vacuum_permeability, o.vacuum_permittivity)
    basis = {"face": 0, "spherical": 1}[o.basis]
    with threadpool_limits(limits=1, user_api="blas"):
        return ol.oracle_gradient(mesh, sm.calderon_params(o), basis, params, kind, ad.susceptibility_scales(o), eta,
                                  rhs, o.quadrature_rule, k0, ph.eta0, prob.dirs, prob.weights)


@pytest.mark.parametrize("level,kind,rule,mode", [(1, 3, 5, "fp64"), (1, 10, 5, "fp64"), (2, 3, 5, "mixed"),
                                                  (2, 10, 7, "fp64"), (2, 3, 17, "mixed")])
def test_gradient_matches_oracle(sm, level, kind, rule, mode):
    """Every gradient entry against the CPU FP64 gradient: <= 1e-9 of the
    largest entry (FP64 and FP32-MFMA + refinement to a normwise backward
    error of 1e-15), and the per-face record sensitivities behind it."""
    from sie_mom_amd import adjoint as ad
    mesh, o, waves, prob, params = make_problem(sm, level=level, kind=kind, rule=rule, mode=mode, ir_tol=1e-15,
                                                ir_stop="backward")
    print(f"oracle gradient level={level} kind={kind} rule={rule} mode={mode}", flush=True)
    J, st, F = J_of(prob, params)
    grad = prob.gradient(st, F).cpu().numpy()
    Jo, go, xo, So = oracle_gradient_of(sm, mesh, o, waves, prob, params, kind)
    assert abs(J - Jo) <= 1e-10 * abs(Jo)
    assert np.linalg.norm(st.xs.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    err = np.abs(grad - go).max() / np.abs(go).max()
    print(f"  max |grad - oracle| / max |oracle| = {err:.2e}", flush=True)
    assert err <= 1e-9
    if mode == "fp64":
        # the contraction's per-face sensitivities of lambda^H Zs x (same lambda, x)
        u = (F * (prob.w_dev / (2 * prob.eta0))[:, None]).contiguous()
        g = ad.farfield_adjoint(mesh, o, u.view(-1), prob.dirs)
        lam = g.clone()
        st.lu.solve(lam, op="C")
        G = ad.mass_sensitivity(mesh, o, lam, st.xs).cpu().numpy().reshape(-1, 32)
        assert np.abs(G - So).max() <= 1e-9 * np.abs(So).max()
