"""The circular-arc spindle of the 72k synthesis workload (a 2 lambda chord,
arc radius 1.5 lambda; the reference's CircularOgive, src/sie_mom/mesh/
circular_ogive.py:13-80) through the GPU path against the oracle: small
spindles in full (assembly, mixed-precision solve, objective, gradient), and
the workload itself (N = 72,048, rule 17, 10 parameters per face) through
sampled rows of Z and the FP64 residual of its refined solution."""
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


def spindle_problem(sm, rings, rule, kind=10, mode="mixed", seed=20260101):
    from sie_mom_amd import adjoint as ad
    mesh = sm.Mesh.spindle(2.0, 1.5, rings)
    k0 = 2 * math.pi
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule)
    waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")]
    params = np.random.default_rng(seed).uniform(-2, 2, size=(mesh.nf, kind)) / k0
    prob = ad.GstcProblem(mesh, o, waves, kind, mode=mode, ir_tol=1e-15, ir_stop="backward", variant="f16x3")
    return mesh, o, waves, prob, params


@pytest.mark.parametrize("rule", [7, 17])
def test_spindle_assembly_matches_oracle(sm, rule):
    from sie_mom_amd import adjoint as ad
    mesh, o, waves, prob, params = spindle_problem(sm, 12, rule)
    print(f"spindle assembly F={mesh.nf} rule={rule}", flush=True)
    rec = ad.records(torch.from_numpy(params), 10, o)
    Z = sm.ops.assemble_gstc(mesh, o, rec).cpu().numpy()
    Zo = ol.oracle_gstc(mesh, sm.calderon_params(o), rec.cpu().numpy(), False, 0)
    assert np.linalg.norm(Z - Zo) / np.linalg.norm(Zo) <= 1e-12


def test_spindle_solve_objective_gradient(sm):
    """Mixed-precision solve, cone objective and every gradient entry of a
    218-face spindle against the CPU FP64 oracle (dense solves, far field,
    record sensitivities), plus central differences of the GPU objective."""
    from sie_mom_amd import adjoint as ad
    from threadpoolctl import threadpool_limits
    mesh, o, waves, prob, params = spindle_problem(sm, 12, 7)
    print(f"spindle solve F={mesh.nf} N={4 * mesh.ne}", flush=True)
    st = prob.solve(torch.from_numpy(params))
    J, F = prob.objective(st)
    grad = prob.gradient(st, F).cpu().numpy()
    ph = sm.physical_parameters(o)
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    eta = math.sqrt(o.vacuum_permeability / o.vacuum_permittivity)
    rhs = ol.oracle_rhs(mesh, 0, 7, waves[0], o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability,
                        o.vacuum_permittivity)
    with threadpool_limits(limits=1, user_api="blas"):
        Jo, go, xo, _ = ol.oracle_gradient(mesh, sm.calderon_params(o), 0, params, 10, ad.susceptibility_scales(o),
                                           eta, rhs, 7, k0, ph.eta0, prob.dirs, prob.weights)
    assert np.linalg.norm(st.xs.cpu().numpy() - xo) <= 1e-10 * np.linalg.norm(xo)
    assert abs(float(J) - Jo) <= 1e-10 * abs(Jo)
    err = np.abs(grad - go).max() / np.abs(go).max()
    print(f"  gradient vs oracle {err:.2e}; refinement steps {st.ir_iterations} / {st.adj_iterations}", flush=True)
    assert err <= 1e-9
    rng = np.random.default_rng(3)
    h = 1e-6
    for _ in range(4):
        f, q = int(rng.integers(0, mesh.nf)), int(rng.integers(0, 10))
        pp, pm = params.copy(), params.copy()
        pp[f, q] += h
        pm[f, q] -= h
        Jp = float(prob.objective(prob.solve(torch.from_numpy(pp)))[0])
        Jm = float(prob.objective(prob.solve(torch.from_numpy(pm)))[0])
        fd = (Jp - Jm) / (2 * h)
        assert abs(grad[f, q] - fd) <= 1e-6 * np.abs(grad).max() + 1e-5 * abs(fd), (f, q, grad[f, q], fd)


def test_spindle_72k_rows_and_residual(sm):
    """The synthesis workload (F = 12,008, N = 72,048, rule 17, 10 parameters
    per face, seed 20260101): 32 sampled rows of the GPU matrix (8 per
    unknown block) against the oracle's rows, integrated for their test faces
    against every source face, and the FP64 residual of the FP32-MFMA + FGMRES-IR
    solution computed by torch's own GEMV and norms."""
    from sie_mom_amd import adjoint as ad
    from threadpoolctl import threadpool_limits
    mesh, o, waves, prob, params = spindle_problem(sm, 91, 17)
    N = 4 * mesh.ne
    assert (mesh.nf, N) == (12008, 72048)
    print(f"spindle 72k: F={mesh.nf} N={N}", flush=True)
    rng = np.random.default_rng(72048)
    rows = np.concatenate([b * mesh.ne + rng.choice(mesh.ne, 8, replace=False) for b in range(4)])
    st = prob.solve(torch.from_numpy(params))
    Zrows = st.Z[torch.from_numpy(rows).cuda()].cpu().numpy()
    with threadpool_limits(limits=1, user_api="blas"):
        Ro = ol.oracle_gstc_rows(mesh, sm.calderon_params(o), st.rec.cpu().numpy(), False, 0, rows)
    rel = np.linalg.norm(Zrows - Ro, axis=1) / np.linalg.norm(Ro, axis=1)
    print(f"  sampled rows: max rel err {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-12
    b = sm.ops.rhs_plane_waves(mesh, o, waves)[:, 0]
    r = torch.mv(st.Z, st.xs) - b
    zf = float(torch.linalg.norm(st.Z))
    be = float(torch.linalg.norm(r)) / (zf * float(torch.linalg.norm(st.xs)) + float(torch.linalg.norm(b)))
    print(f"  refinement: {st.ir_iterations} steps, backward error {st.ir_backward_error:.2e} "
          f"(torch: {be:.2e}), |Z|_F {zf:.4e} vs {st.znorm:.4e}", flush=True)
    assert abs(zf - st.znorm) <= 1e-12 * zf
    assert be <= 2e-15 and st.ir_backward_error <= 1e-15
