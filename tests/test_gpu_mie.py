"""End to end on the GPU: PEC sphere (T-EFIE assembly, plane-wave RHS, LU,
far field) against the Mie series at k a = 1, on meshes of 1,920 and 7,680
edges; the RCS error falls at second order in the mesh size."""
import math

import numpy as np
import pytest

import mie
from test_oracle_mie import cut_directions, rcs_errors

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sm():
    import sie_mom_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return sie_mom_amd


def test_pec_sphere_rcs_matches_mie(sm):
    from sie_mom_amd import adjoint as ad
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(1.0), quadrature_rule=5)
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1))
    th, dirs = cut_directions()
    errs = []
    for level in (3, 4):
        mesh = sm.Mesh.icosphere(1.0, level)
        print(f"pec mie level={level} ne={mesh.ne}", flush=True)
        Z = sm.ops.assemble_pec(mesh, o)
        b = sm.ops.rhs_plane_waves(mesh, o, [w])
        lu = sm.linalg.DenseLuSolver()
        lu.factorize(Z)
        J = lu.solve(b[:, 0].contiguous())
        xs = torch.zeros(4 * mesh.ne, dtype=torch.complex128, device="cuda")
        xs[2 * mesh.ne:3 * mesh.ne] = J
        F = ad.farfield(mesh, o, xs, dirs).cpu().numpy()
        eE, eH, cross = rcs_errors(F, th)
        assert cross < 1e-6
        errs.append(max(eE, eH))
        del Z, lu
    assert errs[0] < 0.01 and errs[1] < 0.003, errs
    assert errs[0] / errs[1] > 3.0, errs


def test_pec_sphere_ka3_181_angles(sm):
    """PEC sphere, k a = 3, icosphere level 3 (1,280 faces, 1,920 edges),
    Dunavant rule 5, x-polarised plane wave along +z, bistatic RCS at 181
    angles of the E-plane against the Mie series (and, at the same angles,
    the CPU oracle's RCS of the same discretisation)."""
    from sie_mom_amd import adjoint as ad
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(3.0), quadrature_rule=5)
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1))
    th = np.radians(np.arange(181.0))
    dirs = np.array([[math.sin(t), 0, math.cos(t), math.cos(t), 0, -math.sin(t), 0, 1, 0] for t in th])
    mesh = sm.Mesh.icosphere(1.0, 3)
    print(f"pec ka=3 level=3 ne={mesh.ne}", flush=True)
    Z = sm.ops.assemble_pec(mesh, o)
    b = sm.ops.rhs_plane_waves(mesh, o, [w])
    lu = sm.linalg.DenseLuSolver()
    lu.factorize(Z)
    J = lu.solve(b[:, 0].contiguous())
    xs = torch.zeros(4 * mesh.ne, dtype=torch.complex128, device="cuda")
    xs[2 * mesh.ne:3 * mesh.ne] = J
    F = ad.farfield(mesh, o, xs, dirs).cpu().numpy()
    ph = sm.physical_parameters(o)
    k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    sE = 4 * math.pi * np.abs(F[:, 0]) ** 2
    mE, _ = mie.bistatic_rcs(3.0, th, k=k0)
    err = np.abs(sE - mE).max() / mE.max()
    rms_db = np.sqrt(np.mean((10 * np.log10(sE / mE)) ** 2))
    print(f"  RCS vs Mie: max {err:.2e} of peak, {rms_db:.3f} dB rms", flush=True)
    assert err < 0.015 and rms_db < 0.1
    assert np.abs(F[:, 1]).max() < 1e-6 * np.abs(F).max()


def _pmchwt_sphere(sm, level, ka=5.0, eps_r=4.0, rule=17):
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(ka), quadrature_rule=rule)
    o.interior.relative_permittivity = eps_r
    return sm.Mesh.icosphere(1.0, level), o


def test_pmchwt_sphere_fp64_mixed_and_mie(sm):
    """T-PMCHWT dielectric sphere, eps_r = 4, k a = 5, rule 17, one plane wave,
    at icosphere levels 3 and 4 (N = 3,840 and 15,360): the FP32-MFMA LU with
    FGMRES-IR (backward error 1e-15) against the plain FP64 LU, currents and
    far field to <= 1e-8; both against the Mie series, whose error falls at
    second order in the mesh size."""
    from sie_mom_amd import adjoint as ad
    errs = []
    for level in (3, 4):
        mesh, o = _pmchwt_sphere(sm, level)
        n = 2 * mesh.ne
        print(f"pmchwt level={level} N={n}", flush=True)
        w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")
        b = sm.ops.rhs_plane_waves(mesh, o, [w])[:, 0].contiguous()
        Z = sm.ops.assemble_pmchwt(mesh, o)
        LU32, zf = sm.linalg.to_precision(Z, 32, fro=True)
        lu32 = sm.linalg.DenseLuSolver(variant="f16x3")
        lu32.factorize(LU32)
        xm, rr, its, be = sm.linalg.ir_solve(Z, LU32, lu32.pivots, b, tol=1e-15, maxit=40, method="fgmres", znorm=zf)
        assert be[0] <= 1e-15
        lu = sm.linalg.DenseLuSolver()
        lu.factorize(Z)
        x64 = lu.solve(b.clone())
        ex = float(torch.linalg.norm(xm - x64) / torch.linalg.norm(x64))
        th, dirs = cut_directions()
        ph = sm.physical_parameters(o)
        k0 = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)

        def far(x):
            xs = torch.zeros(4 * mesh.ne, dtype=torch.complex128, device="cuda")
            xs[2 * mesh.ne:3 * mesh.ne], xs[3 * mesh.ne:] = x[:mesh.ne], x[mesh.ne:]
            return ad.farfield(mesh, o, xs, dirs).cpu().numpy()
        Fm, F64 = far(xm), far(x64)
        ef = np.linalg.norm(Fm - F64) / np.linalg.norm(F64)
        print(f"  mixed vs fp64: x {ex:.2e}, far field {ef:.2e}, {its[0]} FGMRES-IR steps", flush=True)
        assert ex <= 1e-8 and ef <= 1e-8
        m = len(th)
        sE, sH = 4 * math.pi * np.abs(Fm[:m, 0]) ** 2, 4 * math.pi * np.abs(Fm[m:, 1]) ** 2
        mE, mH = mie.bistatic_rcs(5.0, th, k=k0, m=2.0)
        errs.append(max(np.abs(sE - mE).max() / mE.max(), np.abs(sH - mH).max() / mH.max()))
        print(f"  RCS vs Mie: {errs[-1]:.2e} of peak", flush=True)
        del Z, LU32, lu, lu32
        torch.cuda.empty_cache()
    assert errs[0] < 0.05 and errs[1] < 0.015, errs
    assert errs[0] / errs[1] > 2.5, errs


def test_pmchwt_many_waves_and_matrix_free_gmres(sm):
    """256 plane waves on the Fibonacci sphere (alternating theta / phi
    polarisation) on a level-3 dielectric sphere (eps_r = 4, k a = 5, rule
    7): blocked refinement of all 256 columns around the FP32-MFMA LU against
    the FP64 LU, and restarted GMRES(50) to 1e-8 with the matrix-free
    operator (quadrature recomputed every product) against the LU solution."""
    mesh, o = _pmchwt_sphere(sm, 3, rule=7)
    n = 2 * mesh.ne
    waves = sm.fibonacci_plane_waves(o, 256)
    print(f"pmchwt 256 waves level=3 N={n}", flush=True)
    B = sm.ops.rhs_plane_waves(mesh, o, waves)
    Z = sm.ops.assemble_pmchwt(mesh, o)
    LU32, zf = sm.linalg.to_precision(Z, 32, fro=True)
    lu32 = sm.linalg.DenseLuSolver(variant="f16x3")
    lu32.factorize(LU32)
    Xm, rr, its, be = sm.linalg.ir_solve(Z, LU32, lu32.pivots, B, tol=1e-15, maxit=40, znorm=zf)
    assert be.max() <= 1e-15, be.max()
    lu = sm.linalg.DenseLuSolver()
    Zf = Z.clone()
    lu.factorize(Zf)
    X64 = B.clone()
    lu.solve(X64)
    err = (torch.linalg.norm(Xm - X64, dim=0) / torch.linalg.norm(X64, dim=0)).max().item()
    print(f"  256 columns: max rel err vs fp64 {err:.2e}, sweeps {its.max()}", flush=True)
    assert err <= 1e-8

    def matvec(x, y):
        sm.ops.pmchwt_matvec(mesh, o, x, out=y)
    for c in (0, 1):
        x = torch.zeros(n, dtype=torch.complex128, device="cuda")
        r = sm.linalg.GMRESSolver(n, matvec).solve(B[:, c].contiguous(), x, m=50, maxnit=40, tol=1e-8)
        e = float(torch.linalg.norm(x - X64[:, c]) / torch.linalg.norm(X64[:, c]))
        print(f"  matrix-free GMRES(50) column {c}: {r.iterations} steps, converged {r.converged}, "
              f"rel err vs LU {e:.2e}", flush=True)
        assert r.converged and e < 1e-4
