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
