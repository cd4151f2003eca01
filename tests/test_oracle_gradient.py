"""Pins the CPU FP64 gradient of the GSTC cone objective (oracle_lib.
oracle_gradient: the oracle's assembly, a dense solve, the Schelkunoff far
field, the adjoint solve and the per-face record sensitivities read off the
oracle's own mass assembly) by central finite differences of the oracle's
objective, on every parameter. The reference has no gradient; this oracle is
what the GPU gradient is checked against. CPU only."""
import math

import numpy as np
import pytest

import oracle_lib as ol


class Mesh:
    def __init__(self, V, F, E, FE):
        self.vertices, self.faces, self.edges, self.face_edges = V, F, E, FE
        self.nf, self.ne = len(F), len(E)


def setup(level, kind, basis, rule=5, k0=4.0, seed=20260101):
    import sie_mom_amd as sm
    from sie_mom_amd import adjoint as ad
    mesh = Mesh(*ol.icosphere(level))
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule,
                       basis={0: "face", 1: "spherical"}[basis])
    w = sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")
    rhs = ol.oracle_rhs(mesh, 0, rule, w, o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability,
                        o.vacuum_permittivity)
    ph = sm.physical_parameters(o)
    k0v = ph.omega * math.sqrt(o.vacuum_permeability * o.vacuum_permittivity)
    scales = ad.susceptibility_scales(o)
    eta = math.sqrt(o.vacuum_permeability / o.vacuum_permittivity)
    dirs, weights = ad.cone_directions()
    params = np.random.default_rng(seed).uniform(-2, 2, size=(mesh.nf, kind)) / k0
    p = sm.calderon_params(o)
    return mesh, p, rhs, k0v, ph.eta0, scales, eta, dirs, weights, params, rule


@pytest.fixture(autouse=True)
def one_thread():
    # many tiny dense solves: thread start-up would dominate
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1, user_api="blas"):
        yield


@pytest.mark.parametrize("kind,basis", [(3, 0), (10, 0), (3, 1)])
def test_oracle_gradient_vs_central_fd(kind, basis):
    mesh, p, rhs, k0, eta0, scales, eta, dirs, w, params, rule = setup(0, kind, basis)
    J, grad, _, _ = ol.oracle_gradient(mesh, p, basis, params, kind, scales, eta, rhs, rule, k0, eta0, dirs, w)

    def Jof(pp):
        rec = ol.oracle_records(pp, kind, *scales, eta)
        return ol.oracle_cone_objective(mesh, p, basis, rec, rhs, rule, k0, eta0, dirs, w, adjoint=False)[0]
    assert abs(Jof(params) - J) <= 1e-14 * abs(J)
    h = 1e-6
    fd = np.zeros_like(grad)
    for f in range(mesh.nf):
        for q in range(kind):
            pp, pm = params.copy(), params.copy()
            pp[f, q] += h
            pm[f, q] -= h
            fd[f, q] = (Jof(pp) - Jof(pm)) / (2 * h)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    print(f"kind {kind} basis {basis}: J {J:.6e}, max |grad - fd| / max |fd| = {err:.2e}")
    assert err < 1e-6
    # a plausible slip, the adjoint's conjugation, is caught
    assert np.abs(grad).max() > 0 and np.linalg.norm(grad - fd) < 1e-5 * np.linalg.norm(fd)
