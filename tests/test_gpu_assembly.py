"""GPU parity of the assembly kernels against the oracle (own CPU FP64
restatement of the reference, itself pinned against the reference build)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sm():
    import sie_mom_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return sie_mom_amd


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def gstc_opts(sm, k0=4.0, rule=7, basis="face", eps_in=1.0):
    o = sm.GstcOptions(frequency_hz=sm.frequency_for_k0(k0), quadrature_rule=rule, basis=basis)
    o.interior = sm.Medium(eps_in, 1.0)
    return o


@pytest.mark.parametrize("rule", [5, 7, 17])
@pytest.mark.parametrize("eps_in", [1.0, 4.0])
def test_pair_blocks(sm, rule, eps_in):
    mesh = sm.Mesh.icosphere(1.0, 1)
    o = gstc_opts(sm, rule=rule, eps_in=eps_in)
    p = sm.calderon_params(o)
    print(f"pair blocks rule={rule} eps_in={eps_in} nf={mesh.nf}", flush=True)
    got = sm.ops.calderon_pair_blocks(mesh, p).cpu().numpy()
    ref = ol.oracle_pair_blocks_all(mesh, p)
    if eps_in == 1.0:  # one wavenumber: TL(k2) == TL(k1)
        ref = ref.copy()
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    # per-block relative error on the blocks with non-negligible norm
    num = np.linalg.norm((got - ref).reshape(mesh.nf, mesh.nf, -1), axis=2)
    den = np.linalg.norm(ref.reshape(mesh.nf, mesh.nf, -1), axis=2)
    assert np.max(num / den) < 1e-10


@pytest.mark.parametrize("basis", ["face", "spherical"])
@pytest.mark.parametrize("level,rule", [(1, 7), (2, 5)])
def test_gstc_matrix(sm, basis, level, rule):
    mesh = sm.Mesh.icosphere(1.0, level)
    o = gstc_opts(sm, rule=rule, basis=basis)
    t = ol.random_gstc_tensors(mesh.nf, seed=level)
    print(f"gstc matrix level={level} rule={rule} basis={basis}", flush=True)
    Z = sm.ops.assemble_gstc(mesh, o, torch.from_numpy(t)).cpu().numpy()
    Zo = ol.oracle_gstc(mesh, sm.calderon_params(o), t, False, {"face": 0, "spherical": 1}[basis])
    assert rel(Z, Zo) < 1e-12


def test_gstc_uniform_and_parts(sm):
    mesh = sm.Mesh.icosphere(1.0, 1)
    o = gstc_opts(sm, rule=7, eps_in=2.0)
    o.tensors_are_uniform = True
    t = ol.random_gstc_tensors(1, seed=3)
    p = sm.calderon_params(o)
    for cal, mass in ((True, False), (False, True), (True, True)):
        Z = sm.ops.assemble_gstc(mesh, o, torch.from_numpy(t), calderon=cal, mass=mass).cpu().numpy()
        Zo = ol.oracle_gstc(mesh, p, t, True, 0, cal, mass)
        assert rel(Z, Zo) < 1e-12, (cal, mass)


def test_pec_matrix(sm):
    mesh = sm.Mesh.icosphere(1.0, 2)
    o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(3.0, sm.physics.CODATA2018_EPS0,
                                                       sm.physics.CODATA2018_MU0), quadrature_rule=5)
    Z = sm.ops.assemble_pec(mesh, o).cpu().numpy()
    Zo = ol.oracle_pec(mesh, sm.calderon_params(o))
    assert rel(Z, Zo) < 1e-12


def test_pmchwt_matrix(sm):
    mesh = sm.Mesh.icosphere(1.0, 1)
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(5.0, sm.physics.CODATA2018_EPS0,
                                                              sm.physics.CODATA2018_MU0), quadrature_rule=7)
    o.interior = sm.Medium(4.0, 1.0)
    Z = sm.ops.assemble_pmchwt(mesh, o).cpu().numpy()
    Zo = ol.oracle_pmchwt(mesh, sm.calderon_params(o))
    assert rel(Z, Zo) < 1e-12


def test_pmchwt_matrix_per_wavenumber(sm, monkeypatch):
    """T-PMCHWT with two regions: one symmetric pass per wavenumber (the
    default, the second pass adding its terms) against the oracle and against
    the non-symmetric chunks of both wavenumbers together (forced by a scratch
    cap below one wavenumber's scratch)."""
    mesh = sm.Mesh.icosphere(1.0, 3)
    o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(5.0), quadrature_rule=7)
    o.interior = sm.Medium(4.0, 1.0)
    print(f"pmchwt per-wavenumber passes, F={mesh.nf}", flush=True)
    Zw = sm.ops.assemble_pmchwt(mesh, o).cpu().numpy()
    one = mesh.nf * mesh.nf * 18 * 16  # bytes of one wavenumber's scratch
    monkeypatch.setenv("SIE_SYM_SCRATCH_BYTES", str(one // 2))
    Zc = sm.ops.assemble_pmchwt(mesh, o).cpu().numpy()
    monkeypatch.delenv("SIE_SYM_SCRATCH_BYTES")
    assert rel(Zw, Zc) < 1e-14
    Zo = ol.oracle_pmchwt(mesh, sm.calderon_params(o))
    assert rel(Zw, Zo) < 1e-12


@pytest.mark.parametrize("kind", ["gstc", "pmchwt", "efie"])
def test_rhs(sm, kind):
    mesh = sm.Mesh.icosphere(1.0, 2)
    if kind == "gstc":
        o = gstc_opts(sm, rule=7, eps_in=2.0)
        waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior"), sm.plane_wave(o, (0, 1, 0), (1, 0, 0), "interior")]
        k = 0
    elif kind == "pmchwt":
        o = sm.DielectricOptions(frequency_hz=sm.frequency_for_k0(5.0), quadrature_rule=7)
        o.interior = sm.Medium(4.0, 1.0)
        waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior"), sm.plane_wave(o, (0, 0, 1), (0, 1, 0), "interior")]
        k = 1
    else:
        o = sm.PecOptions(frequency_hz=sm.frequency_for_k0(3.0), quadrature_rule=5)
        waves = [sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")]
        k = 2
    B = sm.ops.rhs_plane_waves(mesh, o, waves).cpu().numpy()
    for c, w in enumerate(waves):
        ref = ol.oracle_rhs(mesh, k, o.quadrature_rule, w, o.vacuum_permeability, o.vacuum_permittivity,
                            o.vacuum_permeability, o.vacuum_permittivity)
        assert rel(B[:, c], ref) < 1e-13


def test_assembly_overwrites_whole_matrix(sm):
    """The matrix is written in full (zero-fill + gather) even when the output
    buffer holds garbage, including past the first dispatch's reach."""
    mesh = sm.Mesh.icosphere(1.0, 3)
    o = gstc_opts(sm, rule=1)
    t = torch.from_numpy(ol.random_gstc_tensors(mesh.nf, seed=4))
    Z1 = sm.ops.assemble_gstc(mesh, o, t)
    Z2 = sm.ops.colmajor(4 * mesh.ne)
    Z2.fill_(complex(float("nan"), float("nan")))
    sm.ops.assemble_gstc(mesh, o, t, out=Z2)
    assert torch.equal(Z1, Z2)
