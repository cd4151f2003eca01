"""The C++ interface (include/sie_mom/*.hpp) through a plain C++17 caller,
tests/cpp/sie_cpp_check.cpp: host-side validation here, GPU solves against
the oracle on the GPU."""
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHK = os.path.join(ROOT, "tests", "cpp", "sie_cpp_check")


def _binary():
    if not os.path.exists(CHK):
        pytest.fail(f"{CHK} is missing: build it with sie-gstc-2026_amd/build.sh")
    return CHK


def test_reference_caller_compiles():
    """tests/cpp/contract_check.cpp, a caller written against the reference's
    include paths, namespaces, global host types and signatures, compiles
    against include/ with a plain C++17 compiler."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-D__HIP_PLATFORM_AMD__",
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
                        os.path.join(ROOT, "tests", "cpp", "contract_check.cpp")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr


def test_cpp_validation_and_formats(tmp_path):
    r = subprocess.run([_binary(), "validate", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "validate: 0 failure(s)" in r.stdout


def _read(path):
    return np.fromfile(path, dtype=np.complex128)


@pytest.mark.gpu
def test_cpp_solvers_match_oracle(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import sie_mom_amd as sm
    mesh = sm.Mesh.icosphere(1.0, 1)
    k0, rule = 2.0, 5
    f = sm.frequency_for_k0(k0)
    t = ol.random_gstc_tensors(mesh.nf, seed=11)
    t.astype(np.complex128).tofile(tmp_path / "tensors.bin")
    print(f"C++ solves: F={mesh.nf} rule={rule} f={f!r}", flush=True)
    r = subprocess.run([_binary(), "solve", str(tmp_path), repr(f), str(rule)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # field evaluation of the C++ PEC currents against the oracle
    eh = _read(tmp_path / "pec_fields.bin").reshape(2, 5, 3)
    Eo, Ho = _pec_fields_oracle(tmp_path, sm, f, rule, mesh)
    assert np.abs(eh[0] - Eo).max() < 1e-12 * np.abs(Eo).max()
    assert np.abs(eh[1] - Ho).max() < 1e-12 * np.abs(Ho).max()
    ne = mesh.ne

    def rel(a, b):
        return np.linalg.norm(a - b) / np.linalg.norm(b)

    # GSTC: oracle matrix and RHS, dense solve, M blocks times eta0
    o = sm.GstcOptions(frequency_hz=f, quadrature_rule=rule)
    ph = sm.physical_parameters(o)
    Z = ol.oracle_gstc(mesh, sm.calderon_params(o), t, False, 0)
    ext = sm.plane_wave(o, (1, 0, 0), (0, 0, 1), "exterior")
    inn = sm.plane_wave(o, (0, 0, 2), (0, 1, 0), "interior")
    args = (o.vacuum_permeability, o.vacuum_permittivity, o.vacuum_permeability, o.vacuum_permittivity)
    b_ext = ol.oracle_rhs(mesh, 0, rule, ext, *args)
    b_int = ol.oracle_rhs(mesh, 0, rule, inn, *args)

    def phys(x):
        y = x.copy()
        y[ne:2 * ne] *= ph.eta0
        y[3 * ne:] *= ph.eta0
        return y
    x_ext = phys(scipy.linalg.solve(Z, b_ext))
    x_both = phys(scipy.linalg.solve(Z, b_ext + b_int))
    assert rel(_read(tmp_path / "gstc_fp64.bin"), x_ext) < 1e-10
    assert rel(_read(tmp_path / "gstc_mixed.bin"), x_ext) < 1e-10
    assert rel(_read(tmp_path / "gstc_batch0.bin"), x_ext) < 1e-10
    assert rel(_read(tmp_path / "gstc_batch1.bin"), x_both) < 1e-10
    # PEC T-EFIE
    po = sm.PecOptions(frequency_hz=f, quadrature_rule=rule)
    Zp = ol.oracle_pec(mesh, sm.calderon_params(po))
    wp = sm.plane_wave(po, (1, 0, 0), (0, 0, 1), "exterior")
    bp = ol.oracle_rhs(mesh, 2, rule, wp, po.vacuum_permeability, po.vacuum_permittivity, po.vacuum_permeability,
                       po.vacuum_permittivity)
    assert rel(_read(tmp_path / "pec.bin"), scipy.linalg.solve(Zp, bp)) < 1e-10
    # T-PMCHWT (interior eps_r = 4), M times eta0
    do = sm.DielectricOptions(frequency_hz=f, quadrature_rule=rule)
    do.interior = sm.Medium(4.0, 1.0)
    Zd = ol.oracle_pmchwt(mesh, sm.calderon_params(do))
    wd = sm.plane_wave(do, (1, 0, 0), (0, 0, 1), "exterior")
    bd = ol.oracle_rhs(mesh, 1, rule, wd, do.vacuum_permeability, do.vacuum_permittivity, do.vacuum_permeability,
                       do.vacuum_permittivity)
    xd = scipy.linalg.solve(Zd, bd)
    xd[ne:] *= np.sqrt(do.vacuum_permeability / do.vacuum_permittivity)
    assert rel(_read(tmp_path / "pmchwt.bin"), xd) < 1e-10
    assert "solve: 0 failure(s)" in r.stdout


def _pec_fields_oracle(tmp_path, sm, f, rule, mesh):
    """Oracle fields of the C++ PEC currents at the points the program used."""
    J = _read(tmp_path / "pec.bin")
    pts = np.array([[0, 0, 3], [1.5, -0.5, 0.25], [0.2, 0.1, -0.3], [0, 0, 0.6], [0.5, 0.5, 0.55]], float)
    eps, mu = 8.8541878128e-12, 1.25663706212e-6  # PecOptions defaults (CODATA 2018)
    om = 2 * np.pi * f
    k = om * np.sqrt(mu * eps)
    arrays = (mesh.vertices, mesh.faces, mesh.edges, mesh.face_edges)
    return ol.fields(ol.ORACLE_SO, "oracle_", arrays, rule, J, np.zeros_like(J), pts, om, eps, mu, k)
