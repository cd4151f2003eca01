#!/usr/bin/env python3
"""Writes tests/golden/reference_vectors.npz: input/output vectors produced by
the reference's own host numerics (oracle/_ref/libsie_ref.so, built from the
reference tree by `make -C oracle ref`). Run in a container where the
reference is present; the fixture then pins the oracle everywhere else.

Cases: Dunavant rules 5/7/17 (float32 nodes and weights); icosphere level 1;
per-pair TL/TK blocks (regular and regularised) and singular parts for a
coincident, an edge-adjacent, a vertex-adjacent and a separated pair; Wilton
potentials at on-plane, off-plane and edge-line points; the 8-tensor mass
batch; plane-wave excitation testing; full eta0-scaled GSTC (level 0, both
bases), PEC T-EFIE (level 1) and T-PMCHWT (level 0) matrices; the GSTC and
T-PMCHWT (exterior and interior source) plane-wave right-hand sides; E and H of
random currents on the level-1 sphere at far, near-surface and interior points.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as ol  # noqa: E402


def main():
    r = ol.ref()
    out = {}
    for rule in (5, 7, 17):
        xy, w = r.rule(rule)
        out[f"rule{rule}_xy"] = xy
        out[f"rule{rule}_w"] = w
    V, F, E, FE = ol.ref_icosphere(1)
    out.update(ico1_V=V, ico1_F=F, ico1_E=E, ico1_FE=FE)
    # classify pairs of level-1 faces
    sets = [set(f) for f in F]
    pairs = {"self": (3, 3)}
    for n in range(len(F)):
        s = len(sets[3] & sets[n])
        if s == 2 and "edge" not in pairs:
            pairs["edge"] = (3, n)
        if s == 1 and "vertex" not in pairs:
            pairs["vertex"] = (3, n)
    pairs["far"] = (3, int(np.argmin(V[F][:, :, 2].mean(axis=1) * np.sign(V[F[3]][:, 2].mean()))))
    k1, k2 = 4.0, 9.0
    for name, (m, n) in pairs.items():
        tt, st = V[F[m]], V[F[n]]
        for rule in (5, 7):
            for sing in (0, 1):
                out[f"pair_{name}_r{rule}_s{sing}"] = r.pair_regular(tt, F[m], st, F[n], rule, k1, k2, sing)
            out[f"sing_{name}_r{rule}"] = r.pair_singular(tt, F[m], st, F[n], rule, k1, k2)
        out[f"pairidx_{name}"] = np.array([m, n])
    tri = V[F[3]]
    c = tri.mean(axis=0)
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    n /= np.linalg.norm(n)
    pts = np.array([c, c + 0.05 * n, tri[0], 0.5 * (tri[0] + tri[1]), c + 0.3 * (tri[1] - c) - 0.02 * n,
                    c + 2.0 * (tri[2] - c), tri[0] + 1.5 * (tri[1] - tri[0])])
    out["wilton_tri"] = tri
    out["wilton_pts"] = pts
    out["wilton_out"] = np.array([r.wilton(p, tri) for p in pts])
    rng = np.random.default_rng(7)
    T = rng.normal(size=(8, 3, 3)) + 1j * rng.normal(size=(8, 3, 3))
    out["mass_T"] = T
    out["mass_out"] = r.mass_batch(tri, F[3], 7, T)
    e0 = np.array([1.0, 0.0, 0.0])
    kh = np.array([0.0, 0.0, 1.0])
    ve, vh = r.excitation_planewave(tri, F[3], 7, e0, kh, 4.0, 376.73)
    out["exc_VE"] = ve
    out["exc_VH"] = vh
    # full matrices
    V0, F0, E0, FE0 = ol.ref_icosphere(0)
    out.update(ico0_V=V0, ico0_F=F0, ico0_E=E0, ico0_FE=FE0)
    eps0, mu0 = 8.8541878188e-12, 1.25663706127e-6
    c0 = 1 / np.sqrt(mu0 * eps0)
    f = 4.0 * c0 / (2 * np.pi)
    om = 2 * np.pi * f
    phys = dict(k1=om * np.sqrt(mu0 * eps0 * 2.0), k2=om * np.sqrt(mu0 * eps0), mu_r1=1, mu_r2=1, eps_r1=2.0,
                eps_r2=1, mu0=mu0, eps0=eps0, omega=om, thr=0.1 * c0 / f)
    out["gstc_phys"] = np.array([np.float32(phys[k]) for k in ("k1", "k2", "mu_r1", "mu_r2", "eps_r1", "eps_r2", "mu0",
                                                                "eps0", "omega", "thr")], np.float32)
    tens = ol.random_gstc_tensors(len(F0), seed=11)
    out["gstc_tensors"] = tens
    for basis in (0, 1):
        out[f"gstc_Z_b{basis}"] = ol.assemble_gstc(ol.REF_SO, "ref_", (V0, F0, E0, FE0), 7, phys, tens, False,
                                                   basis)
    import ctypes
    lib = ctypes.CDLL(ol.REF_SO)
    P = ctypes.c_void_p
    ptr = lambda a: a.ctypes.data_as(P)  # noqa: E731
    f32 = np.float32
    # PEC T-EFIE, level 1, rule 5, CODATA-2018 constants
    e18, m18 = 8.8541878128e-12, 1.25663706212e-6
    c18 = 1 / np.sqrt(m18 * e18)
    fp = 3.0 * c18 / (2 * np.pi)
    omp = 2 * np.pi * fp
    k2p = omp * np.sqrt(m18 * e18)
    lam = 2 * np.pi / (omp * np.sqrt(m18 * e18))
    Zp = np.zeros((len(E), len(E)), complex, order="F")
    fn = lib.ref_assemble_pec
    fn.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 5 + [P]
    fn(ptr(V), ptr(F), ptr(FE), len(F), len(E), 5, f32(k2p), f32(1.0), f32(m18), f32(omp), f32(0.1 * lam), ptr(Zp))
    out["pec_phys"] = np.array([k2p, 1.0, m18, omp, 0.1 * lam], np.float32)
    out["pec_Z"] = Zp
    # T-PMCHWT, level 0, rule 7, eps_r1 = 4
    fd = 5.0 * c18 / (2 * np.pi)
    omd = 2 * np.pi * fd
    k1d, k2d = omd * np.sqrt(m18 * e18 * 4.0), omd * np.sqrt(m18 * e18)
    lamd = 2 * np.pi / (omd * np.sqrt(m18 * e18))
    Zd = np.zeros((2 * len(E0), 2 * len(E0)), complex, order="F")
    fn = lib.ref_assemble_pmchwt
    fn.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + [P]
    pd = [k1d, k2d, 1.0, 1.0, 4.0, 1.0, m18, e18, omd, 0.1 * lamd]
    fn(ptr(V0), ptr(F0), ptr(FE0), len(F0), len(E0), 7, *[f32(x) for x in pd], ptr(Zd))
    out["pmchwt_phys"] = np.array(pd, np.float32)
    out["pmchwt_Z"] = Zd
    # GSTC RHS, level 1, exterior x-polarised +z plane wave
    rhs = np.zeros(4 * len(E), complex)
    fn = lib.ref_rhs_gstc_planewave
    fn.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double,
                   ctypes.c_int, ctypes.c_float, ctypes.c_float, P]
    kk = om * np.sqrt(mu0 * eps0)
    et = np.sqrt(mu0 / eps0)
    fn(ptr(V), ptr(F), ptr(FE), len(F), len(E), 7, ptr(e0), ptr(kh), kk, et, 2, f32(mu0), f32(eps0), ptr(rhs))
    out["rhs_gstc"] = rhs
    out["rhs_gstc_k_eta"] = np.array([kk, et])
    # T-PMCHWT RHS, level 1, x-polarised +z plane wave in the exterior (region 2)
    # and, with the interior medium's k and eta (eps_r = 4), in the interior (region 1)
    fn = lib.ref_rhs_pmchwt_planewave
    fn.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, ctypes.c_double, ctypes.c_double,
                   ctypes.c_int, ctypes.c_double, ctypes.c_double, P]
    for region, (kk_, et_) in ((2, (k2d, np.sqrt(m18 / e18))), (1, (k1d, np.sqrt(m18 / e18) / 2.0))):
        rp = np.zeros(2 * len(E), complex)
        fn(ptr(V), ptr(F), ptr(FE), len(F), len(E), 7, ptr(e0), ptr(kh), kk_, et_, region, m18, e18, ptr(rp))
        out[f"rhs_pmchwt_r{region}"] = rp
        out[f"rhs_pmchwt_r{region}_k_eta"] = np.array([kk_, et_])
    out["rhs_pmchwt_vac"] = np.array([m18, e18])
    # fields of random currents on the level-1 sphere at far, near-surface and interior points
    rng = np.random.default_rng(21)
    Jf = rng.normal(size=len(E)) + 1j * rng.normal(size=len(E))
    Mf = rng.normal(size=len(E)) + 1j * rng.normal(size=len(E))
    obs = ol.field_points(V, F, rng)
    omf = 4.0 / np.sqrt(mu0 * eps0)
    kf = omf * np.sqrt(mu0 * eps0)
    Ef, Hf = ol.fields(ol.REF_SO, "ref_", (V, F, E, FE), 7, Jf, Mf, obs, omf, eps0, mu0, kf)
    out.update(fields_J=Jf, fields_M=Mf, fields_obs=obs, fields_phys=np.array([omf, eps0, mu0, kf]),
               fields_E=Ef, fields_H=Hf)
    np.savez_compressed(os.path.join(HERE, "reference_vectors.npz"), **out)
    print("wrote", len(out), "arrays")


if __name__ == "__main__":
    main()
