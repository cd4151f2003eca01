"""ctypes access to the oracle (own CPU restatement) and, where it was built,
oracle/_ref (the reference's host numerics). Test infrastructure only."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle_sie.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsie_ref.so")

_P = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(_P)


class Lib:
    """Uniform wrapper: prefix 'oracle_' or 'ref_'."""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix

    def f(self, name):
        return getattr(self.lib, self.prefix + name)

    def pair_regular(self, tt, ti, st, si, rule, k1, k2, singular):
        out = np.zeros(72)
        fn = self.f("pair_regular")
        fn.argtypes = [_P, _P, _P, _P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, _P]
        fn(_ptr(np.ascontiguousarray(tt, float)), _ptr(np.ascontiguousarray(ti, np.int32)),
           _ptr(np.ascontiguousarray(st, float)), _ptr(np.ascontiguousarray(si, np.int32)),
           rule, k1, k2, int(singular), _ptr(out))
        return out.view(complex).reshape(4, 3, 3)

    def pair_singular(self, tt, ti, st, si, rule, k1, k2):
        out = np.zeros(72)
        fn = self.f("pair_singular")
        fn.argtypes = [_P, _P, _P, _P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _P]
        fn(_ptr(np.ascontiguousarray(tt, float)), _ptr(np.ascontiguousarray(ti, np.int32)),
           _ptr(np.ascontiguousarray(st, float)), _ptr(np.ascontiguousarray(si, np.int32)),
           rule, k1, k2, _ptr(out))
        return out.view(complex).reshape(4, 3, 3)

    def wilton(self, r, tri):
        out = np.zeros(12)
        fn = self.f("wilton")
        fn.argtypes = [_P, _P, _P]
        fn(_ptr(np.ascontiguousarray(r, float)), _ptr(np.ascontiguousarray(tri, float)), _ptr(out))
        return out

    def mass_batch(self, tri, idx, rule, T):
        out = np.zeros(144)
        fn = self.f("mass_batch")
        fn.argtypes = [_P, _P, ctypes.c_int, _P, _P]
        T = np.ascontiguousarray(T, complex)
        fn(_ptr(np.ascontiguousarray(tri, float)), _ptr(np.ascontiguousarray(idx, np.int32)), rule,
           _ptr(T), _ptr(out))
        return out.view(complex).reshape(8, 3, 3)

    def excitation_planewave(self, tri, idx, rule, E0, kh, k, eta):
        ve = np.zeros(6)
        vh = np.zeros(6)
        fn = self.f("excitation_planewave")
        fn.argtypes = [_P, _P, ctypes.c_int, _P, _P, ctypes.c_double, ctypes.c_double, _P, _P]
        fn(_ptr(np.ascontiguousarray(tri, float)), _ptr(np.ascontiguousarray(idx, np.int32)), rule,
           _ptr(np.ascontiguousarray(E0, float)), _ptr(np.ascontiguousarray(kh, float)), k, eta,
           _ptr(ve), _ptr(vh))
        return ve.view(complex), vh.view(complex)

    def face_basis(self, tri, basis):
        u = np.zeros(3)
        v = np.zeros(3)
        fn = self.f("face_basis")
        fn.argtypes = [_P, ctypes.c_int, _P, _P]
        fn(_ptr(np.ascontiguousarray(tri, float)), basis, _ptr(u), _ptr(v))
        return u, v

    def rule(self, r):
        if self.prefix == "oracle_":
            n = self.lib.oracle_rule_order(r)
            xy = np.zeros(2 * n, np.float32)
            w = np.zeros(n, np.float32)
            self.lib.oracle_rule(r, _ptr(xy), _ptr(w))
        else:
            n = self.lib.ref_dunavant_order_num(r)
            xy = np.zeros(2 * n, np.float32)
            w = np.zeros(n, np.float32)
            self.lib.ref_make_quadrature(r, _ptr(xy), _ptr(w))
        return xy.reshape(-1, 2), w


def oracle():
    return Lib(ORACLE_SO, "oracle_")


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    return Lib(REF_SO, "ref_")


def icosphere(level, radius=1.0):
    """Oracle icosphere (restatement of SphericalSurface): V, F, E, FE."""
    lib = ctypes.CDLL(ORACLE_SO)
    counts = np.zeros(3, np.int32)
    lib.oracle_icosphere.argtypes = [ctypes.c_float, ctypes.c_int, _P, _P, _P, _P, _P]
    lib.oracle_icosphere(radius, level, None, None, None, None, _ptr(counts))
    nv, nf, ne = (int(c) for c in counts)
    V = np.zeros((nv, 3))
    F = np.zeros((nf, 3), np.int32)
    E = np.zeros((ne, 2), np.int32)
    FE = np.zeros((nf, 3), np.int32)
    lib.oracle_icosphere(radius, level, _ptr(V), _ptr(F), _ptr(E), _ptr(FE), _ptr(counts))
    return V, F, E, FE


def ref_icosphere(level, radius=1.0):
    lib = ctypes.CDLL(REF_SO)
    nv, nf, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    lib.ref_icosphere_counts(level, ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(ne))
    V = np.zeros((nv.value, 3))
    F = np.zeros((nf.value, 3), np.int32)
    E = np.zeros((ne.value, 2), np.int32)
    FE = np.zeros((nf.value, 3), np.int32)
    lib.ref_icosphere.argtypes = [ctypes.c_float, ctypes.c_int, _P, _P, _P, _P]
    lib.ref_icosphere(radius, level, _ptr(V), _ptr(F), _ptr(E), _ptr(FE))
    return V, F, E, FE


F32 = np.float32


def assemble_gstc(lib_path, prefix, mesh, rule, phys, tensors, uniform, basis, with_calderon=True,
                  with_mass=True):
    """phys: dict of float32-castable k1,k2,mu_r1,mu_r2,eps_r1,eps_r2,mu0,eps0,omega,thr."""
    V, F, E, FE = mesh
    nf, ne = len(F), len(E)
    dim = 4 * ne
    Z = np.zeros((dim, dim), complex, order="F")
    lib = ctypes.CDLL(lib_path)
    fl = [F32(phys[k]) for k in ("k1", "k2", "mu_r1", "mu_r2", "eps_r1", "eps_r2", "mu0", "eps0", "omega", "thr")]
    t = np.ascontiguousarray(tensors, complex)
    if prefix == "oracle_":
        fn = lib.oracle_assemble_gstc
        fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + \
            [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        fn(_ptr(np.ascontiguousarray(V)), _ptr(np.ascontiguousarray(F)), _ptr(np.ascontiguousarray(FE)),
           nf, ne, rule, *fl, _ptr(t), int(uniform), basis, int(with_calderon), int(with_mass), _ptr(Z))
    else:
        fn = lib.ref_assemble_gstc
        fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + \
            [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
        fn(_ptr(np.ascontiguousarray(V)), _ptr(np.ascontiguousarray(F)), _ptr(np.ascontiguousarray(FE)),
           nf, ne, rule, *fl, _ptr(t), int(uniform), basis, int(with_mass), _ptr(Z))
    return Z


def _fl(p, names):
    return [F32(getattr(p, n)) for n in names]


_PNAMES = ("k1", "k2", "mu_r1", "mu_r2", "eps_r1", "eps_r2", "mu0", "eps0", "omega", "threshold")


def oracle_gstc(mesh, p, tensors, uniform, basis, with_calderon=True, with_mass=True, scaled=True):
    """Oracle GSTC matrix from a CalderonParams struct (float fields);
    scaled=False: the unscaled operator (eta0 = 1)."""
    lib = ctypes.CDLL(ORACLE_SO)
    ne, nf = mesh.ne, mesh.nf
    dim = 4 * ne
    Z = np.zeros((dim, dim), complex, order="F")
    fn = lib.oracle_assemble_gstc_ex
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + \
        [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P]
    t = np.ascontiguousarray(tensors, complex)
    fn(_ptr(mesh.vertices), _ptr(mesh.faces), _ptr(mesh.face_edges), nf, ne, p.rule, *_fl(p, _PNAMES),
       _ptr(t), int(uniform), basis, int(with_calderon), int(with_mass), int(scaled), _ptr(Z))
    return Z


def oracle_gstc_rows(mesh, p, tensors, uniform, basis, rows):
    """Rows `rows` (indices into [0, 4E)) of the oracle's eta0-scaled GSTC
    matrix, len(rows) x 4E: only the test faces of those rows' edges are
    integrated, against every source face."""
    lib = ctypes.CDLL(ORACLE_SO)
    rows = np.ascontiguousarray(rows, np.int32)
    out = np.zeros((len(rows), 4 * mesh.ne), complex)
    fn = lib.oracle_gstc_rows
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + \
        [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P]
    t = np.ascontiguousarray(tensors, complex)
    fn(_ptr(np.ascontiguousarray(mesh.vertices, float)), _ptr(np.ascontiguousarray(mesh.faces, np.int32)),
       _ptr(np.ascontiguousarray(mesh.face_edges, np.int32)), mesh.nf, mesh.ne, p.rule, *_fl(p, _PNAMES),
       _ptr(t), int(uniform), basis, len(rows), _ptr(rows), _ptr(out))
    return out


def oracle_pec(mesh, p):
    lib = ctypes.CDLL(ORACLE_SO)
    dim = mesh.ne
    Z = np.zeros((dim, dim), complex, order="F")
    fn = lib.oracle_assemble_pec
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 5 + [_P]
    fn(_ptr(mesh.vertices), _ptr(mesh.faces), _ptr(mesh.face_edges), mesh.nf, mesh.ne, p.rule,
       F32(p.k2), F32(p.mu_r2), F32(p.mu0), F32(p.omega), F32(p.threshold), _ptr(Z))
    return Z


def oracle_pmchwt(mesh, p):
    lib = ctypes.CDLL(ORACLE_SO)
    dim = 2 * mesh.ne
    Z = np.zeros((dim, dim), complex, order="F")
    fn = lib.oracle_assemble_pmchwt
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int] + [ctypes.c_float] * 10 + [_P]
    fn(_ptr(mesh.vertices), _ptr(mesh.faces), _ptr(mesh.face_edges), mesh.nf, mesh.ne, p.rule,
       *_fl(p, _PNAMES), _ptr(Z))
    return Z


def oracle_pair_blocks_all(mesh, p):
    lib = ctypes.CDLL(ORACLE_SO)
    out = np.zeros((mesh.nf, mesh.nf, 4, 3, 3), complex)
    fn = lib.oracle_pair_blocks_all
    fn.argtypes = [_P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]
    fn(_ptr(mesh.vertices), _ptr(mesh.faces), mesh.nf, p.rule, F32(p.k1), F32(p.k2), F32(p.threshold), _ptr(out))
    return out


def oracle_rhs(mesh, kind, rule, wave, mu0, eps0, vac_mu, vac_eps):
    """kind 0 GSTC, 1 PMCHWT, 2 EFIE; wave: PlaneWaveC-like (E0, khat, k, eta, region)."""
    lib = ctypes.CDLL(ORACLE_SO)
    dim = {0: 4, 1: 2, 2: 1}[kind] * mesh.ne
    rhs = np.zeros(dim, complex)
    fn = lib.oracle_rhs_planewave
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.c_double,
                   ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                   ctypes.c_double, ctypes.c_double, _P]
    E0 = np.array(list(wave.E0), float)
    kh = np.array(list(wave.khat), float)
    fn(_ptr(mesh.vertices), _ptr(mesh.faces), _ptr(mesh.face_edges), mesh.nf, mesh.ne, rule, _ptr(E0),
       _ptr(kh), wave.k, wave.eta, wave.region, kind, F32(mu0), F32(eps0), vac_mu, vac_eps, _ptr(rhs))
    return rhs


def random_gstc_tensors(nf, seed=0):
    """Random well-conditioned 4x4 maps and their inverses, 32 complex per face."""
    rng = np.random.default_rng(seed)
    lam = rng.normal(size=(nf, 4, 4)) + 1j * rng.normal(size=(nf, 4, 4)) + 4 * np.eye(4)
    inv = np.linalg.inv(lam)
    return np.concatenate([lam.reshape(nf, 16), inv.reshape(nf, 16)], axis=1).ravel()


# ------------------------------------------------------------------ adjoint
def oracle_records(params, kind, bs, xs, ks, eta):
    """Lambda = [[jB, K], [-K^T, jX]] per face and its inverse (numpy), the
    record layout of src/sie_mom/gstc/tensor_io.py:118-120."""
    nf = len(params)
    L = oracle_lambda(params, kind, bs, xs, ks)
    t = np.array([np.sqrt(eta)] * 2 + [1 / np.sqrt(eta)] * 2)
    bal = L * t[None, :, None] * t[None, None, :]
    inv = np.linalg.inv(bal) * t[None, :, None] * t[None, None, :]
    return np.concatenate([L.reshape(nf, 16), inv.reshape(nf, 16)], axis=1).ravel()


def oracle_lambda(params, kind, bs, xs, ks):
    """The fields-to-currents maps Lambda (nf x 4 x 4) of real parameters:
    kind 3 the omega model (surface_parameters / omega_biisotropic.py:196-219),
    kind 10 the general reciprocal lossless map in the order of the reference's
    _reciprocal_basis (surface_parameters.py:110-127)."""
    nf = len(params)
    L = np.zeros((nf, 4, 4), complex)
    p = np.asarray(params, float)
    if kind == 3:
        L[:, 0, 0] = L[:, 1, 1] = 1j * bs * p[:, 0]
        L[:, 2, 2] = L[:, 3, 3] = 1j * xs * p[:, 1]
        k = ks * p[:, 2]
        L[:, 0, 3], L[:, 1, 2], L[:, 2, 1], L[:, 3, 0] = k, -k, k, -k
    else:
        L[:, 0, 0], L[:, 1, 1] = 1j * bs * p[:, 0], 1j * bs * p[:, 1]
        L[:, 0, 1] = L[:, 1, 0] = 1j * bs * p[:, 2]
        L[:, 2, 2], L[:, 3, 3] = 1j * xs * p[:, 3], 1j * xs * p[:, 4]
        L[:, 2, 3] = L[:, 3, 2] = 1j * xs * p[:, 5]
        K = ks * p[:, 6:10].reshape(nf, 2, 2)
        L[:, 0:2, 2:4] = K
        L[:, 2:4, 0:2] = -np.transpose(K, (0, 2, 1))
    return L


def oracle_farfield_matrix(mesh, rule, k0, eta0, dirs):
    """A (ndir*2 x 4E): [F_theta, F_phi] = A xs for scaled currents xs
    (Schelkunoff, plot_far_field.py:150-219)."""
    xy, w = oracle().rule(rule)
    xy = xy.astype(float)
    w = w.astype(float)
    V, F, FE = mesh.vertices, mesh.faces, mesh.face_edges
    ne = mesh.ne
    nd = len(dirs)
    A = np.zeros((nd, 2, 4 * ne), complex)
    coef = -1j * k0 / (4 * np.pi)
    for f in range(len(F)):
        v = V[F[f]]
        r = (v[0][None, :] * (1 - xy[:, 0:1] - xy[:, 1:2]) + v[1][None, :] * xy[:, 0:1] + v[2][None, :] * xy[:, 1:2])
        ph = np.exp(1j * k0 * (dirs[:, 0:3] @ r.T)) * w[None, :]  # nd x Q
        for le in range(3):
            a, b = F[f, le], F[f, (le + 1) % 3]
            length = np.linalg.norm(v[(le + 1) % 3] - v[le]) * (1.0 if a > b else -1.0)
            rho = r - v[(le + 2) % 3][None, :]          # Q x 3
            n = 0.5 * length * (ph @ rho)                 # nd x 3
            nt = np.einsum("dk,dk->d", n, dirs[:, 3:6])
            npp = np.einsum("dk,dk->d", n, dirs[:, 6:9])
            e = FE[f, le]
            A[:, 0, 2 * ne + e] += coef * eta0 * nt
            A[:, 1, 2 * ne + e] += coef * eta0 * npp
            A[:, 0, 3 * ne + e] += coef * eta0 * npp
            A[:, 1, 3 * ne + e] += -coef * eta0 * nt
    return A.reshape(nd * 2, 4 * ne)


def face_colors(mesh):
    """Greedy colouring of the faces so that two faces sharing an RWG edge
    differ (at most 4 colours: each face has 3 edge neighbours)."""
    FE = np.asarray(mesh.face_edges)
    owners = {}
    for f, es in enumerate(FE):
        for e in es:
            owners.setdefault(int(e), []).append(f)
    col = -np.ones(len(FE), int)
    for f in range(len(FE)):
        used = {col[g] for e in FE[f] for g in owners[int(e)] if g != f}
        col[f] = min(c for c in range(5) if c not in used)
    return col


def oracle_mass_sensitivity(mesh, p, basis, lam, x):
    """S[f, k] = lam^H M(e_{f,k}) x, the derivative of lam^H Z x with respect to
    record entry k (0..31) of face f. The eta0-scaled mass operator is linear
    in each face's record (the map and its inverse enter its Gram blocks
    linearly) and face f's blocks couple only the 12 unknowns of its 3 edges,
    so M(e_{f,k}) is read off the oracle's own mass assembly with unit records:
    one assembly per colour and entry, on faces of that colour only (no two
    share an edge, so their 12 x 12 blocks do not overlap)."""
    nf, ne = mesh.nf, mesh.ne
    FE = np.asarray(mesh.face_edges)
    dofs = np.array([[b * ne + e for b in range(4) for e in FE[f]] for f in range(nf)])
    col = face_colors(mesh)
    S = np.zeros((nf, 32), complex)
    for c in range(col.max() + 1):
        faces = np.flatnonzero(col == c)
        for k in range(32):
            t = np.zeros((nf, 32), complex)
            t[faces, k] = 1.0
            Mk = oracle_gstc(mesh, p, t.ravel(), False, basis, with_calderon=False, with_mass=True)
            for f in faces:
                idx = dofs[f]
                S[f, k] = np.conj(lam[idx]) @ Mk[np.ix_(idx, idx)] @ x[idx]
    return S


def oracle_param_grad(params, kind, scales, eta, S):
    """-2 Re sum_k S[f, k] d rec_{f,k} / d theta_{f,p}: Lambda is linear in
    the parameters (its derivative is the map of a unit parameter vector) and
    d Lambda^-1 = -Lambda^-1 d Lambda Lambda^-1."""
    params = np.asarray(params, float)
    nf = len(params)
    rec = oracle_records(params, kind, *scales, eta).reshape(nf, 2, 4, 4)
    Li = rec[:, 1]
    grad = np.zeros((nf, kind))
    for q in range(kind):
        unit = np.zeros((1, kind))
        unit[0, q] = 1.0
        dL = oracle_lambda(unit, kind, *scales)[0]
        dLi = -Li @ dL @ Li
        drec = np.concatenate([np.broadcast_to(dL.ravel(), (nf, 16)), dLi.reshape(nf, 16)], axis=1)
        grad[:, q] = -2.0 * np.real(np.sum(S * drec, axis=1))
    return grad


def oracle_cone_objective(mesh, p, basis, rec, rhs, rule, k0, eta0, dirs, weights, adjoint=True):
    """J = sum_d w_d |F_d|^2 / (2 eta0) of the scaled GSTC solution (dense
    scipy solve of the oracle matrix) and, with adjoint, lambda = Z^-H A^H u
    with u_d = w_d F_d / (2 eta0). Returns (J, x, F, lambda)."""
    import scipy.linalg
    Z = oracle_gstc(mesh, p, rec, False, basis)
    lu = scipy.linalg.lu_factor(Z)
    x = scipy.linalg.lu_solve(lu, rhs)
    A = oracle_farfield_matrix(mesh, rule, k0, eta0, dirs)
    F = (A @ x).reshape(-1, 2)
    J = float(np.sum(weights * np.sum(np.abs(F) ** 2, axis=1)) / (2 * eta0))
    lam = None
    if adjoint:
        u = (F * (weights / (2 * eta0))[:, None]).ravel()
        lam = scipy.linalg.lu_solve(lu, A.conj().T @ u, trans=2)
    return J, x, F, lam


def oracle_gradient(mesh, p, basis, params, kind, scales, eta_ref, rhs, rule, k0, eta0, dirs, weights):
    """CPU FP64 gradient of the cone objective with respect to the per-face
    real parameters: dJ/dtheta = -2 Re lambda^H (dZ/dtheta) x. Returns
    (J, grad (nf x kind), x, S)."""
    rec = oracle_records(params, kind, *scales, eta_ref)
    J, x, _, lam = oracle_cone_objective(mesh, p, basis, rec, rhs, rule, k0, eta0, dirs, weights)
    S = oracle_mass_sensitivity(mesh, p, basis, lam, x)
    return J, oracle_param_grad(params, kind, scales, eta_ref, S), x, S


# ------------------------------------------------------------------ fields
def fields(so, prefix, mesh, rule, J, M, obs, omega, eps, mu, k):
    """E, H (nobs x 3 complex) radiated by per-edge currents J, M at points
    obs (nobs x 3): the oracle (so = ORACLE_SO, prefix "oracle_") or the
    reference harness (REF_SO, "ref_")."""
    V, F, E, FE = mesh
    lib = ctypes.CDLL(so)
    fn = getattr(lib, prefix + "fields")
    fn.argtypes = [_P, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P, ctypes.c_int] + [ctypes.c_double] * 4 + [_P, _P]
    obs = np.ascontiguousarray(obs, float)
    J = np.ascontiguousarray(J, complex)
    M = np.ascontiguousarray(M, complex)
    Eo = np.zeros((len(obs), 3), complex)
    Ho = np.zeros((len(obs), 3), complex)
    fn(_ptr(np.ascontiguousarray(V, float)), _ptr(np.ascontiguousarray(F, np.int32)),
       _ptr(np.ascontiguousarray(FE, np.int32)), len(F), rule, _ptr(J), _ptr(M), _ptr(obs), len(obs),
       omega, eps, mu, k, _ptr(Eo), _ptr(Ho))
    return Eo, Ho


def field_points(V, F, rng, n_far=6, n_near=6, n_in=4):
    """Observation points: far (|r| = 2.5), just outside and inside a face
    (1% of the radius off its centroid), and interior (|r| = 0.4)."""
    u = rng.normal(size=(n_far, 3))
    far = 2.5 * u / np.linalg.norm(u, axis=1)[:, None]
    fs = rng.choice(len(F), n_near, replace=False)
    c = V[F[fs]].mean(axis=1)
    nrm = c / np.linalg.norm(c, axis=1)[:, None]
    near = c + 0.01 * nrm * np.where(np.arange(n_near) % 2 == 0, 1.0, -1.0)[:, None]
    u = rng.normal(size=(n_in, 3))
    inner = 0.4 * u / np.linalg.norm(u, axis=1)[:, None]
    return np.concatenate([far, near, inner])
