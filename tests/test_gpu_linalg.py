"""GPU dense algebra against numpy/scipy (textbook LAPACK routines)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sm():
    import sie_mom_amd
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return sie_mom_amd


def crandn(rng, *shape):
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def to_dev(a, dtype=torch.complex128):
    """numpy (column-major semantics) -> column-major device tensor."""
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda", dtype)
    return t.T


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "4m"), (32, "3m"), (32, "f16x3")])
@pytest.mark.parametrize("m,n,k", [(128, 128, 16), (300, 200, 70), (513, 65, 256), (1, 1, 1), (1000, 3, 5)])
def test_gemm(sm, prec, variant, m, n, k):
    rng = np.random.default_rng(m + n + k)
    A, B, C = crandn(rng, m, k), crandn(rng, k, n), crandn(rng, m, n)
    dt = torch.complex128 if prec == 64 else torch.complex64
    Cd = to_dev(C, dt)
    print(f"gemm prec={prec} {variant} m={m} n={n} k={k}", flush=True)
    sm.linalg.gemm(to_dev(A, dt), to_dev(B, dt), Cd, alpha=-1.0, beta=1.0, variant=variant)
    ref = C - A @ B
    err = np.abs(Cd.cpu().numpy() - ref).max() / (np.abs(A).max() * np.abs(B).max() * k)
    assert err < (1e-15 if prec == 64 else 2e-7), err


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "4m"), (32, "3m"), (32, "f16x3")])
@pytest.mark.parametrize("alpha,beta", [(0.5, 2.0), (-3.0, 0.25), (2.0, 0.0), (0.0, 1.5)])
@pytest.mark.parametrize("m,n,k", [(130, 70, 16), (300, 200, 70)])
def test_gemm_alpha_beta(sm, prec, variant, alpha, beta, m, n, k):
    # general scalars: C = beta C + alpha A B, including alpha = 0 and beta = 0
    rng = np.random.default_rng(m + n + k + 7)
    A, B, C = crandn(rng, m, k), crandn(rng, k, n), crandn(rng, m, n)
    dt = torch.complex128 if prec == 64 else torch.complex64
    Cd = to_dev(C, dt)
    print(f"gemm prec={prec} {variant} alpha={alpha} beta={beta} m={m} n={n} k={k}", flush=True)
    sm.linalg.gemm(to_dev(A, dt), to_dev(B, dt), Cd, alpha=alpha, beta=beta, variant=variant)
    ref = beta * C + alpha * (A @ B)
    scale = abs(alpha) * np.abs(A).max() * np.abs(B).max() * k + abs(beta) * np.abs(C).max()
    err = np.abs(Cd.cpu().numpy() - ref).max() / scale
    assert err < (1e-15 if prec == 64 else 4e-7), err


@pytest.mark.parametrize("alpha,beta", [(-1.0, 1.0), (0.5, 2.0), (-3.0, 0.25), (2.0, 0.0)])
@pytest.mark.parametrize("m,n,k", [(600, 1100, 512), (300, 2049, 1000), (1100, 1030, 4096)])
def test_gemm_wide(sm, alpha, beta, m, n, k):
    # the wide split-FP16 products (k >= 512, n > 1024: the LU's level updates,
    # 256 x 128 tiles, three LDS stages); ragged tiles and k not a multiple of 16
    variant = "f16x3"
    g = torch.Generator(device="cuda").manual_seed(m + n + k)
    A = torch.randn(k, m, dtype=torch.complex128, device="cuda", generator=g).T
    B = torch.randn(n, k, dtype=torch.complex128, device="cuda", generator=g).T
    C = torch.randn(n, m, dtype=torch.complex128, device="cuda", generator=g).T
    Cd = C.to(torch.complex64)
    print(f"gemm wide {variant} alpha={alpha} beta={beta} m={m} n={n} k={k}", flush=True)
    sm.linalg.gemm(A.to(torch.complex64), B.to(torch.complex64), Cd, alpha=alpha, beta=beta, variant=variant)
    A32, B32, C32 = (x.to(torch.complex64).to(torch.complex128) for x in (A, B, C))
    ref = beta * C32 + alpha * (A32 @ B32)
    scale = abs(alpha) * A32.abs().max() * B32.abs().max() * k + abs(beta) * C32.abs().max()
    err = ((Cd.to(torch.complex128) - ref).abs().max() / scale).item()
    # normwise (Frobenius) error relative to |A| |B|
    fro = (torch.linalg.norm(Cd.to(torch.complex128) - ref) /
           (abs(alpha) * torch.linalg.norm(A32) * torch.linalg.norm(B32) + abs(beta) * torch.linalg.norm(C32))).item()
    print(f"  max err {err:.2e}, normwise {fro:.2e}", flush=True)
    assert err < 2e-7, err
    assert fro < 1e-7, fro


def well_conditioned(rng, n):
    return crandn(rng, n, n) + 2 * np.sqrt(n) * np.eye(n)


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "4m"), (32, "f16x3")])
@pytest.mark.parametrize("n", [16, 100, 300, 513, 1024, 2500, 5000])  # 2500, 5000: several 1024- and 4096-column blocks
def test_lu_solve(sm, prec, variant, n):
    rng = np.random.default_rng(n)
    A = crandn(rng, n, n)  # general (pivoting exercised)
    b = crandn(rng, n, 3)  # 3 columns: the single-RHS sweep path
    dt = torch.complex128 if prec == 64 else torch.complex64
    Ad = to_dev(A, dt)
    print(f"lu prec={prec} {variant} n={n}", flush=True)
    lu = sm.linalg.DenseLuSolver(variant=variant)
    lu.factorize(Ad)
    for op in ("N", "T", "C"):
        Bd = to_dev(b, dt)
        lu.solve(Bd, op=op)
        X = Bd.cpu().numpy()
        M = {"N": A, "T": A.T, "C": A.conj().T}[op]
        res = np.linalg.norm(M @ X - b) / (np.linalg.norm(M) * np.linalg.norm(X))
        assert res < (1e-14 if prec == 64 else 5e-6), (op, res)
    # factors reproduce P A = L U
    F = Ad.cpu().numpy()
    L = np.tril(F, -1) + np.eye(n)
    U = np.triu(F)
    piv = lu.pivots.cpu().numpy()
    PA = A.copy()
    for k in range(n):
        p = piv[k]
        PA[[k, p]] = PA[[p, k]]
    err = np.linalg.norm(PA - L @ U) / np.linalg.norm(A)
    assert err < (4 * n * 1.2e-16 if prec == 64 else 4 * n * 6e-8), err
    # pivots maximise |re| + |im| (LAPACK izamax), so |l_ij| <= sqrt(2)
    assert np.all(np.abs(L) <= np.sqrt(2) * (1 + 1e-6))


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "f16x3")])
def test_lu_deferred_swaps(sm, prec, variant, monkeypatch):
    """n = 12,500: the 4096-column blocks at 0, 4096 and 8192 each receive the
    swaps of every later block after the factorisation (three deferred
    permutations, composed from up to three later blocks), moved through a
    1 MiB temporary so that each block's columns go in many chunks. Checked on
    the GPU: ||PA - LU|| / ||A|| and the backward error of a solve."""
    n = 12500
    monkeypatch.setenv("SIE_DEFERRED_TMP_BYTES", str(1 << 20))
    dt = torch.complex128 if prec == 64 else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(12500)
    A = torch.randn(n, n, dtype=dt, device="cuda", generator=g).T  # column-major view
    Ad = A.clone()
    assert Ad.stride() == (1, n)
    print(f"lu deferred swaps prec={prec} {variant} n={n}", flush=True)
    lu = sm.linalg.DenseLuSolver(variant=variant)
    lu.factorize(Ad)
    piv = lu.pivots.cpu().numpy()
    assert np.all((piv >= np.arange(n)) & (piv < n))
    perm = np.arange(n)
    for k in range(n):
        p = piv[k]
        perm[k], perm[p] = perm[p], perm[k]
    L = torch.tril(Ad, -1) + torch.eye(n, dtype=dt, device="cuda")
    U = torch.triu(Ad)
    PA = A[torch.from_numpy(perm).cuda()]
    err = (torch.linalg.norm(PA - L @ U) / torch.linalg.norm(A)).item()
    del L, U, PA
    eps = 1.2e-16 if prec == 64 else 6e-8
    assert err < 4 * n * eps, err
    b = torch.randn(n, 1, dtype=dt, device="cuda", generator=g)
    x = b.clone()
    lu.solve(x)
    be = (torch.linalg.norm(A @ x - b) / (torch.linalg.norm(A) * torch.linalg.norm(x))).item()
    assert be < (1e-14 if prec == 64 else 5e-6), be


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "f16x3")])
@pytest.mark.parametrize("pattern", ["repeat", "chain"])
def test_lu_pivot_chains(sm, prec, variant, pattern):
    # A = rows of an upper-triangular U with a dominant diagonal, placed so that
    # partial pivoting picks, at every step k, U's row k from a known position:
    # "repeat" - always the last row (the same pivot row for all n - 1 steps,
    # outside every block of interchanges); "chain" - row k + 1 (each
    # interchange moves the row the previous one brought in). The multipliers
    # are zero, so the factors are U exactly and only the row interchanges can
    # be wrong; n = 5000 crosses panels, NB blocks and a 4096-column level.
    n = 5000
    dt = torch.complex128 if prec == 64 else torch.complex64
    g = torch.Generator(device="cuda").manual_seed(5000)
    U = torch.triu(0.01 * torch.randn(n, n, dtype=dt, device="cuda", generator=g), 1)
    U += torch.diag(4.0 + torch.rand(n, dtype=torch.float64, device="cuda", generator=g).to(dt))
    rows = torch.arange(n, device="cuda")
    if pattern == "repeat":
        src = torch.cat([rows[1:], rows[:1]])         # position j holds U row j + 1, the last U row 0
        expect = np.full(n, n - 1)
    else:
        src = torch.cat([rows[-1:], rows[:-1]])       # position 0 holds U row n - 1, position j + 1 U row j
        expect = np.minimum(np.arange(n) + 1, n - 1)
    A = U[src].T.contiguous().T                       # column-major
    print(f"lu pivot chains prec={prec} {variant} {pattern} n={n}", flush=True)
    lu = sm.linalg.DenseLuSolver(variant=variant)
    lu.factorize(A)
    piv = lu.pivots.cpu().numpy()
    assert np.array_equal(piv, expect), np.flatnonzero(piv != expect)[:10]
    assert torch.equal(A, U)


def test_singular_reported(sm):
    n = 64
    rng = np.random.default_rng(1)
    A = crandn(rng, n, n)
    A[:, 10] = 0.0
    lu = sm.linalg.DenseLuSolver()
    with pytest.raises(sm.SieError, match="singular at U diagonal 11"):
        lu.factorize(to_dev(A))


@pytest.mark.parametrize("op", ["N", "C"])
def test_zgemv(sm, op):
    rng = np.random.default_rng(5)
    m, n = 700, 3000
    A = crandn(rng, m, n)
    x = crandn(rng, n if op == "N" else m)
    y0 = crandn(rng, m if op == "N" else n)
    y = torch.from_numpy(y0).cuda()
    sm.linalg.zgemv(to_dev(A), torch.from_numpy(x).cuda(), y, alpha=-1.0, beta=1.0, op=op)
    M = A if op == "N" else A.conj().T
    ref = y0 - M @ x
    assert np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("op", ["N", "C"])
@pytest.mark.parametrize("variant", ["4m", "f16x3"])
def test_ir_solve(sm, op, variant):
    rng = np.random.default_rng(9)
    n = 700
    A = crandn(rng, n, n) + 5 * np.sqrt(n) * np.eye(n)
    b = crandn(rng, n, 2)
    Zd = to_dev(A)
    LU32 = sm.linalg.to_precision(Zd, 32)
    lu = sm.linalg.DenseLuSolver(variant=variant)
    lu.factorize(LU32)
    X, rr, its, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), op=op, tol=1e-13, maxit=20)
    M = A if op == "N" else A.conj().T
    xr = np.linalg.solve(M, b)
    assert np.all(rr < 1e-13)
    assert np.linalg.norm(X.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-12


@pytest.mark.parametrize("op", ["N", "C"])
def test_fgmres_ir_solve(sm, op):
    """FGMRES-IR on a system with condition number 1e7 and b = op(A) x0, where
    the FP32 LU leaves an error factor of ~0.07 per refinement sweep (10-12
    sweeps to 1e-13): both methods reach 1e-13, the Krylov one in fewer steps,
    and both solutions have that residual in FP64 on the host."""
    rng = np.random.default_rng(10)
    n = 600
    U, _ = np.linalg.qr(crandn(rng, n, n))
    V, _ = np.linalg.qr(crandn(rng, n, n))
    A = (U * np.logspace(0, -7, n)) @ V.conj().T
    M = A if op == "N" else A.conj().T
    b = M @ crandn(rng, n, 1)
    Zd = to_dev(A)
    LU32 = sm.linalg.to_precision(Zd, 32)
    lu = sm.linalg.DenseLuSolver(variant="f16x3")
    lu.factorize(LU32)
    Xi, ri, iti, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), op=op, tol=1e-13, maxit=80)
    Xf, rf, itf, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), op=op, tol=1e-13, maxit=80, method="fgmres")
    print(f"op={op}: IR {iti[0]} sweeps ({ri[0]:.1e}), FGMRES-IR {itf[0]} steps ({rf[0]:.1e})", flush=True)
    assert ri[0] < 1e-13 and rf[0] < 1e-13
    assert itf[0] < iti[0]
    for X in (Xi, Xf):
        x = X.cpu().numpy()
        assert np.linalg.norm(M @ x - b) / np.linalg.norm(b) < 1e-13


def test_fgmres_ir_floor(sm):
    """A backward-error target far below FP64 (1e-30): FGMRES-IR refines to the
    residual's FP64 floor and stops at the first restart cycle that does not
    halve the residual, long before maxit, with that floor's backward error."""
    rng = np.random.default_rng(12)
    n = 500
    A = crandn(rng, n, n) + 2 * np.sqrt(n) * np.eye(n)
    b = crandn(rng, n, 1)
    Zd = to_dev(A)
    LU32, zf = sm.linalg.to_precision(Zd, 32, fro=True)
    lu = sm.linalg.DenseLuSolver(variant="f16x3")
    lu.factorize(LU32)
    X, rr, its, be = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), tol=1e-30, maxit=400, method="fgmres",
                                        restart=20, znorm=zf)
    print(f"floor: {its[0]} steps, backward error {be[0]:.2e}", flush=True)
    assert its[0] < 100
    assert be[0] < 1e-16
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(X.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-14


def test_fgmres_ir_restarts(sm):
    """A restart length of 2 forces several cycles (each ending with the true
    residual) and still converges."""
    rng = np.random.default_rng(11)
    n = 300
    A = crandn(rng, n, n) + 3 * np.sqrt(n) * np.eye(n)
    b = crandn(rng, n, 1)
    Zd = to_dev(A)
    LU32 = sm.linalg.to_precision(Zd, 32)
    lu = sm.linalg.DenseLuSolver()
    lu.factorize(LU32)
    X, rr, its, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), tol=1e-14, maxit=40, method="fgmres", restart=2)
    assert rr[0] < 1e-14
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(X.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-13
    with pytest.raises(ValueError):
        sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), method="cg")


@pytest.mark.parametrize("variant", ["4m", "f16x3"])
def test_ir_solve_many_rhs(sm, variant):
    """More than four right-hand sides take the blocked refinement (one FP64
    GEMM residual and one multi-RHS FP32 solve per sweep); columns of very
    different scale converge independently."""
    rng = np.random.default_rng(12)
    n, nrhs = 900, 24
    A = crandn(rng, n, n) + 5 * np.sqrt(n) * np.eye(n)
    b = crandn(rng, n, nrhs) * np.logspace(-20, 20, nrhs)[None, :]
    Zd = to_dev(A)
    LU32 = sm.linalg.to_precision(Zd, 32)
    lu = sm.linalg.DenseLuSolver(variant=variant)
    lu.factorize(LU32)
    X, rr, its, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), tol=1e-13, maxit=30)
    assert np.all(rr < 1e-13), rr
    assert its.max() <= 12
    xr = np.linalg.solve(A, b)
    x = X.cpu().numpy()
    for c in range(nrhs):
        assert np.linalg.norm(x[:, c] - xr[:, c]) / np.linalg.norm(xr[:, c]) < 1e-12


@pytest.mark.parametrize("method,cond", [("ir", 1e6), ("fgmres", 1e7)])
def test_backward_error_stop(sm, method, cond):
    """Ill-conditioned systems with a right-hand side outside the dominant
    range, so |x| is ~cond |b| and no solve reaches |r| / |b| = 1e-12 (FP64
    rounding alone leaves ~u |A| |x| / |b|): that test reports it unreached,
    classic IR after maxit sweeps, FGMRES-IR at the first restart cycle that
    does not halve the residual. With the Frobenius norm from the conversion the refinement
    stops at the normwise backward error |r| / (|A|_F |x| + |b|) <= 1e-15,
    which FP64 LU itself attains, and the result matches the host FP64 solve
    to the conditioning bound. (A host simulation with FP32 LU factors takes
    5-7 sweeps / 20 FGMRES steps; beyond cond ~1e8 the FP32 factors no
    longer precondition.)"""
    rng = np.random.default_rng(13)
    n = 500
    U, _ = np.linalg.qr(crandn(rng, n, n))
    V, _ = np.linalg.qr(crandn(rng, n, n))
    A = (U * np.logspace(0, -np.log10(cond), n)) @ V.conj().T
    b = crandn(rng, n, 1)
    Zd = to_dev(A)
    LU32, fro = sm.linalg.to_precision(Zd, 32, fro=True)
    assert abs(fro - np.linalg.norm(A)) <= 1e-13 * np.linalg.norm(A)
    lu = sm.linalg.DenseLuSolver(variant="f16x3")
    lu.factorize(LU32)
    _, rr0, it0, _ = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), tol=1e-12, maxit=40, method=method)
    assert rr0[0] > 1e-12 and (it0[0] == 40 if method == "ir" else it0[0] < 40)
    X, rr, its, be = sm.linalg.ir_solve(Zd, LU32, lu.pivots, to_dev(b), tol=1e-15, maxit=40, method=method,
                                        znorm=fro)
    x = X.cpu().numpy()
    print(f"{method}: backward error {be[0]:.2e} (rel. residual {rr[0]:.2e}) in {its[0]} steps", flush=True)
    assert be[0] <= 1e-15 and its[0] < 40
    host_be = np.linalg.norm(A @ x - b) / (np.linalg.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))
    assert host_be <= 2e-15
    xr = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < cond * 1e-14


def test_wait_bound_forced_timeout(sm):
    """The panels' grid-wide exchanges are bounded: forced to give up (testing
    switch), a factorisation returns a device error instead of hanging, every
    kernel exits, and the next factorisation (switch off) is correct."""
    rng = np.random.default_rng(14)
    n = 1500  # three panel workgroups: a real grid-wide exchange
    A = crandn(rng, n, n)
    for prec, dt in ((32, torch.complex64), (64, torch.complex128)):
        lu = sm.linalg.DenseLuSolver(variant="f16x3" if prec == 32 else "4m")
        sm.linalg.set_wait_bound(0.01, force_timeout=True)
        try:
            with pytest.raises(sm.SieError, match="device wait gave up") as e:
                lu.factorize(to_dev(A, dt))
            assert e.value.code == 5
        finally:
            sm.linalg.set_wait_bound(5.0, force_timeout=False)
        Ad = to_dev(A, dt)
        lu.factorize(Ad)
        b = crandn(rng, n, 1)
        Bd = to_dev(b, dt)
        lu.solve(Bd)
        x = Bd.cpu().numpy()
        res = np.linalg.norm(A @ x - b) / (np.linalg.norm(A) * np.linalg.norm(x))
        assert res < (1e-14 if prec == 64 else 5e-6), (prec, res)
    with pytest.raises(ValueError):
        sm.linalg.set_wait_bound(0.0)


def test_convert_grid_stride(sm):
    # more elements than one capped dispatch covers: exercises the grid-stride loop
    n = 5000
    A = torch.randn(n, n, dtype=torch.complex128, device="cuda").T
    B = sm.linalg.to_precision(A, 32)
    assert torch.equal(B, A.to(torch.complex64))


@pytest.mark.parametrize("scale", [1e-30, 1.0, 1e30])
def test_gemm_f16x3_scaling(sm, scale):
    # the B operand is rescaled by a power of two before the FP16 split: results
    # keep FP32-level accuracy far outside the FP16 range
    rng = np.random.default_rng(4)
    m, n, k = 200, 150, 96
    A, B = crandn(rng, m, k), crandn(rng, k, n) * scale
    dt = torch.complex64
    Cd = sm.linalg.gemm(to_dev(A, dt), to_dev(B, dt), variant="f16x3")
    ref = A @ B
    err = np.abs(Cd.cpu().numpy() - ref).max() / (np.abs(A).max() * np.abs(B).max() * k)
    assert err < 2e-7, err


@pytest.mark.parametrize("op", ["N", "T", "C"])
def test_lu_solve_many_rhs(sm, op):
    # more than 4 right-hand sides: the blocked TRSM path
    rng = np.random.default_rng(21)
    n, nrhs = 700, 9
    A = crandn(rng, n, n)
    b = crandn(rng, n, nrhs)
    Ad = to_dev(A)
    lu = sm.linalg.DenseLuSolver()
    lu.factorize(Ad)
    Bd = to_dev(b)
    lu.solve(Bd, op=op)
    M = {"N": A, "T": A.T, "C": A.conj().T}[op]
    X = Bd.cpu().numpy()
    assert np.linalg.norm(M @ X - b) / (np.linalg.norm(M) * np.linalg.norm(X)) < 1e-14


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("op", ["N", "C"])
def test_lu_solve_wide_rhs(sm, prec, op):
    # more than 4096 right-hand sides (ragged in 64): the one-lane-per-column
    # small triangular solves, all four triangle forms through ops N and C
    rng = np.random.default_rng(22)
    n, nrhs = 100, 4500
    A = crandn(rng, n, n)
    b = crandn(rng, n, nrhs)
    dt = torch.complex128 if prec == 64 else torch.complex64
    print(f"lu wide rhs prec={prec} op={op}", flush=True)
    lu = sm.linalg.DenseLuSolver()
    lu.factorize(to_dev(A, dt))
    Bd = to_dev(b, dt)
    lu.solve(Bd, op=op)
    M = {"N": A, "T": A.T, "C": A.conj().T}[op]
    X = Bd.cpu().numpy()
    res = np.linalg.norm(M @ X - b, axis=0) / (np.linalg.norm(M) * np.linalg.norm(X, axis=0))
    assert res.max() < (1e-14 if prec == 64 else 5e-6), res.max()


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("op", ["N", "C"])
def test_lu_solve_many_rhs_scaled(sm, prec, op):
    # more than 4 columns over several 256-row blocks (the block products; op C
    # with the factors' rows and columns conjugate-transposed), columns 20
    # orders of magnitude apart: each column's backward error is relative to
    # its own scale
    rng = np.random.default_rng(23)
    n, nrhs = 1000, 12
    A = crandn(rng, n, n)
    b = crandn(rng, n, nrhs) * np.logspace(-10, 10, nrhs)[None, :]
    dt = torch.complex128 if prec == 64 else torch.complex64
    print(f"lu many rhs scaled prec={prec} op={op}", flush=True)
    lu = sm.linalg.DenseLuSolver(variant="4m" if prec == 64 else "f16x3")
    lu.factorize(to_dev(A, dt))
    Bd = to_dev(b, dt)
    lu.solve(Bd, op=op)
    M = A if op == "N" else A.conj().T
    X = Bd.cpu().numpy().astype(np.complex128)
    res = np.linalg.norm(M @ X - b, axis=0) / (np.linalg.norm(M) * np.linalg.norm(X, axis=0))
    assert res.max() < (1e-14 if prec == 64 else 5e-6), res


def _gmres_system(rng, n, shift):
    # complex Gaussian entries: eigenvalues in a disk of radius ~sqrt(2n) about shift sqrt(n)
    A = crandn(rng, n, n) + shift * np.sqrt(n) * np.eye(n)
    return A, crandn(rng, n)


def test_gmres_dense_matches_lu(sm):
    rng = np.random.default_rng(31)
    n = 400
    A, b = _gmres_system(rng, n, 2.0)
    r = sm.linalg.gmres(to_dev(A), torch.from_numpy(b).cuda(), m=40, maxnit=10, tol=1e-12)
    xr = np.linalg.solve(A, b)
    assert r.converged and r.iterations > 0
    assert np.linalg.norm(r.x.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-10
    assert r.residual <= 1e-12 * np.linalg.norm(b) * (1 + 1e-9)
    assert len(r.history) == r.iterations and r.history[-1] <= 1e-12


def test_gmres_restarts_and_callback(sm):
    rng = np.random.default_rng(32)
    n = 300
    # spectrum in a disk of radius ~sqrt(2n) about 2.5 sqrt(n): away from the origin,
    # so GMRES(8) converges, but needs several restarts for 1e-10
    A, b = _gmres_system(rng, n, 2.5)
    Ad = to_dev(A)

    def matvec(x, y):
        y.copy_(Ad @ x)
    x = torch.zeros(n, dtype=torch.complex128, device="cuda")
    r = sm.linalg.GMRESSolver(n, matvec).solve(torch.from_numpy(b).cuda(), x, m=8, maxnit=400, tol=1e-10)
    xr = np.linalg.solve(A, b)
    assert r.converged and r.iterations > 8, (r.info, r.iterations, r.history[-5:])  # needed restarts
    assert np.linalg.norm(x.cpu().numpy() - xr) / np.linalg.norm(xr) < 1e-8


def test_gmres_edge_cases(sm):
    rng = np.random.default_rng(33)
    n = 64
    A, b = _gmres_system(rng, n, 2.0)
    Ad = to_dev(A)
    # zero right-hand side: x = 0 without iterations
    r = sm.linalg.gmres(Ad, torch.zeros(n, dtype=torch.complex128, device="cuda"), x0=torch.ones(n, dtype=torch.complex128, device="cuda"))
    assert r.converged and r.iterations == 0 and r.residual == 0 and torch.all(r.x == 0)
    # exact initial guess: converged at the first residual check
    xr = torch.from_numpy(np.linalg.solve(A, b)).cuda()
    r = sm.linalg.gmres(Ad, torch.from_numpy(b).cuda(), x0=xr, tol=1e-8)
    assert r.converged and r.iterations == 0
    # too few steps: not converged, info 1
    r = sm.linalg.gmres(Ad, torch.from_numpy(b).cuda(), m=2, maxnit=1, tol=1e-14)
    assert r.info == 1 and r.iterations == 2
    with pytest.raises(ValueError):
        sm.linalg.gmres(Ad, torch.from_numpy(b).cuda(), m=0)


@pytest.mark.parametrize("op", ["N", "C"])
def test_single_rhs_graph_replays(sm, op):
    # single-RHS solves replay a captured launch sequence: same factors and
    # vector buffer, new contents each time, eager work in between
    rng = np.random.default_rng(41)
    n = 900
    A = crandn(rng, n, n)
    lu = sm.linalg.DenseLuSolver()
    Ad = to_dev(A, torch.complex64)
    lu.factorize(Ad)
    buf = torch.empty(n, dtype=torch.complex64, device="cuda")
    M = A if op == "N" else A.conj().T
    for rep in range(4):
        b = crandn(rng, n)
        buf.copy_(torch.from_numpy(b).to(torch.complex64))
        noise = torch.randn(256, 256, device="cuda") @ torch.randn(256, 256, device="cuda")  # eager work
        lu.solve(buf, op=op)
        x = buf.cpu().numpy().astype(np.complex128)
        res = np.linalg.norm(M @ x - b) / (np.linalg.norm(M) * np.linalg.norm(x))
        assert res < 5e-6, (rep, res)
        del noise
    # a second factorisation in the same buffers: the replay reads the new factors
    A2 = crandn(rng, n, n)
    Ad.copy_(to_dev(A2, torch.complex64))
    lu.factorize(Ad)
    b = crandn(rng, n)
    buf.copy_(torch.from_numpy(b).to(torch.complex64))
    lu.solve(buf, op=op)
    M2 = A2 if op == "N" else A2.conj().T
    x = buf.cpu().numpy().astype(np.complex128)
    assert np.linalg.norm(M2 @ x - b) / (np.linalg.norm(M2) * np.linalg.norm(x)) < 5e-6


def test_mfma_rate_calibration(sm):
    """The sustained FP16 matrix-core rate lies below the published dense peak
    (2.5 PFLOP/s at 2.4 GHz) and the clock below the peak clock."""
    tf, mhz = sm.linalg.mfma_f16_rate(iters=2000, reps=2)
    print(f"mfma f16 rate {tf:.0f} TFLOP/s at {mhz:.0f} MHz", flush=True)
    assert 200.0 < tf < 2520.0
    assert 500.0 < mhz < 2500.0
    assert tf / 2516.6 <= mhz / 2400.0 * 1.02  # the rate follows the clock


@pytest.mark.parametrize("prec,variant", [(64, "4m"), (32, "f16x3")])
def test_panel_launch_forms(sm, prec, variant):
    """Plain and cooperative launches of the panel kernels give bitwise the
    same factors and pivots (n = 5000: panels of 5000 down to 8 rows)."""
    n = 5000
    rng = np.random.default_rng(77)
    A = crandn(rng, n, n)
    dt = torch.complex128 if prec == 64 else torch.complex64
    out = []
    try:
        for coop in (False, True):
            sm.linalg.set_panel_launch(coop)
            print(f"panel launch cooperative={coop} prec={prec}", flush=True)
            Ad = to_dev(A, dt)
            lu = sm.linalg.DenseLuSolver(variant=variant)
            lu.factorize(Ad)
            out.append((Ad.cpu().numpy(), lu.pivots.cpu().numpy()))
    finally:
        sm.linalg.set_panel_launch(False)
    assert np.array_equal(out[0][1], out[1][1])
    fl = np.float64 if prec == 64 else np.float32
    assert np.array_equal(np.ascontiguousarray(out[0][0]).view(fl), np.ascontiguousarray(out[1][0]).view(fl))
