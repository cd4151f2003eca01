"""sie_zgemv_noalloc (the GEMV whose partial sums live in a static device
array instead of an allocated workspace) against sie_zgemv, through the C
ABI: bit-identical results for op N and op C on two calls in a row, the
second with a smaller operator (the shared scratch keeps the first call's
partial sums, which the second must not read), and the size limit reported
as an invalid argument without touching the device."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def lib():
    from sie_mom_amd import _lib
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return _lib


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _gemv(fn, op, A, x, alpha, beta, y0):
    m, n = A.shape
    y = y0.clone()
    code = fn(op, m, n, alpha, _ptr(A), A.stride(1), _ptr(x), beta, _ptr(y), None)
    return code, y


@pytest.mark.parametrize("op", [0, 2])
def test_noalloc_gemv_matches_gemv(lib, op):
    g = torch.Generator(device="cuda").manual_seed(5 + op)
    for m, n in ((3000, 9000), (700, 1300)):  # the second call smaller: stale partial sums must not leak
        print(f"noalloc gemv op={op} m={m} n={n}", flush=True)
        A = torch.randn(n, m, dtype=torch.complex128, device="cuda", generator=g).T  # column-major
        x = torch.randn(n if op == 0 else m, dtype=torch.complex128, device="cuda", generator=g)
        y0 = torch.randn(m if op == 0 else n, dtype=torch.complex128, device="cuda", generator=g)
        c1, y1 = _gemv(lib.lib.sie_zgemv, op, A, x, -1.5, 0.5, y0)
        c2, y2 = _gemv(lib.lib.sie_zgemv_noalloc, op, A, x, -1.5, 0.5, y0)
        torch.cuda.synchronize()
        assert c1 == 0 and c2 == 0, lib.lib.sie_last_error()
        assert torch.equal(y1, y2)
        ref = 0.5 * y0 - 1.5 * ((A if op == 0 else A.conj().T) @ x)
        assert float(torch.linalg.norm(y2 - ref) / torch.linalg.norm(ref)) < 1e-13


def test_noalloc_gemv_size_limit(lib):
    # op N with m ceil(n / 2048) * 16 bytes above 64 MiB: rejected before any launch
    # (m = 2^20, n = 4 * 2048 + 1: five chunks, 80 MiB); the buffers are never read
    m, n = 1 << 20, 4 * 2048 + 1
    assert m * ((n + 2047) // 2048) * 16 > 64 << 20
    t = torch.zeros(4, dtype=torch.complex128, device="cuda")
    code = lib.lib.sie_zgemv_noalloc(0, m, n, 1.0, _ptr(t), m, _ptr(t), 0.0, _ptr(t), None)
    assert code == lib.SIE_ERR_INVALID
    assert "static scratch" in lib.lib.sie_last_error().decode()
