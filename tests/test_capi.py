"""The C ABI library loads and exports every function declared in include/."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SO = os.path.join(ROOT, "sie-gstc-2026_amd", "sie_mom_amd", "libsie_amd.so")


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "sie_mom", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(sie_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(SO)
    names = declared()
    assert len(names) >= 15
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_invalid_arguments_rejected_without_gpu():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "sie-gstc-2026_amd"))
    from sie_mom_amd import _lib
    rc = _lib.lib.sie_lu_solve(16, 0, 10, 1, None, 10, None, None, 10, None)
    assert rc == _lib.SIE_ERR_INVALID
    assert "precision" in _lib.lib.sie_last_error().decode()
    counts = (ctypes.c_int * 3)()
    f = _lib.lib.sie_icosphere
    f.argtypes = [ctypes.c_float, ctypes.c_int] + [ctypes.c_void_p] * 5
    assert f(1.0, -1, None, None, None, None, counts) == _lib.SIE_ERR_INVALID


def test_gmres_validation_without_gpu():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "sie-gstc-2026_amd"))
    from sie_mom_amd import _lib
    res, it, info, hl = ctypes.c_double(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rc = _lib.lib.sie_gmres_dense(10, None, 10, 5, 5, 1e-8, None, None, ctypes.byref(res), ctypes.byref(it),
                                  ctypes.byref(info), None, 0, ctypes.byref(hl), None)
    assert rc == _lib.SIE_ERR_INVALID
    rc = _lib.lib.sie_gmres(10, 0, 5, 1e-8, _lib.MATVEC_FN(lambda *a: 0), None, None, None, ctypes.byref(res),
                            ctypes.byref(it), ctypes.byref(info), None, 0, ctypes.byref(hl), None)
    assert rc == _lib.SIE_ERR_INVALID
    assert "restart" in _lib.lib.sie_last_error().decode()
