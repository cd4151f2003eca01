"""ctypes binding of libsie_amd.so, the C ABI declared in include/sie_mom/sie_capi.h.

The library is built in place by ``sie-gstc-2026_amd/build.sh`` (or
``__graft_entry__.build()``). There is no fallback: if the shared object is
missing, importing this module raises.
"""
import atexit
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsie_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with sie-gstc-2026_amd/build.sh (no CPU fallback exists)"
    )

lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
I = ctypes.c_int
LL = ctypes.c_longlong
D = ctypes.c_double
Fl = ctypes.c_float

SIE_OK = 0
SIE_ERR_INVALID = 1
SIE_ERR_OVERFLOW = 2
SIE_ERR_SINGULAR = 3
SIE_ERR_NONFINITE = 4
SIE_ERR_DEVICE = 5


class SieError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class SieInvalidArgument(SieError, ValueError):
    pass


class SieOverflow(SieError, OverflowError):
    pass


def check(code):
    if code == SIE_OK:
        return
    msg = lib.sie_last_error().decode()
    if code == SIE_ERR_INVALID:
        raise SieInvalidArgument(code, msg)
    if code == SIE_ERR_OVERFLOW:
        raise SieOverflow(code, msg)
    raise SieError(code, msg)


class CalderonParams(ctypes.Structure):
    _fields_ = [
        ("rule", I),
        ("k1", Fl), ("k2", Fl),
        ("mu_r1", Fl), ("mu_r2", Fl), ("eps_r1", Fl), ("eps_r2", Fl),
        ("mu0", Fl), ("eps0", Fl), ("omega", Fl),
        ("threshold", Fl),
    ]


class PlaneWaveC(ctypes.Structure):
    _fields_ = [("E0", D * 3), ("khat", D * 3), ("k", D), ("eta", D), ("region", I)]


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("sie_last_error", ctypes.c_char_p, [])
_sig("sie_mesh_create", I, [P, I, P, I, P, I, ctypes.POINTER(P)])
_sig("sie_mesh_destroy", I, [P])
_sig("sie_calderon_pair_blocks", I, [P, ctypes.POINTER(CalderonParams), P, P])
_sig("sie_assemble_gstc", I, [P, ctypes.POINTER(CalderonParams), P, I, I, I, P, LL, P])
_sig("sie_assemble_pec", I, [P, ctypes.POINTER(CalderonParams), P, LL, P])
_sig("sie_assemble_pmchwt", I, [P, ctypes.POINTER(CalderonParams), P, LL, P])
_sig("sie_gstc_matvec", I, [P, ctypes.POINTER(CalderonParams), P, I, I, I, P, P, P])
_sig("sie_pec_matvec", I, [P, ctypes.POINTER(CalderonParams), P, P, P])
_sig("sie_pmchwt_matvec", I, [P, ctypes.POINTER(CalderonParams), P, P, P])
_sig("sie_surface_fields", I, [P, I, D, D, D, P, P, P, I, P, P, P])
_sig("sie_rhs_plane_waves", I, [P, I, I, ctypes.POINTER(PlaneWaveC), I, Fl, Fl, D, D, P, LL, P])
_sig("sie_gemm", I, [I, I, I, I, D, P, LL, P, LL, D, P, LL, I, P])
_sig("sie_lu_factor", I, [I, I, P, LL, P, ctypes.POINTER(I), I, P])
_sig("sie_lu_solve", I, [I, I, I, I, P, LL, P, P, LL, P])
_sig("sie_zgemv", I, [I, I, I, D, P, LL, P, D, P, P])
_sig("sie_zgemv_noalloc", I, [I, I, I, D, P, LL, P, D, P, P])
_sig("sie_convert", I, [I, I, LL, LL, P, LL, P, LL, D, P, P])
_sig("sie_ir_solve", I, [I, I, I, P, LL, P, LL, P, P, LL, P, LL, D, D, I, P, P, P, P])
_sig("sie_fgmres_ir_solve", I, [I, I, I, P, LL, P, LL, P, P, LL, P, LL, D, D, I, I, P, P, P, P])
_sig("sie_set_wait_bound", I, [D, I])
_sig("sie_set_panel_launch", I, [I])
_sig("sie_release", I, [])
_sig("sie_norm2", D, [LL, P, P])
_sig("sie_mfma_f16_rate", I, [I, I, ctypes.POINTER(D), ctypes.POINTER(D), P])
MATVEC_FN = ctypes.CFUNCTYPE(I, P, P, P, P)
_sig("sie_gmres", I, [I, I, I, D, MATVEC_FN, P, P, P, ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I), P, I,
                      ctypes.POINTER(I), P])
_sig("sie_gmres_dense", I, [I, P, LL, I, I, D, P, P, ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I), P, I,
                            ctypes.POINTER(I), P])
_sig("sie_gstc_records", I, [I, I, P, D, D, D, D, P, ctypes.POINTER(I), P])
_sig("sie_farfield", I, [P, I, D, D, P, I, P, P, P])
_sig("sie_farfield_adjoint", I, [P, I, D, D, I, P, P, P, P])
_sig("sie_gstc_mass_sensitivity", I, [P, I, I, Fl, Fl, P, P, P, P])
_sig("sie_gstc_param_grad", I, [I, I, P, P, D, D, D, P, P])
_sig("sie_profile_enable", I, [I, I])
_sig("sie_profile_read", I, [I, ctypes.POINTER(I), ctypes.POINTER(D), ctypes.POINTER(D), ctypes.POINTER(D)])


@atexit.register
def _release():
    # device state is released while the HIP runtime is still up (atexit runs
    # before the runtime's own teardown); nothing to do if no device work ran
    try:
        lib.sie_release()
    except Exception:
        pass
