"""ctypes binding of libfrpa.so (the HIP kernels and their C-ABI).

The library is built in-tree by ``python -m fast_rpa_eval_amd.build``.  There is no fallback: if
the library is missing or fails to load, every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# FRPA_LIBRARY may name another build of the library (for example one made with different flags)
LIB_PATH = os.environ.get("FRPA_LIBRARY") or os.path.join(HERE, "libfrpa.so")

FRPA_OK = 0
FRPA_PREC_FP64 = 0
FRPA_PREC_OZAKI = 1
FRPA_PREC_OZAKI4 = 2
FRPA_PREC_FP32 = 3
FRPA_PREC_BF16X3 = 4
FRPA_PREC_FP64_SYM = 5
FRPA_FAR_TERMS = 0
FRPA_FAR_HORNER = 1

_lib = None


class FrpaError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FrpaError(f"{LIB_PATH} not built: run python -m fast_rpa_eval_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        p, i, ll = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)
        L.frpa_last_error.restype = ctypes.c_char_p
        L.frpa_version.restype = ctypes.c_char_p
        L.frpa_rho1_plan_create.argtypes = [ctypes.POINTER(p), i, i, i, i, p, p, p, p, p, p, p]
        L.frpa_rho1_plan_destroy.argtypes = [p]
        L.frpa_rho1_forward.argtypes = [p, i, i, i, p, p, p, p, p, p, p, p]
        L.frpa_rho1_adjoint.argtypes = [p, i, i, p, p, p, p]
        L.frpa_rho1_prepare_wave.argtypes = [p, i, p, ll, p]
        L.frpa_rho1_autotune.argtypes = [p, i, i, p, p, p, i, ctypes.c_char_p, i, p]
        L.frpa_rho1_set_launch.argtypes = [p, i, i, i, i]
        L.frpa_synth_wave.argtypes = [p, i, p, p, p, p, p, p, p, p]
        L.frpa_far_plan_create.argtypes = [ctypes.POINTER(p), i, i, i, i, i, i, i, i] + [p] * 9
        L.frpa_far_plan_destroy.argtypes = [p]
        L.frpa_far_forward.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p, p, p]
        L.frpa_far_adjoint.argtypes = [p, i, i, i, i, p, p, p, p, p, p, p, p, p]
        L.frpa_far_plan_set_near.argtypes = [p, i, p, p, p, p]
        L.frpa_rho1_bias_part2.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p]
        L.frpa_rho1_bias_inline.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p]
        L.frpa_rho1_h1_bias_part2.argtypes = [p, i, i, i, i, i] + [p] * 10
        L.frpa_rho1_h1.argtypes = [p, i, i, i, i, i] + [p] * 16
        L.frpa_near_forward.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p, p]
        L.frpa_far_forward_noalloc.argtypes = [p, i, i, i, i, i, p, p, p, p, p, p, p, p]
        L.frpa_rho1_set_adjoint_wait_limit.argtypes = [p, i]
        L.frpa_rho1_adjoint_status.argtypes = [p, p]
        _lib = L
    return _lib


def check(rc, what):
    if rc != FRPA_OK:
        msg = lib().frpa_last_error().decode()
        raise FrpaError(f"{what} failed (status {rc}): {msg}")


def exported_symbols():
    """Names the loaded library resolves (for the ABI export test)."""
    return lib()
