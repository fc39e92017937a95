"""ctypes binding of libfastsht.so (the C ABI in include/fastsht.h).

The library is built in-tree (fastsht_amd/libfastsht.so, see build.py).  There
is no fallback implementation: if the library is missing the import fails.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# FSHT_LIBRARY selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("FSHT_LIBRARY", os.path.join(_HERE, "libfastsht.so"))

FSHT_OK = 0
FSHT_ERR_ARG = -1
FSHT_ERR_NOMEM = -2
FSHT_ERR_HIP = -3
FSHT_ERR_STATE = -4
FSHT_HOST = 0
FSHT_PREC_FP64 = 0
FSHT_PREC_MIXED = 1
FSHT_DEVICE = 1

# every symbol declared in include/fastsht.h
EXPORTS = (
    "fsht_version", "fsht_last_error", "fsht_create", "fsht_destroy", "fsht_get_stream",
    "fsht_set_stream", "fsht_synchronize", "fsht_info", "fsht_ring_tab", "fsht_set_precision", "fsht_set_batch", "fsht_t2alm", "fsht_t2alm_iter",
    "fsht_get_alm", "fsht_alm2t", "fsht_qu2eb", "fsht_qu2e", "fsht_qu2b", "fsht_eb2qu", "fsht_fix_eb",
    "fsht_fix_eb_noalloc",
    "fsht_profile", "fsht_profile_read", "fsht_norm_alm", "fsht_filt_alm", "fsht_convert_alm_healpy", "fsht_convert_alm_healpy_inv",
    "fsht_alm2cl", "fsht_cl2alm", "fsht_host_alloc", "fsht_host_free",
)

_lib = None


class FastSHTError(RuntimeError):
    """A libfastsht call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"libfastsht error {code}: {msg}")
        self.code = code


def load():
    """Load libfastsht.so and declare its signatures (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m fastsht_amd.build` "
            "(hipcc, gfx950). There is no CPU fallback.")
    lib = ctypes.CDLL(LIB_PATH)
    vp, dp, ip, i, d = ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_double
    H = ctypes.c_void_p
    lib.fsht_version.restype = ctypes.c_char_p
    lib.fsht_version.argtypes = []
    lib.fsht_last_error.restype = ctypes.c_char_p
    lib.fsht_last_error.argtypes = [H]
    lib.fsht_create.argtypes = [i, i, i, i, i, ctypes.POINTER(ctypes.c_void_p)]
    lib.fsht_destroy.argtypes = [H]
    lib.fsht_get_stream.argtypes = [H, ctypes.POINTER(ctypes.c_void_p)]
    lib.fsht_set_stream.argtypes = [H, vp]
    lib.fsht_synchronize.argtypes = [H]
    lib.fsht_info.argtypes = [H, ctypes.POINTER(ctypes.c_longlong), i]
    lib.fsht_ring_tab.argtypes = [H, ip]
    lib.fsht_set_precision.argtypes = [H, i]
    lib.fsht_set_batch.argtypes = [H, i]
    lib.fsht_t2alm.argtypes = [H, dp, dp, i, i]
    lib.fsht_t2alm_iter.argtypes = [H, dp, dp, i, i]
    lib.fsht_get_alm.argtypes = [H, dp, i]
    lib.fsht_alm2t.argtypes = [H, dp, dp, i]
    lib.fsht_qu2eb.argtypes = [H, dp, dp, dp, dp, i, i]
    lib.fsht_qu2e.argtypes = [H, dp, dp, dp, i]
    lib.fsht_qu2b.argtypes = [H, dp, dp, dp, i]
    lib.fsht_eb2qu.argtypes = [H, dp, dp, dp, dp, i]
    lib.fsht_fix_eb.argtypes = [H, dp, dp, dp, dp, dp, dp, dp, i, i]
    lib.fsht_fix_eb_noalloc.argtypes = [H, dp, dp, dp, dp, dp, dp, dp, i]
    lib.fsht_profile.argtypes = [H, i]
    lib.fsht_profile_read.argtypes = [H, ctypes.POINTER(ctypes.c_double), i]
    lib.fsht_norm_alm.argtypes = [dp, dp, dp, dp, d, i, i, vp]
    lib.fsht_filt_alm.argtypes = [dp, dp, i, i, vp]
    lib.fsht_convert_alm_healpy.argtypes = [dp, dp, i, i, i, vp]
    lib.fsht_convert_alm_healpy_inv.argtypes = [dp, dp, i, i, i, vp]
    lib.fsht_alm2cl.argtypes = [dp, dp, dp, dp, i, i, i, vp]
    lib.fsht_host_alloc.argtypes = [ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_void_p)]
    lib.fsht_host_free.argtypes = [vp]
    lib.fsht_cl2alm.argtypes = [dp, dp, dp, dp, i, i, i, ctypes.c_ulonglong, vp]
    for name in EXPORTS:
        if name not in ("fsht_version", "fsht_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def check(rc, handle=None):
    if rc != FSHT_OK:
        msg = load().fsht_last_error(handle)
        raise FastSHTError(rc, msg.decode() if msg else "")
    return rc
