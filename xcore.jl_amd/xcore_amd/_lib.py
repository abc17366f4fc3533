"""ctypes binding of libxcore_hip.so (the C ABI in include/xcore_hip.h).

The library is mandatory: if it is missing or fails to load, every entry point
raises -- there is no CPU fallback on the product path.
"""
import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("XCORE_HIP_LIB", os.path.join(_PKG, "libxcore_hip.so"))

XC_OK = 0
XC_EINVAL_SHAPE = 1
XC_EINVAL_SCHEME = 2
XC_EINVAL_BC = 3
XC_EHIP = 4
XC_ENOCONV = 5
XC_EINVAL_ARG = 6
XC_ENOMEM = 7
XC_CTX_SYNC = 1
XC_CTX_EXACT = 2
XC_CTX_AUTOTUNE = 4

SCHEMES = {"centr": 0, "upwd1": 1, "quick": 2, "fromm": 3, "weno3": 4, "weno5": 5, "tvdim": 6}
BCS = {"periodic": 0, "closed": 1}


class XCMat(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


class XCoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class NoConvergence(XCoreError):
    pass


_lib = None
_lock = threading.Lock()

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_M = XCMat
_PM = ctypes.POINTER(XCMat)

# name -> argtypes (restype int unless listed in _RESTYPE)
_PROTOS = {
    "xc_create": [ctypes.POINTER(_P), _I, ctypes.c_uint],
    "xc_destroy": [_P],
    "xc_set_stream": [_P, _P],
    "xc_use_own_stream": [_P],
    "xc_get_stream": [_P],
    "xc_set_flags": [_P, ctypes.c_uint],
    "xc_sync": [_P],
    "xc_last_error": [_P],
    "xc_malloc": [_P, ctypes.c_size_t, ctypes.POINTER(_P)],
    "xc_free": [_P, _P],
    "xc_scheme_halo": [_I],
    "xc_fill_ghosts": [_P, _M, _I, _I, _I],
    "xc_embed_interior": [_P, _M, _M, _I],
    "xc_advect": [_P, _I, _M, _M, _M, _M, _D],
    "xc_advect_with_flux": [_P, _I, _M, _M, _M, _M, _M, _M, _D, _I, _I],
    "xc_advect_centered": [_P, _I, _M, _PM, _PM, _M, _M, _M, _D, _I, _I],
    "xc_diffus": [_P, _M, _M, _M, _D, _I],
    "xc_diffus_with_flux": [_P, _M, _M, _M, _M, _M, _D, _I, _I, _I],
    "xc_diffus_centered": [_P, _M, _PM, _PM, _M, _M, _D, _I, _I],
    "xc_advect_set_tile": [_P, _I, _I, ctypes.c_int64, ctypes.c_int64, _I, _I],
    "xc_diag_stream": [_P, _I, _P, _I, _P, _I],
    "xc_advect_tile": [_P, _I, _I, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(_I), ctypes.POINTER(_I),
                       ctypes.POINTER(ctypes.c_float)],
}
_RESTYPE = {"xc_last_error": ctypes.c_char_p, "xc_get_stream": _P}
_EXTRA = []  # (name, argtypes, restype) registered by other modules


def register(name, argtypes, restype=ctypes.c_int):
    """Declare one more C entry point (used by the subsystem modules)."""
    _EXTRA.append((name, argtypes, restype))
    if _lib is not None:
        fn = getattr(_lib, name)
        fn.argtypes = argtypes
        fn.restype = restype


def lib():
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f"libxcore_hip.so not found at {LIB_PATH}; build it with "
                        "`python xcore.jl_amd/build.py` (no CPU fallback exists)")
                L = ctypes.CDLL(LIB_PATH)
                for name, args in _PROTOS.items():
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = _RESTYPE.get(name, ctypes.c_int)
                for name, args, res in _EXTRA:
                    fn = getattr(L, name)
                    fn.argtypes = args
                    fn.restype = res
                _lib = L
    return _lib


def declared_symbols():
    return list(_PROTOS) + [n for n, _, _ in _EXTRA]


class Context:
    """An xc_ctx bound to one device; work goes on torch's current stream."""

    def __init__(self, device=0, flags=0):
        self.ptr = _P()
        code = lib().xc_create(ctypes.byref(self.ptr), int(device), int(flags))
        if code != XC_OK:
            raise XCoreError(code, "xc_create: " + lib().xc_last_error(None).decode())
        self.device = device

    def check(self, code, what=""):
        if code == XC_OK:
            return
        msg = lib().xc_last_error(self.ptr).decode(errors="replace")
        if code == XC_EINVAL_SHAPE:
            raise AssertionError(msg)
        if code in (XC_EINVAL_SCHEME, XC_EINVAL_BC):
            raise ValueError(msg)
        if code == XC_ENOCONV:
            raise NoConvergence(code, msg)
        raise XCoreError(code, f"{what}: {msg}")

    def bind_stream(self, stream_handle):
        lib().xc_set_stream(self.ptr, _P(stream_handle))

    flags = 0

    def set_flags(self, flags):
        self.check(lib().xc_set_flags(self.ptr, int(flags)), "xc_set_flags")
        self.flags = int(flags)

    def sync(self):
        self.check(lib().xc_sync(self.ptr), "xc_sync")

    def close(self):
        if self.ptr:
            lib().xc_destroy(self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
