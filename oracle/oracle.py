"""ctypes front-end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module; it checks the HIP path and is never the thing measured or
shipped.  Arrays are numpy float64 in Fortran (column-major) order, i.e. the
memory layout of a Julia ``Matrix{Float64}`` with z as the fast axis.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SCHEMES = {"centr": 0, "upwd1": 1, "quick": 2, "fromm": 3, "weno3": 4, "weno5": 5, "tvdim": 6}
BCS = {"periodic": 0, "closed": 1}
ERRORS = {1: "shape", 2: "scheme", 3: "bc", 5: "noconv", 6: "arg"}


class OMat(ctypes.Structure):
    _fields_ = [("p", ctypes.POINTER(ctypes.c_double)), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


class OracleError(RuntimeError):
    def __init__(self, code, fn):
        super().__init__(f"oracle {fn}: status {code} ({ERRORS.get(code, '?')})")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def zeros(rows, cols):
    return np.zeros((rows, cols), dtype=np.float64, order="F")


def om(a):
    """Wrap a column-major float64 array (or a column-major view) as OMat."""
    if a.dtype != np.float64:
        raise TypeError("oracle arrays are float64")
    if a.ndim != 2:
        raise ValueError("oracle arrays are 2-D")
    s0, s1 = a.strides
    if a.shape[0] > 1 and s0 != 8:
        raise ValueError("oracle arrays must be column-major (z fastest)")
    ld = s1 // 8 if a.shape[1] > 1 else max(a.shape[0], 1)
    ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return OMat(ptr, a.shape[0], a.shape[1], ld)


def _chk(code, fn):
    if code != 0:
        raise OracleError(code, fn)


def set_threads(n):
    lib().or_set_threads(int(n))


def scheme_halo(scheme):
    return lib().or_scheme_halo(SCHEMES[scheme])


def fill_ghosts(f, halo, xbc="periodic", zbc="closed"):
    _chk(lib().or_fill_ghosts(om(f), int(halo), BCS.get(xbc, 9), BCS.get(zbc, 9)), "fill_ghosts")
    return f


def embed_interior(fh, f, halo):
    _chk(lib().or_embed_interior(om(fh), om(f), int(halo)), "embed_interior")
    return fh


def _h(h):
    return ctypes.c_double(float(h))


def advect(adv, fh, u, w, h, scheme):
    _chk(lib().or_advect(om(adv), om(fh), om(u), om(w), _h(h), SCHEMES.get(scheme, 99)), "advect")
    return adv


def advect_with_flux(adv, qx, qz, fh, u, w, h, scheme, xbc="periodic", zbc="closed"):
    _chk(lib().or_advect_with_flux(om(adv), om(qx), om(qz), om(fh), om(u), om(w), _h(h),
                                   SCHEMES.get(scheme, 99), BCS.get(xbc, 9), BCS.get(zbc, 9)),
         "advect_with_flux")
    return adv


def advect_centered(adv, f, u, w, h, scheme, xbc="periodic", zbc="closed", qx=None, qz=None):
    pqx = ctypes.byref(om(qx)) if qx is not None else None
    pqz = ctypes.byref(om(qz)) if qz is not None else None
    _chk(lib().or_advect_centered(om(adv), pqx, pqz, om(f), om(u), om(w), _h(h),
                                  SCHEMES.get(scheme, 99), BCS.get(xbc, 9), BCS.get(zbc, 9)),
         "advect_centered")
    return adv


def diffus(dff, fh, kh, h, halo=1):
    _chk(lib().or_diffus(om(dff), om(fh), om(kh), _h(h), int(halo)), "diffus")
    return dff


def diffus_with_flux(dff, qx, qz, fh, kh, h, halo=1, xbc="periodic", zbc="closed"):
    _chk(lib().or_diffus_with_flux(om(dff), om(qx), om(qz), om(fh), om(kh), _h(h), int(halo),
                                   BCS.get(xbc, 9), BCS.get(zbc, 9)), "diffus_with_flux")
    return dff


def diffus_centered(dff, f, k, h, xbc="periodic", zbc="closed", qx=None, qz=None):
    pqx = ctypes.byref(om(qx)) if qx is not None else None
    pqz = ctypes.byref(om(qz)) if qz is not None else None
    _chk(lib().or_diffus_centered(om(dff), pqx, pqz, om(f), om(k), _h(h),
                                  BCS.get(xbc, 9), BCS.get(zbc, 9)), "diffus_centered")
    return dff
