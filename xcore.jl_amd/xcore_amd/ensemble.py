"""Ensembles of grids of any sizes advanced in one launch (no counterpart in the reference).

Each member is exactly the single-grid ``advect_centered[_with_flux]!`` /
``diffus_centered[_with_flux]!`` of phsevo.jl:21-149 on its own arrays; the
tiles of all members form one grid of workgroups.  ``BatchPlan`` keeps the
members' argument table on the device so that ``run()`` is a single kernel
launch -- the form to record in a HIP graph and replay every step.
"""
import ctypes

from . import _lib
from ._lib import BCS, SCHEMES, XCMat, lib
from .arrays import as_mat
from .kernels import _sym, context

_P = ctypes.c_void_p
_I = ctypes.c_int
_NULL = XCMat(None, 0, 0, 0)


class AdvMember(ctypes.Structure):
    _fields_ = [(n, XCMat) for n in ("adv", "qx", "qz", "f", "u", "w")] + [("h", ctypes.c_double)]


class DifMember(ctypes.Structure):
    _fields_ = [(n, XCMat) for n in ("dff", "qx", "qz", "f", "k")] + [("h", ctypes.c_double)]


_lib.register("xc_advect_centered_batch", [_P, _I, _I, _P, _I, _I])
_lib.register("xc_diffus_centered_batch", [_P, _I, _P, _I, _I])
_lib.register("xc_advect_batch_plan", [_P, _I, _I, _P, _I, _I, ctypes.POINTER(_P)])
_lib.register("xc_diffus_batch_plan", [_P, _I, _P, _I, _I, ctypes.POINTER(_P)])
_lib.register("xc_batch_run", [_P, _P])
_lib.register("xc_batch_destroy", [_P, _P])


def _adv_table(members):
    arr = (AdvMember * len(members))()
    for i, m in enumerate(members):
        qx, qz = m.get("qx"), m.get("qz")
        arr[i] = AdvMember(as_mat(m["adv"], "adv"), as_mat(qx, "qx") if qx is not None else _NULL,
                           as_mat(qz, "qz") if qz is not None else _NULL, as_mat(m["f"], "f"), as_mat(m["u"], "u"),
                           as_mat(m["w"], "w"), float(m["h"]))
    return arr


def _dif_table(members):
    arr = (DifMember * len(members))()
    for i, m in enumerate(members):
        qx, qz = m.get("qx"), m.get("qz")
        arr[i] = DifMember(as_mat(m["dff"], "dff"), as_mat(qx, "qx") if qx is not None else _NULL,
                           as_mat(qz, "qz") if qz is not None else _NULL, as_mat(m["f"], "f"), as_mat(m["k"], "k"),
                           float(m["h"]))
    return arr


def advect_centered_batch_(members, scheme, xBC="periodic", zBC="closed"):
    """members: dicts with adv, f, u, w, h and optionally qx, qz (all members or none)."""
    c = context()
    t = _adv_table(members)
    c.check(lib().xc_advect_centered_batch(c.ptr, _sym(scheme, SCHEMES, "scheme"), len(members), t,
                                           _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")), "advect_centered_batch")


def diffus_centered_batch_(members, xBC="periodic", zBC="closed"):
    """members: dicts with dff, f, k, h and optionally qx, qz (all members or none)."""
    c = context()
    t = _dif_table(members)
    c.check(lib().xc_diffus_centered_batch(c.ptr, len(members), t, _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")),
            "diffus_centered_batch")


class BatchPlan:
    """A reusable one-launch ensemble call (advection when scheme is given, else diffusion)."""

    def __init__(self, members, scheme=None, xBC="periodic", zBC="closed"):
        c = context()
        self.ptr = _P()
        self.members = members  # the arrays must outlive the plan
        if scheme is not None:
            t = _adv_table(members)
            c.check(lib().xc_advect_batch_plan(c.ptr, _sym(scheme, SCHEMES, "scheme"), len(members), t,
                                               _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC"), ctypes.byref(self.ptr)),
                    "advect_batch_plan")
        else:
            t = _dif_table(members)
            c.check(lib().xc_diffus_batch_plan(c.ptr, len(members), t, _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC"),
                                               ctypes.byref(self.ptr)), "diffus_batch_plan")

    def run(self):
        c = context()
        c.check(lib().xc_batch_run(c.ptr, self.ptr), "batch_run")

    def close(self):
        if self.ptr:
            c = context()
            c.check(lib().xc_batch_destroy(c.ptr, self.ptr), "batch_destroy")
            self.ptr = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
