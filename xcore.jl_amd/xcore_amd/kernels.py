"""Stencil entry points with the reference's signatures (src/kernels/*.jl).

Julia's mutating ``f!`` becomes ``f_`` here.  Arguments, their order and their
meaning follow the reference; errors map back to its exception types:
size mismatch -> AssertionError, unknown scheme or BC -> ValueError
(the reference's ArgumentError).
"""
import ctypes

import torch

from . import _lib
from ._lib import BCS, SCHEMES, Context, lib
from .arrays import as_mat

_ctxs = {}


def context(device=None):
    """The per-device default xc_ctx; work is enqueued on torch's current stream."""
    dev = torch.cuda.current_device() if device is None else int(device)
    c = _ctxs.get(dev)
    if c is None:
        c = Context(dev, 0)
        _ctxs[dev] = c
    c.bind_stream(torch.cuda.current_stream(dev).cuda_stream)
    return c


class _flag:
    bit = 0

    def __enter__(self):
        self.c = context()
        self.old = self.c.flags
        self.c.set_flags(self.old | self.bit)
        return self

    def __exit__(self, *a):
        self.c.set_flags(self.old)


class exact_arithmetic(_flag):
    """Context manager: run the stencils in the reference's exact operation order (bit-identical)."""
    bit = _lib.XC_CTX_EXACT


class autotune(_flag):
    """Context manager: advection calls at a new (scheme, fluxes, grid) time every tile shape once and keep
    the fastest (the choice persists in the device context)."""
    bit = _lib.XC_CTX_AUTOTUNE


def advect_tile(scheme, Nz, Nx, with_flux=True):
    """(tz, tx, {tile: us}) the advection kernels use at this size; tuned=False if not measured yet."""
    c = context()
    tz, tx = ctypes.c_int(0), ctypes.c_int(0)
    us = (ctypes.c_float * 6)()
    code = lib().xc_advect_tile(c.ptr, _sym(scheme, SCHEMES, "scheme"), int(bool(with_flux)), int(Nz), int(Nx),
                                ctypes.byref(tz), ctypes.byref(tx), us)
    names = ["64x16", "128x8", "64x8", "32x32", "64x16p", "64x8p"]
    cand = {n: round(float(u), 2) for n, u in zip(names, us) if float(u) < 1e29}
    return tz.value, tx.value, cand, code == 0


def _sym(s, table, kind):
    key = str(s).lstrip(":")
    if key not in table:
        raise ValueError(f"unsupported {kind} = :{key}")
    return table[key]


def scheme_halo(scheme):
    """advect.jl:346-355"""
    return lib().xc_scheme_halo(_sym(scheme, SCHEMES, "scheme"))


def fill_ghosts_(f, halo, zBC="closed", xBC="periodic"):
    """ghosts.jl:18-51 ``fill_ghosts!(f, halo; zBC, xBC)``"""
    c = context()
    c.check(lib().xc_fill_ghosts(c.ptr, as_mat(f, "f"), int(halo), _sym(xBC, BCS, "xBC"),
                                 _sym(zBC, BCS, "zBC")), "fill_ghosts!")
    return f


def embed_interior_(f_halo, f, halo):
    """ghosts.jl:61-64 ``embed_interior!(f_halo, f, halo)``"""
    c = context()
    c.check(lib().xc_embed_interior(c.ptr, as_mat(f_halo, "f_halo"), as_mat(f, "f"), int(halo)),
            "embed_interior!")
    return f_halo


def advect_(adv, f_halo, u, w, h, scheme):
    """advect.jl:374-425 ``advect!(adv, f_halo, u, w, h, scheme)``"""
    c = context()
    s = _sym(scheme, SCHEMES, "scheme")
    c.check(lib().xc_advect(c.ptr, s, as_mat(adv, "adv"), as_mat(f_halo, "f_halo"), as_mat(u, "u"),
                            as_mat(w, "w"), float(h)), "advect!")
    return adv


def advect_with_flux_(adv, qx, qz, f_halo, u, w, h, scheme, xBC="periodic", zBC="closed"):
    """advect.jl:717-798 ``advect_with_flux!(adv, qx, qz, f_halo, u, w, h, scheme; xBC, zBC)``"""
    c = context()
    s = _sym(scheme, SCHEMES, "scheme")
    c.check(lib().xc_advect_with_flux(c.ptr, s, as_mat(adv, "adv"), as_mat(qx, "qx"), as_mat(qz, "qz"),
                                      as_mat(f_halo, "f_halo"), as_mat(u, "u"), as_mat(w, "w"), float(h),
                                      _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")), "advect_with_flux!")
    return adv


def advect_centered_(adv, f, u, w, h, scheme, xBC="periodic", zBC="closed"):
    """phsevo.jl:21-59 ``advect_centered!(adv, f, u, w, h, scheme; xBC, zBC)``"""
    c = context()
    s = _sym(scheme, SCHEMES, "scheme")
    c.check(lib().xc_advect_centered(c.ptr, s, as_mat(adv, "adv"), None, None, as_mat(f, "f"),
                                     as_mat(u, "u"), as_mat(w, "w"), float(h),
                                     _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")), "advect_centered!")
    return adv


def advect_centered_with_flux_(adv, qx, qz, f, u, w, h, scheme, xBC="periodic", zBC="closed"):
    """phsevo.jl:90-128 ``advect_centered_with_flux!(adv, qx, qz, f, u, w, h, scheme; xBC, zBC)``"""
    c = context()
    s = _sym(scheme, SCHEMES, "scheme")
    mqx, mqz = as_mat(qx, "qx"), as_mat(qz, "qz")
    c.check(lib().xc_advect_centered(c.ptr, s, as_mat(adv, "adv"), ctypes.byref(mqx), ctypes.byref(mqz),
                                     as_mat(f, "f"), as_mat(u, "u"), as_mat(w, "w"), float(h),
                                     _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")),
            "advect_centered_with_flux!")
    return adv


def diffus_(dff, f_halo, k_halo, h, halo=1):
    """diffus.jl:45-57 ``diffus!(dff, f_halo, k_halo, h; halo)``"""
    c = context()
    c.check(lib().xc_diffus(c.ptr, as_mat(dff, "dff"), as_mat(f_halo, "f_halo"), as_mat(k_halo, "k_halo"),
                            float(h), int(halo)), "diffus!")
    return dff


def diffus_with_flux_(dff, qx, qz, f_halo, k_halo, h, halo=1, xBC="periodic", zBC="closed"):
    """diffus.jl:105-147 ``diffus_with_flux!(dff, qx, qz, f_halo, k_halo, h; halo, xBC, zBC)``"""
    c = context()
    c.check(lib().xc_diffus_with_flux(c.ptr, as_mat(dff, "dff"), as_mat(qx, "qx"), as_mat(qz, "qz"),
                                      as_mat(f_halo, "f_halo"), as_mat(k_halo, "k_halo"), float(h), int(halo),
                                      _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")), "diffus_with_flux!")
    return dff


def diffus_centered_(dff, f, k, h, xBC="periodic", zBC="closed"):
    """phsevo.jl:67-79 ``diffus_centered!(dff, f, k, h; xBC, zBC)``"""
    c = context()
    c.check(lib().xc_diffus_centered(c.ptr, as_mat(dff, "dff"), None, None, as_mat(f, "f"), as_mat(k, "k"),
                                     float(h), _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")),
            "diffus_centered!")
    return dff


def diffus_centered_with_flux_(dff, qx, qz, f, k, h, xBC="periodic", zBC="closed"):
    """phsevo.jl:135-149 ``diffus_centered_with_flux!(dff, qx, qz, f, k, h; xBC, zBC)``"""
    c = context()
    mqx, mqz = as_mat(qx, "qx"), as_mat(qz, "qz")
    c.check(lib().xc_diffus_centered(c.ptr, as_mat(dff, "dff"), ctypes.byref(mqx), ctypes.byref(mqz),
                                     as_mat(f, "f"), as_mat(k, "k"), float(h),
                                     _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")),
            "diffus_centered_with_flux!")
    return dff


def set_advect_tile(scheme, Nz, Nx, tz, tx, with_flux=True):
    """Pin the advection tile shape for this (scheme, fluxes, grid): one of 64x16, 128x8, 64x8, 32x32, or
    (64, -16) / (64, -8) for the persistent pipelined kernels."""
    c = context()
    c.check(lib().xc_advect_set_tile(c.ptr, _sym(scheme, SCHEMES, "scheme"), int(bool(with_flux)), int(Nz), int(Nx),
                                     int(tz), int(tx)), "advect_set_tile")


def diag_stream_(inputs, outputs, nts=False):
    """Diagnostic (no reference counterpart): the stencil traffic without its arithmetic --
    outputs[k][i + (k > 0), j] = sum of inputs[r][i, j] over the cells of outputs[0]'s shape, with
    non-temporal stores if nts bit 0 is set; nts bit 1: every stream aligned and unshifted, 16-byte accesses;
    nts bit 2: non-temporal loads.
    Timing it gives the bandwidth HBM sustains for that stream mix."""
    c = context()
    ins = (_lib.XCMat * max(1, len(inputs)))(*[as_mat(t, "in") for t in inputs])
    outs = (_lib.XCMat * len(outputs))(*[as_mat(t, "out") for t in outputs])
    c.check(lib().xc_diag_stream(c.ptr, len(inputs), ins, len(outputs), outs, int(nts)), "diag_stream")
