"""Field updates with the reference's signatures (phsevo.jl, update.jl, store.jl, timing.jl)."""
import ctypes

from . import _lib
from ._lib import BCS, SCHEMES, lib
from .kernels import _sym, context
from .state import grid_struct, par_struct, scales_struct

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_V = ctypes.c_void_p  # struct pointers are passed as byref(...)

_lib.register("xc_phsevo", [_P, _V, _V, _V, _V, _V, _I, _I, _I, ctypes.POINTER(_D), _D, _D])
_lib.register("xc_update", [_P, _V, _V, _V, _V, _V, _I, _I])
_lib.register("xc_update_dt", [_P, _V, _V, _V, _V, _D, ctypes.POINTER(_D)])
_lib.register("xc_store_previous", [_P, _V, _V])
_lib.register("xc_update_phase_velocities", [_P, _V, _V, _V, _V, _V, _I, _I])
_lib.register("xc_time_coefs", [_I, ctypes.c_int64, ctypes.POINTER(_D)])
_lib.register("xc_noise_step", [_P, _V, _V, _V, _V, _V, _D, _I])
_lib.register("xc_store_noise", [_P, _V])
_lib.register("xc_normal_fill", [_P, _lib.XCMat, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32])
_lib.register("xc_fft_filter", [_P, _lib.XCMat, _lib.XCMat, ctypes.c_int64, _D, _D])

TINTS = {"be1im": 0, "bd2im": 1, "cn2si": 2, "bd2si": 3}


def _r(s):
    return ctypes.byref(s)


def time_coefs(TINT, step):
    """timing.jl:18-36 ``time_coefs(T, TINT, step) -> (a1, a2, a3, b1, b2, b3)``"""
    key = str(TINT).lstrip(":")
    if key not in TINTS:
        raise ValueError(f"time_coefs: unsupported TINT scheme {TINT} (use :be1im, :bd2im, :cn2si, or :bd2si)")
    out = (_D * 6)()
    code = lib().xc_time_coefs(TINTS[key], int(step), out)
    if code != 0:
        raise ValueError(f"time_coefs: {TINT}")
    return tuple(out)


def phsevo_(phase, fluid, grid, par, scales, ADVN="weno5", xBC="periodic", zBC="closed",
            a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0, dt=1.0, alpha=None):
    """phsevo.jl:164-214 ``phsevo!(phase, fluid, grid, par, scales; ADVN, xBC, zBC, a1..b3, dt, alpha)``"""
    c = context()
    alpha = par.alpha if alpha is None else alpha
    coef = (_D * 6)(a1, a2, a3, b1, b2, b3)
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_phsevo(c.ptr, _r(ps), _r(fs), _r(grid_struct(grid)), _r(par_struct(par)),
                            _r(scales_struct(scales)), _sym(ADVN, SCHEMES, "scheme"), _sym(xBC, BCS, "xBC"),
                            _sym(zBC, BCS, "zBC"), coef, float(dt), float(alpha)), "phsevo!")


def update_(phase, fluid, grid, par, scales, xBC="periodic", zBC="closed"):
    """update.jl:525-538 ``update!(phase, fluid, grid, par, scales; xBC, zBC)``"""
    c = context()
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_update(c.ptr, _r(ps), _r(fs), _r(grid_struct(grid)), _r(par_struct(par)),
                            _r(scales_struct(scales)), _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")), "update!")


def update_dt(phase, fluid, grid, par, dt_prev):
    """update.jl:494-502 ``update_dt(phase, fluid, grid, par, dt_prev) -> dt``"""
    c = context()
    out = _D(0.0)
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_update_dt(c.ptr, _r(ps), _r(fs), _r(grid_struct(grid)), _r(par_struct(par)),
                               float(dt_prev), ctypes.byref(out)), "update_dt")
    return out.value


def store_previous_(fluid, phase):
    """store.jl:13-30 ``store_previous!(fluid, phase)``"""
    c = context()
    fs, ps = fluid.pack(), phase.pack()
    c.check(lib().xc_store_previous(c.ptr, _r(fs), _r(ps)), "store_previous!")


def update_phase_velocities_(phase, fluid, ns, grid, par, xBC="periodic", zBC="closed"):
    """store.jl:80-105 ``update_phase_velocities!(phase, fluid, ns, grid, par; xBC, zBC)``"""
    c = context()
    ps, fs, nss = phase.pack(), fluid.pack(), ns.pack()
    c.check(lib().xc_update_phase_velocities(c.ptr, _r(ps), _r(fs), _r(nss), _r(grid_struct(grid)),
                                             _r(par_struct(par)), _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")),
            "update_phase_velocities!")


def noise_(ns, phase, grid, par, scales, dt, first_iter=True):
    """noise.jl:138-239 ``noise!(ns, phase, grid, par, scales, dt; first_iter)``"""
    c = context()
    nss, ps = ns.pack(), phase.pack()
    c.check(lib().xc_noise_step(c.ptr, _r(nss), _r(ps), _r(grid_struct(grid)), _r(par_struct(par)),
                                _r(scales_struct(scales)), float(dt), int(bool(first_iter))), "noise!")
    ns.draws = int(nss.draws)


def store_noise_(ns):
    """noise.jl:115-120 ``store_noise!(ns)``"""
    c = context()
    nss = ns.pack()
    c.check(lib().xc_store_noise(c.ptr, _r(nss)), "store_noise!")


def normal_fill_(a, seed, stream=0, draw=0):
    """Standard normals from the Philox4x32-10 stream (seed, stream, draw)."""
    from .arrays import as_mat
    c = context()
    c.check(lib().xc_normal_fill(c.ptr, as_mat(a, "a"), int(seed), int(stream), int(draw)), "normal_fill")
    return a


def compute_fft_filter(psi, pad, h, sigma):
    """noise.jl:244-270 ``compute_fft_filter(psi, Gk, pad)`` with Gk = exp(-sigma^2 |k|^2)."""
    from .arrays import as_mat, xcore_zeros
    c = context()
    out = xcore_zeros(psi.shape[0], psi.shape[1], psi.device)
    c.check(lib().xc_fft_filter(c.ptr, as_mat(psi, "psi"), as_mat(out, "out"), int(pad), float(h), float(sigma)),
            "compute_fft_filter")
    return out
