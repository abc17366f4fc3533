"""Stokes solve with the reference's signature (fluidmech.jl:71-80).

``fluidmech_(state, grid, par; phase, bnchm_data, sds, top_cnv, bot_cnv, open_cnv,
xBC, zBC, dt, a1..b3)`` updates W, U, P in place like the reference; the direct
UMFPACK solve becomes a matrix-free FGMRES on the HIP device (tolerance, iteration
cap and multigrid controls are extra keywords).  Returns the solve statistics.
"""
import ctypes

import torch

from . import _lib
from ._lib import BCS, SCHEMES, XCMat, lib
from .arrays import as_mat
from .kernels import _sym, context
from .state import grid_struct, par_struct

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double


class FmOpts(ctypes.Structure):
    _fields_ = [(n, _I) for n in ("sds", "top_cnv", "bot_cnv", "open_cnv", "xbc", "zbc", "scheme")] + \
               [(n, _D) for n in ("dt", "a1", "a2", "a3", "b1", "b2", "b3", "tol")] + \
               [(n, _I) for n in ("maxit", "restart", "mg_pre", "mg_post", "mg_coarse")] + [("mg_omega", _D)] + \
               [("learned", _I), ("lc_level", _I), ("mg_sub", _I), ("mg_levels", _I), ("mg_fp64", _I),
                ("xmean", _I), ("mg_jacobi", _I)]


class BnchmStruct(ctypes.Structure):
    _fields_ = [(n, XCMat) for n in ("src_W", "src_U", "src_P", "W_exact", "U_exact", "P_exact")]


class SolveStats(ctypes.Structure):
    _fields_ = [("iters", _I), ("converged", _I), ("rel_resid", _D), ("init_rel_resid", _D), ("mg_levels", _I),
                ("fallback_used", _I), ("inertia_ratio", _D), ("learned_steps", _I)]


_lib.register("xc_fluidmech", [_P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("xc_stokes_apply", [_P, _P, _P, _P, _P, _I, _P, _P])
_lib.register("xc_stokes_rhs", [_P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("xc_stokes_precond", [_P, _P, _P, _P, _P, _I, _P, _P])
_lib.register("xc_stokes_solve", [_P, _P, _P, _P, _P, _I, _P, _P, _P])
_lib.register("xc_stokes_lc_restrict", [_P, _P, _P, _P, _I, _P, _P, _P, _I])
_lib.register("xc_stokes_krylov_vector", [_P, _I, _P])
_lib.register("xc_stokes_lc_prolong", [_P, _P, _P, _I, _P, _P, _P])


class BnchmData:
    """state.jl:131-134 ``BnchmData``: MMS source terms and exact solutions (device matrices)."""

    def __init__(self, src_W, src_U, src_P, W_exact, U_exact, P_exact):
        self.src_W, self.src_U, self.src_P = src_W, src_U, src_P
        self.W_exact, self.U_exact, self.P_exact = W_exact, U_exact, P_exact

    def pack(self):
        return BnchmStruct(*(as_mat(getattr(self, n), n) for n in
                             ("src_W", "src_U", "src_P", "W_exact", "U_exact", "P_exact")))


def fm_opts(par, sds=-1, top_cnv=1, bot_cnv=1, open_cnv=False, xBC="periodic", zBC="closed", dt=1e32,
            a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0, tol=1e-8, maxit=1000, restart=40,
            mg_pre=3, mg_post=3, mg_coarse=0, mg_omega=0.0, learned=False, lc_level=2, mg_sub=1, mg_levels=0,
            mg_fp64=False, xmean=0, mg_jacobi=False):
    return FmOpts(int(sds), int(top_cnv), int(bot_cnv), int(bool(open_cnv)), _sym(xBC, BCS, "xBC"),
                  _sym(zBC, BCS, "zBC"), _sym(par.ADVN, SCHEMES, "scheme"), float(dt), float(a1), float(a2),
                  float(a3), float(b1), float(b2), float(b3), float(tol), int(maxit), int(restart), int(mg_pre),
                  int(mg_post), int(mg_coarse), float(mg_omega), int(learned), int(lc_level), int(mg_sub),
                  int(mg_levels), int(bool(mg_fp64)), int(xmean), int(bool(mg_jacobi)))


def fluidmech_(state, grid, par, phase=None, bnchm_data=None, raise_on_noconv=True, **kw):
    """fluidmech.jl:71-534 ``fluidmech!(state, grid, par; phase, bnchm_data, ...)``."""
    c = context()
    o = fm_opts(par, **kw)
    fs = state.pack()
    ps = phase.pack() if phase is not None else None
    bs = bnchm_data.pack() if bnchm_data is not None else None
    st = SolveStats()
    code = lib().xc_fluidmech(c.ptr, ctypes.byref(fs), ctypes.byref(ps) if ps is not None else None,
                              ctypes.byref(grid_struct(grid)), ctypes.byref(par_struct(par)), ctypes.byref(o),
                              ctypes.byref(bs) if bs is not None else None, ctypes.byref(st))
    stats = {k: getattr(st, k) for k, _ in SolveStats._fields_}
    if code == _lib.XC_ENOCONV and not raise_on_noconv:
        return stats
    c.check(code, "fluidmech!")
    return stats


def stokes_vector_size(grid):
    Nz, Nx = grid.Nz, grid.Nx
    return (Nz + 1) * (Nx + 2) + (Nz + 2) * (Nx + 1) + (Nz + 2) * (Nx + 2)


def stokes_apply(state, grid, par, x, bnchm=False, **kw):
    """y = LL x with the matrix-free operator (x: 1-D float64 device tensor)."""
    c = context()
    o = fm_opts(par, **kw)
    y = torch.empty_like(x)
    fs = state.pack()
    c.check(lib().xc_stokes_apply(c.ptr, ctypes.byref(fs), ctypes.byref(grid_struct(grid)),
                                  ctypes.byref(par_struct(par)), ctypes.byref(o), int(bool(bnchm)),
                                  _P(x.data_ptr()), _P(y.data_ptr())), "stokes_apply")
    return y


def stokes_rhs(state, grid, par, phase=None, bnchm_data=None, **kw):
    """RR of fluidmech.jl (performs the MFS update when phase is given)."""
    c = context()
    o = fm_opts(par, **kw)
    rr = torch.empty(stokes_vector_size(grid), dtype=torch.float64, device="cuda")
    fs = state.pack()
    ps = phase.pack() if phase is not None else None
    bs = bnchm_data.pack() if bnchm_data is not None else None
    c.check(lib().xc_stokes_rhs(c.ptr, ctypes.byref(fs), ctypes.byref(ps) if ps is not None else None,
                                ctypes.byref(grid_struct(grid)), ctypes.byref(par_struct(par)), ctypes.byref(o),
                                ctypes.byref(bs) if bs is not None else None, _P(rr.data_ptr())), "stokes_rhs")
    return rr


def stokes_precond(state, grid, par, v, bnchm=False, **kw):
    """z = M^-1 v with the plain block preconditioner (no reference counterpart)."""
    c = context()
    o = fm_opts(par, **kw)
    z = torch.empty_like(v)
    fs = state.pack()
    c.check(lib().xc_stokes_precond(c.ptr, ctypes.byref(fs), ctypes.byref(grid_struct(grid)),
                                    ctypes.byref(par_struct(par)), ctypes.byref(o), int(bool(bnchm)),
                                    _P(v.data_ptr()), _P(z.data_ptr())), "stokes_precond")
    return z


def stokes_solve(state, grid, par, rhs, x=None, bnchm=False, raise_on_noconv=True, **kw):
    """FGMRES on LL x = rhs for any right-hand side (no reference counterpart); returns (x, stats)."""
    c = context()
    o = fm_opts(par, **kw)
    x = torch.zeros_like(rhs) if x is None else x
    fs = state.pack()
    st = SolveStats()
    code = lib().xc_stokes_solve(c.ptr, ctypes.byref(fs), ctypes.byref(grid_struct(grid)),
                                 ctypes.byref(par_struct(par)), ctypes.byref(o), int(bool(bnchm)),
                                 _P(rhs.data_ptr()), _P(x.data_ptr()), ctypes.byref(st))
    stats = {k: getattr(st, k) for k, _ in SolveStats._fields_}
    if not (code == _lib.XC_ENOCONV and not raise_on_noconv):
        c.check(code, "stokes_solve")
    return x, stats


def lc_restrict(state, grid, par, level, vec, sigma=None):
    """The learned correction's input map of a stacked vector: (out (1, Wc, Hc, 4) float32, sigma (1,))."""
    c = context()
    f = 1 << level
    Hc, Wc = grid.Nz // f, grid.Nx // f
    out = torch.empty((1, Wc, Hc, 4), dtype=torch.float32, device=vec.device)
    compute = sigma is None
    sg = torch.empty(1, dtype=torch.float32, device=vec.device) if compute else sigma
    fs = state.pack()
    c.check(lib().xc_stokes_lc_restrict(c.ptr, ctypes.byref(fs), ctypes.byref(grid_struct(grid)),
                                        ctypes.byref(par_struct(par)), int(level), _P(vec.data_ptr()),
                                        _P(out.data_ptr()), _P(sg.data_ptr()), int(compute)), "stokes_lc_restrict")
    return out, sg


def krylov_vector(grid, j):
    """Basis vector j of the last FGMRES cycle (no reference counterpart)."""
    c = context()
    out = torch.empty(stokes_vector_size(grid), dtype=torch.float64, device="cuda")
    c.check(lib().xc_stokes_krylov_vector(c.ptr, int(j), _P(out.data_ptr())), "stokes_krylov_vector")
    return out


def lc_prolong_(z, state, grid, level, e3, sigma):
    """z += P e3 (the learned correction's prolongation; no reference counterpart)."""
    c = context()
    fs = state.pack()
    c.check(lib().xc_stokes_lc_prolong(c.ptr, ctypes.byref(fs), ctypes.byref(grid_struct(grid)), int(level),
                                       _P(e3.data_ptr()), _P(sigma.data_ptr()), _P(z.data_ptr())), "stokes_lc_prolong")
    return z
