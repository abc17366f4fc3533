"""Time loop with the reference's signatures: initialize! (run.jl:71-140), run!
(run.jl:167-277) and the history sums of record_history! (history.jl:101-181).

One outer step runs inside the library (xc_run_step); this module only keeps
the host-side bookkeeping the reference keeps in Julia (time, dt, step,
History vectors).
"""
import ctypes
import math
from dataclasses import dataclass, field

import torch

from . import _lib
from ._lib import BCS, SCHEMES, lib
from .fields import TINTS, noise_, normal_fill_, store_noise_, update_, update_phase_velocities_
from .kernels import _sym, context
from .state import grid_struct, par_struct, scales_struct
from .stokes import FmOpts, fm_opts

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
EPS = 2.220446049250313e-16


class StepOpts(ctypes.Structure):
    _fields_ = [("tint", _I), ("scheme", _I), ("xbc", _I), ("zbc", _I), ("rtol", _D), ("atol", _D),
                ("maxit", _I), ("fixed_sweeps", _I), ("fm", FmOpts)]


class StepState(ctypes.Structure):
    _fields_ = [("time", _D), ("dt", _D), ("step", ctypes.c_int64)]


class StepStats(ctypes.Structure):
    _fields_ = [("sweeps", _I), ("resnorm", _D), ("stokes_iters", _I), ("stokes_worst_rel_resid", _D),
                ("stokes_fallbacks", _I), ("stokes_mg_levels", _I), ("stokes_inertia_ratio", _D),
                ("stokes_iters_first", _I), ("stokes_init_rel_first", _D), ("stokes_init_rel_later", _D)]


_lib.register("xc_run_step", [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("xc_history_sums", [_P, _P, _P, ctypes.POINTER(_D)])
_lib.register("xc_snapshot", [_P, _P, _P])
_lib.register("xc_compute_resnorm", [_P, _P, _P, _D, ctypes.POINTER(_D)])


def step_opts(par, ADVN=None, xBC="periodic", zBC="closed", fixed_sweeps=0, sds=-1, top_cnv=1, bot_cnv=1,
              open_cnv=False, **fmkw):
    o = StepOpts()
    o.tint = TINTS[str(par.TINT).lstrip(":")]
    o.scheme = _sym(ADVN or par.ADVN, SCHEMES, "scheme")
    o.xbc, o.zbc = _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC")
    o.rtol, o.atol, o.maxit = float(par.rtol), float(par.atol), int(par.maxit)
    o.fixed_sweeps = int(fixed_sweeps)
    o.fm = fm_opts(par, sds=sds, top_cnv=top_cnv, bot_cnv=bot_cnv, open_cnv=open_cnv, xBC=xBC, zBC=zBC, **fmkw)
    return o


@dataclass
class History:
    """history.jl:21-93 (the conservation part of it)."""
    time: list = field(default_factory=list)
    dt: list = field(default_factory=list)
    sumB: list = field(default_factory=list)
    sumM: list = field(default_factory=list)
    sumX: list = field(default_factory=list)
    DB: list = field(default_factory=list)
    DM: list = field(default_factory=list)
    DX: list = field(default_factory=list)
    EB: list = field(default_factory=list)
    EM: list = field(default_factory=list)
    EX: list = field(default_factory=list)
    x_mean: list = field(default_factory=list)
    dsumBdt: float = 0.0
    dsumBdto: float = 0.0
    dsumBdtoo: float = 0.0
    dsumMdt: float = 0.0
    dsumMdto: float = 0.0
    dsumMdtoo: float = 0.0
    dsumXdt: float = 0.0
    dsumXdto: float = 0.0
    dsumXdtoo: float = 0.0


def history_sums(phase, fluid):
    c = context()
    out = (_D * 12)()
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_history_sums(c.ptr, ctypes.byref(ps), ctypes.byref(fs), out), "record_history!")
    return list(out)


def record_history_(hst, time, dt, phase, fluid, ns, grid, a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0):
    """history.jl:101-181 (mass sums, boundary-flux drift, conservation errors EB/EM/EX)."""
    h = grid.h
    h2 = h * h
    s = history_sums(phase, fluid)
    hst.dsumBdtoo, hst.dsumBdto = hst.dsumBdto, hst.dsumBdt
    hst.dsumMdtoo, hst.dsumMdto = hst.dsumMdto, hst.dsumMdt
    hst.dsumXdtoo, hst.dsumXdto = hst.dsumXdto, hst.dsumXdt
    hst.time.append(time)
    hst.dt.append(dt)
    sumB, sumM, sumX = s[0] * h2 + EPS, s[1] * h2 + EPS, s[2] * h2 + EPS
    hst.sumB.append(sumB); hst.sumM.append(sumM); hst.sumX.append(sumX)
    top_aX, bot_aX, top_aM, bot_aM, top_dX, bot_dX, top_dM, bot_dM = (v * h for v in s[4:12])
    sumGx = s[3] * h2
    dB = (top_aX - bot_aX) + (top_aM - bot_aM) + (top_dX - bot_dX) + (top_dM - bot_dM)
    dM = -sumGx + (top_aM - bot_aM) + (top_dM - bot_dM)
    dX = +sumGx + (top_aX - bot_aX) + (top_dX - bot_dX)
    hst.dsumBdt, hst.dsumMdt, hst.dsumXdt = dB, dM, dX
    stp = len(hst.time)
    if stp >= 2:
        def drift(D, new, o, oo):
            return (a2 * D[stp - 2] + a3 * D[max(0, stp - 3)] + (b1 * new + b2 * o + b3 * oo) * dt) / a1
        hst.DB.append(drift(hst.DB, dB, hst.dsumBdto, hst.dsumBdtoo))
        hst.DM.append(drift(hst.DM, dM, hst.dsumMdto, hst.dsumMdtoo))
        hst.DX.append(drift(hst.DX, dX, hst.dsumXdto, hst.dsumXdtoo))
    else:
        hst.DB.append(0.0); hst.DM.append(0.0); hst.DX.append(0.0)
    sumB0 = hst.sumB[0]
    hst.EB.append((sumB - hst.DB[-1] - sumB0) / sumB0)
    hst.EM.append((sumM - hst.DM[-1] - hst.sumM[0]) / sumB0)
    hst.EX.append((sumX - hst.DX[-1] - hst.sumX[0]) / sumB0)
    hst.x_mean.append(float(phase.x.mean()))


def initialize_(phase, fluid, ns, hst, grid, par, scales, xBC="periodic", zBC="closed", x_init=None):
    """run.jl:71-140 ``initialize!`` (fresh start; a restart loads a checkpoint with ``load_checkpoint_``).

    The random perturbation ``rp`` comes from the device generator (seed par.seed, stream 2) in place of
    Julia's MersenneTwister.  ``x_init`` (a device matrix) replaces the perturbed profile when given.
    """
    Nz, Nx, h, D, L = grid.Nz, grid.Nx, grid.h, grid.D, grid.L
    dev = phase.x.device
    dd = torch.float64
    zf = torch.arange(Nz + 1, dtype=dd, device=dev) * h
    osgr = 1.0 if par.open_sgr else 0.0
    taper = 1.0 - (torch.exp(-zf / scales.l0h) + torch.exp(-(D - zf) / scales.l0h)) * (1.0 - osgr)
    phase.bndtaperw.copy_(taper[:, None].expand(Nz + 1, Nx + 2))
    zc = (torch.arange(Nz, dtype=dd, device=dev) + 0.5) * h
    phase.bndshape.copy_(torch.exp((-zc + h / 2) / scales.bnd_w)[:, None].expand(Nz, Nx))
    if x_init is None:
        rp = torch.zeros((Nx, Nz), dtype=dd, device=dev).t()
        normal_fill_(rp, par.seed, stream=2, draw=0)
        rp = (rp - rp.mean()) / rp.std()
        xcn = (torch.arange(Nx, dtype=dd, device=dev) + 0.5) * h / L
        zcn = (torch.arange(Nz, dtype=dd, device=dev) + 0.5) * h / D
        gp = torch.exp(-((zcn[:, None] - 0.5) / 0.125) ** 2) * torch.exp(-((xcn[None, :] - 0.5) / 0.125) ** 2)
        xin = par.x0 + (par.Da - par.x0) * phase.bndshape
        phase.x.copy_(xin * (1.0 + par.dxr * rp + par.dxg * gp))
    else:
        phase.x.copy_(x_init)
    phase.m.copy_(1.0 - phase.x)
    fluid.eta.fill_(par.etam0)
    fluid.etaco.fill_(par.etam0)
    phase.etas.fill_(par.etam0)
    update_(phase, fluid, grid, par, scales, xBC=xBC, zBC=zBC)
    phase.X.copy_(fluid.rho * phase.x)
    phase.M.copy_(fluid.rho * phase.m)
    noise_(ns, phase, grid, par, scales, scales.dt0, first_iter=True)
    update_phase_velocities_(phase, fluid, ns, grid, par, xBC=xBC, zBC=zBC)
    phase.Xo.copy_(phase.X); phase.Xoo.copy_(phase.X)
    phase.Mo.copy_(phase.M); phase.Moo.copy_(phase.M)
    fluid.rhoo.copy_(fluid.rho); fluid.rhooo.copy_(fluid.rho)
    store_noise_(ns)
    if hst is not None:
        record_history_(hst, 0.0, scales.dt0, phase, fluid, ns, grid)
    return 0.0, scales.dt0, 0


def run_step_(phase, fluid, ns, grid, par, scales, state, opts):
    """One outer step of run! (run.jl:205-274) on the device; ``state`` is a StepState."""
    c = context()
    ps, fs, nss = phase.pack(), fluid.pack(), ns.pack()
    st = StepStats()
    code = lib().xc_run_step(c.ptr, ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss),
                             ctypes.byref(grid_struct(grid)), ctypes.byref(par_struct(par)),
                             ctypes.byref(scales_struct(scales)), ctypes.byref(opts), ctypes.byref(state),
                             ctypes.byref(st))
    ns.draws = int(nss.draws)
    c.check(code, "run!")
    return {k: getattr(st, k) for k, _ in StepStats._fields_}


def run_(phase, fluid, ns, hst, grid, par, scales, nsteps=None, time=0.0, dt=None, step=0, ADVN=None,
         xBC="periodic", zBC="closed", sds=-1, top_cnv=1, bot_cnv=1, open_cnv=False, verbose=False, **fmkw):
    """run.jl:167-277 ``run!`` (history every par.nrh steps; figures are out of scope, checkpoints are
    ``save_checkpoint`` / ``load_checkpoint_`` between calls)."""
    nsteps = nsteps if nsteps is not None else 10 ** 9
    state = StepState(float(time), float(scales.dt0 if dt is None else dt), int(step))
    opts = step_opts(par, ADVN=ADVN, xBC=xBC, zBC=zBC, sds=sds, top_cnv=top_cnv, bot_cnv=bot_cnv,
                     open_cnv=open_cnv, **fmkw)
    log = []
    while state.step < nsteps:
        dt_used = state.dt
        stats = run_step_(phase, fluid, ns, grid, par, scales, state, opts)
        log.append(stats)
        if hst is not None and state.step % par.nrh == 0:
            from .fields import time_coefs
            a1, a2, a3, b1, b2, b3 = time_coefs(par.TINT, state.step)
            record_history_(hst, state.time, dt_used, phase, fluid, ns, grid, a1, a2, a3, b1, b2, b3)
        if verbose:
            print(f"step {state.step}: t={state.time:.4e} dt={dt_used:.4e} {stats}")
    return state.time, state.dt, log
