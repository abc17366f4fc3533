"""Time loop with the reference's signatures: initialize! (run.jl:71-140), run!
(run.jl:167-277) and record_history! (history.jl:101-222).

One outer step runs inside the library (xc_run_step); this module only keeps
the host-side bookkeeping the reference keeps in Julia (time, dt, step,
History vectors).
"""
import ctypes
import math

from . import _lib
from ._lib import BCS, SCHEMES, lib
from .fields import TINTS
from .kernels import _sym, context
from .arrays import as_mat
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
                ("stokes_iters_first", _I), ("stokes_init_rel_first", _D), ("stokes_init_rel_later", _D),
                ("stokes_learned_steps", _I)]


# history.jl:184-221: the statistics record_history! keeps, in the order of xc_history_stats
HIST_FIELDS = ("x", "V", "vx", "vm", "xie", "xix", "xis", "Ra", "Rc", "ReD", "Red", "ks", "ke", "kx", "eta",
               "etamix", "etae", "etas", "etat")
# what each field records: x min/mean/max/std; speeds min/rms/max; the rest min/geometric mean/max
HIST_KIND = {"x": "std", "V": "rms", "vx": "rms", "vm": "rms", "xie": "rms", "xix": "rms", "xis": "rms"}
XC_HIST_NSTATS = 12 + 7 * len(HIST_FIELDS)


def _hist_vectors():
    names = ["time", "dt", "sumB", "sumM", "sumX", "EB", "EM", "EX", "DB", "DM", "DX"]
    for f in HIST_FIELDS:
        kind = HIST_KIND.get(f, "gm")
        names += [f"{f}_min", f"{f}_mean" if kind == "std" else f"{f}_{kind}", f"{f}_max"]
        if kind == "std":
            names.append(f"{f}_std")
    return names


HIST_VECTORS = _hist_vectors()


class History:
    """history.jl:21-93 ``History(T)``: every vector the reference appends to, plus the cached rates."""

    def __init__(self):
        for n in HIST_VECTORS:
            setattr(self, n, [])
        for q in ("B", "M", "X"):
            for sfx in ("", "o", "oo"):
                setattr(self, f"dsum{q}dt{sfx}", 0.0)

    def vectors(self):
        return {n: getattr(self, n) for n in HIST_VECTORS}

    def rates(self):
        return {f"dsum{q}dt{sfx}": getattr(self, f"dsum{q}dt{sfx}") for q in ("B", "M", "X") for sfx in ("", "o", "oo")}


_lib.register("xc_run_step", [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_lib.register("xc_history_sums", [_P, _P, _P, ctypes.POINTER(_D)])
_lib.register("xc_history_stats", [_P, _P, _P, _P, ctypes.POINTER(_D)])
_lib.register("xc_snapshot", [_P, _P, _P])
_lib.register("xc_compute_resnorm", [_P, _P, _P, _D, ctypes.POINTER(_D)])
_lib.register("xc_initialize", [_P, _P, _P, _P, _P, _P, _P, _I, _I, _P])


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


def history_sums(phase, fluid):
    c = context()
    out = (_D * 12)()
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_history_sums(c.ptr, ctypes.byref(ps), ctypes.byref(fs), out), "record_history!")
    return list(out)


def history_stats(phase, fluid, ns):
    """The XC_HIST_NSTATS reductions of record_history! (include/xcore_hip.h, xc_history_stats)."""
    c = context()
    out = (_D * XC_HIST_NSTATS)()
    ps, fs, nss = phase.pack(), fluid.pack(), ns.pack()
    c.check(lib().xc_history_stats(c.ptr, ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss), out),
            "record_history!")
    return list(out)


_lib.register("xc_history_stats_noalloc", [_P, _P, _P, _P, _P])


def history_stats_noalloc(phase, fluid, ns, out, stream=None):
    """xc_history_stats_noalloc: the same reductions into the device tensor ``out`` (XC_HIST_NSTATS float64)
    on ``stream`` (torch's current one by default), without a context or any allocation; asynchronous."""
    import torch
    if out.dtype != torch.float64 or out.device.type != "cuda" or out.numel() < XC_HIST_NSTATS:
        raise AssertionError("out must be a float64 device tensor of XC_HIST_NSTATS elements")
    s = torch.cuda.current_stream() if stream is None else stream
    ps, fs, nss = phase.pack(), fluid.pack(), ns.pack()
    code = lib().xc_history_stats_noalloc(_P(s.cuda_stream), ctypes.byref(ps), ctypes.byref(fs),
                                          ctypes.byref(nss), _P(out.data_ptr()))
    if code != 0:
        msg = lib().xc_last_error(None).decode(errors="replace")
        if code == _lib.XC_EINVAL_SHAPE:
            raise AssertionError(msg)
        raise _lib.XCoreError(code, "record_history! (no-alloc): " + msg)


def snapshot_(phase, fluid):
    """diagnose.jl:42-49 ``snapshot!`` into the context (one snapshot per context)."""
    c = context()
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_snapshot(c.ptr, ctypes.byref(ps), ctypes.byref(fs)), "snapshot!")


def compute_resnorm(phase, fluid, dt):
    """diagnose.jl:57-71 -> (resnorm, res_mass, res_mmnt) against the context's snapshot."""
    c = context()
    out = (_D * 3)()
    ps, fs = phase.pack(), fluid.pack()
    c.check(lib().xc_compute_resnorm(c.ptr, ctypes.byref(ps), ctypes.byref(fs), float(dt), out), "compute_resnorm")
    return tuple(out)


def record_history_(hst, time, dt, phase, fluid, ns, grid, a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0):
    """history.jl:101-222 ``record_history!``: the reductions on the device, the bookkeeping here."""
    h = grid.h
    h2 = h * h
    s = history_stats(phase, fluid, ns)
    hst.dsumBdtoo, hst.dsumBdto = hst.dsumBdto, hst.dsumBdt
    hst.dsumMdtoo, hst.dsumMdto = hst.dsumMdto, hst.dsumMdt
    hst.dsumXdtoo, hst.dsumXdto = hst.dsumXdto, hst.dsumXdt
    hst.time.append(float(time))
    hst.dt.append(float(dt))
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
    n = float(grid.Nz * grid.Nx)
    for i, f in enumerate(HIST_FIELDS):
        mn, mx, s1, s2, d1, d2, sl = s[12 + 7 * i: 19 + 7 * i]
        kind = HIST_KIND.get(f, "gm")
        getattr(hst, f"{f}_min").append(mn)
        getattr(hst, f"{f}_max").append(mx)
        if kind == "std":
            hst.x_mean.append(s1 / n)
            hst.x_std.append(math.sqrt(max(d2 - d1 * d1 / n, 0.0) / (n - 1.0)))
        elif kind == "rms":
            getattr(hst, f"{f}_rms").append(math.sqrt(s2 / n))
        else:
            getattr(hst, f"{f}_gm").append(math.exp(sl / n))


def initialize_(phase, fluid, ns, hst, grid, par, scales, xBC="periodic", zBC="closed", x_init=None, ctx=None):
    """run.jl:71-140 ``initialize!`` (fresh start; a restart loads a checkpoint with ``load_checkpoint_``).

    The random perturbation ``rp`` comes from the library's generator (seed par.seed, stream 2) in place of
    Julia's MersenneTwister.  ``x_init`` (a device matrix) replaces the perturbed profile when given.
    """
    c = ctx or context()
    ps, fs, nss = phase.pack(), fluid.pack(), ns.pack()
    xi = ctypes.byref(as_mat(x_init, "x_init")) if x_init is not None else None
    code = lib().xc_initialize(c.ptr, ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss),
                               ctypes.byref(grid_struct(grid)), ctypes.byref(par_struct(par)),
                               ctypes.byref(scales_struct(scales)), _sym(xBC, BCS, "xBC"), _sym(zBC, BCS, "zBC"),
                               xi)
    ns.draws = int(nss.draws)
    c.check(code, "initialize!")
    if ctx is not None:
        ctx.sync()
    if hst is not None:
        record_history_(hst, 0.0, scales.dt0, phase, fluid, ns, grid)
    return 0.0, scales.dt0, 0


def run_step_(phase, fluid, ns, grid, par, scales, state, opts, ctx=None):
    """One outer step of run! (run.jl:205-274) on the device; ``state`` is a StepState."""
    c = ctx or context()
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
