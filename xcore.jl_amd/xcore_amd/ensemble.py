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


# ---- ensembles of whole runs: every member advances one step of run! per call -------------------------

class EnsembleMemberStruct(ctypes.Structure):
    _fields_ = [("ctx", _P), ("phase", _P), ("fluid", _P), ("ns", _P), ("grid", _P), ("par", _P), ("scales", _P),
                ("opts", _P), ("state", _P), ("stats", _P), ("status", _I)]


_lib.register("xc_ensemble_run_step", [_I, _P, _I])


class RunMember:
    """One run of an ensemble: its own context (stream, solver hierarchy), grid, parameters, scales, noise
    seed and states, set up by initialize! on its context."""

    def __init__(self, par, grid, scales, seed, x_init=None, device=None):
        import torch
        from .driver import StepState
        from .state import FluidState, NoiseState, PhaseState
        self.par, self.grid, self.scales = par, grid, scales
        dev = torch.cuda.current_device() if device is None else int(device)
        self.ctx = _lib.Context(dev, 0)   # its own stream
        self.phase, self.fluid = PhaseState(grid.Nz, grid.Nx), FluidState(grid.Nz, grid.Nx)
        self.ns = NoiseState(grid, scales, seed=seed)
        torch.cuda.synchronize()          # the zero-filled states are written on torch's stream
        from .driver import initialize_
        t, dt, step = initialize_(self.phase, self.fluid, self.ns, None, grid, par, scales, x_init=x_init,
                                  ctx=self.ctx)
        self.state = StepState(t, dt, step)

    @property
    def cells(self):
        return self.grid.Nz * self.grid.Nx


def ensemble_members(n, sizes=(128, 256, 384), seed=0, base=None, drho_span=0.2, log10_etax0=(20.0, 30.0)):
    """n members cycling through the grid sizes, each with its own noise seed, crystal-melt density contrast
    drawn uniformly within +-drho_span of base's and crystal viscosity etax0 = 10^e, e uniform in log10_etax0
    (parameters.jl:50-57); scales of all members from one compute_scales call."""
    import dataclasses

    import numpy as np
    from .state import Grid, Parameters, compute_scales_batch
    base = base or Parameters()
    rng = np.random.default_rng(seed)
    pars = []
    for i in range(n):
        N = sizes[i % len(sizes)]
        f = rng.uniform(1.0 - drho_span, 1.0 + drho_span)
        e = rng.uniform(*log10_etax0)
        pars.append(dataclasses.replace(base, N=N, L=base.D, rhox0=base.rhom0 + (base.rhox0 - base.rhom0) * f,
                                        etax0=10.0 ** e, seed=int(seed * 1000 + i)))
    grids = [Grid.from_par(p) for p in pars]
    scales = compute_scales_batch(pars, grids)
    return [RunMember(p, g, s, seed=p.seed) for p, g, s in zip(pars, grids, scales)]


def ensemble_run_step_(members, opts, nthreads=0):
    """One step of run! for every member (xc_ensemble_run_step): the members' work runs concurrently on the
    device (nthreads host threads; 0: the library's default, 4).  opts: one StepOpts for all members or a list.
    Returns the members' step statistics."""
    from .driver import StepStats
    from .state import grid_struct, par_struct, scales_struct
    n = len(members)
    keep = []
    arr = (EnsembleMemberStruct * n)()
    stats = [StepStats() for _ in range(n)]
    nss = []
    for i, m in enumerate(members):
        o = opts[i] if isinstance(opts, (list, tuple)) else opts
        ps, fs, ns = m.phase.pack(), m.fluid.pack(), m.ns.pack()
        gs, pr, sc = grid_struct(m.grid), par_struct(m.par), scales_struct(m.scales)
        keep += [ps, fs, ns, gs, pr, sc, o]
        nss.append(ns)
        arr[i] = EnsembleMemberStruct(m.ctx.ptr, ctypes.addressof(ps), ctypes.addressof(fs), ctypes.addressof(ns),
                                      ctypes.addressof(gs), ctypes.addressof(pr), ctypes.addressof(sc),
                                      ctypes.addressof(o), ctypes.addressof(m.state), ctypes.addressof(stats[i]), 0)
    code = lib().xc_ensemble_run_step(n, ctypes.cast(arr, _P), int(nthreads))
    for m, ns in zip(members, nss):
        m.ns.draws = int(ns.draws)
    if code != 0:
        bad = next(i for i in range(n) if arr[i].status != 0)
        members[bad].ctx.check(arr[bad].status, f"ensemble member {bad}")
    del keep
    return [{k: getattr(s, k) for k, _ in StepStats._fields_} for s in stats]
