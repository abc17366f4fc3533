"""Seeded, physically plausible states shared by the field tests (host numpy side).

Builds a HostState pair (oracle containers) from a Parameters/Grid/Scales
triple; GPU tests copy the same arrays to the device.
"""
import math

import numpy as np


def plausible_state(orc, xc, Nz, Nx, seed=0, D=10.0):
    """Return (par, grid, scales, phase, fluid, noise) on the host."""
    par = xc.Parameters(D=D, N=Nz, L=D * Nx / Nz, etam0=10.0, x0=0.1)
    grid = xc.Grid(Nz, Nx, D / Nz, D, D * Nx / Nz)
    sc = xc.compute_scales(par, grid)
    rng = np.random.default_rng(seed)
    ph = orc.phase_state(Nz, Nx)
    fl = orc.fluid_state(Nz, Nx)
    h = grid.h
    zc = (np.arange(Nz) + 0.5) * h
    xcn = (np.arange(Nx) + 0.5) * h
    x = 0.1 + 0.05 * np.outer(np.sin(2 * np.pi * zc / D), np.cos(2 * np.pi * xcn / grid.L)) \
        + 0.01 * rng.standard_normal((Nz, Nx))
    ph.x[:] = x
    ph.m[:] = 1.0 - x
    rho = 1.0 / (ph.m / par.rhom0 + ph.x / par.rhox0)
    ph.X[:] = rho * ph.x
    ph.M[:] = rho * ph.m
    ph.Xo[:] = ph.X * (1 + 1e-3 * rng.standard_normal((Nz, Nx)))
    ph.Xoo[:] = ph.X * (1 + 1e-3 * rng.standard_normal((Nz, Nx)))
    ph.Mo[:] = ph.M
    ph.Moo[:] = ph.M
    fl.rho[:] = rho
    fl.rhoo[:] = rho
    fl.rhooo[:] = rho
    fl.eta[:] = par.etam0 * np.exp(0.3 * rng.standard_normal((Nz, Nx)))
    fl.etaco[:] = par.etam0
    ph.etas[:] = par.etam0 * np.exp(0.3 * rng.standard_normal((Nz, Nx)))
    fl.W[:] = 1e-3 * rng.standard_normal(fl.W.shape)
    fl.U[:] = 1e-3 * rng.standard_normal(fl.U.shape)
    fl.P[:] = 10.0 * rng.standard_normal(fl.P.shape)
    ph.wx[:] = 1e-4 * rng.standard_normal(ph.wx.shape)
    ph.wm[:] = 1e-5 * rng.standard_normal(ph.wm.shape)
    ph.kx[:] = 1e-3 * np.exp(rng.standard_normal((Nz, Nx)))
    ph.chi[:] = np.clip(ph.x * rho / par.rhox0, 1e-16, 1)
    for n in ("Wx", "Wm"):
        getattr(ph, n)[:] = 1e-3 * rng.standard_normal(ph.Wx.shape)
    for n in ("Ux", "Um"):
        getattr(ph, n)[:] = 1e-3 * rng.standard_normal(ph.Ux.shape)
    for n in ("dXdt", "dXdto", "dXdtoo"):
        getattr(ph, n)[:] = 1e-4 * rng.standard_normal((Nz, Nx))
    # boundary taper / shape (run.jl:24-55)
    zf = np.arange(Nz + 1) * h
    ph.bndtaperw[:] = (1 - (np.exp(-zf / sc.l0h) + np.exp(-(D - zf) / sc.l0h)))[:, None]
    ph.bndshape[:] = np.exp((-zc + h / 2) / sc.bnd_w)[:, None]
    padL0 = 4 * math.ceil(sc.L0h / h)
    padl0 = 4 * math.ceil(sc.l0h / h)
    fL = 2 / math.sqrt(1 - math.exp(-h ** 2 / (2 * sc.L0h ** 2)))
    fl_ = 2 / math.sqrt(1 - math.exp(-h ** 2 / (2 * sc.l0h ** 2)))
    ns = orc.HostNoise(Nz, Nx, padL0, padl0, fL, fl_)
    for n in ("xiew", "xixw", "xisw", "xieu", "xixu", "xisu"):
        a = getattr(ns, n)
        a[:] = 1e-4 * rng.standard_normal(a.shape)
    return par, grid, sc, ph, fl, ns


def to_device_state(xc, host, kind, grid=None, sc=None):
    """Copy a host container into a fresh device container of the same kind."""
    import torch
    if kind == "phase":
        dev = xc.PhaseState(host.Nz, host.Nx)
    elif kind == "fluid":
        dev = xc.FluidState(host.Nz, host.Nx)
    else:
        dev = xc.NoiseState(grid, sc)
    for n, a in host.arrays().items():
        t = getattr(dev, n)
        t.copy_(torch.as_tensor(np.ascontiguousarray(a)))
    return dev
