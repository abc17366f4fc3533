"""Pins of oracle functions that the reference's own tests leave open, by what the mathematics fixes:

- fluidmech! in production mode: a manufactured solution with every production term active (inertia,
  gamma, the drunken-sailor terms, momentum advection, buoyancy, the previous-momentum terms, the closed
  bottom and the pressure pin), second-order convergence of W, U, P (each of those terms, sign-flipped in
  the manufactured forcing, stops the convergence);
- the closed-bottom row and the pin: rho_w W = mean(MFS) D on the bottom face and P(np0) = 0 exactly, which
  follow from summing the divergence rows (fluidmech.jl:170-191, 367-395, 480-487);
- the MFS update (fluidmech.jl:111-127), the segregation viscosity (update.jl:345-409) and the stresses
  (update.jl:432-476) on fields whose results are known in closed form;
- noise! (noise.jl:138-239): equivariance under the x and z mirrors of the grid (the Gaussian filter is even,
  the grid's x is periodic and its z padding symmetric), which fixes the index alignment of the fluxes and
  of the cell magnitudes.
"""
import copy
import math

import numpy as np
import pytest

import xcore_amd as xc
from fieldsetup import plausible_state
from mms_inertial import GAMMA, G0, inertial_case, load_fluid

EPS = 2.220446049250313e-16


def _inertial_solve(orc, N):
    case, kw = inertial_case(N)
    par = xc.Parameters(D=1.0, N=N, L=1.0, g0=G0, gamma=GAMMA)
    g = xc.Grid(N, N, 1.0 / N, 1.0, 1.0)
    fl, ph = orc.fluid_state(N, N), orc.phase_state(N, N)
    load_fluid(fl, case)
    orc.fluidmech(fl, ph, g, orc.fm_opts(par, **kw))
    return case, fl


def _mms_errors(case, fl):
    out = {}
    for k, sl in (("W", np.s_[1:-1, 1:-1]), ("U", np.s_[1:-1, :]), ("P", np.s_[1:-1, 1:-1])):
        ex = case[k + "_exact"][sl]
        out[k] = np.linalg.norm(getattr(fl, k)[sl] - ex) / np.linalg.norm(ex)
    return out


def test_inertial_mms_orders(orc):
    errs = {"W": [], "U": [], "P": []}
    for N in (32, 64, 128):
        case, fl = _inertial_solve(orc, N)
        for k, e in _mms_errors(case, fl).items():
            errs[k].append(e)
        # the pin's pressure is exactly zero (see test_closed_bottom_and_pin)
        iz = ix = round(N / 2) - 1
        assert abs(fl.P[iz, ix]) <= 1e-11 * np.abs(fl.P).max()
    for k, e in errs.items():
        assert e[2] < 2e-4, (k, e)
        assert math.log2(e[0] / e[1]) > 1.9 and math.log2(e[1] / e[2]) > 1.9, (k, e)


@pytest.mark.parametrize("shape", [(24, 32), (40, 24)])
def test_closed_bottom_and_pin(orc, shape):
    # production solve of a plausible state whose MFS has a nonzero mean: the bottom face carries the whole
    # divergence (rho_w W = mean(MFS) D in every column) and the pinned pressure is exactly zero
    Nz, Nx = shape
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=4)
    orc.update(ph, fl, g, par, sc)
    fl.MFS[:] = 1e-3 + 1e-3 * np.random.default_rng(1).standard_normal((Nz, Nx))
    ph.advn_rho[:] = 0.0
    fl.rhoo[:] = fl.rho
    fl.rhooo[:] = fl.rho
    for k in ("drhodto", "drhodtoo"):
        getattr(fl, k)[:] = 0.0
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5)
    orc.fluidmech(fl, ph, g, orc.fm_opts(par, **kw))
    want = fl.MFS.mean() * g.D
    np.testing.assert_allclose(fl.rhow[-1, :] * fl.W[-1, 1:-1], want, rtol=1e-9)   # LU round-off ~1e-10
    assert np.abs(fl.W[0, 1:-1]).max() <= 1e-10 * np.abs(fl.W).max()   # top W = 0 (scaled-update round-off)
    iz, ix = round(Nz / 2) - 1, round(Nx / 2) - 1
    assert abs(fl.P[iz, ix]) <= 1e-9 * np.abs(fl.P).max()


def test_mfs_update_closed_form(orc):
    # uniform fields: drhodt = -advn_rho; MFS -= alpha ((a1 rho - a2 rhoo - a3 rhooo)/dt - (b1 drhodt + b2 drhodto
    # + b3 drhodtoo)) (fluidmech.jl:113-125)
    N = 12
    par = xc.Parameters(D=1.0, N=N, L=1.0, alpha=0.7)
    g = xc.Grid(N, N, 1.0 / N, 1.0, 1.0)
    fl, ph = orc.fluid_state(N, N), orc.phase_state(N, N)
    fl.eta[:] = 1.0
    fl.etaco[:] = 1.0
    for k in ("rho", "rhow", "rhou"):
        getattr(fl, k)[:] = 2.0
    fl.rhoo[:], fl.rhooo[:] = 1.5, 1.25
    fl.drhodto[:], fl.drhodtoo[:] = 0.1, -0.05
    fl.MFS[:] = 0.2
    ph.advn_rho[:] = 0.3
    dt, (a1, a2, a3, b1, b2, b3) = 0.5, (1.5, 2.0, -0.5, 0.6, 0.3, 0.1)
    orc.fluidmech_assemble(fl, ph, g, orc.fm_opts(par, dt=dt, a1=a1, a2=a2, a3=a3, b1=b1, b2=b2, b3=b3))
    res = (a1 * 2.0 - a2 * 1.5 - a3 * 1.25) / dt - (b1 * -0.3 + b2 * 0.1 + b3 * -0.05)
    np.testing.assert_allclose(fl.drhodt, -0.3, rtol=0, atol=0)
    np.testing.assert_allclose(fl.MFS, 0.2 - 0.7 * res, rtol=1e-15)


def _seg_state(orc, Nz, Nx):
    ph, fl = orc.phase_state(Nz, Nx), orc.fluid_state(Nz, Nx)
    return ph, fl


def test_segregation_viscosity_closed_form(orc):
    Nz, Nx = 6, 5
    par = xc.Parameters(D=1.0, N=Nz, L=1.0, etacntr=1e3, d0=0.02)
    sc = xc.compute_scales(par, xc.Grid(Nz, Nx, 1.0 / Nz, 1.0, 1.0 * Nx / Nz))
    ph, fl = _seg_state(orc, Nz, Nx)
    v, r, em, ke, fReL, V, prev = 3e-3, 2800.0, 40.0, 2e-4, 0.3, 5e-3, 55.0
    ph.vx[:], fl.rho[:], ph.etamix[:], ph.ke[:], ph.fReL[:], ph.V[:], ph.etas[:] = v, r, em, ke, fReL, V, prev
    orc.update_stage("segregation_viscosity", ph, fl, None, par, sc)
    l0 = sc.l0
    Rel = v * l0 * r / em
    fRel = 1 - math.exp(-Rel)
    ks = v * l0
    etat = fRel * ks * r
    etasi = em + etat
    capped = 1 / (1 / (etasi * 1e3) + 1 / etasi)
    etas = 0.5 * capped + 0.5 * prev
    for name, want in (("Rel", Rel), ("fRel", fRel), ("ks", ks), ("etat", etat), ("kx", ks + fReL * ke),
                       ("Rc", V / v), ("etas", etas), ("Red", v * par.d0 / (etas / r)), ("etasw", etas)):
        np.testing.assert_allclose(getattr(ph, name), want, rtol=1e-14, err_msg=name)
    # the cap uses the global minimum; the z-face values are geometric means with the boundary rows repeated
    ph.etamix[:] = em
    ph.etamix[2, :] = 4 * em
    ph.vx[:] = 0.0
    ph.etas[:] = 0.0
    orc.update_stage("segregation_viscosity", ph, fl, None, par, sc)
    e_lo = 0.5 / (1 / (em * 1e3) + 1 / em)
    e_hi = 0.5 / (1 / (em * 1e3) + 1 / (4 * em))
    col = np.array([e_lo, e_lo, e_hi, e_lo, e_lo, e_lo])
    np.testing.assert_allclose(ph.etas, np.repeat(col[:, None], Nx, 1), rtol=1e-14)
    zf = np.sqrt(np.r_[col[:1], col] * np.r_[col, col[-1:]])
    np.testing.assert_allclose(ph.etasw, np.repeat(zf[:, None], Nx, 1), rtol=1e-14)
    np.testing.assert_array_equal(ph.Rc, V / EPS)   # V / max(vx, eps) at vx = 0


def test_stresses_closed_form(orc):
    Nz, Nx = 5, 7
    par = xc.Parameters(D=1.0, N=Nz, L=1.0 * Nx / Nz)
    g = xc.Grid(Nz, Nx, 1.0 / Nz, 1.0, 1.0 * Nx / Nz)
    sc = xc.compute_scales(par, g)
    ph, fl = orc.phase_state(Nz, Nx), orc.fluid_state(Nz, Nx)
    fl.eta[:], fl.etaco[:] = 3.0, 2.0
    ph.exx[:], ph.ezz[:] = 0.5, -0.25
    ph.exz[:] = 0.0
    ph.exz[2, 3] = 4.0              # one corner: only the four cells around it see it, with weight 1/4
    ph.V[:], ph.kx[:] = 2e-3, 4e-4
    orc.update_stage("stresses", ph, fl, g, par, sc)
    np.testing.assert_array_equal(ph.txx, 1.5)
    np.testing.assert_array_equal(ph.tzz, -0.75)
    np.testing.assert_array_equal(ph.txz, 2.0 * ph.exz)
    base = math.sqrt(EPS + (1.5 ** 2 + 0.75 ** 2) / 2)
    touched = math.sqrt(EPS + (1.5 ** 2 + 0.75 ** 2) / 2 + 8.0 ** 2 / 4)
    want = np.full((Nz, Nx), base)
    want[1:3, 2:4] = touched
    np.testing.assert_allclose(ph.tII, want, rtol=1e-15)
    np.testing.assert_allclose(ph.Ra, 2e-3 * sc.D0 / 4e-4, rtol=1e-15)


# ---- noise! mirror equivariance ----------------------------------------------------------------------

_CELL = ("V", "vx", "fReL", "ke", "chi", "ks", "x", "m")
_WT = ("chiw", "muw")
_UT = ("chiu", "muu")
# flux components that are differences along the mirrored axis change sign
_FLIP = {"x": {"xiew", "xixu", "xisu"}, "z": {"xieu", "xixw", "xisw"}}


def _mirror(a, axis):
    return np.asfortranarray(a[:, ::-1] if axis == "x" else a[::-1, :])


def _noise_inputs(orc, Nz, Nx, seed):
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=seed)
    orc.update(ph, fl, g, par, sc)
    rng = np.random.default_rng(seed)
    for n in ("re", "rs", "psieo", "psixo", "psiso"):
        getattr(ns, n)[:] = rng.standard_normal((Nz, Nx))
    return par, g, sc, ph, ns


@pytest.mark.parametrize("axis", ["x", "z"])
def test_noise_mirror_equivariance(orc, axis):
    Nz, Nx = 40, 48
    par, g, sc, ph, ns = _noise_inputs(orc, Nz, Nx, 3)
    assert ns.padL0 % 2 == 0 and ns.padl0 % 2 == 0
    ph2, ns2 = copy.deepcopy(ph), copy.deepcopy(ns)
    for n in _CELL + _WT + _UT:
        getattr(ph2, n)[:] = _mirror(getattr(ph, n), axis)
    for n in ("re", "rs", "psieo", "psixo", "psiso"):
        getattr(ns2, n)[:] = _mirror(getattr(ns, n), axis)
    dt = 0.3 * sc.dt0
    orc.noise(ns, ph, g, par, sc, dt, first_iter=False)
    orc.noise(ns2, ph2, g, par, sc, dt, first_iter=False)
    for n in ("psie", "psix", "psis", "xiew", "xieu", "xixw", "xixu", "xisw", "xisu", "xie", "xix", "xis"):
        a, b = getattr(ns, n), getattr(ns2, n)
        s = -1.0 if n in _FLIP[axis] else 1.0
        scale = max(np.abs(a).max(), 1e-300)
        assert np.abs(s * _mirror(a, axis) - b).max() <= 1e-12 * scale, (n, axis)
        assert np.abs(a).max() > 0, n


def test_noise_constant_potential_gives_no_flux(orc):
    # a uniform field survives the filter unchanged (std = 0) and has no differences: every flux and
    # magnitude is zero (noise.jl:195-236)
    Nz, Nx = 24, 32
    par, g, sc, ph, ns = _noise_inputs(orc, Nz, Nx, 5)
    for n in ("re", "rs"):
        getattr(ns, n)[:] = 0.0
    for n in ("psieo", "psixo", "psiso"):
        getattr(ns, n)[:] = 0.7
    ph.V[:], ph.vx[:] = 1e-3, 1e-4   # uniform decay factors
    orc.noise(ns, ph, g, par, sc, 0.3 * sc.dt0, first_iter=False)
    for n in ("xiew", "xieu", "xixw", "xixu", "xisw", "xisu", "xie", "xix", "xis"):
        assert np.abs(getattr(ns, n)).max() <= 1e-12, n   # round-off of the mean of a constant
