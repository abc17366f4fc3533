"""Pins for the oracle's field updates (update.jl, store.jl, phsevo.jl, timing.jl).

Known answers from the reference's own tests (time_coefs table, naive rheology,
strain-rate orders, store_previous! chain, update! smoke) plus closed forms.
"""
import math

import numpy as np
import pytest

from fieldsetup import plausible_state


def test_time_coefs_table(orc):
    # test_run.jl:5-25
    assert orc.time_coefs("bd2im", 1) == (1, 1, 0, 1, 0, 0)
    assert orc.time_coefs("bd2im", 2) == (1, 1, 0, 1, 0, 0)
    assert orc.time_coefs("bd2im", 3) == (1.5, 2, -0.5, 1, 0, 0)
    assert orc.time_coefs("bd2si", 1) == (1, 1, 0, 1, 0, 0)
    assert orc.time_coefs("bd2si", 2) == (1, 1, 0, 0.5, 0.5, 0)
    assert orc.time_coefs("bd2si", 3) == (1.5, 2, -0.5, 0.75, 0.5, -0.25)
    assert orc.time_coefs("be1im", 7) == (1, 1, 0, 1, 0, 0)
    assert orc.time_coefs("cn2si", 9) == (1, 1, 0, 0.5, 0.5, 0)
    with pytest.raises(orc.OracleError):
        orc.time_coefs("bogus", 1)


def _naive_rheology(chi, mu, par):
    # test_update.jl:8-28 -- the reference's own independent implementation
    f = (chi, mu)
    AA = np.array(par.AA).reshape(2, 2)
    BB = np.array(par.BB).reshape(2, 2)
    CC = np.array(par.CC).reshape(2, 2)
    Sf = np.array([[(f[j] / BB[i, j]) ** (1.0 / CC[i, j]) for j in range(2)] for i in range(2)])
    Sf = Sf / Sf.sum(axis=1, keepdims=True)
    Xf = np.zeros((2, 2))
    for i in range(2):
        a = AA[i, 0] * Sf[i, 0] + AA[i, 1] * Sf[i, 1]
        Xf[i, 0] = a * f[0] + (1 - a) * Sf[i, 0]
        Xf[i, 1] = a * f[1] + (1 - a) * Sf[i, 1]
    kv = (par.etax0, par.etam0)
    thtv = [(kv[0] / kv[i]) ** Xf[i, 0] * (kv[1] / kv[i]) ** Xf[i, 1] for i in range(2)]
    etaf = (kv[0] * thtv[0], kv[1] * thtv[1])
    return chi * etaf[0] + mu * etaf[1]


@pytest.mark.parametrize("chi,mu", [(0.30, 0.70), (0.50, 0.50), (0.10, 0.90), (0.90, 0.10)])
def test_rheology_vs_naive(orc, chi, mu):
    import xcore_amd as xc
    par = xc.Parameters()
    got = orc.rheology_cell(chi, mu, par)
    want = _naive_rheology(chi, mu, par)
    assert got == pytest.approx(want, rel=1e-12)


def test_rheology_pure_melt_limit(orc):
    # all melt: the melt row of the permission weights gives eta ~ etam0 scale;
    # and a swapped (etax0, etam0) pair would give ~1e18 instead
    import xcore_amd as xc
    par = xc.Parameters()
    v = orc.rheology_cell(1e-12, 1 - 1e-12, par)
    assert 1.0 < v < 1e3


def test_kinematics_orders(orc):
    # test_update.jl:49-94
    import xcore_amd as xc
    L = 2 * math.pi
    errs_exx, errs_ezz, errs_div = [], [], []
    for N in (16, 32, 64):
        h = L / N
        k = 2 * math.pi / L
        ph = orc.phase_state(N, N)
        fl = orc.fluid_state(N, N)
        jj, ii = np.meshgrid(np.arange(N + 1), np.arange(N + 2), indexing="ij")
        fl.W[:] = np.sin(k * jj * h) * np.cos(k * (ii - 0.5) * h)
        jj, ii = np.meshgrid(np.arange(N + 2), np.arange(N + 1), indexing="ij")
        fl.U[:] = -np.cos(k * (jj - 0.5) * h) * np.sin(k * ii * h)
        orc.update_stage("kinematics", ph, fl, xc.Grid(N, N, h, L, L), None, None)
        c = (np.arange(N) + 0.5) * h
        exact_exx = -k * np.outer(np.cos(k * c), np.cos(k * c))
        errs_exx.append(np.linalg.norm(ph.exx - exact_exx) / np.linalg.norm(exact_exx))
        errs_ezz.append(np.linalg.norm(ph.ezz + exact_exx) / np.linalg.norm(exact_exx))
        errs_div.append(np.linalg.norm(ph.Div_V))
    for e in (errs_exx, errs_ezz):
        assert math.log2(e[0] / e[1]) > 1.8 and math.log2(e[1] / e[2]) > 1.8
    assert errs_div[-1] < 1e-10


def test_update_smoke_and_invariants(orc):
    # test_update.jl:96-131 plus closed-form consequences of the pipeline
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 24, 32, seed=1)
    eta_prev = fl.eta.copy()
    orc.update(ph, fl, g, par, sc)
    for a in (fl.rho, ph.etamix, fl.eta, ph.eII, ph.tII):
        assert np.all(np.isfinite(a)) and np.all(a >= 0)
    rho = 1.0 / (ph.m / par.rhom0 + ph.x / par.rhox0)
    np.testing.assert_allclose(fl.rho, rho, rtol=1e-15)
    # closed z: boundary face densities equal the boundary cells'
    np.testing.assert_array_equal(fl.rhow[0], fl.rho[0])
    np.testing.assert_array_equal(fl.rhow[-1], fl.rho[-1])
    # periodic x: rhou faces 0 and Nx coincide
    np.testing.assert_array_equal(fl.rhou[:, 0], fl.rhou[:, -1])
    # Drho has zero row mean
    np.testing.assert_allclose(fl.Drho.mean(axis=1), 0, atol=1e-12 * fl.rho.max())
    # lithostatic pressure: increments are rowmean(rhow)*g*h
    dPl = np.diff(ph.Pl[:, 0])
    np.testing.assert_allclose(dPl, fl.rhow[1:-1].mean(axis=1) * par.g0 * g.h, rtol=1e-12)
    # eta is the 50/50 blend of the capped value with the previous iterate
    etai = ph.etamix + ph.etae
    emax = etai.min() * par.etacntr
    np.testing.assert_allclose(fl.eta, 0.5 / (1 / emax + 1 / etai) + 0.5 * eta_prev, rtol=1e-14)
    # corner eta is a geometric mean: between min and max of the 4 neighbours
    assert np.all(fl.etaco[1:-1, 1:-1] <= np.maximum.reduce([fl.eta[:-1, :-1], fl.eta[1:, :-1], fl.eta[:-1, 1:],
                                                             fl.eta[1:, 1:]]) * (1 + 1e-14))
    # kx = ks + fReL*ke and Rc = V / max(vx, eps)
    np.testing.assert_allclose(ph.kx, ph.ks + ph.fReL * ph.ke, rtol=1e-15)
    assert ph.etamix.shape == (24, 32) and ph.exz.shape == (25, 33)


def test_store_previous_chain(orc):
    # test_run.jl:27-46
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 8, 8, seed=2)
    ph.X[0, 0] = 7.0
    orc.store_previous(fl, ph)
    assert ph.Xo[0, 0] == 7.0
    ph.X[0, 0] = 9.0
    orc.store_previous(fl, ph)
    assert ph.Xo[0, 0] == 9.0 and ph.Xoo[0, 0] == 7.0
    np.testing.assert_array_equal(fl.rhoWo, fl.rhow * fl.W[:, 1:-1])
    np.testing.assert_array_equal(fl.rhoUo, fl.rhou * fl.U[1:-1, :])


def test_phase_velocities_closed_form(orc):
    # uniform etas and x: wx = d0^2/etas * (rhox0 - rhow) * g0 * taper, wm = -x/m * wx;
    # with zero noise Wx = W + wx, Wm = W + wm, Ux = Um = U.
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 12, 10, seed=3)
    for n in ("xiew", "xixw", "xisw", "xieu", "xixu", "xisu"):
        getattr(ns, n)[:] = 0
    ph.etasw[:] = 7.0
    ph.x_w[:] = 0.2
    ph.m_w[:] = 0.8
    ph.Drhox[:] = 500.0
    orc.update_phase_velocities(ph, fl, ns, g, par)
    wx = par.d0 ** 2 / 7.0 * 500.0 * par.g0 * ph.bndtaperw
    np.testing.assert_allclose(ph.wx, wx, rtol=1e-14)
    np.testing.assert_allclose(ph.wm, -0.25 * wx, rtol=1e-14)
    np.testing.assert_array_equal(ph.Ux, fl.U)
    np.testing.assert_array_equal(ph.Um, fl.U)
    np.testing.assert_allclose(ph.Wx, fl.W + wx, rtol=1e-14, atol=1e-20)


def test_update_dt_closed_form(orc):
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 10, 10, seed=4)
    ph.kx[:] = 1e-3
    ph.kx[3, 4] = 2e-3
    for n in ("Wx", "Wm", "Ux", "Um"):
        getattr(ph, n)[:] = 0.0
    ph.Um[5, 5] = -0.5
    hh = g.h / 2
    dtk = hh * hh / 2e-3
    dta = hh / (0.5 + 2.220446049250313e-16)
    assert orc.update_dt(ph, fl, g, par, 1.0) == min(1.5, min(dtk, par.CFL * dta), par.dtmax)
    assert orc.update_dt(ph, fl, g, par, 1e-9) == 1.5 * 1e-9


def test_phsevo_wiring(orc):
    # alpha = 1, BE coefficients: X_new = clamp(Xo + dt*(-advn_X + dffn_X + Gx)); advn_X is the
    # WENO5 divergence of X by the *face views* of Ux/Wx, dffn_X diffuses chi with k = X.*kx.
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 16, 20, seed=5)
    X0 = ph.X.copy(order="F")
    Xo = ph.Xo.copy(order="F")
    dt = 1e-2
    orc.phsevo(ph, fl, g, par, sc, coef=(1, 1, 0, 1, 0, 0), dt=dt, alpha=1.0)
    adv = orc.zeros(16, 20)
    orc.advect_centered(adv, X0, np.asfortranarray(ph.Ux[1:-1, :]), np.asfortranarray(ph.Wx[:, 1:-1]), g.h, "weno5")
    np.testing.assert_array_equal(ph.advn_X, adv)
    dff = orc.zeros(16, 20)
    orc.diffus_centered(dff, ph.chi, np.asfortranarray(X0 * ph.kx), g.h)
    np.testing.assert_array_equal(ph.dffn_X, dff)
    want = np.clip(Xo + dt * (-ph.advn_X + ph.dffn_X + ph.Gx), fl.rho * 2.220446049250313e-16,
                   fl.rho * (1 - 2.220446049250313e-16))
    np.testing.assert_allclose(ph.X, want, rtol=1e-13)
    np.testing.assert_allclose(ph.X + ph.M, fl.rho, rtol=1e-15)
    np.testing.assert_allclose(ph.advn_rho, ph.advn_X + ph.advn_M, rtol=0)


def test_phsevo_no_flow_no_change(orc):
    # zero velocities, zero diffusivity, G0 = 0, BE: X returns to Xo exactly
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 10, 12, seed=6)
    for n in ("Wx", "Wm", "Ux", "Um", "kx"):
        getattr(ph, n)[:] = 0
    sc.vals["G0"] = 0.0
    ph.Xo[:] = ph.X * 0.999
    orc.phsevo(ph, fl, g, par, sc, coef=(1, 1, 0, 1, 0, 0), dt=0.1, alpha=1.0)
    np.testing.assert_allclose(ph.X, np.clip(ph.Xo, fl.rho * 2.2e-16, fl.rho), rtol=1e-15)
