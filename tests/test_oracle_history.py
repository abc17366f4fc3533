"""The oracle's initialize! (run.jl:71-140), record_history! (history.jl:101-222) and compute_resnorm
(diagnose.jl:57-71), pinned by the reference's own tests (test_history.jl, test_diagnose.jl) and by closed
forms of the initial profile."""
import math

import numpy as np
import pytest

import xcore_amd as xc


def _small(orc, N=8, **kw):
    # test_history.jl:5-15 / test_diagnose.jl:6-16: D = 10, N = 8, L = 10, etam0 = 10, x0 = 0.1, Da = 0
    kw = {"Da": 0.0, **kw}
    par = xc.Parameters(D=10.0, N=N, L=10.0, etam0=10.0, x0=0.1, **kw)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = orc.phase_state(g.Nz, g.Nx), orc.fluid_state(g.Nz, g.Nx)
    ns = orc.HostNoise(g.Nz, g.Nx, 4 * math.ceil(sc.L0h / g.h), 4 * math.ceil(sc.l0h / g.h),
                       2 / math.sqrt(1 - math.exp(-g.h ** 2 / (2 * sc.L0h ** 2))),
                       2 / math.sqrt(1 - math.exp(-g.h ** 2 / (2 * sc.l0h ** 2))))
    orc.initialize(ph, fl, ns, g, par, sc, seed=1)
    return par, g, sc, ph, fl, ns


def test_history_first_record_exact_zero(orc):
    # test_history.jl:29-48: one record; EB = EM = EX = 0 exactly
    par, g, sc, ph, fl, ns = _small(orc)
    hst = orc.new_history()
    orc.record_history(hst, 0.0, 1.0, ph, fl, ns, g)
    assert all(len(v) == 1 for k, v in hst.items() if isinstance(v, list))
    assert hst["EB"][0] == 0.0 and hst["EM"][0] == 0.0 and hst["EX"][0] == 0.0
    # 69 vectors, as the reference's History (history.jl:21-72)
    assert sum(isinstance(v, list) for v in hst.values()) == 69


def test_history_ordering_finite_sorted(orc):
    # test_history.jl:50-88: min <= mean/rms <= max, all finite, time sorted
    par, g, sc, ph, fl, ns = _small(orc)
    hst = orc.new_history()
    for k in range(1, 6):
        orc.record_history(hst, k * sc.dt0, sc.dt0, ph, fl, ns, g)
    eps = math.sqrt(2.220446049250313e-16)
    for k in range(5):
        assert hst["x_min"][k] <= hst["x_mean"][k] <= hst["x_max"][k]
        assert hst["V_min"][k] <= hst["V_rms"][k] <= hst["V_max"][k] + eps
        assert hst["vx_min"][k] <= hst["vx_rms"][k] <= hst["vx_max"][k] + eps
    for k, v in hst.items():
        assert np.all(np.isfinite(v)), k
    assert hst["time"] == sorted(hst["time"])
    # the same state recorded again: no mass change and no flux drift beyond the recorded rates
    assert hst["sumB"][1:] == hst["sumB"][:-1]


def test_history_drift_closed_form(orc):
    # a state with a single known boundary flux: qz_advn_X = 2 on the top row's interior columns, 0 elsewhere,
    # Gx = 0.  Then dsumXdt = +2 L (top inflow), and with BE coefficients D_k = D_{k-1} + dsumXdt dt
    par, g, sc, ph, fl, ns = _small(orc)
    for n in ("qz_advn_X", "qz_advn_M", "qz_dffn_X", "qz_dffn_M", "Gx"):
        getattr(ph, n)[:] = 0.0
    ph.qz_advn_X[0, 1:-1] = 2.0
    ph.qz_advn_X[0, 0] = ph.qz_advn_X[0, -1] = 99.0   # ghost columns are not part of the line integral
    hst = orc.new_history()
    dt = 0.25
    for k in range(3):
        orc.record_history(hst, k * dt, dt, ph, fl, ns, g)
    L = g.Nx * g.h
    assert hst["dsumXdt"] == pytest.approx(2.0 * L, rel=1e-15)
    assert hst["DX"] == pytest.approx([0.0, 2.0 * L * dt, 4.0 * L * dt], rel=1e-15)
    assert hst["DB"] == pytest.approx(hst["DX"], rel=1e-15)
    assert hst["DM"] == [0.0, 0.0, 0.0]
    # BD2 coefficients (4/3, -1/3 history, 2/3 b1 in time_coefs' normalisation): the drift recursion
    a1, a2, a3, b1, b2, b3 = orc.time_coefs("bd2im", 3)
    orc.record_history(hst, 3 * dt, dt, ph, fl, ns, g, a1, a2, a3, b1, b2, b3)
    want = (a2 * hst["DX"][2] + a3 * hst["DX"][1] + (b1 + b2 + b3) * 2.0 * L * dt) / a1
    assert hst["DX"][3] == pytest.approx(want, rel=1e-15)


def test_history_statistics_closed_forms(orc):
    # geometric mean of a two-valued field, rms of a sign-alternating field, n-1 std of a two-valued field
    par, g, sc, ph, fl, ns = _small(orc)
    n = g.Nz * g.Nx
    a = np.where(np.arange(n).reshape(g.Nz, g.Nx, order="F") % 2 == 0, 4.0, 16.0)
    ph.Ra[:] = a
    ph.V[:] = np.where(a == 4.0, -3.0, 3.0)
    ph.x[:] = np.where(a == 4.0, 0.1, 0.3)
    hst = orc.new_history()
    orc.record_history(hst, 0.0, 1.0, ph, fl, ns, g)
    assert hst["Ra_gm"][0] == pytest.approx(8.0, rel=1e-14)
    assert hst["V_rms"][0] == pytest.approx(3.0, rel=1e-15)
    assert hst["x_mean"][0] == pytest.approx(0.2, rel=1e-14)
    assert hst["x_std"][0] == pytest.approx(0.1 * math.sqrt(n / (n - 1)), rel=1e-13)
    ph.Ra[0, 0] = 0.0   # log(max(a, eps)) keeps a zero finite
    orc.record_history(hst, 1.0, 1.0, ph, fl, ns, g)
    assert np.isfinite(hst["Ra_gm"][1]) and hst["Ra_min"][1] == 0.0


def test_initialize_profile_closed_form(orc):
    # dxr = dxg = 0: x = xin = x0 + (Da - x0) bndshape, bndshape = exp((-zc + h/2)/bnd_w), taper closed form
    par, g, sc, ph, fl, ns = _small(orc, N=12)
    par2 = xc.Parameters(D=10.0, N=12, L=10.0, etam0=10.0, x0=0.1, Da=0.02, dxr=0.0, dxg=0.0)
    ph2, fl2 = orc.phase_state(g.Nz, g.Nx), orc.fluid_state(g.Nz, g.Nx)
    ns2 = orc.HostNoise(g.Nz, g.Nx, ns.padL0, ns.padl0, ns.fL, ns.fl)
    sc2 = xc.compute_scales(par2, g)
    orc.initialize(ph2, fl2, ns2, g, par2, sc2, seed=1)
    zc = (np.arange(g.Nz) + 0.5) * g.h
    shape = np.exp((-zc + g.h / 2) / sc2.bnd_w)
    np.testing.assert_allclose(ph2.bndshape, np.repeat(shape[:, None], g.Nx, 1), rtol=1e-15)
    np.testing.assert_allclose(ph2.x, np.repeat((0.1 + (0.02 - 0.1) * shape)[:, None], g.Nx, 1), rtol=1e-14)
    np.testing.assert_allclose(ph2.m, 1.0 - ph2.x, rtol=1e-15)
    zf = np.arange(g.Nz + 1) * g.h
    taper = 1 - (np.exp(-zf / sc2.l0h) + np.exp(-(10.0 - zf) / sc2.l0h))
    np.testing.assert_allclose(ph2.bndtaperw, np.repeat(taper[:, None], g.Nx + 2, 1), rtol=1e-14, atol=1e-15)
    # X = rho x, M = rho m, the old levels equal the new, rho is the mixture density (update.jl:24-33)
    np.testing.assert_array_equal(ph2.X, fl2.rho * ph2.x)
    np.testing.assert_array_equal(ph2.Xo, ph2.X)
    np.testing.assert_array_equal(ph2.Moo, ph2.M)
    np.testing.assert_array_equal(fl2.rhooo, fl2.rho)
    np.testing.assert_allclose(fl2.rho, 1.0 / (ph2.m / par2.rhom0 + ph2.x / par2.rhox0), rtol=1e-14)


def test_initialize_perturbation_normalised(orc):
    # x = x0 (1 + dxr rp) with Da = x0: rp = (x / x0 - 1) / dxr has zero mean and unit n-1 std
    par, g, sc, ph, fl, ns = _small(orc, N=16, Da=0.1)
    rp = (ph.x / par.x0 - 1.0) / par.dxr
    assert abs(rp.mean()) < 1e-12
    assert rp.std(ddof=1) == pytest.approx(1.0, rel=1e-12)
    # and it is the library generator's stream 2 of seed par.seed, normalised
    raw = orc.normal_fill(orc.zeros(g.Nz, g.Nx), par.seed, 2, 0)
    np.testing.assert_allclose(rp, (raw - raw.mean()) / raw.std(ddof=1), rtol=1e-9, atol=1e-12)


def test_resnorm_zero_and_positive(orc):
    # test_diagnose.jl:52-85: unchanged state -> exactly 0; a perturbed W -> positive momentum part only
    par, g, sc, ph, fl, ns = _small(orc)
    snap = orc.snapshot(ph, fl)
    assert orc.compute_resnorm(snap, ph, fl, sc.dt0) == (0.0, 0.0, 0.0)
    fl.W[2, 2] += 1e-3
    rn, rm, rp = orc.compute_resnorm(snap, ph, fl, sc.dt0)
    assert rm == 0.0 and rp > 0 and rn == rp
    # closed form of the W term: |dW| / (|W| + eps)
    want = 1e-3 / (np.linalg.norm(fl.W) + 2.220446049250313e-16)
    assert rp == pytest.approx(want, rel=1e-14)
