"""GPU Stokes (fluidmech!) against the oracle's assembled system and sparse-LU solution.

Bars: the matrix-free operator and right-hand side equal the oracle's assembled
LL x and RR to 1e-12 (normwise); the FGMRES solution has relative residual
||RR - LL x|| / ||RR|| <= 1e-8 on the oracle's own assembled system, and the
MMS convergence orders of test_mms.jl (> 1.8) hold on the device path.
"""
import copy
import math

import numpy as np
import pytest
import torch

from fieldsetup import plausible_state, to_device_state
from mms import mms_sources

pytestmark = pytest.mark.gpu


def _rel(g, w):
    return np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300)


def _mms(orc, xc, N, L=10.0):
    m = mms_sources(N, N, L)
    par = xc.Parameters(D=L, N=N, L=L, etam0=1e3, g0=10.0, rhom0=2700, rhox0=3200, d0=1e-2, Da=0.0, Xi=0.0,
                        gamma=0.0)
    g = xc.Grid(N, N, L / N, L, L)
    fl = orc.fluid_state(N, N)
    for k in ("eta", "etaco", "rho", "rhow", "rhou", "Drho", "MFS"):
        getattr(fl, k)[:] = m[k]
    kw = dict(sds=-1, top_cnv=-1, bot_cnv=-1, open_cnv=False, dt=1e32, a1=1, a2=1, a3=0)
    return m, par, g, fl, kw


def _dev_bnchm(xc, m):
    return xc.BnchmData(*(xc.to_device(m[k]) for k in ("src_W", "src_U", "src_P", "W_exact", "U_exact", "P_exact")))


def test_operator_and_rhs_mms(xc, orc):
    m, par, g, fl, kw = _mms(orc, xc, 24)
    LL, RR = orc.fluidmech_assemble(fl, None, g, orc.fm_opts(par, **{k: v for k, v in kw.items()}), bnchm=m)
    dfl = to_device_state(xc, fl, "fluid")
    x = np.random.default_rng(0).standard_normal(LL.shape[0])
    y = xc.stokes_apply(dfl, g, par, torch.as_tensor(x, device="cuda"), bnchm=True, **kw).cpu().numpy()
    assert _rel(y, LL @ x) <= 1e-12
    rr = xc.stokes_rhs(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), **kw).cpu().numpy()
    assert _rel(rr, RR) <= 1e-12


@pytest.mark.parametrize("open_cnv", [False, True])
def test_operator_and_rhs_production(xc, orc, open_cnv):
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 20, 24, seed=4)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0, open_cnv=open_cnv)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    LL, RR = orc.fluidmech_assemble(fl, ph, g, orc.fm_opts(par, **kw))
    x = np.random.default_rng(1).standard_normal(LL.shape[0])
    y = xc.stokes_apply(dfl, g, par, torch.as_tensor(x, device="cuda"), bnchm=False, **kw).cpu().numpy()
    assert _rel(y, LL @ x) <= 1e-12
    rr = xc.stokes_rhs(dfl, g, par, phase=dph, **kw).cpu().numpy()
    assert _rel(rr, RR) <= 1e-12
    # the MFS update happened identically on both sides
    assert _rel(xc.to_host(dfl.MFS), fl.MFS) <= 1e-13


@pytest.mark.parametrize("N", [32, 64])
def test_mms_solve_vs_lu(xc, orc, N):
    m, par, g, fl, kw = _mms(orc, xc, N)
    dfl = to_device_state(xc, fl, "fluid")
    stats = xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-8, **kw)
    print(f"MMS N={N}: {stats}", flush=True)
    assert stats["converged"] and stats["rel_resid"] <= 1e-8
    LL, RR = orc.fluidmech(fl, None, g, orc.fm_opts(par, **kw), bnchm=m)
    sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
    # residual judged on the oracle's assembled system
    assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8
    lu = orc.stokes_vector(fl)
    print(f"  |x - x_LU| / |x_LU| = {_rel(sol, lu):.3e}", flush=True)
    assert _rel(sol, lu) <= 1e-5


def test_mms_orders_on_device(xc, orc):
    errs = {"W": [], "U": [], "P": []}
    for N in (32, 64, 128):
        m, par, g, fl, kw = _mms(orc, xc, N)
        dfl = to_device_state(xc, fl, "fluid")
        stats = xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-10, **kw)
        print(f"MMS N={N}: {stats}", flush=True)
        for k in errs:
            ex = m[k + "_exact"]
            errs[k].append(np.linalg.norm(xc.to_host(getattr(dfl, k)) - ex) / np.linalg.norm(ex))
    for k, e in errs.items():
        assert math.log2(e[0] / e[1]) > 1.8 and math.log2(e[1] / e[2]) > 1.8, (k, e)


def test_inertial_mms_orders_on_device(xc, orc):
    # production mode with every term active (tests/mms_inertial.py): second-order convergence of the device
    # solution to the manufactured W, U, P, and agreement with sparse LU within the residual's bound
    from mms_inertial import GAMMA, G0, inertial_case, load_fluid
    errs = {"W": [], "U": [], "P": []}
    for N in (32, 64, 128):
        case, kw = inertial_case(N)
        par = xc.Parameters(D=1.0, N=N, L=1.0, g0=G0, gamma=GAMMA)
        g = xc.Grid(N, N, 1.0 / N, 1.0, 1.0)
        dfl, dph = xc.FluidState(N, N), xc.PhaseState(N, N)
        load_fluid(dfl, case)
        stats = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-10, maxit=3000, **kw)
        print(f"inertial MMS N={N}: {stats}", flush=True)
        assert stats["converged"]
        for k, sl in (("W", np.s_[1:-1, 1:-1]), ("U", np.s_[1:-1, :]), ("P", np.s_[1:-1, 1:-1])):
            ex = case[k + "_exact"][sl]
            errs[k].append(np.linalg.norm(xc.to_host(getattr(dfl, k))[sl] - ex) / np.linalg.norm(ex))
        if N == 64:
            fl, ph = orc.fluid_state(N, N), orc.phase_state(N, N)
            load_fluid(fl, case)
            LL, RR = orc.fluidmech(fl, ph, g, orc.fm_opts(par, **kw))
            sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
            _assert_forward_error(LL, RR, sol, orc.stokes_vector(fl))
    print("inertial MMS errors:", errs, flush=True)
    for k, e in errs.items():
        assert math.log2(e[0] / e[1]) > 1.9 and math.log2(e[1] / e[2]) > 1.9, (k, e)


@pytest.mark.parametrize("shape,open_cnv", [((32, 40), False), ((64, 64), False), ((48, 64), True)])
def test_production_solve(xc, orc, shape, open_cnv):
    # open_cnv pins the pressure on row Nz, ghost columns included (the Schur solve's ghost ring)
    Nz, Nx = shape
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=7)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0, open_cnv=open_cnv)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    stats = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, **kw)
    print(f"production {shape}: {stats}", flush=True)
    assert stats["converged"]
    LL, RR = orc.fluidmech(fl, ph, g, orc.fm_opts(par, **kw))
    sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
    assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8
    _assert_forward_error(LL, RR, sol, orc.stokes_vector(fl))
    if open_cnv:
        # the pinned row Nz is part of the Schur solve: about as many FGMRES iterations as the closed top
        closed = xc.fluidmech_(to_device_state(xc, fl, "fluid"), g, par, phase=dph, tol=1e-8, **dict(kw, open_cnv=False))
        print(f"  closed top on the same state: {closed['iters']} iterations", flush=True)
        assert stats["iters"] <= 2 * closed["iters"], (stats["iters"], closed["iters"])
    # the closed bottom carries the whole divergence and the pinned pressure is zero (fluidmech.jl:170-191,
    # 480-487; both follow from summing the divergence rows)
    W, P, rhow, MFS = (xc.to_host(getattr(dfl, k)) for k in ("W", "P", "rhow", "MFS"))
    if not open_cnv:
        want = MFS.mean() * g.D
        assert np.abs(rhow[-1, :] * W[-1, 1:-1] - want).max() <= 1e-6 * abs(want) + 1e-8 * np.abs(rhow * W[:, 1:-1]).max()
        iz, ix = round(Nz / 2) - 1, round(Nx / 2) - 1   # MapP[round(Nz/2), round(Nx/2)], 1-based
        assert abs(P[iz, ix]) <= 1e-8 * np.abs(P).max()


def _assert_forward_error(LL, RR, sol, lu):
    # forward error against sparse LU within the bound the residual allows: |x - x*| / |x*| <= k(LL) |r| / |b|
    # in the 1-norm, with k estimated from the LU factors (Hager / Higham, scipy's onenormest)
    import scipy.sparse.linalg as spla
    A = LL.tocsc()
    F = spla.splu(A)
    Ainv = spla.LinearOperator(A.shape, matvec=F.solve, rmatvec=lambda v: F.solve(v, trans="T"), dtype=float)
    kappa = spla.onenormest(A) * spla.onenormest(Ainv)
    r = np.abs(A @ sol - RR).sum() / np.abs(RR).sum()
    fe = np.abs(sol - lu).sum() / np.abs(lu).sum()
    print(f"  |x - x_LU|_1 / |x_LU|_1 = {fe:.3e}, kappa_1 ~ {kappa:.2e}, |r|_1/|b|_1 = {r:.2e}", flush=True)
    assert fe <= 2.0 * kappa * r + 1e-12, (fe, kappa, r)


@pytest.mark.parametrize("open_cnv", [False, True])
def test_fine_smoother_precision(xc, orc, open_cnv):
    # the fine level's smoother in single precision (default) or double precision (mg_fp64): the same
    # residual bar against the oracle's assembled system, and about the same iteration count
    Nz, Nx = 64, 64
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=11)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0, open_cnv=open_cnv)
    LL, RR = orc.fluidmech_assemble(copy.deepcopy(fl), copy.deepcopy(ph), g, orc.fm_opts(par, **kw))
    its = {}
    for fp64 in (False, True):
        dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
        stats = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, mg_fp64=fp64, **kw)
        print(f"mg_fp64={fp64} open_cnv={open_cnv}: {stats}", flush=True)
        assert stats["converged"]
        sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
        assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8
        its[fp64] = stats["iters"]
    assert abs(its[False] - its[True]) <= max(2, its[True] // 10), its


@pytest.mark.parametrize("shape", [(64, 64), (128, 96)])
def test_chebyshev_smoothing(xc, orc, shape):
    # the fine level's Chebyshev passes (default where its inertia ratio is >= 1) and damped-Jacobi sweeps
    # (mg_jacobi): both meet the residual bar on the oracle's assembled system, Chebyshev in no more iterations
    Nz, Nx = shape
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=13)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0)
    LL, RR = orc.fluidmech_assemble(copy.deepcopy(fl), copy.deepcopy(ph), g, orc.fm_opts(par, **kw))
    its = {}
    for jacobi in (True, False):
        dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
        stats = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, mg_jacobi=jacobi, **kw)
        print(f"{shape} mg_jacobi={jacobi}: {stats}", flush=True)
        assert stats["converged"] and stats["inertia_ratio"] >= 1.0 and stats["mg_levels"] >= 2
        sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
        assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8
        its[jacobi] = stats["iters"]
    assert its[False] <= its[True], its


@pytest.mark.parametrize("shape", [(48, 64), (256, 32), (512, 16), (2048, 8)])
def test_schur_poisson_solve(xc, orc, shape):
    # the inertial part of the pressure block of z = M^-1 v: with inertia minus (the weighted) without is
    # (a1 h^2 / dt) T^-1 rP on the interior cells, T = -d_zz + a1 / (a1 + gamma) (-d_xx) in grid units
    # (zero flux at the top and bottom rows, periodic in x) on the zero-mean part of rP; checked against a
    # sparse direct solve of T bordered by the constant mode.  (48, 64): mixed-radix transforms; (256, 32):
    # the power-of-two z path; (512, 16): more (k, Nz-k) pairs than lanes
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    Nz, Nx = shape
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=5)
    orc.update(ph, fl, g, par, sc)
    dfl = to_device_state(xc, fl, "fluid")
    n = xc.stokes_vector_size(g)
    v = torch.randn(n, generator=torch.Generator().manual_seed(3), dtype=torch.float64).cuda()
    kw = dict(a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0)
    z_in = xc.stokes_precond(dfl, g, par, v, dt=sc.dt0, **kw)
    z_vis = xc.stokes_precond(dfl, g, par, v, dt=1e32, **kw)
    nv = (Nz + 1) * (Nx + 2) + (Nz + 2) * (Nx + 1)
    blk = lambda t: t[nv:].cpu().numpy().reshape(Nx + 2, Nz + 2).T[1:-1, 1:-1]
    # with inertia the viscous part rP eta / rho carries the weight 0.75 (the library's SCHUR_CV), except at the
    # pressure pin (MapP[round(Nz/2), round(Nx/2)], 1-based, whose row keeps rP / (1 + rho / eta))
    wv = np.full((Nz, Nx), 0.75)
    wv[round(Nz / 2) - 2, round(Nx / 2) - 2] = 1.0
    d = (blk(z_in) - wv * blk(z_vis)).ravel(order="F")
    rP = blk(v).ravel(order="F")
    ratio = 1.5 / (1.5 + par.gamma)
    Dz = sp.diags([-np.ones(Nz - 1), 2 * np.ones(Nz), -np.ones(Nz - 1)], [-1, 0, 1]).tolil()
    Dz[0, 0] = Dz[Nz - 1, Nz - 1] = 1
    Dx = sp.diags([-np.ones(Nx - 1), 2 * np.ones(Nx), -np.ones(Nx - 1)], [-1, 0, 1]).tolil()
    Dx[0, Nx - 1] = Dx[Nx - 1, 0] = -1
    T = sp.kron(sp.eye(Nx), Dz) + ratio * sp.kron(Dx, sp.eye(Nz))
    one = sp.csr_matrix(np.ones((Nz * Nx, 1)))
    y = spl.spsolve(sp.bmat([[T, one], [one.T, None]]).tocsc(), np.concatenate([rP, [0.0]]))[:-1]
    want = 1.5 * g.h ** 2 / sc.dt0 * y
    # both solves are backward stable: the bar is the rounding unit times T's condition number on the
    # zero-mean space (largest eigenvalue over the least nonzero one), with a margin of ~10
    lam = (2 - 2 * np.cos(np.pi * np.arange(Nz) / Nz))[:, None] + \
        ratio * (2 - 2 * np.cos(2 * np.pi * np.arange(Nx) / Nx))[None, :]
    kappa = lam.max() / lam[lam > 1e-14].min()
    print(f"schur poisson {shape}: rel err {_rel(d, want):.3e}, kappa {kappa:.3g}", flush=True)
    assert _rel(d, want) <= 2e-16 * kappa


def test_noconv_reports(xc, orc):
    m, par, g, fl, kw = _mms(orc, xc, 32)
    dfl = to_device_state(xc, fl, "fluid")
    stats = xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-14, maxit=3, raise_on_noconv=False,
                          **kw)
    assert not stats["converged"] and stats["iters"] == 3
    with pytest.raises(xc.NoConvergence):
        xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-14, maxit=2, **kw)


def test_vcycle_depth_rule(xc, orc):
    # with inertia the V-cycle stops at the first level l >= 1 with 4^l R0 >= 6 (R0 = a1 min(rho) h^2 /
    # (dt max(eta))); the solution meets the same bar against the oracle's assembled system whatever the depth
    Nz, Nx = 64, 64
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=9)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0)
    # the assembly updates MFS in the state it is given (fluidmech.jl:115-127): assemble on a copy
    LL, RR = orc.fluidmech_assemble(copy.deepcopy(fl), copy.deepcopy(ph), g, orc.fm_opts(par, **kw))
    R0 = 1.5 * fl.rho.min() * g.h ** 2 / (sc.dt0 * fl.eta.max())
    full = 5  # 64 -> 32 -> 16 -> 8 -> 4
    want = 2  # at least two levels
    while want < full and R0 * 4.0 ** (want - 1) < 6.0:
        want += 1
    for levels in (0, 1, 3, full):
        dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
        stats = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, mg_levels=levels, **kw)
        print(f"mg_levels={levels}: {stats} (R0 {R0:.3g}, rule {want})", flush=True)
        assert stats["converged"]
        if levels == 0:
            assert stats["mg_levels"] == want
            assert abs(stats["inertia_ratio"] - R0) <= 1e-12 * R0
        else:
            assert stats["mg_levels"] == levels
        sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
        assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8


def test_vcycle_full_depth_without_inertia(xc, orc):
    m, par, g, fl, kw = _mms(orc, xc, 32)
    dfl = to_device_state(xc, fl, "fluid")
    stats = xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-8, **kw)
    assert stats["converged"] and stats["mg_levels"] == 4 and stats["inertia_ratio"] == 0.0
    with pytest.raises(xc.XCoreError):
        xc.fluidmech_(dfl, g, par, bnchm_data=_dev_bnchm(xc, m), tol=1e-8, mg_levels=-1, **kw)
