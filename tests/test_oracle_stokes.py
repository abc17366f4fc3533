"""Pins for the oracle's Stokes assembly + sparse LU solve (fluidmech.jl).

test_mms.jl:4-71: manufactured solution, free-slip top/bottom, periodic sides,
W/U/P relative L2 error order > 1.8 over N = 32, 64, 128.  Plus structural
facts of the assembled operator.
"""
import math

import numpy as np
import pytest

from mms import mms_sources


def _mms_case(orc, xc, N, L=10.0):
    m = mms_sources(N, N, L)
    par = xc.Parameters(D=L, N=N, L=L, etam0=1e3, g0=10.0, rhom0=2700, rhox0=3200, d0=1e-2, Da=0.0, Xi=0.0,
                        gamma=0.0)
    g = xc.Grid(N, N, L / N, L, L)
    fl = orc.fluid_state(N, N)
    for k in ("eta", "etaco", "rho", "rhow", "rhou", "Drho", "MFS"):
        getattr(fl, k)[:] = m[k]
    opts = orc.fm_opts(par, sds=-1, top_cnv=-1, bot_cnv=-1, open_cnv=False, dt=1e32, a1=1, a2=1, a3=0)
    return m, par, g, fl, opts


def test_mms_convergence(orc):
    import xcore_amd as xc
    errs = {"W": [], "U": [], "P": []}
    for N in (32, 64, 128):
        m, par, g, fl, opts = _mms_case(orc, xc, N)
        orc.fluidmech(fl, None, g, opts, bnchm=m)
        for k in errs:
            ex = m[k + "_exact"]
            errs[k].append(np.linalg.norm(getattr(fl, k) - ex) / np.linalg.norm(ex))
    for k, e in errs.items():
        assert math.log2(e[0] / e[1]) > 1.8, (k, e)
        assert math.log2(e[1] / e[2]) > 1.8, (k, e)


def test_solution_satisfies_assembled_system(orc):
    # the Newton-style scaled update returns SOL = LL \ RR (fluidmech.jl:503-532)
    import xcore_amd as xc
    m, par, g, fl, opts = _mms_case(orc, xc, 24)
    LL, RR = orc.fluidmech(fl, None, g, opts, bnchm=m)
    sol = orc.stokes_vector(fl)
    assert np.linalg.norm(LL @ sol - RR) <= 1e-10 * np.linalg.norm(RR)


def test_operator_structure(orc):
    # BC rows: periodic W ghost columns, free-slip U rows, P ghost rows; pins
    import xcore_amd as xc
    N = 12
    m, par, g, fl, opts = _mms_case(orc, xc, N)
    LL, RR = orc.fluidmech_assemble(fl, None, g, opts, bnchm=m)
    NW, NU = (N + 1) * (N + 2), (N + 2) * (N + 1)
    MW = lambda r, c: r + c * (N + 1)
    MU = lambda r, c: NW + r + c * (N + 2)
    MP = lambda r, c: NW + NU + r + c * (N + 2)
    L = LL.toarray()
    # W[:,0] - W[:,Nx] = 0 (sds = -1)
    assert L[MW(3, 0), MW(3, 0)] == 1 and L[MW(3, 0), MW(3, N)] == -1
    # free-slip: U[0,c] - U[1,c] = 0
    assert L[MU(0, 4), MU(0, 4)] == 1 and L[MU(0, 4), MU(1, 4)] == -1
    # P ghost rows
    assert L[MP(0, 5), MP(0, 5)] == 1 and L[MP(0, 5), MP(1, 5)] == -1
    # MMS pins: P at (round(34*3/8), round(14/2)) 1-based, U at (7, 7)
    nzp, nxp = round((N + 2) * 3 / 8) - 1, round((N + 2) / 2) - 1
    row = L[MP(nzp, nxp)]
    assert row[MP(nzp, nxp)] == 1 and np.count_nonzero(row) == 1
    assert RR[MP(nzp, nxp)] == m["P_exact"][nzp, nxp]
    # interior divergence row is rho-weighted and sums (over unit velocities) to the face-density difference
    r, c = 5, 6
    drow = L[MP(r, c), :NW + NU]
    h = g.h
    np.testing.assert_allclose(drow[MU(r, c)], fl.rhou[r - 1, c] / h)
    np.testing.assert_allclose(drow[MW(r - 1, c)], -fl.rhow[r - 1, c - 1] / h)
    # viscous W row: zero row sum over W couplings for constant viscosity ... (here eta varies);
    # symmetric part: KV[W_i, W_j] == KV[W_j, W_i] for interior neighbours
    i1, i2 = MW(5, 5), MW(6, 5)
    assert abs(L[i1, i2] - L[i2, i1]) <= 1e-12 * abs(L[i1, i2])


def test_production_pin_keeps_divergence_row(orc):
    # non-benchmark mode: KP[np0,np0] = 1 with the DM row kept (fluidmech.jl:478-487)
    import xcore_amd as xc
    from fieldsetup import plausible_state
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 10, 10, seed=3)
    orc.update(ph, fl, g, par, sc)
    opts = orc.fm_opts(par, dt=sc.dt0)
    LL, RR = orc.fluidmech_assemble(fl, ph, g, opts)
    N = 10
    NW, NU = (N + 1) * (N + 2), (N + 2) * (N + 1)
    np0 = NW + NU + (5 - 1) + (5 - 1) * (N + 2)
    row = LL.getrow(np0).toarray().ravel()
    assert row[np0] == 1 and np.count_nonzero(row[:NW + NU]) == 4
