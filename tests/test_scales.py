"""compute_scales (scales.jl:85-185) in the library against the oracle's restatement, the reference's own
fixed-point check (test_scales.jl:4-37) and the ensemble form (one call for many parameter sets)."""
import math

import pytest

import xcore_amd as xc
from xcore_amd.state import SCALES_FULL

EPS = 2.220446049250313e-16


def _ref_par():
    # test_scales.jl:4-9 (mirrors usr/run_D1_dm2_e1.m)
    return xc.Parameters(D=10.0, N=200, L=15.0, d0=1e-2, etam0=1e1, rhom0=2700, rhox0=3200, g0=10.0, L0=10.0 / 100,
                         l0=1e-2 * 10, Da=0.01, Xi=0.5, x0=EPS)


def test_library_equals_oracle(orc):
    for par in (_ref_par(), xc.Parameters(), xc.Parameters(N=512, etacntr=1e3, rhox0=2900.0),
                xc.Parameters(D=1e4, N=100, L=1e4, Da=0.05, x0=0.05)):
        g = xc.Grid.from_par(par)
        lib = xc.compute_scales(par, g)
        ref = orc.compute_scales(par, g)
        assert set(ref) == set(SCALES_FULL)
        for k in SCALES_FULL:
            a, b = getattr(lib, k), ref[k]
            assert abs(a - b) <= 4 * EPS * abs(b), (k, a, b)


def test_reference_fixed_point_and_sanity():
    # test_scales.jl:13-31: the W0 fixed-point residual is zero to rtol 1e-12; positivity and ramp bounds
    par = _ref_par()
    s = xc.compute_scales(par, xc.Grid.from_par(par))
    W0_check = ((math.sqrt(4 / s.Ri0 ** 2 * s.Dchi0 * s.Drho0 * par.g0 * s.rho0 * s.fReL0 * s.L0 ** 2 * s.D0
                           + s.eta0 ** 2) - s.eta0) * s.D0 / (2 * s.fReL0 * s.L0 ** 2 * s.rho0 / s.Ri0 ** 2))
    assert W0_check == pytest.approx(s.W0, rel=1e-12)
    assert s.W0l > 0 and s.W0 > 0 and s.w0 > 0
    assert 0 < s.fReL0 <= 1 and 0 < s.fRel0 <= 1
    assert s.Ra0 > 0 and s.ReD0 >= 0 and s.Red0 >= 0
    # the closed forms derived from W0, w0 (scales.jl:135-176)
    assert s.dt0 == min((s.h0 / 2) ** 2 / s.kx0, (s.h0 / 2) / (s.W0 + s.w0))
    assert s.L0h == (s.L0 + s.h0) / 2 and s.bnd_w == s.l0 / 2 + par.D / 100


def test_ensemble_members_in_one_call():
    # each member's density contrast and viscosity change its scales; the batch equals member-by-member calls
    pars = [xc.Parameters(N=128, rhox0=2700.0 + 200.0 * f, etam0=10.0 ** e)
            for f, e in ((0.8, 1.0), (1.0, 1.5), (1.2, 2.0), (0.9, 0.5))]
    grids = [xc.Grid.from_par(p) for p in pars]
    batch = xc.state.compute_scales_batch(pars, grids)
    for p, g, b in zip(pars, grids, batch):
        one = xc.compute_scales(p, g)
        assert all(getattr(one, k) == getattr(b, k) for k in SCALES_FULL)
    assert len({b.dt0 for b in batch}) == 4


def test_invalid_arguments():
    par = xc.Parameters(D=-1.0)
    with pytest.raises(ValueError):
        xc.compute_scales(par, xc.Grid(8, 8, 0.1, 1.0, 1.0))
