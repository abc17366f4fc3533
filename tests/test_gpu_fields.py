"""GPU parity of the field updates against the CPU oracle.

Bar: every output array within 1e-12 normwise relative of the oracle (pow/exp
and the chunked row means may differ in the last bits; min/max reductions are
exact).
"""
import numpy as np
import pytest
import torch

from fieldsetup import plausible_state, to_device_state

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _cmp(host, dev, names=None, tol=TOL):
    bad = []
    for n, a in host.arrays().items():
        if names is not None and n not in names:
            continue
        g = getattr(dev, n).cpu().numpy()
        if a.dtype == np.uint8:
            if not np.array_equal(g, a):
                bad.append((n, "bool mismatch"))
            continue
        s = np.abs(a).max()
        d = np.abs(g - a).max()
        if (d / s if s > 0 else d) > tol:
            bad.append((n, d / s if s > 0 else d))
    assert not bad, bad


@pytest.mark.parametrize("shape,bcs", [((24, 32), ("periodic", "closed")), ((37, 50), ("periodic", "closed")),
                                       ((64, 64), ("closed", "periodic")), ((130, 96), ("periodic", "periodic"))])
def test_update_parity(xc, orc, shape, bcs):
    Nz, Nx = shape
    xbc, zbc = bcs
    print(f"update! {shape} {bcs}", flush=True)
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=Nz)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    orc.update(ph, fl, g, par, sc, xbc=xbc, zbc=zbc)
    xc.update_(dph, dfl, g, par, sc, xBC=xbc, zBC=zbc)
    torch.cuda.synchronize()
    _cmp(ph, dph)
    _cmp(fl, dfl)


@pytest.mark.parametrize("scheme", ["weno5", "quick", "centr", "tvdim"])
def test_phsevo_parity(xc, orc, scheme):
    Nz, Nx = 40, 56
    print(f"phsevo! {scheme}", flush=True)
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=11)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    _, _, _, ph0, fl0, _ = plausible_state(orc, xc, Nz, Nx, seed=11)
    coef = orc.time_coefs("bd2im", 3)
    orc.phsevo(ph, fl, g, par, sc, scheme=scheme, coef=coef, dt=0.05)
    a1, a2, a3, b1, b2, b3 = coef
    dph2, dfl2 = to_device_state(xc, ph0, "phase"), to_device_state(xc, fl0, "fluid")
    with xc.exact_arithmetic():
        xc.phsevo_(dph, dfl, g, par, sc, ADVN=scheme, a1=a1, a2=a2, a3=a3, b1=b1, b2=b2, b3=b3, dt=0.05)
    xc.phsevo_(dph2, dfl2, g, par, sc, ADVN=scheme, a1=a1, a2=a2, a3=a3, b1=b1, b2=b2, b3=b3, dt=0.05)
    torch.cuda.synchronize()
    # the stencil parts are bit-exact; the pointwise tail keeps the reference order too
    _cmp(ph, dph, tol=0.0)
    _cmp(fl, dfl, tol=0.0)
    # default arithmetic (one-division WENO5 weights): within the 1e-12 stencil bar
    _cmp(ph, dph2, tol=1e-12)


def test_store_and_phase_velocities_parity(xc, orc):
    Nz, Nx = 33, 47
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=21)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    dns = to_device_state(xc, ns, "noise", g, sc)
    orc.store_previous(fl, ph)
    xc.store_previous_(dfl, dph)
    orc.update_phase_velocities(ph, fl, ns, g, par)
    xc.update_phase_velocities_(dph, dfl, dns, g, par)
    torch.cuda.synchronize()
    _cmp(ph, dph, tol=0.0)
    _cmp(fl, dfl, tol=0.0)


def test_update_dt_parity(xc, orc):
    Nz, Nx = 50, 70
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=31)
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    for dtp in (1e-6, 1.0, 1e3):
        assert xc.update_dt(dph, dfl, g, par, dtp) == orc.update_dt(ph, fl, g, par, dtp)


def test_time_coefs_device_table(xc, orc):
    for t in ("be1im", "bd2im", "cn2si", "bd2si"):
        for s in (1, 2, 3, 10):
            assert xc.time_coefs(t, s) == orc.time_coefs(t, s)
    with pytest.raises(ValueError):
        xc.time_coefs("bogus", 1)


def test_shape_errors(xc):
    g = xc.Grid(8, 8, 1.0, 8.0, 8.0)
    par = xc.Parameters()
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(8, 8), xc.FluidState(8, 9)  # wrong fluid size
    with pytest.raises(AssertionError):
        xc.update_(ph, fl, g, par, sc)
    with pytest.raises(ValueError):
        xc.update_(ph, xc.FluidState(8, 8), g, par, sc, xBC="bogus")
