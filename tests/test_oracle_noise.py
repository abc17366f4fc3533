"""Pins for the oracle's noise restatement (noise.jl) and its white-noise generator.

The reference's noise tests check shapes, finiteness and reuse only
(test_noise.jl); its FFT values are pinned here by numpy's FFT (a library
routine), committed golden vectors, eigen-mode and constant-field closed forms,
and the discrete divergence-free property of the eddy noise.
"""
import json
import math
import os

import numpy as np
import pytest

from fieldsetup import plausible_state

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_philox_known_answers(orc):
    # Random123 philox4x32-10 known-answer vectors
    assert orc.philox4x32_10((0, 0, 0, 0), (0, 0)) == (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)
    assert orc.philox4x32_10((0xffffffff,) * 4, (0xffffffff,) * 2) == (0x408f276d, 0x41c83b0e, 0xa20bc7c6,
                                                                       0x6d5451fd)
    assert orc.philox4x32_10((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0)) == (
        0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)


def test_normal_fill_statistics(orc):
    from scipy import stats
    a = orc.zeros(301, 333)
    orc.normal_fill(a, seed=42, stream=0, draw=3)
    v = a.ravel(order="F")
    assert abs(v.mean()) < 5 / math.sqrt(v.size)
    assert abs(v.var() - 1) < 0.02
    assert stats.kstest(v, "norm").pvalue > 1e-3
    b = orc.zeros(301, 333)
    orc.normal_fill(b, seed=42, stream=0, draw=3)
    assert np.array_equal(a, b)
    orc.normal_fill(b, seed=42, stream=1, draw=3)
    assert abs(np.corrcoef(a.ravel(), b.ravel())[0, 1]) < 0.02


def _numpy_filter(psi, pad, h, sigma):
    Nz, Nx = psi.shape
    hp = pad // 2
    a = np.zeros((Nz + pad, Nx), dtype=complex)
    a[hp:hp + Nz] = psi
    kz = 2 * np.pi * np.fft.fftfreq(Nz + pad, d=h)
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=h)
    G = np.exp(-sigma ** 2 * (kz[:, None] ** 2 + kx[None, :] ** 2))
    out = np.real(np.fft.ifft2(np.fft.fft2(a) * G))[hp:hp + Nz]
    rs, rm = psi.std(ddof=1), psi.mean()
    fs, fm = out.std(ddof=1), out.mean()
    if fs > 2.220446049250313e-16:
        out = (out - fm) * rs / fs + rm
    return out


@pytest.mark.parametrize("Nz,Nx,pad,sigma_h", [(24, 20, 8, 1.5), (100, 100, 16, 3.1), (31, 26, 12, 2.0)])
def test_fft_filter_vs_numpy(orc, Nz, Nx, pad, sigma_h):
    h = 0.1
    rng = np.random.default_rng(Nz)
    psi = np.asfortranarray(rng.standard_normal((Nz, Nx)))
    got = orc.fft_filter(psi, pad, h, sigma_h * h)
    want = _numpy_filter(psi, pad, h, sigma_h * h)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_fft_filter_golden(orc):
    # committed vectors (tests/golden/make_noise_golden.py, numpy pocketfft)
    with open(os.path.join(GOLDEN, "fft_filter_24x20.json")) as fh:
        g = json.load(fh)
    psi = np.asfortranarray(np.array(g["psi"]))
    got = orc.fft_filter(psi, g["pad"], g["h"], g["sigma"])
    want = np.array(g["out"])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_fft_filter_constant_field_unchanged(orc):
    psi = np.full((20, 16), 3.25, order="F")
    out = orc.fft_filter(psi, 8, 0.1, 0.2)
    assert np.array_equal(out, psi)


def test_fft_filter_cosine_mode(orc):
    # psi = cos(2 pi m x / L), constant in z: x is periodic and unpadded, so every
    # output row stays proportional to the same cosine (plus the global rescale offset)
    Nz, Nx, m = 18, 32, 3
    h = 1.0 / Nx
    x = (np.arange(Nx) + 0.5) * h
    mode = np.cos(2 * np.pi * m * x)
    psi = np.asfortranarray(np.repeat(mode[None, :], Nz, axis=0))
    out = orc.fft_filter(psi, 8, h, 2 * h)
    off = out.mean(axis=1, keepdims=True)
    coef = (out - off) @ mode / (mode @ mode)
    resid = out - off - coef[:, None] * mode[None, :]
    assert np.abs(resid).max() < 1e-12 * np.abs(out).max()
    # the zero padding damps rows near the top/bottom more than the middle
    assert coef[Nz // 2] > coef[0] > 0


def test_noise_step_invariants(orc):
    import xcore_amd as xc
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 24, 28, seed=7)
    orc.update(ph, fl, g, par, sc)
    orc.store_noise(ns)
    orc.noise(ns, ph, g, par, sc, sc.dt0, first_iter=True, seed=5, draw=0)
    re1 = ns.re.copy()
    # test_noise.jl:78-92: no redraw on later sweeps
    orc.noise(ns, ph, g, par, sc, sc.dt0, first_iter=False, seed=5, draw=1)
    assert np.array_equal(ns.re, re1)
    for n in ("xie", "xix", "xis"):
        a = getattr(ns, n)
        assert np.all(np.isfinite(a)) and np.all(a >= 0)
    # eddy noise comes from a stream function: discrete divergence is zero
    U, W = ns.xieu, ns.xiew
    div = (U[1:-1, 1:] - U[1:-1, :-1]) + (W[1:, 1:-1] - W[:-1, 1:-1])
    assert np.abs(div).max() < 1e-12 * max(np.abs(U).max(), np.abs(W).max())
    # OU update formula at the cell level (before filtering), noise.jl:160-183
    V, dt = ph.V, sc.dt0
    taue = sc.L0h / (2 * (V + 2.220446049250313e-16))
    sge = par.Xi * np.sqrt(np.maximum(0, ph.fReL * ph.ke / taue))
    taue2 = 1 / (1 / taue + 1 / (100 * dt)) + dt
    Fte = np.exp(-dt / taue2)
    want = Fte * ns.psieo + np.sqrt(np.maximum(0, 1 - Fte ** 2)) * sge * ns.fL * ns.re
    np.testing.assert_allclose(ns.psie, want, rtol=1e-12, atol=1e-300)


def test_store_noise(orc):
    # test_noise.jl:44-59
    ns = orc.HostNoise(6, 5, 8, 8, 1.0, 1.0)
    ns.psie[:] = 1.0
    ns.psix[:] = 2.0
    ns.psis[:] = 3.0
    orc.store_noise(ns)
    assert np.all(ns.psieo == 1.0) and np.all(ns.psixo == 2.0) and np.all(ns.psiso == 3.0)
    ns.psie[:] = 7.0
    orc.store_noise(ns)
    assert np.all(ns.psieo == 7.0)
