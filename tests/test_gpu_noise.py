"""GPU parity of the OU noise (noise.jl) against the CPU oracle.

Bars: white noise -- the uniforms are integer-exact Philox outputs, the normals
match to 1e-14 (log/cos/sin last-bit differences); FFT filter and every noise!
output -- 1e-12 normwise relative.
"""
import numpy as np
import pytest
import torch

from fieldsetup import plausible_state, to_device_state

pytestmark = pytest.mark.gpu


def _rel(g, w):
    s = np.abs(w).max()
    return np.abs(g - w).max() / (s if s > 0 else 1.0)


def test_normal_fill_parity(xc, orc):
    for (r, c), seed, stream, draw in (((37, 41), 7, 0, 0), ((256, 256), 2**40 + 3, 1, 5), ((1, 1), 1, 0, 2)):
        want = orc.zeros(r, c)
        orc.normal_fill(want, seed, stream, draw)
        got = xc.xcore_zeros(r, c)
        xc.normal_fill_(got, seed, stream, draw)
        assert _rel(xc.to_host(got), want) < 1e-14


# Filters several cells wide run as direct convolutions (both passes at (100, 100) and (512, 512); x only
# at (40, 200); z only at (200, 40); z over two 512-row segments at (600, 24)); the rest as FFT passes.
@pytest.mark.parametrize("Nz,Nx,pad,sig", [(24, 20, 8, 1.5), (100, 100, 16, 3.1), (31, 26, 12, 2.0),
                                           (104, 100, 4, 0.7), (512, 512, 16, 3.06), (40, 200, 8, 3.0),
                                           (200, 40, 16, 3.0), (600, 24, 16, 3.0)])
def test_fft_filter_parity(xc, orc, Nz, Nx, pad, sig):
    print(f"fft_filter {Nz}x{Nx} pad {pad}", flush=True)
    h = 1.0 / Nz
    rng = np.random.default_rng(Nz + Nx)
    psi = np.asfortranarray(rng.standard_normal((Nz, Nx)))
    want = orc.fft_filter(psi, pad, h, sig * h)
    got = xc.compute_fft_filter(xc.to_device(psi), pad, h, sig * h)
    assert _rel(xc.to_host(got), want) <= 1e-12


def test_fft_filter_constant_unchanged(xc):
    psi = xc.to_device(np.full((40, 32), -1.75))
    out = xc.compute_fft_filter(psi, 8, 0.1, 0.3)
    assert np.array_equal(xc.to_host(out), np.full((40, 32), -1.75))


@pytest.mark.parametrize("shape", [(24, 28), (64, 64), (100, 100)])
def test_noise_step_parity(xc, orc, shape):
    Nz, Nx = shape
    print(f"noise! {shape}", flush=True)
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=Nz)
    orc.update(ph, fl, g, par, sc)
    rng = np.random.default_rng(1)
    for n in ("psieo", "psixo", "psiso"):
        getattr(ns, n)[:] = 1e-3 * rng.standard_normal((Nz, Nx))
    dph = to_device_state(xc, ph, "phase")
    dns = to_device_state(xc, ns, "noise", g, sc)
    dns.seed = 99
    xc.noise_(dns, dph, g, par, sc, sc.dt0, first_iter=True)
    assert dns.draws == 1
    torch.cuda.synchronize()
    # identical white noise for the oracle: the device draw vs the oracle's own generator
    re = xc.to_host(dns.re)
    ref = orc.zeros(Nz, Nx)
    orc.normal_fill(ref, 99, 0, 0)
    assert _rel(re, ref) < 1e-14
    ns.re[:] = re
    ns.rs[:] = xc.to_host(dns.rs)
    orc.noise(ns, ph, g, par, sc, sc.dt0, first_iter=False)
    bad = [(n, _rel(xc.to_host(getattr(dns, n)), a)) for n, a in ns.arrays().items()
           if _rel(xc.to_host(getattr(dns, n)), a) > 1e-12]
    assert not bad, bad
    # later Picard sweeps reuse the draw (test_noise.jl:78-92)
    xc.noise_(dns, dph, g, par, sc, sc.dt0, first_iter=False)
    assert dns.draws == 1
    assert np.array_equal(xc.to_host(dns.re), re)


def test_store_noise_device(xc):
    g = xc.Grid(6, 5, 1.0, 6.0, 5.0)
    sc = xc.compute_scales(xc.Parameters(D=6.0, N=6, L=5.0), g)
    ns = xc.NoiseState(g, sc)
    ns.psie.fill_(1.0)
    ns.psix.fill_(2.0)
    ns.psis.fill_(3.0)
    xc.store_noise_(ns)
    assert bool((ns.psieo == 1).all()) and bool((ns.psixo == 2).all()) and bool((ns.psiso == 3).all())
