"""GPU parity of the advection / diffusion stencils against the CPU oracle.

Bar: <= 1e-12 normwise relative (north-star stencil tolerance); the kernels
keep the reference's operation order without FP contraction, so they are also
checked for bitwise equality.
"""
import numpy as np
import pytest
import torch

SCHEMES = ["centr", "upwd1", "quick", "fromm", "weno3", "weno5", "tvdim"]
SHAPES = [(13, 17), (64, 64), (70, 33), (129, 257)]
BCS = [("periodic", "closed"), ("closed", "periodic"), ("periodic", "periodic"), ("closed", "closed")]

pytestmark = pytest.mark.gpu


def _rel(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d


def _inputs(Nz, Nx, seed, periodic_x=True):
    rng = np.random.default_rng(seed)
    f = np.asfortranarray(rng.standard_normal((Nz, Nx)) + 2.0)
    u = np.asfortranarray(rng.standard_normal((Nz, Nx + 1)))
    w = np.asfortranarray(rng.standard_normal((Nz + 1, Nx)))
    if periodic_x:
        u[:, Nx] = u[:, 0]
    return f, u, w


@pytest.mark.parametrize("scheme", SCHEMES)
@pytest.mark.parametrize("shape", SHAPES)
def test_advect_centered_with_flux_parity(xc, orc, scheme, shape):
    Nz, Nx = shape
    h = 1.0 / Nx
    for bi, (xbc, zbc) in enumerate(BCS):
        print(f"advect_centered_with_flux {scheme} {shape} {xbc}/{zbc}", flush=True)
        f, u, w = _inputs(Nz, Nx, 100 * bi + Nz)
        adv_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
        orc.advect_centered(adv_o, f, u, w, h, scheme, xbc=xbc, zbc=zbc, qx=qx_o, qz=qz_o)
        fd, ud, wd = xc.to_device(f), xc.to_device(u), xc.to_device(w)
        for exact in (True, False):
            adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
            if exact:
                with xc.exact_arithmetic():
                    xc.advect_centered_with_flux_(adv, qx, qz, fd, ud, wd, h, scheme, xBC=xbc, zBC=zbc)
            else:
                xc.advect_centered_with_flux_(adv, qx, qz, fd, ud, wd, h, scheme, xBC=xbc, zBC=zbc)
            torch.cuda.synchronize()
            for got, want, nm in ((adv, adv_o, "adv"), (qx, qx_o, "qx"), (qz, qz_o, "qz")):
                g = xc.to_host(got)
                assert _rel(g, want) <= 1e-12, (nm, exact, _rel(g, want))
                if exact or scheme != "weno5":
                    assert np.array_equal(g, want), (nm, "not bitwise", np.abs(g - want).max())


@pytest.mark.parametrize("scale", [1e-200, 1e30, 1e40, 1e60])
def test_weno5_large_and_small_magnitudes(xc, orc, scale):
    # the default WENO5 weights against the reference's (the oracle) for fields of magnitude 1e-200 .. 1e60:
    # the reference's own weights stay finite up to |df| ~ 1e77; the default form must follow them
    Nz, Nx = 64, 48
    h = 1.0 / Nx
    print(f"WENO5 at field magnitude {scale:g}", flush=True)
    f, u, w = _inputs(Nz, Nx, 7)
    f = np.asfortranarray(f * scale)
    adv_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
    orc.advect_centered(adv_o, f, u, w, h, "weno5", qx=qx_o, qz=qz_o)
    assert np.all(np.isfinite(adv_o))
    adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
    xc.advect_centered_with_flux_(adv, qx, qz, xc.to_device(f), xc.to_device(u), xc.to_device(w), h, "weno5")
    torch.cuda.synchronize()
    for got, want, nm in ((adv, adv_o, "adv"), (qx, qx_o, "qx"), (qz, qz_o, "qz")):
        g = xc.to_host(got)
        assert np.all(np.isfinite(g)), nm
        assert _rel(g, want) <= 1e-12, (nm, scale, _rel(g, want))


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advect_centered_no_flux(xc, orc, scheme):
    Nz, Nx = 47, 90
    h = 0.37
    f, u, w = _inputs(Nz, Nx, 5)
    adv_o = orc.zeros(Nz, Nx)
    orc.advect_centered(adv_o, f, u, w, h, scheme, xbc="periodic", zbc="closed")
    adv = xc.xcore_zeros(Nz, Nx)
    with xc.exact_arithmetic():
        xc.advect_centered_(adv, xc.to_device(f), xc.to_device(u), xc.to_device(w), h, scheme)
    assert np.array_equal(xc.to_host(adv), adv_o)
    adv2 = xc.xcore_zeros(Nz, Nx)
    xc.advect_centered_(adv2, xc.to_device(f), xc.to_device(u), xc.to_device(w), h, scheme)
    assert _rel(xc.to_host(adv2), adv_o) <= 1e-12


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advect_halo_input(xc, orc, scheme):
    # advect! / advect_with_flux! on an explicit halo buffer (ghosts filled by
    # fill_ghosts! on the device), including tvdim's padded velocities.
    Nz, Nx = 40, 75
    h = 0.01
    g = orc.scheme_halo(scheme)
    f, u, w = _inputs(Nz, Nx, 11)
    if scheme == "tvdim":
        rng = np.random.default_rng(12)
        u = np.asfortranarray(rng.standard_normal((Nz, Nx + 3)))
        w = np.asfortranarray(rng.standard_normal((Nz + 3, Nx)))
    fh_o = orc.zeros(Nz + 2 * g, Nx + 2 * g)
    orc.embed_interior(fh_o, f, g)
    orc.fill_ghosts(fh_o, g, xbc="periodic", zbc="closed")
    adv_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
    orc.advect_with_flux(adv_o, qx_o, qz_o, fh_o, u, w, h, scheme, xbc="periodic", zbc="closed")
    fh = xc.xcore_zeros(Nz + 2 * g, Nx + 2 * g)
    xc.embed_interior_(fh, xc.to_device(f), g)
    xc.fill_ghosts_(fh, g, zBC="closed", xBC="periodic")
    assert np.array_equal(xc.to_host(fh), fh_o)
    adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
    with xc.exact_arithmetic():
        xc.advect_with_flux_(adv, qx, qz, fh, xc.to_device(u), xc.to_device(w), h, scheme)
        adv2 = xc.xcore_zeros(Nz, Nx)
        xc.advect_(adv2, fh, xc.to_device(u), xc.to_device(w), h, scheme)
    assert np.array_equal(xc.to_host(adv), adv_o)
    assert np.array_equal(xc.to_host(qx), qx_o)
    assert np.array_equal(xc.to_host(qz), qz_o)
    assert np.array_equal(xc.to_host(adv2), adv_o)


def test_advect_strided_views(xc, orc):
    # phsevo.jl:177-180 passes views of the ghosted phase velocities
    Nz, Nx = 33, 41
    h = 0.1
    rng = np.random.default_rng(9)
    f = np.asfortranarray(rng.standard_normal((Nz, Nx)))
    Ux = np.asfortranarray(rng.standard_normal((Nz + 2, Nx + 1)))
    Wx = np.asfortranarray(rng.standard_normal((Nz + 1, Nx + 2)))
    adv_o = orc.zeros(Nz, Nx)
    orc.advect_centered(adv_o, f, np.asfortranarray(Ux[1:-1, :]), np.asfortranarray(Wx[:, 1:-1]), h, "weno5")
    Uxd, Wxd = xc.to_device(Ux), xc.to_device(Wx)
    adv = xc.xcore_zeros(Nz, Nx)
    with xc.exact_arithmetic():
        xc.advect_centered_(adv, xc.to_device(f), Uxd[1:-1, :], Wxd[:, 1:-1], h, "weno5")
    assert np.array_equal(xc.to_host(adv), adv_o)


@pytest.mark.parametrize("shape", SHAPES)
def test_diffus_parity(xc, orc, shape):
    Nz, Nx = shape
    h = 2.0 / Nz
    rng = np.random.default_rng(Nz)
    for xbc, zbc in BCS:
        f = np.asfortranarray(rng.standard_normal((Nz, Nx)))
        k = np.asfortranarray(np.exp(rng.standard_normal((Nz, Nx))))
        dff_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
        orc.diffus_centered(dff_o, f, k, h, xbc=xbc, zbc=zbc, qx=qx_o, qz=qz_o)
        dff, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
        xc.diffus_centered_with_flux_(dff, qx, qz, xc.to_device(f), xc.to_device(k), h, xBC=xbc, zBC=zbc)
        assert np.array_equal(xc.to_host(dff), dff_o)
        assert np.array_equal(xc.to_host(qx), qx_o)
        assert np.array_equal(xc.to_host(qz), qz_o)
        # halo-buffer entry points
        fh, kh = orc.zeros(Nz + 2, Nx + 2), orc.zeros(Nz + 2, Nx + 2)
        orc.embed_interior(fh, f, 1); orc.embed_interior(kh, k, 1)
        orc.fill_ghosts(fh, 1, xbc=xbc, zbc=zbc); orc.fill_ghosts(kh, 1, xbc=xbc, zbc=zbc)
        dff2 = xc.xcore_zeros(Nz, Nx)
        xc.diffus_(dff2, xc.to_device(fh), xc.to_device(kh), h)
        assert np.array_equal(xc.to_host(dff2), dff_o)
        dff3, qx3, qz3 = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
        xc.diffus_with_flux_(dff3, qx3, qz3, xc.to_device(fh), xc.to_device(kh), h, xBC=xbc, zBC=zbc)
        assert np.array_equal(xc.to_host(qx3), qx_o) and np.array_equal(xc.to_host(qz3), qz_o)


def test_error_mapping(xc):
    adv = xc.xcore_zeros(8, 8)
    f = xc.xcore_zeros(8, 8)
    u = xc.xcore_zeros(8, 9)
    w = xc.xcore_zeros(9, 8)
    with pytest.raises(ValueError):
        xc.advect_centered_(adv, f, u, w, 0.1, "bogus")
    with pytest.raises(ValueError):
        xc.advect_centered_(adv, f, u, w, 0.1, "weno5", xBC="bogus")
    with pytest.raises(AssertionError):
        xc.advect_centered_(adv, f, xc.xcore_zeros(8, 8), w, 0.1, "weno5")
    with pytest.raises(AssertionError):
        xc.advect_(adv, xc.xcore_zeros(12, 12), u, w, 0.1, "weno5")
    with pytest.raises(AssertionError):
        # row-major tensor: wrong layout for the column-major ABI
        xc.advect_centered_(adv, torch.zeros(8, 8, dtype=torch.float64, device="cuda"), u, w, 0.1, "weno5")


def test_weno5_order_on_device(xc):
    # test_advect.jl:38-78 run through the device path
    import math
    U0, W0 = 0.7, -0.3
    errs = []
    for N in (32, 64, 128):
        h = 2 * math.pi / N
        c = (np.arange(1, N + 1) - 0.5) * h
        XX = np.repeat(c[None, :], N, axis=0)
        ZZ = np.repeat(c[:, None], N, axis=1)
        f = xc.to_device(np.cos(XX) * np.cos(ZZ))
        u = xc.to_device(np.full((N, N + 1), U0))
        w = xc.to_device(np.full((N + 1, N), W0))
        adv = xc.xcore_zeros(N, N)
        xc.advect_centered_(adv, f, u, w, h, "weno5", xBC="periodic", zBC="periodic")
        exact = -U0 * np.sin(XX) * np.cos(ZZ) + W0 * np.cos(XX) * (-np.sin(ZZ))
        errs.append(math.sqrt(np.sum((xc.to_host(adv) - exact) ** 2) / (N * N)))
    assert math.log2(errs[0] / errs[1]) > 4.0
    assert math.log2(errs[1] / errs[2]) > 3.5


@pytest.mark.parametrize("scheme", ["weno5", "quick", "tvdim"])
def test_every_tile_shape_bitwise(xc, orc, scheme):
    # the autotuner's candidates compute the same faces in different tilings: all bit-identical
    Nz, Nx = 200, 136
    h = 1.0 / Nx
    f, u, w = _inputs(Nz, Nx, 77)
    adv_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
    orc.advect_centered(adv_o, f, u, w, h, scheme, xbc="periodic", zbc="closed", qx=qx_o, qz=qz_o)
    fd, ud, wd = xc.to_device(f), xc.to_device(u), xc.to_device(w)
    tiles = [(64, 16), (128, 8), (64, 8), (32, 32)] + ([(64, -16), (64, -8)] if scheme != "tvdim" else [])
    if scheme == "weno5":  # the fast form: the same bits in every tiling too
        first = None
        for tz, tx in tiles:
            xc.set_advect_tile(scheme, Nz, Nx, tz, tx)
            adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
            xc.advect_centered_with_flux_(adv, qx, qz, fd, ud, wd, h, scheme)
            got = [xc.to_host(t) for t in (adv, qx, qz)]
            assert _rel(got[0], adv_o) <= 1e-12
            if first is None:
                first = got
            for g, g0 in zip(got, first):
                assert np.array_equal(g, g0), ("fast", tz, tx)
    with xc.exact_arithmetic():
        for tz, tx in tiles:
            print(f"tile {tz}x{tx} {scheme}", flush=True)
            xc.set_advect_tile(scheme, Nz, Nx, tz, tx)
            adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
            xc.advect_centered_with_flux_(adv, qx, qz, fd, ud, wd, h, scheme)
            torch.cuda.synchronize()
            for got, want in ((adv, adv_o), (qx, qx_o), (qz, qz_o)):
                assert np.array_equal(xc.to_host(got), want), (tz, tx)
        with xc.autotune():
            Nz2 = Nz + 8  # a grid not tuned yet
            f2, u2, w2 = _inputs(Nz2, Nx, 78)
            ref = orc.zeros(Nz2, Nx)
            orc.advect_centered(ref, f2, u2, w2, h, scheme, xbc="periodic", zbc="closed")
            adv2 = xc.xcore_zeros(Nz2, Nx)
            xc.advect_centered_(adv2, xc.to_device(f2), xc.to_device(u2), xc.to_device(w2), h, scheme)
            assert np.array_equal(xc.to_host(adv2), ref)
        tz, tx, cand, tuned = xc.advect_tile(scheme, Nz + 8, Nx, with_flux=False)
    print("tuned", scheme, tz, tx, cand, flush=True)
    assert tuned and (tz, tx) in ((64, 16), (128, 8), (64, 8), (32, 32), (64, -16), (64, -8))


@pytest.mark.parametrize("scheme", ["weno5", "quick", "tvdim"])
def test_streaming_store_path(xc, scheme):
    # outputs above 64 MiB per launch are written with non-temporal stores; a one-member batch launch
    # (plain stores, 64 x 8 tiles) gives the reference bits for every tile shape
    Nz, Nx = 2048, 1500
    h = 1.0 / Nx
    print(f"streaming stores {scheme} {Nz}x{Nx}", flush=True)
    f, u, w = _inputs(Nz, Nx, 91)
    fd, ud, wd = xc.to_device(f), xc.to_device(u), xc.to_device(w)
    ref = dict(adv=xc.xcore_zeros(Nz, Nx), qx=xc.xcore_zeros(Nz + 2, Nx + 1), qz=xc.xcore_zeros(Nz + 1, Nx + 2),
               f=fd, u=ud, w=wd, h=h)
    xc.advect_centered_batch_([ref], scheme)
    want = [xc.to_host(ref[k]) for k in ("adv", "qx", "qz")]
    tiles = [(64, 16), (128, 8), (64, 8), (32, 32)] + ([(64, -16), (64, -8)] if scheme != "tvdim" else [])
    for tz, tx in tiles:
        xc.set_advect_tile(scheme, Nz, Nx, tz, tx)
        adv, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
        xc.advect_centered_with_flux_(adv, qx, qz, fd, ud, wd, h, scheme)
        torch.cuda.synchronize()
        for got, wnt in zip((adv, qx, qz), want):
            assert np.array_equal(xc.to_host(got), wnt), (scheme, tz, tx)


def test_streaming_store_path_diffus(xc):
    Nz, Nx = 2048, 1500
    h = 1.0 / Nx
    rng = np.random.default_rng(92)
    f = xc.to_device(np.asfortranarray(rng.standard_normal((Nz, Nx))))
    k = xc.to_device(np.asfortranarray(np.exp(rng.standard_normal((Nz, Nx)))))
    ref = dict(dff=xc.xcore_zeros(Nz, Nx), qx=xc.xcore_zeros(Nz + 2, Nx + 1), qz=xc.xcore_zeros(Nz + 1, Nx + 2),
               f=f, k=k, h=h)
    xc.diffus_centered_batch_([ref])
    dff, qx, qz = xc.xcore_zeros(Nz, Nx), xc.xcore_zeros(Nz + 2, Nx + 1), xc.xcore_zeros(Nz + 1, Nx + 2)
    xc.diffus_centered_with_flux_(dff, qx, qz, f, k, h)
    torch.cuda.synchronize()
    for got, key in ((dff, "dff"), (qx, "qx"), (qz, "qz")):
        assert np.array_equal(xc.to_host(got), xc.to_host(ref[key])), key
