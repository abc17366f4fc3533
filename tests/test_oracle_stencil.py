"""Pins for the oracle's advection / diffusion restatement.

The reference (xcore.jl) ships no golden data; its stencil tests are
convergence and fixture tests.  They are reproduced here on the oracle,
together with closed-form facts a plausible slip (swapped operand, wrong sign,
wrong index, dropped term) would break.
"""
import math

import numpy as np
import pytest

SCHEMES = ["centr", "upwd1", "quick", "fromm", "weno3", "weno5", "tvdim"]


def _grid(N, L=2 * math.pi):
    h = L / N
    c = (np.arange(1, N + 1) - 0.5) * h
    XX = np.asfortranarray(np.repeat(c[None, :], N, axis=0))
    ZZ = np.asfortranarray(np.repeat(c[:, None], N, axis=1))
    return h, XX, ZZ


def _advect_halo(orc, f_int, u, w, h, scheme, zbc="periodic", xbc="periodic"):
    g = orc.scheme_halo(scheme)
    N = f_int.shape[0]
    fh = orc.zeros(N + 2 * g, f_int.shape[1] + 2 * g)
    orc.embed_interior(fh, f_int, g)
    orc.fill_ghosts(fh, g, xbc=xbc, zbc=zbc)
    adv = orc.zeros(*f_int.shape)
    orc.advect(adv, fh, u, w, h, scheme)
    return adv


def test_fill_ghosts_fixture(orc):
    # test_stencils.jl:37-56 -- 4x4 periodic-x / closed-z
    f = np.asfortranarray(np.arange(1, 17, dtype=np.float64).reshape(4, 4))
    fh = orc.zeros(6, 6)
    orc.embed_interior(fh, f, 1)
    orc.fill_ghosts(fh, 1, xbc="periodic", zbc="closed")
    assert np.array_equal(fh[1:5, 0], fh[1:5, 4])
    assert np.array_equal(fh[1:5, 5], fh[1:5, 1])
    assert np.array_equal(fh[0, 1:5], fh[1, 1:5])
    assert np.array_equal(fh[5, 1:5], fh[4, 1:5])
    # and the interior values themselves: column 0 holds f[:, 3]
    assert np.array_equal(fh[1:5, 0], f[:, 3])
    # corners: x first then z, so fh[0,0] = f[0, 3]
    assert fh[0, 0] == f[0, 3] and fh[5, 5] == f[3, 0]


def test_fill_ghosts_wide_halo_periodic(orc):
    f = np.asfortranarray(np.random.default_rng(1).standard_normal((7, 9)))
    g = 3
    fh = orc.zeros(7 + 2 * g, 9 + 2 * g)
    orc.embed_interior(fh, f, g)
    orc.fill_ghosts(fh, g, xbc="periodic", zbc="periodic")
    ext = np.pad(f, g, mode="wrap")
    assert np.array_equal(fh, ext)
    fh2 = orc.zeros(7 + 2 * g, 9 + 2 * g)
    orc.embed_interior(fh2, f, g)
    orc.fill_ghosts(fh2, g, xbc="closed", zbc="closed")
    assert np.array_equal(fh2, np.pad(f, g, mode="edge"))


def test_fill_ghosts_bad_bc(orc):
    fh = orc.zeros(6, 6)
    with pytest.raises(orc.OracleError):
        orc.fill_ghosts(fh, 1, xbc="bogus")


@pytest.mark.parametrize("scheme,lo,order_min", [("centr", None, 1.9), ("weno5", None, 4.0)])
def test_advect_convergence(orc, scheme, lo, order_min):
    # test_advect.jl:1-78
    U0, W0 = 0.7, -0.3
    errs = []
    for N in (32, 64, 128):
        h, XX, ZZ = _grid(N)
        f = np.asfortranarray(np.cos(XX) * np.cos(ZZ))
        u = np.full((N, N + 1), U0, order="F")
        w = np.full((N + 1, N), W0, order="F")
        adv = _advect_halo(orc, f, u, w, h, scheme)
        exact = -U0 * np.sin(XX) * np.cos(ZZ) + W0 * np.cos(XX) * (-np.sin(ZZ))
        errs.append(math.sqrt(np.sum((adv - exact) ** 2) / (N * N)))
    assert math.log2(errs[0] / errs[1]) > order_min
    assert math.log2(errs[1] / errs[2]) > (3.5 if scheme == "weno5" else order_min)


@pytest.mark.parametrize("scheme,expect", [("upwd1", 0.9), ("quick", 2.5), ("fromm", 1.9),
                                           ("weno3", 1.9), ("tvdim", 1.4)])
def test_advect_convergence_other_schemes(orc, scheme, expect):
    # no reference test: observed order on a smooth cosine, both velocity signs
    errs = []
    for N in (64, 128, 256):
        h, XX, ZZ = _grid(N)
        f = np.asfortranarray(np.cos(XX) * np.cos(ZZ))
        u = np.full((N, N + 1), -0.6, order="F")
        w = np.full((N + 1, N), 0.4, order="F")
        adv = orc.zeros(N, N)
        orc.advect_centered(adv, f, u, w, h, scheme, xbc="periodic", zbc="periodic")
        exact = 0.6 * np.sin(XX) * np.cos(ZZ) + 0.4 * np.cos(XX) * (-np.sin(ZZ))
        errs.append(math.sqrt(np.sum((adv - exact) ** 2) / (N * N)))
    assert math.log2(errs[1] / errs[2]) > expect


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advect_linear_profile_exact(orc, scheme):
    # f linear in x, u uniform, periodic-z/closed-x: every scheme reconstructs the
    # face value of a linear profile exactly in the interior, so adv = u * slope.
    N = 24
    h = 0.1
    g = orc.scheme_halo(scheme)
    xs = (np.arange(N) + 0.5) * h
    f = np.asfortranarray(np.repeat((3.0 + 2.0 * xs)[None, :], N, axis=0))
    for U0 in (0.5, -0.5):
        u = np.full((N, N + 1), U0, order="F")
        w = np.zeros((N + 1, N), order="F")
        if scheme == "tvdim":
            u = np.full((N, N + 3), U0, order="F")
            w = np.zeros((N + 3, N), order="F")
        adv = _advect_halo(orc, f, u, w, h, scheme, xbc="closed", zbc="periodic")
        inner = adv[:, g + 1:N - g - 1]
        np.testing.assert_allclose(inner, U0 * 2.0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_advect_flux_identity_and_conservation(orc, scheme):
    # advect.jl:714-715: adv == div(q)/h bitwise; telescoping sum of adv*h^2
    # equals the boundary flux line integral.
    rng = np.random.default_rng(7)
    Nz, Nx = 13, 17
    h = 0.25
    f = np.asfortranarray(rng.standard_normal((Nz, Nx)))
    u = np.asfortranarray(rng.standard_normal((Nz, Nx + 1)))
    u[:, Nx] = u[:, 0]  # periodic x: face Nx is face 0
    w = np.asfortranarray(rng.standard_normal((Nz + 1, Nx)))
    adv = orc.zeros(Nz, Nx)
    qx = orc.zeros(Nz + 2, Nx + 1)
    qz = orc.zeros(Nz + 1, Nx + 2)
    orc.advect_centered(adv, f, u, w, h, scheme, xbc="periodic", zbc="closed", qx=qx, qz=qz)
    qxi = qx[1:Nz + 1, :]
    qzi = qz[:, 1:Nx + 1]
    div = (qxi[:, 1:] - qxi[:, :-1] + qzi[1:, :] - qzi[:-1, :]) * (1.0 / h)
    assert np.array_equal(div, adv)
    # padding: closed z repeats rows, periodic x wraps columns
    assert np.array_equal(qx[0], qx[1]) and np.array_equal(qx[-1], qx[-2])
    assert np.array_equal(qz[:, 0], qz[:, Nx]) and np.array_equal(qz[:, -1], qz[:, 1])
    # periodic x: the wrap faces 0 and Nx carry the same flux
    np.testing.assert_array_equal(qxi[:, 0], qxi[:, Nx])
    tot = adv.sum() * h * h
    bnd = (qzi[-1].sum() - qzi[0].sum()) * h
    assert abs(tot - bnd) < 1e-12 * max(1.0, np.abs(adv).sum() * h * h)


def test_advect_errors(orc):
    adv = orc.zeros(8, 8)
    fh = orc.zeros(14, 14)
    u = orc.zeros(8, 9)
    w = orc.zeros(9, 8)
    with pytest.raises(orc.OracleError) as e:
        orc.advect(adv, fh, u, w, 0.1, "bogus")
    assert e.value.code == 2
    with pytest.raises(orc.OracleError) as e:
        orc.advect(adv, orc.zeros(12, 12), u, w, 0.1, "weno5")
    assert e.value.code == 1


def test_diffus_convergence(orc):
    # test_diffus.jl:1-35 -- k == 1: dff -> Laplacian f = -2 f
    errs = []
    for N in (32, 64, 128):
        h, XX, ZZ = _grid(N)
        f = np.asfortranarray(np.sin(XX) * np.sin(ZZ))
        k = np.ones((N, N), order="F")
        fh = orc.zeros(N + 2, N + 2)
        kh = orc.zeros(N + 2, N + 2)
        orc.embed_interior(fh, f, 1)
        orc.embed_interior(kh, k, 1)
        orc.fill_ghosts(fh, 1, xbc="periodic", zbc="periodic")
        orc.fill_ghosts(kh, 1, xbc="periodic", zbc="periodic")
        dff = orc.zeros(N, N)
        orc.diffus(dff, fh, kh, h, 1)
        errs.append(math.sqrt(np.sum((dff + 2 * f) ** 2) / (N * N)))
    assert math.log2(errs[0] / errs[1]) > 1.9
    assert math.log2(errs[1] / errs[2]) > 1.9


def test_diffus_variable_k_closed_form(orc):
    # variable k, closed form on the 5-point stencil: with k = a + b*x (x only)
    # and f = x^2 the discrete operator is exact: d/dx(k df/dx) = 2a + 4bx.
    N = 20
    h = 0.05
    xs = (np.arange(N) + 0.5) * h
    a, b = 1.5, 0.75
    f = np.asfortranarray(np.repeat((xs ** 2)[None, :], N, axis=0))
    k = np.asfortranarray(np.repeat((a + b * xs)[None, :], N, axis=0))
    dff = orc.zeros(N, N)
    orc.diffus_centered(dff, f, k, h, xbc="closed", zbc="periodic")
    exact = 2 * a + 4 * b * xs
    np.testing.assert_allclose(dff[:, 1:-1], np.repeat(exact[None, 1:-1], N, axis=0), rtol=1e-10)


def test_diffus_flux_telescoping(orc):
    rng = np.random.default_rng(3)
    Nz, Nx = 11, 9
    h = 0.3
    f = np.asfortranarray(rng.standard_normal((Nz, Nx)))
    k = np.asfortranarray(np.exp(rng.standard_normal((Nz, Nx))))
    dff = orc.zeros(Nz, Nx)
    qx = orc.zeros(Nz + 2, Nx + 1)
    qz = orc.zeros(Nz + 1, Nx + 2)
    orc.diffus_centered(dff, f, k, h, xbc="periodic", zbc="closed", qx=qx, qz=qz)
    qxi = qx[1:Nz + 1, :]
    qzi = qz[:, 1:Nx + 1]
    # dff = -div(q) up to rounding (different operation order in the reference)
    div = (qxi[:, 1:] - qxi[:, :-1] + qzi[1:, :] - qzi[:-1, :]) / h
    np.testing.assert_allclose(dff, -div, rtol=1e-12, atol=1e-12 * np.abs(dff).max())
    # closed z: boundary faces carry zero flux (repeat ghost => zero gradient)
    assert np.all(qzi[0] == 0) and np.all(qzi[-1] == 0)
    assert abs(dff.sum()) < 1e-12 * np.abs(dff).sum()
