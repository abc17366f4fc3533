"""Pins of the FP64 U-Net oracle (oracle/unet.py): the convolution against
scipy.ndimage.correlate, the hand-written backward pass against central
finite differences of the loss along random directions, the adjoint identities
of pooling and upsampling, and the first Adam step against its closed form."""
import numpy as np
from scipy.ndimage import correlate

from oracle import unet as U


def test_conv_matches_scipy_correlate():
    rng = np.random.default_rng(0)
    H, W = 10, 12
    x = rng.standard_normal((H, W, 3))
    w = rng.standard_normal((2, 9, 3))
    ref = np.zeros((H, W, 2))
    for co in range(2):
        for ci in range(3):
            k = np.zeros((3, 3))
            for dx in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    k[dz + 1, dx + 1] = w[co, 3 * (dx + 1) + (dz + 1), ci]
            xp = np.pad(x[:, :, ci], ((1, 1), (0, 0)))          # zero in z
            ref[:, :, co] += correlate(xp, k, mode="wrap")[1:-1]  # periodic in x
    assert np.abs(U.conv3(x, w) - ref).max() < 1e-13


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    H, W = 16, 24
    th = U.init_params(3)
    inp = rng.standard_normal((H, W, U.CIN))
    tgt = rng.standard_normal((H, W, U.COUT))
    L, g = U.backward(th, inp, tgt)
    assert abs(L - U.loss(th, inp, tgt)) <= 1e-12 * abs(L)
    def fd_err(d):
        # ReLU kinks make large steps inexact; the best of two small steps is compared
        errs = []
        for e in (1e-6, 1e-7):
            fd = (U.loss(th + e * d, inp, tgt) - U.loss(th - e * d, inp, tgt)) / (2 * e)
            errs.append(abs(fd - g @ d) / max(abs(fd), 1e-9))
        return min(errs)

    for k in range(4):
        d = rng.standard_normal(th.size)
        d /= np.linalg.norm(d)
        assert fd_err(d) <= 1e-6, k
    # a transposed operand or a dropped skip would break single-layer directions too
    for name, shp in U.layer_shapes():
        d = np.zeros(th.size)
        P = U.unflatten(d)
        P[name][...] = rng.standard_normal(shp)
        assert fd_err(d) <= 1e-6, name


def test_pool_up_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((8, 12, 3))
    y = rng.standard_normal((4, 6, 3))
    assert abs(np.sum(U.pool(x) * y) - np.sum(x * U.pool_backward(y))) < 1e-12
    assert abs(np.sum(U.up(y) * x) - np.sum(y * U.up_backward(x))) < 1e-12


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    th = rng.standard_normal(50)
    g = rng.standard_normal(50)
    m, v = np.zeros(50), np.zeros(50)
    th1 = th.copy()
    U.adam_step(th1, m, v, g, 1, lr=0.01)
    assert np.allclose(th1, th - 0.01 * g / (np.abs(g) + 1e-8), rtol=0, atol=1e-14)


def test_positive_homogeneity_in_residual_channels():
    # bias-free ReLU layers: scaling the residual channels (fixed log-viscosity channel 0 = 0) scales the output
    rng = np.random.default_rng(4)
    th = U.init_params(5)
    inp = rng.standard_normal((8, 8, U.CIN))
    inp[..., 0] = 0.0
    assert np.allclose(U.forward(th, 3.0 * inp), 3.0 * U.forward(th, inp), rtol=1e-12, atol=1e-12)
