"""FP64 numpy restatement of the learned coarse-correction U-Net (test infrastructure only).

The network has no counterpart in the reference (it is the north star's learned
Stokes preconditioner); this module is the CPU oracle its HIP forward/backward
kernels are checked against.  Only tests/ import it.  It is itself pinned by
central finite differences of the loss (tests/test_oracle_unet.py) and by
scipy.ndimage.correlate for the 3x3 convolution.

Layout (shared with the device code):
  activations  (H, W, C) with H = Nz (z) and W = Nx (x); boundary: x periodic,
               z zero-padded;
  3x3 weights  (Cout, 9, Cin), tap t = 3 * (dx + 1) + (dz + 1);
  1x1 head     (Cout, Cin).
Architecture: 4 levels, channels (16, 32, 48, 64), two bias-free 3x3 conv +
ReLU per level, 2x2 average pooling down, nearest-neighbour upsampling and
channel concatenation [up, skip] on the way up, a 1x1 head to 3 outputs.
Bias-free ReLU layers make the map positively homogeneous in its residual
inputs once the log-viscosity channel is held fixed.
"""
import numpy as np

CH = (16, 32, 48, 64)
CIN = 4     # log-viscosity, r_W, r_U, r_P (cell centred)
COUT = 3    # e_W, e_U, e_P (cell centred)


def layer_shapes():
    """(name, shape) of every weight tensor, in the flat parameter order."""
    c0, c1, c2, c3 = CH
    return [
        ("e0a", (c0, 9, CIN)), ("e0b", (c0, 9, c0)),
        ("e1a", (c1, 9, c0)), ("e1b", (c1, 9, c1)),
        ("e2a", (c2, 9, c1)), ("e2b", (c2, 9, c2)),
        ("e3a", (c3, 9, c2)), ("e3b", (c3, 9, c3)),
        ("d2a", (c2, 9, c3 + c2)), ("d2b", (c2, 9, c2)),
        ("d1a", (c1, 9, c2 + c1)), ("d1b", (c1, 9, c1)),
        ("d0a", (c0, 9, c1 + c0)), ("d0b", (c0, 9, c0)),
        ("head", (COUT, c0)),
    ]


def n_params():
    return sum(int(np.prod(s)) for _, s in layer_shapes())


def unflatten(theta):
    out, o = {}, 0
    for name, shp in layer_shapes():
        k = int(np.prod(shp))
        out[name] = theta[o:o + k].reshape(shp)
        o += k
    return out


def init_params(seed=0):
    """He-normal initialisation (fan-in), flat FP64 vector."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in layer_shapes():
        fan_in = shp[1] * shp[2] if len(shp) == 3 else shp[1]
        parts.append(rng.standard_normal(int(np.prod(shp))) * np.sqrt(2.0 / fan_in))
    return np.concatenate(parts)


# ---- primitives ------------------------------------------------------------
def _shift(x, dz, dx):
    """y[z, x] = x[z + dz, x + dx] with x periodic and zero outside z."""
    H = x.shape[0]
    y = np.roll(x, -dx, axis=1)
    out = np.zeros_like(y)
    if dz == 0:
        out[:] = y
    elif dz > 0:
        out[:H - dz] = y[dz:]
    else:
        out[-dz:] = y[:H + dz]
    return out


def conv3(x, w):
    H, W, _ = x.shape
    out = np.zeros((H, W, w.shape[0]))
    for dx in (-1, 0, 1):
        for dz in (-1, 0, 1):
            t = 3 * (dx + 1) + (dz + 1)
            out += _shift(x, dz, dx) @ w[:, t, :].T
    return out


def conv3_backward(x, w, gy):
    """gradients of conv3 w.r.t. its input and its weights."""
    gx = np.zeros_like(x)
    gw = np.zeros_like(w)
    for dx in (-1, 0, 1):
        for dz in (-1, 0, 1):
            t = 3 * (dx + 1) + (dz + 1)
            xs = _shift(x, dz, dx)
            gw[:, t, :] = np.einsum("hwo,hwi->oi", gy, xs)
            gx += _shift(gy @ w[:, t, :], -dz, -dx)
    return gx, gw


def pool(x):
    return 0.25 * (x[0::2, 0::2] + x[1::2, 0::2] + x[0::2, 1::2] + x[1::2, 1::2])


def pool_backward(gy):
    g = 0.25 * gy
    return np.repeat(np.repeat(g, 2, axis=0), 2, axis=1)


def up(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up_backward(gy):
    return gy[0::2, 0::2] + gy[1::2, 0::2] + gy[0::2, 1::2] + gy[1::2, 1::2]


def relu(x):
    return np.maximum(x, 0.0)


# ---- network -----------------------------------------------------------------
def forward(theta, inp, keep=False):
    """inp (H, W, CIN) -> out (H, W, COUT); with keep=True also the tape for backward."""
    P = unflatten(theta)
    tape = {}

    def block(x, a, b, tag):
        ya = conv3(x, P[a]); ra = relu(ya)
        yb = conv3(ra, P[b]); rb = relu(yb)
        tape[tag] = (x, ya, ra, yb)
        return rb

    e0 = block(inp, "e0a", "e0b", "e0")
    e1 = block(pool(e0), "e1a", "e1b", "e1")
    e2 = block(pool(e1), "e2a", "e2b", "e2")
    e3 = block(pool(e2), "e3a", "e3b", "e3")
    d2 = block(np.concatenate([up(e3), e2], axis=2), "d2a", "d2b", "d2")
    d1 = block(np.concatenate([up(d2), e1], axis=2), "d1a", "d1b", "d1")
    d0 = block(np.concatenate([up(d1), e0], axis=2), "d0a", "d0b", "d0")
    out = d0 @ P["head"].T
    tape["head"] = d0
    tape["e"] = (e0, e1, e2, e3, d2, d1)
    return (out, tape) if keep else out


def loss(theta, inp, tgt):
    out = forward(theta, inp)
    return 0.5 * np.sum((out - tgt) ** 2) / (inp.shape[0] * inp.shape[1])


def backward(theta, inp, tgt):
    """(loss, flat gradient) of 0.5 * sum((out - tgt)^2) / (H W)."""
    P = unflatten(theta)
    out, tape = forward(theta, inp, keep=True)
    npx = inp.shape[0] * inp.shape[1]
    L = 0.5 * np.sum((out - tgt) ** 2) / npx
    G = {k: np.zeros_like(v) for k, v in P.items()}
    gout = (out - tgt) / npx
    d0 = tape["head"]
    G["head"] = np.einsum("hwo,hwi->oi", gout, d0)
    g = gout @ P["head"]

    def block_back(g, a, b, tag):
        x, ya, ra, yb = tape[tag]
        g = g * (yb > 0)
        g, G[b] = conv3_backward(ra, P[b], g)
        g = g * (ya > 0)
        gx, G[a] = conv3_backward(x, P[a], g)
        return gx

    c0, c1, c2, c3 = CH
    g = block_back(g, "d0a", "d0b", "d0")
    g_up, g_e0 = g[..., :c1], g[..., c1:]
    g = block_back(up_backward(g_up), "d1a", "d1b", "d1")
    g_up, g_e1 = g[..., :c2], g[..., c2:]
    g = block_back(up_backward(g_up), "d2a", "d2b", "d2")
    g_up, g_e2 = g[..., :c3], g[..., c3:]
    g_e3 = up_backward(g_up)
    g = block_back(g_e3, "e3a", "e3b", "e3")
    g = block_back(pool_backward(g) + g_e2, "e2a", "e2b", "e2")
    g = block_back(pool_backward(g) + g_e1, "e1a", "e1b", "e1")
    block_back(pool_backward(g) + g_e0, "e0a", "e0b", "e0")
    return L, np.concatenate([G[name].ravel() for name, _ in layer_shapes()])


def adam_step(theta, m, v, grad, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1) in place; t counts from 1."""
    m[:] = b1 * m + (1 - b1) * grad
    v[:] = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    theta -= lr * mh / (np.sqrt(vh) + eps)
