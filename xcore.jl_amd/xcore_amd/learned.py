"""The learned coarse-correction U-Net (no reference counterpart): parameters,
inference and training through the C ABI (xc_unet_*).

Layouts match the device code: inputs FP32 [pixel][4] and outputs FP32
[pixel][3] with pixel p = (b * W + x) * H + z, i.e. torch tensors of shape
(B, W, H, C) in C order.  The flat FP32 parameter vector keeps the layers in
the order e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b head with
3x3 weights [cout][9][cin] (tap t = 3 (dx + 1) + (dz + 1)) and the head
[3][16].
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import lib
from .kernels import context

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_I64 = ctypes.c_int64

_lib.register("xc_unet_num_params", [ctypes.POINTER(_I64)])
_lib.register("xc_unet_set_params", [_P, _P, _I64])
_lib.register("xc_unet_get_params", [_P, _P, _I64])
_lib.register("xc_unet_forward", [_P, _P, _I, _I, _I, _P])
_lib.register("xc_unet_loss_grad", [_P, _P, _P, _I, _I, _I, ctypes.POINTER(_D), _P])
_lib.register("xc_unet_train_step", [_P, _P, _P, _I, _I, _I, ctypes.c_float, ctypes.POINTER(_D)])
_lib.register("xc_unet_set_io_scales", [_P, _P, _P])

CH = (16, 32, 48, 64)
CIN, COUT = 4, 3


def layer_shapes():
    c0, c1, c2, c3 = CH
    return [("e0a", (c0, 9, CIN)), ("e0b", (c0, 9, c0)), ("e1a", (c1, 9, c0)), ("e1b", (c1, 9, c1)),
            ("e2a", (c2, 9, c1)), ("e2b", (c2, 9, c2)), ("e3a", (c3, 9, c2)), ("e3b", (c3, 9, c3)),
            ("d2a", (c2, 9, c3 + c2)), ("d2b", (c2, 9, c2)), ("d1a", (c1, 9, c2 + c1)), ("d1b", (c1, 9, c1)),
            ("d0a", (c0, 9, c1 + c0)), ("d0b", (c0, 9, c0)), ("head", (COUT, c0))]


def num_params():
    n = _I64(0)
    lib().xc_unet_num_params(ctypes.byref(n))
    return n.value


def he_init(seed=0):
    """He-normal (fan-in) initial parameters, flat float32."""
    rng = np.random.default_rng(seed)
    parts = []
    for _, shp in layer_shapes():
        fan_in = shp[1] * shp[2] if len(shp) == 3 else shp[1]
        parts.append(rng.standard_normal(int(np.prod(shp))) * np.sqrt(2.0 / fan_in))
    return np.concatenate(parts).astype(np.float32)


def set_params(theta):
    c = context()
    th = np.ascontiguousarray(theta, dtype=np.float32)
    if th.size != num_params():
        raise ValueError(f"expected {num_params()} parameters, got {th.size}")
    c.check(lib().xc_unet_set_params(c.ptr, th.ctypes.data, th.size), "unet_set_params")


def get_params():
    c = context()
    th = np.zeros(num_params(), dtype=np.float32)
    c.check(lib().xc_unet_get_params(c.ptr, th.ctypes.data, th.size), "unet_get_params")
    return th


def _chk(t, C):
    if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.dim() != 4 or t.shape[3] != C:
        raise AssertionError(f"expected a contiguous CUDA float32 tensor (B, W, H, {C})")


def forward(inp):
    """inp (B, W, H, 4) float32 on the device -> (B, W, H, 3)."""
    _chk(inp, CIN)
    B, W, H, _ = inp.shape
    out = torch.empty((B, W, H, COUT), dtype=torch.float32, device=inp.device)
    c = context()
    c.check(lib().xc_unet_forward(c.ptr, inp.data_ptr(), B, H, W, out.data_ptr()), "unet_forward")
    return out


def loss_grad(inp, tgt):
    """(loss, flat float32 gradient) of 0.5 * mean_pixels sum_channels (out - tgt)^2."""
    _chk(inp, CIN)
    _chk(tgt, COUT)
    B, W, H, _ = inp.shape
    g = torch.empty(num_params(), dtype=torch.float32, device=inp.device)
    loss = _D(0.0)
    c = context()
    c.check(lib().xc_unet_loss_grad(c.ptr, inp.data_ptr(), tgt.data_ptr(), B, H, W, ctypes.byref(loss),
                                    g.data_ptr()), "unet_loss_grad")
    return loss.value, g


def train_step(inp, tgt, lr=1e-3, want_loss=True):
    """One Adam step on a batch; returns the batch loss before the step (synchronises) or None."""
    _chk(inp, CIN)
    _chk(tgt, COUT)
    B, W, H, _ = inp.shape
    loss = _D(0.0)
    c = context()
    c.check(lib().xc_unet_train_step(c.ptr, inp.data_ptr(), tgt.data_ptr(), B, H, W, float(lr),
                                     ctypes.byref(loss) if want_loss else None), "unet_train_step")
    return loss.value if want_loss else None


def set_io_scales(in_scale, out_scale):
    """Fixed per-channel scales of the trained network (inputs multiplied, outputs multiplied)."""
    c = context()
    a = np.ascontiguousarray(in_scale, dtype=np.float32)
    b = np.ascontiguousarray(out_scale, dtype=np.float32)
    if a.size != 4 or b.size != 3:
        raise ValueError("in_scale has 4 entries, out_scale 3")
    c.check(lib().xc_unet_set_io_scales(c.ptr, a.ctypes.data, b.ctypes.data), "unet_set_io_scales")
