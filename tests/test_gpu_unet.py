"""GPU U-Net (BF16 MFMA forward/backward, FP32 accumulate) against the FP64 oracle.

The oracle is fed the same BF16-rounded parameters and inputs, so the remaining
differences come from BF16 storage of activations and gradients between
layers.  Bars (normwise relative): output 3e-2, loss 3e-2, every layer's weight
gradient 8e-2.  Training: Adam on a fixed batch must cut the loss in half.
"""
import numpy as np
import pytest
import torch

from oracle import unet as U

pytestmark = pytest.mark.gpu


def _bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def _data(B, H, W, seed):
    rng = np.random.default_rng(seed)
    inp = rng.standard_normal((B, W, H, U.CIN)).astype(np.float32)
    tgt = rng.standard_normal((B, W, H, U.COUT)).astype(np.float32)
    return inp, tgt


def test_forward_and_gradients_vs_oracle(xc):
    from xcore_amd import learned as L
    assert L.num_params() == U.n_params()
    th = L.he_init(7)
    L.set_params(th)
    thb = _bf16(th)
    B, H, W = 2, 16, 24
    inp, tgt = _data(B, H, W, 11)
    print(f"unet forward/grad B={B} H={H} W={W}", flush=True)
    out = L.forward(torch.as_tensor(inp, device="cuda")).cpu().numpy()
    loss, g = L.loss_grad(torch.as_tensor(inp, device="cuda"), torch.as_tensor(tgt, device="cuda"))
    g = g.cpu().numpy().astype(np.float64)
    # head weights stay FP32 on the device
    P = U.unflatten(thb.copy())
    P["head"][...] = th[-U.COUT * U.CH[0]:].astype(np.float64).reshape(U.COUT, U.CH[0])
    thm = np.concatenate([P[n].ravel() for n, _ in U.layer_shapes()])
    ref_out, ref_L, ref_g = [], 0.0, np.zeros_like(thm)
    for b in range(B):
        x = _bf16(inp[b]).transpose(1, 0, 2)
        t = tgt[b].astype(np.float64).transpose(1, 0, 2)
        ref_out.append(U.forward(thm, x).transpose(1, 0, 2))
        Lb, gb = U.backward(thm, x, t)
        ref_L += Lb / B
        ref_g += gb / B
    ref_out = np.stack(ref_out)
    print("out rel", _rel(out, ref_out), "loss", loss, ref_L, flush=True)
    assert _rel(out, ref_out) <= 3e-2
    assert abs(loss - ref_L) <= 3e-2 * ref_L
    o = 0
    for name, shp in U.layer_shapes():
        k = int(np.prod(shp))
        r = _rel(g[o:o + k], ref_g[o:o + k])
        print(f"  grad {name}: rel {r:.3e}", flush=True)
        assert r <= 8e-2, (name, r)
        o += k


def test_training_reduces_loss(xc):
    from xcore_amd import learned as L
    L.set_params(L.he_init(3))
    B, H, W = 4, 32, 32
    rng = np.random.default_rng(5)
    inp = rng.standard_normal((B, W, H, U.CIN)).astype(np.float32)
    # a smooth, learnable target: local averages of the input channels
    tgt = np.stack([0.5 * inp[..., 1] + 0.25 * np.roll(inp[..., 2], 1, axis=1), inp[..., 3] - inp[..., 1],
                    0.3 * inp[..., 0]], axis=-1).astype(np.float32)
    ti, tt = torch.as_tensor(inp, device="cuda"), torch.as_tensor(tgt, device="cuda")
    l0 = L.train_step(ti, tt, lr=2e-3)
    for _ in range(80):
        L.train_step(ti, tt, lr=2e-3, want_loss=False)
    l1, _ = L.loss_grad(ti, tt)
    print(f"training loss {l0:.4e} -> {l1:.4e}", flush=True)
    assert l1 < 0.5 * l0


def test_shape_errors(xc):
    from xcore_amd import learned as L
    with pytest.raises(AssertionError):
        L.forward(torch.zeros((1, 12, 16, 4), dtype=torch.float32, device="cuda"))  # W not a multiple of 8
    with pytest.raises(ValueError):
        L.set_params(np.zeros(10, dtype=np.float32))
