"""Training of the learned coarse correction (no reference counterpart).

For a Stokes system LL (random lognormal viscosity, the reference's operator
with inertia, buoyancy forcing) the inputs are the directions v a plain FGMRES
solve feeds its preconditioner (its Krylov basis) and the targets their exact
corrections e = LL^-1 v (plain FGMRES); inside FGMRES the network's
coarse estimate of e is applied first and the plain block preconditioner
cleans up what it leaves.  Inputs and targets are block averages
on a grid 2^level times coarser (xc_stokes_lc_restrict), normalised per
sample by the RMS of the residual channels and per channel by dataset
statistics (the io scales stored with the weights).  Every solve, network
forward/backward pass and Adam step runs on the GPU through the library.

    python -m xcore_amd.train_lcc --samples 5000 --n 256 --out lcc_weights.npz
"""
import argparse
import math
import os
import time

import numpy as np
import torch

from . import learned as L
from .fields import compute_fft_filter, normal_fill_
from .state import FluidState, Grid, Parameters, compute_scales
from .stokes import krylov_vector, lc_restrict, stokes_precond, stokes_solve, stokes_vector_size

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lcc_weights.npz")


def _smooth_field(N, seed, stream, sigma_cells):
    """Unit-variance random field: white noise (device Philox) through the device Gaussian FFT filter."""
    a = torch.zeros((N, N), dtype=torch.float64, device="cuda").t()
    normal_fill_(a, seed, stream=stream, draw=0)
    f = compute_fft_filter(a, 0, 1.0, float(sigma_cells))
    f = f - f.mean()
    return f / f.std()


def random_problem(N, seed, max_contrast=1e4, par=None):
    """A Stokes system at N x N with a lognormal viscosity of random contrast (up to max_contrast)."""
    par = par or Parameters(N=N)
    g = Grid.from_par(par)
    sc = compute_scales(par, g)
    fl = FluidState(N, N)
    rng = np.random.default_rng(seed)
    contrast = math.exp(rng.uniform(0.0, math.log(max_contrast)))
    f = _smooth_field(N, seed, 11, rng.uniform(2.0, N / 8))
    span = float(f.max() - f.min())
    logeta = math.log(par.etam0) + f * (math.log(contrast) / max(span, 1e-12))
    eta = torch.exp(logeta)
    fl.eta.copy_(eta)
    # corner viscosity: geometric mean of the four cells, x periodic, z clamped (update.jl:411-426)
    ez = torch.cat([eta[:1], eta, eta[-1:]], dim=0)
    ex = torch.cat([ez[:, -1:], ez], dim=1)
    ex = torch.cat([ex, ex[:, 1:2]], dim=1)
    fl.etaco.copy_((ex[:-1, :-1] * ex[1:, :-1] * ex[:-1, 1:] * ex[1:, 1:]) ** 0.25)
    rho = par.rhom0 + 0.05 * (par.rhox0 - par.rhom0) * _smooth_field(N, seed, 12, N / 16)
    fl.rho.copy_(rho)
    rz = torch.cat([rho[:1], rho, rho[-1:]], dim=0)
    fl.rhow.copy_(0.5 * (rz[:-1] + rz[1:]))
    rx = torch.cat([rho[:, -1:], rho, rho[:, :1]], dim=1)
    fl.rhou.copy_(0.5 * (rx[:, :-1] + rx[:, 1:]))
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1.0, b2=0.0, b3=0.0)
    return par, g, fl, kw


def random_residual(g, seed):
    """A unit-norm stacked vector with smooth and rough parts in every block."""
    n = stokes_vector_size(g)
    N = g.Nz
    parts = []
    for k, (r, c) in enumerate(((N + 1, N + 2), (N + 2, N + 1), (N + 2, N + 2))):
        a = torch.zeros((c, r), dtype=torch.float64, device="cuda").t()
        normal_fill_(a, seed, stream=20 + k, draw=0)
        sm = compute_fft_filter(a, 0, 1.0, float(N) / 16)
        sm = (sm - sm.mean()) / sm.std()
        parts.append((0.8 * sm + 0.2 * a).t().reshape(-1))
    v = torch.cat(parts)
    assert v.numel() == n
    return v / v.norm()


def buoyancy_rhs(g, fl, par, seed):
    """A buoyancy-driven right-hand side: W rows carry a smooth density anomaly times g0, the rest 0."""
    N = g.Nz
    n = stokes_vector_size(g)
    rhs = torch.zeros(n, dtype=torch.float64, device="cuda")
    drho = 0.2 * (par.rhox0 - par.rhom0) * _smooth_field(N, seed, 13, N / 12)
    w = torch.zeros((N + 2, N + 1), dtype=torch.float64, device="cuda").t()   # W-type (N+1, N+2)
    w[1:N, 1:N + 1] = 0.5 * (drho[:-1] + drho[1:]) * par.g0
    rhs[:(N + 1) * (N + 2)] = w.t().reshape(-1)
    return rhs


def zero_mean_pressure(x, g):
    """x with the interior mean of its pressure block removed.  The pressure pin (a +1 on one
    continuity row) makes the constant pressure mode a near-null direction of LL, so LL^-1 v carries
    a large constant pressure set by v at the pinned cell alone; that mode is left to FGMRES and
    removed from the targets."""
    Nz, Nx = g.Nz, g.Nx
    nv = (Nz + 1) * (Nx + 2) + (Nz + 2) * (Nx + 1)
    P = x[nv:].view(Nx + 2, Nz + 2)
    y = x.clone()
    y[nv:] -= P[1:-1, 1:-1].mean()
    return y


def generate(samples, N=256, level=2, seed=0, per_system=12, tol=1e-8, log_every=500, maxit=1000, mode=2):
    """Training pairs from the library's own solves: for random lognormal-viscosity systems driven by
    buoyancy, the directions FGMRES feeds its preconditioner (the Krylov basis of a plain solve) and
    their exact corrections e = LL^-1 v (plain FGMRES) -- for mode 2 what the plain block preconditioner
    M misses, e - M v; returned as coarse-grid block averages."""
    L.set_io_scales([1, 1, 1, 1], [1, 1, 1])
    f = 1 << level
    Hc, Wc = N // f, N // f
    le = torch.empty((samples, Wc, Hc), dtype=torch.float32, device="cuda")
    rr = torch.empty((samples, Wc, Hc, 3), dtype=torch.float32, device="cuda")
    ee = torch.empty((samples, Wc, Hc, 3), dtype=torch.float32, device="cuda")
    its = []
    keep = torch.ones(samples, dtype=torch.bool)
    t0 = time.time()
    s = 0
    system = 0
    while s < samples:
        par, g, fl, kw = random_problem(N, seed * 100003 + system)
        rhs = buoyancy_rhs(g, fl, par, seed * 100003 + system)
        system += 1
        _, st = stokes_solve(fl, g, par, rhs, tol=tol, maxit=maxit, restart=40, raise_on_noconv=False, **kw)
        its.append(st["iters"])
        nv = min(per_system, st["iters"], 40)
        basis = [krylov_vector(g, j) for j in range(nv)]
        for v in basis:
            if s >= samples:
                break
            e, st2 = stokes_solve(fl, g, par, v, tol=1e-6, maxit=maxit, raise_on_noconv=False, **kw)
            if not st2["rel_resid"] <= 1e-5:
                keep[s] = False
                s += 1
                continue
            if mode == 2:
                e = e - stokes_precond(fl, g, par, v, **kw)
            e = zero_mean_pressure(e, g)
            a, sg = lc_restrict(fl, g, par, level, v)
            b, _ = lc_restrict(fl, g, par, level, e, sigma=sg)
            le[s] = a[0, ..., 0]
            rr[s] = a[0, ..., 1:] * sg
            ee[s] = b[0, ..., 1:] * sg
            s += 1
            if log_every and s % log_every == 0:
                print(f"  generated {s}/{samples} pairs from {system} systems ({time.time() - t0:.1f} s, mean "
                      f"FGMRES iterations of the buoyancy solves {np.mean(its):.1f})", flush=True)
    k = keep.to(le.device)
    return le[k], rr[k], ee[k], its


def normalise(le, rr, ee):
    """Dataset scales and the normalised network inputs/targets (the device maps, reproduced)."""
    is0 = 1.0 / max(float(le.std()), 1e-12)
    isr = [1.0 / max(float(rr[..., k].pow(2).mean().sqrt()), 1e-30) for k in range(3)]
    c = rr * torch.tensor(isr, device=rr.device, dtype=torch.float32)
    sig = c.pow(2).mean(dim=(1, 2, 3), keepdim=True).sqrt().clamp_min(1e-30)
    inp = torch.cat([(le * is0)[..., None], c / sig], dim=-1).contiguous()
    t = ee / sig
    osr = [max(float(t[..., k].pow(2).mean().sqrt()), 1e-30) for k in range(3)]
    tgt = (t / torch.tensor(osr, device=t.device, dtype=torch.float32)).contiguous()
    return inp, tgt, np.array([is0] + isr, dtype=np.float32), np.array(osr, dtype=np.float32)


def train(inp, tgt, steps=2000, batch=32, lr=1e-3, seed=0, log_every=250):
    L.set_params(L.he_init(seed))
    n = inp.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    t0 = time.time()
    for it in range(steps):
        idx = torch.randint(0, n, (batch,), generator=g).to(inp.device)
        lr_t = lr * (0.1 ** (it / max(steps, 1)))  # exponential decay to lr / 10
        want = log_every and (it % log_every == 0 or it == steps - 1)
        loss = L.train_step(inp[idx].contiguous(), tgt[idx].contiguous(), lr=lr_t, want_loss=bool(want))
        if want:
            hist.append((it, loss))
            print(f"  step {it}: loss {loss:.4e} ({time.time() - t0:.1f} s)", flush=True)
    return hist


def save(path, in_scale, out_scale, level, meta=None, mode=2):
    np.savez(path, theta=L.get_params(), in_scale=in_scale, out_scale=out_scale, level=np.int32(level),
             mode=np.int32(mode), meta=np.array(str(meta or {})))


_MODE = [1]


def load(path=WEIGHTS):
    """Load trained weights into the device context; returns the correction level (the form the network was
    trained for -- the fm_opts `learned` value -- is then ``mode()``)."""
    d = np.load(path, allow_pickle=False)
    L.set_params(d["theta"])
    L.set_io_scales(d["in_scale"], d["out_scale"])
    _MODE[0] = int(d["mode"]) if "mode" in d.files else 1
    return int(d["level"])


def mode():
    """The learned form of the loaded weights: 1 coarse estimate of LL^-1 v, 2 estimate of what M misses."""
    return _MODE[0]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=5000)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--level", type=int, default=2)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--mode", type=int, default=2)
    a = ap.parse_args()
    torch.cuda.set_device(0)
    print(f"generating {a.samples} Stokes solves at {a.n}^2", flush=True)
    t0 = time.time()
    le, rr, ee, its = generate(a.samples, a.n, a.level, mode=a.mode)
    tg = time.time() - t0
    inp, tgt, is_, os_ = normalise(le, rr, ee)
    n_val = max(1, a.samples // 20)
    print(f"training on {a.samples - n_val} samples, {n_val} held out", flush=True)
    t1 = time.time()
    hist = train(inp[n_val:], tgt[n_val:], steps=a.steps, batch=a.batch, lr=a.lr)
    tt = time.time() - t1
    vl = []
    for i in range(0, n_val, 16):
        l, _ = L.loss_grad(inp[i:i + 16].contiguous(), tgt[i:i + 16].contiguous())
        vl.append(l)
    meta = dict(samples=a.samples, n=a.n, level=a.level, mode=a.mode, steps=a.steps, gen_s=round(tg, 1), train_s=round(tt, 1),
                mean_fgmres_iters=float(np.mean(its)), final_train_loss=hist[-1][1], val_loss=float(np.mean(vl)))
    print("result", meta, flush=True)
    save(a.out, is_, os_, a.level, meta, mode=a.mode)


if __name__ == "__main__":
    main()
