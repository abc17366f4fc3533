"""The learned coarse correction inside FGMRES, end to end on the GPU.

Training data come from the library's own Stokes solves (random lognormal
viscosity); the network is trained by the library's BF16 forward/backward
kernels; the learned solve is judged like the plain one -- relative residual
<= 1e-8 on the oracle's assembled system, and the same solution as the plain
solve and as sparse LU.  The iteration counts of both are printed (the gain,
or the lack of it, at this small training budget).
"""
import numpy as np
import pytest
import torch

from fieldsetup import plausible_state, to_device_state

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_learned_correction_end_to_end(xc, orc):
    from xcore_amd import learned as L
    from xcore_amd import train_lcc as T
    print("generating 48 training solves at 64^2", flush=True)
    le, rr, ee, its = T.generate(48, N=64, level=2, log_every=0, mode=2)
    assert min(its) > 0
    inp, tgt, is_, os_ = T.normalise(le, rr, ee)
    L.set_params(L.he_init(0))
    l0, _ = L.loss_grad(inp, tgt)
    hist = T.train(inp, tgt, steps=300, batch=16, lr=1e-3, seed=0, log_every=100)  # starts from he_init(0) too
    l1, _ = L.loss_grad(inp, tgt)
    print(f"training: full-set loss {l0:.4e} -> {l1:.4e}", hist, flush=True)
    assert l1 < 0.7 * l0
    L.set_io_scales(is_, os_)
    Nz = Nx = 64
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, Nz, Nx, seed=7)
    orc.update(ph, fl, g, par, sc)
    orc.phsevo(ph, fl, g, par, sc, coef=orc.time_coefs("bd2im", 3), dt=sc.dt0)
    kw = dict(dt=sc.dt0, a1=1.5, a2=2.0, a3=-0.5, b1=1, b2=0, b3=0)
    res = {}
    for learned in (False, True):
        dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
        st = xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, learned=2 if learned else 0, lc_level=2, **kw)
        sol = np.concatenate([xc.to_host(getattr(dfl, k)).ravel(order="F") for k in ("W", "U", "P")])
        res[learned] = (st, sol)
        print(f"learned={learned}: {st}", flush=True)
    LL, RR = orc.fluidmech(fl, ph, g, orc.fm_opts(par, **kw))
    for learned, (st, sol) in res.items():
        assert st["converged"]
        assert np.linalg.norm(LL @ sol - RR) / np.linalg.norm(RR) <= 1e-8, learned
    assert _rel(res[True][1], orc.stokes_vector(fl)) <= 1e-6
    assert _rel(res[True][1], res[False][1]) <= 1e-6


def test_learned_requires_weights_divisible_grid(xc, orc):
    from xcore_amd import learned as L
    L.set_params(L.he_init(0))
    par, g, sc, ph, fl, ns = plausible_state(orc, xc, 36, 36, seed=3)  # 36 / 4 = 9: not a multiple of 8
    dph, dfl = to_device_state(xc, ph, "phase"), to_device_state(xc, fl, "fluid")
    with pytest.raises(AssertionError):
        xc.fluidmech_(dfl, g, par, phase=dph, tol=1e-8, learned=True, lc_level=2, dt=sc.dt0, a1=1.5, a2=2.0,
                      a3=-0.5)
