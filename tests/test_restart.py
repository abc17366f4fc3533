"""Checkpoints (restart.jl): the file layout and a save/load round trip of every state array on the CPU.

The resumed device run is checked bit for bit in test_gpu_step.py.
"""
import os

import numpy as np
import pytest
import torch

import xcore_amd as xc


def _states(N, seed):
    par = xc.Parameters(D=10.0, N=N, L=10.0)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(N, N, device="cpu"), xc.FluidState(N, N, device="cpu")
    ns = xc.NoiseState(g, sc, seed=seed, device="cpu")
    gen = torch.Generator().manual_seed(seed)
    for st in (ph, fl, ns):
        for n, a in st.arrays().items():
            if a.dtype == torch.uint8:
                a.copy_(torch.randint(0, 2, a.shape, generator=gen, dtype=torch.uint8))
            else:
                a.copy_(torch.randn(a.shape, generator=gen, dtype=a.dtype))
    hst = xc.History()
    hst.time, hst.EB, hst.dsumBdt = [0.0, 0.5, 1.25], [0.0, 1e-17, -3e-16], 0.375
    return ph, fl, ns, hst


def test_restart_path_and_resolve(tmp_path):
    assert xc.restart_path("out", "run1") == os.path.join("out", "run1", "run1_cont.npz")
    assert xc.restart_path("out", "run1", frame=7) == os.path.join("out", "run1", "run1_0007.npz")
    assert xc.resolve_restart(str(tmp_path), "r", 0) is None
    with pytest.raises(FileNotFoundError):
        xc.resolve_restart(str(tmp_path), "r", -1)
    with pytest.raises(FileNotFoundError):
        xc.resolve_restart(str(tmp_path), "r", 3)


def test_checkpoint_round_trip(tmp_path):
    ph, fl, ns, hst = _states(12, 1)
    ns.draws = 5
    p = xc.restart_path(str(tmp_path), "r", frame=2)
    xc.save_checkpoint(p, ph, fl, ns, hst, time=1.5, dt=0.25, step=9)
    assert xc.resolve_restart(str(tmp_path), "r", 2) == p
    with np.load(p, allow_pickle=False) as f:
        keys = set(f.files)
    assert {"fluid/W", "phase/X", "noise/psieo", "noise/seed", "noise/draws", "hst/EB", "meta/time",
            "meta/dt", "meta/step"} <= keys
    ph2, fl2, ns2, hst2 = _states(12, 2)
    t, dt, step = xc.load_checkpoint_(p, ph2, fl2, ns2, hst2)
    assert (t, dt, step) == (1.5, 0.25, 9)
    for a, b in ((ph, ph2), (fl, fl2), (ns, ns2)):
        for n, x in a.arrays().items():
            assert torch.equal(x, b.arrays()[n]), n
    assert (ns2.seed, ns2.draws) == (1, 5)
    assert hst2.time == hst.time and hst2.EB == hst.EB and hst2.dsumBdt == hst.dsumBdt


def test_checkpoint_grid_mismatch(tmp_path):
    ph, fl, ns, hst = _states(12, 1)
    p = str(tmp_path / "c.npz")
    xc.save_checkpoint(p, ph, fl, ns, hst, time=0.0, dt=1.0, step=0)
    ph2, fl2, ns2, hst2 = _states(16, 1)
    with pytest.raises(ValueError):
        xc.load_checkpoint_(p, ph2, fl2, ns2, hst2)
    with pytest.raises(FileNotFoundError):
        xc.load_checkpoint_(str(tmp_path / "missing.npz"), ph2, fl2, ns2, hst2)
