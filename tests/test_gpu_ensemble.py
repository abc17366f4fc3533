"""Ensembles of ragged grids in one launch, and their HIP-graph replay.

Each member must be bit-identical (exact arithmetic) to its own single-grid call
and to the oracle; a captured graph of the plan launches, replayed several times
with the program's own eager work between replays (the inputs change), must
match the eager result after every replay.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(48, 48), (64, 40), (96, 96), (33, 70), (128, 64)]


def _member(xc, Nz, Nx, seed, flux=True):
    rng = np.random.default_rng(seed)
    f = np.asfortranarray(rng.standard_normal((Nz, Nx)) + 2.0)
    u = np.asfortranarray(rng.standard_normal((Nz, Nx + 1)))
    u[:, Nx] = u[:, 0]
    w = np.asfortranarray(rng.standard_normal((Nz + 1, Nx)))
    k = np.asfortranarray(np.exp(rng.standard_normal((Nz, Nx))))
    d = dict(f=xc.to_device(f), u=xc.to_device(u), w=xc.to_device(w), k=xc.to_device(k), h=1.0 / Nx,
             adv=xc.xcore_zeros(Nz, Nx), dff=xc.xcore_zeros(Nz, Nx))
    if flux:
        d.update(qx=xc.xcore_zeros(Nz + 2, Nx + 1), qz=xc.xcore_zeros(Nz + 1, Nx + 2),
                 dqx=xc.xcore_zeros(Nz + 2, Nx + 1), dqz=xc.xcore_zeros(Nz + 1, Nx + 2))
    return d, (f, u, w, k)


@pytest.mark.parametrize("scheme", ["weno5", "quick"])
def test_batch_matches_single_and_oracle(xc, orc, scheme):
    mem = [_member(xc, Nz, Nx, 10 + i) for i, (Nz, Nx) in enumerate(SIZES)]
    print(f"ensemble {scheme}: {len(mem)} members", flush=True)
    with xc.exact_arithmetic():
        xc.advect_centered_batch_([m for m, _ in mem], scheme)
        xc.diffus_centered_batch_([dict(dff=m["dff"], qx=m["dqx"], qz=m["dqz"], f=m["f"], k=m["k"], h=m["h"])
                                   for m, _ in mem])
        torch.cuda.synchronize()
        for (m, (f, u, w, k)), (Nz, Nx) in zip(mem, SIZES):
            adv_o, qx_o, qz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
            orc.advect_centered(adv_o, f, u, w, m["h"], scheme, qx=qx_o, qz=qz_o)
            for got, want in ((m["adv"], adv_o), (m["qx"], qx_o), (m["qz"], qz_o)):
                assert np.array_equal(xc.to_host(got), want), (Nz, Nx)
            dff_o, dqx_o, dqz_o = orc.zeros(Nz, Nx), orc.zeros(Nz + 2, Nx + 1), orc.zeros(Nz + 1, Nx + 2)
            orc.diffus_centered(dff_o, f, k, m["h"], qx=dqx_o, qz=dqz_o)
            for got, want in ((m["dff"], dff_o), (m["dqx"], dqx_o), (m["dqz"], dqz_o)):
                assert np.array_equal(xc.to_host(got), want), (Nz, Nx)
            # the single-grid call gives the same bits
            a1 = xc.xcore_zeros(Nz, Nx)
            xc.advect_centered_(a1, m["f"], m["u"], m["w"], m["h"], scheme)
            assert torch.equal(a1, m["adv"])


def test_plan_graph_replay(xc, orc):
    mem = [_member(xc, Nz, Nx, 40 + i) for i, (Nz, Nx) in enumerate(SIZES)]
    members = [m for m, _ in mem]
    plan_a = xc.BatchPlan(members, scheme="weno5")
    plan_d = xc.BatchPlan([dict(dff=m["dff"], qx=m["dqx"], qz=m["dqz"], f=m["f"], k=m["k"], h=m["h"])
                           for m in members])
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):  # warm-up outside capture
        plan_a.run()
        plan_d.run()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan_a.run()
        plan_d.run()
    for rep in range(3):
        print(f"graph replay {rep}", flush=True)
        for m in members:  # eager work between replays: new inputs
            m["f"].mul_(1.01).add_(0.1)
        g.replay()
        torch.cuda.synchronize()
        got = [(m["adv"].clone(), m["qx"].clone(), m["dff"].clone()) for m in members]
        for m in members:
            m["adv"].zero_(), m["qx"].zero_(), m["dff"].zero_()
        plan_a.run()
        plan_d.run()
        torch.cuda.synchronize()
        for (a, q, d), m in zip(got, members):
            assert torch.equal(a, m["adv"]) and torch.equal(q, m["qx"]) and torch.equal(d, m["dff"])
    plan_a.close()
    plan_d.close()


def test_batch_shape_errors(xc):
    m, _ = _member(xc, 32, 32, 1)
    bad = dict(m)
    bad["u"] = xc.xcore_zeros(32, 32)
    with pytest.raises(AssertionError):
        xc.advect_centered_batch_([m, bad], "weno5")
    with pytest.raises(ValueError):
        xc.advect_centered_batch_([m], "bogus")
