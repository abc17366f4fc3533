"""Device initialize! (run.jl:71-140), record_history! (history.jl:101-222), snapshot! and compute_resnorm
(diagnose.jl:42-71) against the oracle, and the error codes of compute_resnorm."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _host(orc, dev, kind):
    if kind == "phase":
        h = orc.phase_state(dev.Nz, dev.Nx)
    elif kind == "fluid":
        h = orc.fluid_state(dev.Nz, dev.Nx)
    else:
        h = orc.HostNoise(dev.Nz, dev.Nx, dev.padL0, dev.padl0, dev.fL, dev.fl)
    for n, a in h.arrays().items():
        a[:] = getattr(dev, n).cpu().numpy()
    return h


def _setup(xc, Nz, Nx, **kw):
    par = xc.Parameters(D=10.0, N=Nz, L=10.0 * Nx / Nz, etam0=10.0, **kw)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    return par, g, sc


def _fresh_host(orc, g, ns_dev):
    return (orc.phase_state(g.Nz, g.Nx), orc.fluid_state(g.Nz, g.Nx),
            orc.HostNoise(g.Nz, g.Nx, ns_dev.padL0, ns_dev.padl0, ns_dev.fL, ns_dev.fl))


def _compare_states(pairs, rtol):
    worst = {}
    for dev, host in pairs:
        for n, want in host.arrays().items():
            got = getattr(dev, n).cpu().numpy()
            if want.dtype == np.uint8:
                assert np.array_equal(got, want), n
                continue
            scale = max(np.abs(want).max(), 1e-300)
            worst[n] = float(np.abs(got - want).max() / scale)
    bad = {k: v for k, v in worst.items() if v > rtol}
    assert not bad, bad
    return worst


@pytest.mark.parametrize("shape,kw", [((48, 40), dict(x0=0.1, Da=0.05, dxg=0.3)),
                                      ((100, 100), dict(x0=0.05, Da=0.05))])
def test_initialize_vs_oracle(xc, orc, shape, kw):
    Nz, Nx = shape
    par, g, sc = _setup(xc, Nz, Nx, **kw)
    print(f"initialize! {shape} {kw}", flush=True)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=7)
    hst = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc)
    assert (t, dt, step) == (0.0, sc.dt0, 0) and ns.draws == 1
    hph, hfl, hns = _fresh_host(orc, g, ns)
    orc.initialize(hph, hfl, hns, g, par, sc, seed=7)
    torch.cuda.synchronize()
    worst = _compare_states(((ph, hph), (fl, hfl), (ns, hns)), 1e-12)
    print("initialize! worst relative differences:", sorted(worst.items(), key=lambda kv: -kv[1])[:5], flush=True)
    # the t = 0 record: conservation errors exactly zero (test_history.jl:44-47)
    assert hst.EB == [0.0] and hst.EM == [0.0] and hst.EX == [0.0]


def test_initialize_x_init_vs_oracle(xc, orc):
    Nz, Nx = 40, 56
    par, g, sc = _setup(xc, Nz, Nx, x0=0.05, Da=0.05)
    rng = np.random.default_rng(11)
    x0 = np.asfortranarray(rng.uniform(0.04, 0.06, (Nz, Nx)))
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=2)
    xc.initialize_(ph, fl, ns, None, g, par, sc, x_init=xc.to_device(x0))
    hph, hfl, hns = _fresh_host(orc, g, ns)
    orc.initialize(hph, hfl, hns, g, par, sc, x_init=x0, seed=2)
    torch.cuda.synchronize()
    assert np.array_equal(ph.x.cpu().numpy(), x0)
    _compare_states(((ph, hph), (fl, hfl), (ns, hns)), 1e-12)


def _compare_history(hd, ho, rtol=1e-12, atol_E=1e-14):
    for n, got in hd.vectors().items():
        want = ho[n]
        assert len(got) == len(want), n
        g, w = np.asarray(got), np.asarray(want)
        if n in ("EB", "EM", "EX", "DB", "DM", "DX"):
            scale = max(np.abs(np.asarray(ho["sumB"])).max() if n[0] == "D" else 1.0, 1e-300)
            err = np.abs(g - w).max() / scale
            assert err <= atol_E, (n, got, want)
        else:
            err = np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-300))
            assert err <= rtol, (n, got, want, err)


def test_history_record_vs_oracle(xc, orc):
    # every History vector after the t = 0 record and two steps of the oracle-fed device state
    Nz, Nx = 48, 40
    par, g, sc = _setup(xc, Nz, Nx, x0=0.1, Da=0.02)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=3)
    hd = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hd, g, par, sc)
    hph, hfl, hns = _host(orc, ph, "phase"), _host(orc, fl, "fluid"), _host(orc, ns, "noise")
    ho = orc.new_history()
    orc.record_history(ho, 0.0, sc.dt0, hph, hfl, hns, g)
    _compare_history(hd, ho)
    state_d = xc.StepState(t, dt, step)
    opts = xc.step_opts(par, fixed_sweeps=3, tol=1e-12, maxit=2000)
    for k in range(2):
        xc.run_step_(ph, fl, ns, g, par, sc, state_d, opts)
        coef = orc.time_coefs("bd2im", state_d.step)
        # record the device state and the same arrays copied to the host: the reductions are compared
        hph, hfl, hns = _host(orc, ph, "phase"), _host(orc, fl, "fluid"), _host(orc, ns, "noise")
        xc.record_history_(hd, state_d.time, dt, ph, fl, ns, g, *coef)
        orc.record_history(ho, state_d.time, dt, hph, hfl, hns, g, *coef)
        dt = state_d.dt
    _compare_history(hd, ho)
    print("EB", hd.EB, "EX", hd.EX, flush=True)


def test_history_sums_match_stats(xc):
    Nz, Nx = 32, 48
    par, g, sc = _setup(xc, Nz, Nx, x0=0.1)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=1)
    xc.initialize_(ph, fl, ns, None, g, par, sc)
    s12 = xc.history_sums(ph, fl)
    s = xc.driver.history_stats(ph, fl, ns)
    assert s12 == s[:12]


def test_resnorm_vs_oracle(xc, orc):
    Nz, Nx = 40, 32
    par, g, sc = _setup(xc, Nz, Nx, x0=0.1)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=4)
    xc.initialize_(ph, fl, ns, None, g, par, sc)
    gen = torch.Generator(device="cuda").manual_seed(5)
    for t in (fl.W, fl.U, fl.P):
        t.copy_(torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.float64))
    xc.driver.snapshot_(ph, fl)
    # unchanged: exactly zero (test_diagnose.jl:52-63)
    assert xc.driver.compute_resnorm(ph, fl, sc.dt0) == (0.0, 0.0, 0.0)
    snap = orc.snapshot(_host(orc, ph, "phase"), _host(orc, fl, "fluid"))
    for t, s in ((fl.W, 1e-3), (fl.U, 2e-3), (fl.P, 5e-4), (fl.MFS, 1e-2), (ph.X, 1e-4)):
        t.add_(s * torch.randn(t.shape, generator=gen, device="cuda", dtype=torch.float64))
    got = xc.driver.compute_resnorm(ph, fl, sc.dt0)
    want = orc.compute_resnorm(snap, _host(orc, ph, "phase"), _host(orc, fl, "fluid"), sc.dt0)
    print("resnorm device", got, "oracle", want, flush=True)
    for a, b in zip(got, want):
        assert abs(a - b) <= 1e-13 * abs(b)


def test_resnorm_errors(xc):
    # a fresh context has no snapshot; a snapshot of another grid size is rejected
    from xcore_amd import _lib
    c = _lib.Context(torch.cuda.current_device(), 0)
    c.bind_stream(torch.cuda.current_stream().cuda_stream)
    Nz, Nx = 24, 16
    par, g, sc = _setup(xc, Nz, Nx, x0=0.1)
    ph, fl = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx)
    ps, fs = ph.pack(), fl.pack()
    out = (ctypes.c_double * 3)()
    L = _lib.lib()
    code = L.xc_compute_resnorm(c.ptr, ctypes.byref(ps), ctypes.byref(fs), 1.0, out)
    assert code == _lib.XC_EINVAL_ARG and b"no snapshot" in L.xc_last_error(c.ptr)
    ph2, fl2 = xc.PhaseState(Nz + 8, Nx), xc.FluidState(Nz + 8, Nx)
    ps2, fs2 = ph2.pack(), fl2.pack()
    assert L.xc_snapshot(c.ptr, ctypes.byref(ps2), ctypes.byref(fs2)) == 0
    code = L.xc_compute_resnorm(c.ptr, ctypes.byref(ps), ctypes.byref(fs), 1.0, out)
    assert code == _lib.XC_EINVAL_ARG and b"grid" in L.xc_last_error(c.ptr)
    # a state whose W does not match its X: a shape error
    fs.W.rows = fs.W.rows - 1
    assert L.xc_snapshot(c.ptr, ctypes.byref(ps), ctypes.byref(fs)) == _lib.XC_EINVAL_SHAPE
    # history sums before anything else on a fresh context: allowed (no snapshot involved)
    s = (ctypes.c_double * 12)()
    assert L.xc_history_sums(c.ptr, ctypes.byref(ph.pack()), ctypes.byref(fl.pack()), s) == 0
    torch.cuda.synchronize()
    c.close()


def test_failed_step_leaves_counter(xc):
    # a Stokes solve that cannot meet its tolerance in 1 iteration fails the step: step, time and dt stay
    Nz, Nx = 32, 32
    par, g, sc = _setup(xc, Nz, Nx, x0=0.1)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=1)
    t, dt, step = xc.initialize_(ph, fl, ns, None, g, par, sc)
    st = xc.StepState(t, dt, step)
    opts = xc.step_opts(par, fixed_sweeps=3, tol=1e-14, maxit=1)
    with pytest.raises(xc.NoConvergence):
        xc.run_step_(ph, fl, ns, g, par, sc, st, opts)
    assert (st.time, st.dt, st.step) == (t, dt, step)
