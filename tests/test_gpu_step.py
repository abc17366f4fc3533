"""The device time loop (run! / initialize!) against the oracle's loop with sparse-LU Stokes.

The Picard sweep count is fixed (the adaptive count can differ between two
implementations whose Stokes solutions differ at the 1e-8 residual level);
fields are compared after two steps.  Conservation: EB/EM/EX are exactly 0 at
the first record (test_history.jl:44-47) and stay at round-off.
"""
import numpy as np
import pytest
import torch

from fieldsetup import to_device_state

pytestmark = pytest.mark.gpu


def _host_copy(orc, xc, dev, kind, g=None, sc=None):
    if kind == "phase":
        h = orc.phase_state(dev.Nz, dev.Nx)
    elif kind == "fluid":
        h = orc.fluid_state(dev.Nz, dev.Nx)
    else:
        h = orc.HostNoise(dev.Nz, dev.Nx, dev.padL0, dev.padl0, dev.fL, dev.fl)
    for n, a in h.arrays().items():
        a[:] = getattr(dev, n).cpu().numpy()
    return h


def _rel(g, w):
    return np.linalg.norm(g - w) / max(np.linalg.norm(w), 1e-300)


@pytest.mark.parametrize("shape", [(32, 32), (48, 40)])
def test_two_steps_vs_oracle(xc, orc, shape):
    Nz, Nx = shape
    par = xc.Parameters(D=10.0, N=Nz, L=10.0 * Nx / Nz, etam0=10.0, x0=0.1, Da=0.0)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx)
    ns = xc.NoiseState(g, sc, seed=3)
    hst = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc)
    assert hst.EB[0] == 0 and hst.EM[0] == 0 and hst.EX[0] == 0
    hph, hfl, hns = _host_copy(orc, xc, ph, "phase"), _host_copy(orc, xc, fl, "fluid"), _host_copy(orc, xc, ns, "noise")
    state_h = dict(time=t, dt=dt, step=step, draws=ns.draws)
    state_d = xc.StepState(t, dt, step)
    opts = xc.step_opts(par, fixed_sweeps=3, tol=1e-12, maxit=2000)
    for k in range(2):
        stats = xc.run_step_(ph, fl, ns, g, par, sc, state_d, opts)
        print(f"step {k + 1}: {stats}", flush=True)
        orc.run_step(hph, hfl, hns, g, par, sc, state_h, fixed_sweeps=3, seed=3)
        assert state_d.step == state_h["step"] and ns.draws == state_h["draws"]
        assert abs(state_d.time - state_h["time"]) <= 1e-12 * state_h["time"]
        assert abs(state_d.dt - state_h["dt"]) <= 1e-6 * state_h["dt"]
    torch.cuda.synchronize()
    worst = {}
    for n in ("X", "x", "W", "U", "P", "eta", "MFS"):
        src = hph if hasattr(hph, n) else hfl
        worst[n] = _rel(getattr(ph if hasattr(ph, n) else fl, n).cpu().numpy(), getattr(src, n))
    print("relative differences after 2 steps:", worst, flush=True)
    # ~10x the differences measured on an MI355X (P 1.3e-12, MFS 1.1e-12, W/U 8e-14, X 1e-15)
    assert max(worst.values()) <= 2e-11 and worst["X"] <= 1e-14, worst


def test_run_conservation(xc):
    Nz = Nx = 32
    par = xc.Parameters(D=10.0, N=Nz, L=10.0, etam0=10.0, x0=0.1, Da=0.0)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx)
    ns = xc.NoiseState(g, sc, seed=1)
    hst = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc)
    t, dt, log = xc.run_(ph, fl, ns, hst, g, par, sc, nsteps=4, time=t, dt=dt, step=step)
    print("EB", hst.EB, "EX", hst.EX, flush=True)
    assert all(np.isfinite(hst.EB))
    assert max(abs(e) for e in hst.EX) < 1e-6
    for k in ("X", "x"):
        assert bool(torch.isfinite(getattr(ph, k)).all())


def test_config1_box_free_slip_vs_oracle(xc, orc):
    # BASELINE config 1 mapped onto the reference: 10 km square box, 100 x 100 cells, free-slip walls
    # (top_cnv = bot_cnv = -1), crystal fraction uniform random in [0.04, 0.06], WENO5, CFL 0.5.
    # Three time steps on the device (FGMRES) against the oracle (sparse LU): fields and the
    # conservation diagnostics agree.
    N = 100
    par = xc.Parameters(D=1e4, N=N, L=1e4, CFL=0.5, Da=0.05, x0=0.05)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(N, N), xc.FluidState(N, N)
    ns = xc.NoiseState(g, sc, seed=5)
    rng = np.random.default_rng(2024)
    x_init = xc.to_device(np.asfortranarray(rng.uniform(0.04, 0.06, (N, N))))
    hst = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc, x_init=x_init)
    hph, hfl, hns = (_host_copy(orc, xc, ph, "phase"), _host_copy(orc, xc, fl, "fluid"),
                     _host_copy(orc, xc, ns, "noise"))
    state_h = dict(time=t, dt=dt, step=step, draws=ns.draws)
    state_d = xc.StepState(t, dt, step)
    opts = xc.step_opts(par, fixed_sweeps=3, top_cnv=-1, bot_cnv=-1, tol=1e-12, maxit=3000)
    hst_h = xc.History()
    for k in range(3):
        print(f"config-1 step {k + 1}", flush=True)
        stats = xc.run_step_(ph, fl, ns, g, par, sc, state_d, opts)
        orc.run_step(hph, hfl, hns, g, par, sc, state_h, fixed_sweeps=3, top_cnv=-1, bot_cnv=-1, seed=5)
        print(f"  {stats}", flush=True)
        assert abs(state_d.dt - state_h["dt"]) <= 1e-6 * state_h["dt"]
    torch.cuda.synchronize()
    worst = {}
    for n in ("X", "x", "W", "U", "P", "rho"):
        src = hph if hasattr(hph, n) else hfl
        worst[n] = _rel(getattr(ph if hasattr(ph, n) else fl, n).cpu().numpy(), getattr(src, n))
    print("config-1 relative differences after 3 steps:", worst, flush=True)
    # ~10x the differences measured on an MI355X (P 3.6e-9, U 1.2e-11, W 6.7e-12, X 4.6e-13)
    assert worst["P"] <= 5e-8 and max(v for k, v in worst.items() if k != "P") <= 2e-10, worst
    # mass sums of the device state equal the oracle's
    s_dev = xc.history_sums(ph, fl)
    assert abs(s_dev[0] - float(hfl.rho.sum())) <= 1e-12 * abs(float(hfl.rho.sum()))
    assert abs(s_dev[2] - float(hph.X.sum())) <= 1e-10 * abs(float(hph.X.sum()))


def test_checkpoint_resume_bitwise(xc, tmp_path):
    # restart.jl: a run saved after two steps and resumed in fresh states takes the third step exactly as
    # the uninterrupted run does (every fluid, phase and noise array and the history, bit for bit)
    Nz = Nx = 32
    par = xc.Parameters(D=10.0, N=Nz, L=10.0, etam0=10.0, x0=0.1)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)

    def fresh():
        return xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=3), xc.History()

    ph, fl, ns, hst = fresh()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc)
    t, dt, _ = xc.run_(ph, fl, ns, hst, g, par, sc, nsteps=2, time=t, dt=dt, step=step)
    p = xc.restart_path(str(tmp_path), "r")
    xc.save_checkpoint(p, ph, fl, ns, hst, time=t, dt=dt, step=2)
    ta, dta, _ = xc.run_(ph, fl, ns, hst, g, par, sc, nsteps=3, time=t, dt=dt, step=2)
    ph2, fl2, ns2, hst2 = fresh()
    t2, dt2, step2 = xc.load_checkpoint_(p, ph2, fl2, ns2, hst2)
    assert (t2, dt2, step2) == (t, dt, 2)
    tb, dtb, _ = xc.run_(ph2, fl2, ns2, hst2, g, par, sc, nsteps=3, time=t2, dt=dt2, step=step2)
    torch.cuda.synchronize()
    assert (ta, dta) == (tb, dtb)
    for a, b in ((ph, ph2), (fl, fl2), (ns, ns2)):
        for n, x in a.arrays().items():
            assert torch.equal(x, b.arrays()[n]), n
    assert ns.draws == ns2.draws and hst.EB == hst2.EB and hst.time == hst2.time


def test_run_step_equals_its_entry_points(xc):
    # xc_run_step puts noise! (and update!'s lithostatic chain) on a second stream beside phsevo! and
    # fluidmech!; the same step composed from the public entry points, one after another on one stream in
    # the reference's order (run.jl:204-275), must give the same bits in every field
    Nz, Nx = 48, 40
    par = xc.Parameters(D=10.0, N=Nz, L=10.0 * Nx / Nz, etam0=10.0, x0=0.1, Da=0.0)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    states = []
    for _ in range(2):
        ph, fl = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx)
        ns = xc.NoiseState(g, sc, seed=3)
        t, dt, step = xc.initialize_(ph, fl, ns, xc.History(), g, par, sc)
        states.append((ph, fl, ns, xc.StepState(t, dt, step)))
    sweeps = 3
    fm = dict(tol=1e-10, maxit=2000)
    (ph, fl, ns, st), (ph2, fl2, ns2, st2) = states
    xc.run_step_(ph, fl, ns, g, par, sc, st, xc.step_opts(par, fixed_sweeps=sweeps, **fm))
    # the same step by hand
    step = st2.step + 1
    a1, a2, a3, b1, b2, b3 = xc.time_coefs(par.TINT, step)
    dt = st2.dt
    tc = dict(a1=a1, a2=a2, a3=a3, b1=b1, b2=b2, b3=b3)
    xc.store_previous_(fl2, ph2)
    xc.store_noise_(ns2)
    for it in range(1, sweeps + 1):
        xc.driver.snapshot_(ph2, fl2)
        xc.phsevo_(ph2, fl2, g, par, sc, ADVN=par.ADVN, dt=dt, **tc)
        xc.fluidmech_(fl2, g, par, phase=ph2, dt=dt, xmean=1 if it == 1 else 2, **tc, **fm)
        xc.noise_(ns2, ph2, g, par, sc, dt, first_iter=it <= 1)
        xc.update_phase_velocities_(ph2, fl2, ns2, g, par)
        xc.update_(ph2, fl2, g, par, sc)
        xc.driver.compute_resnorm(ph2, fl2, dt)
    dtn = xc.update_dt(ph2, fl2, g, par, dt)
    torch.cuda.synchronize()
    assert st.step == step and st.dt == dtn and st.time == st2.time + dt
    bad = []
    for a, b in ((ph, ph2), (fl, fl2), (ns, ns2)):
        for n, x in a.arrays().items():
            if not torch.equal(x, b.arrays()[n]):
                bad.append(n)
    assert ns.draws == ns2.draws and not bad, bad
