"""Benchmark of the xcore.jl hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--no-cpu] [--no-stencil]

Workload (one "step"): one outer time step of run! at 512 x 512, FP64 --
store_previous!, store_noise!, then the adaptive Picard loop (>= 3 sweeps of
phsevo!, fluidmech! solved by FGMRES + block multigrid to a 1e-8 relative
residual, noise!, update_phase_velocities!, update!, compute_resnorm) and
update_dt.  Parameters are the reference defaults with viscosity contrast cap
etacntr = 1e3 and crystal-melt density contrast rhox0 - rhom0 = 200 kg/m^3;
the state comes from initialize! (seeded perturbation), x periodic, z closed.
All state is resident in HBM before timing; value = steps / wall time with
the plain block preconditioner (Cahouet-Chabard Schur + multigrid).  When the
trained learned-correction weights are present (xcore.jl_amd/lcc_weights.npz)
the same steps are timed with the learned preconditioner too and reported in
config.learned with learned_over_plain.

The stencil microbench (2048^2, four fields, WENO5 + QUICK advection with
fluxes and lognormal-k diffusion with fluxes) runs after the step bench and
gives each stencil's algorithmic GB/s from HIP events on the launch stream;
its WENO5 kernel is the `roofline` object (the HBM-bound kernel family of the
path; the 512^2 step itself is dominated by many small, latency-bound Stokes
kernels whose working set is largely cache-resident).

The ensemble section (64 members of 128^2 / 256^2 / 384^2) times one batched launch per
operator -- replayed from a HIP graph -- against one launch per member and operator.

`cpu_baseline` times the CPU oracle (the C restatement of the reference with
SciPy's sparse LU for the Stokes solve) on one Picard sweep of the same 512^2
step (~2.5 minutes, most of it the sparse LU) and converts it to steps/s with
the mean sweep count of the GPU run.  Config 1 (100^2, 20 steps) runs on both
the device and the oracle, the 2048^2 stencil calls on the oracle too, and
noise! at 1024^2 is timed as the OU microbench.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "xcore.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "steps/s at 512² learned vs plain Stokes @1e-8 resid; stencil % HBM peak"
HBM_PEAK_GBS = 8000.0
N_STENCIL = 2048
NFIELDS = 4
# advection tile shapes (rows x columns); (64, -16) and (64, -8) are the persistent software-pipelined kernels
# the library writes a launch's outputs with non-temporal stores above this size
NT_STORE_BYTES = 64 << 20
TILE_SHAPES = {"64x16": (64, 16), "128x8": (128, 8), "64x8": (64, 8), "32x32": (32, 32), "64x16p": (64, -16),
               "64x8p": (64, -8)}
# HBM bytes per launch of the WENO5 stencil at 2048^2 (64 x 16 tiles, non-temporal loads and stores), measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE in separate runs of `bench.py --no-step --no-cpu --no-ensemble
# --calib-copy-mb 1024`: 51,778 KiB x 2.0 read + 103,914 KiB x 1.0 written, the factors being what a 1 GiB
# device copy reads on the same counters (524,302 and 1,048,576 KiB).  1.055 x the algorithmic bytes.
TRAFFIC_WENO5_2048 = 212446368


def make_inputs(N, seed_base=1001):
    """Synthetic C2 inputs on the host (numpy, seeded)."""
    import numpy as np
    h = 1.0 / N
    xc = (np.arange(N) + 0.5) * h
    zc = (np.arange(N) + 0.5) * h
    fields = []
    for s in range(NFIELDS):
        rng = np.random.default_rng(seed_base + s)
        f = np.zeros((N, N))
        for _ in range(32):
            kx, kz = rng.integers(1, 33, size=2)
            amp = 1.0 / (kx * kx + kz * kz)
            phi = rng.uniform(0, 2 * np.pi)
            f += amp * np.outer(np.cos(np.pi * kz * zc), np.ones(N)) * np.cos(2 * np.pi * kx * xc + phi)[None, :]
        fields.append(np.asfortranarray(f + 1.0))
    # corner streamfunction psi (N+1, N+1) at (z faces, x faces)
    rng = np.random.default_rng(2001)
    zf = np.arange(N + 1) * h
    xf = np.arange(N + 1) * h
    psi = np.outer(np.sin(np.pi * zf), np.sin(np.pi * xf))
    for _ in range(16):
        m, l = rng.integers(1, 17, size=2)
        a = rng.normal(0.0, 0.3)
        psi += a * np.outer(np.sin(l * np.pi * zf), np.sin(m * np.pi * xf))
    u = np.asfortranarray((psi[1:, :] - psi[:-1, :]) / h)      # (N, N+1)  U = d psi/dz
    w = np.asfortranarray(-(psi[:, 1:] - psi[:, :-1]) / h)     # (N+1, N)  W = -d psi/dx
    rng = np.random.default_rng(3001)
    k = np.asfortranarray(np.exp(rng.standard_normal((N, N))) * 1e-4)
    return fields, u, w, k, h


def c3_setup(N, seed=15):
    """The 512^2 full-step workload: reference defaults, etacntr 1e3, rhox0 - rhom0 = 200."""
    import torch
    import xcore_amd as xc
    par = xc.Parameters(N=N, etacntr=1e3, rhox0=2700.0 + 200.0, seed=seed)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl = xc.PhaseState(g.Nz, g.Nx), xc.FluidState(g.Nz, g.Nx)
    ns = xc.NoiseState(g, sc, seed=seed)
    hst = xc.History()
    t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc)
    torch.cuda.synchronize()
    return par, g, sc, ph, fl, ns, hst, xc.StepState(t, dt, step)


def step_bench(args, learned=False):
    import torch
    import xcore_amd as xc
    xc.load()
    torch.cuda.set_device(0)
    par, g, sc, ph, fl, ns, hst, state = c3_setup(args.n)
    lc_level = 2
    if learned:
        from xcore_amd import train_lcc
        lc_level = train_lcc.load(args.weights)
        learned = train_lcc.mode()
    with xc.autotune():  # tile shapes for the step's advection calls, chosen once before timing
        xc.run_step_(ph, fl, ns, g, par, sc, xc.StepState(state.time, state.dt, state.step),
                     xc.step_opts(par, tol=1e-8, maxit=2000, fixed_sweeps=1))
    par, g, sc, ph, fl, ns, hst, state = c3_setup(args.n)
    opts = xc.step_opts(par, tol=1e-8, maxit=2000, mg_pre=args.mg_pre, mg_post=args.mg_post, mg_coarse=args.mg_coarse,
                        mg_sub=args.mg_sub, mg_levels=args.mg_levels, mg_omega=args.mg_omega, learned=learned,
                        lc_level=lc_level, mg_jacobi=args.mg_jacobi)
    def one_step():
        # run.jl:205-275: the step and its record_history! (nrh = 1, parameters.jl:18)
        dt_step = state.dt
        st = xc.run_step_(ph, fl, ns, g, par, sc, state, opts)
        xc.record_history_(hst, state.time, dt_step, ph, fl, ns, g, *xc.time_coefs(par.TINT, state.step))
        return st

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    log = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        log.append(one_step())
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    sweeps = sum(r["sweeps"] for r in log)
    its = sum(r["stokes_iters"] for r in log)
    info = {"sweeps_per_step": sweeps / len(log), "fgmres_iters_per_solve": its / max(sweeps, 1),
            "worst_stokes_rel_resid": max(r["stokes_worst_rel_resid"] for r in log),
            "stokes_fallbacks": sum(r["stokes_fallbacks"] for r in log),
            "stokes_learned_steps": sum(r["stokes_learned_steps"] for r in log),
            "vcycle_levels": [min(r["stokes_mg_levels"] for r in log), max(r["stokes_mg_levels"] for r in log)],
            "inertia_ratio": [min(r["stokes_inertia_ratio"] for r in log), max(r["stokes_inertia_ratio"] for r in log)],
            "fgmres_iters_first_sweep": sum(r["stokes_iters_first"] for r in log) / len(log),
            "init_rel_resid_first_sweep": max(r["stokes_init_rel_first"] for r in log),
            "init_rel_resid_later_sweeps": max(r["stokes_init_rel_later"] for r in log),
            "sim_time": state.time, "dt": state.dt, "step": int(state.step),
            "history_EB_max": max(abs(e) for e in hst.EB), "history_records": len(hst.EB),
            "x_finite": bool(torch.isfinite(ph.x).all())}
    return 1e3 * elapsed / args.steps, info


def stencil_bench(N, steps, warmup):
    """C2 stencil microbench: per-kernel HIP-event timings on the launch stream.  Each of the
    NFIELDS rotating sets holds its own field, velocities, diffusivity and outputs (about 300 MB a
    set at 2048^2), so no launch finds its inputs in the 256 MB last-level cache left by the one
    before: the timings are HBM-bound.  The advection tile shape is chosen the same way: each
    candidate is timed over the rotating sets and the fastest is pinned."""
    import torch
    import xcore_amd as xc
    fields_h, u_h, w_h, k_h, h = make_inputs(N)
    sets = []
    for f in fields_h:
        sets.append(dict(f=xc.to_device(f), u=xc.to_device(u_h), w=xc.to_device(w_h), k=xc.to_device(k_h),
                         adv=xc.xcore_zeros(N, N), qx=xc.xcore_zeros(N + 2, N + 1), qz=xc.xcore_zeros(N + 1, N + 2),
                         dff=xc.xcore_zeros(N, N), dqx=xc.xcore_zeros(N + 2, N + 1),
                         dqz=xc.xcore_zeros(N + 1, N + 2)))
    torch.cuda.synchronize()

    def launch(kind, s):
        if kind == "diffus":
            xc.diffus_centered_with_flux_(s["dff"], s["dqx"], s["dqz"], s["f"], s["k"], h)
        else:
            xc.advect_centered_with_flux_(s["adv"], s["qx"], s["qz"], s["f"], s["u"], s["w"], h, kind)

    stream = torch.cuda.current_stream()

    def timed_calls(call, reps, pre):
        """us per call of `reps` calls (call(r), r = 0 .. reps-1) timed with HIP events, after `pre` untimed calls
        enqueued right before them: the device is busy when the timed region starts.  (After the stream has
        been idle the clocks ramp up over the first few milliseconds of load: the WENO5 launch averaged 49 us
        over 8 launches started from an idle device, 45 over 80 and 42 over 400.)"""
        for r in range(pre):
            call(r)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for r in range(reps):
            call(r)
        e1.record(stream)
        e1.synchronize()
        return 1e3 * e0.elapsed_time(e1) / reps

    def timed(kind, reps, pre):
        return timed_calls(lambda r: launch(kind, sets[r % NFIELDS]), reps, pre)

    tiles = {}
    for kind in ("weno5", "quick"):
        cand = {}
        for name, (tz, tx) in TILE_SHAPES.items():
            xc.set_advect_tile(kind, N, N, tz, tx)
            cand[name] = round(timed(kind, 2 * NFIELDS, 4 * NFIELDS), 2)
        best = min(cand, key=cand.get)
        xc.set_advect_tile(kind, N, N, *TILE_SHAPES[best])
        tiles[kind] = (best, cand)
    for _ in range(warmup):
        for s in sets:
            for kind in ("weno5", "quick", "diffus"):
                launch(kind, s)
    torch.cuda.synchronize()
    reps = max(8, 4 * steps)
    pre = max(8, 2 * steps)  # untimed launches of the same kernel right before each timed loop
    adv_bytes = 8 * (N * N + N * (N + 1) + (N + 1) * N + N * N + (N + 2) * (N + 1) + (N + 1) * (N + 2))
    dif_bytes = 8 * (N * N + N * N + N * N + (N + 2) * (N + 1) + (N + 1) * (N + 2))
    out = {}
    for kind, nbytes, name in (("weno5", adv_bytes, "weno5_advect_with_flux"),
                               ("quick", adv_bytes, "quick_advect_with_flux"),
                               ("diffus", dif_bytes, "diffus_with_flux")):
        us = timed(kind, reps, pre)
        gbs = nbytes / (us * 1e-6) / 1e9
        out[name] = {"avg_launch_us": round(us, 2), "bytes_per_launch": nbytes, "GB/s": round(gbs, 1),
                     "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    for k in ("weno5", "quick"):
        best, cand = tiles[k]
        out[f"{k}_advect_with_flux"]["tile"] = best
        out[f"{k}_advect_with_flux"]["tile_candidates_us"] = cand
    # the same traffic without the arithmetic (xc.diag_stream_, same arrays, same store kind): the
    # bandwidth HBM sustains for this mix of streams, which bounds what the stencil can reach
    def stream_us(ins, outs, mode):
        def call(r):
            s = sets[r % NFIELDS]
            xc.diag_stream_([s[a] for a in ins], [s[a] for a in outs], mode)
        return timed_calls(call, reps, pre)

    # each floor is the faster of plain and non-temporal loads (mode bit 2)
    floors = {}
    for kind, ins, outs, nbytes in (("advect", ("f", "u", "w"), ("adv", "qx", "qz"), adv_bytes),
                                    ("diffus", ("f", "k"), ("dff", "dqx", "dqz"), dif_bytes)):
        nts = 1 if 8 * (N * N + 2 * (N + 2) * (N + 1)) > NT_STORE_BYTES else 0  # the outputs of one launch
        floors[kind] = min(stream_us(ins, outs, nts), stream_us(ins, outs, nts | 4))
    # the best case of the same traffic: N x N streams (3 + 3, 2 + 3, and a copy), every access an aligned
    # 16-byte one and no output shifted a row -- what HBM sustains when every line is written whole
    al = [[xc.xcore_zeros(N, N) for _ in range(6)] for _ in range(NFIELDS)]
    best = {}
    for kind, nr, nw in (("advect", 3, 3), ("diffus", 2, 3), ("copy", 1, 1)):
        tb = []
        for mode in (1 | 2, 1 | 2 | 4):
            def call(r):
                a6 = al[r % NFIELDS]
                xc.diag_stream_(a6[:nr], a6[3:3 + nw], mode)
            tb.append(timed_calls(call, reps, pre))
        best[kind] = (min(tb), 8 * N * N * (nr + nw))
    del al
    for name, kind in (("weno5_advect_with_flux", "advect"), ("quick_advect_with_flux", "advect"),
                       ("diffus_with_flux", "diffus")):
        o = out[name]
        o["stream_floor_us"] = round(floors[kind], 2)
        o["frac_of_stream_floor"] = round(floors[kind] / o["avg_launch_us"], 4)
        us_b, by_b = best[kind]
        o["aligned_stream_GBs"] = round(by_b / (us_b * 1e-6) / 1e9, 1)
        o["frac_of_aligned_stream"] = round(o["GB/s"] / o["aligned_stream_GBs"], 4)
    out["aligned_copy_GBs"] = round(best["copy"][1] / (best["copy"][0] * 1e-6) / 1e9, 1)
    wk = out["weno5_advect_with_flux"]
    roofline = {"bound": "hbm", "achieved": wk["GB/s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": wk["hbm_frac"], "traffic": TRAFFIC_WENO5_2048,
                "kernel": "k_advect<WENO5> (advect_centered_with_flux!, %d^2, %d rotating sets of field, velocities "
                          "and outputs)" % (N, NFIELDS),
                "bytes_per_launch": adv_bytes, "avg_launch_us": wk["avg_launch_us"], "launches_timed": reps,
                "same_traffic_stream_us": wk["stream_floor_us"],
                "same_traffic_stream_GBs": round(adv_bytes / (wk["stream_floor_us"] * 1e-6) / 1e9, 1),
                "frac_of_same_traffic_stream": wk["frac_of_stream_floor"],
                "aligned_stream_GBs": wk["aligned_stream_GBs"], "frac_of_aligned_stream": wk["frac_of_aligned_stream"],
                "aligned_copy_GBs": out["aligned_copy_GBs"]}
    return roofline, out


def ensemble_bench(reps=20, seed=4001):
    """Config-4 style ensemble: 64 members of 128^2, 256^2 and 384^2 (cycled), each with its own
    fields, advanced by WENO5 and QUICK advection and diffusion (all with fluxes) -- one launch per
    operator for the whole ensemble (a device-resident batch plan, replayed from a HIP graph) against
    one launch per member and operator."""
    import numpy as np
    import torch
    import xcore_amd as xc
    sizes = [128, 256, 384]
    rng = np.random.default_rng(seed)
    mem = []
    for i in range(64):
        N = sizes[i % 3]
        f = xc.to_device(np.asfortranarray(rng.standard_normal((N, N)) + 2.0))
        u = xc.to_device(np.asfortranarray(rng.standard_normal((N, N + 1))))
        w = xc.to_device(np.asfortranarray(rng.standard_normal((N + 1, N))))
        k = xc.to_device(np.asfortranarray(np.exp(rng.standard_normal((N, N)))))
        mem.append(dict(f=f, u=u, w=w, k=k, h=1.0 / N, adv=xc.xcore_zeros(N, N), qx=xc.xcore_zeros(N + 2, N + 1),
                        qz=xc.xcore_zeros(N + 1, N + 2), dff=xc.xcore_zeros(N, N), dqx=xc.xcore_zeros(N + 2, N + 1),
                        dqz=xc.xcore_zeros(N + 1, N + 2)))
    cells = sum(m["f"].shape[0] * m["f"].shape[1] for m in mem)
    dmem = [dict(dff=m["dff"], qx=m["dqx"], qz=m["dqz"], f=m["f"], k=m["k"], h=m["h"]) for m in mem]
    plans = [xc.BatchPlan(mem, scheme="weno5"), xc.BatchPlan(mem, scheme="quick"), xc.BatchPlan(dmem)]

    def batched():
        for p in plans:
            p.run()

    def single():
        for m in mem:
            xc.advect_centered_with_flux_(m["adv"], m["qx"], m["qz"], m["f"], m["u"], m["w"], m["h"], "weno5")
            xc.advect_centered_with_flux_(m["adv"], m["qx"], m["qz"], m["f"], m["u"], m["w"], m["h"], "quick")
            xc.diffus_centered_with_flux_(m["dff"], m["dqx"], m["dqz"], m["f"], m["k"], m["h"])

    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        batched()
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        batched()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    t_graph = timed(g.replay)
    t_batch = timed(batched)
    t_single = timed(single)
    for p in plans:
        p.close()
    return {"members": 64, "sizes": "128^2, 256^2, 384^2 (cycled)", "cells": cells,
            "operators": "WENO5 + QUICK advection with fluxes, diffusion with fluxes",
            "ms_per_ensemble_step_graph": round(1e3 * t_graph, 3), "ms_per_ensemble_step_batched": round(1e3 * t_batch, 3),
            "ms_per_ensemble_step_per_member_launches": round(1e3 * t_single, 3),
            "cell_steps_per_s_graph": round(cells / t_graph, 1), "speedup_vs_per_member": round(t_single / t_graph, 3),
            "launches_per_step": {"batched": 3, "per_member": 192}}


def ensemble_steps_bench(n=64, steps=2, threads=4, seed=7):
    """Config-4 ensemble of whole time steps: n members cycling through 128^2, 256^2 and 384^2, each with its own
    noise seed, density contrast (+-20 %) and crystal viscosity 10^e (e uniform in 20..30), each member one full
    step of run! per ensemble step (adaptive Picard sweeps, FGMRES to 1e-8).  The members' steps run
    concurrently (xc_ensemble_run_step, `threads` host threads over per-member streams) and, for comparison,
    one member after another (one thread)."""
    import torch
    import xcore_amd as xc
    from xcore_amd.ensemble import ensemble_members, ensemble_run_step_
    out = {"members": n, "sizes": "128^2, 256^2, 384^2 (cycled)", "threads": threads}
    res = {}
    for mode, nt in (("concurrent", threads), ("one_after_another", 1)):
        members = ensemble_members(n, seed=seed)
        cells = sum(m.cells for m in members)
        opts = xc.step_opts(members[0].par, tol=1e-8, maxit=2000)
        ensemble_run_step_(members, opts, nthreads=nt)   # warm-up step (solver hierarchies, FFT plans)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        sweeps = 0
        for _ in range(steps):
            st = ensemble_run_step_(members, opts, nthreads=nt)
            sweeps += sum(s["sweeps"] for s in st)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        res[mode] = dt
        out[mode] = {"ms_per_ensemble_step": round(1e3 * dt, 2), "member_steps_per_s": round(n / dt, 1),
                     "cell_steps_per_s": round(cells / dt, 1), "picard_sweeps_per_member_step": sweeps / (n * steps)}
        out["cells"] = cells
        del members
        torch.cuda.empty_cache()
    out["concurrent_speedup"] = round(res["one_after_another"] / res["concurrent"], 3)
    return out


def _cores():
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    return max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))


def c1_bench(steps=20):
    """Config 1, the reference's CPU-runnable case: 10 km box, 100 x 100, free-slip walls, x ~ U(0.04, 0.06),
    WENO5, CFL 0.5, 20 steps with the adaptive Picard loop -- on the device (FGMRES to 1e-8; timed on a second
    run from the same initial state, after one untimed) and on the C oracle (sparse LU), from the same initial
    state; steps/s of each and the conservation errors."""
    import numpy as np
    import torch
    import xcore_amd as xc
    from oracle import oracle as orc
    N = 100
    par = xc.Parameters(D=1e4, N=N, L=1e4, CFL=0.5, Da=0.05, x0=0.05)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    x_init = np.asfortranarray(np.random.default_rng(2024).uniform(0.04, 0.06, (N, N)))
    opts = xc.step_opts(par, top_cnv=-1, bot_cnv=-1, tol=1e-8, maxit=2000)

    def start():
        ph, fl = xc.PhaseState(N, N), xc.FluidState(N, N)
        ns = xc.NoiseState(g, sc, seed=5)
        hst = xc.History()
        t, dt, step = xc.initialize_(ph, fl, ns, hst, g, par, sc, x_init=xc.to_device(x_init))
        return ph, fl, ns, hst, t, dt, step

    # warm-up: the same 20 steps once untimed (the first run of a process also pays the solver's set-up)
    ph, fl, ns, hst, t, dt, step = start()
    state = xc.StepState(t, dt, step)
    for _ in range(steps):
        xc.run_step_(ph, fl, ns, g, par, sc, state, opts)
    ph, fl, ns, hst, t, dt, step = start()
    hosts = []
    for kind, dev in (("phase", ph), ("fluid", fl), ("noise", ns)):
        h = (orc.phase_state(N, N) if kind == "phase" else orc.fluid_state(N, N) if kind == "fluid" else
             orc.HostNoise(N, N, ns.padL0, ns.padl0, ns.fL, ns.fl))
        for n, a in h.arrays().items():
            a[:] = getattr(dev, n).cpu().numpy()
        hosts.append(h)
    st_h = dict(time=t, dt=dt, step=step, draws=ns.draws)
    state = xc.StepState(t, dt, step)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sweeps_d = 0
    for _ in range(steps):
        dt_step = state.dt
        r = xc.run_step_(ph, fl, ns, g, par, sc, state, opts)
        xc.record_history_(hst, state.time, dt_step, ph, fl, ns, g, *xc.time_coefs(par.TINT, state.step))
        sweeps_d += r["sweeps"]
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    cores = _cores()
    orc.set_threads(cores)
    hph, hfl, hns = hosts
    t0 = time.perf_counter()
    sweeps_h = 0
    for _ in range(steps):
        sweeps_h += orc.run_step(hph, hfl, hns, g, par, sc, st_h, top_cnv=-1, bot_cnv=-1, seed=5)
    t_cpu = time.perf_counter() - t0
    x_d = ph.x.cpu().numpy()
    return {"workload": "config 1: 100^2, 10 km box, free slip, 20 steps, adaptive Picard",
            "gpu_steps_per_s": round(steps / t_gpu, 2), "gpu_picard_sweeps": sweeps_d,
            "cpu_steps_per_s": round(steps / t_cpu, 3), "cpu_picard_sweeps": sweeps_h, "cpu_cores": cores,
            "cpu_kind": "port (C oracle, OpenMP; SciPy SuperLU for the Stokes solve)",
            "gpu_over_cpu": round(t_cpu / t_gpu, 2), "EB_max_gpu": max(abs(e) for e in hst.EB),
            "x_rel_diff_after_20_steps": float(np.linalg.norm(x_d - hph.x) / np.linalg.norm(hph.x))}


def stencil_cpu_bench(N, reps=10):
    """The 2048^2 stencil calls on the C oracle (OpenMP over the host cores): median seconds per call."""
    import numpy as np
    from oracle import oracle as orc
    fields, u, w, k, h = make_inputs(N)
    f = fields[0]
    cores = _cores()
    orc.set_threads(cores)
    out = {"cores": cores, "kind": "port (C oracle)", "grid": [N, N], "calls": reps}
    bufs = (orc.zeros(N, N), orc.zeros(N + 2, N + 1), orc.zeros(N + 1, N + 2))
    for name, fn in (("weno5_with_flux", lambda: orc.advect_centered(bufs[0], f, u, w, h, "weno5", qx=bufs[1],
                                                                     qz=bufs[2])),
                     ("quick_with_flux", lambda: orc.advect_centered(bufs[0], f, u, w, h, "quick", qx=bufs[1],
                                                                     qz=bufs[2])),
                     ("diffus_with_flux", lambda: orc.diffus_centered(bufs[0], f, k, h, qx=bufs[1], qz=bufs[2]))):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        out[name + "_ms"] = round(1e3 * float(np.median(ts)), 2)
    return out


def ou_setup(N=1024, seed=4001):
    """noise! at N^2 with L0h = l0h = 4h (pad 16: FFT length N + 16) and a correlation time of 10 dt after
    noise.jl:171-172's update (pre-update tau = L0h / (2 V) = 9 / 0.91 dt)."""
    import torch
    import xcore_amd as xc
    D = 1.0
    h = D / N
    par = xc.Parameters(D=D, N=N, L=D, L0=7 * h, l0=7 * h, x0=0.05, Da=0.0)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    assert abs(sc.L0h - 4 * h) < 1e-15 * D and 4 * math.ceil(sc.L0h / h) == 16
    dt = 1e-3
    tau0 = dt / (1.0 / 9.0 - 0.01)
    ph = xc.PhaseState(N, N)
    ns = xc.NoiseState(g, sc, seed=seed)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    ph.V.fill_(sc.L0h / (2 * tau0))
    ph.vx.fill_(sc.l0h / (2 * tau0))
    ph.fReL.fill_(1.0)
    ph.ke.fill_(1e-3)
    ph.ks.fill_(1e-4)
    ph.chi.copy_(0.05 + 0.01 * torch.rand((N, N), generator=gen, device="cuda", dtype=torch.float64))
    for n in ("psieo", "psixo", "psiso"):
        getattr(ns, n).copy_(1e-3 * torch.randn((N, N), generator=gen, device="cuda", dtype=torch.float64))
    ph.bndtaperw.fill_(1.0)
    return par, g, sc, ph, ns, dt


def ou_bench(N=1024, reps=20, batches=5, warm=5):
    """noise! (OU update of three potentials, their Gaussian filters with FFT length 1040 x 1024, the noise fluxes
    and magnitudes) at 1024^2 with fresh white noise every call: time per call from HIP events on the launch
    stream, the median of `batches` batches of `reps` calls after `warm` untimed calls (single batches of a fresh
    process have read anywhere from 178 to 230 us per call); algorithmic bytes = every array it reads or writes
    once."""
    import torch
    import xcore_amd as xc
    par, g, sc, ph, ns, dt = ou_setup(N)
    for _ in range(warm):
        xc.noise_(ns, ph, g, par, sc, dt, first_iter=True)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    per = []
    for _ in range(batches):
        torch.cuda.synchronize()
        e0.record(s)
        for _ in range(reps):
            xc.noise_(ns, ph, g, par, sc, dt, first_iter=True)
        e1.record(s)
        torch.cuda.synchronize()
        per.append(1e3 * e0.elapsed_time(e1) / reps)
    us = sorted(per)[len(per) // 2]
    cell, wt, ut = N * N, (N + 1) * (N + 2), (N + 2) * (N + 1)
    # read: psi*o (3), V, vx, fReL, ke, chi, ks (cells), chiw, muw (W faces), chiu, muu (U faces);
    # written: re, rs, psi* (3), xie/xix/xis (cells), 3 W-face and 3 U-face fluxes
    nbytes = 8 * (cell * (3 + 6) + 2 * wt + 2 * ut + cell * (2 + 3 + 3) + 3 * wt + 3 * ut)
    return {"workload": "noise! at 1024^2, Gaussian width 4h (FFT 1040 x 1024), correlation time 10 dt",
            "us_per_call": round(us, 2), "cells_per_s": round(cell / (us * 1e-6), 1),
            "algorithmic_bytes_per_call": nbytes, "gbs": round(nbytes / (us * 1e-6) / 1e9, 1),
            "frac_of_hbm_peak": round(nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}


def cpu_baseline(N, sweeps_per_step):
    """One Picard sweep of the same step on the CPU oracle (C, OpenMP; SciPy sparse LU for Stokes)."""
    import numpy as np
    from oracle import oracle as orc
    import xcore_amd as xc
    import torch
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    orc.set_threads(cores)
    # same initial state as the device run at this size (initialize! on the device, copied to the host)
    par, g, sc, ph, fl, ns, hst, state = c3_setup(N)
    hph, hfl = orc.phase_state(N, N), orc.fluid_state(N, N)
    hns = orc.HostNoise(N, N, ns.padL0, ns.padl0, ns.fL, ns.fl)
    for host, dev in ((hph, ph), (hfl, fl), (hns, ns)):
        for n, a in host.arrays().items():
            a[:] = getattr(dev, n).cpu().numpy()
    st = dict(time=state.time, dt=state.dt, step=int(state.step), draws=ns.draws)
    del ph, fl, ns
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    orc.run_step(hph, hfl, hns, g, par, sc, st, fixed_sweeps=1, seed=par.seed)
    t_sweep = time.perf_counter() - t0
    if not np.isfinite(hph.x).all():
        raise RuntimeError("oracle sweep produced non-finite values")
    spp = max(sweeps_per_step, 1.0)
    return {"value": round(1.0 / (t_sweep * spp), 5), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"one Picard sweep (phsevo!, fluidmech! by sparse LU, noise!, update_phase_velocities!, "
                      f"update!) of the {N}^2 step on the C oracle took {t_sweep:.2f} s; steps/s = 1 / (sweep time x "
                      f"{spp:.2f} sweeps/step of the GPU run)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--stencil-n", type=int, default=N_STENCIL)
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--weights", default=os.path.join(ROOT, "xcore.jl_amd", "lcc_weights.npz"),
                    help="trained learned-correction weights; when present the learned step is timed too")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--mg-pre", type=int, default=3)
    ap.add_argument("--mg-post", type=int, default=3)
    ap.add_argument("--mg-coarse", type=int, default=0, help="coarsest-level sweeps (0: the library's choice)")
    ap.add_argument("--mg-sub", type=int, default=1)
    ap.add_argument("--mg-omega", type=float, default=0.0,
                    help="damped-Jacobi weight of the multigrid smoother (0: chosen per level from the inertia ratio)")
    ap.add_argument("--mg-levels", type=int, default=0, help="most multigrid levels (0: down to 4 x 4)")
    ap.add_argument("--mg-jacobi", action="store_true",
                    help="damped-Jacobi fine-level smoothing instead of the default Chebyshev passes")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stencil", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-step", action="store_true", help="skip the full-step bench (profiling runs)")
    ap.add_argument("--no-ou", action="store_true", help="skip the noise! microbench (profiling runs)")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="before the stencil bench, copy a buffer of this many MiB three times (a known-size "
                         "kernel that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU runs are not part of it", file=sys.stderr)
        sys.exit(2)
    if args.no_step:
        import torch
        import xcore_amd as xc
        xc.load()
        torch.cuda.set_device(0)
        ms, info = float("nan"), {"sweeps_per_step": 1.0}
    else:
        print(f"bench: full step {args.n}^2, {args.warmup} warmup + {args.steps} timed", file=sys.stderr, flush=True)
        ms, info = step_bench(args)
        print(f"bench: plain {ms:.2f} ms/step {info}", file=sys.stderr, flush=True)
        plain = {"steps_per_s": round(1e3 / ms, 4), "ms_per_step": round(ms, 3), **info}
        learned = None
        if not args.no_learned and os.path.exists(args.weights):
            lms, linfo = step_bench(args, learned=True)
            print(f"bench: learned {lms:.2f} ms/step {linfo}", file=sys.stderr, flush=True)
            learned = {"steps_per_s": round(1e3 / lms, 4), "ms_per_step": round(lms, 3), **linfo}
    if args.calib_copy_mb > 0:
        import torch
        n = args.calib_copy_mb * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"bench: calibration copy of {args.calib_copy_mb} MiB x 3 done", file=sys.stderr, flush=True)
        del a, b
    variant = "multigrid"
    if not args.no_step:
        info = {"plain": plain}
        if learned is not None:  # value stays the faster, plain, variant; the learned one is reported beside it
            info["learned"] = learned
            info["learned_over_plain"] = round(learned["steps_per_s"] / plain["steps_per_s"], 4)
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms, 4),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (initialize! with a seeded perturbation; no dataset)",
        "config": {"workload": f"full run! time step at {args.n}^2 FP64, FGMRES Stokes to 1e-8 relative "
                               "residual, etacntr 1e3, rhox0 - rhom0 = 200, WENO5, bd2im, x periodic / z closed",
                   "grid": [args.n, args.n], "stokes_preconditioner": variant, **info},
    }
    if not args.no_stencil:
        print(f"bench: stencil microbench {args.stencil_n}^2", file=sys.stderr, flush=True)
        roof, kernels = stencil_bench(args.stencil_n, 20, 3)
        out["roofline"] = roof
        out["stencil_kernels"] = kernels
    if not args.no_ensemble:
        print("bench: ensemble of 64 members", file=sys.stderr, flush=True)
        out["ensemble"] = ensemble_bench()
        print("bench: ensemble of 64 whole runs", file=sys.stderr, flush=True)
        out["ensemble_steps"] = ensemble_steps_bench()
    if not args.no_ou:
        print("bench: noise! at 1024^2 (OU microbench)", file=sys.stderr, flush=True)
        out["ou_noise_1024"] = ou_bench()
    if not args.no_cpu:
        print("bench: config 1 (100^2, 20 steps) on the device and the CPU oracle", file=sys.stderr, flush=True)
        out["config1"] = c1_bench()
        if not args.no_stencil:
            print(f"bench: CPU oracle stencil calls at {args.stencil_n}^2", file=sys.stderr, flush=True)
            out["stencil_cpu"] = stencil_cpu_bench(args.stencil_n)
        print(f"bench: CPU oracle sweep at {args.cpu_n}^2", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, info["plain"]["sweeps_per_step"] if "plain" in info else 1.0)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
