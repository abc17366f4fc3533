"""Benchmark of the xcore.jl hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload (one "step"): the stencil + noise microbench at 2048^2, FP64 --
four scalar fields (sums of 32 random Fourier modes, amplitude ~ k^-2) are
each advected with WENO5 and with QUICK through advect_centered_with_flux!
by a divergence-free staggered velocity (streamfunction sin(pi x) sin(pi z)
plus 16 seeded modes), and diffused through diffus_centered_with_flux! with a
lognormal (sigma = 1) diffusivity; x periodic, z closed.  Inputs are resident
in HBM before timing; the four field/output sets (604 MB) rotate so the
256 MB Infinity Cache cannot hold one step's working set.

Prints one JSON line (rank 0).  `roofline` is for the dominant kernel
(WENO5 advect_centered_with_flux!): algorithmic bytes per launch / its
average duration measured here with HIP events on the launch stream.
`cpu_baseline` times the CPU oracle (the C restatement of the reference,
OpenMP) on one full step of the same workload.
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


def gpu_bench(args):
    import torch
    import xcore_amd as xc
    xc.load()
    N = args.n
    fields_h, u_h, w_h, k_h, h = make_inputs(N)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    u = xc.to_device(u_h)
    w = xc.to_device(w_h)
    k = xc.to_device(k_h)
    sets = []
    for f in fields_h:
        sets.append(dict(f=xc.to_device(f), adv=xc.xcore_zeros(N, N), qx=xc.xcore_zeros(N + 2, N + 1),
                         qz=xc.xcore_zeros(N + 1, N + 2), dff=xc.xcore_zeros(N, N),
                         dqx=xc.xcore_zeros(N + 2, N + 1), dqz=xc.xcore_zeros(N + 1, N + 2)))
    torch.cuda.synchronize()

    def step():
        for s in sets:
            xc.advect_centered_with_flux_(s["adv"], s["qx"], s["qz"], s["f"], u, w, h, "weno5")
            xc.advect_centered_with_flux_(s["adv"], s["qx"], s["qz"], s["f"], u, w, h, "quick")
            xc.diffus_centered_with_flux_(s["dff"], s["dqx"], s["dqz"], s["f"], k, h)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps

    # per-launch duration of the dominant kernel, HIP events on its stream
    stream = torch.cuda.current_stream()
    reps = max(8, 4 * args.steps)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for r in range(reps):
        s = sets[r % NFIELDS]
        xc.advect_centered_with_flux_(s["adv"], s["qx"], s["qz"], s["f"], u, w, h, "weno5")
    e1.record(stream)
    e1.synchronize()
    weno_us = 1e3 * e0.elapsed_time(e1) / reps
    e0.record(stream)
    for r in range(reps):
        s = sets[r % NFIELDS]
        xc.advect_centered_with_flux_(s["adv"], s["qx"], s["qz"], s["f"], u, w, h, "quick")
    e1.record(stream)
    e1.synchronize()
    quick_us = 1e3 * e0.elapsed_time(e1) / reps
    e0.record(stream)
    for r in range(reps):
        s = sets[r % NFIELDS]
        xc.diffus_centered_with_flux_(s["dff"], s["dqx"], s["dqz"], s["f"], k, h)
    e1.record(stream)
    e1.synchronize()
    diff_us = 1e3 * e0.elapsed_time(e1) / reps

    # algorithmic bytes per launch: every input and output array touched once
    adv_bytes = 8 * (N * N + N * (N + 1) + (N + 1) * N + N * N + (N + 2) * (N + 1) + (N + 1) * (N + 2))
    dif_bytes = 8 * (N * N + N * N + N * N + (N + 2) * (N + 1) + (N + 1) * (N + 2))
    ach = adv_bytes / (weno_us * 1e-6) / 1e9
    kernels = {
        "weno5_advect_with_flux": {"us": weno_us, "GB/s": ach, "hbm_frac": ach / HBM_PEAK_GBS},
        "quick_advect_with_flux": {"us": quick_us, "GB/s": adv_bytes / (quick_us * 1e-6) / 1e9,
                                   "hbm_frac": adv_bytes / (quick_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
        "diffus_with_flux": {"us": diff_us, "GB/s": dif_bytes / (diff_us * 1e-6) / 1e9,
                             "hbm_frac": dif_bytes / (diff_us * 1e-6) / 1e9 / HBM_PEAK_GBS},
    }
    roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": "k_advect<WENO5,centered,flux>", "bytes_per_launch": adv_bytes,
                "avg_launch_us": round(weno_us, 2)}
    return ms_per_step, roofline, kernels


def cpu_baseline(N):
    """One full step of the same workload on the CPU oracle (OpenMP)."""
    from oracle import oracle as orc
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, int(os.environ.get("OMP_NUM_THREADS", cores))))
    orc.set_threads(cores)
    fields, u, w, k, h = make_inputs(N)
    adv, qx, qz = orc.zeros(N, N), orc.zeros(N + 2, N + 1), orc.zeros(N + 1, N + 2)
    dff = orc.zeros(N, N)
    t0 = time.perf_counter()
    for f in fields:
        orc.advect_centered(adv, f, u, w, h, "weno5", qx=qx, qz=qz)
        orc.advect_centered(adv, f, u, w, h, "quick", qx=qx, qz=qz)
        orc.diffus_centered(dff, f, k, h, qx=qx, qz=qz)
    dt = time.perf_counter() - t0
    return {"value": round(1.0 / dt, 4), "unit": "steps/s", "cores": cores, "kind": "port",
            "sample": f"one full step ({NFIELDS} fields x WENO5+QUICK advect_centered_with_flux! + "
                      f"diffus_centered_with_flux!) at {N}^2 on the C oracle, {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_STENCIL)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: this build measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the hot path would shard as described in DESIGN.md", file=sys.stderr)
        sys.exit(2)
    ms, roof, kernels = gpu_bench(args)
    out = {
        "metric": METRIC,
        "value": round(1e3 / ms, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"stencil microbench {args.n}^2 FP64: {NFIELDS} fields x (WENO5 + QUICK "
                               "advect_centered_with_flux!, lognormal-k diffus_centered_with_flux!), "
                               "x periodic / z closed",
                   "grid": [args.n, args.n], "fields": NFIELDS},
        "roofline": roof,
        "kernels": kernels,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
