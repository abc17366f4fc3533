"""restart.jl: checkpoints of the device state, for resuming a run.

The reference writes JLD2 (HDF5) files; neither is available here, so a
checkpoint is a NumPy ``.npz`` archive holding the same keys: ``fluid/<field>``,
``phase/<field>``, ``noise/<field>`` (the noise's dynamic arrays), ``hst/<field>``
and ``meta/time``, ``meta/dt``, ``meta/step``.  The noise generator is
counter-based (Philox keyed by seed, a draw counter and a stream), so its state
is ``noise/seed`` and ``noise/draws`` in place of the reference's serialised
MersenneTwister.  Files are read with ``allow_pickle=False``: a checkpoint is
arrays and numbers only.  Device arrays are staged through host memory; the
restored run continues bit-for-bit (the step is deterministic given the state).
"""
import os

import numpy as np
import torch

from .arrays import to_host


def restart_path(outdir, runID, frame=-1):
    """restart.jl:37-46 ``restart_path(outdir, runID; frame)``: ``<runID>_cont`` for frame < 0, else
    ``<runID>_<frame:04d>``, in ``outdir/runID`` (``.npz`` in place of ``.jld2``)."""
    d = os.path.join(outdir, runID)
    if frame < 0:
        return os.path.join(d, f"{runID}_cont.npz")
    return os.path.join(d, f"{runID}_{int(frame):04d}.npz")


def save_checkpoint(path, phase, fluid, ns, hst, time, dt, step):
    """restart.jl:53-84 ``save_checkpoint(path, phase, fluid, ns, hst; time, dt, step)``.

    The archive is written next to ``path`` and renamed into place, so an interrupted write never
    leaves a truncated checkpoint under the final name."""
    out = {}
    for prefix, st in (("fluid", fluid), ("phase", phase), ("noise", ns)):
        for n, a in st.arrays().items():
            out[f"{prefix}/{n}"] = to_host(a)
    out["noise/seed"] = np.int64(ns.seed)
    out["noise/draws"] = np.int64(ns.draws)
    for n, v in list(hst.vectors().items()) + list(hst.rates().items()):
        out[f"hst/{n}"] = np.asarray(v, dtype=np.float64)
    out["meta/time"] = np.float64(time)
    out["meta/dt"] = np.float64(dt)
    out["meta/step"] = np.int64(step)
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp.npz"
    np.savez(tmp, **out)
    os.replace(tmp, path)


def load_checkpoint_(path, phase, fluid, ns, hst):
    """restart.jl:94-127 ``load_checkpoint!(path, phase, fluid, ns, hst) -> (time, dt, step)``: overwrite the
    arrays of the pre-allocated states (same grid as when the checkpoint was made) in place."""
    if not os.path.isfile(path):
        raise FileNotFoundError(f"load_checkpoint!: file not found: {path}")
    with np.load(path, allow_pickle=False) as f:
        for prefix, st in (("fluid", fluid), ("phase", phase), ("noise", ns)):
            for n, a in st.arrays().items():
                v = f[f"{prefix}/{n}"]
                if tuple(v.shape) != tuple(a.shape):
                    raise ValueError(f"load_checkpoint!: {prefix}/{n} has shape {v.shape}, the state {tuple(a.shape)}")
                a.copy_(torch.from_numpy(np.ascontiguousarray(v)).to(dtype=a.dtype))
        ns.seed = int(f["noise/seed"])
        ns.draws = int(f["noise/draws"])
        for n in hst.vectors():
            setattr(hst, n, [float(x) for x in f[f"hst/{n}"].reshape(-1)])
        for n in hst.rates():
            setattr(hst, n, float(f[f"hst/{n}"]))
        return float(f["meta/time"]), float(f["meta/dt"]), int(f["meta/step"])


def resolve_restart(outdir, runID, restart):
    """restart.jl:140-146 ``resolve_restart(outdir, runID, restart)``: None for restart == 0, the rolling
    ``_cont`` file for restart < 0, frame ``restart`` otherwise; FileNotFoundError if it is missing."""
    if restart == 0:
        return None
    p = restart_path(outdir, runID, frame=-1 if restart < 0 else restart)
    if not os.path.isfile(p):
        raise FileNotFoundError(f"resolve_restart: checkpoint not found: {p}")
    return p
