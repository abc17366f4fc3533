"""xcore_amd -- MI355X (gfx950) implementation of xcore.jl's per-timestep hot path.

Host-side mirror of the reference's Julia interface over the C ABI of
libxcore_hip.so (include/xcore_hip.h).  Device fields are float64 torch
tensors in the reference's column-major layout (see ``arrays``).
"""
from ._lib import LIB_PATH, NoConvergence, XCoreError, declared_symbols, lib
from .arrays import as_mat, interior, to_device, to_host, xcore_zeros
from .kernels import (advect_, advect_centered_, advect_centered_with_flux_, advect_with_flux_, context,
                      diffus_, diffus_centered_, diffus_centered_with_flux_, diffus_with_flux_,
                      embed_interior_, exact_arithmetic, autotune, advect_tile, set_advect_tile, fill_ghosts_,
                      scheme_halo, diag_stream_)
from .state import (FluidState, Grid, NoiseState, Parameters, PhaseState, Scales, compute_scales)
from .fields import (compute_fft_filter, noise_, normal_fill_, phsevo_, store_noise_, store_previous_, time_coefs,
                     update_, update_dt, update_phase_velocities_)
from .stokes import (BnchmData, fluidmech_, lc_restrict, stokes_apply, stokes_precond, stokes_rhs, stokes_solve,
                     stokes_vector_size)
from .driver import (History, StepState, history_sums, initialize_, record_history_, run_, run_step_,
                     step_opts)
from . import learned
from .restart import load_checkpoint_, resolve_restart, restart_path, save_checkpoint
from .ensemble import BatchPlan, advect_centered_batch_, diffus_centered_batch_


def load():
    """Load libxcore_hip.so; raises if it is missing (there is no fallback)."""
    return lib()
