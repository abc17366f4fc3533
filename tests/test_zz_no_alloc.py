"""xc_history_stats_noalloc (no allocation; partial sums in the library's fixed device array) against
xc_history_stats through the C ABI: two calls in a row, the second on a smaller grid, both equal to the
allocating entry point's results bit for bit, plus the argument checks."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _state(xc, Nz, Nx, seed):
    par = xc.Parameters(D=10.0, N=Nz, L=10.0 * Nx / Nz, etam0=10.0, x0=0.1, Da=0.02)
    g = xc.Grid.from_par(par)
    sc = xc.compute_scales(par, g)
    ph, fl, ns = xc.PhaseState(Nz, Nx), xc.FluidState(Nz, Nx), xc.NoiseState(g, sc, seed=seed)
    xc.initialize_(ph, fl, ns, None, g, par, sc)
    return ph, fl, ns


def test_noalloc_matches_history_stats(xc):
    from xcore_amd import _lib
    L = _lib.lib()
    L.xc_history_stats_noalloc.argtypes = [_P, _P, _P, _P, _P]
    L.xc_history_stats_noalloc.restype = ctypes.c_int
    n = xc.driver.XC_HIST_NSTATS
    stream = torch.cuda.current_stream().cuda_stream
    for shape, seed in (((96, 80), 1), ((40, 24), 2)):
        print(f"no-alloc history stats {shape}", flush=True)
        ph, fl, ns = _state(xc, *shape, seed)
        want = np.array(xc.driver.history_stats(ph, fl, ns))
        out = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
        ps, fs, nss = ph.pack(), fl.pack(), ns.pack()
        code = L.xc_history_stats_noalloc(_P(stream), ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss),
                                          _P(out.data_ptr()))
        assert code == 0, L.xc_last_error(None)
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        assert np.array_equal(got, want), np.abs(got - want).max()
    # refused: NULL output, and a W that does not match the grid
    assert L.xc_history_stats_noalloc(_P(stream), ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss),
                                      None) == _lib.XC_EINVAL_ARG
    fs.W.rows -= 1
    assert L.xc_history_stats_noalloc(_P(stream), ctypes.byref(ps), ctypes.byref(fs), ctypes.byref(nss),
                                      _P(out.data_ptr())) == _lib.XC_EINVAL_SHAPE
    assert b"shapes" in L.xc_last_error(None)
