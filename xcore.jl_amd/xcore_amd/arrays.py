"""Device matrices with the reference's memory layout.

xcore.jl stores every field as a column-major ``Matrix{Float64}`` with z as the
leading (contiguous) axis (grid.jl:17-19, state.jl:57-68).  Here a field is a
float64 torch tensor on the HIP device with shape ``(rows, cols)`` and strides
``(1, ld)``: ``f[iz, ix]`` addresses the same element as the Julia ``f[iz+1, ix+1]``,
and slices such as ``qx[1:Nz+1, :]`` are valid strided views, exactly like
``view(qx, 2:Nz+1, :)`` in the reference.
"""
import torch

from ._lib import XCMat

DTYPE = torch.float64


def _device(device=None):
    if device is None:
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def xcore_zeros(rows, cols, device=None, dtype=DTYPE):
    """backend.jl:12-13 ``xcore_zeros(backend, T, dims...)``: a zeroed column-major device matrix."""
    return torch.zeros((int(cols), int(rows)), dtype=dtype, device=_device(device)).t()


def to_device(a, device=None):
    """Copy a host array (numpy, any order) into a column-major device matrix."""
    t = torch.as_tensor(a, dtype=DTYPE)
    out = xcore_zeros(t.shape[0], t.shape[1], device)
    out.copy_(t)
    return out


def to_host(t):
    """Device matrix -> numpy array in Fortran order (the oracle's layout)."""
    import numpy as np
    return np.asfortranarray(t.detach().cpu().numpy())


def interior(f, halo):
    """backend.jl:22-23 ``interior(f, halo)``: view excluding ``halo`` rings."""
    return f[halo:f.shape[0] - halo, halo:f.shape[1] - halo]


def as_mat(t, name="array"):
    """torch tensor -> XCMat, checking dtype, device and the column-major layout."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a device tensor, got {type(t).__name__}")
    if t.dtype != DTYPE:
        raise TypeError(f"{name}: expected float64, got {t.dtype}")
    if t.device.type != "cuda":
        raise TypeError(f"{name}: expected a HIP device tensor, got {t.device}")
    if t.dim() != 2:
        raise AssertionError(f"{name}: expected a matrix, got {t.dim()} dims")
    rows, cols = t.shape
    s0, s1 = t.stride()
    if rows > 1 and s0 != 1:
        raise AssertionError(f"{name}: must be column-major (z contiguous); strides {t.stride()}")
    ld = s1 if cols > 1 else max(rows, 1)
    if cols > 1 and ld < rows:
        raise AssertionError(f"{name}: leading dimension {ld} < rows {rows}")
    return XCMat(t.data_ptr(), rows, cols, ld)
