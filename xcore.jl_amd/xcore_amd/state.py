"""Host-side mirrors of the reference's parameter, grid, scale and state types.

Parameters/Grid/Scales restate parameters.jl, grid.jl and scales.jl (host
scalars, cheap, computed once).  FluidState/PhaseState/NoiseState hold device
matrices with the exact shapes of state.jl:101-122, 218-259 and
noise.jl:61-93, and pack themselves into the C structs of include/xcore_hip.h
(same field order) for the ABI calls.
"""
import ctypes
import math
from dataclasses import dataclass, field, fields

import torch

from . import _lib
from ._lib import XCMat
from .arrays import DTYPE, as_mat, xcore_zeros

EPS = 2.220446049250313e-16


@dataclass
class Parameters:
    """parameters.jl:87-164 (defaults of ``Parameters{Float64}()``)."""
    D: float = 10.0
    N: int = 100
    L: float = 10.0
    x0: float = EPS
    dxr: float = 0.1
    dxg: float = 0.0
    seed: int = 15
    rhom0: float = 2700.0
    rhox0: float = 3200.0
    d0: float = 0.01
    g0: float = 10.0
    etam0: float = 1e1
    etax0: float = 1e18
    AA: tuple = (0.72, 0.19, 0.81, 0.20)   # row-major [11, 12, 21, 22]
    BB: tuple = (0.63, 0.37, 0.999, 0.001)
    CC: tuple = (2.09, 0.09, 0.37, 1.45)
    L0: float = 0.1
    l0: float = 0.1
    Da: float = 0.01
    Xi: float = 0.5
    Ptop: float = 1e5
    open_cnv: bool = False
    open_sgr: bool = False
    TINT: str = "bd2im"
    ADVN: str = "weno5"
    CFL: float = 0.5
    rtol: float = 1e-4
    atol: float = 1e-9
    maxit: int = 15
    alpha: float = 0.9
    gamma: float = 1e-3
    dtmax: float = 1e32
    etacntr: float = 1e8
    nrh: int = 1


@dataclass
class Grid:
    """grid.jl:52-75: h = D/N, Nz = N, Nx = max(1, round(L/h))."""
    Nz: int
    Nx: int
    h: float
    D: float
    L: float

    @classmethod
    def from_par(cls, par):
        h = par.D / par.N
        Nx = max(1, _round_half_even(par.L / h))
        return cls(par.N, Nx, h, par.D, par.L)


def _round_half_even(v):
    return int(round(v))  # Python rounds half to even, like Julia's round(Int, x)


@dataclass
class Scales:
    """scales.jl:349-462 (all characteristic scales; the kernels read a subset)."""
    vals: dict = field(default_factory=dict)

    def __getattr__(self, k):
        v = self.__dict__.get("vals", {})
        if k in v:
            return v[k]
        raise AttributeError(k)


SCALES_FULL = ("h0", "D0", "d0", "L0", "L0h", "l0", "l0h", "bnd_w", "rho0", "Drho0", "chi0", "Dchi0", "eta0", "W0",
               "w0", "W0l", "w0l", "W0t", "w0t", "W0i", "Ri0", "ReL0", "Rel0", "fReL0", "fRel0", "eII0", "ke0", "ks0",
               "kx0", "tW0", "tw0", "tk0", "ti0", "txi0", "t0", "dt0", "taue0", "taus0", "St0", "xie0", "xix0",
               "xis0", "G0", "etae0", "etat0", "p0", "Noe0", "Nox0", "Nos0", "Rc0", "Ra0", "ReD0", "Red0")
SCALES_IN = ("D", "d0", "L0", "l0", "x0", "Da", "rhom0", "rhox0", "etam0", "g0", "Xi", "h")


class ScalesIn(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in SCALES_IN]


class ScalesFull(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in SCALES_FULL]


_lib.register("xc_compute_scales", [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_void_p])


def compute_scales_batch(pars, grids, tol=1e-9, maxiter=200):
    """scales.jl:85-185 ``compute_scales`` of several (par, grid) pairs in one library call (ensemble members)."""
    n = len(pars)
    ins = (ScalesIn * n)()
    for i, (par, g) in enumerate(zip(pars, grids)):
        for k in SCALES_IN:
            setattr(ins[i], k, float(g.h if k == "h" else getattr(par, k)))
    outs = (ScalesFull * n)()
    code = _lib.lib().xc_compute_scales(n, ctypes.cast(ins, ctypes.c_void_p), float(tol), int(maxiter),
                                        ctypes.cast(outs, ctypes.c_void_p))
    if code != 0:
        raise ValueError("compute_scales: " + _lib.lib().xc_last_error(None).decode(errors="replace"))
    return [Scales({k: getattr(o, k) for k in SCALES_FULL}) for o in outs]


def compute_scales(par, grid, tol=1e-9, maxiter=200):
    """scales.jl:85-185 ``compute_scales(par, grid; tol, maxiter)`` -- fixed point on (W0, w0), computed by the
    library (host code, no device work)."""
    return compute_scales_batch([par], [grid], tol, maxiter)[0]


# ---- state containers -------------------------------------------------------

def _shape(code, Nz, Nx):
    return {"c": (Nz, Nx), "w": (Nz + 1, Nx + 2), "u": (Nz + 2, Nx + 1), "p": (Nz + 2, Nx + 2),
            "k": (Nz + 1, Nx + 1), "zf": (Nz + 1, Nx), "xf": (Nz, Nx + 1), "b": (Nz, Nx)}[code]


# state.jl:101-122
FLUID_FIELDS = [("W", "w"), ("U", "u"), ("P", "p"), ("eta", "c"), ("etaco", "k"), ("rho", "c"), ("rhow", "zf"),
                ("rhou", "xf"), ("Drho", "zf"), ("rhoo", "c"), ("rhooo", "c"), ("drhodt", "c"), ("drhodto", "c"),
                ("drhodtoo", "c"), ("MFS", "c"), ("rhoWo", "zf"), ("rhoWoo", "zf"), ("rhoUo", "xf"),
                ("rhoUoo", "xf")]
# state.jl:222-257
PHASE_FIELDS = [("x", "c"), ("m", "c"), ("X", "c"), ("M", "c"), ("Xo", "c"), ("Xoo", "c"), ("Mo", "c"),
                ("Moo", "c"), ("chi", "c"), ("mu", "c"), ("hasx", "b"), ("hasm", "b"), ("chiw", "w"), ("muw", "w"),
                ("chiu", "u"), ("muu", "u"), ("x_w", "zf"), ("m_w", "zf"), ("x_u", "xf"), ("m_u", "xf"),
                ("Drhom", "zf"), ("Drhox", "zf"), ("Pl", "c"), ("Pt", "c"), ("etamix", "c"), ("etae", "c"),
                ("ke", "c"), ("kx", "c"), ("etat", "c"), ("etas", "c"), ("etasw", "zf"), ("ks", "c"), ("V", "c"),
                ("Div_V", "c"), ("exx", "c"), ("ezz", "c"), ("exz", "k"), ("eII", "c"), ("vx", "c"), ("vm", "c"),
                ("txx", "c"), ("tzz", "c"), ("txz", "k"), ("tII", "c"), ("fReL", "c"), ("ReL", "c"), ("ReD", "c"),
                ("Ra", "c"), ("fRel", "c"), ("Rel", "c"), ("Red", "c"), ("Rc", "c"), ("Wx", "w"), ("Wm", "w"),
                ("Ux", "u"), ("Um", "u"), ("wx", "w"), ("wm", "w"), ("bndtaperw", "w"), ("bndshape", "c"),
                ("advn_X", "c"), ("advn_M", "c"), ("advn_rho", "c"), ("qx_advn_X", "u"), ("qz_advn_X", "w"),
                ("qx_advn_M", "u"), ("qz_advn_M", "w"), ("qx_dffn_X", "u"), ("qz_dffn_X", "w"),
                ("qx_dffn_M", "u"), ("qz_dffn_M", "w"), ("dffn_X", "c"), ("Gx", "c"), ("dXdt", "c"),
                ("dXdto", "c"), ("dXdtoo", "c")]
# noise.jl:82-92 (matrices only; padL0/padl0/fL/fl are scalars in the struct)
NOISE_MATS_A = [("psie", "c"), ("psix", "c"), ("psis", "c"), ("psieo", "c"), ("psixo", "c"), ("psiso", "c"),
                ("re", "c"), ("rs", "c")]
NOISE_MATS_B = [("xiew", "w"), ("xieu", "u"), ("xixw", "w"), ("xixu", "u"), ("xisw", "w"), ("xisu", "u"),
                ("xie", "c"), ("xix", "c"), ("xis", "c")]


class XCBMat(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


def _struct(name, flist):
    return type(name, (ctypes.Structure,),
                {"_fields_": [(n, XCBMat if c == "b" else XCMat) for n, c in flist]})


FluidStruct = _struct("xc_fluid", FLUID_FIELDS)
PhaseStruct = _struct("xc_phase", PHASE_FIELDS)


class NoiseStruct(ctypes.Structure):
    _fields_ = ([(n, XCMat) for n, _ in NOISE_MATS_A] +
                [("padL0", ctypes.c_int64), ("padl0", ctypes.c_int64), ("fL", ctypes.c_double),
                 ("fl", ctypes.c_double)] + [(n, XCMat) for n, _ in NOISE_MATS_B] +
                [("seed", ctypes.c_uint64), ("draws", ctypes.c_uint64)])


class GridStruct(ctypes.Structure):
    _fields_ = [("Nz", ctypes.c_int64), ("Nx", ctypes.c_int64), ("h", ctypes.c_double), ("D", ctypes.c_double),
                ("L", ctypes.c_double)]


class ParStruct(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("rhom0", "rhox0", "d0", "g0", "etam0", "etax0")] + \
               [("AA", ctypes.c_double * 4), ("BB", ctypes.c_double * 4), ("CC", ctypes.c_double * 4)] + \
               [(n, ctypes.c_double) for n in ("Xi", "Ptop", "alpha", "gamma", "CFL", "dtmax", "etacntr")] + \
               [(n, ctypes.c_double) for n in ("x0", "Da", "dxr", "dxg")] + \
               [("seed", ctypes.c_uint64), ("open_sgr", ctypes.c_int)]


SCALES_FIELDS = ("L0", "L0h", "l0", "l0h", "D0", "G0", "txi0", "bnd_w", "dt0")


class ScalesStruct(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in SCALES_FIELDS]


def grid_struct(g):
    return GridStruct(int(g.Nz), int(g.Nx), float(g.h), float(g.D), float(g.L))


def par_struct(p):
    s = ParStruct()
    for n in ("rhom0", "rhox0", "d0", "g0", "etam0", "etax0", "Xi", "Ptop", "alpha", "gamma", "CFL", "dtmax",
              "etacntr"):
        setattr(s, n, float(getattr(p, n)))
    s.AA = (ctypes.c_double * 4)(*map(float, p.AA))
    s.BB = (ctypes.c_double * 4)(*map(float, p.BB))
    s.CC = (ctypes.c_double * 4)(*map(float, p.CC))
    for n in ("x0", "Da", "dxr", "dxg"):
        setattr(s, n, float(getattr(p, n)))
    s.seed = int(p.seed)
    s.open_sgr = int(bool(p.open_sgr))
    return s


def scales_struct(sc):
    return ScalesStruct(*(float(getattr(sc, n)) for n in SCALES_FIELDS))


def _bmat(t):
    if t.dtype != torch.uint8 or t.device.type != "cuda" or (t.shape[0] > 1 and t.stride(0) != 1):
        raise AssertionError("Bool matrix must be a column-major uint8 device tensor")
    ld = t.stride(1) if t.shape[1] > 1 else max(t.shape[0], 1)
    return XCBMat(t.data_ptr(), t.shape[0], t.shape[1], ld)


class _State:
    FIELDS = []

    def __init__(self, Nz, Nx, device=None):
        self.Nz, self.Nx = int(Nz), int(Nx)
        for n, c in self.FIELDS:
            r, k = _shape(c, self.Nz, self.Nx)
            if c == "b":
                setattr(self, n, xcore_zeros(r, k, device, dtype=torch.uint8))
            else:
                setattr(self, n, xcore_zeros(r, k, device))

    def pack(self):
        s = self.STRUCT()
        for n, c in self.FIELDS:
            t = getattr(self, n)
            setattr(s, n, _bmat(t) if c == "b" else as_mat(t, n))
        return s

    def arrays(self):
        return {n: getattr(self, n) for n, _ in self.FIELDS}


class FluidState(_State):
    """state.jl:74-122 ``FluidState(T, backend, Nz, Nx)`` on the HIP device."""
    FIELDS = FLUID_FIELDS
    STRUCT = FluidStruct


class PhaseState(_State):
    """state.jl:165-259 ``PhaseState(T, backend, Nz, Nx)`` on the HIP device."""
    FIELDS = PHASE_FIELDS
    STRUCT = PhaseStruct


class NoiseState:
    """noise.jl:38-93 ``NoiseState(T, grid, scales; seed)`` on the HIP device.

    The Gaussian filters are separable (exp(-s^2(kz^2+kx^2)) = exp(-s^2 kz^2) exp(-s^2 kx^2)), so the
    library keeps two 1-D tables per filter in its context instead of the padded 2-D Gkpe/Gkps.
    """

    def __init__(self, grid, scales, seed=0, device=None):
        Nz, Nx = grid.Nz, grid.Nx
        self.Nz, self.Nx = Nz, Nx
        h, L0h, l0h = grid.h, scales.L0h, scales.l0h
        self.padL0 = 4 * math.ceil(L0h / h)
        self.padl0 = 4 * math.ceil(l0h / h)
        self.fL = 2 / math.sqrt(1 - math.exp(-h ** 2 / (2 * L0h ** 2)))
        self.fl = 2 / math.sqrt(1 - math.exp(-h ** 2 / (2 * l0h ** 2)))
        self.seed = int(seed)
        self.draws = 0   # how many white-noise draws were made (the RNG counter)
        for n, c in NOISE_MATS_A + NOISE_MATS_B:
            setattr(self, n, xcore_zeros(*_shape(c, Nz, Nx), device))

    def pack(self):
        s = NoiseStruct()
        for n, _ in NOISE_MATS_A + NOISE_MATS_B:
            setattr(s, n, as_mat(getattr(self, n), n))
        s.padL0, s.padl0, s.fL, s.fl = int(self.padL0), int(self.padl0), float(self.fL), float(self.fl)
        s.seed, s.draws = self.seed, self.draws
        return s

    def arrays(self):
        return {n: getattr(self, n) for n, _ in NOISE_MATS_A + NOISE_MATS_B}
