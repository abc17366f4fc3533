"""ctypes front-end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this module; it checks the HIP path and is never the thing measured or
shipped.  Arrays are numpy float64 in Fortran (column-major) order, i.e. the
memory layout of a Julia ``Matrix{Float64}`` with z as the fast axis.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

SCHEMES = {"centr": 0, "upwd1": 1, "quick": 2, "fromm": 3, "weno3": 4, "weno5": 5, "tvdim": 6}
BCS = {"periodic": 0, "closed": 1}
ERRORS = {1: "shape", 2: "scheme", 3: "bc", 5: "noconv", 6: "arg"}


class OMat(ctypes.Structure):
    _fields_ = [("p", ctypes.POINTER(ctypes.c_double)), ("rows", ctypes.c_int64),
                ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


class OracleError(RuntimeError):
    def __init__(self, code, fn):
        super().__init__(f"oracle {fn}: status {code} ({ERRORS.get(code, '?')})")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def zeros(rows, cols):
    return np.zeros((rows, cols), dtype=np.float64, order="F")


def om(a):
    """Wrap a column-major float64 array (or a column-major view) as OMat."""
    if a.dtype != np.float64:
        raise TypeError("oracle arrays are float64")
    if a.ndim != 2:
        raise ValueError("oracle arrays are 2-D")
    s0, s1 = a.strides
    if a.shape[0] > 1 and s0 != 8:
        raise ValueError("oracle arrays must be column-major (z fastest)")
    ld = s1 // 8 if a.shape[1] > 1 else max(a.shape[0], 1)
    ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    return OMat(ptr, a.shape[0], a.shape[1], ld)


def _chk(code, fn):
    if code != 0:
        raise OracleError(code, fn)


def set_threads(n):
    lib().or_set_threads(int(n))


def scheme_halo(scheme):
    return lib().or_scheme_halo(SCHEMES[scheme])


def fill_ghosts(f, halo, xbc="periodic", zbc="closed"):
    _chk(lib().or_fill_ghosts(om(f), int(halo), BCS.get(xbc, 9), BCS.get(zbc, 9)), "fill_ghosts")
    return f


def embed_interior(fh, f, halo):
    _chk(lib().or_embed_interior(om(fh), om(f), int(halo)), "embed_interior")
    return fh


def _h(h):
    return ctypes.c_double(float(h))


def advect(adv, fh, u, w, h, scheme):
    _chk(lib().or_advect(om(adv), om(fh), om(u), om(w), _h(h), SCHEMES.get(scheme, 99)), "advect")
    return adv


def advect_with_flux(adv, qx, qz, fh, u, w, h, scheme, xbc="periodic", zbc="closed"):
    _chk(lib().or_advect_with_flux(om(adv), om(qx), om(qz), om(fh), om(u), om(w), _h(h),
                                   SCHEMES.get(scheme, 99), BCS.get(xbc, 9), BCS.get(zbc, 9)),
         "advect_with_flux")
    return adv


def advect_centered(adv, f, u, w, h, scheme, xbc="periodic", zbc="closed", qx=None, qz=None):
    pqx = ctypes.byref(om(qx)) if qx is not None else None
    pqz = ctypes.byref(om(qz)) if qz is not None else None
    _chk(lib().or_advect_centered(om(adv), pqx, pqz, om(f), om(u), om(w), _h(h),
                                  SCHEMES.get(scheme, 99), BCS.get(xbc, 9), BCS.get(zbc, 9)),
         "advect_centered")
    return adv


def diffus(dff, fh, kh, h, halo=1):
    _chk(lib().or_diffus(om(dff), om(fh), om(kh), _h(h), int(halo)), "diffus")
    return dff


def diffus_with_flux(dff, qx, qz, fh, kh, h, halo=1, xbc="periodic", zbc="closed"):
    _chk(lib().or_diffus_with_flux(om(dff), om(qx), om(qz), om(fh), om(kh), _h(h), int(halo),
                                   BCS.get(xbc, 9), BCS.get(zbc, 9)), "diffus_with_flux")
    return dff


def diffus_centered(dff, f, k, h, xbc="periodic", zbc="closed", qx=None, qz=None):
    pqx = ctypes.byref(om(qx)) if qx is not None else None
    pqz = ctypes.byref(om(qz)) if qz is not None else None
    _chk(lib().or_diffus_centered(om(dff), pqx, pqz, om(f), om(k), _h(h),
                                  BCS.get(xbc, 9), BCS.get(zbc, 9)), "diffus_centered")
    return dff


# ---- state containers (numpy, Fortran order) and field updates --------------
# Field order mirrors or_fluid / or_phase / or_noise in oracle.h (state.jl order).

def _shape(code, Nz, Nx):
    return {"c": (Nz, Nx), "w": (Nz + 1, Nx + 2), "u": (Nz + 2, Nx + 1), "p": (Nz + 2, Nx + 2),
            "k": (Nz + 1, Nx + 1), "zf": (Nz + 1, Nx), "xf": (Nz, Nx + 1), "b": (Nz, Nx)}[code]


FLUID = [("W", "w"), ("U", "u"), ("P", "p"), ("eta", "c"), ("etaco", "k"), ("rho", "c"), ("rhow", "zf"),
         ("rhou", "xf"), ("Drho", "zf"), ("rhoo", "c"), ("rhooo", "c"), ("drhodt", "c"), ("drhodto", "c"),
         ("drhodtoo", "c"), ("MFS", "c"), ("rhoWo", "zf"), ("rhoWoo", "zf"), ("rhoUo", "xf"), ("rhoUoo", "xf")]
PHASE = [("x", "c"), ("m", "c"), ("X", "c"), ("M", "c"), ("Xo", "c"), ("Xoo", "c"), ("Mo", "c"), ("Moo", "c"),
         ("chi", "c"), ("mu", "c"), ("hasx", "b"), ("hasm", "b"), ("chiw", "w"), ("muw", "w"), ("chiu", "u"),
         ("muu", "u"), ("x_w", "zf"), ("m_w", "zf"), ("x_u", "xf"), ("m_u", "xf"), ("Drhom", "zf"), ("Drhox", "zf"),
         ("Pl", "c"), ("Pt", "c"), ("etamix", "c"), ("etae", "c"), ("ke", "c"), ("kx", "c"), ("etat", "c"),
         ("etas", "c"), ("etasw", "zf"), ("ks", "c"), ("V", "c"), ("Div_V", "c"), ("exx", "c"), ("ezz", "c"),
         ("exz", "k"), ("eII", "c"), ("vx", "c"), ("vm", "c"), ("txx", "c"), ("tzz", "c"), ("txz", "k"),
         ("tII", "c"), ("fReL", "c"), ("ReL", "c"), ("ReD", "c"), ("Ra", "c"), ("fRel", "c"), ("Rel", "c"),
         ("Red", "c"), ("Rc", "c"), ("Wx", "w"), ("Wm", "w"), ("Ux", "u"), ("Um", "u"), ("wx", "w"), ("wm", "w"),
         ("bndtaperw", "w"), ("bndshape", "c"), ("advn_X", "c"), ("advn_M", "c"), ("advn_rho", "c"),
         ("qx_advn_X", "u"), ("qz_advn_X", "w"), ("qx_advn_M", "u"), ("qz_advn_M", "w"), ("qx_dffn_X", "u"),
         ("qz_dffn_X", "w"), ("qx_dffn_M", "u"), ("qz_dffn_M", "w"), ("dffn_X", "c"), ("Gx", "c"), ("dXdt", "c"),
         ("dXdto", "c"), ("dXdtoo", "c")]
NOISE_A = [("psie", "c"), ("psix", "c"), ("psis", "c"), ("psieo", "c"), ("psixo", "c"), ("psiso", "c"),
           ("re", "c"), ("rs", "c")]
NOISE_B = [("xiew", "w"), ("xieu", "u"), ("xixw", "w"), ("xixu", "u"), ("xisw", "w"), ("xisu", "u"),
           ("xie", "c"), ("xix", "c"), ("xis", "c")]


class OBMat(ctypes.Structure):
    _fields_ = [("p", ctypes.c_void_p), ("rows", ctypes.c_int64), ("cols", ctypes.c_int64), ("ld", ctypes.c_int64)]


def _mk(name, flist):
    return type(name, (ctypes.Structure,), {"_fields_": [(n, OBMat if c == "b" else OMat) for n, c in flist]})


OFluid = _mk("or_fluid", FLUID)
OPhase = _mk("or_phase", PHASE)


class ONoise(ctypes.Structure):
    _fields_ = ([(n, OMat) for n, _ in NOISE_A] + [("padL0", ctypes.c_int64), ("padl0", ctypes.c_int64),
                                                   ("fL", ctypes.c_double), ("fl", ctypes.c_double)] +
                [(n, OMat) for n, _ in NOISE_B])


class OGrid(ctypes.Structure):
    _fields_ = [("Nz", ctypes.c_int64), ("Nx", ctypes.c_int64), ("h", ctypes.c_double), ("D", ctypes.c_double),
                ("L", ctypes.c_double)]


class OPar(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("rhom0", "rhox0", "d0", "g0", "etam0", "etax0")] + \
               [("AA", ctypes.c_double * 4), ("BB", ctypes.c_double * 4), ("CC", ctypes.c_double * 4)] + \
               [(n, ctypes.c_double) for n in ("Xi", "Ptop", "alpha", "gamma", "CFL", "dtmax", "etacntr")]


class OScales(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("L0", "L0h", "l0", "l0h", "D0", "G0", "txi0")]


def _obm(a):
    assert a.dtype == np.uint8 and a.flags.f_contiguous
    return OBMat(a.ctypes.data, a.shape[0], a.shape[1], max(a.shape[0], 1))


class HostState:
    """A numpy (Fortran-order) state container with the reference's shapes."""

    def __init__(self, flist, Nz, Nx, struct):
        self.flist, self.Nz, self.Nx, self.struct = flist, Nz, Nx, struct
        for n, c in flist:
            r, k = _shape(c, Nz, Nx)
            setattr(self, n, np.zeros((r, k), dtype=np.uint8 if c == "b" else np.float64, order="F"))

    def pack(self):
        s = self.struct()
        for n, c in self.flist:
            a = getattr(self, n)
            setattr(s, n, _obm(a) if c == "b" else om(a))
        return s

    def arrays(self):
        return {n: getattr(self, n) for n, _ in self.flist}


def fluid_state(Nz, Nx):
    return HostState(FLUID, Nz, Nx, OFluid)


def phase_state(Nz, Nx):
    return HostState(PHASE, Nz, Nx, OPhase)


class HostNoise:
    def __init__(self, Nz, Nx, padL0, padl0, fL, fl):
        self.Nz, self.Nx = Nz, Nx
        self.padL0, self.padl0, self.fL, self.fl = padL0, padl0, fL, fl
        for n, c in NOISE_A + NOISE_B:
            setattr(self, n, np.zeros(_shape(c, Nz, Nx), order="F"))

    def pack(self):
        s = ONoise()
        for n, _ in NOISE_A + NOISE_B:
            setattr(s, n, om(getattr(self, n)))
        s.padL0, s.padl0, s.fL, s.fl = int(self.padL0), int(self.padl0), float(self.fL), float(self.fl)
        return s

    def arrays(self):
        return {n: getattr(self, n) for n, _ in NOISE_A + NOISE_B}


def ogrid(g):
    return OGrid(int(g.Nz), int(g.Nx), float(g.h), float(g.D), float(g.L))


def opar(p):
    s = OPar()
    for n in ("rhom0", "rhox0", "d0", "g0", "etam0", "etax0", "Xi", "Ptop", "alpha", "gamma", "CFL", "dtmax",
              "etacntr"):
        setattr(s, n, float(getattr(p, n)))
    s.AA = (ctypes.c_double * 4)(*map(float, p.AA))
    s.BB = (ctypes.c_double * 4)(*map(float, p.BB))
    s.CC = (ctypes.c_double * 4)(*map(float, p.CC))
    return s


def oscales(sc):
    return OScales(*(float(getattr(sc, n)) for n in ("L0", "L0h", "l0", "l0h", "D0", "G0", "txi0")))


_R = ctypes.byref
TINTS = {"be1im": 0, "bd2im": 1, "cn2si": 2, "bd2si": 3}


def time_coefs(tint, step):
    c = (ctypes.c_double * 6)()
    _chk(lib().or_time_coefs(TINTS.get(str(tint).lstrip(":"), 9), ctypes.c_int64(int(step)), c), "time_coefs")
    return tuple(c)


def update(ph, fl, g, par, sc, xbc="periodic", zbc="closed"):
    p, f = ph.pack(), fl.pack()
    _chk(lib().or_update(_R(p), _R(f), _R(ogrid(g)), _R(opar(par)), _R(oscales(sc)), BCS.get(xbc, 9),
                         BCS.get(zbc, 9)), "update")


def update_stage(name, ph, fl, g, par, sc, xbc="periodic", zbc="closed"):
    """Run one update! sub-step (for per-stage tests)."""
    p, f = ph.pack(), fl.pack()
    L = lib()
    if name == "volume_fractions":
        st = L.or_update_volume_fractions(_R(p), _R(f), _R(opar(par)))
    elif name == "interpolations":
        st = L.or_update_interpolations(_R(p), _R(f), _R(ogrid(g)), _R(opar(par)), BCS[xbc], BCS[zbc])
    elif name == "pressure":
        st = L.or_update_pressure(_R(p), _R(f), _R(ogrid(g)), _R(opar(par)))
    elif name == "rheology":
        st = L.or_update_rheology(_R(p), _R(opar(par)))
    elif name == "kinematics":
        st = L.or_update_kinematics(_R(p), _R(f), _R(ogrid(g)))
    elif name == "viscosity":
        st = L.or_update_viscosity(_R(p), _R(f), _R(ogrid(g)), _R(opar(par)), _R(oscales(sc)))
    elif name == "segregation_viscosity":
        st = L.or_update_segregation_viscosity(_R(p), _R(f), _R(opar(par)), _R(oscales(sc)), BCS[zbc])
    elif name == "stresses":
        st = L.or_update_stresses(_R(p), _R(f), _R(ogrid(g)), _R(oscales(sc)))
    else:
        raise ValueError(name)
    _chk(st, name)


def phsevo(ph, fl, g, par, sc, scheme="weno5", xbc="periodic", zbc="closed", coef=(1, 1, 0, 1, 0, 0), dt=1.0,
           alpha=None):
    alpha = par.alpha if alpha is None else alpha
    c = (ctypes.c_double * 6)(*map(float, coef))
    p, f = ph.pack(), fl.pack()
    _chk(lib().or_phsevo(_R(p), _R(f), _R(ogrid(g)), _R(opar(par)), _R(oscales(sc)), SCHEMES.get(scheme, 99),
                         BCS.get(xbc, 9), BCS.get(zbc, 9), c, ctypes.c_double(dt), ctypes.c_double(alpha)),
         "phsevo")


def store_previous(fl, ph):
    f, p = fl.pack(), ph.pack()
    _chk(lib().or_store_previous(_R(f), _R(p)), "store_previous")


def update_phase_velocities(ph, fl, ns, g, par, xbc="periodic", zbc="closed"):
    p, f, n = ph.pack(), fl.pack(), ns.pack()
    _chk(lib().or_update_phase_velocities(_R(p), _R(f), _R(n), _R(ogrid(g)), _R(opar(par)), BCS.get(xbc, 9),
                                          BCS.get(zbc, 9)), "update_phase_velocities")


def update_dt(ph, fl, g, par, dt_prev):
    f = lib().or_update_dt
    f.restype = ctypes.c_double
    p, fs = ph.pack(), fl.pack()
    return f(_R(p), _R(fs), _R(ogrid(g)), _R(opar(par)), ctypes.c_double(dt_prev))


def rheology_cell(chi, mu, par):
    """etamix of one cell (update.jl:144-177) via a 1x1 PhaseState."""
    ph = phase_state(1, 1)
    ph.chi[0, 0], ph.mu[0, 0] = chi, mu
    update_stage("rheology", ph, fluid_state(1, 1), None, par, None)
    return ph.etamix[0, 0]


# ---- noise (noise.jl) -------------------------------------------------------

def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return tuple(o)


def normal_fill(a, seed, stream, draw):
    _chk(lib().or_normal_fill(om(a), ctypes.c_uint64(seed), ctypes.c_uint32(stream), ctypes.c_uint32(draw)),
         "normal_fill")
    return a


def fft_filter(psi, pad, h, sigma):
    out = zeros(*psi.shape)
    _chk(lib().or_fft_filter(om(psi), om(out), ctypes.c_int64(pad), ctypes.c_double(h), ctypes.c_double(sigma)),
         "fft_filter")
    return out


def noise(ns, ph, g, par, sc, dt, first_iter=True, seed=0, draw=0):
    n, p = ns.pack(), ph.pack()
    _chk(lib().or_noise_step(_R(n), _R(p), _R(ogrid(g)), _R(opar(par)), _R(oscales(sc)), ctypes.c_double(dt),
                        int(bool(first_iter)), ctypes.c_uint64(seed), ctypes.c_uint32(draw)), "noise")


def store_noise(ns):
    n = ns.pack()
    _chk(lib().or_store_noise(_R(n)), "store_noise")


# ---- Stokes (fluidmech.jl) --------------------------------------------------

class OFmOpts(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("sds", "top_cnv", "bot_cnv", "open_cnv", "xbc", "zbc", "scheme")] + \
               [(n, ctypes.c_double) for n in ("dt", "a1", "a2", "a3", "b1", "b2", "b3", "alpha", "gamma", "g0")]


def fm_opts(par, sds=-1, top_cnv=1, bot_cnv=1, open_cnv=False, xbc="periodic", zbc="closed", dt=1e32,
            a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0, scheme=None):
    return OFmOpts(int(sds), int(top_cnv), int(bot_cnv), int(bool(open_cnv)), BCS[xbc], BCS[zbc],
                   SCHEMES[scheme or par.ADVN], float(dt), float(a1), float(a2), float(a3), float(b1), float(b2),
                   float(b3), float(par.alpha), float(par.gamma), float(par.g0))


def fluidmech_assemble(fl, ph, g, opts, bnchm=None):
    """LL (scipy CSR, duplicates summed) and RR of fluidmech.jl after the pins; mutates MFS/drhodt."""
    import scipy.sparse as sp
    Nz, Nx = g.Nz, g.Nx
    n = (Nz + 1) * (Nx + 2) + (Nz + 2) * (Nx + 1) + (Nz + 2) * (Nx + 2)
    cap = 32 * n
    I = np.zeros(cap, dtype=np.int64)
    J = np.zeros(cap, dtype=np.int64)
    V = np.zeros(cap)
    RR = np.zeros(n)
    nnz = ctypes.c_int64(0)
    f = fl.pack()
    p = ph.pack() if ph is not None else phase_state(Nz, Nx).pack()
    bn = None
    if bnchm is not None:
        arr = (OMat * 6)(*[om(bnchm[k]) for k in ("src_W", "src_U", "src_P", "W_exact", "U_exact", "P_exact")])
        bn = arr
    _chk(lib().or_fluidmech_assemble(_R(f), _R(p), _R(ogrid(g)), _R(opts), bn,
                                     I.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     J.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     V.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_int64(cap),
                                     _R(nnz), RR.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "fluidmech_assemble")
    k = nnz.value
    LL = sp.coo_matrix((V[:k], (I[:k], J[:k])), shape=(n, n)).tocsr()
    LL.sum_duplicates()
    return LL, RR


def stokes_vector(fl):
    return np.concatenate([fl.W.ravel(order="F"), fl.U.ravel(order="F"), fl.P.ravel(order="F")])


def set_stokes_vector(fl, sol):
    nw, nu = fl.W.size, fl.U.size
    fl.W[:] = sol[:nw].reshape(fl.W.shape, order="F")
    fl.U[:] = sol[nw:nw + nu].reshape(fl.U.shape, order="F")
    fl.P[:] = sol[nw + nu:].reshape(fl.P.shape, order="F")


def fluidmech(fl, ph, g, opts, bnchm=None):
    """fluidmech.jl:71-534: assemble, scale (495-507), sparse LU solve (SuperLU in place of
    UMFPACK), Newton-style update.  Returns (LL, RR) of the solved system."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    LL, RR = fluidmech_assemble(fl, ph, g, opts, bnchm)
    Nz, Nx = g.Nz, g.Nx
    nv = (Nz + 1) * (Nx + 2) + (Nz + 2) * (Nx + 1)
    diagLL = np.abs(LL.diagonal())
    scl_p = np.ones((Nz + 2, Nx + 2), order="F")
    scl_p[1:-1, 1:-1] = fl.rho / fl.eta
    extra = np.concatenate([np.zeros(nv), np.sqrt(1.0 / scl_p.ravel(order="F"))])
    scl = 1.0 / (np.sqrt(diagLL) + extra)
    SOL = stokes_vector(fl)
    FF = scl * (LL @ SOL - RR)
    S = sp.diags(scl)
    LLs = (S @ LL @ S).tocsc()
    UPD = scl * spla.splu(LLs).solve(FF)
    set_stokes_vector(fl, SOL - UPD)
    return LL, RR


# ---- time loop (run.jl:204-275) with the sparse-LU Stokes solve ---------------

def snapshot(ph, fl):
    return {k: getattr(fl, k).copy(order="F") for k in ("W", "U", "P", "MFS")} | {"X": ph.X.copy(order="F")}


def compute_resnorm(snap, ph, fl, dt):
    """diagnose.jl:57-71"""
    ep = 2.220446049250313e-16
    nrho = np.linalg.norm(fl.rho)
    rm = np.linalg.norm(ph.X - snap["X"]) / (nrho + ep) + np.linalg.norm(fl.MFS - snap["MFS"]) / (nrho / dt + ep)
    rp = sum(np.linalg.norm(getattr(fl, k) - snap[k]) / (np.linalg.norm(getattr(fl, k)) + ep) for k in ("W", "U", "P"))
    return rm + rp, rm, rp


def run_step(ph, fl, ns, g, par, sc, state, tint="bd2im", scheme="weno5", xbc="periodic", zbc="closed",
             fixed_sweeps=0, sds=-1, top_cnv=1, bot_cnv=1, open_cnv=False, seed=0):
    """One outer step of run! on the oracle; state = dict(time, dt, step, draws).  Noise draws use the same
    counter-based generator as the HIP path (seed, draw = state['draws'])."""
    state["step"] += 1
    c = time_coefs(tint, state["step"])
    dt = state["dt"]
    store_previous(fl, ph)
    store_noise(ns)
    resnorm, resnorm0, it = 1.0, 1.0, 0
    while (it < fixed_sweeps) if fixed_sweeps > 0 else (
            (resnorm / resnorm0 >= par.rtol and resnorm >= par.atol and it < par.maxit) or it < 3):
        it += 1
        snap = snapshot(ph, fl)
        phsevo(ph, fl, g, par, sc, scheme=scheme, xbc=xbc, zbc=zbc, coef=c, dt=dt)
        opts = fm_opts(par, sds=sds, top_cnv=top_cnv, bot_cnv=bot_cnv, open_cnv=open_cnv, xbc=xbc, zbc=zbc, dt=dt,
                       a1=c[0], a2=c[1], a3=c[2], b1=c[3], b2=c[4], b3=c[5], scheme=scheme)
        fluidmech(fl, ph, g, opts)
        first = it <= 1
        noise(ns, ph, g, par, sc, dt, first_iter=first, seed=seed, draw=state["draws"])
        if first:
            state["draws"] += 1
        update_phase_velocities(ph, fl, ns, g, par, xbc, zbc)
        update(ph, fl, g, par, sc, xbc, zbc)
        resnorm = compute_resnorm(snap, ph, fl, dt)[0]
        if it == 1 or resnorm > resnorm0:
            resnorm0 = resnorm + 1e-32
    state["time"] += dt
    state["dt"] = update_dt(ph, fl, g, par, dt)
    return it


# ---- initialize! (run.jl:71-140) and record_history! (history.jl:101-222) ------

def initialize(ph, fl, ns, g, par, sc, xbc="periodic", zbc="closed", x_init=None, seed=0):
    """run.jl:71-140 fresh-start ``initialize!`` (without the t = 0 history record, see record_history).

    rp is drawn from the counter-based generator (seed par.seed, stream 2, draw 0) in place of Julia's
    MersenneTwister.  Returns (time, dt, step, draws)."""
    Nz, Nx, h, D, L = g.Nz, g.Nx, g.h, g.D, g.L
    osgr = 1.0 if par.open_sgr else 0.0
    # run.jl:24-34 compute_bndtaperw!
    for j in range(Nz + 1):
        z = float(j) * h
        ph.bndtaperw[j, :] = 1.0 - (np.exp(-z / sc.l0h) + np.exp(-(D - z) / sc.l0h)) * (1.0 - osgr)
    # run.jl:47-55 compute_bndshape!
    for j in range(Nz):
        zc = (float(j + 1) - 0.5) * h
        ph.bndshape[j, :] = np.exp((-zc + h / 2.0) / sc.bnd_w)
    if x_init is None:
        rp = normal_fill(zeros(Nz, Nx), par.seed, 2, 0)
        rp = (rp - np.sum(rp) / rp.size) / np.std(rp, ddof=1)   # run.jl:99
        for j in range(Nz):                                      # run.jl:100-107
            for i in range(Nx):
                xc = (float(i + 1) - 0.5) * h / L
                zc = (float(j + 1) - 0.5) * h / D
                gp = np.exp(-((xc - 0.5) / 0.125) ** 2) * np.exp(-((zc - 0.5) / 0.125) ** 2)
                xin = par.x0 + (par.Da - par.x0) * ph.bndshape[j, i]
                ph.x[j, i] = xin * (1.0 + par.dxr * rp[j, i] + par.dxg * gp)
    else:
        ph.x[:] = x_init
    ph.m[:] = 1.0 - ph.x
    fl.eta[:] = par.etam0
    fl.etaco[:] = par.etam0
    ph.etas[:] = par.etam0
    update(ph, fl, g, par, sc, xbc, zbc)
    ph.X[:] = fl.rho * ph.x
    ph.M[:] = fl.rho * ph.m
    noise(ns, ph, g, par, sc, sc.dt0, first_iter=True, seed=seed, draw=0)
    update_phase_velocities(ph, fl, ns, g, par, xbc, zbc)
    ph.Xo[:] = ph.X; ph.Xoo[:] = ph.X
    ph.Mo[:] = ph.M; ph.Moo[:] = ph.M
    fl.rhoo[:] = fl.rho; fl.rhooo[:] = fl.rho
    store_noise(ns)
    return 0.0, sc.dt0, 0, 1


HISTORY_STATS = (("x", "std"), ("V", "rms"), ("vx", "rms"), ("vm", "rms"), ("xie", "rms"), ("xix", "rms"),
                 ("xis", "rms"), ("Ra", "gm"), ("Rc", "gm"), ("ReD", "gm"), ("Red", "gm"), ("ks", "gm"),
                 ("ke", "gm"), ("kx", "gm"), ("eta", "gm"), ("etamix", "gm"), ("etae", "gm"), ("etas", "gm"),
                 ("etat", "gm"))


def new_history():
    """history.jl:74-93 ``History(T)`` as a dict of lists plus the nine cached rates."""
    hst = {k: [] for k in ("time", "dt", "sumB", "sumM", "sumX", "EB", "EM", "EX", "DB", "DM", "DX")}
    for f, kind in HISTORY_STATS:
        for s in ("min", "mean" if kind == "std" else kind, "max") + (("std",) if kind == "std" else ()):
            hst[f"{f}_{s}"] = []
    for q in "BMX":
        for sfx in ("", "o", "oo"):
            hst[f"dsum{q}dt{sfx}"] = 0.0
    return hst


def record_history(hst, time, dt, ph, fl, ns, g, a1=1.0, a2=1.0, a3=0.0, b1=1.0, b2=0.0, b3=0.0):
    """history.jl:101-222 ``record_history!`` on the host arrays (numpy reductions)."""
    ep = 2.220446049250313e-16
    h = g.h
    h2 = h ** 2
    for q in "BMX":
        hst[f"dsum{q}dtoo"] = hst[f"dsum{q}dto"]
        hst[f"dsum{q}dto"] = hst[f"dsum{q}dt"]
    hst["time"].append(float(time))
    hst["dt"].append(float(dt))
    sumB = np.sum(fl.rho) * h2 + ep
    sumM = np.sum(ph.M) * h2 + ep
    sumX = np.sum(ph.X) * h2 + ep
    hst["sumB"].append(sumB); hst["sumM"].append(sumM); hst["sumX"].append(sumX)
    Nz1 = ph.qz_advn_X.shape[0]

    def line(q, r):
        return np.sum(q[r, 1:-1]) * h

    top_aX, bot_aX = line(ph.qz_advn_X, 0), line(ph.qz_advn_X, Nz1 - 1)
    top_aM, bot_aM = line(ph.qz_advn_M, 0), line(ph.qz_advn_M, Nz1 - 1)
    top_dX, bot_dX = line(ph.qz_dffn_X, 0), line(ph.qz_dffn_X, Nz1 - 1)
    top_dM, bot_dM = line(ph.qz_dffn_M, 0), line(ph.qz_dffn_M, Nz1 - 1)
    sumGx = np.sum(ph.Gx) * h2
    dB = (top_aX - bot_aX) + (top_aM - bot_aM) + (top_dX - bot_dX) + (top_dM - bot_dM)
    dM = -sumGx + (top_aM - bot_aM) + (top_dM - bot_dM)
    dX = +sumGx + (top_aX - bot_aX) + (top_dX - bot_dX)
    hst["dsumBdt"], hst["dsumMdt"], hst["dsumXdt"] = dB, dM, dX
    stp = len(hst["time"])
    for q, new in (("B", dB), ("M", dM), ("X", dX)):
        Dv = hst[f"D{q}"]
        if stp >= 2:
            Dv.append((a2 * Dv[stp - 2] + a3 * Dv[max(1, stp - 2) - 1] +
                       (b1 * new + b2 * hst[f"dsum{q}dto"] + b3 * hst[f"dsum{q}dtoo"]) * dt) / a1)
        else:
            Dv.append(0.0)
    sumB0 = hst["sumB"][0]
    hst["EB"].append((sumB - hst["DB"][stp - 1] - sumB0) / sumB0)
    hst["EM"].append((sumM - hst["DM"][stp - 1] - hst["sumM"][0]) / sumB0)
    hst["EX"].append((sumX - hst["DX"][stp - 1] - hst["sumX"][0]) / sumB0)
    for f, kind in HISTORY_STATS:
        a = getattr(ns, f) if f in ("xie", "xix", "xis") else (getattr(ph, f) if hasattr(ph, f) else getattr(fl, f))
        hst[f"{f}_min"].append(float(np.min(a)))
        hst[f"{f}_max"].append(float(np.max(a)))
        if kind == "std":
            hst[f"{f}_mean"].append(float(np.mean(a)))
            hst[f"{f}_std"].append(float(np.std(a, ddof=1)))
        elif kind == "rms":
            hst[f"{f}_rms"].append(float(np.sqrt(np.mean(a ** 2))))
        else:
            hst[f"{f}_gm"].append(float(np.exp(np.mean(np.log(np.maximum(a, ep))))))
    return hst


# ---- compute_scales (scales.jl:85-185) ----------------------------------------

def compute_scales(par, grid, tol=1e-9, maxiter=200):
    """scales.jl:85-185 ``compute_scales`` restated in Python (the checker of the library's xc_compute_scales);
    returns a dict of every Scales field."""
    D, d0, L0, l0 = par.D, par.d0, par.L0, par.l0
    x0, Da, rhom0, rhox0 = par.x0, par.Da, par.rhom0, par.rhox0
    etam0, g0, Xi, h = par.etam0, par.g0, par.Xi, grid.h
    h0 = h
    D0 = D / 10
    L0h = (L0 + h0) / 2
    l0h = (l0 + h0) / 2
    bnd_w = l0 / 2 + D / 100
    rho0 = rhom0
    Drho0 = rhox0 - rhom0
    chi0, Dchi0 = (Da, Da / 10) if Da > 0 else (x0, x0 / 10)
    eta0 = etam0
    W0l = Dchi0 * Drho0 * g0 * D0 ** 2 / eta0
    w0l = Drho0 * g0 * d0 ** 2 / eta0
    w0t = math.sqrt(Drho0 * g0 * d0 ** 2 / (l0 * rho0))
    W0t = math.sqrt(Dchi0 * Drho0 * g0 * D0 ** 3 / (L0 ** 2 * rho0))
    W0i = math.sqrt(Dchi0 * Drho0 * g0 * D / rho0)
    Ri0 = W0i / W0t
    W0, w0 = W0l, w0l
    ReL0 = W0 * L0 / (eta0 / rho0)
    Rel0 = w0 * l0 / (eta0 / rho0)
    fReL0 = 1 - math.exp(-ReL0)
    fRel0 = 1 - math.exp(-Rel0)
    for _ in range(maxiter):
        W0prv, w0prv = W0, w0
        ReL0 = W0 * L0 / (eta0 / rho0)
        Rel0 = w0 * l0 / (eta0 / rho0)
        fReL0 = 1 - math.exp(-ReL0)
        fRel0 = 1 - math.exp(-Rel0)
        W0 = ((math.sqrt(4 / Ri0 ** 2 * Dchi0 * Drho0 * g0 * rho0 * fReL0 * L0 ** 2 * D0 + eta0 ** 2) - eta0)
              * D0 / (2 * fReL0 * L0 ** 2 * rho0 / Ri0 ** 2))
        w0 = ((math.sqrt(4 * Drho0 * g0 * rho0 * fRel0 * l0 * d0 ** 2 + eta0 ** 2) - eta0)
              / (2 * fRel0 * l0 * rho0))
        res = abs(W0 - W0prv) / W0 + abs(w0 - w0prv) / w0
        if res <= tol:
            break
    eII0 = W0 / D0
    ke0 = eII0 * L0 ** 2
    ks0 = w0 * l0
    kx0 = ks0 + fReL0 * ke0
    tW0 = D / W0
    tw0 = D / w0
    tk0 = D ** 2 / kx0
    ti0 = D / W0i
    txi0 = rhox0 * d0 ** 2 / 18 / eta0
    t0 = 1 / (1 / (ti0 + tW0) + 1 / tw0 + 1 / tk0)
    dt0 = min((h0 / 2) ** 2 / kx0, (h0 / 2) / (W0 + w0))
    taue0 = L0 / 2 / W0
    taus0 = l0 / 2 / w0
    St0 = txi0 / taue0
    xie0 = Xi * math.sqrt(fReL0 * ke0 / taue0)
    xix0 = Xi * math.sqrt(chi0 * fReL0 * ke0 / taue0 * St0 / (1 + St0 ** 2))
    xis0 = Xi * math.sqrt(chi0 * ks0 / taus0)
    G0 = Da * rho0 / t0 * D / D0
    etae0 = fReL0 * ke0 * rho0
    etat0 = fRel0 * ks0 * rho0
    p0 = (eta0 + etae0) * eII0
    vals = dict(h0=h0, D0=D0, d0=d0, L0=L0, L0h=L0h, l0=l0, l0h=l0h, bnd_w=bnd_w, rho0=rho0, Drho0=Drho0,
                chi0=chi0, Dchi0=Dchi0, eta0=eta0, W0=W0, w0=w0, W0l=W0l, w0l=w0l, W0t=W0t, w0t=w0t, W0i=W0i,
                Ri0=Ri0, ReL0=ReL0, Rel0=Rel0, fReL0=fReL0, fRel0=fRel0, eII0=eII0, ke0=ke0, ks0=ks0, kx0=kx0,
                tW0=tW0, tw0=tw0, tk0=tk0, ti0=ti0, txi0=txi0, t0=t0, dt0=dt0, taue0=taue0, taus0=taus0,
                St0=St0, xie0=xie0, xix0=xix0, xis0=xis0, G0=G0, etae0=etae0, etat0=etat0, p0=p0,
                Noe0=xie0 / W0, Nox0=xix0 / w0, Nos0=xis0 / w0, Rc0=W0 / w0, Ra0=W0 * D0 / kx0,
                ReD0=W0 * D0 / ((eta0 + etae0) / rho0), Red0=w0 * d0 / ((eta0 + etat0) / rho0))
    return vals
