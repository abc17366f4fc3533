"""A manufactured solution for fluidmech!'s production mode (no benchmark data): every term the MMS of
test/mms_sources.jl leaves out is active -- the inertia a1 rho W / dt and (a1 + gamma) rho U / dt, the
"drunken sailor" terms d(rho)/dz g0 dt W and d(rho)/dx g0 dt U (fluidmech.jl:218-222, 285-289), the momentum
advection of rho W and rho U (fluidmech.jl:236-241, 310-318), the buoyancy Drho g0, the previous-momentum
terms (a2 rho Wo + a3 rho Woo) / dt, the closed-bottom row rho W / h = mean(MFS) D / h (fluidmech.jl:170-191)
and the production pressure pin with its divergence row kept (fluidmech.jl:480-487).

Production mode takes no source terms, so the manufactured forcing enters through the inputs the reference
reads: rhoWo, rhoWoo, rhoUo, rhoUoo carry what the continuous equations need beyond the operator, Drho a
given buoyancy, MFS the continuous divergence of rho v.  The state's W, U are the exact fields, so the
momentum advection (explicit, from the current velocities) approximates the exact advection.  The MFS update
(fluidmech.jl:111-127) leaves MFS unchanged: rhoo = rhooo = rho, advn_rho = 0, old drhodt = 0 and
a1 = a2 + a3.  Unit box, periodic x, free-slip walls (W = 0, dU/dz = 0, dP/dz = 0 at z = 0, 1).

The continuous operator is the reference's (test/mms_sources.jl:60-120): tau_xx = eta (dU/dx - div/3),
tau_zz = eta (dW/dz - div/3), tau_xz = eta (dU/dz + dW/dx) / 2, W rows -(d tau_zz/dz + d tau_xz/dx) + dP/dz,
U rows -(d tau_xx/dx + d tau_xz/dz) + dP/dx; z grows with the row index (depth).
"""
import functools

import numpy as np

AW, AU, AP = 1.0, 0.8, 5.0
ETA0, RHO0 = 1.0, 1.0
G0 = 400.0
DT = 0.02
GAMMA = 0.3
COEF = (1.5, 2.0, -0.5, 1.0, 0.0, 0.0)   # BD2 (timing.jl:18-36): a1 = a2 + a3


@functools.lru_cache(maxsize=None)
def _symbolic():
    import sympy as sp
    x, z, cpin = sp.symbols("x z cpin", real=True)
    kx, kz = 2 * sp.pi, sp.pi
    a1, a2, a3 = COEF[:3]
    W = AW * sp.cos(kx * x) * sp.sin(kz * z)
    U = AU * sp.sin(kx * x) * sp.cos(kz * z)
    P = AP * (sp.cos(kx * x) * sp.cos(kz * z) - cpin)
    eta = ETA0 * sp.exp(sp.Rational(1, 2) * sp.sin(kx * x) * sp.cos(kz * z))
    rho = RHO0 * (1 + sp.Rational(1, 20) * sp.cos(kx * x + sp.Rational(3, 10)) * sp.cos(kz * z))
    drho = sp.Rational(3, 100) * sp.sin(kx * x) * sp.sin(kz * z)
    div = sp.diff(U, x) + sp.diff(W, z)
    txx = eta * (sp.diff(U, x) - div / 3)
    tzz = eta * (sp.diff(W, z) - div / 3)
    txz = eta * (sp.diff(U, z) + sp.diff(W, x)) / 2
    lhs_w = (a1 * rho / DT + sp.diff(rho, z) * G0 * DT) * W - (sp.diff(tzz, z) + sp.diff(txz, x)) + sp.diff(P, z)
    lhs_u = ((a1 + GAMMA) * rho / DT + sp.diff(rho, x) * G0 * DT) * U - (sp.diff(txx, x) + sp.diff(txz, z)) \
        + sp.diff(P, x)
    adv_w = sp.diff(rho * W * U, x) + sp.diff(rho * W * W, z)
    adv_u = sp.diff(rho * U * U, x) + sp.diff(rho * U * W, z)
    # (a2 rho_o + a3 rho_oo) / dt = lhs - Drho g0 + adv (W rows) and lhs + adv (U rows)
    qw = lhs_w - drho * G0 + adv_w
    qu = lhs_u + adv_u
    mfs = sp.diff(rho * U, x) + sp.diff(rho * W, z)
    f = lambda e: sp.lambdify((x, z, cpin), e, "numpy")
    return dict(W=f(W), U=f(U), P=f(P), eta=f(eta), rho=f(rho), drho=f(drho), qw=f(qw), qu=f(qu), mfs=f(mfs))


def _ev(fn, zs, xs, cpin=0.0):
    v = fn(xs[None, :], zs[:, None], cpin)
    return np.asfortranarray(np.broadcast_to(v, (len(zs), len(xs))).astype(np.float64))


def inertial_case(N):
    """Host arrays of a FluidState, the exact (W, U, P) and the fm_opts keywords for an N x N unit box."""
    S = _symbolic()
    h = 1.0 / N
    a1, a2, a3 = COEF[:3]
    zc = (np.arange(N) + 0.5) * h           # cell centres
    xc = (np.arange(N) + 0.5) * h
    zf = np.arange(N + 1) * h               # z faces / corners
    xf = np.arange(N + 1) * h
    zp = (np.arange(N + 2) - 0.5) * h       # P / U rows with the ghost ring
    xp = (np.arange(N + 2) - 0.5) * h
    # the pin: P(np0) of the production system is exactly zero, so P* vanishes there
    # (np0 = MapP[round(Nz/2), round(Nx/2)], 1-based; fluidmech.jl:480-487)
    iz, ix = round(N / 2) - 1, round(N / 2) - 1
    cpin = float(np.cos(2 * np.pi * xp[ix]) * np.cos(np.pi * zp[iz]))
    out = {}
    out["W_exact"] = _ev(S["W"], zf, xp)
    out["U_exact"] = _ev(S["U"], zp, xf)
    out["P_exact"] = _ev(S["P"], zp, xp, cpin)
    out["eta"] = _ev(S["eta"], zc, xc)
    out["etaco"] = _ev(S["eta"], zf, xf)
    out["rho"] = _ev(S["rho"], zc, xc)
    out["rhow"] = _ev(S["rho"], zf, xc)
    out["rhou"] = _ev(S["rho"], zc, xf)
    out["Drho"] = _ev(S["drho"], zf, xc)
    qw = _ev(S["qw"], zf, xc) * DT           # a2 rhoWo + a3 rhoWoo = qw dt, split so both terms count
    qu = _ev(S["qu"], zc, xf) * DT
    out["rhoWoo"] = np.asfortranarray(0.5 * qw)
    out["rhoWo"] = np.asfortranarray((qw - a3 * out["rhoWoo"]) / a2)
    out["rhoUoo"] = np.asfortranarray(0.5 * qu)
    out["rhoUo"] = np.asfortranarray((qu - a3 * out["rhoUoo"]) / a2)
    out["MFS"] = _ev(S["mfs"], zc, xc)
    out["h"] = h
    kw = dict(dt=DT, a1=COEF[0], a2=COEF[1], a3=COEF[2], b1=COEF[3], b2=COEF[4], b3=COEF[5], sds=-1, top_cnv=-1,
              bot_cnv=-1, open_cnv=False)
    return out, kw


def load_fluid(fl, case):
    """Fill a (host or device) FluidState's arrays from inertial_case's dict (W, U start at the exact fields)."""
    for k in ("eta", "etaco", "rho", "rhow", "rhou", "Drho", "rhoWo", "rhoWoo", "rhoUo", "rhoUoo", "MFS"):
        _set(getattr(fl, k), case[k])
    for k in ("rhoo", "rhooo"):
        _set(getattr(fl, k), case["rho"])
    _set(fl.W, case["W_exact"])
    _set(fl.U, case["U_exact"])
    for k in ("drhodt", "drhodto", "drhodtoo"):
        _set(getattr(fl, k), np.zeros_like(case["rho"]))


def _set(dst, src):
    if isinstance(dst, np.ndarray):
        dst[:] = src
    else:
        import torch
        dst.copy_(torch.as_tensor(np.ascontiguousarray(src)))
