"""Manufactured Stokes solution of the reference's test helper (test/mms_sources.jl:36-145),
restated in numpy: analytic W/U/P, the Stokes residual source terms and the material fields,
on the staggered grids of a FluidState."""
import numpy as np


def mms_sources(Nz, Nx, L):
    assert Nz == Nx
    h = L / Nz
    aW, aU, aP = 5e-5, 4e-5, -3e3
    eta0, aeta = 1e3, -9e2
    rho0, arho = 3e3, -5e1
    asrc, rhoref, g0 = -1e3, 3e3, 10.0
    k = 4 * np.pi / L
    x_mms = -h / 2 + np.arange(Nx + 2) * h
    z_mms = -h / 2 + np.arange(Nz + 2) * h
    xu = (x_mms[:-1] + x_mms[1:]) / 2
    zw = (z_mms[:-1] + z_mms[1:]) / 2
    C = lambda x, z: np.cos(k * x) * np.sin(k * z)
    S = lambda x, z: np.sin(k * x) * np.cos(k * z)
    CC = lambda x, z: np.cos(k * x) * np.cos(k * z)
    SS = lambda x, z: np.sin(k * x) * np.sin(k * z)
    W_ = lambda x, z: aW * C(x, z)
    U_ = lambda x, z: aU * S(x, z)
    P_ = lambda x, z: aP * CC(x, z)
    eta_ = lambda x, z: eta0 + aeta * C(x, z)
    rho_ = lambda x, z: rho0 + arho * C(x, z)
    src_ = lambda x, z: asrc * C(x, z)
    a_exx = (2 * aU - aW) / 3
    a_ezz = (2 * aW - aU) / 3
    a_exz = -(aW + aU) / 2
    exx = lambda x, z: k * CC(x, z) * a_exx
    ezz = lambda x, z: k * CC(x, z) * a_ezz
    exz = lambda x, z: k * SS(x, z) * a_exz
    eta_x = lambda x, z: -aeta * k * SS(x, z)
    eta_z = lambda x, z: aeta * k * CC(x, z)
    exx_x = lambda x, z: -k ** 2 * S(x, z) * a_exx
    exx_z = lambda x, z: -k ** 2 * C(x, z) * a_exx
    ezz_x = lambda x, z: -k ** 2 * S(x, z) * a_ezz
    ezz_z = lambda x, z: -k ** 2 * C(x, z) * a_ezz
    exz_x = lambda x, z: k ** 2 * C(x, z) * a_exz
    exz_z = lambda x, z: k ** 2 * S(x, z) * a_exz
    dx_txx = lambda x, z: eta_x(x, z) * exx(x, z) + eta_(x, z) * exx_x(x, z)
    dz_tzz = lambda x, z: eta_z(x, z) * ezz(x, z) + eta_(x, z) * ezz_z(x, z)
    dx_txz = lambda x, z: eta_x(x, z) * exz(x, z) + eta_(x, z) * exz_x(x, z)
    dz_txz = lambda x, z: eta_z(x, z) * exz(x, z) + eta_(x, z) * exz_z(x, z)
    Px = lambda x, z: -aP * k * S(x, z)
    Pz = lambda x, z: -aP * k * C(x, z)
    res_W = lambda x, z: -(dz_tzz(x, z) + dx_txz(x, z)) + Pz(x, z) - (rho_(x, z) - rhoref) * g0
    res_U = lambda x, z: -(dx_txx(x, z) + dz_txz(x, z)) + Px(x, z)
    rho_z = lambda x, z: arho * k * CC(x, z)
    rho_x = lambda x, z: -arho * k * SS(x, z)
    W_z = lambda x, z: aW * k * CC(x, z)
    U_x = lambda x, z: aU * k * CC(x, z)
    res_P = lambda x, z: (rho_z(x, z) * W_(x, z) + rho_(x, z) * W_z(x, z) + rho_x(x, z) * U_(x, z)
                          + rho_(x, z) * U_x(x, z) - src_(x, z))

    def ev(f, zs, xs):
        return np.asfortranarray(f(xs[None, :], zs[:, None]))

    out = dict(W_exact=ev(W_, zw, x_mms), U_exact=ev(U_, z_mms, xu), P_exact=ev(P_, z_mms, x_mms),
               src_W=ev(res_W, zw, x_mms), src_U=ev(res_U, z_mms, xu), src_P=ev(res_P, z_mms, x_mms))
    out["eta"] = np.asfortranarray(ev(eta_, z_mms, x_mms)[1:-1, 1:-1])
    out["etaco"] = ev(eta_, zw, xu)
    out["rho"] = np.asfortranarray(ev(rho_, z_mms, x_mms)[1:-1, 1:-1])
    out["rhow"] = np.asfortranarray(ev(rho_, zw, x_mms)[:, 1:-1])
    out["rhou"] = np.asfortranarray(ev(rho_, z_mms, xu)[1:-1, :])
    out["MFS"] = np.asfortranarray(ev(src_, z_mms, x_mms)[1:-1, 1:-1])
    out["Drho"] = np.asfortranarray(out["rhow"] - out["rhow"].mean(axis=1, keepdims=True))
    out["h"] = h
    return out
