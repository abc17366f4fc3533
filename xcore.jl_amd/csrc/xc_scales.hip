// xc_scales.hip -- compute_scales (src/scales.jl:85-185) on the host, for one parameter set or for every
// member of an ensemble at once (each member's density contrast and viscosity change its W0, t0, dt0, ...).
//
// Host code only.  The fixed point on (W0, w0) and every derived scale follow scales.jl's operation order
// without contraction into fused multiply-adds, so the result is the reference's to the last bit up to the
// libm's exp/sqrt.
#include <math.h>

#include "xc_internal.h"

#pragma clang fp contract(off)

namespace {

inline double sq(double v) { return v * v; }

// scales.jl:88-185 _compute_scales_inner
void scales_inner(const xc_scales_in& p, double tol, int maxiter, xc_scales_full* s) {
    const double D = p.D, d0 = p.d0, h = p.h, x0 = p.x0, Da = p.Da, rhom0 = p.rhom0, rhox0 = p.rhox0,
                 etam0 = p.etam0, g0 = p.g0, Xi = p.Xi;
    const double h0 = h;
    const double D0 = D / 10;
    const double L0 = p.L0, L0h = (L0 + h0) / 2;
    const double l0 = p.l0, l0h = (l0 + h0) / 2;
    const double bnd_w = l0 / 2 + D / 100;
    const double rho0 = rhom0;
    const double Drho0 = rhox0 - rhom0;
    double chi0, Dchi0;
    if (Da > 0) {
        chi0 = Da;
        Dchi0 = Da / 10;
    } else {
        chi0 = x0;
        Dchi0 = x0 / 10;
    }
    const double eta0 = etam0;
    const double W0l = Dchi0 * Drho0 * g0 * sq(D0) / eta0;
    const double w0l = Drho0 * g0 * sq(d0) / eta0;
    const double w0t = sqrt(Drho0 * g0 * sq(d0) / (l0 * rho0));
    const double W0t = sqrt(Dchi0 * Drho0 * g0 * (D0 * D0 * D0) / (sq(L0) * rho0));
    const double W0i = sqrt(Dchi0 * Drho0 * g0 * D / rho0);
    const double Ri0 = W0i / W0t;
    double W0 = W0l, w0 = w0l;
    double ReL0 = W0 * L0 / (eta0 / rho0);
    double Rel0 = w0 * l0 / (eta0 / rho0);
    double fReL0 = 1 - exp(-ReL0);
    double fRel0 = 1 - exp(-Rel0);
    for (int it = 0; it < maxiter; ++it) {
        const double W0prv = W0, w0prv = w0;
        ReL0 = W0 * L0 / (eta0 / rho0);
        Rel0 = w0 * l0 / (eta0 / rho0);
        fReL0 = 1 - exp(-ReL0);
        fRel0 = 1 - exp(-Rel0);
        W0 = (sqrt(4 / sq(Ri0) * Dchi0 * Drho0 * g0 * rho0 * fReL0 * sq(L0) * D0 + sq(eta0)) - eta0) * D0 /
             (2 * fReL0 * sq(L0) * rho0 / sq(Ri0));
        w0 = (sqrt(4 * Drho0 * g0 * rho0 * fRel0 * l0 * sq(d0) + sq(eta0)) - eta0) / (2 * fRel0 * l0 * rho0);
        const double res = fabs(W0 - W0prv) / W0 + fabs(w0 - w0prv) / w0;
        if (res <= tol) break;
    }
    const double eII0 = W0 / D0;
    const double ke0 = eII0 * sq(L0);
    const double ks0 = w0 * l0;
    const double kx0 = ks0 + fReL0 * ke0;
    const double tW0 = D / W0;
    const double tw0 = D / w0;
    const double tk0 = sq(D) / kx0;
    const double ti0 = D / W0i;
    const double txi0 = rhox0 * sq(d0) / 18 / eta0;
    const double t0 = 1.0 / (1 / (ti0 + tW0) + 1 / tw0 + 1 / tk0);
    const double dta = sq(h0 / 2) / kx0, dtb = (h0 / 2) / (W0 + w0);
    const double dt0 = dtb < dta ? dtb : dta;
    const double taue0 = L0 / 2 / W0;
    const double taus0 = l0 / 2 / w0;
    const double St0 = txi0 / taue0;
    const double xie0 = Xi * sqrt(fReL0 * ke0 / taue0);
    const double xix0 = Xi * sqrt(chi0 * fReL0 * ke0 / taue0 * St0 / (1 + sq(St0)));
    const double xis0 = Xi * sqrt(chi0 * ks0 / taus0);
    const double G0 = Da * rho0 / t0 * D / D0;
    const double etae0 = fReL0 * ke0 * rho0;
    const double etat0 = fRel0 * ks0 * rho0;
    const double p0 = (eta0 + etae0) * eII0;
    *s = xc_scales_full{h0,    D0,    d0,    L0,    L0h,   l0,    l0h,   bnd_w, rho0,  Drho0, chi0,
                        Dchi0, eta0,  W0,    w0,    W0l,   w0l,   W0t,   w0t,   W0i,   Ri0,   ReL0,
                        Rel0,  fReL0, fRel0, eII0,  ke0,   ks0,   kx0,   tW0,   tw0,   tk0,   ti0,
                        txi0,  t0,    dt0,   taue0, taus0, St0,   xie0,  xix0,  xis0,  G0,    etae0,
                        etat0, p0,    xie0 / W0, xix0 / w0, xis0 / w0, W0 / w0, W0 * D0 / kx0,
                        W0 * D0 / ((eta0 + etae0) / rho0), w0 * d0 / ((eta0 + etat0) / rho0)};
}

}  // namespace

extern "C" int xc_compute_scales(int n, const xc_scales_in* in, double tol, int maxiter, xc_scales_full* out) {
    if (n < 0 || (n > 0 && (!in || !out)) || maxiter < 0 || !(tol >= 0))
        return xc_fail(nullptr, XC_EINVAL_ARG, "compute_scales: n %d, tol %g, maxiter %d", n, tol, maxiter);
    for (int i = 0; i < n; ++i) {
        const xc_scales_in& p = in[i];
        if (!(p.h > 0) || !(p.D > 0) || !(p.L0 > 0) || !(p.l0 > 0) || !(p.etam0 > 0) || !(p.rhom0 > 0))
            return xc_fail(nullptr, XC_EINVAL_ARG, "compute_scales: member %d: h, D, L0, l0, etam0, rhom0 must be > 0",
                           i);
        scales_inner(p, tol, maxiter, &out[i]);
    }
    return XC_OK;
}
