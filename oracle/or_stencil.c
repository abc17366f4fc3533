/*
 * or_stencil.c -- oracle (test infrastructure only): restatement of the
 * advection and diffusion stencils of xcore.jl, src/kernels/{advect,diffus,
 * ghosts}.jl and the halo-allocating wrappers of src/phsevo.jl:21-149.
 *
 * Structure deliberately follows the reference: a divergence kernel that
 * reconstructs all 8 face states per cell, separate per-face flux kernels,
 * then the MATLAB-style boundary padding of the flux arrays.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define AT(m, i, j) ((m).p[(int64_t)(i) + (int64_t)(j) * (m).ld])

static int g_threads = 1;
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int  or_get_threads(void) { return g_threads; }

/* Julia max/min on non-NaN doubles: -0.0 < +0.0 */
static inline double jl_max(double a, double b) {
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}
static inline double jl_min(double a, double b) {
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}

/* advect.jl:346-355 */
int or_scheme_halo(int scheme) {
    switch (scheme) {
        case OR_CENTR: return 1;
        case OR_UPWD1: return 1;
        case OR_QUICK: return 2;
        case OR_FROMM: return 2;
        case OR_WENO3: return 2;
        case OR_WENO5: return 3;
        case OR_TVDIM: return 2;
        default: return -1;
    }
}

/* ghosts.jl:18-51 -- x first (periodic wrap or repeat), then z. */
int or_fill_ghosts(omat f, int halo, int xbc, int zbc) {
    int64_t Nzt = f.rows, Nxt = f.cols;
    int64_t Nz = Nzt - 2 * halo, Nx = Nxt - 2 * halo;
    if (Nz < 1 || Nx < 1) return OR_ESHAPE;
    if (xbc != OR_PERIODIC && xbc != OR_CLOSED) return OR_EBC;
    if (zbc != OR_PERIODIC && zbc != OR_CLOSED) return OR_EBC;
    double* tmp = (double*)malloc(sizeof(double) * (size_t)(Nzt > Nxt ? Nzt : Nxt) * (size_t)halo);
    if (xbc == OR_PERIODIC) {
        /* f[:, 1:halo] .= f[:, Nx+1:Nx+halo]; broadcast unaliases its source */
        for (int k = 0; k < halo; ++k)
            for (int64_t j = 0; j < Nzt; ++j) tmp[k * Nzt + j] = AT(f, j, Nx + k);
        for (int k = 0; k < halo; ++k)
            for (int64_t j = 0; j < Nzt; ++j) AT(f, j, k) = tmp[k * Nzt + j];
        /* f[:, Nx+halo+1:end] .= f[:, halo+1:2halo] */
        for (int k = 0; k < halo; ++k)
            for (int64_t j = 0; j < Nzt; ++j) tmp[k * Nzt + j] = AT(f, j, halo + k);
        for (int k = 0; k < halo; ++k)
            for (int64_t j = 0; j < Nzt; ++j) AT(f, j, Nx + halo + k) = tmp[k * Nzt + j];
    } else {
        for (int k = 0; k < halo; ++k)
            for (int64_t j = 0; j < Nzt; ++j) {
                AT(f, j, k) = AT(f, j, halo);
                AT(f, j, Nx + halo + k) = AT(f, j, Nx + halo - 1);
            }
    }
    if (zbc == OR_PERIODIC) {
        for (int k = 0; k < halo; ++k)
            for (int64_t i = 0; i < Nxt; ++i) tmp[k * Nxt + i] = AT(f, Nz + k, i);
        for (int k = 0; k < halo; ++k)
            for (int64_t i = 0; i < Nxt; ++i) AT(f, k, i) = tmp[k * Nxt + i];
        for (int k = 0; k < halo; ++k)
            for (int64_t i = 0; i < Nxt; ++i) tmp[k * Nxt + i] = AT(f, halo + k, i);
        for (int k = 0; k < halo; ++k)
            for (int64_t i = 0; i < Nxt; ++i) AT(f, Nz + halo + k, i) = tmp[k * Nxt + i];
    } else {
        for (int k = 0; k < halo; ++k)
            for (int64_t i = 0; i < Nxt; ++i) {
                AT(f, k, i) = AT(f, halo, i);
                AT(f, Nz + halo + k, i) = AT(f, Nz + halo - 1, i);
            }
    }
    free(tmp);
    return OR_OK;
}

/* ghosts.jl:61-64 */
int or_embed_interior(omat fh, omat f, int halo) {
    if (fh.rows != f.rows + 2 * halo || fh.cols != f.cols + 2 * halo) return OR_ESHAPE;
    for (int64_t i = 0; i < f.cols; ++i)
        for (int64_t j = 0; j < f.rows; ++j) AT(fh, j + halo, i + halo) = AT(f, j, i);
    return OR_OK;
}

/* ---- face reconstructions (advect.jl) ------------------------------------ */

/* advect.jl:35-53 */
static inline double weno5_poly(double fmm, double fm, double fc, double fp, double fpp, double e) {
    const double sixth = 1.0 / 6.0, quarter = 0.25, tt = 13.0 / 12.0;
    double p1 = sixth * (2.0 * fmm - 7.0 * fm + 11.0 * fc);
    double p2 = sixth * (-fm + 5.0 * fc + 2.0 * fp);
    double p3 = sixth * (2.0 * fc + 5.0 * fp - fpp);
    double s1 = fmm - 2.0 * fm + fc, t1 = fmm - 4.0 * fm + 3.0 * fc;
    double s2 = fm - 2.0 * fc + fp,  t2 = fm - fp;
    double s3 = fc - 2.0 * fp + fpp, t3 = 3.0 * fc - 4.0 * fp + fpp;
    double b1 = tt * (s1 * s1) + quarter * (t1 * t1);
    double b2 = tt * (s2 * s2) + quarter * (t2 * t2);
    double b3 = tt * (s3 * s3) + quarter * (t3 * t3);
    double a1 = b1 + e, a2 = b2 + e, a3 = b3 + e;
    double w1 = 0.1 / (a1 * a1);
    double w2 = 0.6 / (a2 * a2);
    double w3 = 0.3 / (a3 * a3);
    return (w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3);
}

/* advect.jl:216-229 */
static inline double weno3_poly(double fm, double fc, double fp, double e) {
    double p1 = 0.5 * (fc + fp);
    double p2 = 0.5 * (3.0 * fc - fm);
    double b1 = (fp - fc) * (fp - fc);
    double b2 = (fc - fm) * (fc - fm);
    double w1 = (1.0 / 3.0) / (b1 + e);
    double w2 = (2.0 / 3.0) / (b2 + e);
    return (w1 * p1 + w2 * p2) / (w1 + w2);
}

/* advect.jl:275-283 */
static inline double tvd_flux(double fm, double fc, double fp, double vm, double vp) {
    const double eps = 2.220446049250313e-16;
    double num = vm * (fc - fm);
    double den = vp * (fp - fc);
    double R = num / (den + copysign(eps, den));
    double l = jl_max(0.0, jl_max(jl_min(1.0, 2.0 * R), jl_min(2.0, R)));
    return fc + 0.5 * l * (fp - fc);
}

static inline double vpos(double v) { return (v + fabs(v)) * 0.5; }
static inline double vneg(double v) { return (v - fabs(v)) * 0.5; }

/* ---- divergence kernels (advect.jl:11-337) ------------------------------- */

static void advect_div(omat adv, omat f, omat u, omat w, double h, int scheme, int g) {
    const int64_t Nz = adv.rows, Nx = adv.cols;
    const double invh = 1.0 / h, half = 0.5, sixth = 1.0 / 6.0, quarter = 0.25;
    const double eps5 = 2.220446049250313e-16, eps3 = 1e-6;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t ix = 0; ix < Nx; ++ix) {
        for (int64_t iz = 0; iz < Nz; ++iz) {
            int64_t j = iz + g, i = ix + g;
            double fc = AT(f, j, i);
            double qxr, qxl, qzb, qzt;
            if (scheme == OR_CENTR) {  /* advect.jl:11-31 */
                double fmx = AT(f, j, i - 1), fpx = AT(f, j, i + 1);
                double fmz = AT(f, j - 1, i), fpz = AT(f, j + 1, i);
                double ul = AT(u, iz, ix), ur = AT(u, iz, ix + 1);
                double wt = AT(w, iz, ix), wb = AT(w, iz + 1, ix);
                qxr = ur * (fc + fpx) * half;
                qxl = ul * (fc + fmx) * half;
                qzb = wb * (fc + fpz) * half;
                qzt = wt * (fc + fmz) * half;
            } else if (scheme == OR_TVDIM) {  /* advect.jl:285-337 */
                double fxmm = AT(f, j, i - 2), fxm = AT(f, j, i - 1);
                double fxp = AT(f, j, i + 1), fxpp = AT(f, j, i + 2);
                double fzmm = AT(f, j - 2, i), fzm = AT(f, j - 1, i);
                double fzp = AT(f, j + 1, i), fzpp = AT(f, j + 2, i);
                double ull = AT(u, iz, ix), ul = AT(u, iz, ix + 1);
                double ur = AT(u, iz, ix + 2), urr = AT(u, iz, ix + 3);
                double wtt = AT(w, iz, ix), wt = AT(w, iz + 1, ix);
                double wb = AT(w, iz + 2, ix), wbb = AT(w, iz + 3, ix);
                double fxppos = tvd_flux(fxm, fc, fxp, vpos(ul), vpos(ur));
                double fxpneg = tvd_flux(fxpp, fxp, fc, vneg(urr), vneg(ur));
                double fxmpos = tvd_flux(fxmm, fxm, fc, vpos(ull), vpos(ul));
                double fxmneg = tvd_flux(fxp, fc, fxm, vneg(ur), vneg(ul));
                double fzppos = tvd_flux(fzm, fc, fzp, vpos(wt), vpos(wb));
                double fzpneg = tvd_flux(fzpp, fzp, fc, vneg(wbb), vneg(wb));
                double fzmpos = tvd_flux(fzmm, fzm, fc, vpos(wtt), vpos(wt));
                double fzmneg = tvd_flux(fzp, fc, fzm, vneg(wb), vneg(wt));
                qxr = vpos(ur) * fxppos + vneg(ur) * fxpneg;
                qxl = vpos(ul) * fxmpos + vneg(ul) * fxmneg;
                qzb = vpos(wb) * fzppos + vneg(wb) * fzpneg;
                qzt = vpos(wt) * fzmpos + vneg(wt) * fzmneg;
            } else {
                double fxppos, fxpneg, fxmpos, fxmneg, fzppos, fzpneg, fzmpos, fzmneg;
                if (scheme == OR_WENO5) {  /* advect.jl:57-96 */
                    double fxmmm = AT(f, j, i - 3), fxmm = AT(f, j, i - 2), fxm = AT(f, j, i - 1);
                    double fxp = AT(f, j, i + 1), fxpp = AT(f, j, i + 2), fxppp = AT(f, j, i + 3);
                    fxppos = weno5_poly(fxmm, fxm, fc, fxp, fxpp, eps5);
                    fxpneg = weno5_poly(fxppp, fxpp, fxp, fc, fxm, eps5);
                    fxmpos = weno5_poly(fxmmm, fxmm, fxm, fc, fxp, eps5);
                    fxmneg = weno5_poly(fxpp, fxp, fc, fxm, fxmm, eps5);
                    double fzmmm = AT(f, j - 3, i), fzmm = AT(f, j - 2, i), fzm = AT(f, j - 1, i);
                    double fzp = AT(f, j + 1, i), fzpp = AT(f, j + 2, i), fzppp = AT(f, j + 3, i);
                    fzppos = weno5_poly(fzmm, fzm, fc, fzp, fzpp, eps5);
                    fzpneg = weno5_poly(fzppp, fzpp, fzp, fc, fzm, eps5);
                    fzmpos = weno5_poly(fzmmm, fzmm, fzm, fc, fzp, eps5);
                    fzmneg = weno5_poly(fzpp, fzp, fc, fzm, fzmm, eps5);
                } else if (scheme == OR_UPWD1) {  /* advect.jl:100-126 */
                    double fmx = AT(f, j, i - 1), fpx = AT(f, j, i + 1);
                    double fmz = AT(f, j - 1, i), fpz = AT(f, j + 1, i);
                    fxppos = fc; fxpneg = fpx; fxmpos = fmx; fxmneg = fc;
                    fzppos = fc; fzpneg = fpz; fzmpos = fmz; fzmneg = fc;
                } else {
                    double fxmm = AT(f, j, i - 2), fxm = AT(f, j, i - 1);
                    double fxp = AT(f, j, i + 1), fxpp = AT(f, j, i + 2);
                    double fzmm = AT(f, j - 2, i), fzm = AT(f, j - 1, i);
                    double fzp = AT(f, j + 1, i), fzpp = AT(f, j + 2, i);
                    if (scheme == OR_QUICK) {  /* advect.jl:130-170 */
                        fxppos = sixth * (2 * fxp + 5 * fc - fxm);
                        fxpneg = sixth * (2 * fc + 5 * fxp - fxpp);
                        fxmpos = sixth * (2 * fc + 5 * fxm - fxmm);
                        fxmneg = sixth * (2 * fxm + 5 * fc - fxp);
                        fzppos = sixth * (2 * fzp + 5 * fc - fzm);
                        fzpneg = sixth * (2 * fc + 5 * fzp - fzpp);
                        fzmpos = sixth * (2 * fc + 5 * fzm - fzmm);
                        fzmneg = sixth * (2 * fzm + 5 * fc - fzp);
                    } else if (scheme == OR_FROMM) {  /* advect.jl:174-212 */
                        fxppos = fc + quarter * (fxp - fxm);
                        fxpneg = fxp + quarter * (fc - fxpp);
                        fxmpos = fxm + quarter * (fc - fxmm);
                        fxmneg = fc + quarter * (fxm - fxp);
                        fzppos = fc + quarter * (fzp - fzm);
                        fzpneg = fzp + quarter * (fc - fzpp);
                        fzmpos = fzm + quarter * (fc - fzmm);
                        fzmneg = fc + quarter * (fzm - fzp);
                    } else {  /* OR_WENO3, advect.jl:231-269 */
                        fxppos = weno3_poly(fxm, fc, fxp, eps3);
                        fxpneg = weno3_poly(fxpp, fxp, fc, eps3);
                        fxmpos = weno3_poly(fxmm, fxm, fc, eps3);
                        fxmneg = weno3_poly(fxp, fc, fxm, eps3);
                        fzppos = weno3_poly(fzm, fc, fzp, eps3);
                        fzpneg = weno3_poly(fzpp, fzp, fc, eps3);
                        fzmpos = weno3_poly(fzmm, fzm, fc, eps3);
                        fzmneg = weno3_poly(fzp, fc, fzm, eps3);
                    }
                }
                double ul = AT(u, iz, ix), ur = AT(u, iz, ix + 1);
                double wt = AT(w, iz, ix), wb = AT(w, iz + 1, ix);
                qxr = vpos(ur) * fxppos + vneg(ur) * fxpneg;
                qxl = vpos(ul) * fxmpos + vneg(ul) * fxmneg;
                qzb = vpos(wb) * fzppos + vneg(wb) * fzpneg;
                qzt = vpos(wt) * fzmpos + vneg(wt) * fzmneg;
            }
            AT(adv, iz, ix) = (qxr - qxl + qzb - qzt) * invh;
        }
    }
}

/* ---- per-face flux kernels (advect.jl:444-695) ---------------------------
 * dir 0: x-faces, out (Nz, Nx+1), face ix between cells ix-1 and ix;
 * dir 1: z-faces, out (Nz+1, Nx).  `off(k)` reads f at face-relative cell k
 * (k = 0 is the cell after the face). */
static void advect_faces(omat q, int64_t q_r0, int64_t q_c0, int64_t nr, int64_t nc,
                         omat f, omat v, int scheme, int g, int dir) {
    const double half = 0.5, sixth = 1.0 / 6.0, quarter = 0.25;
    const double eps5 = 2.220446049250313e-16, eps3 = 1e-6;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t ix = 0; ix < nc; ++ix) {
        for (int64_t iz = 0; iz < nr; ++iz) {
            /* cell index of "fc" (the cell after the face) in halo coordinates */
            int64_t j, i, dj, di;
            if (dir == 0) { j = iz + g; i = ix + g; dj = 0; di = 1; }
            else          { j = iz + g; i = ix + g; dj = 1; di = 0; }
#define F(k) AT(f, j + (k) * dj, i + (k) * di)
            double out;
            if (scheme == OR_TVDIM) {
                /* velocity halo: v(face-1), v(face), v(face+1) */
                double vl, vc, vr;
                if (dir == 0) { vl = AT(v, iz, ix); vc = AT(v, iz, ix + 1); vr = AT(v, iz, ix + 2); }
                else          { vl = AT(v, iz, ix); vc = AT(v, iz + 1, ix); vr = AT(v, iz + 2, ix); }
                double fppos = tvd_flux(F(-2), F(-1), F(0), vpos(vl), vpos(vc));
                double fpneg = tvd_flux(F(1), F(0), F(-1), vneg(vr), vneg(vc));
                out = vpos(vc) * fppos + vneg(vc) * fpneg;
            } else {
                double vv = AT(v, iz, ix);
                if (scheme == OR_CENTR) {
                    out = vv * (F(-1) + F(0)) * half;
                } else {
                    double fppos, fpneg;
                    if (scheme == OR_UPWD1)      { fppos = F(-1); fpneg = F(0); }
                    else if (scheme == OR_QUICK) {
                        fppos = sixth * (2 * F(0) + 5 * F(-1) - F(-2));
                        fpneg = sixth * (2 * F(-1) + 5 * F(0) - F(1));
                    } else if (scheme == OR_FROMM) {
                        fppos = F(-1) + quarter * (F(0) - F(-2));
                        fpneg = F(0) + quarter * (F(-1) - F(1));
                    } else if (scheme == OR_WENO3) {
                        fppos = weno3_poly(F(-2), F(-1), F(0), eps3);
                        fpneg = weno3_poly(F(1), F(0), F(-1), eps3);
                    } else {
                        fppos = weno5_poly(F(-3), F(-2), F(-1), F(0), F(1), eps5);
                        fpneg = weno5_poly(F(2), F(1), F(0), F(-1), F(-2), eps5);
                    }
                    out = vpos(vv) * fppos + vneg(vv) * fpneg;
                }
            }
#undef F
            AT(q, q_r0 + iz, q_c0 + ix) = out;
        }
    }
}

static int check_advect_shapes(omat adv, omat fh, omat u, omat w, int scheme, int* g) {
    int halo = or_scheme_halo(scheme);
    if (halo < 0) return OR_ESCHEME;
    int64_t Nz = adv.rows, Nx = adv.cols;
    if (fh.rows != Nz + 2 * halo || fh.cols != Nx + 2 * halo) return OR_ESHAPE;
    if (scheme == OR_TVDIM) {
        if (u.rows != Nz || u.cols != Nx + 3) return OR_ESHAPE;
        if (w.rows != Nz + 3 || w.cols != Nx) return OR_ESHAPE;
    } else {
        if (u.rows != Nz || u.cols != Nx + 1) return OR_ESHAPE;
        if (w.rows != Nz + 1 || w.cols != Nx) return OR_ESHAPE;
    }
    *g = halo;
    return OR_OK;
}

/* advect.jl:374-425 */
int or_advect(omat adv, omat fh, omat u, omat w, double h, int scheme) {
    int g, st = check_advect_shapes(adv, fh, u, w, scheme, &g);
    if (st) return st;
    advect_div(adv, fh, u, w, h, scheme, g);
    return OR_OK;
}

/* MATLAB padding of flux arrays, advect.jl:779-795 / diffus.jl:130-144 */
static void pad_fluxes(omat qx, omat qz, int64_t Nz, int64_t Nx, int xbc, int zbc) {
    if (xbc == OR_PERIODIC && Nx > 1) {
        for (int64_t j = 0; j < qz.rows; ++j) { AT(qz, j, 0) = AT(qz, j, Nx); AT(qz, j, Nx + 1) = AT(qz, j, 1); }
    } else {
        for (int64_t j = 0; j < qz.rows; ++j) { AT(qz, j, 0) = AT(qz, j, 1); AT(qz, j, Nx + 1) = AT(qz, j, Nx); }
    }
    if (zbc == OR_PERIODIC && Nz > 1) {
        for (int64_t i = 0; i < qx.cols; ++i) { AT(qx, 0, i) = AT(qx, Nz, i); AT(qx, Nz + 1, i) = AT(qx, 1, i); }
    } else {
        for (int64_t i = 0; i < qx.cols; ++i) { AT(qx, 0, i) = AT(qx, 1, i); AT(qx, Nz + 1, i) = AT(qx, Nz, i); }
    }
}

/* advect.jl:717-798 */
int or_advect_with_flux(omat adv, omat qx, omat qz, omat fh, omat u, omat w,
                        double h, int scheme, int xbc, int zbc) {
    int64_t Nz = adv.rows, Nx = adv.cols;
    if (qx.rows != Nz + 2 || qx.cols != Nx + 1) return OR_ESHAPE;
    if (qz.rows != Nz + 1 || qz.cols != Nx + 2) return OR_ESHAPE;
    int g, st = check_advect_shapes(adv, fh, u, w, scheme, &g);
    if (st) return st;
    advect_div(adv, fh, u, w, h, scheme, g);
    advect_faces(qx, 1, 0, Nz, Nx + 1, fh, u, scheme, g, 0);
    advect_faces(qz, 0, 1, Nz + 1, Nx, fh, w, scheme, g, 1);
    pad_fluxes(qx, qz, Nz, Nx, xbc, zbc);
    return OR_OK;
}

static omat or_alloc(int64_t r, int64_t c) {
    omat m; m.rows = r; m.cols = c; m.ld = r;
    m.p = (double*)calloc((size_t)(r * c), sizeof(double));
    return m;
}

/* phsevo.jl:21-59 (advect_centered!) and 90-128 (advect_centered_with_flux!) */
int or_advect_centered(omat adv, const omat* qx, const omat* qz, omat f, omat u, omat w,
                       double h, int scheme, int xbc, int zbc) {
    int halo = or_scheme_halo(scheme);
    if (halo < 0) return OR_ESCHEME;
    if (xbc != OR_PERIODIC && xbc != OR_CLOSED) return OR_EBC;
    if (zbc != OR_PERIODIC && zbc != OR_CLOSED) return OR_EBC;
    int64_t Nz = adv.rows, Nx = adv.cols;
    if (f.rows != Nz || f.cols != Nx) return OR_ESHAPE;
    if (u.rows != Nz || u.cols != Nx + 1 || w.rows != Nz + 1 || w.cols != Nx) return OR_ESHAPE;
    omat fh = or_alloc(Nz + 2 * halo, Nx + 2 * halo);
    or_embed_interior(fh, f, halo);
    int st = or_fill_ghosts(fh, halo, xbc, zbc);
    omat up = u, wp = w;
    int padded = 0;
    if (!st && scheme == OR_TVDIM) {
        up = or_alloc(Nz, Nx + 3); wp = or_alloc(Nz + 3, Nx); padded = 1;
        for (int64_t i = 0; i < Nx + 1; ++i) for (int64_t j = 0; j < Nz; ++j) AT(up, j, i + 1) = AT(u, j, i);
        for (int64_t i = 0; i < Nx; ++i) for (int64_t j = 0; j < Nz + 1; ++j) AT(wp, j + 1, i) = AT(w, j, i);
        for (int64_t j = 0; j < Nz; ++j) {
            if (xbc == OR_PERIODIC) { AT(up, j, 0) = AT(u, j, Nx - 1); AT(up, j, Nx + 2) = AT(u, j, 1); }
            else                    { AT(up, j, 0) = AT(u, j, 0);      AT(up, j, Nx + 2) = AT(u, j, Nx); }
        }
        for (int64_t i = 0; i < Nx; ++i) {
            if (zbc == OR_PERIODIC) { AT(wp, 0, i) = AT(w, Nz - 1, i); AT(wp, Nz + 2, i) = AT(w, 1, i); }
            else                    { AT(wp, 0, i) = AT(w, 0, i);      AT(wp, Nz + 2, i) = AT(w, Nz, i); }
        }
    }
    if (!st) {
        if (qx && qz) st = or_advect_with_flux(adv, *qx, *qz, fh, up, wp, h, scheme, xbc, zbc);
        else          st = or_advect(adv, fh, up, wp, h, scheme);
    }
    free(fh.p);
    if (padded) { free(up.p); free(wp.p); }
    return st;
}

/* ---- diffusion (diffus.jl) ----------------------------------------------- */

/* diffus.jl:9-57 */
int or_diffus(omat dff, omat fh, omat kh, double h, int halo) {
    if (fh.rows != kh.rows || fh.cols != kh.cols) return OR_ESHAPE;
    if (dff.rows != fh.rows - 2 * halo || dff.cols != fh.cols - 2 * halo) return OR_ESHAPE;
    const double invh = 1.0 / h, invh2 = invh * invh, half = 0.5;
    const int64_t Nz = dff.rows, Nx = dff.cols;
#pragma omp parallel for num_threads(g_threads) schedule(static)
    for (int64_t ix = 0; ix < Nx; ++ix)
        for (int64_t iz = 0; iz < Nz; ++iz) {
            int64_t j = iz + halo, i = ix + halo;
            double fc = AT(fh, j, i), fmx = AT(fh, j, i - 1), fpx = AT(fh, j, i + 1);
            double fmz = AT(fh, j - 1, i), fpz = AT(fh, j + 1, i);
            double kc = AT(kh, j, i), kmx = AT(kh, j, i - 1), kpx = AT(kh, j, i + 1);
            double kmz = AT(kh, j - 1, i), kpz = AT(kh, j + 1, i);
            double flux_xr = (kpx + kc) * (fpx - fc);
            double flux_xl = (kc + kmx) * (fc - fmx);
            double flux_zb = (kpz + kc) * (fpz - fc);
            double flux_zt = (kc + kmz) * (fc - fmz);
            AT(dff, iz, ix) = half * invh2 * (flux_xr - flux_xl + flux_zb - flux_zt);
        }
    return OR_OK;
}

/* diffus.jl:105-147 */
int or_diffus_with_flux(omat dff, omat qx, omat qz, omat fh, omat kh, double h,
                        int halo, int xbc, int zbc) {
    int64_t Nz = dff.rows, Nx = dff.cols;
    if (qx.rows != Nz + 2 || qx.cols != Nx + 1) return OR_ESHAPE;
    if (qz.rows != Nz + 1 || qz.cols != Nx + 2) return OR_ESHAPE;
    int st = or_diffus(dff, fh, kh, h, halo);
    if (st) return st;
    const double invh = 1.0 / h, half = 0.5;
    /* diffus.jl:68-79: qx_int (Nz, Nx+1) at qx[2:Nz+1, 1:Nx+1] */
    for (int64_t ix = 0; ix < Nx + 1; ++ix)
        for (int64_t iz = 0; iz < Nz; ++iz) {
            int64_t j = iz + halo;
            double fL = AT(fh, j, ix + halo - 1), fR = AT(fh, j, ix + halo);
            double kL = AT(kh, j, ix + halo - 1), kR = AT(kh, j, ix + halo);
            AT(qx, iz + 1, ix) = -(kL + kR) * half * (fR - fL) * invh;
        }
    /* diffus.jl:81-91: qz_int (Nz+1, Nx) at qz[1:Nz+1, 2:Nx+1] */
    for (int64_t ix = 0; ix < Nx; ++ix)
        for (int64_t iz = 0; iz < Nz + 1; ++iz) {
            int64_t i = ix + halo;
            double fT = AT(fh, iz + halo - 1, i), fB = AT(fh, iz + halo, i);
            double kT = AT(kh, iz + halo - 1, i), kB = AT(kh, iz + halo, i);
            AT(qz, iz, ix + 1) = -(kT + kB) * half * (fB - fT) * invh;
        }
    pad_fluxes(qx, qz, Nz, Nx, xbc, zbc);
    return OR_OK;
}

/* phsevo.jl:67-79 and 135-149 */
int or_diffus_centered(omat dff, const omat* qx, const omat* qz, omat f, omat k,
                       double h, int xbc, int zbc) {
    if (xbc != OR_PERIODIC && xbc != OR_CLOSED) return OR_EBC;
    if (zbc != OR_PERIODIC && zbc != OR_CLOSED) return OR_EBC;
    int64_t Nz = dff.rows, Nx = dff.cols;
    if (f.rows != Nz || f.cols != Nx || k.rows != Nz || k.cols != Nx) return OR_ESHAPE;
    omat fh = or_alloc(Nz + 2, Nx + 2), kh = or_alloc(Nz + 2, Nx + 2);
    or_embed_interior(fh, f, 1); or_embed_interior(kh, k, 1);
    int st = or_fill_ghosts(fh, 1, xbc, zbc);
    if (!st) st = or_fill_ghosts(kh, 1, xbc, zbc);
    if (!st) {
        if (qx && qz) st = or_diffus_with_flux(dff, *qx, *qz, fh, kh, h, 1, xbc, zbc);
        else          st = or_diffus(dff, fh, kh, h, 1);
    }
    free(fh.p); free(kh.p);
    return st;
}
