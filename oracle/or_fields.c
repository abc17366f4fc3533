/*
 * or_fields.c -- oracle (test infrastructure only): restatement of the field
 * updates of xcore.jl: src/update.jl (constitutive pipeline, update_dt),
 * src/store.jl (history shift, segregation speed, phase velocities),
 * src/phsevo.jl:164-214 (phase evolution) and src/timing.jl.
 *
 * Index maps: the reference's ghost vectors icx/icz (grid.jl:68-71,
 * update.jl:78-79) become cx()/cz() below: 0-based index -1 and N map to the
 * wrap (periodic) or to the boundary cell (closed).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define AT(m, i, j) ((m).p[(int64_t)(i) + (int64_t)(j) * (m).ld])
#define EPS 2.220446049250313e-16

static inline int64_t gidx(int64_t i, int64_t n, int bc) {
    if (i < 0) return bc == OR_PERIODIC ? n - 1 : 0;
    if (i >= n) return bc == OR_PERIODIC ? 0 : n - 1;
    return i;
}
static inline double jl_max(double a, double b) {
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}
static inline double jl_clamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* timing.jl:18-36 */
int or_time_coefs(int tint, int64_t step, double c[6]) {
    static const double be[6] = {1, 1, 0, 1, 0, 0};
    static const double bd2[6] = {1.5, 2.0, -0.5, 1, 0, 0};
    static const double cn[6] = {1, 1, 0, 0.5, 0.5, 0};
    static const double bd2si[6] = {1.5, 2.0, -0.5, 0.75, 0.5, -0.25};
    const double* src = NULL;
    if (tint == 0 || tint == 1) {           /* *im */
        if (tint == 0 || step <= 2) src = be;
        else src = bd2;
    } else if (tint == 2 || tint == 3) {    /* *si */
        if (step == 1) src = be;
        else if (tint == 2 || step == 2) src = cn;
        else src = bd2si;
    } else {
        return OR_EARG;
    }
    memcpy(c, src, sizeof(double) * 6);
    return OR_OK;
}

/* update.jl:18-58 */
int or_update_volume_fractions(or_phase* ph, or_fluid* fl, const or_par* par) {
    const int64_t Nz = ph->x.rows, Nx = ph->x.cols;
    const double sq = sqrt(EPS);
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double x = AT(ph->x, j, i), m = AT(ph->m, j, i);
            double rho = 1.0 / (m / par->rhom0 + x / par->rhox0);
            AT(fl->rho, j, i) = rho;
            AT(ph->chi, j, i) = jl_clamp(x * rho / par->rhox0, EPS, 1.0 - EPS);
            AT(ph->mu, j, i) = jl_clamp(m * rho / par->rhom0, EPS, 1.0 - EPS);
            ph->hasx.p[j + i * ph->hasx.ld] = x >= sq;
            ph->hasm.p[j + i * ph->hasm.ld] = m >= sq;
        }
    return OR_OK;
}

/* Statistics.mean(rhow, dims=2): sequential sum over columns, times 1/Nx
 * (mean! scales by the Rational length(R)//length(A), converted to Float64). */
static void row_mean(omat a, double* out) {
    const double s = 1.0 / (double)a.cols;
    for (int64_t j = 0; j < a.rows; ++j) {
        double acc = 0.0;
        for (int64_t i = 0; i < a.cols; ++i) acc += AT(a, j, i);
        out[j] = acc * s;
    }
}

/* update.jl:73-107 */
int or_update_interpolations(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, int xbc, int zbc) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j)
            AT(fl->rhow, j, i) = (AT(fl->rho, gidx(j - 1, Nz, zbc), i) + AT(fl->rho, gidx(j, Nz, zbc), i)) / 2.0;
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j)
            AT(fl->rhou, j, i) = (AT(fl->rho, j, gidx(i - 1, Nx, xbc)) + AT(fl->rho, j, gidx(i, Nx, xbc))) / 2.0;
    double* ref = (double*)malloc(sizeof(double) * (size_t)(Nz + 1));
    row_mean(fl->rhow, ref);
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            AT(fl->Drho, j, i) = AT(fl->rhow, j, i) - ref[j];
            AT(ph->Drhom, j, i) = par->rhom0 - AT(fl->rhow, j, i);
            AT(ph->Drhox, j, i) = par->rhox0 - AT(fl->rhow, j, i);
        }
    free(ref);
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t ci = gidx(i - 1, Nx, xbc);
            int64_t z0 = gidx(j - 1, Nz, zbc), z1 = gidx(j, Nz, zbc);
            AT(ph->chiw, j, i) = (AT(ph->chi, z0, ci) + AT(ph->chi, z1, ci)) / 2.0;
            AT(ph->muw, j, i) = (AT(ph->mu, z0, ci) + AT(ph->mu, z1, ci)) / 2.0;
        }
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz + 2; ++j) {
            int64_t zj = gidx(j - 1, Nz, zbc);
            int64_t x0 = gidx(i - 1, Nx, xbc), x1 = gidx(i, Nx, xbc);
            AT(ph->chiu, j, i) = (AT(ph->chi, zj, x0) + AT(ph->chi, zj, x1)) / 2.0;
            AT(ph->muu, j, i) = (AT(ph->mu, zj, x0) + AT(ph->mu, zj, x1)) / 2.0;
        }
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t z0 = gidx(j - 1, Nz, zbc), z1 = gidx(j, Nz, zbc);
            AT(ph->x_w, j, i) = (AT(ph->x, z0, i) + AT(ph->x, z1, i)) / 2.0;
            AT(ph->m_w, j, i) = (AT(ph->m, z0, i) + AT(ph->m, z1, i)) / 2.0;
        }
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            int64_t x0 = gidx(i - 1, Nx, xbc), x1 = gidx(i, Nx, xbc);
            AT(ph->x_u, j, i) = (AT(ph->x, j, x0) + AT(ph->x, j, x1)) / 2.0;
            AT(ph->m_u, j, i) = (AT(ph->m, j, x0) + AT(ph->m, j, x1)) / 2.0;
        }
    return OR_OK;
}

/* update.jl:119-136 */
int or_update_pressure(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double h = g->h;
    double* ref = (double*)malloc(sizeof(double) * (size_t)(Nz + 1));
    double* col = (double*)malloc(sizeof(double) * (size_t)Nz);
    row_mean(fl->rhow, ref);
    col[0] = ref[0] * par->g0 * h / 2.0 + par->Ptop;
    for (int64_t j = 1; j < Nz; ++j) col[j] = col[j - 1] + ref[j] * par->g0 * h;
    const double floor_p = par->Ptop / 100.0;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            AT(ph->Pl, j, i) = col[j];
            AT(ph->Pt, j, i) = jl_max(floor_p, col[j] + AT(fl->P, j + 1, i + 1));
        }
    free(ref);
    free(col);
    return OR_OK;
}

/* update.jl:144-177: permission-weight mixture viscosity of one cell */
static double etamix_cell(double f1, double f2, const or_par* p) {
    const double* AA = p->AA; const double* BB = p->BB; const double* CC = p->CC;
    double s11 = pow(f1 / BB[0], 1.0 / CC[0]), s12 = pow(f2 / BB[1], 1.0 / CC[1]);
    double s21 = pow(f1 / BB[2], 1.0 / CC[2]), s22 = pow(f2 / BB[3], 1.0 / CC[3]);
    double n1 = s11 + s12, n2 = s21 + s22;
    double Sf11 = s11 / n1, Sf12 = s12 / n1, Sf21 = s21 / n2, Sf22 = s22 / n2;
    double a1 = AA[0] * Sf11 + AA[1] * Sf12;
    double a2 = AA[2] * Sf21 + AA[3] * Sf22;
    double Xf11 = a1 * f1 + (1.0 - a1) * Sf11;
    double Xf12 = a1 * f2 + (1.0 - a1) * Sf12;
    double Xf21 = a2 * f1 + (1.0 - a2) * Sf21;
    double Xf22 = a2 * f2 + (1.0 - a2) * Sf22;
    double ex = p->etax0, em = p->etam0;
    double thtv1 = pow(ex / ex, Xf11) * pow(em / ex, Xf12);
    double thtv2 = pow(ex / em, Xf21) * pow(em / em, Xf22);
    return f1 * (ex * thtv1) + f2 * (em * thtv2);
}

int or_update_rheology(or_phase* ph, const or_par* par) {
    const int64_t Nz = ph->chi.rows, Nx = ph->chi.cols;
#pragma omp parallel for num_threads(or_get_threads())
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j)
            AT(ph->etamix, j, i) = etamix_cell(AT(ph->chi, j, i), AT(ph->mu, j, i), par);
    return OR_OK;
}

/* update.jl:202-280 */
int or_update_kinematics(or_phase* ph, or_fluid* fl, const or_grid* g) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double invh = 1.0 / g->h;
    omat W = fl->W, U = fl->U;
    for (int64_t i = 0; i < Nx; ++i)      /* compute_strain_kernel! 202-213 */
        for (int64_t j = 0; j < Nz; ++j) {
            double Uxc = (AT(U, j + 1, i + 1) - AT(U, j + 1, i)) * invh;
            double Wzc = (AT(W, j + 1, i + 1) - AT(W, j, i + 1)) * invh;
            double div = Uxc + Wzc;
            AT(ph->Div_V, j, i) = div;
            AT(ph->exx, j, i) = Uxc - div / 3.0;
            AT(ph->ezz, j, i) = Wzc - div / 3.0;
        }
    for (int64_t i = 0; i <= Nx; ++i)     /* compute_shear_kernel! 215-223 */
        for (int64_t j = 0; j <= Nz; ++j) {
            double dUdz = (AT(U, j + 1, i) - AT(U, j, i)) * invh;
            double dWdx = (AT(W, j, i + 1) - AT(W, j, i)) * invh;
            AT(ph->exz, j, i) = (dUdz + dWdx) / 2.0;
        }
    for (int64_t i = 0; i < Nx; ++i)      /* compute_eII_V_kernel! 225-241 */
        for (int64_t j = 0; j < Nz; ++j) {
            double e1 = AT(ph->exz, j, i), e2 = AT(ph->exz, j + 1, i);
            double e3 = AT(ph->exz, j, i + 1), e4 = AT(ph->exz, j + 1, i + 1);
            double sq = (e1 * e1 + e2 * e2 + e3 * e3 + e4 * e4) / 4.0;
            double exx = AT(ph->exx, j, i), ezz = AT(ph->ezz, j, i);
            AT(ph->eII, j, i) = sqrt(EPS + (exx * exx + ezz * ezz) / 2.0 + sq);
            double wc = (AT(W, j, i + 1) + AT(W, j + 1, i + 1)) / 2.0;
            double uc = (AT(U, j + 1, i) + AT(U, j + 1, i + 1)) / 2.0;
            AT(ph->V, j, i) = sqrt(wc * wc + uc * uc);
        }
    for (int64_t i = 0; i < Nx; ++i)      /* compute_segspeed_kernel! 243-253 */
        for (int64_t j = 0; j < Nz; ++j) {
            double w1 = AT(ph->wx, j, i + 1), w2 = AT(ph->wx, j + 1, i + 1);
            double a = (w1 + w2) / 2.0;
            AT(ph->vx, j, i) = sqrt(a * a + EPS);
            double m1 = AT(ph->wm, j, i + 1), m2 = AT(ph->wm, j + 1, i + 1);
            double b = (m1 + m2) / 2.0;
            AT(ph->vm, j, i) = sqrt(b * b + EPS);
        }
    return OR_OK;
}

/* update.jl:411-426 */
int or_compute_corner_eta(omat etaco, omat eta, int xbc, int zbc) {
    const int64_t Nz = eta.rows, Nx = eta.cols;
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t z0 = gidx(j - 1, Nz, zbc), z1 = gidx(j, Nz, zbc);
            int64_t x0 = gidx(i - 1, Nx, xbc), x1 = gidx(i, Nx, xbc);
            double e00 = AT(eta, z0, x0), e10 = AT(eta, z1, x0), e01 = AT(eta, z0, x1), e11 = AT(eta, z1, x1);
            AT(etaco, j, i) = pow(e00 * e10 * e01 * e11, 0.25);
        }
    return OR_OK;
}

static double min_all(omat a) {
    double m = INFINITY;
    for (int64_t i = 0; i < a.cols; ++i)
        for (int64_t j = 0; j < a.rows; ++j) {
            double v = AT(a, j, i);
            if (v < m) m = v;
        }
    return m;
}

/* update.jl:286-343 */
int or_update_viscosity(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double L0 = sc->L0;
    double* prev = (double*)malloc(sizeof(double) * (size_t)(Nz * Nx));
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            prev[j + i * Nz] = AT(fl->eta, j, i);
            double eII = AT(ph->eII, j, i), rho = AT(fl->rho, j, i);
            double ke = eII * L0 * L0;
            double ReL = AT(ph->V, j, i) * L0 / (AT(ph->etamix, j, i) / rho);
            double fReL = 1.0 - exp(-ReL);
            double etae = fReL * ke * rho;
            AT(fl->eta, j, i) = AT(ph->etamix, j, i) + etae;
            AT(ph->etae, j, i) = etae;
            AT(ph->ke, j, i) = ke;
            AT(ph->kx, j, i) = ke;
            AT(ph->ReL, j, i) = ReL;
            AT(ph->fReL, j, i) = fReL;
        }
    const double emax = min_all(fl->eta) * par->etacntr;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double e = 1.0 / (1.0 / emax + 1.0 / AT(fl->eta, j, i));
            e = 0.5 * e + (1.0 - 0.5) * prev[j + i * Nz];
            AT(fl->eta, j, i) = e;
            AT(ph->ReD, j, i) = AT(ph->V, j, i) * sc->D0 / (e / AT(fl->rho, j, i));
        }
    free(prev);
    return or_compute_corner_eta(fl->etaco, fl->eta, OR_PERIODIC, OR_CLOSED);
}

/* update.jl:345-409 */
int or_update_segregation_viscosity(or_phase* ph, or_fluid* fl, const or_par* par, const or_scales* sc, int zbc) {
    const int64_t Nz = ph->etas.rows, Nx = ph->etas.cols;
    const double l0 = sc->l0;
    double* prev = (double*)malloc(sizeof(double) * (size_t)(Nz * Nx));
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            prev[j + i * Nz] = AT(ph->etas, j, i);
            double vx = AT(ph->vx, j, i), rho = AT(fl->rho, j, i), em = AT(ph->etamix, j, i);
            double Rel = vx * l0 * rho / em;
            double fRel = 1.0 - exp(-Rel);
            double ks = vx * l0;
            double etat = fRel * ks * rho;
            AT(ph->etas, j, i) = em + etat;
            AT(ph->Rel, j, i) = Rel;
            AT(ph->fRel, j, i) = fRel;
            AT(ph->ks, j, i) = ks;
            AT(ph->etat, j, i) = etat;
            AT(ph->kx, j, i) = ks + AT(ph->fReL, j, i) * AT(ph->ke, j, i);
            AT(ph->Rc, j, i) = AT(ph->V, j, i) / jl_max(vx, EPS);
        }
    const double emax = min_all(ph->etas) * par->etacntr;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double e = 1.0 / (1.0 / emax + 1.0 / AT(ph->etas, j, i));
            e = 0.5 * e + (1.0 - 0.5) * prev[j + i * Nz];
            AT(ph->etas, j, i) = e;
            AT(ph->Red, j, i) = AT(ph->vx, j, i) * par->d0 / (e / AT(fl->rho, j, i));
        }
    free(prev);
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j)
            AT(ph->etasw, j, i) = sqrt(AT(ph->etas, gidx(j - 1, Nz, zbc), i) * AT(ph->etas, gidx(j, Nz, zbc), i));
    return OR_OK;
}

/* update.jl:432-476 */
int or_update_stresses(or_phase* ph, or_fluid* fl, const or_grid* g, const or_scales* sc) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j) AT(ph->txz, j, i) = AT(fl->etaco, j, i) * AT(ph->exz, j, i);
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double eta = AT(fl->eta, j, i);
            double txx = eta * AT(ph->exx, j, i), tzz = eta * AT(ph->ezz, j, i);
            AT(ph->txx, j, i) = txx;
            AT(ph->tzz, j, i) = tzz;
            double t1 = AT(ph->txz, j, i), t2 = AT(ph->txz, j + 1, i);
            double t3 = AT(ph->txz, j, i + 1), t4 = AT(ph->txz, j + 1, i + 1);
            double sq = (t1 * t1 + t2 * t2 + t3 * t3 + t4 * t4) / 4.0;
            AT(ph->tII, j, i) = sqrt(EPS + (txx * txx + tzz * tzz) / 2.0 + sq);
            AT(ph->Ra, j, i) = AT(ph->V, j, i) * sc->D0 / AT(ph->kx, j, i);
        }
    return OR_OK;
}

/* update.jl:525-538 */
int or_update(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc, int xbc, int zbc) {
    if (xbc != OR_PERIODIC && xbc != OR_CLOSED) return OR_EBC;
    if (zbc != OR_PERIODIC && zbc != OR_CLOSED) return OR_EBC;
    or_update_volume_fractions(ph, fl, par);
    or_update_interpolations(ph, fl, g, par, xbc, zbc);
    or_update_pressure(ph, fl, g, par);
    or_update_rheology(ph, par);
    or_update_kinematics(ph, fl, g);
    or_update_viscosity(ph, fl, g, par, sc);
    or_update_segregation_viscosity(ph, fl, par, sc, zbc);
    or_update_stresses(ph, fl, g, sc);
    return OR_OK;
}

static double maxabs(omat a) {
    double m = 0.0;
    for (int64_t i = 0; i < a.cols; ++i)
        for (int64_t j = 0; j < a.rows; ++j) {
            double v = fabs(AT(a, j, i));
            if (v > m) m = v;
        }
    return m;
}

/* update.jl:494-502 */
double or_update_dt(const or_phase* ph, const or_fluid* fl, const or_grid* g, const or_par* par, double dt_prev) {
    const double h = g->h;
    double kmax = -INFINITY;
    for (int64_t i = 0; i < ph->kx.cols; ++i)
        for (int64_t j = 0; j < ph->kx.rows; ++j)
            if (AT(ph->kx, j, i) > kmax) kmax = AT(ph->kx, j, i);
    double hh = h / 2.0;
    double dtk = hh * hh / kmax;
    double vmax = jl_max(jl_max(maxabs(ph->Wx), maxabs(ph->Wm)), jl_max(maxabs(ph->Ux), maxabs(ph->Um))) + EPS;
    double dta = hh / vmax;
    double a = 1.5 * dt_prev, b = dtk < par->CFL * dta ? dtk : par->CFL * dta;
    double r = a < b ? a : b;
    return r < par->dtmax ? r : par->dtmax;
}

/* store.jl:13-30 */
int or_store_previous(or_fluid* fl, or_phase* ph) {
#define CPY(dst, src) do { for (int64_t i_ = 0; i_ < (dst).cols; ++i_) for (int64_t j_ = 0; j_ < (dst).rows; ++j_) AT(dst, j_, i_) = AT(src, j_, i_); } while (0)
    CPY(ph->Xoo, ph->Xo); CPY(ph->Xo, ph->X);
    CPY(ph->Moo, ph->Mo); CPY(ph->Mo, ph->M);
    CPY(ph->dXdtoo, ph->dXdto); CPY(ph->dXdto, ph->dXdt);
    CPY(fl->rhooo, fl->rhoo); CPY(fl->rhoo, fl->rho);
    CPY(fl->drhodtoo, fl->drhodto); CPY(fl->drhodto, fl->drhodt);
    CPY(fl->rhoWoo, fl->rhoWo);
    for (int64_t i = 0; i < fl->rhoWo.cols; ++i)
        for (int64_t j = 0; j < fl->rhoWo.rows; ++j) AT(fl->rhoWo, j, i) = AT(fl->rhow, j, i) * AT(fl->W, j, i + 1);
    CPY(fl->rhoUoo, fl->rhoUo);
    for (int64_t i = 0; i < fl->rhoUo.cols; ++i)
        for (int64_t j = 0; j < fl->rhoUo.rows; ++j) AT(fl->rhoUo, j, i) = AT(fl->rhou, j, i) * AT(fl->U, j + 1, i);
#undef CPY
    return OR_OK;
}

/* store.jl:47-68 */
int or_update_segregation_speed(or_phase* ph, const or_grid* g, const or_par* par, int xbc) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double d02 = par->d0 * par->d0;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j)
            AT(ph->wx, j, i + 1) = d02 / AT(ph->etasw, j, i) * AT(ph->Drhox, j, i) * par->g0;
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) AT(ph->wx, j, i) *= AT(ph->bndtaperw, j, i);
    for (int64_t j = 0; j <= Nz; ++j) {
        AT(ph->wx, j, 0) = AT(ph->wx, j, Nx);
        AT(ph->wx, j, Nx + 1) = AT(ph->wx, j, 1);
    }
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t ci = gidx(i - 1, Nx, xbc);
            AT(ph->wm, j, i) = -AT(ph->x_w, j, ci) / jl_max(AT(ph->m_w, j, ci), EPS) * AT(ph->wx, j, i);
        }
    return OR_OK;
}

/* store.jl:80-105 */
int or_update_phase_velocities(or_phase* ph, or_fluid* fl, const or_noise* ns, const or_grid* g,
                               const or_par* par, int xbc, int zbc) {
    if (xbc != OR_PERIODIC && xbc != OR_CLOSED) return OR_EBC;
    if (zbc != OR_PERIODIC && zbc != OR_CLOSED) return OR_EBC;
    or_update_segregation_speed(ph, g, par, xbc);
    const int64_t Nz = g->Nz, Nx = g->Nx;
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t ci = gidx(i - 1, Nx, xbc);
            double xw = AT(ph->x_w, j, ci), mw = AT(ph->m_w, j, ci);
            double sx = AT(ns->xisw, j, i) + AT(ns->xixw, j, i);
            AT(ph->Wx, j, i) = AT(fl->W, j, i) + AT(ph->wx, j, i) + sx + AT(ns->xiew, j, i);
            AT(ph->Wm, j, i) = AT(fl->W, j, i) + AT(ph->wm, j, i) - xw / jl_max(mw, EPS) * sx + AT(ns->xiew, j, i);
        }
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz + 2; ++j) {
            int64_t zj = gidx(j - 1, Nz, zbc);
            double xu = AT(ph->x_u, zj, i), mu = AT(ph->m_u, zj, i);
            double sx = AT(ns->xisu, j, i) + AT(ns->xixu, j, i);
            AT(ph->Ux, j, i) = AT(fl->U, j, i) + sx + AT(ns->xieu, j, i);
            AT(ph->Um, j, i) = AT(fl->U, j, i) + (-xu) / jl_max(mu, EPS) * sx + AT(ns->xieu, j, i);
        }
    return OR_OK;
}

static omat sub(omat a, int64_t r0, int64_t c0, int64_t nr, int64_t nc) {
    omat v; v.p = a.p + r0 + c0 * a.ld; v.rows = nr; v.cols = nc; v.ld = a.ld;
    return v;
}

/* phsevo.jl:164-214 */
int or_phsevo(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc,
              int scheme, int xbc, int zbc, const double coef[6], double dt, double alpha) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double h = g->h;
    omat Wx = sub(ph->Wx, 0, 1, Nz + 1, Nx), Wm = sub(ph->Wm, 0, 1, Nz + 1, Nx);
    omat Ux = sub(ph->Ux, 1, 0, Nz, Nx + 1), Um = sub(ph->Um, 1, 0, Nz, Nx + 1);
    int st = or_advect_centered(ph->advn_X, &ph->qx_advn_X, &ph->qz_advn_X, ph->X, Ux, Wx, h, scheme, xbc, zbc);
    if (st) return st;
    st = or_advect_centered(ph->advn_M, &ph->qx_advn_M, &ph->qz_advn_M, ph->M, Um, Wm, h, scheme, xbc, zbc);
    if (st) return st;
    omat Xkx; Xkx.rows = Nz; Xkx.cols = Nx; Xkx.ld = Nz;
    Xkx.p = (double*)malloc(sizeof(double) * (size_t)(Nz * Nx));
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            AT(ph->advn_rho, j, i) = AT(ph->advn_X, j, i) + AT(ph->advn_M, j, i);
            AT(Xkx, j, i) = AT(ph->X, j, i) * AT(ph->kx, j, i);
        }
    st = or_diffus_centered(ph->dffn_X, &ph->qx_dffn_X, &ph->qz_dffn_X, ph->chi, Xkx, h, xbc, zbc);
    free(Xkx.p);
    if (st) return st;
    const double a1 = coef[0], a2 = coef[1], a3 = coef[2], b1 = coef[3], b2 = coef[4], b3 = coef[5];
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double Gx = sc->G0 * (1.0 - AT(ph->x, j, i)) * AT(ph->bndshape, j, i);
            AT(ph->Gx, j, i) = Gx;
            double dXdt = -AT(ph->advn_X, j, i) + AT(ph->dffn_X, j, i) + Gx;
            AT(ph->dXdt, j, i) = dXdt;
            double X = AT(ph->X, j, i);
            double res = (a1 * X - a2 * AT(ph->Xo, j, i) - a3 * AT(ph->Xoo, j, i)) -
                         (b1 * dXdt + b2 * AT(ph->dXdto, j, i) + b3 * AT(ph->dXdtoo, j, i)) * dt;
            double upd = -alpha * res / a1;
            X = X + upd;
            double rho = AT(fl->rho, j, i);
            X = jl_clamp(X, rho * EPS, rho * (1.0 - EPS));
            AT(ph->X, j, i) = X;
            double M = rho - X;
            AT(ph->M, j, i) = M;
            AT(ph->x, j, i) = X / rho;
            AT(ph->m, j, i) = M / rho;
        }
    return OR_OK;
}
