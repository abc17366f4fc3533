/*
 * or_stokes.c -- oracle (test infrastructure only): restatement of the Stokes
 * assembly of src/fluidmech.jl:71-534.  Produces the global matrix
 * LL = [KV GG; DM KP] (after the P/U pins) as 0-based COO triplets with
 * duplicates (summed by the consumer, as `sparse(I, J, V)` does) and the
 * dense right-hand side RR; the direct solve (the reference uses UMFPACK) is
 * done by the caller with a sparse LU.  Also performs the MFS update of
 * fluidmech.jl:115-127 in non-benchmark mode.
 *
 * Unknown numbering (fluidmech.jl:95-100): W (Nz+1, Nx+2) column-major first,
 * then U (Nz+2, Nx+1), then P (Nz+2, Nx+2).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define AT(m, i, j) ((m).p[(int64_t)(i) + (int64_t)(j) * (m).ld])

typedef struct {
    int64_t *I, *J;
    double* V;
    int* tag;
    int64_t n, cap;
} trip;

static void put(trip* t, int tag, int64_t i, int64_t j, double v) {
    if (t->n < t->cap) {
        t->I[t->n] = i; t->J[t->n] = j; t->V[t->n] = v; t->tag[t->n] = tag;
    }
    t->n++;
}

enum { KV = 0, GG = 1, DM = 2, KP = 3 };

/* Julia round(Int, x): nearest, ties to even */
static int64_t jround(double x) { return (int64_t)nearbyint(x); }

static omat mk(int64_t r, int64_t c) {
    omat m; m.rows = r; m.cols = c; m.ld = r;
    m.p = (double*)calloc((size_t)(r * c), sizeof(double));
    return m;
}

int or_fluidmech_assemble(or_fluid* fl, const or_phase* ph, const or_grid* g, const or_fm_opts* o,
                          const omat* bn, int64_t* I, int64_t* J, double* V, int64_t cap, int64_t* nnz_out,
                          double* RR) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double h = g->h, invh = 1.0 / h, invh2 = (1.0 / h) * (1.0 / h);
    const double g0 = o->g0, gamma = o->gamma, dt = o->dt;
    const double a1 = o->a1, a2 = o->a2, a3 = o->a3, b1 = o->b1, b2 = o->b2, b3 = o->b3;
    const int bnchm = bn != NULL;
    const int64_t NP = (Nz + 2) * (Nx + 2), NW = (Nz + 1) * (Nx + 2), NU = (Nz + 2) * (Nx + 1);
    /* 0-based maps */
#define MW(r, c) ((int64_t)(r) + (int64_t)(c) * (Nz + 1))
#define MU(r, c) (NW + (int64_t)(r) + (int64_t)(c) * (Nz + 2))
#define MP(r, c) (NW + NU + (int64_t)(r) + (int64_t)(c) * (Nz + 2))
    int* tag = (int*)malloc(sizeof(int) * (size_t)cap);
    trip t = {I, J, V, tag, 0, cap};
    for (int64_t k = 0; k < NW + NU + NP; ++k) RR[k] = 0.0;

    /* ---- MFS update (fluidmech.jl:115-127) */
    if (!bnchm) {
        for (int64_t i = 0; i < Nx; ++i)
            for (int64_t j = 0; j < Nz; ++j) {
                double drhodt = -AT(ph->advn_rho, j, i);
                AT(fl->drhodt, j, i) = drhodt;
                double res = (a1 * AT(fl->rho, j, i) - a2 * AT(fl->rhoo, j, i) - a3 * AT(fl->rhooo, j, i)) / dt -
                             (b1 * drhodt + b2 * AT(fl->drhodto, j, i) + b3 * AT(fl->drhodtoo, j, i));
                AT(fl->MFS, j, i) = AT(fl->MFS, j, i) - o->alpha * res;
            }
    }
    double msum = 0.0;
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) msum += AT(fl->MFS, j, i);
    const double MFSmean = msum / (double)(Nz * Nx);

    /* ghost index maps (fluidmech.jl:132-134), 0-based, argument is the 1-based position - 1 */
#define ICX(k) ((k) == 0 ? Nx - 1 : ((k) == Nx + 1 ? 0 : (k) - 1))      /* [Nx; 1:Nx; 1] */
#define IFX(k) ((k) == 0 ? Nx - 1 : ((k) == Nx + 2 ? 1 : (k) - 1))      /* [Nx; 1:Nx+1; 2] */
    omat W = fl->W, U = fl->U, eta = fl->eta, etaco = fl->etaco, rho = fl->rho, rhow = fl->rhow, rhou = fl->rhou;

    /* ---- W rows: boundaries (fluidmech.jl:145-191) */
    for (int64_t r = 0; r <= Nz; ++r) {        /* left: W[:,1] + sds W[:,end-1] = 0 */
        put(&t, KV, MW(r, 0), MW(r, 0), 1.0);
        put(&t, KV, MW(r, 0), MW(r, Nx), (double)o->sds);
    }
    for (int64_t r = 0; r <= Nz; ++r) {        /* right: W[:,end] + sds W[:,2] = 0 */
        put(&t, KV, MW(r, Nx + 1), MW(r, Nx + 1), 1.0);
        put(&t, KV, MW(r, Nx + 1), MW(r, 1), (double)o->sds);
    }
    for (int64_t c = 1; c <= Nx; ++c) put(&t, KV, MW(0, c), MW(0, c), 1.0);  /* top: W = 0 */
    {
        const double oc = o->open_cnv ? 1.0 : 0.0;
        const double mfbg = MFSmean * g->D / h;
        for (int64_t c = 1; c <= Nx; ++c) {  /* bottom */
            int64_t ii = MW(Nz, c);
            put(&t, KV, ii, ii, AT(rhow, Nz, c - 1) * invh);
            put(&t, KV, ii, MW(Nz - 1, c), -oc * AT(rhow, Nz - 1, c - 1) * invh);
            put(&t, KV, ii, MU(Nz, c), oc * AT(rhou, Nz - 1, c) * invh);
            put(&t, KV, ii, MU(Nz, c - 1), -oc * AT(rhou, Nz - 1, c - 1) * invh);
            RR[ii] = oc * AT(fl->MFS, Nz - 1, c - 1) + (1.0 - oc) * mfbg;
        }
    }
    /* ---- W interior rows (fluidmech.jl:194-247) */
    {
        /* momentum advection: f = rhow[2:end-1,:] .* W[2:end-1,2:end-1] (Nz-1, Nx) */
        omat f = mk(Nz - 1, Nx), um = mk(Nz - 1, Nx + 1), wm = mk(Nz, Nx), adv = mk(Nz - 1, Nx);
        for (int64_t i = 0; i < Nx; ++i)
            for (int64_t j = 0; j < Nz - 1; ++j) AT(f, j, i) = AT(rhow, j + 1, i) * AT(W, j + 1, i + 1);
        for (int64_t i = 0; i <= Nx; ++i)
            for (int64_t j = 0; j < Nz - 1; ++j) AT(um, j, i) = (AT(U, j + 1, i) + AT(U, j + 2, i)) / 2.0;
        for (int64_t i = 0; i < Nx; ++i)
            for (int64_t j = 0; j < Nz; ++j) AT(wm, j, i) = (AT(W, j, i + 1) + AT(W, j + 1, i + 1)) / 2.0;
        int st = or_advect_centered(adv, NULL, NULL, f, um, wm, h, o->scheme, o->xbc, o->zbc);
        if (st) { free(tag); free(f.p); free(um.p); free(wm.p); free(adv.p); return st; }
        for (int64_t c = 1; c <= Nx; ++c)
            for (int64_t r = 1; r <= Nz - 1; ++r) {
                const int64_t ii = MW(r, c);
                const double C1 = AT(etaco, r, c - 1), C2 = AT(etaco, r, c);
                const double P1 = AT(eta, r - 1, c - 1), P2 = AT(eta, r, c - 1);
                put(&t, KV, ii, ii, a1 * AT(rhow, r, c - 1) / dt);
                put(&t, KV, ii, ii, (2.0 / 3.0) * (P1 + P2) * invh2 + (1.0 / 2.0) * (C1 + C2) * invh2);
                put(&t, KV, ii, MW(r - 1, c), -(2.0 / 3.0) * P1 * invh2);
                put(&t, KV, ii, MW(r + 1, c), -(2.0 / 3.0) * P2 * invh2);
                put(&t, KV, ii, MW(r, c - 1), -(1.0 / 2.0) * C1 * invh2);
                put(&t, KV, ii, MW(r, c + 1), -(1.0 / 2.0) * C2 * invh2);
                if (!bnchm) {
                    double ddz = (AT(rho, r, c - 1) - AT(rho, r - 1, c - 1)) * invh;
                    put(&t, KV, ii, ii, ddz * g0 * dt);
                }
                put(&t, KV, ii, MU(r, c - 1), -((1.0 / 2.0) * C1 - (1.0 / 3.0) * P1) * invh2);
                put(&t, KV, ii, MU(r + 1, c - 1), +((1.0 / 2.0) * C1 - (1.0 / 3.0) * P2) * invh2);
                put(&t, KV, ii, MU(r, c), +((1.0 / 2.0) * C2 - (1.0 / 3.0) * P1) * invh2);
                put(&t, KV, ii, MU(r + 1, c), -((1.0 / 2.0) * C2 - (1.0 / 3.0) * P2) * invh2);
                double rr = AT(fl->Drho, r, c - 1) * g0 +
                            (a2 * AT(fl->rhoWo, r, c - 1) + a3 * AT(fl->rhoWoo, r, c - 1)) / dt -
                            AT(adv, r - 1, c - 1);
                if (bnchm) rr += AT(bn[0], r, c);
                RR[ii] = rr;
            }
        free(f.p); free(um.p); free(wm.p); free(adv.p);
    }
    /* ---- U rows: boundaries (fluidmech.jl:252-267) */
    for (int64_t c = 0; c <= Nx; ++c) {
        put(&t, KV, MU(0, c), MU(0, c), 1.0);
        put(&t, KV, MU(0, c), MU(1, c), (double)o->top_cnv);
        put(&t, KV, MU(Nz + 1, c), MU(Nz + 1, c), 1.0);
        put(&t, KV, MU(Nz + 1, c), MU(Nz, c), (double)o->bot_cnv);
    }
    /* ---- U interior rows (fluidmech.jl:270-324) */
    {
        omat f = mk(Nz, Nx + 1), um = mk(Nz, Nx + 2), wm = mk(Nz + 1, Nx + 1), adv = mk(Nz, Nx + 1);
        for (int64_t i = 0; i <= Nx; ++i)
            for (int64_t j = 0; j < Nz; ++j) AT(f, j, i) = AT(rhou, j, i) * AT(U, j + 1, i);
        for (int64_t k = 0; k < Nx + 2; ++k)
            for (int64_t j = 0; j < Nz; ++j) AT(um, j, k) = (AT(U, j + 1, IFX(k)) + AT(U, j + 1, IFX(k + 1))) / 2.0;
        for (int64_t i = 0; i <= Nx; ++i)
            for (int64_t j = 0; j <= Nz; ++j) AT(wm, j, i) = (AT(W, j, i) + AT(W, j, i + 1)) / 2.0;
        int st = or_advect_centered(adv, NULL, NULL, f, um, wm, h, o->scheme, o->xbc, o->zbc);
        if (st) { free(tag); free(f.p); free(um.p); free(wm.p); free(adv.p); return st; }
        for (int64_t j = 0; j < Nz; ++j) {  /* fluidmech.jl:316-318 */
            double avg = (AT(adv, j, 0) + AT(adv, j, Nx)) / 2.0;
            AT(adv, j, 0) = avg;
            AT(adv, j, Nx) = avg;
        }
        for (int64_t c = 0; c <= Nx; ++c)
            for (int64_t r = 1; r <= Nz; ++r) {
                const int64_t ii = MU(r, c);
                const double C1 = AT(etaco, r - 1, c), C2 = AT(etaco, r, c);
                const double P1 = AT(eta, r - 1, ICX(c)), P2 = AT(eta, r - 1, ICX(c + 1));
                put(&t, KV, ii, ii, (a1 + gamma) * AT(rhou, r - 1, c) / dt);
                put(&t, KV, ii, ii, (2.0 / 3.0) * (P1 + P2) * invh2 + (1.0 / 2.0) * (C1 + C2) * invh2);
                put(&t, KV, ii, MU(r, IFX(c)), -(2.0 / 3.0) * P1 * invh2);
                put(&t, KV, ii, MU(r, IFX(c + 2)), -(2.0 / 3.0) * P2 * invh2);
                put(&t, KV, ii, MU(r - 1, IFX(c + 1)), -(1.0 / 2.0) * C1 * invh2);
                put(&t, KV, ii, MU(r + 1, IFX(c + 1)), -(1.0 / 2.0) * C2 * invh2);
                if (!bnchm) {
                    double ddx = (AT(rho, r - 1, ICX(c + 1)) - AT(rho, r - 1, ICX(c))) * invh;
                    put(&t, KV, ii, ii, ddx * g0 * dt);
                }
                put(&t, KV, ii, MW(r - 1, c), -((1.0 / 2.0) * C1 - (1.0 / 3.0) * P1) * invh2);
                put(&t, KV, ii, MW(r - 1, c + 1), +((1.0 / 2.0) * C1 - (1.0 / 3.0) * P2) * invh2);
                put(&t, KV, ii, MW(r, c), +((1.0 / 2.0) * C2 - (1.0 / 3.0) * P1) * invh2);
                put(&t, KV, ii, MW(r, c + 1), -((1.0 / 2.0) * C2 - (1.0 / 3.0) * P2) * invh2);
                double rr = (a2 * AT(fl->rhoUo, r - 1, c) + a3 * AT(fl->rhoUoo, r - 1, c)) / dt - AT(adv, r - 1, c);
                if (bnchm) rr += AT(bn[1], r, c);
                RR[ii] = rr;
            }
        free(f.p); free(um.p); free(wm.p); free(adv.p);
    }
    /* ---- GG (fluidmech.jl:343-356) */
    for (int64_t c = 1; c <= Nx; ++c)
        for (int64_t r = 1; r <= Nz - 1; ++r) {
            put(&t, GG, MW(r, c), MP(r, c), -invh);
            put(&t, GG, MW(r, c), MP(r + 1, c), +invh);
        }
    for (int64_t c = 0; c <= Nx; ++c)
        for (int64_t r = 1; r <= Nz; ++r) {
            put(&t, GG, MU(r, c), MP(r, c), -invh);
            put(&t, GG, MU(r, c), MP(r, c + 1), +invh);
        }
    /* ---- DM (fluidmech.jl:373-386) */
    for (int64_t c = 1; c <= Nx; ++c)
        for (int64_t r = 1; r <= Nz; ++r) {
            const int64_t ii = MP(r, c);
            put(&t, DM, ii, MU(r, c - 1), -AT(rhou, r - 1, c - 1) * invh);
            put(&t, DM, ii, MU(r, c), +AT(rhou, r - 1, c) * invh);
            put(&t, DM, ii, MW(r - 1, c), -AT(rhow, r - 1, c - 1) * invh);
            put(&t, DM, ii, MW(r, c), +AT(rhow, r, c - 1) * invh);
        }
    /* ---- KP boundary rows (fluidmech.jl:405-417) */
    for (int64_t c = 0; c < Nx + 2; ++c) {
        put(&t, KP, MP(0, c), MP(0, c), 1.0);
        put(&t, KP, MP(0, c), MP(1, c), -1.0);
    }
    for (int64_t c = 0; c < Nx + 2; ++c) {
        put(&t, KP, MP(Nz + 1, c), MP(Nz + 1, c), 1.0);
        put(&t, KP, MP(Nz + 1, c), MP(Nz, c), -1.0);
    }
    for (int64_t r = 1; r <= Nz; ++r) {
        put(&t, KP, MP(r, 0), MP(r, 0), 1.0);
        put(&t, KP, MP(r, 0), MP(r, Nx), -1.0);
    }
    for (int64_t r = 1; r <= Nz; ++r) {
        put(&t, KP, MP(r, Nx + 1), MP(r, Nx + 1), 1.0);
        put(&t, KP, MP(r, Nx + 1), MP(r, 1), -1.0);
    }
    /* ---- pressure RHS (fluidmech.jl:450-459) */
    for (int64_t c = 1; c <= Nx; ++c)
        for (int64_t r = 1; r <= Nz; ++r) {
            double v = AT(fl->MFS, r - 1, c - 1);
            if (bnchm) v += AT(bn[2], r, c);
            RR[MP(r, c)] = v;
        }
    if (t.n > cap) { free(tag); *nnz_out = t.n; return OR_EARG; }
    /* ---- pins (fluidmech.jl:462-488): drop the zeroed rows' entries, then set diagonals */
    int64_t pinP[4096];
    int npin = 0;
    int64_t nu0 = -1;
    if (bnchm) {
        int64_t nzp = jround((double)(Nz + 2) * 3.0 / 8.0) - 1, nxp = jround((double)(Nx + 2) / 2.0) - 1;
        pinP[npin++] = MP(nzp, nxp);
        RR[MP(nzp, nxp)] = AT(bn[5], nzp, nxp);
        int64_t nzu = jround((double)(Nz + 2) / 2.0) - 1, nxu = jround((double)(Nx + 2) / 2.0) - 1;
        nu0 = MU(nzu, nxu);
        RR[nu0] = AT(bn[4], nzu, nxu);
    } else if (o->open_cnv) {
        for (int64_t c = 0; c < Nx + 2 && npin < 4096; ++c) pinP[npin++] = MP(Nz, c);
    } else {
        pinP[npin++] = MP(jround((double)Nz / 2.0) - 1, jround((double)Nx / 2.0) - 1);
    }
    int64_t w = 0;
    for (int64_t k = 0; k < t.n; ++k) {
        int drop = 0;
        for (int q = 0; q < npin; ++q)
            if (I[k] == pinP[q] && (tag[k] == KP || (bnchm && tag[k] == DM))) drop = 1;
        if (bnchm && I[k] == nu0 && (tag[k] == KV || tag[k] == GG)) drop = 1;
        if (!drop) { I[w] = I[k]; J[w] = J[k]; V[w] = V[k]; ++w; }
    }
    for (int q = 0; q < npin; ++q) { I[w] = pinP[q]; J[w] = pinP[q]; V[w] = 1.0; ++w; }
    if (bnchm) { I[w] = nu0; J[w] = nu0; V[w] = 1.0; ++w; }
    *nnz_out = w;
    free(tag);
    return OR_OK;
#undef MW
#undef MU
#undef MP
#undef ICX
#undef IFX
}
