/*
 * or_noise.c -- oracle (test infrastructure only): restatement of
 * src/noise.jl (OU noise with the zero-padded Gaussian FFT filter) and of the
 * counter-based white-noise generator the HIP path uses in place of Julia's
 * MersenneTwister (noise.jl:154-157; that stream cannot be reproduced without
 * Julia).
 *
 * FFT: textbook recursive mixed-radix Cooley-Tukey (radix = smallest prime
 * factor, naive DFT for a prime length), used as the 2-D fft/ifft of
 * noise.jl:254-256.  Pinned against numpy.fft in the tests.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define AT(m, i, j) ((m).p[(int64_t)(i) + (int64_t)(j) * (m).ld])
#define EPS 2.220446049250313e-16
#define PI 3.14159265358979323846

/* ---- Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as 1, 2, 3") */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Standard normals for a column-major (rows x cols) matrix: element pair p =
 * e/2 of linear index e uses counter {p_lo, p_hi, draw, stream} and key {seed_lo, seed_hi};
 * two 53-bit uniforms u1 in (0,1], u2 in [0,1) feed Box-Muller. */
int or_normal_fill(omat a, uint64_t seed, uint32_t stream, uint32_t draw) {
    const int64_t n = a.rows * a.cols;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int64_t p = 0; p < (n + 1) / 2; ++p) {
        uint32_t ctr[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), draw, stream}, o[4];
        or_philox4x32_10(ctr, key, o);
        uint64_t a64 = ((uint64_t)o[1] << 32) | o[0], b64 = ((uint64_t)o[3] << 32) | o[2];
        double u1 = ((double)(a64 >> 11) + 1.0) * 0x1.0p-53;
        double u2 = (double)(b64 >> 11) * 0x1.0p-53;
        double r = sqrt(-2.0 * log(u1));
        double th = 2.0 * PI * u2;
        int64_t e0 = 2 * p, e1 = 2 * p + 1;
        AT(a, e0 % a.rows, e0 / a.rows) = r * cos(th);
        if (e1 < n) AT(a, e1 % a.rows, e1 / a.rows) = r * sin(th);
    }
    return OR_OK;
}

/* ---- FFT ------------------------------------------------------------------ */
static void fft_rec(double complex* x, int64_t n, int64_t stride, double complex* out, double complex* tmp, int sign) {
    if (n == 1) { out[0] = x[0]; return; }
    int64_t p = 2;
    while (p * p <= n && n % p) ++p;
    if (n % p) p = n;  /* prime length */
    const int64_t m = n / p;
    if (m == 1) {      /* naive DFT of prime length */
        for (int64_t k = 0; k < n; ++k) {
            double complex s = 0;
            for (int64_t j = 0; j < n; ++j) {
                double ang = sign * 2.0 * PI * (double)((j * k) % n) / (double)n;
                s += x[j * stride] * (cos(ang) + I * sin(ang));
            }
            out[k] = s;
        }
        return;
    }
    /* sub-transforms of the p decimated sequences */
    for (int64_t r = 0; r < p; ++r) fft_rec(x + r * stride, m, stride * p, tmp + r * m, out, sign);
    for (int64_t k = 0; k < m; ++k)
        for (int64_t q = 0; q < p; ++q) {
            double complex s = 0;
            for (int64_t r = 0; r < p; ++r) {
                int64_t t = (r * (k + q * m)) % n;
                double ang = sign * 2.0 * PI * (double)t / (double)n;
                s += tmp[r * m + k] * (cos(ang) + I * sin(ang));
            }
            out[k + q * m] = s;
        }
}

/* in-place 1-D transform of n values at stride `st`; sign -1 forward, +1 inverse (unnormalised) */
static void fft1(double complex* a, int64_t n, int64_t st, int sign, double complex* w1, double complex* w2,
                 double complex* w3) {
    for (int64_t i = 0; i < n; ++i) w1[i] = a[i * st];
    fft_rec(w1, n, 1, w2, w3, sign);
    for (int64_t i = 0; i < n; ++i) a[i * st] = w2[i];
}

/* 2-D fft (dims 1 and 2) of a column-major nz x nx complex array */
int or_fft2(double complex* a, int64_t nz, int64_t nx, int sign) {
    int64_t nmax = nz > nx ? nz : nx;
    double complex* w = (double complex*)malloc(sizeof(double complex) * 3 * (size_t)nmax);
    for (int64_t i = 0; i < nx; ++i) fft1(a + i * nz, nz, 1, sign, w, w + nmax, w + 2 * nmax);
    for (int64_t j = 0; j < nz; ++j) fft1(a + j, nx, nz, sign, w, w + nmax, w + 2 * nmax);
    free(w);
    if (sign > 0) {
        double s = 1.0 / (double)(nz * nx);
        for (int64_t k = 0; k < nz * nx; ++k) a[k] *= s;
    }
    return OR_OK;
}

/* AbstractFFTs fftfreq(n, fs)[j] = (j < ceil(n/2) ? j : j - n) * (fs/n) */
static double fftfreq(int64_t j, int64_t n, double fs) {
    int64_t nn = (n + 1) / 2;
    double mult = fs / (double)n;
    return (double)(j < nn ? j : j - n) * mult;
}

/* Statistics.mean / std (corrected) of a matrix */
static void mean_std(omat a, double* mean, double* sd) {
    const int64_t n = a.rows * a.cols;
    double s = 0.0;
    for (int64_t i = 0; i < a.cols; ++i)
        for (int64_t j = 0; j < a.rows; ++j) s += AT(a, j, i);
    double m = s / (double)n, q = 0.0;
    for (int64_t i = 0; i < a.cols; ++i)
        for (int64_t j = 0; j < a.rows; ++j) { double d = AT(a, j, i) - m; q += d * d; }
    *mean = m;
    *sd = sqrt(q / (double)(n - 1));
}

/* noise.jl:98-107 + 244-270: filter psi with exp(-sigma^2 (kz^2+kx^2)) on the
 * pad-row zero-padded grid, unpad, rescale to the raw mean/std.  out may alias nothing. */
int or_fft_filter(omat psi, omat out, int64_t pad, double h, double sigma) {
    const int64_t Nz = psi.rows, Nx = psi.cols, hp = pad / 2, Nzp = Nz + pad;
    double complex* a = (double complex*)calloc((size_t)(Nzp * Nx), sizeof(double complex));
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) a[(hp + j) + i * Nzp] = AT(psi, j, i);
    or_fft2(a, Nzp, Nx, -1);
    const double twopi = 2.0 * PI, s2 = sigma * sigma;
    for (int64_t i = 0; i < Nx; ++i) {
        double kx = twopi * fftfreq(i, Nx, 1.0 / h);
        for (int64_t j = 0; j < Nzp; ++j) {
            double kz = twopi * fftfreq(j, Nzp, 1.0 / h);
            a[j + i * Nzp] *= exp(-s2 * (kz * kz + kx * kx));
        }
    }
    or_fft2(a, Nzp, Nx, +1);
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) AT(out, j, i) = creal(a[(hp + j) + i * Nzp]);
    free(a);
    double rm, rs, fm, fs;
    mean_std(psi, &rm, &rs);
    mean_std(out, &fm, &fs);
    if (fs > EPS)
        for (int64_t i = 0; i < Nx; ++i)
            for (int64_t j = 0; j < Nz; ++j) AT(out, j, i) = (AT(out, j, i) - fm) * rs / fs + rm;
    return OR_OK;
}

static omat mk(int64_t r, int64_t c) {
    omat m; m.rows = r; m.cols = c; m.ld = r;
    m.p = (double*)calloc((size_t)(r * c), sizeof(double));
    return m;
}

/* noise.jl:138-239 (noise!), with re/rs drawn by or_normal_fill when first_iter */
int or_noise_step(or_noise* ns, const or_phase* ph, const or_grid* g, const or_par* par, const or_scales* sc,
             double dt, int first_iter, uint64_t seed, uint32_t draw) {
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const double Xi = par->Xi, L0h = sc->L0h, l0h = sc->l0h, h = g->h, txi0 = sc->txi0;
    if (first_iter) {
        or_normal_fill(ns->re, seed, 0u, draw);
        or_normal_fill(ns->rs, seed, 1u, draw);
    }
    const double fL = 2.0 / sqrt(1.0 - exp(-(h * h) / (2 * (L0h * L0h))));
    const double fl = 2.0 / sqrt(1.0 - exp(-(h * h) / (2 * (l0h * l0h))));
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double taue = L0h / (2.0 * (AT(ph->V, j, i) + EPS));
            double taus = l0h / (2.0 * (AT(ph->vx, j, i) + EPS));
            double St = txi0 / (taue + EPS);
            double fReL = AT(ph->fReL, j, i), ke = AT(ph->ke, j, i), chi = AT(ph->chi, j, i);
            double a = fReL * ke / taue;
            double sge = Xi * sqrt(a > 0.0 ? a : 0.0);
            double b = chi * fReL * ke / taue * St / (1.0 + St * St);
            double sgx = Xi * sqrt(b > 0.0 ? b : 0.0);
            double c = chi * AT(ph->ks, j, i) / taus;
            double sgs = Xi * sqrt(c > 0.0 ? c : 0.0);
            taue = 1.0 / (1.0 / taue + 1.0 / (1e2 * dt)) + dt;
            taus = 1.0 / (1.0 / taus + 1.0 / (1e2 * dt)) + dt;
            double Fte = exp(-dt / taue), Fts = exp(-dt / taus);
            double ge = 1.0 - Fte * Fte, gs = 1.0 - Fts * Fts;
            ge = ge > 0.0 ? ge : 0.0;
            gs = gs > 0.0 ? gs : 0.0;
            AT(ns->psie, j, i) = Fte * AT(ns->psieo, j, i) + sqrt(ge) * sge * fL * AT(ns->re, j, i);
            AT(ns->psix, j, i) = Fte * AT(ns->psixo, j, i) + sqrt(ge) * sgx * fL * AT(ns->re, j, i);
            AT(ns->psis, j, i) = Fts * AT(ns->psiso, j, i) + sqrt(gs) * sgs * fl * AT(ns->rs, j, i);
        }
    omat fe = mk(Nz, Nx), fx = mk(Nz, Nx), fs = mk(Nz, Nx);
    or_fft_filter(ns->psie, fe, ns->padL0, h, L0h);
    or_fft_filter(ns->psix, fx, ns->padL0, h, L0h);
    or_fft_filter(ns->psis, fs, ns->padl0, h, l0h);
    /* grid.icz = [1; 1:Nz; Nz] (closed), grid.icx = [Nx; 1:Nx; 1] (periodic): grid.jl:68-69 */
#define CZ(j) ((j) < 0 ? 0 : ((j) >= Nz ? Nz - 1 : (j)))
#define CX(i) ((i) < 0 ? Nx - 1 : ((i) >= Nx ? 0 : (i)))
#define EXT(m, j, i) AT(m, CZ((j) - 1), CX((i) - 1))
    /* psiec (Nz+1, Nx+1) corner averages (noise.jl:195-197) */
    omat pc = mk(Nz + 1, Nx + 1);
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j <= Nz; ++j)
            AT(pc, j, i) = (EXT(fe, j, i) + EXT(fe, j, i + 1) + EXT(fe, j + 1, i) + EXT(fe, j + 1, i + 1)) / 4.0;
    /* noise.jl:200-205 */
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz + 2; ++j) {
            int64_t z = CZ(j - 1);
            AT(ns->xieu, j, i) = AT(pc, z + 1, i) - AT(pc, z, i);
        }
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            int64_t x = CX(i - 1);
            AT(ns->xiew, j, i) = -(AT(pc, j, x + 1) - AT(pc, j, x));
        }
    /* noise.jl:208-220 */
    for (int64_t i = 0; i <= Nx; ++i)
        for (int64_t j = 0; j < Nz + 2; ++j) {
            double mu = AT(ph->muu, j, i) - 0.4;
            double tap = (1.0 - exp(-AT(ph->chiu, j, i) / 1e-4)) * (1.0 - exp(-(mu > 0.0 ? mu : 0.0) / 0.05));
            AT(ns->xixu, j, i) = (-(EXT(fx, j, i + 1) - EXT(fx, j, i))) * tap;
            AT(ns->xisu, j, i) = (-(EXT(fs, j, i + 1) - EXT(fs, j, i))) * tap;
        }
    for (int64_t i = 0; i < Nx + 2; ++i)
        for (int64_t j = 0; j <= Nz; ++j) {
            double mu = AT(ph->muw, j, i) - 0.4;
            double tap = (1.0 - exp(-AT(ph->chiw, j, i) / 1e-4)) * (1.0 - exp(-(mu > 0.0 ? mu : 0.0) / 0.05));
            AT(ns->xixw, j, i) = (-(EXT(fx, j + 1, i) - EXT(fx, j, i))) * tap;
            AT(ns->xisw, j, i) = (-(EXT(fs, j + 1, i) - EXT(fs, j, i))) * tap;
        }
    /* noise.jl:223-236 */
    for (int64_t i = 0; i < Nx; ++i)
        for (int64_t j = 0; j < Nz; ++j) {
            double wz = (AT(ns->xisw, j, i + 1) + AT(ns->xisw, j + 1, i + 1)) / 2.0;
            double wx = (AT(ns->xisu, j + 1, i) + AT(ns->xisu, j + 1, i + 1)) / 2.0;
            AT(ns->xis, j, i) = sqrt(wz * wz + wx * wx);
            wz = (AT(ns->xixw, j, i + 1) + AT(ns->xixw, j + 1, i + 1)) / 2.0;
            wx = (AT(ns->xixu, j + 1, i) + AT(ns->xixu, j + 1, i + 1)) / 2.0;
            AT(ns->xix, j, i) = sqrt(wz * wz + wx * wx);
            wz = (AT(ns->xiew, j, i + 1) + AT(ns->xiew, j + 1, i + 1)) / 2.0;
            wx = (AT(ns->xieu, j + 1, i) + AT(ns->xieu, j + 1, i + 1)) / 2.0;
            AT(ns->xie, j, i) = sqrt(wz * wz + wx * wx);
        }
#undef EXT
#undef CZ
#undef CX
    free(fe.p); free(fx.p); free(fs.p); free(pc.p);
    return OR_OK;
}

/* noise.jl:115-120 */
int or_store_noise(or_noise* ns) {
    for (int64_t i = 0; i < ns->psie.cols; ++i)
        for (int64_t j = 0; j < ns->psie.rows; ++j) {
            AT(ns->psieo, j, i) = AT(ns->psie, j, i);
            AT(ns->psixo, j, i) = AT(ns->psix, j, i);
            AT(ns->psiso, j, i) = AT(ns->psis, j, i);
        }
    return OR_OK;
}
