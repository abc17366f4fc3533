// xc_stokes.hip -- the Stokes saddle-point solve of src/fluidmech.jl on gfx950.
//
// The reference assembles LL = [KV GG; DM KP] as a sparse matrix and factors it
// with UMFPACK every Picard sweep (fluidmech.jl:329-522).  Here LL is never
// stored: one kernel evaluates any row of LL x from the coefficient fields
// (eta, etaco, rho, rhow, rhou) with exactly the reference's stencil, BC rows
// and pins, and the system is solved by right-preconditioned FGMRES.
//
// Preconditioner (applied to a residual (r_v, r_p)):
//   z_p = S^-1 r_p   Cahouet-Chabard form of the Schur complement
//                    S = KP - DM A^-1 GG with A = a1 rho / dt - div(eta grad):
//                    S^-1 ~ eta/rho (viscous part, diagonal; the reference's own
//                    scaling uses sqrt(eta/rho) there, fluidmech.jl:497-500)
//                         + T^-1 (inertial part, T = -DM (a1 rho/dt)^-1 GG, a
//                    constant-coefficient Laplacian solved exactly by FFT in x
//                    and a tridiagonal sweep in z); ghost rows solved exactly
//                    from KP;
//   z_v = A^-1 (r_v - GG z_p) by one geometric-multigrid V-cycle on the
//                    velocity block A = KV: coarse operators rediscretised from
//                    geometric-mean viscosities, staggered full-weighting
//                    restriction, bilinear prolongation, damped-Jacobi smoothing.
// All levels, vectors and the Krylov basis live in the context; nothing is
// allocated after the first call at a given grid size.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#include <vector>

#include "xc_fft.h"
#include "xc_internal.h"

namespace {

constexpr int NT = 256;

constexpr int XM_CH = 64;      // columns per partial x-sum of the horizontal-mean correction
constexpr int XM_PROBES = 6;   // comb probes of its Galerkin operator (W and P x 3 residues of z mod 3)

struct Flags {
    int bnchm, sds, top_cnv, bot_cnv, open_cnv;
    double a1, gamma, dt, g0;
    double a1_dt, a1g_dt;  // a1 / dt and (a1 + gamma) / dt, for the preconditioner's rows
    int64_t nu0;          // MMS U-pin velocity index (fine level) or -1
    int pin_kind;         // 1: single P pin (pr, pc); 2: open_cnv pins on P row Nz
    int pr, pc;
};

// Global element index of a 1-D launch of NT-lane workgroups, with the workgroups remapped so that each
// XCD (which receives workgroup b when b % 8 is its number) walks a contiguous run of elements: the
// neighbours a row reads, and the same elements in the next kernel (the next sweep of a level, the next
// pass over a vector), are then in that XCD's L2.
__device__ __forceinline__ int64_t gidx() {
    return (int64_t)xcd_tile(blockIdx.x, gridDim.x) * NT + threadIdx.x;
}

struct LevD {  // coefficients of one level, dense column-major (ld = rows)
    typedef double real;
    int Nz, Nx;
    double h, invh, invh2;
    const double *eta, *etaco, *rho, *rhow, *rhou;
    __device__ __forceinline__ int64_t nW() const { return (int64_t)(Nz + 1) * (Nx + 2); }
    __device__ __forceinline__ int64_t nU() const { return (int64_t)(Nz + 2) * (Nx + 1); }
    __device__ __forceinline__ int64_t iW(int r, int c) const { return r + (int64_t)c * (Nz + 1); }
    __device__ __forceinline__ int64_t iU(int r, int c) const { return nW() + r + (int64_t)c * (Nz + 2); }
    __device__ __forceinline__ int64_t iP(int r, int c) const { return nW() + nU() + r + (int64_t)c * (Nz + 2); }
    __device__ __forceinline__ double ETA(int r, int c) const { return eta[r + (int64_t)c * Nz]; }
    __device__ __forceinline__ double ETACO(int r, int c) const { return etaco[r + (int64_t)c * (Nz + 1)]; }
    __device__ __forceinline__ double RHO(int r, int c) const { return rho[r + (int64_t)c * Nz]; }
    __device__ __forceinline__ double RHOW(int r, int c) const { return rhow[r + (int64_t)c * (Nz + 1)]; }
    __device__ __forceinline__ double RHOU(int r, int c) const { return rhou[r + (int64_t)c * Nz]; }
    __device__ __forceinline__ int icx(int k) const { return k == 0 ? Nx - 1 : (k == Nx + 1 ? 0 : k - 1); }
    __device__ __forceinline__ int ifx(int k) const { return k == 0 ? Nx - 1 : (k == Nx + 2 ? 1 : k - 1); }
};

// the fine level's coefficients in single precision: the multigrid smoother's copy (the operator and the
// coarse levels keep double precision)
struct LevF {
    typedef float real;
    int Nz, Nx;
    double h, invh, invh2;
    const float *eta, *etaco, *rho, *rhow, *rhou;
    __device__ __forceinline__ float RHOU(int r, int c) const { return rhou[r + (int64_t)c * Nz]; }
};

// ---- rows of the velocity block KV (fluidmech.jl:145-324).  Return sum_j KV_ij x_j, diag in *dg.
// PRE: the multigrid smoother's form of the row (precomputed a1 / dt, products with 1/3 instead of
// quotients): the same values to rounding, without the divisions.  The operator itself (ll_row) keeps
// the reference's operations.
template <bool PRE = false>
__device__ double w_row(const LevD& L, const Flags& F, const double* x, int r, int c, double* dg, bool* bc) {
    const int Nz = L.Nz, Nx = L.Nx;
    *bc = true;
    if (c == 0) { *dg = 1.0; return x[L.iW(r, 0)] + F.sds * x[L.iW(r, Nx)]; }
    if (c == Nx + 1) { *dg = 1.0; return x[L.iW(r, Nx + 1)] + F.sds * x[L.iW(r, 1)]; }
    if (r == 0) { *dg = 1.0; return x[L.iW(0, c)]; }
    if (r == Nz) {
        const double oc = F.open_cnv ? 1.0 : 0.0;
        const double d = L.RHOW(Nz, c - 1) * L.invh;
        *dg = d;
        return d * x[L.iW(Nz, c)] - oc * L.RHOW(Nz - 1, c - 1) * L.invh * x[L.iW(Nz - 1, c)] +
               oc * L.RHOU(Nz - 1, c) * L.invh * x[L.iU(Nz, c)] - oc * L.RHOU(Nz - 1, c - 1) * L.invh * x[L.iU(Nz, c - 1)];
    }
    *bc = false;
    const double i2 = L.invh2;
    const double C1 = L.ETACO(r, c - 1), C2 = L.ETACO(r, c), P1 = L.ETA(r - 1, c - 1), P2 = L.ETA(r, c - 1);
    double d = (PRE ? L.RHOW(r, c - 1) * F.a1_dt : F.a1 * L.RHOW(r, c - 1) / F.dt) +
               ((2.0 / 3.0) * (P1 + P2) * i2 + 0.5 * (C1 + C2) * i2);
    if (!F.bnchm) d += (L.RHO(r, c - 1) - L.RHO(r - 1, c - 1)) * L.invh * F.g0 * F.dt;
    *dg = d;
    const double T1 = PRE ? P1 * (1.0 / 3.0) : P1 / 3.0, T2 = PRE ? P2 * (1.0 / 3.0) : P2 / 3.0;
    return d * x[L.iW(r, c)] - (2.0 / 3.0) * P1 * i2 * x[L.iW(r - 1, c)] - (2.0 / 3.0) * P2 * i2 * x[L.iW(r + 1, c)] -
           0.5 * C1 * i2 * x[L.iW(r, c - 1)] - 0.5 * C2 * i2 * x[L.iW(r, c + 1)] -
           (0.5 * C1 - T1) * i2 * x[L.iU(r, c - 1)] + (0.5 * C1 - T2) * i2 * x[L.iU(r + 1, c - 1)] +
           (0.5 * C2 - T1) * i2 * x[L.iU(r, c)] - (0.5 * C2 - T2) * i2 * x[L.iU(r + 1, c)];
}

template <bool PRE = false>
__device__ double u_row(const LevD& L, const Flags& F, const double* x, int r, int c, double* dg, bool* bc) {
    const int Nz = L.Nz;
    *bc = true;
    if (r == 0) { *dg = 1.0; return x[L.iU(0, c)] + F.top_cnv * x[L.iU(1, c)]; }
    if (r == Nz + 1) { *dg = 1.0; return x[L.iU(Nz + 1, c)] + F.bot_cnv * x[L.iU(Nz, c)]; }
    *bc = false;
    const double i2 = L.invh2;
    const double C1 = L.ETACO(r - 1, c), C2 = L.ETACO(r, c);
    const double P1 = L.ETA(r - 1, L.icx(c)), P2 = L.ETA(r - 1, L.icx(c + 1));
    double d = (PRE ? L.RHOU(r - 1, c) * F.a1g_dt : (F.a1 + F.gamma) * L.RHOU(r - 1, c) / F.dt) +
               ((2.0 / 3.0) * (P1 + P2) * i2 + 0.5 * (C1 + C2) * i2);
    if (!F.bnchm) d += (L.RHO(r - 1, L.icx(c + 1)) - L.RHO(r - 1, L.icx(c))) * L.invh * F.g0 * F.dt;
    *dg = d;
    const double T1 = PRE ? P1 * (1.0 / 3.0) : P1 / 3.0, T2 = PRE ? P2 * (1.0 / 3.0) : P2 / 3.0;
    return d * x[L.iU(r, c)] - (2.0 / 3.0) * P1 * i2 * x[L.iU(r, L.ifx(c))] -
           (2.0 / 3.0) * P2 * i2 * x[L.iU(r, L.ifx(c + 2))] - 0.5 * C1 * i2 * x[L.iU(r - 1, c)] -
           0.5 * C2 * i2 * x[L.iU(r + 1, c)] - (0.5 * C1 - T1) * i2 * x[L.iW(r - 1, c)] +
           (0.5 * C1 - T2) * i2 * x[L.iW(r - 1, c + 1)] + (0.5 * C2 - T1) * i2 * x[L.iW(r, c)] -
           (0.5 * C2 - T2) * i2 * x[L.iW(r, c + 1)];
}

// velocity row k of KV (index into the W|U vector): value and diagonal
template <bool PRE = false>
__device__ __forceinline__ double vel_row(const LevD& L, const Flags& F, const double* x, int64_t k, double* dg,
                                          bool* bc, bool fine) {
    if (fine && k == F.nu0) { *dg = 1.0; *bc = true; return x[k]; }
    const int64_t nW = L.nW();
    if (k < nW) {
        const int r = ((int)k % (L.Nz + 1)), c = ((int)k / (L.Nz + 1));
        return w_row<PRE>(L, F, x, r, c, dg, bc);
    }
    const int64_t q = k - nW;
    const int r = ((int)q % (L.Nz + 2)), c = ((int)q / (L.Nz + 2));
    return u_row<PRE>(L, F, x, r, c, dg, bc);
}

__device__ __forceinline__ bool is_pin(const Flags& F, int r, int c, int Nz) {
    if (F.pin_kind == 1) return r == F.pr && c == F.pc;
    if (F.pin_kind == 2) return r == Nz;
    return false;
}

// damped Jacobi on the velocity block: xo = x + w (b - A x) / diag (BC rows with weight 1)
// 1 / d for the smoother: hardware reciprocal refined by two Newton-Raphson steps (d: a nonzero normal
// diagonal entry)
__device__ __forceinline__ double recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(r, fma(-d, r, 1.0), r);
}
__device__ __forceinline__ void jacobi_at(const LevD& L, const Flags& F, const double* __restrict__ x,
                                          const double* __restrict__ b, double* __restrict__ xo, double omega,
                                          int fine, int zero_guess, int64_t k) {
    double dg;
    bool bc;
    if (zero_guess) {  // x == 0: A x == 0, diag from the row formula evaluated on x (all zeros)
        vel_row<true>(L, F, x, k, &dg, &bc, fine != 0);
        xo[k] = (bc ? 1.0 : omega) * b[k] * recip(dg);
        return;
    }
    const double ax = vel_row<true>(L, F, x, k, &dg, &bc, fine != 0);
    xo[k] = x[k] + (bc ? 1.0 : omega) * (b[k] - ax) * recip(dg);
}
__device__ __forceinline__ void vel_resid_at(const LevD& L, const Flags& F, const double* __restrict__ x,
                                             const double* __restrict__ b, double* __restrict__ r, int fine,
                                             int64_t k) {
    double dg;
    bool bc;
    r[k] = b[k] - vel_row<true>(L, F, x, k, &dg, &bc, fine != 0);
}
// ---- the smoother and the level residual on the grid of cells: one lane per (r, c), r in [0, Nz+1],
// c in [0, Nx+1], computing the W row (r, c) (r <= Nz) and the U row (r, c) (c <= Nx) together.  A wave
// holds 64 consecutive rows of one column, so every column offset is a scalar and every operand is a
// scalar base plus the lane's row offset (clamped into the array); the rows are those of vel_row.
// LV: the level's coefficients (LevD, or LevF: single precision, arithmetic in single precision).
// MODE 0: damped Jacobi sweep out = x + w (b - A x) / diag (BC rows weight 1); 1: the same from x = 0
// (x is not read); 2: residual out = b - A x; 3: the first sweep of the fine level from x = 0 with the
// right-hand side b = r_v - GG z_p formed here from the Krylov vector's velocity part (b) and pressure
// part (zP), and stored to bout; 4: the first sweep of a coarse level from x = 0 with its right-hand side
// restricted here from the finer level's residual (b, of that level's size), and stored to bout.
template <typename T>
__device__ __forceinline__ T ldr(const T* __restrict__ col, int row) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(col) + (size_t)((unsigned)row * (unsigned)sizeof(T)));
}
__device__ __forceinline__ float recip(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    return fmaf(r, fmaf(-d, r, 1.0f), r);
}
template <int MODE, class LV, typename TX, typename TB, typename TO>
__global__ __launch_bounds__(NT) void k_smooth_cells(LV L, Flags F, const TX* __restrict__ x,
                                                     const TB* __restrict__ b, TO* out, double omega,
                                                     int fine, const double* __restrict__ zP, TO* __restrict__ bout,
                                                     const TX* prev = nullptr, double alpha = 0.0) {
    using R = typename LV::real;
    const int Nz = L.Nz, Nx = L.Nx;
    const int segs = (Nz + 2 + 63) >> 6;
    const int task = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * (NT / 64) + (threadIdx.x >> 6));
    const int c = task / segs;
    if (c > Nx + 1) return;
    const int lane = threadIdx.x & 63;
    const int r = ((task - c * segs) << 6) + lane;
    const bool live = r <= Nz + 1;
    const int ru = r <= Nz + 1 ? r : Nz + 1;   // U row
    const int rw = ru <= Nz ? ru : Nz;         // W row
    const int um = ru > 0 ? ru - 1 : 0, up = ru <= Nz ? ru + 1 : Nz + 1;
    const int wm = rw > 0 ? rw - 1 : 0, wp = rw < Nz ? rw + 1 : Nz;
    const int em = rw > 0 ? (rw - 1 < Nz - 1 ? rw - 1 : Nz - 1) : 0;  // cell row r - 1
    const int e0 = rw < Nz - 1 ? rw : Nz - 1;                          // cell row r
    const int ldW = Nz + 1, ldU = Nz + 2;
    // columns (wave-uniform)
    const int cw = c;                                        // W column, [0, Nx+1]
    const int cwm = c > 0 ? c - 1 : 0, cwp = c <= Nx ? c + 1 : Nx + 1;
    const int cu = c <= Nx ? c : Nx;                         // U column, [0, Nx]
    const int cul = c == 0 ? Nx - 1 : (c - 1 <= Nx ? c - 1 : Nx);   // periodic left U column / W row's c - 1
    const int cur = c + 1 >= Nx + 1 ? (c + 1 == Nx + 1 ? 1 : Nx) : c + 1;  // periodic right U column (c + 2 -> 1)
    const int cl = c == 0 ? Nx - 1 : (c - 1 < Nx ? c - 1 : Nx - 1);  // cell column left of the face
    const int cr = c < Nx ? c : 0;                                    // cell column right of the U face
    const int ceco = c <= Nx ? c : Nx, cecm = c > 0 ? (c - 1 <= Nx ? c - 1 : Nx) : 0;
    const int64_t offU = (int64_t)ldW * (Nx + 2);
    // operands
    const TX* Wc = x + (int64_t)cw * ldW;
    const TX* Uc = x + offU + (int64_t)cu * ldU;
    R Wr = 0, Wn = 0, Ws = 0, Ww = 0, We = 0, Wen = 0, Wg = 0;
    R Ur = 0, Un = 0, Us = 0, Ul = 0, Uls = 0, Urt = 0;
    if (MODE == 0 || MODE == 2) {
        Wr = (R)ldr(Wc, rw); Wn = (R)ldr(Wc, wm); Ws = (R)ldr(Wc, wp);
        Ww = (R)ldr(x + (int64_t)cwm * ldW, rw);
        const TX* Wpc = x + (int64_t)cwp * ldW;
        We = (R)ldr(Wpc, rw); Wen = (R)ldr(Wpc, wm);
        Ur = (R)ldr(Uc, ru); Un = (R)ldr(Uc, um); Us = (R)ldr(Uc, up);
        const TX* Ulc = x + offU + (int64_t)cul * ldU;
        Ul = (R)ldr(Ulc, ru); Uls = (R)ldr(Ulc, up);
        Urt = (R)ldr(x + offU + (int64_t)cur * ldU, ru);
        if (c == 0 || c == Nx + 1) Wg = (R)ldr(x + (int64_t)(c == 0 ? Nx : 1) * ldW, rw);
    }
    double bw, bu;
    if (MODE == 4) {  // b = the full-weighting restriction of the finer level's residual (restrict_at)
        const int fz = 2 * Nz, fx = 2 * Nx, ldWf = fz + 1, ldUf = fz + 2;
        const int64_t offUf = (int64_t)ldWf * (fx + 2);
        const int rr = rw < 1 ? 1 : (rw > Nz - 1 ? Nz - 1 : rw), cc = c < 1 ? 1 : (c > Nx ? Nx : c);
        const TB* F0 = b + (int64_t)(2 * cc - 1) * ldWf;
        const TB* F1 = b + (int64_t)(2 * cc) * ldWf;
        const double vw = (2.0 * (double)ldr(F0, 2 * rr) + 2.0 * (double)ldr(F1, 2 * rr) + (double)ldr(F0, 2 * rr - 1) +
                           (double)ldr(F1, 2 * rr - 1) + (double)ldr(F0, 2 * rr + 1) + (double)ldr(F1, 2 * rr + 1)) *
                          0.125;
        bw = r >= 1 && r <= Nz - 1 && c >= 1 && c <= Nx ? vw : 0.0;
        const int ur = ru < 1 ? 1 : (ru > Nz ? Nz : ru);
        const int f = 2 * cu, fl = f == 0 ? fx - 1 : f - 1, fr = f == fx ? 1 : f + 1;
        const TB* G0 = b + offUf + (int64_t)f * ldUf;
        const TB* Gl = b + offUf + (int64_t)fl * ldUf;
        const TB* Gr = b + offUf + (int64_t)fr * ldUf;
        const double vu = (2.0 * (double)ldr(G0, 2 * ur - 1) + 2.0 * (double)ldr(G0, 2 * ur) + (double)ldr(Gl, 2 * ur - 1) +
                           (double)ldr(Gl, 2 * ur) + (double)ldr(Gr, 2 * ur - 1) + (double)ldr(Gr, 2 * ur)) *
                          0.125;
        bu = r >= 1 && r <= Nz ? vu : 0.0;
    } else {
        bw = (double)ldr(b + (int64_t)cw * ldW, rw);
        bu = (double)ldr(b + offU + (int64_t)cu * ldU, ru);
    }
    if (MODE == 3) {  // b = r_v - GG z_p (fluidmech.jl:343-356): interior W and U rows, not at the MMS pin
        const int ldP = Nz + 2;
        const double* Pc = zP + (int64_t)cu * ldP;
        const double p0 = ldr(Pc, ru), p1 = ldr(Pc, ru <= Nz ? ru + 1 : Nz + 1);
        const double pe = ldr(zP + (int64_t)(cu + 1) * ldP, ru);
        const double ih = L.invh;
        const int64_t kw = rw + (int64_t)cw * ldW, ku = offU + ru + (int64_t)cu * ldU;
        if (r >= 1 && r <= Nz - 1 && c >= 1 && c <= Nx && kw != F.nu0) bw -= -ih * p0 + ih * p1;
        if (r >= 1 && r <= Nz && c <= Nx && ku != F.nu0) bu -= -ih * p0 + ih * pe;
    }
    const auto* ECm = L.etaco + (int64_t)cecm * (Nz + 1);
    const auto* EC0 = L.etaco + (int64_t)ceco * (Nz + 1);
    const auto* El = L.eta + (int64_t)cl * Nz;
    const auto* Er = L.eta + (int64_t)cr * Nz;
    const auto* Rl = L.rho + (int64_t)cl * Nz;
    const auto* Rr = L.rho + (int64_t)cr * Nz;
    const auto* RWc = L.rhow + (int64_t)cl * (Nz + 1);
    const auto* RUc = L.rhou + (int64_t)cu * Nz;
    const R eC1w = ldr(ECm, rw), eC2w = ldr(EC0, rw), eC1u = ldr(EC0, wm);  // etaco(r, c-1), (r, c), (r-1, c)
    const R eP1w = ldr(El, em), eP2w = ldr(El, e0), eP2u = ldr(Er, em);    // eta(r-1, cl), (r, cl), (r-1, cr)
    const R rl0 = ldr(Rl, e0), rlm = ldr(Rl, em), rrm = ldr(Rr, em);      // rho(r, cl), (r-1, cl), (r-1, cr)
    const R rw0 = ldr(RWc, rw), rwm = ldr(RWc, wm);                       // rhow(r, c-1), (r-1, c-1)
    const R ru0 = ldr(RUc, em);                                           // rhou(r-1, c)
    const R i2 = (R)L.invh2, ih = (R)L.invh;
    const R gz = F.bnchm ? (R)0 : (R)(L.invh * F.g0 * F.dt);
    const R a1w = (R)F.a1_dt, a1u = (R)F.a1g_dt, two3 = (R)(2.0 / 3.0), third = (R)(1.0 / 3.0), half = (R)0.5;
    // W row (r, c)
    R dw, vw;
    {
        const R C1 = eC1w, C2 = eC2w, P1 = eP1w, P2 = eP2w;
        dw = rw0 * a1w + (two3 * (P1 + P2) * i2 + half * (C1 + C2) * i2) + (rl0 - rlm) * gz;
        const R T1 = P1 * third, T2 = P2 * third;
        vw = dw * Wr - two3 * P1 * i2 * Wn - two3 * P2 * i2 * Ws - half * C1 * i2 * Ww - half * C2 * i2 * We -
             (half * C1 - T1) * i2 * Ul + (half * C1 - T2) * i2 * Uls + (half * C2 - T1) * i2 * Ur -
             (half * C2 - T2) * i2 * Us;
    }
    const bool wghost = c == 0 || c == Nx + 1;
    const bool wtop = !wghost && r == 0, wbot = !wghost && r == Nz;
    if (wbot) {
        dw = rw0 * ih;
        if (F.open_cnv) {
            const R RU0 = L.RHOU(Nz - 1, cul), RU1 = L.RHOU(Nz - 1, cu);
            vw = dw * Wr - rwm * ih * Wn + RU1 * ih * Ur - RU0 * ih * Ul;
        } else {
            vw = dw * Wr;
        }
    }
    if (wtop) { dw = 1; vw = Wr; }
    if (wghost) { dw = 1; vw = Wr + (R)F.sds * Wg; }
    bool bcw = wghost || wtop || wbot;
    // U row (r, c)
    R du, vu;
    {
        const R C1 = eC1u, C2 = eC2w, P1 = eP1w, P2 = eP2u;
        du = ru0 * a1u + (two3 * (P1 + P2) * i2 + half * (C1 + C2) * i2) + (rrm - rlm) * gz;
        const R T1 = P1 * third, T2 = P2 * third;
        vu = du * Ur - two3 * P1 * i2 * Ul - two3 * P2 * i2 * Urt - half * C1 * i2 * Un - half * C2 * i2 * Us -
             (half * C1 - T1) * i2 * Wn + (half * C1 - T2) * i2 * Wen + (half * C2 - T1) * i2 * Wr -
             (half * C2 - T2) * i2 * We;
    }
    const bool utop = r == 0, ubot = r == Nz + 1;
    if (utop) { du = 1; vu = Ur + (R)F.top_cnv * Us; }
    if (ubot) { du = 1; vu = Ur + (R)F.bot_cnv * Un; }
    bool bcu = utop || ubot;
    const int64_t kw = rw + (int64_t)cw * ldW, ku = offU + ru + (int64_t)cu * ldU;
    if (fine && kw == F.nu0) { dw = 1; vw = Wr; bcw = true; }
    if (fine && ku == F.nu0) { du = 1; vu = Ur; bcu = true; }
    if (!live) return;
    const R om = (R)omega;
    R ow, ou;
    if (MODE == 2) {
        ow = (R)bw - vw;
        ou = (R)bu - vu;
    } else if (MODE == 1 || MODE == 3 || MODE == 4) {
        ow = (bcw ? (R)1 : om) * (R)bw * recip(dw);
        ou = (bcu ? (R)1 : om) * (R)bu * recip(du);
    } else if (MODE == 0 && alpha != 0.0) {
        // Chebyshev step (BC rows as in a Jacobi sweep of weight 1): x + alpha (x - x_prev) + omega D^-1 (b - A x),
        // the previous iterate read at the lane's own rows only (it may be the output buffer); prev == nullptr:
        // the previous iterate is zero
        const R al = (R)alpha;
        const R pw = prev ? (R)prev[kw] : (R)0, pu = prev ? (R)prev[ku] : (R)0;
        const R gw = ((R)bw - vw) * recip(dw), gu = ((R)bu - vu) * recip(du);
        ow = bcw ? Wr + gw : Wr + al * (Wr - pw) + om * gw;
        ou = bcu ? Ur + gu : Ur + al * (Ur - pu) + om * gu;
    } else {
        ow = Wr + (bcw ? (R)1 : om) * ((R)bw - vw) * recip(dw);
        ou = Ur + (bcu ? (R)1 : om) * ((R)bu - vu) * recip(du);
    }
    if (r <= Nz) {
        out[kw] = (TO)ow;
        if (MODE >= 3) bout[kw] = (TO)bw;
    }
    if (c <= Nx) {
        out[ku] = (TO)ou;
        if (MODE >= 3) bout[ku] = (TO)bu;
    }
}
// ---- LL x (RESID: b - LL x) on the grid of cells: one lane per (r, c), r in [0, Nz+1], c in [0, Nx+1],
// computing the W row (r <= Nz), the U row (c <= Nx) and the P row of that position with the operations
// of the reference's rows (fluidmech.jl:145-488) in the order vel_row and the pressure rows use, from
// scalar column bases plus the lane's row offset.
// sq: when given, wave w (of 4 gridDim.x) writes the sum of squares of the values its lanes stored to sq[w]
// bq (RESID): likewise the sums of squares of the b entries of those rows (||b||^2 over the whole vector)
template <bool RESID>
__global__ __launch_bounds__(NT) void k_ll_cells(LevD L, Flags F, const double* __restrict__ x,
                                                 const double* __restrict__ b, double* __restrict__ y,
                                                 double* __restrict__ sq = nullptr, double* __restrict__ pinc = nullptr,
                                                 double* __restrict__ bq = nullptr) {
    const int Nz = L.Nz, Nx = L.Nx;
    const int segs = (Nz + 2 + 63) >> 6;
    const int task = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * (NT / 64) + (threadIdx.x >> 6));
    const int c = task / segs;
    const int lane = threadIdx.x & 63;
    if (c > Nx + 1) {
        if (sq && lane == 0) sq[task] = 0.0;
        if (bq && lane == 0) bq[task] = 0.0;
        return;
    }
    const int r0 = ((task - c * segs) << 6) + lane;
    const bool live = r0 <= Nz + 1;
    const int r = live ? r0 : Nz + 1;  // lanes past the last row repeat it and store nothing
    double ss = 0.0, sb = 0.0;
    const int ru = r;                            // U and P row
    const int rw = r <= Nz ? r : Nz;             // W row
    const int um = ru > 0 ? ru - 1 : 0, up = ru <= Nz ? ru + 1 : Nz + 1;
    const int wm = rw > 0 ? rw - 1 : 0, wp = rw < Nz ? rw + 1 : Nz;
    const int em = rw > 0 ? (rw - 1 < Nz - 1 ? rw - 1 : Nz - 1) : 0;
    const int e0 = rw < Nz - 1 ? rw : Nz - 1;
    const int ldW = Nz + 1, ldU = Nz + 2, ldP = Nz + 2;
    const int cw = c;
    const int cwm = c > 0 ? c - 1 : 0, cwp = c <= Nx ? c + 1 : Nx + 1;
    const int cu = c <= Nx ? c : Nx;
    const int cul = c == 0 ? Nx - 1 : (c - 1 <= Nx ? c - 1 : Nx);
    const int cur = c + 1 >= Nx + 1 ? (c + 1 == Nx + 1 ? 1 : Nx) : c + 1;
    const int cl = c == 0 ? Nx - 1 : (c - 1 < Nx ? c - 1 : Nx - 1);
    const int cr = c < Nx ? c : 0;
    const int ceco = c <= Nx ? c : Nx, cecm = c > 0 ? (c - 1 <= Nx ? c - 1 : Nx) : 0;
    const int64_t offU = (int64_t)ldW * (Nx + 2), offP = offU + (int64_t)ldU * (Nx + 1);
    const double* xU = x + offU;
    const double* xP = x + offP;
    const double* Wc = x + (int64_t)cw * ldW;
    const double* Uc = xU + (int64_t)cu * ldU;
    const double* Pc = xP + (int64_t)c * ldP;
    const double Wr = ldr(Wc, rw), Wn = ldr(Wc, wm), Ws = ldr(Wc, wp);
    const double Ww = ldr(x + (int64_t)cwm * ldW, rw);
    const double* Wpc = x + (int64_t)cwp * ldW;
    const double We = ldr(Wpc, rw), Wen = ldr(Wpc, wm);
    const double Ur = ldr(Uc, ru), Un = ldr(Uc, um), Us = ldr(Uc, up);
    const double* Ulc = xU + (int64_t)cul * ldU;
    const double Ul = ldr(Ulc, ru), Uls = ldr(Ulc, up);
    const double Urt = ldr(xU + (int64_t)cur * ldU, ru);
    const double Pr = ldr(Pc, ru), Pn = ldr(Pc, um), Ps = ldr(Pc, up);
    const double Pe = ldr(xP + (int64_t)(cu + 1) * ldP, ru);
    double Wg = 0.0, Pg = 0.0;
    if (c == 0 || c == Nx + 1) {
        Wg = ldr(x + (int64_t)(c == 0 ? Nx : 1) * ldW, rw);
        Pg = ldr(xP + (int64_t)(c == 0 ? Nx : 1) * ldP, ru);
    }
    const double* EC0 = L.etaco + (int64_t)ceco * (Nz + 1);
    const double eC1w = ldr(L.etaco + (int64_t)cecm * (Nz + 1), rw), eC2w = ldr(EC0, rw), eC1u = ldr(EC0, wm);
    const double* El = L.eta + (int64_t)cl * Nz;
    const double eP1w = ldr(El, em), eP2w = ldr(El, e0), eP2u = ldr(L.eta + (int64_t)cr * Nz, em);
    const double* Rl = L.rho + (int64_t)cl * Nz;
    const double rl0 = ldr(Rl, e0), rlm = ldr(Rl, em), rrm = ldr(L.rho + (int64_t)cr * Nz, em);
    const double* RWc = L.rhow + (int64_t)cl * (Nz + 1);
    const double rw0 = ldr(RWc, rw), rwm = ldr(RWc, wm);
    const double ru0 = ldr(L.rhou + (int64_t)cu * Nz, em), rul = ldr(L.rhou + (int64_t)cul * Nz, em);
    const double i2 = L.invh2, ih = L.invh;
    const int64_t kw = rw + (int64_t)cw * ldW, ku = offU + ru + (int64_t)cu * ldU, kp = offP + ru + (int64_t)c * ldP;
    // W row (r, c) (w_row), then GG on the interior rows (fluidmech.jl:343-356)
    if (r <= Nz) {
        double v;
        bool bc = true;
        if (kw == F.nu0) v = Wr;
        else if (c == 0 || c == Nx + 1) v = Wr + F.sds * Wg;
        else if (r == 0) v = Wr;
        else if (r == Nz) {
            const double oc = F.open_cnv ? 1.0 : 0.0;
            const double d = rw0 * ih;
            v = d * Wr - oc * rwm * ih * Wn + oc * ru0 * ih * Ur - oc * rul * ih * Ul;
        } else {
            bc = false;
            const double C1 = eC1w, C2 = eC2w, P1 = eP1w, P2 = eP2w;
            double d = F.a1 * rw0 / F.dt + ((2.0 / 3.0) * (P1 + P2) * i2 + 0.5 * (C1 + C2) * i2);
            if (!F.bnchm) d += (rl0 - rlm) * ih * F.g0 * F.dt;
            const double T1 = P1 / 3.0, T2 = P2 / 3.0;
            v = d * Wr - (2.0 / 3.0) * P1 * i2 * Wn - (2.0 / 3.0) * P2 * i2 * Ws - 0.5 * C1 * i2 * Ww -
                0.5 * C2 * i2 * We - (0.5 * C1 - T1) * i2 * Ul + (0.5 * C1 - T2) * i2 * Uls +
                (0.5 * C2 - T1) * i2 * Ur - (0.5 * C2 - T2) * i2 * Us;
        }
        if (!bc) v += -ih * Pr + ih * Ps;
        const double bv = RESID ? ldr(b + (int64_t)cw * ldW, rw) : 0.0;
        const double o = RESID ? bv - v : v;
        if (live) {
            y[kw] = o;
            ss = fma(o, o, ss);
            if (RESID) sb = fma(bv, bv, sb);
        }
    }
    // U row (r, c) (u_row), then GG on the interior rows
    if (c <= Nx) {
        double v;
        bool bc = true;
        if (ku == F.nu0) v = Ur;
        else if (r == 0) v = Ur + F.top_cnv * Us;
        else if (r == Nz + 1) v = Ur + F.bot_cnv * Un;
        else {
            bc = false;
            const double C1 = eC1u, C2 = eC2w, P1 = eP1w, P2 = eP2u;
            double d = (F.a1 + F.gamma) * ru0 / F.dt + ((2.0 / 3.0) * (P1 + P2) * i2 + 0.5 * (C1 + C2) * i2);
            if (!F.bnchm) d += (rrm - rlm) * ih * F.g0 * F.dt;
            const double T1 = P1 / 3.0, T2 = P2 / 3.0;
            v = d * Ur - (2.0 / 3.0) * P1 * i2 * Ul - (2.0 / 3.0) * P2 * i2 * Urt - 0.5 * C1 * i2 * Un -
                0.5 * C2 * i2 * Us - (0.5 * C1 - T1) * i2 * Wn + (0.5 * C1 - T2) * i2 * Wen +
                (0.5 * C2 - T1) * i2 * Wr - (0.5 * C2 - T2) * i2 * We;
        }
        if (!bc) v += -ih * Pr + ih * Pe;
        const double bv = RESID ? ldr(b + offU + (int64_t)cu * ldU, ru) : 0.0;
        const double o = RESID ? bv - v : v;
        if (live) {
            y[ku] = o;
            ss = fma(o, o, ss);
            if (RESID) sb = fma(bv, bv, sb);
        }
    }
    // P row (r, c): DM, the BC rows and the pins (fluidmech.jl:357-488)
    {
        double v;
        const bool pin = is_pin(F, r, c, Nz);
        const bool interior = r >= 1 && r <= Nz && c >= 1 && c <= Nx;
        if (pin && (F.bnchm || !interior)) v = Pr;
        else if (r == 0) v = Pr - Ps;
        else if (r == Nz + 1) v = Pr - Pn;
        else if (c == 0 || c == Nx + 1) v = Pr - Pg;
        else {
            v = -rul * ih * Ul + ru0 * ih * Ur - rwm * ih * Wn + rw0 * ih * Wr;
            if (pin) v += Pr;
        }
        const double bv = RESID ? ldr(b + offP + (int64_t)c * ldP, ru) : 0.0;
        double o = RESID ? bv - v : v;
        if (RESID && live) sb = fma(bv, bv, sb);
        if (!RESID && pinc && pin && live) {
            // constant-pressure fix of an FGMRES step (b = the step's Krylov vector): LL 1_P = e_pin, so with
            // c = b_pin - (LL x)_pin the step uses x + c 1_P, whose image has b's value in the pin row
            const double bp = b[kp];
            *pinc = bp - v;
            o = bp;
        }
        if (live) {
            y[kp] = o;
            ss = fma(o, o, ss);
        }
    }
    if (sq) {
        const double t = wave_total(ss);
        if (lane == 0) sq[task] = t;
    }
    if (RESID && bq) {
        const double t = wave_total(sb);
        if (lane == 0) bq[task] = t;
    }
}
inline unsigned ll_cells_blocks(int Nz, int Nx) {
    const int64_t tasks = (int64_t)(Nx + 2) * ((Nz + 2 + 63) / 64);
    return (unsigned)((tasks + NT / 64 - 1) / (NT / 64));
}

// bilinear prolongation of the coarse correction ec added to the fine iterate xf (prolong_at), on the fine
// level's grid of cells: the coarse columns are wave-uniform, the coarse rows per lane
template <typename TF>
__global__ __launch_bounds__(NT) void k_prolong_cells(int Nz, int Nx, Flags F, const double* __restrict__ ec,
                                                      TF* __restrict__ xf) {
    const int segs = (Nz + 2 + 63) >> 6;
    const int task = __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * (NT / 64) + (threadIdx.x >> 6));
    const int c = task / segs;
    if (c > Nx + 1) return;
    const int r = ((task - c * segs) << 6) + (threadIdx.x & 63);
    if (r > Nz + 1) return;
    const int Zc = Nz / 2, Xc = Nx / 2, ldWc = Zc + 1, ldUc = Zc + 2;
    const int64_t offUc = (int64_t)ldWc * (Xc + 2), offUf = (int64_t)(Nz + 1) * (Nx + 2);
    if (r >= 1 && r <= Nz - 1 && c >= 1 && c <= Nx) {  // W (r, c)
        // coarse W column value with the periodic/mirror ghost map of the BC rows (wave-uniform columns)
        const int I = (c + 1) / 2, In = (c & 1) ? I - 1 : I + 1;
        const int cI = I == 0 ? Xc : (I == Xc + 1 ? 1 : I), cN = In == 0 ? Xc : (In == Xc + 1 ? 1 : In);
        const double fI = (I == 0 || I == Xc + 1) ? -F.sds : 1.0, fN = (In == 0 || In == Xc + 1) ? -F.sds : 1.0;
        const double* eI = ec + (int64_t)cI * ldWc;
        const double* eN = ec + (int64_t)cN * ldWc;
        auto EX = [&](int J) -> double {  // x-interpolation at fine column c, coarse row J (0 outside 1..Zc-1)
            const int Jc = J < 0 ? 0 : (J > Zc ? Zc : J);
            const double a = fI * ldr(eI, Jc), b = fN * ldr(eN, Jc);
            return J <= 0 || J >= Zc ? 0.0 : 0.75 * a + 0.25 * b;
        };
        const int J = r / 2;
        const double v = (r & 1) ? 0.5 * (EX(J) + EX(J + 1)) : EX(J);
        const int64_t k = r + (int64_t)c * (Nz + 1);
        xf[k] += v;
    }
    if (r >= 1 && r <= Nz && c <= Nx) {  // U (r, c)
        const int I = c / 2;
        const double* u0 = ec + offUc + (int64_t)I * ldUc;
        const double* u1 = ec + offUc + (int64_t)(I + 1 <= Xc ? I + 1 : Xc) * ldUc;
        auto E = [&](const double* col, int J) -> double {
            if (J == 0) return -F.top_cnv * ldr(col, 1);
            if (J == Zc + 1) return -F.bot_cnv * ldr(col, Zc);
            return ldr(col, J);
        };
        const int J = (r + 1) / 2, Jn = (r & 1) ? J - 1 : J + 1;
        auto EZ = [&](const double* col) -> double { return 0.75 * E(col, J) + 0.25 * E(col, Jn); };
        const double v = (c & 1) ? 0.5 * (EZ(u0) + EZ(u1)) : EZ(u0);
        const int64_t k = offUf + r + (int64_t)c * (Nz + 2);
        xf[k] += v;
    }
}

#define SMOOTH_D(M) (k_smooth_cells<M, LevD, double, double, double>)
#define SMOOTH_F(M, TB, TO) (k_smooth_cells<M, LevF, float, TB, TO>)
// single -> double precision copy
__global__ __launch_bounds__(NT) void k_f2d(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = gidx();
    if (i < n) b[i] = (double)a[i];
}
inline unsigned smooth_cells_blocks(int Nz, int Nx) {
    const int64_t tasks = (int64_t)(Nx + 2) * ((Nz + 2 + 63) / 64);
    return (unsigned)((tasks + NT / 64 - 1) / (NT / 64));
}

// staggered full-weighting restriction of a velocity residual rf (level f) into bc (level c)
template <typename TF>
__device__ __forceinline__ void restrict_at(const LevD& Lf, const LevD& Lc, const TF* __restrict__ rf,
                                            double* __restrict__ bcv, int64_t k) {
    const int64_t nWc = Lc.nW();
    double v = 0.0;
    if (k < nWc) {
        const int J = ((int)k % (Lc.Nz + 1)), I = ((int)k / (Lc.Nz + 1));
        if (J >= 1 && J <= Lc.Nz - 1 && I >= 1 && I <= Lc.Nx) {
            const int r = 2 * J, c0 = 2 * I - 1, c1 = 2 * I;
            v = (2.0 * (double)rf[Lf.iW(r, c0)] + 2.0 * (double)rf[Lf.iW(r, c1)] + (double)rf[Lf.iW(r - 1, c0)] +
                 (double)rf[Lf.iW(r - 1, c1)] + (double)rf[Lf.iW(r + 1, c0)] + (double)rf[Lf.iW(r + 1, c1)]) * 0.125;
        }
    } else {
        const int64_t q = k - nWc;
        const int J = ((int)q % (Lc.Nz + 2)), I = ((int)q / (Lc.Nz + 2));
        if (J >= 1 && J <= Lc.Nz) {
            const int r0 = 2 * J - 1, r1 = 2 * J, f = 2 * I;
            const int fl = f == 0 ? Lf.Nx - 1 : f - 1;
            const int fr = f == Lf.Nx ? 1 : f + 1;
            v = (2.0 * (double)rf[Lf.iU(r0, f)] + 2.0 * (double)rf[Lf.iU(r1, f)] + (double)rf[Lf.iU(r0, fl)] +
                 (double)rf[Lf.iU(r1, fl)] + (double)rf[Lf.iU(r0, fr)] + (double)rf[Lf.iU(r1, fr)]) * 0.125;
        }
    }
    bcv[k] = v;
}
template <typename TF>
__global__ __launch_bounds__(NT) void k_restrict(LevD Lf, LevD Lc, const TF* __restrict__ rf,
                                                 double* __restrict__ bcv) {
    const int64_t k = gidx();
    if (k < Lc.nW() + Lc.nU()) restrict_at(Lf, Lc, rf, bcv, k);
}

// bilinear prolongation of a coarse correction ec, added to the fine iterate xf (interior rows)
template <typename TF>
__device__ __forceinline__ void prolong_at(const LevD& Lf, const LevD& Lc, const Flags& F,
                                           const double* __restrict__ ec, TF* __restrict__ xf, int64_t k) {
    const int64_t nWf = Lf.nW();
    if (k < nWf) {
        const int r = ((int)k % (Lf.Nz + 1)), c = ((int)k / (Lf.Nz + 1));
        if (r < 1 || r > Lf.Nz - 1 || c < 1 || c > Lf.Nx) return;
        // coarse W column value with the periodic/mirror ghost map of the BC rows
        auto E = [&](int J, int I) -> double {
            if (J <= 0 || J >= Lc.Nz) return 0.0;
            if (I == 0) return -F.sds * ec[Lc.iW(J, Lc.Nx)];
            if (I == Lc.Nx + 1) return -F.sds * ec[Lc.iW(J, 1)];
            return ec[Lc.iW(J, I)];
        };
        auto EX = [&](int J) -> double {  // x-interpolation at fine column c, coarse row J
            const int I = (c + 1) / 2;
            const int In = (c & 1) ? I - 1 : I + 1;
            return 0.75 * E(J, I) + 0.25 * E(J, In);
        };
        const int J = r / 2;
        const double v = (r & 1) ? 0.5 * (EX(J) + EX(J + 1)) : EX(J);
        xf[k] += v;
    } else {
        const int64_t q = k - nWf;
        const int r = ((int)q % (Lf.Nz + 2)), c = ((int)q / (Lf.Nz + 2));
        if (r < 1 || r > Lf.Nz) return;
        auto E = [&](int J, int I) -> double {
            if (J == 0) return -F.top_cnv * ec[Lc.iU(1, I)];
            if (J == Lc.Nz + 1) return -F.bot_cnv * ec[Lc.iU(Lc.Nz, I)];
            return ec[Lc.iU(J, I)];
        };
        auto EZ = [&](int I) -> double {  // z-interpolation at fine row r, coarse face I
            const int J = (r + 1) / 2;
            const int Jn = (r & 1) ? J - 1 : J + 1;
            return 0.75 * E(J, I) + 0.25 * E(Jn, I);
        };
        const int I = c / 2;
        const double v = (c & 1) ? 0.5 * (EZ(I) + EZ(I + 1)) : EZ(I);
        xf[k] += v;
    }
}

// The coarse end of the V-cycle (levels of at most TAIL_CELLS cells) in one
// workgroup: the sweeps, residuals and transfers of these small levels are
// latency-bound, so they run back to back in one launch instead of one launch
// each, on copies of the coefficients and vectors held in LDS.  Levels of more
// than TAIL_WAVE_PTS unknowns use the whole workgroup with a barrier after each
// phase; smaller ones are run by wave 0 alone, wave-synchronously (no workgroup
// barrier), while the other waves wait at the next barrier.  The coefficients
// of the tail levels are one contiguous block of the solver's pool (carved in
// the LDS layout), staged with every load in flight.  Same arithmetic as the
// per-level kernels.
constexpr int TAIL_NT = 1024;
constexpr int TAIL_MAX = 6;
constexpr int64_t TAIL_CELLS = 16 * 16;
constexpr int TAIL_WAVE_PTS = 256;  // at most 4 unknowns per lane of one wave
constexpr int TAIL_IMG_PER_THREAD = 8;
struct TailArgs {
    LevD d[TAIL_MAX];      // the levels' sizes (coefficient pointers are replaced by LDS ones)
    const double* img;     // the tail levels' coefficients, contiguous, in the LDS layout
    int T;                 // its length (doubles) <= TAIL_IMG_PER_THREAD * TAIL_NT
    const double* bg;      // right-hand side of the first tail level (global)
    double* xg;            // its result (global)
    int off[TAIL_MAX][9];  // LDS offsets (doubles): eta, etaco, rho, rhow, rhou, x, b, r, t
    int nlev, nbig, pre, post, coarse;
    double omega;
    Flags F;
};
__device__ __forceinline__ LevD tail_lev(const TailArgs& a, double* lds, int l) {
    LevD L = a.d[l];
    L.eta = lds + a.off[l][0];
    L.etaco = lds + a.off[l][1];
    L.rho = lds + a.off[l][2];
    L.rhow = lds + a.off[l][3];
    L.rhou = lds + a.off[l][4];
    return L;
}
// completes this wave's LDS accesses: after it, every lane of the wave sees the others' LDS writes
// (a wave's LDS operations execute in order)
__device__ __forceinline__ void wave_sync() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}
__global__ __launch_bounds__(TAIL_NT) void k_mg_tail(TailArgs a) {
    extern __shared__ double lds[];
    __shared__ unsigned swp_shared;
    const int tid = threadIdx.x;
    // stage the coefficient image and the first level's right-hand side, every load in flight
    {
        double v[TAIL_IMG_PER_THREAD];
#pragma unroll
        for (int i = 0; i < TAIL_IMG_PER_THREAD; ++i) v[i] = a.img[min(tid + i * TAIL_NT, a.T - 1)];
        const int nv0 = (int)(a.d[0].nW() + a.d[0].nU());  // <= TAIL_NT (host check)
        const double bv = a.bg[min(tid, nv0 - 1)];
#pragma unroll
        for (int i = 0; i < TAIL_IMG_PER_THREAD; ++i)
            if (tid + i * TAIL_NT < a.T) lds[tid + i * TAIL_NT] = v[i];
        if (tid < nv0) lds[a.off[0][6] + tid] = bv;
    }
    __syncthreads();
    unsigned swp = 0;  // bit l: x and t of level l exchanged
    auto X = [&](int l) { return lds + a.off[l][((swp >> l) & 1) ? 8 : 5]; };
    auto T = [&](int l) { return lds + a.off[l][((swp >> l) & 1) ? 5 : 8]; };
    // level l down: sweeps from zero, then (not coarsest) residual and restriction into l + 1.
    // k0 / dk: this thread's first point and stride; sync: the barrier of the phase.
    auto down = [&](int l, int k0, int dk, auto sync) -> bool {
        const LevD L = tail_lev(a, lds, l);
        const int64_t nv = L.nW() + L.nU();
        const bool coarsest = l + 1 == a.nlev;
        const int sweeps = coarsest ? a.coarse : a.pre;
        double* bl = lds + a.off[l][6];
        for (int it = 0; it < sweeps; ++it) {
            double* xl = X(l);
            double* tl = T(l);
            for (int64_t k = k0; k < nv; k += dk) jacobi_at(L, a.F, xl, bl, tl, a.omega, 0, it == 0, k);
            sync();
            swp ^= 1u << l;
        }
        if (coarsest) return false;
        double* rl = lds + a.off[l][7];
        const double* xl = X(l);
        for (int64_t k = k0; k < nv; k += dk) vel_resid_at(L, a.F, xl, bl, rl, 0, k);
        sync();
        const LevD C = tail_lev(a, lds, l + 1);
        const int64_t nc = C.nW() + C.nU();
        double* bc = lds + a.off[l + 1][6];
        for (int64_t k = k0; k < nc; k += dk) restrict_at(L, C, rl, bc, k);
        sync();
        return true;
    };
    // level l up: prolongation of level l + 1's correction, then post-sweeps
    auto up = [&](int l, int k0, int dk, auto sync) {
        const LevD L = tail_lev(a, lds, l);
        const LevD C = tail_lev(a, lds, l + 1);
        const int64_t nv = L.nW() + L.nU();
        double* bl = lds + a.off[l][6];
        {
            double* xl = X(l);
            const double* xc = X(l + 1);
            for (int64_t k = k0; k < nv; k += dk) prolong_at(L, C, a.F, xc, xl, k);
        }
        sync();
        for (int it = 0; it < a.post; ++it) {
            double* xl = X(l);
            double* tl = T(l);
            for (int64_t k = k0; k < nv; k += dk) jacobi_at(L, a.F, xl, bl, tl, a.omega, 0, 0, k);
            sync();
            swp ^= 1u << l;
        }
    };
    auto block_sync = [] { __syncthreads(); };
    for (int l = 0; l < a.nbig; ++l)
        if (!down(l, tid, TAIL_NT, block_sync)) break;
    if (a.nbig < a.nlev) {
        if (tid < 64) {  // the small levels: wave 0 alone
            for (int l = a.nbig; l < a.nlev; ++l)
                if (!down(l, tid, 64, wave_sync)) break;
            for (int l = a.nlev - 2; l >= a.nbig; --l) up(l, tid, 64, wave_sync);
            if (tid == 0) swp_shared = swp;
        }
        __syncthreads();
        swp = swp_shared;  // the other waves did not follow the small levels' exchanges
    }
    for (int l = min(a.nbig, a.nlev - 1) - 1; l >= 0; --l) up(l, tid, TAIL_NT, block_sync);
    const int nv0 = (int)(a.d[0].nW() + a.d[0].nU());
    const double* x0 = X(0);
    for (int k = tid; k < nv0; k += TAIL_NT) a.xg[k] = x0[k];
}

// coarse coefficients from fine ones
__global__ __launch_bounds__(NT) void k_coarsen(LevD Lf, double* eta, double* rho, double* rhow, double* rhou,
                                                int Nzc, int Nxc) {
    const int64_t k = gidx();
    const int64_t nmax = (int64_t)(Nzc + 1) * (Nxc + 1);
    if (k >= nmax) return;
    const int J = ((int)k % (Nzc + 1)), I = ((int)k / (Nzc + 1));
    if (J < Nzc && I < Nxc) {
        const double e00 = Lf.ETA(2 * J, 2 * I), e10 = Lf.ETA(2 * J + 1, 2 * I);
        const double e01 = Lf.ETA(2 * J, 2 * I + 1), e11 = Lf.ETA(2 * J + 1, 2 * I + 1);
        eta[J + (int64_t)I * Nzc] = sqrt(sqrt(e00 * e10) * sqrt(e01 * e11));
        rho[J + (int64_t)I * Nzc] = 0.25 * (Lf.RHO(2 * J, 2 * I) + Lf.RHO(2 * J + 1, 2 * I) + Lf.RHO(2 * J, 2 * I + 1) +
                                            Lf.RHO(2 * J + 1, 2 * I + 1));
    }
    if (I < Nxc) rhow[J + (int64_t)I * (Nzc + 1)] = 0.5 * (Lf.RHOW(2 * J, 2 * I) + Lf.RHOW(2 * J, 2 * I + 1));
    if (J < Nzc) rhou[J + (int64_t)I * Nzc] = 0.5 * (Lf.RHOU(2 * J, 2 * I) + Lf.RHOU(2 * J + 1, 2 * I));
}

// corner viscosity of a coarse level: geometric mean, periodic x / closed z (update.jl:411-426)
__global__ __launch_bounds__(NT) void k_corner_eta(const double* eta, double* etaco, int Nz, int Nx) {
    const int64_t k = gidx();
    if (k >= (int64_t)(Nz + 1) * (Nx + 1)) return;
    const int j = ((int)k % (Nz + 1)), i = ((int)k / (Nz + 1));
    const int z0 = j - 1 < 0 ? 0 : j - 1, z1 = j >= Nz ? Nz - 1 : j;
    const int x0 = i - 1 < 0 ? Nx - 1 : i - 1, x1 = i >= Nx ? 0 : i;
    const double p = eta[z0 + (int64_t)x0 * Nz] * eta[z1 + (int64_t)x0 * Nz] * eta[z0 + (int64_t)x1 * Nz] *
                     eta[z1 + (int64_t)x1 * Nz];
    etaco[k] = sqrt(sqrt(p));
}

// Schur diagonal on the P block + GG coupling into the velocity residual:
//   zp = S^-1 rp (interior: eta/rho; pins: see below), then rv' = rv - GG zp
__global__ __launch_bounds__(NT) void k_schur_int(LevD L, Flags F, const double* __restrict__ rP,
                                                  double* __restrict__ zP) {
    const int64_t k = gidx();
    const int Nz = L.Nz, Nx = L.Nx;
    if (k >= (int64_t)(Nz + 2) * (Nx + 2)) return;
    const int r = ((int)k % (Nz + 2)), c = ((int)k / (Nz + 2));
    if (r < 1 || r > Nz || c < 1 || c > Nx) return;
    // S = KP - DM A^-1 GG ~ +rho/eta (DM ~ rho div, GG ~ grad, A ~ -eta lap)
    const double s = L.RHO(r - 1, c - 1) / L.ETA(r - 1, c - 1);
    if (is_pin(F, r, c, Nz)) zP[k] = F.bnchm ? rP[k] : rP[k] / (1.0 + s);
    else zP[k] = rP[k] / s;
}
// ghost rows of KP solved exactly given the interior values (fluidmech.jl:405-417)
__global__ __launch_bounds__(NT) void k_schur_ghost(LevD L, Flags F, const double* __restrict__ rP,
                                                    double* __restrict__ zP) {
    const int64_t k = gidx();
    const int Nz = L.Nz, Nx = L.Nx;
    if (k >= (int64_t)(Nz + 2) * (Nx + 2)) return;
    const int r = ((int)k % (Nz + 2)), c = ((int)k / (Nz + 2));
    auto P = [&](int rr, int cc) { return zP[rr + (int64_t)cc * (Nz + 2)]; };
    const bool pin = is_pin(F, r, c, Nz);
    double v;
    if (r >= 1 && r <= Nz && c >= 1 && c <= Nx) return;
    if (pin) { zP[k] = rP[k]; return; }
    if (c == 0 && r >= 1 && r <= Nz) v = rP[k] + P(r, Nx);
    else if (c == Nx + 1 && r >= 1 && r <= Nz) v = rP[k] + P(r, 1);
    else {
        // rows 0 / Nz+1 reference row 1 / Nz at the same column (which may itself be a ghost column)
        const int rn = r == 0 ? 1 : Nz;
        double inner;
        if (c == 0) inner = rP[rn + (int64_t)0 * (Nz + 2)] + P(rn, Nx);
        else if (c == Nx + 1) inner = rP[rn + (int64_t)(Nx + 1) * (Nz + 2)] + P(rn, 1);
        else inner = P(rn, c);
        v = rP[k] + inner;
    }
    zP[k] = v;
}
__global__ __launch_bounds__(NT) void k_sub_grad(LevD L, Flags F, const double* __restrict__ rV,
                                                 const double* __restrict__ zP, double* __restrict__ tV) {
    const int64_t k = gidx();
    const int64_t nW = L.nW(), n = nW + L.nU();
    if (k >= n) return;
    const int Nz = L.Nz, Nx = L.Nx;
    double g = 0.0;
    if (k != F.nu0) {
        if (k < nW) {
            const int r = ((int)k % (Nz + 1)), c = ((int)k / (Nz + 1));
            if (r >= 1 && r <= Nz - 1 && c >= 1 && c <= Nx)
                g = -L.invh * zP[r + (int64_t)c * (Nz + 2)] + L.invh * zP[r + 1 + (int64_t)c * (Nz + 2)];
        } else {
            const int64_t q = k - nW;
            const int r = ((int)q % (Nz + 2)), c = ((int)q / (Nz + 2));
            if (r >= 1 && r <= Nz)
                g = -L.invh * zP[r + (int64_t)c * (Nz + 2)] + L.invh * zP[r + (int64_t)(c + 1) * (Nz + 2)];
        }
    }
    tV[k] = rV[k] - g;
}

// ---- inertial part of the Schur complement: a pressure Poisson solve ------
// With the mass terms M = a1 rho_f / dt on the interior velocity rows, the
// face densities cancel exactly in -DM M^-1 GG, leaving the constant-coefficient
// 5-point operator T = cz (-d_zz) + cx (-d_xx), cz = dt / (a1 h^2),
// cx = dt / ((a1 + gamma) h^2), periodic in x and zero-flux at the top and
// bottom rows (W there is a BC row).  T is diagonalised in x by a real FFT and
// in z by a cosine transform; the kx = 0 system is singular (constant
// pressure) and is solved with the mean removed.

// rows of the interior P residual -> half spectra S[r][k], k = 0..Nx/2 (two rows per complex line)
struct PoisXArgs {
    const double* rP;   // P block of the residual, (Nz+2) x (Nx+2)
    double2* S;         // Nz x K, k fastest
    Plan p;
    int Nz, Nx, K, nl;
};
template <int LG>  // 8..10: the power-of-two path with prefetched twiddles (P2Fast); 0: lds_fft
__global__ __launch_bounds__(FFT_NT) void k_pois_xfwd(PoisXArgs a) {
    extern __shared__ double2 sm[];
    double2 w[P2Fast<LG>::NST][3];
    if constexpr (LG > 0) fft_p2_twiddles<LG>(a.nl, a.p.tw, w);
    const int n = a.Nx, nl = a.nl, rows = 2 * nl;
    // neighbouring row blocks share the cache lines of the column-major P arrays: keep them on one XCD
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * rows;
    double2* A = sm;
    double2* B = sm + nl * n;
    for (int k = threadIdx.x; k < rows * n; k += FFT_NT) {
        const int rr = k % rows, c = k / rows;
        const int r = r0 + rr;
        const double v = r < a.Nz ? a.rP[(r + 1) + (int64_t)(c + 1) * (a.Nz + 2)] : 0.0;
        reinterpret_cast<double*>(&A[(rr >> 1) * n + c])[rr & 1] = v;
    }
    __syncthreads();
    const double2* X;
    if constexpr (LG > 0) X = lds_fft_p2<false, LG>(A, B, nl, w);
    else X = lds_fft<false>(A, B, nl, a.p);
    for (int k = threadIdx.x; k < nl * a.K; k += FFT_NT) {
        const int l = k / a.K, f = k - l * a.K;
        const double2 zk = X[l * n + f], zm = X[l * n + (f == 0 ? 0 : n - f)];
        const int ra = r0 + 2 * l, rb = ra + 1;
        if (ra < a.Nz) a.S[(int64_t)ra * a.K + f] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
        if (rb < a.Nz) a.S[(int64_t)rb * a.K + f] = make_double2(0.5 * (zk.y + zm.y), -0.5 * (zk.x - zm.x));
    }
}

// z solve of every wavenumber column, in place.  The zero-flux rows make the
// cell-centred z operator the periodic one on the column mirrored about its end
// faces, so the cosine transform (DCT-II) of the column diagonalises it with
// eigenvalues lam_z[kz] = 2 - 2 cos(pi kz / Nz): transform, divide by
// lam_x[k] + lam_z[kz] (the kx = kz = 0 mode, the constant, is dropped:
// zero-mean solution), transform back (DCT-III).  Both cosine transforms are
// Nz-point complex FFTs (Makhoul's reordering, J. Makhoul, "A fast cosine
// transform in one and two dimensions", IEEE Trans. ASSP 28 (1980) 27-34):
//   v[n] = s[2n], v[Nz-1-n] = s[2n+1];  V = FFT(v);
//   C[k] = (w_k V[k] + conj(w_k) V[Nz-k]) / 2,  w_k = exp(-i pi k / (2 Nz)),
// which holds for complex columns because the transform is real-linear; the
// inverse is V'[k] = conj(w_k) (E[k] - i E[Nz-k]) (E[Nz] = 0), v' = IFFT(V'),
// y[src(n)] = v'[n] / Nz.  A lane takes the pair (k, Nz-k) through the
// forward post-twiddle, the division and the inverse pre-twiddle in registers.
// This is the mirrored 2 Nz-point FFT solve at half the transform length.
//
// open_cnv: the P row Nz is pinned (its DM row plus the pin; the DM row vanishes on the velocities
// A^-1 GG p produces, because the W row Nz is the same mass balance as a BC row), so the Schur complement
// is the identity on row Nz and, on rows 1..Nz-1, the operator with the value of row Nz as a fixed
// (Dirichlet) neighbour: z(Nz) = r(Nz), and rows 1..Nz-1 solve T with y(Nz) held at cz r(Nz) in the
// spectral units y = cz FFT_x(z).  Per wavenumber that is the zero-flux solve y0 of the column with its
// last entry zeroed plus a multiple of g_k = (T_N + lam_x[k])^-1 e_last (tabulated on the host) that
// brings y(last) to the target; for k = 0 (T_N singular: g_0 = T_N^+ e_last) the right-hand side's sum s
// is moved into the last row (-s g_0) and the constant mode sets y(last).
constexpr int PZ_NLMAX = 4;  // lines a workgroup of k_pois_zfft takes at most
struct PoisZArgs {
    double2* S;          // Nz x K, k fastest
    Plan p;              // length Nz
    const double2* wk;   // Nz: exp(-i pi k / (2 Nz))
    const double* lamx;  // K: (cx/cz) (2 - 2 cos(2 pi k / Nx))
    const double* lamz;  // Nz: 2 - 2 cos(pi kz / Nz)
    int Nz, K, nl;
    const double* g;     // open_cnv: K x Nz, g[k * Nz + i] = g_k(i); null otherwise
    double cz;           // dt / (a1 h^2)
};
constexpr int ZD_PER = 5;  // (k, Nz-k) pairs a lane handles: nl (Nz/2 + 1) <= ZD_PER FFT_NT (Nz <= MAXLINE)
template <int LG>
__global__ __launch_bounds__(FFT_NT) void k_pois_zfft(PoisZArgs a) {
    extern __shared__ double2 sm[];
    double2 w[P2Fast<LG>::NST][3];
    if constexpr (LG > 0) fft_p2_twiddles<LG>(a.nl, a.p.tw, w);
    const int n = a.Nz, nl = a.nl, K = a.K, H = n / 2 + 1;
    const int k0 = xcd_tile(blockIdx.x, gridDim.x) * nl;  // neighbouring wavenumbers share S's cache lines
    double2* A = sm;
    double2* B = sm + nl * n;
    auto src = [&](int m) { return m < (n + 1) / 2 ? 2 * m : 2 * n - 2 * m - 1; };
    const bool open = a.g != nullptr;
    const int mlast = (n - 1) % 2 == 0 ? (n - 1) / 2 : n / 2;  // src(mlast) = n - 1
    __shared__ double2 s_tgt[PZ_NLMAX], s_sum[PZ_NLMAX], s_t[PZ_NLMAX], s_c[PZ_NLMAX];
    for (int i = threadIdx.x; i < nl * n; i += FFT_NT) {
        const int l = i / n, m = i - l * n;
        const int k = k0 + l;
        const double2 v = k < K ? a.S[(int64_t)src(m) * K + k] : make_double2(0.0, 0.0);
        if (open && m == mlast) {  // open_cnv: the pinned row leaves the column; its target value
            s_tgt[l] = make_double2(a.cz * v.x, a.cz * v.y);
            A[i] = make_double2(0.0, 0.0);
        } else {
            A[i] = v;
        }
    }
    // this lane's pairs: the twiddle and both scaled inverse eigenvalues, fetched before the transform
    double2 wq[ZD_PER];
    double fa[ZD_PER], fb[ZD_PER];
    const double inv_n = 1.0 / (double)n;
#pragma unroll
    for (int q = 0; q < ZD_PER; ++q) {
        const int t = threadIdx.x + q * FFT_NT;
        const int tc = t < nl * H ? t : 0;
        const int l = tc / H, kz = tc - l * H;
        const double lx = k0 + l < K ? a.lamx[k0 + l] : 1.0;
        const double la = lx + a.lamz[kz], lb = lx + a.lamz[kz == 0 ? 0 : n - kz];
        wq[q] = a.wk[kz];
        fa[q] = la > 0.0 ? inv_n / la : 0.0;
        fb[q] = lb > 0.0 ? inv_n / lb : 0.0;
    }
    __syncthreads();
    double2* X;
    if constexpr (LG > 0) X = lds_fft_p2<false, LG>(A, B, nl, w);
    else X = lds_fft<false>(A, B, nl, a.p);
#pragma unroll
    for (int q = 0; q < ZD_PER; ++q) {
        const int t = threadIdx.x + q * FFT_NT;
        if (t < nl * H) {
            const int l = t / H, kz = t - l * H, kb = kz == 0 ? 0 : n - kz;
            double2* Xl = X + l * n;
            const double2 va = Xl[kz], vb = Xl[kb], wa = wq[q];
            // w_{Nz-k} = -i conj(w_k)
            const double2 wb = make_double2(-wa.y, -wa.x);
            // C[k] = (w_k V[k] + conj(w_k) V[Nz-k]) / 2, C[Nz-k] = (w_{Nz-k} V[Nz-k] + conj(w_{Nz-k}) V[k]) / 2
            const double2 ca = make_double2(0.5 * ((wa.x * va.x - wa.y * va.y) + (wa.x * vb.x + wa.y * vb.y)),
                                            0.5 * ((wa.x * va.y + wa.y * va.x) + (wa.x * vb.y - wa.y * vb.x)));
            const double2 cb = make_double2(0.5 * ((wb.x * vb.x - wb.y * vb.y) + (wb.x * va.x + wb.y * va.y)),
                                            0.5 * ((wb.x * vb.y + wb.y * vb.x) + (wb.x * va.y - wb.y * va.x)));
            if (open && kz == 0) s_sum[l] = ca;  // C[0]: the column's sum
            const double2 ea = make_double2(ca.x * fa[q], ca.y * fa[q]);
            const double2 eb = kz == 0 ? make_double2(0.0, 0.0) : make_double2(cb.x * fb[q], cb.y * fb[q]);
            // V'[k] = conj(w_k) (E[k] - i E[Nz-k]), V'[Nz-k] = conj(w_{Nz-k}) (E[Nz-k] - i E[k])
            const double2 da = make_double2(ea.x + eb.y, ea.y - eb.x);
            const double2 db = make_double2(eb.x + ea.y, eb.y - ea.x);
            Xl[kz] = make_double2(wa.x * da.x + wa.y * da.y, wa.x * da.y - wa.y * da.x);
            if (kb != kz && kz != 0) Xl[kb] = make_double2(wb.x * db.x + wb.y * db.y, wb.x * db.y - wb.y * db.x);
        }
    }
    __syncthreads();
    const double2* Y;
    if constexpr (LG > 0) Y = lds_fft_p2<true, LG>(X, X == A ? B : A, nl, w);
    else Y = lds_fft<true>(X, X == A ? B : A, nl, a.p);
    if (!open) {
        for (int i = threadIdx.x; i < nl * n; i += FFT_NT) {
            const int l = i / n, m = i - l * n;
            const int k = k0 + l;
            if (k < K) a.S[(int64_t)src(m) * K + k] = Y[i];
        }
        return;
    }
    // open_cnv: y = y0 + t g_k + c per line (c = 0 for k > 0)
    if (threadIdx.x < nl && k0 + (int)threadIdx.x < K) {
        const int l = threadIdx.x, k = k0 + l;
        const double2 y0 = Y[l * n + mlast], tg = s_tgt[l];
        const double gl = a.g[(int64_t)k * n + n - 1];
        if (k > 0) {
            s_t[l] = make_double2((tg.x - y0.x) / gl, (tg.y - y0.y) / gl);
            s_c[l] = make_double2(0.0, 0.0);
        } else {
            const double2 sm0 = s_sum[l];
            s_t[l] = make_double2(-sm0.x, -sm0.y);
            s_c[l] = make_double2(tg.x - (y0.x - sm0.x * gl), tg.y - (y0.y - sm0.y * gl));
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nl * n; i += FFT_NT) {
        const int l = i / n, m = i - l * n;
        const int k = k0 + l;
        if (k < K) {
            const double gi = a.g[(int64_t)k * n + src(m)];
            const double2 t = s_t[l], c = s_c[l], y = Y[i];
            a.S[(int64_t)src(m) * K + k] = make_double2(y.x + t.x * gi + c.x, y.y + t.y * gi + c.y);
        }
    }
}

// inverse x transform of two rows per complex line and the whole interior Schur apply:
//   zP = CV rP / s (s = rho/eta; 1 + s at a kept-DM pin; identity at an MMS pin) + scale * T^-1 rP
// CV weighs the viscous part against the inertial one: measured on the 512^2 step, 0.75 gives 8.6 FGMRES
// iterations per solve against 9.0 with 1 (0.5: 9.4, 1.33: 10.9, 2: 13.9; the inertial weight is best at 1)
constexpr double SCHUR_CV = 0.75;
struct PoisInvArgs {
    const double2* S;
    const double* rP;
    double* zP;
    Plan p;
    LevD L;
    Flags F;
    double scale;   // 1 / (cz Nx)
    int Nz, Nx, K, nl;
};
template <int LG>
__global__ __launch_bounds__(FFT_NT) void k_pois_xinv(PoisInvArgs a) {
    extern __shared__ double2 sm[];
    double2 w[P2Fast<LG>::NST][3];
    if constexpr (LG > 0) fft_p2_twiddles<LG>(a.nl, a.p.tw, w);
    const int n = a.Nx, nl = a.nl, rows = 2 * nl, K = a.K;
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * rows;  // row blocks of one XCD share cache lines
    double2* A = sm;
    double2* B = sm + nl * n;
    for (int k = threadIdx.x; k < nl * n; k += FFT_NT) {
        const int l = k / n, f = k - l * n;
        const int ra = r0 + 2 * l, rb = ra + 1;
        const bool lo = f < K;
        const int fs = lo ? f : n - f;
        double2 av = ra < a.Nz ? a.S[(int64_t)ra * K + fs] : make_double2(0.0, 0.0);
        double2 bv = rb < a.Nz ? a.S[(int64_t)rb * K + fs] : make_double2(0.0, 0.0);
        if (!lo) { av.y = -av.y; bv.y = -bv.y; }
        A[k] = make_double2(av.x - bv.y, av.y + bv.x);
    }
    __syncthreads();
    const double2* Y;
    if constexpr (LG > 0) Y = lds_fft_p2<true, LG>(A, B, nl, w);
    else Y = lds_fft<true>(A, B, nl, a.p);
    double* Zs = reinterpret_cast<double*>(Y == A ? B : A);  // rows x n interior results for the ghost ring
    const int Nz = a.Nz;
    for (int k = threadIdx.x; k < rows * n; k += FFT_NT) {
        const int rr = k % rows, c = k / rows;
        const int r = r0 + rr;
        if (r >= Nz) continue;
        const double pv = reinterpret_cast<const double*>(&Y[(rr >> 1) * n + c])[rr & 1];
        const int64_t q = (r + 1) + (int64_t)(c + 1) * (Nz + 2);
        const double sv = a.L.RHO(r, c) / a.L.ETA(r, c);
        double z;
        if (a.F.open_cnv && is_pin(a.F, r + 1, c + 1, Nz)) z = a.rP[q];  // identity row of the Schur complement
        else if (is_pin(a.F, r + 1, c + 1, Nz)) z = (a.F.bnchm ? a.rP[q] : a.rP[q] / (1.0 + sv)) + a.scale * pv;
        else z = SCHUR_CV * a.rP[q] / sv + a.scale * pv;
        a.zP[q] = z;
        Zs[rr * n + c] = z;
    }
    __syncthreads();
    // the ghost ring of zP, as k_schur_ghost computes it (fluidmech.jl:405-417): columns 0 and Nx+1 of
    // this workgroup's rows; row 0 / Nz+1 when this workgroup holds row 1 / Nz
    const int Nx = n;
    auto P = [&](int rr, int c) { return Zs[rr * n + c]; };  // interior (r0 + rr + 1, c + 1)
    auto rPv = [&](int r, int c) { return a.rP[r + (int64_t)c * (Nz + 2)]; };
    for (int k = threadIdx.x; k < 2 * rows; k += FFT_NT) {
        const int rr = k >> 1, r = r0 + rr;
        if (r >= Nz) continue;
        const int zr = r + 1, c = (k & 1) ? Nx + 1 : 0;
        const int64_t q = zr + (int64_t)c * (Nz + 2);
        a.zP[q] = is_pin(a.F, zr, c, Nz) ? a.rP[q] : a.rP[q] + ((k & 1) ? P(rr, 0) : P(rr, Nx - 1));
    }
    for (int side = 0; side < 2; ++side) {
        const int zr = side ? Nz + 1 : 0;              // ghost row
        const int rn = side ? Nz : 1;                  // the interior row it reads
        const int rr = rn - 1 - r0;                    // that row in this workgroup
        if (rr < 0 || rr >= rows) continue;
        for (int c = threadIdx.x; c < Nx + 2; c += FFT_NT) {
            const int64_t q = zr + (int64_t)c * (Nz + 2);
            if (is_pin(a.F, zr, c, Nz)) { a.zP[q] = a.rP[q]; continue; }
            double inner;
            if (c == 0) inner = rPv(rn, 0) + P(rr, Nx - 1);
            else if (c == Nx + 1) inner = rPv(rn, Nx + 1) + P(rr, 0);
            else inner = P(rr, c - 1);
            a.zP[q] = a.rP[q] + inner;
        }
    }
}

// ---- vector kernels for FGMRES ------------------------------------------
// partial dot products of w against K basis vectors V[0..K) (stride n) and, as entry K, w.w
constexpr int DOT_BLOCKS = 1024;
// partial dot products of w against K basis vectors V[0..K) (stride n) and, as entry K, w.w
__global__ __launch_bounds__(NT) void k_multidot(const double* __restrict__ V, int64_t n, int K,
                                                 const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double red[NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q <= K; ++q) {
        const double* v = q < K ? V + (int64_t)q * n : w;
        double acc = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += (int64_t)gridDim.x * NT) acc += v[i] * w[i];
        acc = wave_total(acc);
        if (lane == 0) red[wave] = acc;
        __syncthreads();
        if (threadIdx.x == 0) part[(int64_t)q * gridDim.x + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
        __syncthreads();
    }
}
// out[q] = sum of the nb partials of q (fixed order); one wavefront per q
__global__ __launch_bounds__(NT) void k_dot_final(const double* __restrict__ part, int nb, int K1, double* out) {
    const int q = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    if (q >= K1) return;
    const double t = xc_rows_reduce<false>(part, nb, q);
    if ((threadIdx.x & 63) == 0) out[q] = t;
}
// out[0] = the sum of nb partials (fixed order) by the whole workgroup, sixteen loads in flight per lane: for
// the thousands of per-wave partials of the residual kernel, which one wave would read in many dependent rounds
__device__ __forceinline__ void block_sum1(const double* __restrict__ part, int nb, double* out) {
    __shared__ double red[NT / 64];
    constexpr int U = 16;
    double acc = 0.0;
    for (int b0 = threadIdx.x; b0 < nb; b0 += NT * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + NT * u;
            v[u] = part[b < nb ? b : nb - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + NT * u < nb) acc += v[u];
    }
    acc = wave_total(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ __launch_bounds__(NT) void k_sum_final1(const double* __restrict__ part, int nb, double* out,
                                                   const double* __restrict__ part2 = nullptr, double* out2 = nullptr) {
    block_sum1(part, nb, out);
    if (part2) {
        __syncthreads();  // block_sum1's LDS is read by thread 0 before it is reused
        block_sum1(part2, nb, out2);
    }
}
// ---- learned coarse correction: restriction to the network's grid and prolongation back ----
// Cell-centred (W, U, P) of a stacked vector at interior cell (r, c), 0-based
__device__ __forceinline__ void cell_vals(const LevD& L, const double* v, int r, int c, double* w, double* u,
                                          double* p) {
    *w = 0.5 * (v[L.iW(r, c + 1)] + v[L.iW(r + 1, c + 1)]);
    *u = 0.5 * (v[L.iU(r + 1, c)] + v[L.iU(r + 1, c + 1)]);
    *p = v[L.iP(r + 1, c + 1)];
}
struct LcArgs {
    LevD L;
    int f, Hc, Wc;
    const double* v;
    float* out;        // [p][4], p = Xc * Hc + Zc
    double* part;      // per-block sums of squares of channels 1..3
    double eta_ref;
    float s[4];        // fixed channel scales
};
__global__ __launch_bounds__(NT) void k_lc_restrict(LcArgs a) {
    __shared__ double red[NT];
    const int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x;
    double ss = 0.0;
    if (q < (int64_t)a.Hc * a.Wc) {
        const int Zc = ((int)q % a.Hc), Xc = ((int)q / a.Hc);
        double sw = 0.0, su = 0.0, sp = 0.0, se = 0.0;
        for (int j = 0; j < a.f; ++j)
            for (int i = 0; i < a.f; ++i) {
                const int r = Zc * a.f + i, c = Xc * a.f + j;
                double w, u, p;
                cell_vals(a.L, a.v, r, c, &w, &u, &p);
                sw += w;
                su += u;
                sp += p;
                se += log(a.L.ETA(r, c) / a.eta_ref);
            }
        const double inv = 1.0 / ((double)a.f * a.f);
        const double cw = sw * inv * a.s[1], cu = su * inv * a.s[2], cp = sp * inv * a.s[3];
        a.out[q * 4 + 0] = (float)(se * inv * a.s[0]);
        a.out[q * 4 + 1] = (float)cw;
        a.out[q * 4 + 2] = (float)cu;
        a.out[q * 4 + 3] = (float)cp;
        ss = cw * cw + cu * cu + cp * cp;
    }
    red[threadIdx.x] = ss;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}
// sigma = RMS of channels 1..3 (or the given one); divide channels 1..3 by it (one workgroup)
__global__ __launch_bounds__(1024) void k_lc_norm(float* inp, const double* part, int nb, int64_t ncell, float* sigma,
                                                  int compute) {
    __shared__ double red[1024];
    __shared__ float sg;
    if (compute) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < nb; b += 1024) acc += part[b];
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int st = 512; st > 0; st >>= 1) {
            if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const double rms = sqrt(red[0] / (3.0 * (double)ncell));
            sigma[0] = rms > 0.0 ? (float)rms : 1.0f;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) sg = 1.0f / sigma[0];
    __syncthreads();
    const float is = sg;
    for (int64_t k = threadIdx.x; k < ncell * 3; k += 1024) {
        const int64_t c = k / 3;
        inp[c * 4 + 1 + (k - c * 3)] *= is;
    }
}
// z += P e: coarse cell values (times sigma and the output scales) interpolated bilinearly to the
// fine cell centres, averaged onto the interior W and U faces; BC rows and P ghosts receive nothing
struct LcProArgs {
    LevD L;
    int f, Hc, Wc;
    const float* e;    // [p][3]
    const float* sigma;
    float os[3];
    double* z;
    int64_t n;
};
__global__ __launch_bounds__(NT) void k_lc_prolong(LcProArgs a) {
    const int64_t k = gidx();
    if (k >= a.n) return;
    const LevD& L = a.L;
    const int Nz = L.Nz, Nx = L.Nx;
    const double sg = (double)a.sigma[0];
    // fine cell (r, c): bilinear in the coarse cell centres, x periodic, z clamped
    auto E = [&](int r, int c, int ch) {
        const double zc = ((double)r + 0.5) / a.f - 0.5, xc = ((double)c + 0.5) / a.f - 0.5;
        const int z0 = (int)floor(zc), x0 = (int)floor(xc);
        const double tz = zc - z0, tx = xc - x0;
        const int za = z0 < 0 ? 0 : z0, zb = z0 + 1 > a.Hc - 1 ? a.Hc - 1 : z0 + 1;
        const int xa = x0 < 0 ? x0 + a.Wc : x0, xb = x0 + 1 >= a.Wc ? x0 + 1 - a.Wc : x0 + 1;
        auto at = [&](int z, int x) { return (double)a.e[((int64_t)x * a.Hc + z) * 3 + ch]; };
        return (1.0 - tz) * ((1.0 - tx) * at(za, xa) + tx * at(za, xb)) + tz * ((1.0 - tx) * at(zb, xa) + tx * at(zb, xb));
    };
    const int64_t nW = L.nW(), nU = L.nU();
    double v = 0.0;
    if (k < nW) {
        const int r = ((int)k % (Nz + 1)), c = ((int)k / (Nz + 1));
        if (r < 1 || r > Nz - 1 || c < 1 || c > Nx) return;
        v = 0.5 * (E(r - 1, c - 1, 0) + E(r, c - 1, 0)) * a.os[0];
    } else if (k < nW + nU) {
        const int64_t q = k - nW;
        const int r = ((int)q % (Nz + 2)), c = ((int)q / (Nz + 2));
        if (r < 1 || r > Nz) return;
        const int cl = c == 0 ? Nx - 1 : c - 1, cr = c == Nx ? 0 : c;
        v = 0.5 * (E(r - 1, cl, 1) + E(r - 1, cr, 1)) * a.os[1];
    } else {
        const int64_t q = k - nW - nU;
        const int r = ((int)q % (Nz + 2)), c = ((int)q / (Nz + 2));
        if (r < 1 || r > Nz || c < 1 || c > Nx) return;
        v = E(r - 1, c - 1, 2) * a.os[2];
    }
    a.z[k] += sg * v;
}

// ---- Arnoldi bookkeeping on the device (one lane; m <= 60) -------------
// kry layout: H (m+1) x m column-major | cs[m] | sn[m] | g[m+1] | y[m]
struct Kry {
    double *H, *cs, *sn, *g, *y;
};
// the start of a cycle: v_0 = r / beta, and g = beta e_0 (by the first workgroup)
__global__ __launch_bounds__(NT) void k_cycle_start(const double* __restrict__ r, double inv_beta,
                                                    double* __restrict__ V, int64_t n, Kry K, int m, double beta) {
    const int64_t i = gidx();
    if (i < n) V[i] = r[i] * inv_beta;
    if (blockIdx.x == 0)
        for (int q = threadIdx.x; q <= m; q += NT) K.g[q] = q == 0 ? beta : 0.0;
}
// the same with beta^2 = ||r||^2 still on the device (b2), so that the host need not wait for it; the first
// workgroup also hands beta^2 to the host (hb2)
__global__ __launch_bounds__(NT) void k_cycle_start_dev(const double* __restrict__ r, const double* __restrict__ b2,
                                                        double* __restrict__ V, int64_t n, Kry K, int m,
                                                        double* hb2) {
    const double bb = *b2;
    const double beta = sqrt(bb), inv_beta = 1.0 / beta;
    const int64_t i = gidx();
    if (i < n) V[i] = r[i] * inv_beta;
    if (blockIdx.x == 0) {
        for (int q = threadIdx.x; q <= m; q += NT) K.g[q] = q == 0 ? beta : 0.0;
        if (threadIdx.x == 0) *hb2 = bb;
    }
}
// ---- Arnoldi with delayed reorthogonalisation (DCGS2, classical Gram-Schmidt twice with the second
// pass of v_j folded into step j+1: Bielich, Langou, Thomas, Swirydowicz, Yamazaki, Boman, "Low-synch
// Gram-Schmidt with delayed reorthogonalization for Krylov solvers", Parallel Computing 2022).
// Step j holds the provisional vector p_j (one Gram-Schmidt pass, ~unit norm) in basis slot j and
// w = LL M^-1 p_j.  One reduction gives a = V_j^T p_j, b = V_j^T w, p.p, p.w, w.w (V_j = v_0..v_{j-1},
// orthonormal); then v_j = (p_j - V_j a) / rho with rho^2 = p.p - |a|^2 (p_j's second pass), column j-1
// of H is final (its last vector is now v_j), and w's projections give the provisional column j and
// p_{j+1} = (w - V_j b - h_jj v_j) / nu_{j+1}.  FGMRES needs only LL z_j in span(V): z_j = M^-1 p_j is a
// direction like any other.
//
// The reduction of one DCGS2 step: per block, partials of V_q.p (row q) and V_q.w (row K + q) for q < K,
// then p.p, p.w, w.w (rows 2K .. 2K+2); k_arnoldi2 adds the nb = gridDim.x partials of each row in a
// fixed order.  Each thread keeps its EPT elements of p and w in registers and streams the basis vectors
// MD_QC at a time, their loads in flight together (p and w are read once, not once per basis vector).
constexpr int MD_QC = 4;
constexpr int MD_BLOCKS = 512;  // target number of blocks (two per CU); more when n needs them
constexpr int MD_ROWS = 2 * 64 + 3;  // rows of the reduction for a restart length up to 64
template <int EPT>
__global__ __launch_bounds__(NT) void k_multidot2(const double* __restrict__ V, int64_t n, int K,
                                                  const double* __restrict__ p, const double* __restrict__ w,
                                                  double* __restrict__ part) {
    __shared__ double red[MD_ROWS][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int64_t i0 = (int64_t)blk * (EPT * NT) + threadIdx.x;
    // indices past the end are clamped (in bounds) and their p, w entries zeroed, so that every load is
    // unconditional and the V entries read there contribute nothing
    int64_t ix[EPT];
    double pr[EPT], wr[EPT];
#pragma unroll
    for (int e = 0; e < EPT; ++e) {
        const int64_t i = i0 + (int64_t)e * NT;
        ix[e] = i < n ? i : n - 1;
        const double pv = p[ix[e]], wv = w[ix[e]];
        pr[e] = i < n ? pv : 0.0;
        wr[e] = i < n ? wv : 0.0;
    }
    {
        double a = 0.0, b = 0.0, c = 0.0;
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
            a = fma(pr[e], pr[e], a);
            b = fma(pr[e], wr[e], b);
            c = fma(wr[e], wr[e], c);
        }
        a = wave_total(a);
        b = wave_total(b);
        c = wave_total(c);
        if (lane == 0) { red[2 * K][wave] = a; red[2 * K + 1][wave] = b; red[2 * K + 2][wave] = c; }
    }
    for (int q0 = 0; q0 < K; q0 += MD_QC) {
        double v[MD_QC][EPT];
#pragma unroll
        for (int c = 0; c < MD_QC; ++c) {
            const double* Vq = V + (int64_t)(q0 + c < K ? q0 + c : q0) * n;
#pragma unroll
            for (int e = 0; e < EPT; ++e) v[c][e] = Vq[ix[e]];
        }
#pragma unroll
        for (int c = 0; c < MD_QC; ++c) {
            double s0 = 0.0, s1 = 0.0;
#pragma unroll
            for (int e = 0; e < EPT; ++e) {
                s0 = fma(v[c][e], pr[e], s0);
                s1 = fma(v[c][e], wr[e], s1);
            }
            s0 = wave_total(s0);
            s1 = wave_total(s1);
            if (lane == 0 && q0 + c < K) { red[q0 + c][wave] = s0; red[K + q0 + c][wave] = s1; }
        }
    }
    __syncthreads();
    const int R = 2 * K + 3, nb = gridDim.x;
    for (int r = threadIdx.x; r < R; r += NT)
        part[(int64_t)r * nb + blk] = (red[r][0] + red[r][1]) + (red[r][2] + red[r][3]);
}
// elements per thread and blocks of k_multidot2 for a vector of n elements
int md_ept(int64_t n) {
    const int64_t need = (n + (int64_t)MD_BLOCKS * NT - 1) / ((int64_t)MD_BLOCKS * NT);
    return need <= 1 ? 1 : (need <= 2 ? 2 : (need <= 4 ? 4 : 8));
}
int64_t md_blocks(int64_t n) {
    const int64_t c = (int64_t)md_ept(n) * NT;
    return (n + c - 1) / c;
}
// part: k_multidot2's nb partials of each row of step j's reduction; their sums d (a[0..j), b[0..j),
// p.p, p.w, w.w) are formed here (in a fixed order) and stored for the update kernel.  cf: 1/rho, h_jj,
// 1/nu_{j+1} for the update kernel; cf[3] carries nu from step to step.  host_out[2j], [2j+1]: the
// provisional |g[j+1]| and nu_{j+1} (the convergence and breakdown tests read them).  The block sums the
// partials, sixteen lanes a row (one DPP row), with every lane's loads in flight together; then lane 0
// does the arithmetic (m <= ARN_MMAX).
constexpr int ARN_MMAX = 64;
constexpr int ARN_NT = 256;
__global__ __launch_bounds__(ARN_NT) void k_arnoldi2(Kry K, int m, int j, const double* __restrict__ part, int nb,
                                                     double* __restrict__ d, double* __restrict__ cf,
                                                     double* __restrict__ host_out) {
    __shared__ double sd[2 * ARN_MMAX + 3], sc[ARN_MMAX + 2], scs[ARN_MMAX], ssn[ARN_MMAX], sg[2], snu;
    const int lane = threadIdx.x;
    double* Hc = K.H + (int64_t)(j > 0 ? j - 1 : 0) * (m + 1);  // column j - 1
    // the rotations, column j - 1, g and nu staged first: their loads overlap the partial sums'
    for (int i = lane; i < j; i += ARN_NT) {
        scs[i] = K.cs[i];
        ssn[i] = K.sn[i];
    }
    if (j > 0)
        for (int i = lane; i < j; i += ARN_NT) sc[i] = Hc[i];
    if (lane == 0) {
        sg[0] = j > 0 ? K.g[j - 1] : 0.0;
        sg[1] = K.g[j];
        snu = cf[3];
    }
    {
        const int R = 2 * j + 3, l16 = lane & 15;
        constexpr int U = 16;
        for (int r0 = 0; r0 < R; r0 += ARN_NT / 16) {
            const int r = r0 + (lane >> 4);
            const double* pr = part + (int64_t)(r < R ? r : R - 1) * nb;
            double acc = 0.0;
            for (int b0 = l16; b0 < nb; b0 += 16 * U) {
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int b = b0 + 16 * u;
                    v[u] = pr[b < nb ? b : nb - 1];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) acc += b0 + 16 * u < nb ? v[u] : 0.0;
            }
            acc += dpp_shift<0x111, 0xf>(acc);  // row_shr 1, 2, 4, 8: lane 15 of the row holds its sum
            acc += dpp_shift<0x112, 0xf>(acc);
            acc += dpp_shift<0x114, 0xf>(acc);
            acc += dpp_shift<0x118, 0xf>(acc);
            if (l16 == 15 && r < R) {
                sd[r] = acc;
                d[r] = acc;
            }
        }
    }
    __syncthreads();
    if (lane != 0) return;
    const double* a = sd;
    const double* b = sd + j;
    const double pp = sd[2 * j], pw = sd[2 * j + 1], ww = sd[2 * j + 2];
    double rho = 1.0;  // v_0 = r / beta is final as it is
    double gj = sg[1];
    if (j > 0) {
        double sa = 0.0;
        for (int q = 0; q < j; ++q) sa += a[q] * a[q];
        const double r2 = pp - sa;
        rho = r2 > 0.0 ? sqrt(r2) : 0.0;
        // column j-1: LL z_{j-1} = V_{j-1} h + nu_j p_j = V_j (h + nu_j a) + nu_j rho v_j; then its Givens
        // rotations (0..j-2 applied, j-1 formed)
        const double nu = snu;
        for (int q = 0; q < j; ++q) sc[q] += nu * a[q];
        sc[j] = nu * rho;
        for (int i = 0; i < j - 1; ++i) {
            const double x = sc[i], y = sc[i + 1];
            sc[i] = scs[i] * x + ssn[i] * y;
            sc[i + 1] = -ssn[i] * x + scs[i] * y;
        }
        const double x = sc[j - 1], y = sc[j];
        const double rr = sqrt(x * x + y * y);
        const double c = rr > 0 ? x / rr : 1.0, sv = rr > 0 ? y / rr : 0.0;
        scs[j - 1] = c;
        ssn[j - 1] = sv;
        sc[j - 1] = rr;
        sc[j] = 0.0;
        for (int q = 0; q <= j; ++q) Hc[q] = sc[q];
        K.cs[j - 1] = c;
        K.sn[j - 1] = sv;
        const double gp = sg[0];
        gj = -sv * gp;
        K.g[j] = gj;
        K.g[j - 1] = c * gp;
    }
    double ab = 0.0, bb = 0.0;
    for (int q = 0; q < j; ++q) {
        ab += a[q] * b[q];
        bb += b[q] * b[q];
    }
    const double hjj = rho > 0.0 ? (pw - ab) / rho : 0.0;
    const double n2 = ww - bb - hjj * hjj;
    const double nu = n2 > 0.0 ? sqrt(n2) : 0.0;
    // the provisional column j (corrected by step j + 1) and its residual estimate, rotated on a copy
    double* Hj = K.H + (int64_t)j * (m + 1);
    for (int q = 0; q < j; ++q) {
        Hj[q] = b[q];
        sc[q] = b[q];
    }
    Hj[j] = hjj;
    Hj[j + 1] = nu;
    sc[j] = hjj;
    for (int i = 0; i < j; ++i) {
        const double x = sc[i], y = sc[i + 1];
        sc[i] = scs[i] * x + ssn[i] * y;
        sc[i + 1] = -ssn[i] * x + scs[i] * y;
    }
    const double rr = sqrt(sc[j] * sc[j] + nu * nu);
    host_out[2 * j] = rr > 0 ? fabs(nu / rr * gj) : 0.0;
    host_out[2 * j + 1] = nu;
    cf[0] = rho > 0.0 ? 1.0 / rho : 0.0;
    cf[1] = hjj;
    cf[2] = nu > 0.0 ? 1.0 / nu : 0.0;
    cf[3] = nu;
}
// v_j = (p_j - sum_q a_q V_q) / rho in place, p_{j+1} = (w - sum_q b_q V_q - h_jj v_j) / nu_{j+1}
__global__ __launch_bounds__(NT) void k_dcgs_update(const double* __restrict__ V, int64_t n, int K,
                                                    const double* __restrict__ d, const double* __restrict__ cf,
                                                    double* vj, const double* __restrict__ w,
                                                    double* __restrict__ pnext) {
    const int64_t i = gidx();
    if (i >= n) return;
    double sa = 0.0, sb = 0.0;
    for (int q = 0; q < K; ++q) {
        const double v = V[(int64_t)q * n + i];
        sa += d[q] * v;
        sb += d[K + q] * v;
    }
    const double v = (vj[i] - sa) * cf[0];
    vj[i] = v;
    pnext[i] = (w[i] - sb - cf[1] * v) * cf[2];
}
// The end of a cycle in one kernel: every workgroup forms y = R^-1 g for the used x used upper triangle R of H
// after the rotations, then adds Z y to x over its elements.  close: column used-1 still holds its provisional
// values (the last column of a full cycle, or a cycle that converged before its next step was enqueued) and is
// closed first: rotated by rotations 0..used-2, rotation used-1 formed from its provisional subdiagonal,
// g[used-1] scaled by its cosine.  Every workgroup runs the same operations in the same
// order, so all hold the same y; the rotations and g are not written back (the next cycle starts from the true
// residual).  Back substitution by columns on one wave (used <= ARN_MMAX = 64): lane k holds g_k;
// y_i = g_i / R_ii, then g_k -= R_ki y_i for k < i.  Dynamic LDS: R packed by columns (column k, rows 0..k, at
// k (k+1) / 2), then y.
__global__ __launch_bounds__(NT) void k_update_x_solve(Kry K, int m, int used, int close,
                                                       const double* __restrict__ Z, int64_t n,
                                                       double* __restrict__ x, const double* __restrict__ pinc,
                                                       int64_t offP) {
    extern __shared__ double sh[];
    __shared__ double s_c;
    double* R = sh;
    double* y = sh + used * (used + 1) / 2;
    for (int k = 0; k < used; ++k)
        for (int i = threadIdx.x; i <= k; i += NT) R[k * (k + 1) / 2 + i] = K.H[i + (int64_t)k * (m + 1)];
    if (threadIdx.x == 0) s_c = 1.0;
    __syncthreads();
    if (close && threadIdx.x == 0) {
        const int j = used - 1;
        double* Hc = R + j * (j + 1) / 2;
        for (int i = 0; i < j; ++i) {
            const double a = Hc[i], b = Hc[i + 1];
            Hc[i] = K.cs[i] * a + K.sn[i] * b;
            Hc[i + 1] = -K.sn[i] * a + K.cs[i] * b;
        }
        const double a = Hc[j], b = K.H[(j + 1) + (int64_t)j * (m + 1)];
        const double rr = sqrt(a * a + b * b);
        Hc[j] = rr;
        s_c = rr > 0 ? a / rr : 1.0;
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        double g = lane < used ? K.g[lane] : 0.0;
        if (lane == used - 1) g *= s_c;
        for (int i = used - 1; i >= 0; --i) {
            const double yi = __shfl(g, i) / R[i * (i + 1) / 2 + i];
            if (lane == 0) y[i] = yi;
            if (lane < i) g -= R[i * (i + 1) / 2 + lane] * yi;
        }
    }
    __syncthreads();
    // the steps' constant-pressure fixes (pinc: c_q of step q, or none): x_P += sum_q c_q y_q
    double pc = 0.0;
    if (pinc)
        for (int q = 0; q < used; ++q) pc += pinc[q] * y[q];
    // UX elements per thread, their loads of each Z column in flight together (each element's sum in the
    // order x + y_0 z_0 + y_1 z_1 + ...)
    constexpr int UX = 4;
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i0 = (int64_t)blockIdx.x * NT + threadIdx.x; i0 < n; i0 += UX * stride) {
        double sx[UX];
        int64_t ix[UX];
#pragma unroll
        for (int e = 0; e < UX; ++e) {
            ix[e] = i0 + e * stride < n ? i0 + e * stride : n - 1;
            sx[e] = x[ix[e]];
        }
        for (int q = 0; q < used; ++q) {
            const double yq = y[q];
            double zq[UX];
#pragma unroll
            for (int e = 0; e < UX; ++e) zq[e] = Z[(int64_t)q * n + ix[e]];
#pragma unroll
            for (int e = 0; e < UX; ++e) sx[e] += yq * zq[e];
        }
#pragma unroll
        for (int e = 0; e < UX; ++e) {
            if (i0 + e * stride >= n) continue;
            double v = sx[e];
            if (pinc && ix[e] >= offP) v += pc;
            x[ix[e]] = v;
        }
    }
}
// x_P += c for every pressure entry (LL 1_P = e_pin: only the pin row's residual moves, by -c)
__global__ __launch_bounds__(NT) void k_p_shift(double* __restrict__ xP, int64_t nP, double c) {
    const int64_t i = gidx();
    if (i < nP) xP[i] += c;
}
__global__ __launch_bounds__(NT) void k_zero(double* a, int64_t n) {
    const int64_t i = gidx();
    if (i < n) a[i] = 0.0;
}
// y += x
__global__ __launch_bounds__(NT) void k_axpy1(const double* __restrict__ x, double* __restrict__ y, int64_t n) {
    const int64_t i = gidx();
    if (i < n) y[i] += x[i];
}
__global__ __launch_bounds__(NT) void k_copyv(const double* __restrict__ a, double* __restrict__ b, int64_t n) {
    const int64_t i = gidx();
    if (i < n) b[i] = a[i];
}

// ---- state <-> stacked vector
__global__ __launch_bounds__(NT) void k_gather(xc_mat W, xc_mat U, xc_mat P, double* x, int Nz, int Nx) {
    const int64_t k = gidx();
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1), nP = (int64_t)(Nz + 2) * (Nx + 2);
    if (k < nW) x[k] = W.p[(k % (Nz + 1)) + (k / (Nz + 1)) * W.ld];
    else if (k < nW + nU) { const int64_t q = k - nW; x[k] = U.p[(q % (Nz + 2)) + (q / (Nz + 2)) * U.ld]; }
    else if (k < nW + nU + nP) { const int64_t q = k - nW - nU; x[k] = P.p[(q % (Nz + 2)) + (q / (Nz + 2)) * P.ld]; }
}
// shift: P = x_P + cP (the pressure level of a pin-fixed solve, fgmres_run's pshift)
__global__ __launch_bounds__(NT) void k_scatter(const double* x, xc_mat W, xc_mat U, xc_mat P, int Nz, int Nx,
                                                int shift = 0, double cP = 0.0) {
    const int64_t k = gidx();
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1), nP = (int64_t)(Nz + 2) * (Nx + 2);
    if (k < nW) W.p[(k % (Nz + 1)) + (k / (Nz + 1)) * W.ld] = x[k];
    else if (k < nW + nU) { const int64_t q = k - nW; U.p[(q % (Nz + 2)) + (q / (Nz + 2)) * U.ld] = x[k]; }
    else if (k < nW + nU + nP) {
        const int64_t q = k - nW - nU;
        P.p[(q % (Nz + 2)) + (q / (Nz + 2)) * P.ld] = shift ? x[k] + cP : x[k];
    }
}
// dense copy of a strided matrix

// the fine level's five coefficient fields packed dense (eta, rho, etaco, rhow, rhou), their single-
// precision copies at the same offsets from the first (fo = f_coef - dst[0]), and per-block max(eta),
// min(rho) partials (part[2b], part[2b+1]) for the V-cycle depth
struct Pack5 {
    xc_mat a[5];
    double* dst[5];
    float* f;        // single-precision copy of dst[0]'s region
    const double* base;
    double* part;
};
__global__ __launch_bounds__(NT) void k_pack5(Pack5 p) {
    const int64_t k = gidx();
    double emax = 0.0, rmin = 1e308;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const xc_mat& a = p.a[q];
        if (k < a.rows * a.cols) {
            const double v = xc_mat_at(a, k);
            p.dst[q][k] = v;
            p.f[(p.dst[q] - p.base) + k] = (float)v;
            if (q == 0) emax = v;
            if (q == 1) rmin = v;
        }
    }
    emax = -wave_min(-emax);
    rmin = wave_min(rmin);
    __shared__ double se[NT / 64], sr[NT / 64];
    if ((threadIdx.x & 63) == 0) { se[threadIdx.x >> 6] = emax; sr[threadIdx.x >> 6] = rmin; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < NT / 64; ++w) { emax = fmax(emax, se[w]); rmin = fmin(rmin, sr[w]); }
        p.part[2 * blockIdx.x] = emax;
        p.part[2 * blockIdx.x + 1] = rmin;
    }
}

// ---- right-hand side (fluidmech.jl:111-127, 188-190, 236-246, 310-323, 450-476)
struct RhsArgs {
    xc_mat rho, rhoo, rhooo, drhodt, drhodto, drhodtoo, MFS, advn_rho;  // MFS update
    double a1, a2, a3, b1, b2, b3, dt, alpha;
    int do_mfs;
};
__global__ __launch_bounds__(NT) void k_mfs(RhsArgs a, int Nz, int Nx, double* part) {
    const int64_t k = gidx();
    double v = 0.0;
    if (k < (int64_t)Nz * Nx) {
        const int j = ((int)k % Nz), i = ((int)k / Nz);
        auto at = [&](const xc_mat& m) -> double& { return m.p[j + (int64_t)i * m.ld]; };
        if (a.do_mfs) {
#pragma clang fp contract(off)
            const double drhodt = -at(a.advn_rho);
            at(a.drhodt) = drhodt;
            const double res = (a.a1 * at(a.rho) - a.a2 * at(a.rhoo) - a.a3 * at(a.rhooo)) / a.dt -
                               (a.b1 * drhodt + a.b2 * at(a.drhodto) + a.b3 * at(a.drhodtoo));
            at(a.MFS) = at(a.MFS) - a.alpha * res;
        }
        v = at(a.MFS);
    }
    __shared__ double red[NT];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ __launch_bounds__(NT) void k_sum_final(const double* part, int nb, double scale, double* out) {
    __shared__ double red[NT];
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) acc += part[b];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// momentum-advection inputs (fluidmech.jl:236-238, 310-312)
__global__ __launch_bounds__(NT) void k_mom_prep(xc_mat W, xc_mat U, xc_mat rhow, xc_mat rhou, double* fz, double* uz,
                                                 double* wz, double* fx, double* ux, double* wx, int Nz, int Nx) {
    const int64_t k = gidx();
    auto Wv = [&](int r, int c) { return W.p[r + (int64_t)c * W.ld]; };
    auto Uv = [&](int r, int c) { return U.p[r + (int64_t)c * U.ld]; };
    auto ifx = [&](int kk) { return kk == 0 ? Nx - 1 : (kk == Nx + 2 ? 1 : kk - 1); };
    // fz (Nz-1, Nx), uz (Nz-1, Nx+1), wz (Nz, Nx), fx (Nz, Nx+1), ux (Nz, Nx+2), wx (Nz+1, Nx+1)
    if (k < (int64_t)(Nz - 1) * Nx) {
        const int j = ((int)k % (Nz - 1)), i = ((int)k / (Nz - 1));
        fz[k] = rhow.p[(j + 1) + (int64_t)i * rhow.ld] * Wv(j + 1, i + 1);
    }
    if (k < (int64_t)(Nz - 1) * (Nx + 1)) {
        const int j = ((int)k % (Nz - 1)), i = ((int)k / (Nz - 1));
        uz[k] = (Uv(j + 1, i) + Uv(j + 2, i)) / 2.0;
    }
    if (k < (int64_t)Nz * Nx) {
        const int j = ((int)k % Nz), i = ((int)k / Nz);
        wz[k] = (Wv(j, i + 1) + Wv(j + 1, i + 1)) / 2.0;
    }
    if (k < (int64_t)Nz * (Nx + 1)) {
        const int j = ((int)k % Nz), i = ((int)k / Nz);
        fx[k] = rhou.p[j + (int64_t)i * rhou.ld] * Uv(j + 1, i);
    }
    if (k < (int64_t)Nz * (Nx + 2)) {
        const int j = ((int)k % Nz), i = ((int)k / Nz);
        ux[k] = (Uv(j + 1, ifx(i)) + Uv(j + 1, ifx(i + 1))) / 2.0;
    }
    if (k < (int64_t)(Nz + 1) * (Nx + 1)) {
        const int j = ((int)k % (Nz + 1)), i = ((int)k / (Nz + 1));
        wx[k] = (Wv(j, i) + Wv(j, i + 1)) / 2.0;
    }
}

struct RRArgs {
    xc_mat Drho, rhoWo, rhoWoo, rhoUo, rhoUoo, MFS, rhow;
    const double *advz, *advx;   // (Nz-1, Nx), (Nz, Nx+1) dense
    const double* mfsmean;
    xc_mat srcW, srcU, srcP;     // bnchm only
    double a2, a3, dt, g0, D, h, oc;
    int bnchm;
    // optional: the starting guess gathered from the state in the same pass (k_gather's work)
    xc_mat gW, gU, gP;
    double* gx;
};
__global__ __launch_bounds__(NT) void k_rr(RRArgs a, int Nz, int Nx, double* RR) {
#pragma clang fp contract(off)
    const int64_t k = gidx();
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1), nP = (int64_t)(Nz + 2) * (Nx + 2);
    if (k >= nW + nU + nP) return;
    auto at = [](const xc_mat& m, int r, int c) { return m.p[r + (int64_t)c * m.ld]; };
    double v = 0.0;
    if (k < nW) {
        const int r = ((int)k % (Nz + 1)), c = ((int)k / (Nz + 1));
        if (c >= 1 && c <= Nx) {
            if (r == Nz) {
                const double mfbg = a.mfsmean[0] * a.D / a.h;
                v = a.oc * at(a.MFS, Nz - 1, c - 1) + (1.0 - a.oc) * mfbg;
            } else if (r >= 1) {
                const double adv = (c - 1 >= 0 && c - 1 < Nx) ? a.advz[(r - 1) + (int64_t)(c - 1) * (Nz - 1)] : 0.0;
                v = at(a.Drho, r, c - 1) * a.g0 + (a.a2 * at(a.rhoWo, r, c - 1) + a.a3 * at(a.rhoWoo, r, c - 1)) / a.dt - adv;
                if (a.bnchm) v += at(a.srcW, r, c);
            }
        }
    } else if (k < nW + nU) {
        const int64_t q = k - nW;
        const int r = ((int)q % (Nz + 2)), c = ((int)q / (Nz + 2));
        if (r >= 1 && r <= Nz) {
            double adv = a.advx[(r - 1) + (int64_t)c * Nz];
            if (c == 0 || c == Nx) adv = (a.advx[(r - 1)] + a.advx[(r - 1) + (int64_t)Nx * Nz]) / 2.0;
            v = (a.a2 * at(a.rhoUo, r - 1, c) + a.a3 * at(a.rhoUoo, r - 1, c)) / a.dt - adv;
            if (a.bnchm) v += at(a.srcU, r, c);
        }
    } else {
        const int64_t q = k - nW - nU;
        const int r = ((int)q % (Nz + 2)), c = ((int)q / (Nz + 2));
        if (r >= 1 && r <= Nz && c >= 1 && c <= Nx) {
            v = at(a.MFS, r - 1, c - 1);
            if (a.bnchm) v += at(a.srcP, r, c);
        }
    }
    RR[k] = v;
    if (a.gx) {
        if (k < nW) a.gx[k] = at(a.gW, (int)(k % (Nz + 1)), (int)(k / (Nz + 1)));
        else if (k < nW + nU) { const int64_t q = k - nW; a.gx[k] = at(a.gU, (int)(q % (Nz + 2)), (int)(q / (Nz + 2))); }
        else { const int64_t q = k - nW - nU; a.gx[k] = at(a.gP, (int)(q % (Nz + 2)), (int)(q / (Nz + 2))); }
    }
}
// MMS pins of the right-hand side (fluidmech.jl:468, 475)
__global__ void k_rr_pins(double* RR, int64_t kp, xc_mat Pex, int pr, int pc, int64_t ku, xc_mat Uex, int ur, int uc) {
    RR[kp] = Pex.p[pr + (int64_t)pc * Pex.ld];
    RR[ku] = Uex.p[ur + (int64_t)uc * Uex.ld];
}

// ---------------------------------------------------------------- host side
struct Level {
    int Nz, Nx;
    double h;
    double *eta, *etaco, *rho, *rhow, *rhou;  // device (owned for coarse levels; fine level: packed copies)
    double *x, *b, *r, *t;                    // velocity vectors
    int64_t nv;
    LevD d() const {
        LevD L;
        L.Nz = Nz; L.Nx = Nx; L.h = h; L.invh = 1.0 / h; L.invh2 = (1.0 / h) * (1.0 / h);
        L.eta = eta; L.etaco = etaco; L.rho = rho; L.rhow = rhow; L.rhou = rhou;
        return L;
    }
};

struct StokesCtx {
    int Nz = -1, Nx = -1, m = -1, maxlev = 0;
    std::vector<Level> lev;
    // the fine level's smoother in single precision: its coefficients (the level's five arrays at the same
    // offsets from eta) and vectors x, b, t (ping-pong), r
    float *f_coef = nullptr, *fx = nullptr, *fb = nullptr, *ft = nullptr, *fr = nullptr;
    int64_t f_ncoef = 0;
    unsigned pack_blocks = 0;  // blocks of the last k_pack5 launch (its max(eta) / min(rho) partials)
    LevF levf() const {
        const Level& L = lev[0];
        LevF f;
        f.Nz = L.Nz; f.Nx = L.Nx; f.h = L.h; f.invh = 1.0 / L.h; f.invh2 = (1.0 / L.h) * (1.0 / L.h);
        f.eta = f_coef; f.etaco = f_coef + (L.etaco - L.eta); f.rho = f_coef + (L.rho - L.eta);
        f.rhow = f_coef + (L.rhow - L.eta); f.rhou = f_coef + (L.rhou - L.eta);
        return f;
    }
    double* pool = nullptr;   // everything below carved from one allocation
    size_t pool_bytes = 0;
    double *xs, *bs, *rs, *ws, *V, *Z, *part, *dots, *tmp, *adv_buf;
    double* host = nullptr;   // pinned host scratch for reductions (mapped, coherent)
    double* host_dev = nullptr;  // the same memory as seen by kernels
    double* pk_host = nullptr;   // k_pack5's per-block max(eta) / min(rho) partials (mapped, coherent), read by
    double* pk_dev = nullptr;    // choose_depth on the host after the solve's first synchronization
    double* kry = nullptr;    // Arnoldi H, Givens, g, y (device)
    double* d3 = nullptr;     // dot-product results: 2m + 3 of one Arnoldi step, then 4 coefficients
    hipEvent_t ev[2] = {nullptr, nullptr};
    // pressure Poisson (inertial Schur part)
    int pK = 0;                 // Nx/2 + 1
    double2* pS = nullptr;      // Nz x pK half spectra
    double* plamx = nullptr;    // pK x-eigenvalues (normalised by cz)
    double* plamz = nullptr;    // Nz z-eigenvalues of the cosine transform
    double* pg = nullptr;       // open_cnv: K x Nz correction vectors g_k of the pinned bottom row
    double pgtab = -1.0;        // cx / cz the g table holds
    double2* ptw = nullptr;     // Nx twiddles
    double2* ptwz = nullptr;    // Nz twiddles of the z plan, then Nz cosine-transform twiddles
    std::vector<int> prad, pradz;
    double pratio = -1.0;       // cx / cz of this solve; <= 0: inertial part off
    double ptab = -1.0;         // cx / cz the device eigenvalue table holds
    double pscale = 0.0;        // a1 h^2 / (dt Nx)
    // learned coarse correction
    float* lc_in = nullptr;     // [cells][4] network input
    float* lc_sigma = nullptr;  // per-call input scale (device)
    double* lc_part = nullptr;  // restriction partial sums
    int lc_cap = 0;             // capacity in coarse cells
    // the V-cycle's one-workgroup tail: from level tail_l on (-1: none), its launch arguments and LDS bytes
    int tail_l = -1;
    TailArgs tail{};
    size_t tail_lds = 0;
    // depth of the V-cycle for the current solve (levels 0 .. leff-1) and the inertia ratio it came from
    int leff = 0;
    double inertia = 0.0;
    // the plain preconditioner's mean residual reduction per FGMRES iteration over recent solves (0: none yet),
    // the yardstick a learned step is measured against; learned solves dropped at their first step in a row
    // and the number of solves still to run plain before the learned correction is tried again
    double plain_rate = 0.0;
    int lc_fail_streak = 0, lc_skip = 0;
    int lc_backoff = 0;  // skip periods in a row that ended in another dropped retry (the next one is doubled)
    // horizontal-mean correction of each FGMRES cycle's starting guess (see xm_build): the Galerkin operator
    // R^T LL R on the x-uniform fields, factorized on the host (banded LU, rows ordered W, U, P per z), its
    // mapped host scratch (q = R^T r, delta, the probe sums) and device buffers (partial sums, delta)
    int xm_n0 = 0, xm_nch = 0;
    double* xm_host = nullptr;
    double* xm_hdev = nullptr;
    double* xm_part = nullptr;
    double* xm_delta = nullptr;
    double* xm_pinc = nullptr;  // the FGMRES steps' constant-pressure fixes (m)
    double* xm_b2 = nullptr;    // ||r||^2 of the corrected starting guess (device; the cycle start reads it)
    std::vector<double> xm_lu;
    std::vector<int> xm_piv;
    bool xm_built = false;   // a probing solve has run on this grid
    bool xm_ok = false;      // ... and its factorization is held
    bool xm_active = false;  // this solve corrects its cycles' starting guesses
};

int pois_open_table(xc_ctx* ctx, StokesCtx* s, int Nz, int Nx);

// Decide whether the inertial Schur part applies to this solve and (re)build its eigenvalue tables.
int pois_table(xc_ctx* ctx, StokesCtx* s, const Flags& F, int Nz, int Nx, double h) {
    s->pratio = -1.0;
    if (F.bnchm || !(F.dt < 1e30) || !(F.a1 > 0) || s->prad.empty() || s->pradz.empty()) return XC_OK;
    const double ratio = F.a1 / (F.a1 + F.gamma);
    s->pscale = F.a1 * h * h / (F.dt * (double)Nx);
    s->pratio = ratio;
    if (ratio == s->ptab) return F.open_cnv ? pois_open_table(ctx, s, Nz, Nx) : XC_OK;
    const int K = s->pK;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<double> lam((size_t)K + (size_t)Nz);
    for (int k = 0; k < K; ++k)
        lam[k] = (double)((long double)ratio * (2.0L - 2.0L * cosl(2.0L * pi * (long double)k / (long double)Nx)));
    for (int kz = 0; kz < Nz; ++kz)
        lam[K + kz] = (double)(2.0L - 2.0L * cosl(pi * (long double)kz / (long double)Nz));
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->plamx, lam.data(), lam.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->ptab = ratio;
    return F.open_cnv ? pois_open_table(ctx, s, Nz, Nx) : XC_OK;
}

// open_cnv's correction vectors: g_k = (T_N + lam_x[k])^-1 e_last for k > 0 (tridiagonal elimination,
// diagonally dominant), g_0 = T_N^+ e_last (zero-flux Laplacian, pseudo-inverse: the zero-mean solution of
// T_N y = e_last - 1/Nz, by its face fluxes f_i = sum_{j <= i} b_j, y_{i+1} = y_i - f_i), in long double
int pois_open_table(xc_ctx* ctx, StokesCtx* s, int Nz, int Nx) {
    if (s->pgtab == s->ptab) return XC_OK;
    const int K = s->pK;
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<double> g((size_t)K * Nz);
    std::vector<long double> y((size_t)Nz), cp((size_t)Nz);
    for (int k = 0; k < K; ++k) {
        if (k == 0) {
            long double f = 0.0L, mean = 0.0L;
            y[0] = 0.0L;
            for (int i = 0; i + 1 < Nz; ++i) {
                f += (i == Nz - 1 ? 1.0L : 0.0L) - 1.0L / (long double)Nz;
                y[i + 1] = y[i] - f;
            }
            for (int i = 0; i < Nz; ++i) mean += y[i];
            mean /= (long double)Nz;
            for (int i = 0; i < Nz; ++i) g[i] = (double)(y[i] - mean);
            continue;
        }
        const long double lam = (long double)s->ptab * (2.0L - 2.0L * cosl(2.0L * pi * (long double)k / (long double)Nx));
        // Thomas: diagonal 1 + lam at both ends (2 + lam inside), off-diagonals -1, right-hand side e_last
        auto dg = [&](int i) { return (i == 0 || i == Nz - 1 ? 1.0L : 2.0L) + lam; };
        if (Nz == 1) { g[(size_t)k * Nz] = (double)(1.0L / dg(0)); continue; }
        long double den = dg(0);
        cp[0] = -1.0L / den;
        y[0] = 0.0L;
        for (int i = 1; i < Nz; ++i) {
            den = dg(i) + cp[i - 1];  // dg(i) - (-1) cp[i-1]
            cp[i] = -1.0L / den;
            y[i] = ((i == Nz - 1 ? 1.0L : 0.0L) + y[i - 1]) / den;
        }
        for (int i = Nz - 2; i >= 0; --i) y[i] -= cp[i] * y[i + 1];
        for (int i = 0; i < Nz; ++i) g[(size_t)k * Nz + i] = (double)y[i];
    }
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->pg, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    s->pgtab = s->ptab;
    return XC_OK;
}

void stokes_free(xc_ctx*, void* p) {
    StokesCtx* s = (StokesCtx*)p;
    if (s->pool) (void)hipFree(s->pool);
    if (s->host) (void)hipHostFree(s->host);
    if (s->pk_host) (void)hipHostFree(s->pk_host);
    if (s->xm_host) (void)hipHostFree(s->xm_host);
    if (s->xm_part) (void)hipFree(s->xm_part);
    for (auto& e : s->ev)
        if (e) (void)hipEventDestroy(e);
    delete s;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

// the tail of the V-cycle: the first level l >= 1 of at most TAIL_CELLS cells with at most TAIL_MAX
// levels from it to the coarsest, if its coefficient image, vectors and right-hand side fit
void tail_setup(StokesCtx* s) {
    s->tail_l = -1;
    const int nl = (int)s->lev.size();
    int l0 = -1;
    for (int l = 1; l < nl; ++l)
        if ((int64_t)s->lev[l].Nz * s->lev[l].Nx <= TAIL_CELLS && nl - l <= TAIL_MAX) { l0 = l; break; }
    if (l0 < 0) return;
    TailArgs& ta = s->tail;
    ta = TailArgs{};
    ta.nlev = nl - l0;
    const double* img = s->lev[l0].eta;
    const Level& last = s->lev[nl - 1];
    ta.T = (int)((last.rhou + (((size_t)last.Nz * (last.Nx + 1) + 1) & ~(size_t)1)) - img);
    int o = (ta.T + 1) & ~1;
    ta.nbig = 0;
    for (int q = 0; q < ta.nlev; ++q) {
        const Level& L = s->lev[l0 + q];
        ta.d[q] = L.d();
        ta.off[q][0] = (int)(L.eta - img);
        ta.off[q][1] = (int)(L.etaco - img);
        ta.off[q][2] = (int)(L.rho - img);
        ta.off[q][3] = (int)(L.rhow - img);
        ta.off[q][4] = (int)(L.rhou - img);
        for (int k = 5; k < 9; ++k) {
            ta.off[q][k] = o;
            o += ((int)L.nv + 1) & ~1;
        }
        if (L.nv > TAIL_WAVE_PTS && ta.nbig == q) ta.nbig = q + 1;
    }
    ta.img = img;
    const size_t bytes = (size_t)o * sizeof(double);
    if (bytes > 65536 || ta.T > TAIL_IMG_PER_THREAD * TAIL_NT || s->lev[l0].nv > TAIL_NT) return;
    s->tail_lds = bytes;
    s->tail_l = l0;
}

// maxlev: most multigrid levels (0: all; < 0: whatever the existing hierarchy has)
int setup(xc_ctx* ctx, int Nz, int Nx, int m, StokesCtx** out, int maxlev = -1) {
    StokesCtx* s = (StokesCtx*)ctx->ext[XC_EXT_STOKES];
    if (s && s->Nz == Nz && s->Nx == Nx && s->m >= m && (maxlev < 0 || s->maxlev == maxlev)) {
        *out = s;
        return XC_OK;
    }
    if (maxlev < 0) maxlev = s ? s->maxlev : 0;
    if (s) { XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream)); stokes_free(ctx, s); ctx->ext[XC_EXT_STOKES] = nullptr; }
    s = new StokesCtx();
    s->Nz = Nz; s->Nx = Nx; s->m = m; s->maxlev = maxlev;
    // levels: halve while both sizes stay even and >= 8 (coarsest at least 4x4), at most maxlev of them
    int nz = Nz, nx = Nx;
    s->lev.push_back(Level{});
    s->lev.back().Nz = nz; s->lev.back().Nx = nx;
    const size_t cap = maxlev > 0 && maxlev < 12 ? (size_t)maxlev : 12;
    while (nz % 2 == 0 && nx % 2 == 0 && nz >= 8 && nx >= 8 && s->lev.size() < cap) {
        nz /= 2; nx /= 2;
        Level L{};
        L.Nz = nz; L.Nx = nx;
        s->lev.push_back(L);
    }
    const int64_t n = (int64_t)(Nz + 1) * (Nx + 2) + (int64_t)(Nz + 2) * (Nx + 1) + (int64_t)(Nz + 2) * (Nx + 2);
    size_t tot = 0;
    for (auto& L : s->lev) {
        L.nv = (int64_t)(L.Nz + 1) * (L.Nx + 2) + (int64_t)(L.Nz + 2) * (L.Nx + 1);
        tot += (size_t)L.Nz * L.Nx * 2 + (size_t)(L.Nz + 1) * (L.Nx + 1) + (size_t)(L.Nz + 1) * L.Nx +
               (size_t)L.Nz * (L.Nx + 1) + 5 /* even rounding */ + 4 * (size_t)L.nv;
    }
    const int K = Nx / 2 + 1;
    const size_t lcc = (size_t)(Nz / 2) * (Nx / 2);  // learned correction at level >= 1
    tot += 3 * lcc + 16 + (lcc + NT - 1) / NT + 8;
    size_t part_cols = md_blocks(n) > DOT_BLOCKS ? (size_t)md_blocks(n) : (size_t)DOT_BLOCKS;
    {  // also room for k_ll_cells' per-wave partials behind the first DOT_BLOCKS (a solve's first residual is
       // reduced while ||RR||'s partials are)
        const size_t waves = 2 * (size_t)ll_cells_blocks(Nz, Nx) * (NT / 64) + DOT_BLOCKS;  // r and b sums
        const size_t rows = 2 * (size_t)m + 4;
        if (rows * part_cols < waves) part_cols = (waves + rows - 1) / rows;
    }
    tot += 4 * (size_t)n + (size_t)(m + 1) * n + (size_t)m * n + (size_t)(2 * m + 4) * part_cols + 128 /* dots */ +
           (size_t)n /* tmp */ + 8 * (size_t)(Nz + 2) * (Nx + 3) /* advection scratch */ +
           3 * (size_t)Nz * K + 2 * (size_t)Nx + 4 * (size_t)Nz + K + 2 * (size_t)Nz + 16 /* Poisson */ +
           (size_t)(m + 1) * m + 4 * (size_t)m + 8 /* Arnoldi */ + 4 * (size_t)(m + 2) /* dots */;
    const Level& F0 = s->lev[0];
    const size_t ncoef0 = 2 * (size_t)F0.Nz * F0.Nx + (size_t)(F0.Nz + 1) * (F0.Nx + 1) + (size_t)(F0.Nz + 1) * F0.Nx +
                          (size_t)F0.Nz * (F0.Nx + 1) + 10 /* even rounding */;
    tot += (ncoef0 + 4 * (size_t)F0.nv) / 2 + 16;  /* single-precision fine-level smoother */
    void* pool;
    hipError_t e = hipMalloc(&pool, tot * sizeof(double));
    if (e != hipSuccess) { delete s; return xc_fail(ctx, XC_ENOMEM, "fluidmech!: %zu bytes: %s", tot * 8, hipGetErrorString(e)); }
    s->pool = (double*)pool;
    s->pool_bytes = tot * sizeof(double);
    double* p = s->pool;
    auto take = [&](size_t k) { double* q = p; p += k; return q; };
    // coefficients of every level first, in the order and even padding of the tail's LDS layout, so
    // that the tail levels' coefficients are one contiguous block
    auto take2 = [&](size_t k) { return take((k + 1) & ~(size_t)1); };
    for (auto& L : s->lev) {
        L.eta = take2((size_t)L.Nz * L.Nx);
        L.etaco = take2((size_t)(L.Nz + 1) * (L.Nx + 1));
        L.rho = take2((size_t)L.Nz * L.Nx);
        L.rhow = take2((size_t)(L.Nz + 1) * L.Nx);
        L.rhou = take2((size_t)L.Nz * (L.Nx + 1));
    }
    for (auto& L : s->lev) {
        L.x = take(L.nv); L.b = take(L.nv); L.r = take(L.nv); L.t = take(L.nv);
    }
    s->xs = take(n); s->bs = take(n); s->rs = take(n); s->ws = take(n);
    s->V = take((size_t)(m + 1) * n);
    s->Z = take((size_t)m * n);
    s->part = take((size_t)(2 * m + 4) * part_cols);
    s->dots = take(128);
    s->tmp = take(n);
    s->adv_buf = take(8 * (size_t)(Nz + 2) * (Nx + 3));
    s->lc_in = reinterpret_cast<float*>(take(2 * lcc + 8));   // 4 floats per cell
    s->lc_sigma = reinterpret_cast<float*>(take(8));
    s->lc_part = take(lcc / NT + 8 + (lcc + NT - 1) / NT);
    s->lc_cap = (int)lcc;
    s->kry = take((size_t)(m + 1) * m + 4 * (size_t)m + 8);
    s->d3 = take(4 * (size_t)(m + 2));
    {
        const Level& L0 = s->lev[0];
        s->f_ncoef = (L0.rhou + (size_t)L0.Nz * (L0.Nx + 1)) - L0.eta;
        const size_t fl = ((size_t)s->f_ncoef + 3) & ~(size_t)3, fv = ((size_t)L0.nv + 3) & ~(size_t)3;
        float* f = reinterpret_cast<float*>(take((fl + 4 * fv) / 2));
        s->f_coef = f;
        s->fx = f + fl;
        s->fb = s->fx + fv;
        s->ft = s->fb + fv;
        s->fr = s->ft + fv;
    }
    if ((p - s->pool) & 1) ++p;  // 16-byte alignment for the complex arrays
    s->pK = K;
    s->pS = reinterpret_cast<double2*>(take(2 * (size_t)Nz * K));
    s->ptw = reinterpret_cast<double2*>(take(2 * (size_t)Nx));
    s->ptwz = reinterpret_cast<double2*>(take(4 * (size_t)Nz));
    s->plamx = take((size_t)K + 2 * (size_t)Nz);
    s->plamz = s->plamx + K;
    s->pg = take((size_t)Nz * K);
    if ((size_t)(p - s->pool) > tot) {  // carving overran the allocation: a sizing bug, never a user error
        stokes_free(ctx, s);
        return xc_fail(ctx, XC_EHIP, "fluidmech!: internal pool sizing (%zu > %zu)", (size_t)(p - s->pool), tot);
    }
    tail_setup(s);
    if (Nx > MAXLINE || Nz > MAXLINE || (Nz / 2 + 1) > ZD_PER * FFT_NT || !fft_factorize(Nx, s->prad) ||
        !fft_factorize(Nz, s->pradz)) {
        s->prad.clear();  // inertial Schur part unavailable at this size: the viscous part alone is used
        s->pradz.clear();
    } else {
        std::vector<double2> tw;
        fft_twiddles(Nx, tw);
        e = hipMemcpy(s->ptw, tw.data(), Nx * sizeof(double2), hipMemcpyHostToDevice);
        if (e == hipSuccess) {  // the Nz-point plan's twiddles, then the cosine transform's w_k
            fft_twiddles(Nz, tw);
            tw.resize(2 * (size_t)Nz);
            const long double pi = 3.141592653589793238462643383279502884L;
            for (int k = 0; k < Nz; ++k) {
                const long double ang = -pi * (long double)k / (2.0L * (long double)Nz);
                tw[Nz + k] = make_double2((double)cosl(ang), (double)sinl(ang));
            }
            e = hipMemcpy(s->ptwz, tw.data(), 2 * (size_t)Nz * sizeof(double2), hipMemcpyHostToDevice);
        }
        if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_EHIP, "fluidmech!: %s", hipGetErrorString(e)); }
    }
    {  // every k_pack5 launch of this grid: its blocks cover the largest of the five coefficient arrays
        const int64_t most = (int64_t)(Nz + 1) * (Nx + 2) > (int64_t)(Nz + 2) * (Nx + 1) ? (int64_t)(Nz + 1) * (Nx + 2)
                                                                                         : (int64_t)(Nz + 2) * (Nx + 1);
        e = hipHostMalloc((void**)&s->pk_host, 2 * (size_t)nblk(most) * sizeof(double),
                          hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer((void**)&s->pk_dev, s->pk_host, 0);
        if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_ENOMEM, "fluidmech!: pinned scratch"); }
    }
    e = hipHostMalloc((void**)&s->host, 256 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_ENOMEM, "fluidmech!: pinned scratch"); }
    e = hipHostGetDevicePointer((void**)&s->host_dev, s->host, 0);
    if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_EHIP, "fluidmech!: %s", hipGetErrorString(e)); }
    s->xm_n0 = 3 * Nz + 5;
    s->xm_nch = (Nx + 2 + XM_CH - 1) / XM_CH;
    e = hipHostMalloc((void**)&s->xm_host, (size_t)(2 + XM_PROBES) * s->xm_n0 * sizeof(double),
                      hipHostMallocMapped | hipHostMallocCoherent);
    if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_ENOMEM, "fluidmech!: pinned scratch"); }
    e = hipHostGetDevicePointer((void**)&s->xm_hdev, s->xm_host, 0);
    if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_EHIP, "fluidmech!: %s", hipGetErrorString(e)); }
    e = hipMalloc((void**)&s->xm_part, ((size_t)(XM_PROBES * s->xm_nch + 1) * s->xm_n0 + (size_t)m + 8) * sizeof(double));
    if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_ENOMEM, "fluidmech!: horizontal-mean scratch"); }
    s->xm_delta = s->xm_part + (size_t)XM_PROBES * s->xm_nch * s->xm_n0;
    s->xm_pinc = s->xm_delta + s->xm_n0;
    s->xm_b2 = s->xm_pinc + m;
    for (auto& ev : s->ev) {
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) { stokes_free(ctx, s); return xc_fail(ctx, XC_EHIP, "fluidmech!: event"); }
    }
    ctx->ext[XC_EXT_STOKES] = s;
    ctx->ext_free[XC_EXT_STOKES] = stokes_free;
    *out = s;
    return XC_OK;
}

Flags make_flags(const xc_fm_opts* o, const xc_par* par, int Nz, int Nx, bool bnchm) {
    Flags F{};
    F.bnchm = bnchm;
    F.sds = o->sds; F.top_cnv = o->top_cnv; F.bot_cnv = o->bot_cnv; F.open_cnv = o->open_cnv;
    F.a1 = o->a1; F.gamma = par->gamma; F.dt = o->dt; F.g0 = par->g0;
    F.a1_dt = F.a1 / F.dt;
    F.a1g_dt = (F.a1 + F.gamma) / F.dt;
    F.nu0 = -1;
    auto jround = [](double v) { return (int)nearbyint(v); };
    if (bnchm) {
        F.pin_kind = 1;
        F.pr = jround((double)(Nz + 2) * 3.0 / 8.0) - 1;
        F.pc = jround((double)(Nx + 2) / 2.0) - 1;
        const int ur = jround((double)(Nz + 2) / 2.0) - 1, uc = jround((double)(Nx + 2) / 2.0) - 1;
        F.nu0 = (int64_t)(Nz + 1) * (Nx + 2) + ur + (int64_t)uc * (Nz + 2);
    } else if (o->open_cnv) {
        F.pin_kind = 2;
    } else {
        F.pin_kind = 1;
        F.pr = jround((double)Nz / 2.0) - 1;
        F.pc = jround((double)Nx / 2.0) - 1;
    }
    return F;
}

// copy the fine-level coefficients densely and build the coarse hierarchy
void build_coarse(xc_ctx* ctx, StokesCtx* s, int upto);
// the finest level's coefficients (and its single-precision copy); all: the coarser levels too
void build_levels(xc_ctx* ctx, StokesCtx* s, const xc_fluid* fl, double h, bool all = true) {
    s->leff = (int)s->lev.size();  // the full V-cycle unless choose_depth says otherwise
    s->inertia = 0.0;
    Level& L0 = s->lev[0];
    L0.h = h;
    Pack5 pk{{fl->eta, fl->rho, fl->etaco, fl->rhow, fl->rhou}, {L0.eta, L0.rho, L0.etaco, L0.rhow, L0.rhou},
             s->f_coef, L0.eta, s->pk_dev};
    int64_t most = 0;
    for (int q = 0; q < 5; ++q) most = pk.a[q].rows * pk.a[q].cols > most ? pk.a[q].rows * pk.a[q].cols : most;
    s->pack_blocks = nblk(most);
    hipLaunchKernelGGL(k_pack5, dim3(s->pack_blocks), dim3(NT), 0, ctx->stream, pk);
    if (all) build_coarse(ctx, s, (int)s->lev.size());
}

// coefficients of levels 1 .. upto-1 from the finer ones (the levels a cycle of depth upto reads)
void build_coarse(xc_ctx* ctx, StokesCtx* s, int upto) {
    for (size_t l = 1; l < (size_t)upto && l < s->lev.size(); ++l) {
        Level& Lf = s->lev[l - 1];
        Level& Lc = s->lev[l];
        Lc.h = 2.0 * Lf.h;
        hipLaunchKernelGGL(k_coarsen, dim3(nblk((int64_t)(Lc.Nz + 1) * (Lc.Nx + 1))), dim3(NT), 0, ctx->stream, Lf.d(),
                           Lc.eta, Lc.rho, Lc.rhow, Lc.rhou, Lc.Nz, Lc.Nx);
        hipLaunchKernelGGL(k_corner_eta, dim3(nblk((int64_t)(Lc.Nz + 1) * (Lc.Nx + 1))), dim3(NT), 0, ctx->stream,
                           Lc.eta, Lc.etaco, Lc.Nz, Lc.Nx);
    }
}

// V-cycle depth for this solve.  The velocity block is a1 rho / dt (inertia) + eta k^2 (viscous).  With
// R = a1 rho h^2 / (dt eta) a W or U row's diagonal is (R + 7/3) eta / h^2 and the smoothest error mode's
// eigenvalue R eta / h^2, so a damped-Jacobi sweep scales that mode by 1 - omega R / (R + 7/3): where R is
// large the smoother alone removes the smooth error and coarser levels add work but not convergence.  On
// level l (spacing 2^l h) R_l = 4^l R_0 with R_0 = a1 min(rho) h^2 / (dt max(eta)) (the least inertial
// cell); the cycle stops at the first level below the finest where two sweeps reduce the smoothest
// mode at least 4x (omega R / (R + 7/3) >= 1/2, R >= 2.5 * 7/3, MG_INERTIA_STOP).  It keeps at least
// two levels: with the finest alone that level would be the coarsest and lose its pre- and post-sweeps
// (22-26 FGMRES iterations instead of 15-16 on the 32 x 40 and 64^2 production systems of the tests).
// Without inertia (dt = Inf, the benchmark mode) every level is kept.  mg_levels > 0 in the options caps
// the depth instead.
constexpr double MG_INERTIA_STOP = 6.0;
// whether choose_depth needs the fine level's largest viscosity and smallest density (k_pack5 writes its per-block
// partials into mapped host memory; choose_depth reduces them once the stream has passed it)
bool depth_launch(xc_ctx* ctx, StokesCtx* s, const Flags& F, const xc_fm_opts* o) {
    (void)ctx;
    (void)s;
    return !(o->mg_levels > 0 || F.bnchm || !(F.dt < 1e30) || !(F.a1 > 0));
}
// sets s->leff; launched: depth_launch enqueued the reduction; synced: the stream has been synchronized since
int choose_depth(xc_ctx* ctx, StokesCtx* s, const Flags& F, const xc_fm_opts* o, double h, bool launched,
                 bool synced) {
    const int nl = (int)s->lev.size();
    s->leff = nl;
    s->inertia = 0.0;
    if (o->mg_levels > 0) {
        s->leff = o->mg_levels < nl ? o->mg_levels : nl;
        return XC_OK;
    }
    if (!launched) return XC_OK;
    if (!synced) XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double emax = 0.0, rmin = 1e308;  // max and min are exact in any order
    for (unsigned b = 0; b < s->pack_blocks; ++b) {
        emax = fmax(emax, s->pk_host[2 * b]);
        rmin = fmin(rmin, s->pk_host[2 * b + 1]);
    }
    if (!(emax > 0) || !(rmin > 0)) return XC_OK;
    const double R0 = F.a1 * rmin * h * h / (F.dt * emax);
    s->inertia = R0;
    int L = nl < 2 ? nl : 2;  // level 0 keeps its pre- and post-smoothing, level 1 costs a quarter of it
    for (double R = 4.0 * R0; L < nl && R < MG_INERTIA_STOP; R *= 4.0) ++L;
    s->leff = L;
    return XC_OK;
}

// sub: pre = post sweeps on levels >= 1 (0: as level 0); f32: the fine level's smoother in single precision;
// om[l]: the damped-Jacobi weight on level l
constexpr int MG_MAXLEV = 16;
// cheb: the fine level's pre- and post-smoothing as Chebyshev passes on [ca, cb] (for D^-1 A) instead of damped
// Jacobi sweeps
struct MGOpts { int pre, post, coarse; double om[MG_MAXLEV]; int sub; bool f32; bool cheb; double ca, cb; };

// The weights of a Chebyshev smoothing pass of n <= CHEB_MAX sweeps for D^-1 A with spectrum in [a, b] (the
// Chebyshev iteration of Y. Saad, "Iterative Methods for Sparse Linear Systems", 2nd ed., Alg. 12.1, with the
// Jacobi preconditioner): sweep 0 is a Jacobi sweep of weight 1/theta; sweep k >= 1 adds al[k] (x_k - x_{k-1})
// to a Jacobi sweep of weight om[k] (theta = (a+b)/2, delta = (b-a)/2, sigma = theta/delta, rho_0 = 1/sigma,
// rho_k = 1/(2 sigma - rho_{k-1}), al[k] = rho_k rho_{k-1}, om[k] = 2 rho_k / delta)
constexpr int CHEB_MAX = 8;
struct Cheb { double om[CHEB_MAX], al[CHEB_MAX]; };
Cheb cheb_coefs(double a, double b, int n) {
    const double th = 0.5 * (b + a), de = 0.5 * (b - a), sg = th / de;
    Cheb c{};
    double rho = 1.0 / sg;
    c.om[0] = 1.0 / th;
    for (int k = 1; k < n && k < CHEB_MAX; ++k) {
        const double rn = 1.0 / (2.0 * sg - rho);
        c.al[k] = rn * rho;
        c.om[k] = 2.0 * rn / de;
        rho = rn;
    }
    return c;
}

// Jacobi weight of a level from its inertia ratio R = a1 rho h^2 / (dt eta).  The eigenvalues of D^-1 A of the
// velocity block lie in about [R / (R + 7/3), lam_max]: the smoothest mode, and a top that power iteration on the
// 256^2 production system put at 1.61 for R = 2.4 (its smallest eigenvalue 0.478 by inverse iteration, against
// 0.509 for R / (R + 7/3)).  Where R >= 1 the whole spectrum is smoothable (smallest eigenvalue >= 0.3) and the
// weight 0.95 (2 / (lam_min + lam_max) for that range, capped) reduces both ends: 9.0 against 9.4 FGMRES
// iterations per solve with 0.8 on the 512^2 step (1.0 and above lose).  Where R < 1 the smooth modes are left
// to the coarser levels and the weight stays the classical 0.8 (0.95 there loses convergence altogether on
// the inertial manufactured solution, R ~ 0.003-0.04).
double auto_omega(double R) { return R >= 1.0 ? 0.95 : 0.8; }

// solver controls with the defaults measured best on the 512^2 production step (zero fields: default); after
// choose_depth, which measures the inertia ratio the automatic Jacobi weights come from (R_l = 4^l R_0)
MGOpts mg_opts(const xc_fm_opts* o, const StokesCtx* s) {
    MGOpts m;
    m.pre = o->mg_pre > 0 ? o->mg_pre : 3;
    m.post = o->mg_post > 0 ? o->mg_post : 3;
    m.coarse = o->mg_coarse > 0 ? o->mg_coarse : 0;  // 0: coarse_sweeps decides per solve
    double R = s->inertia;
    for (int l = 0; l < MG_MAXLEV; ++l, R *= 4.0) m.om[l] = o->mg_omega > 0 ? o->mg_omega : auto_omega(R);
    m.sub = o->mg_sub > 0 ? o->mg_sub : (o->mg_sub == 0 ? 1 : 0);
    m.f32 = o->mg_fp64 == 0;
    // The fine level's Chebyshev interval, from its inertia ratio R (the smallest, choose_depth): the bottom at
    // 1.2 R / (R + 7/3), above the smoothest mode's R / (R + 7/3) (the coarse level takes the modes below), the
    // top at 1 + 1.1 x 3 / (R + 7/3), the uniform-viscosity Gershgorin bound plus 10 %.  On the 512^2 step
    // (R ~ 2.5) 3 + 3 sweeps on it take 4.26 FGMRES iterations per solve against 4.63 with damped Jacobi (first
    // sweep 6.0 against 7.0); bottoms of 0.5, 0.6, 0.9, 1.0, 1.35 and 1.5 R / (R + 7/3) gave 5.33, 5.12, 4.41,
    // 4.31, 4.37 and 4.41, tops of 0.9 and 1.3 x the bound 4.48 and 4.25, and 2 + 2 sweeps 5.2 (Jacobi 5.4).
    const double R0 = s->inertia;
    m.cheb = o->mg_jacobi == 0 && R0 >= 1.0 && m.pre <= CHEB_MAX && m.post <= CHEB_MAX;
    m.ca = 1.2 * R0 / (R0 + 7.0 / 3.0);
    m.cb = 1.0 + 1.1 * 3.0 / (R0 + 7.0 / 3.0);
    return m;
}

// sweeps on the coarsest level of the cycle: as set, else 20 on a 4 x 4 bottom level and 2 on a level where
// the cycle stopped because inertia dominates (the two sweeps choose_depth's criterion is stated for; on the
// 512^2 step 1, 2 and 4 sweeps there gave 10.4, 10.3 and 9.6 FGMRES iterations per solve)
int coarse_sweeps(const StokesCtx* s, const MGOpts& mo) {
    if (mo.coarse > 0) return mo.coarse;
    return s->leff < (int)s->lev.size() ? 2 : 20;
}

void vcycle(xc_ctx* ctx, StokesCtx* s, const Flags& F, size_t l, const MGOpts& mo, bool first_done = false);

// level l's residual rf restricted to level l + 1 and the cycle from there; the restriction is fused into
// level l + 1's first sweep (from zero) unless that level starts the one-workgroup tail
template <typename TF>
void descend(xc_ctx* ctx, StokesCtx* s, const Flags& F, size_t l, const MGOpts& mo, const TF* rf) {
    Level& C = s->lev[l + 1];
    if ((int)(l + 1) == s->tail_l && s->leff > s->tail_l) {
        hipLaunchKernelGGL(k_restrict<TF>, dim3(nblk(C.nv)), dim3(NT), 0, ctx->stream, s->lev[l].d(), C.d(), rf, C.b);
        vcycle(ctx, s, F, l + 1, mo);
        return;
    }
    hipLaunchKernelGGL((k_smooth_cells<4, LevD, double, TF, double>), dim3(smooth_cells_blocks(C.Nz, C.Nx)), dim3(NT), 0,
                       ctx->stream, C.d(), F, C.x, rf, C.t, mo.om[l + 1], 0, nullptr, C.b);
    std::swap(C.x, C.t);
    vcycle(ctx, s, F, l + 1, mo, true);
}

// one V-cycle on level l: solve A x = b approximately from x = 0
void vcycle(xc_ctx* ctx, StokesCtx* s, const Flags& F, size_t l, const MGOpts& mo, bool first_done) {
    Level& L = s->lev[l];
    if ((int)l == s->tail_l && s->leff > s->tail_l) {
        TailArgs ta = s->tail;
        ta.nlev = s->leff - s->tail_l;
        if (ta.nbig > ta.nlev) ta.nbig = ta.nlev;
        for (int q = 0; q < ta.nlev; ++q) ta.d[q] = s->lev[l + q].d();  // h is set per solve
        ta.bg = L.b;
        ta.xg = L.x;
        ta.pre = mo.sub > 0 ? mo.sub : mo.pre;
        ta.post = mo.sub > 0 ? mo.sub : mo.post;
        ta.coarse = coarse_sweeps(s, mo); ta.omega = mo.om[l];
        ta.F = F;
        hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(TAIL_NT), s->tail_lds, ctx->stream, ta);
        return;
    }
    const LevD d = L.d();
    const unsigned nbc = smooth_cells_blocks(L.Nz, L.Nx);
    const int fine = l == 0;
    const bool coarsest = (int)l + 1 == s->leff;
    const int pre = coarsest ? coarse_sweeps(s, mo) : (l > 0 && mo.sub > 0 ? mo.sub : mo.pre);
    const int post = l > 0 && mo.sub > 0 ? mo.sub : mo.post;
    // first sweep from zero (already done by the caller when first_done)
    if (!first_done) {
        hipLaunchKernelGGL(SMOOTH_D(1), dim3(nbc), dim3(NT), 0, ctx->stream, d, F, L.x, L.b, L.t, mo.om[l], fine, nullptr, nullptr);
        std::swap(L.x, L.t);
    }
    for (int it = 1; it < pre; ++it) {
        hipLaunchKernelGGL(SMOOTH_D(0), dim3(nbc), dim3(NT), 0, ctx->stream, d, F, L.x, L.b, L.t, mo.om[l], fine, nullptr, nullptr);
        std::swap(L.x, L.t);
    }
    if (coarsest) return;
    hipLaunchKernelGGL(SMOOTH_D(2), dim3(nbc), dim3(NT), 0, ctx->stream, d, F, L.x, L.b, L.r, 0.0, fine, nullptr, nullptr);
    descend(ctx, s, F, l, mo, (const double*)L.r);
    hipLaunchKernelGGL(k_prolong_cells<double>, dim3(nbc), dim3(NT), 0, ctx->stream, L.Nz, L.Nx, F, s->lev[l + 1].x, L.x);
    for (int it = 0; it < post; ++it) {
        hipLaunchKernelGGL(SMOOTH_D(0), dim3(nbc), dim3(NT), 0, ctx->stream, d, F, L.x, L.b, L.t, mo.om[l], fine, nullptr, nullptr);
        std::swap(L.x, L.t);
    }
}

// The V-cycle with the fine level in single precision: the first sweep from zero with the right-hand side
// b = r_v - GG z_p formed from v, the cycle (levels >= 1 in double precision), and the last fine-level
// sweep written to z's velocity part in double precision.  The preconditioner's precision does not bound
// the solution's: FGMRES forms its residual and basis in double precision from LL, and the smoother only
// needs to reduce the error by a constant factor per sweep.
void vcycle_f32(xc_ctx* ctx, StokesCtx* s, const Flags& F, const MGOpts& mo, const double* v, const double* zP,
                double* z) {
    Level& L = s->lev[0];
    const LevD d = L.d();
    const LevF f = s->levf();
    const unsigned nbc = smooth_cells_blocks(L.Nz, L.Nx);
    float *x = s->fx, *t = s->ft;
    const bool coarsest = s->leff <= 1;
    const int pre = coarsest ? coarse_sweeps(s, mo) : mo.pre, post = mo.post;
    const bool cheb = mo.cheb && !coarsest;
    const Cheb cp = cheb_coefs(mo.ca, mo.cb, pre), cq = cheb_coefs(mo.ca, mo.cb, post);
    hipLaunchKernelGGL(SMOOTH_F(3, double, float), dim3(nbc), dim3(NT), 0, ctx->stream, f, F, x, v, x,
                       cheb ? cp.om[0] : mo.om[0], 1, zP, s->fb);
    bool in_z = false;
    // sweep k of a pass: Jacobi (weight mo.om[0]), or Chebyshev sweep k with the previous iterate in t (k = 1 of
    // the pre-smoothing: zero, the pass having started from zero)
    auto sweep = [&](bool to_z, const Cheb& cc, int k, bool from_zero) {
        const double om = cheb ? cc.om[k] : mo.om[0], al = cheb ? cc.al[k] : 0.0;
        const float* prev = (cheb && k > 0 && !(from_zero && k == 1)) ? t : nullptr;
        if (to_z) {
            hipLaunchKernelGGL(SMOOTH_F(0, float, double), dim3(nbc), dim3(NT), 0, ctx->stream, f, F, x, s->fb, z,
                               om, 1, nullptr, nullptr, prev, al);
            in_z = true;
        } else {
            hipLaunchKernelGGL(SMOOTH_F(0, float, float), dim3(nbc), dim3(NT), 0, ctx->stream, f, F, x, s->fb, t,
                               om, 1, nullptr, nullptr, prev, al);
            std::swap(x, t);
        }
    };
    for (int it = 1; it < pre; ++it) sweep(coarsest && it == pre - 1, cp, it, true);
    if (!coarsest) {
        hipLaunchKernelGGL(SMOOTH_F(2, float, float), dim3(nbc), dim3(NT), 0, ctx->stream, f, F, x, s->fb, s->fr, 0.0,
                           1, nullptr, nullptr);
        descend(ctx, s, F, 0, mo, (const float*)s->fr);
        hipLaunchKernelGGL(k_prolong_cells<float>, dim3(nbc), dim3(NT), 0, ctx->stream, L.Nz, L.Nx, F, s->lev[1].x, x);
        for (int it = 0; it < post; ++it) sweep(it == post - 1, cq, it, false);
    }
    if (!in_z) hipLaunchKernelGGL(k_f2d, dim3(nblk(L.nv)), dim3(NT), 0, ctx->stream, x, z, L.nv);
}

// z = M^-1 v  (block-triangular preconditioner)
void precond(xc_ctx* ctx, StokesCtx* s, const Flags& F, const double* v, double* z, const MGOpts& mo) {
    Level& L = s->lev[0];
    const LevD d = L.d();
    const int64_t nv = L.nv, nP = (int64_t)(L.Nz + 2) * (L.Nx + 2);
    const double* vP = v + nv;
    double* zP = z + nv;
    if (s->pratio > 0) {  // inertial Schur part active (see pois_table)
        const int Nz = L.Nz, Nx = L.Nx, K = s->pK;
        // enough workgroups to spread over the CUs: one complex line (two rows) each at 512 x 512
        int nl = Nz / 512;
        const int nlmax = MAXLINE / Nx < 8 ? MAXLINE / Nx : 8;
        nl = nl < 1 ? 1 : (nl > nlmax ? nlmax : nl);
        Plan pl{};
        pl.n = Nx;
        pl.nst = (int)s->prad.size();
        for (int q = 0; q < pl.nst; ++q) pl.rad[q] = s->prad[q];
        pl.tw = s->ptw;
        plan_mark_pow2(pl);
        const unsigned nbx = (unsigned)((Nz + 2 * nl - 1) / (2 * nl));
        const size_t lds = 2 * (size_t)nl * Nx * sizeof(double2);
        PoisXArgs xa{vP, s->pS, pl, Nz, Nx, K, nl};
        const int lgx = p2_fast_lg(pl, nl);
        switch (lgx) {
            case 8: hipLaunchKernelGGL(k_pois_xfwd<8>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, xa); break;
            case 9: hipLaunchKernelGGL(k_pois_xfwd<9>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, xa); break;
            case 10: hipLaunchKernelGGL(k_pois_xfwd<10>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, xa); break;
            default: hipLaunchKernelGGL(k_pois_xfwd<0>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, xa); break;
        }
        Plan pz{};
        pz.n = Nz;
        pz.nst = (int)s->pradz.size();
        for (int q = 0; q < pz.nst; ++q) pz.rad[q] = s->pradz[q];
        pz.tw = s->ptwz;
        plan_mark_pow2(pz);
        // one line per workgroup up to K = 512 lines; at most ZD_PER (k, Nz-k) pairs a lane
        int nlz = K / 512;
        int nlzmax = MAXLINE / Nz < PZ_NLMAX ? MAXLINE / Nz : PZ_NLMAX;
        while (nlzmax > 1 && nlzmax * (Nz / 2 + 1) > ZD_PER * FFT_NT) --nlzmax;
        nlz = nlz < 1 ? 1 : (nlz > nlzmax ? nlzmax : nlz);
        PoisZArgs za{s->pS, pz, s->ptwz + Nz, s->plamx, s->plamz, Nz, K, nlz, F.open_cnv ? s->pg : nullptr,
                     F.dt / (F.a1 * L.h * L.h)};
        {
            const dim3 gz((unsigned)((K + nlz - 1) / nlz));
            const size_t ldz = 2 * (size_t)nlz * Nz * sizeof(double2);
            switch (p2_fast_lg(pz, nlz)) {
                case 8: hipLaunchKernelGGL(k_pois_zfft<8>, gz, dim3(FFT_NT), ldz, ctx->stream, za); break;
                case 9: hipLaunchKernelGGL(k_pois_zfft<9>, gz, dim3(FFT_NT), ldz, ctx->stream, za); break;
                case 10: hipLaunchKernelGGL(k_pois_zfft<10>, gz, dim3(FFT_NT), ldz, ctx->stream, za); break;
                default: hipLaunchKernelGGL(k_pois_zfft<0>, gz, dim3(FFT_NT), ldz, ctx->stream, za); break;
            }
        }
        PoisInvArgs ia{s->pS, vP, zP, pl, d, F, s->pscale, Nz, Nx, K, nl};
        switch (lgx) {  // with the ghost ring
            case 8: hipLaunchKernelGGL(k_pois_xinv<8>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, ia); break;
            case 9: hipLaunchKernelGGL(k_pois_xinv<9>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, ia); break;
            case 10: hipLaunchKernelGGL(k_pois_xinv<10>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, ia); break;
            default: hipLaunchKernelGGL(k_pois_xinv<0>, dim3(nbx), dim3(FFT_NT), lds, ctx->stream, ia); break;
        }
    } else {
        hipLaunchKernelGGL(k_schur_int, dim3(nblk(nP)), dim3(NT), 0, ctx->stream, d, F, vP, zP);
        hipLaunchKernelGGL(k_schur_ghost, dim3(nblk(nP)), dim3(NT), 0, ctx->stream, d, F, vP, zP);
    }
    // velocity: b = r_v - GG z_p and the first sweep in one kernel, then the V-cycle.  With an even
    // number of level-0 sweeps the iterates alternate between z and the level's scratch and end in z.
    if (mo.f32) {
        vcycle_f32(ctx, s, F, mo, v, zP, z);
        return;
    }
    double* own_x = L.x;
    const bool direct = s->leff > 1 && (mo.pre + mo.post) % 2 == 0;
    if (direct) L.x = z;
    hipLaunchKernelGGL(SMOOTH_D(3), dim3(smooth_cells_blocks(L.Nz, L.Nx)), dim3(NT), 0, ctx->stream, d, F, L.x, v, L.t,
                       mo.om[0], 1, zP, L.b);
    std::swap(L.x, L.t);
    vcycle(ctx, s, F, 0, mo, true);
    if (!direct) hipLaunchKernelGGL(k_copyv, dim3(nblk(nv)), dim3(NT), 0, ctx->stream, L.x, z, nv);
    else L.x = own_x;  // L.x is z here; hand the level its own buffer back
}

// dots[0..K] = (V[q] . w) for q < K and w.w; synchronous, results in s->host
int multidot(xc_ctx* ctx, StokesCtx* s, const double* V, int K, const double* w, int64_t n) {
    const int nbk = DOT_BLOCKS;
    hipLaunchKernelGGL(k_multidot, dim3(nbk), dim3(NT), 0, ctx->stream, V, n, K, w, s->part);
    // the sums are written straight into the mapped pinned host buffer (no copy launch)
    hipLaunchKernelGGL(k_dot_final, dim3((K + 4) / 4), dim3(NT), 0, ctx->stream, s->part, nbk, K + 1, s->host_dev);
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return XC_OK;
}


// x0: also gather the state's (W, U, P) into x0 (the solve's starting guess) in the right-hand side's pass
int build_rhs(xc_ctx* ctx, StokesCtx* s, xc_fluid* fl, const xc_phase* ph, const xc_grid* g, const xc_par* par,
              const xc_fm_opts* o, const xc_bnchm* bn, const Flags& F, double* RR, double* x0 = nullptr) {
    const int Nz = (int)g->Nz, Nx = (int)g->Nx;
    // MFS update + mean (fluidmech.jl:115-127)
    RhsArgs ra{};
    ra.rho = fl->rho; ra.rhoo = fl->rhoo; ra.rhooo = fl->rhooo; ra.drhodt = fl->drhodt; ra.drhodto = fl->drhodto;
    ra.drhodtoo = fl->drhodtoo; ra.MFS = fl->MFS;
    if (ph) ra.advn_rho = ph->advn_rho;
    ra.a1 = o->a1; ra.a2 = o->a2; ra.a3 = o->a3; ra.b1 = o->b1; ra.b2 = o->b2; ra.b3 = o->b3; ra.dt = o->dt;
    ra.alpha = par->alpha;
    ra.do_mfs = bn == nullptr;
    const unsigned nbc = nblk((int64_t)Nz * Nx);
    hipLaunchKernelGGL(k_mfs, dim3(nbc), dim3(NT), 0, ctx->stream, ra, Nz, Nx, s->part);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(NT), 0, ctx->stream, s->part, (int)nbc, 1.0 / ((double)Nz * Nx), s->dots + 100);
    // momentum advection (fluidmech.jl:236-240, 310-318)
    double* fz = s->adv_buf;
    double* uz = fz + (size_t)(Nz + 2) * (Nx + 3);
    double* wz = uz + (size_t)(Nz + 2) * (Nx + 3);
    double* fx = wz + (size_t)(Nz + 2) * (Nx + 3);
    double* ux = fx + (size_t)(Nz + 2) * (Nx + 3);
    double* wx = ux + (size_t)(Nz + 2) * (Nx + 3);
    double* az = wx + (size_t)(Nz + 2) * (Nx + 3);
    double* ax = az + (size_t)(Nz + 2) * (Nx + 3);
    hipLaunchKernelGGL(k_mom_prep, dim3(nblk((int64_t)(Nz + 2) * (Nx + 3))), dim3(NT), 0, ctx->stream, fl->W, fl->U,
                       fl->rhow, fl->rhou, fz, uz, wz, fx, ux, wx, Nz, Nx);
    const unsigned saved = ctx->flags;
    ctx->flags &= ~XC_CTX_SYNC;
    const xc_mat none{nullptr, 0, 0, 0};
    const xc_adv_member mm[2] = {
        {xc_mat{az, Nz - 1, Nx, Nz - 1}, none, none, xc_mat{fz, Nz - 1, Nx, Nz - 1}, xc_mat{uz, Nz - 1, Nx + 1, Nz - 1},
         xc_mat{wz, Nz, Nx, Nz}, g->h},
        {xc_mat{ax, Nz, Nx + 1, Nz}, none, none, xc_mat{fx, Nz, Nx + 1, Nz}, xc_mat{ux, Nz, Nx + 2, Nz},
         xc_mat{wx, Nz + 1, Nx + 1, Nz + 1}, g->h}};
    int st = xc_advect_centered_pair(ctx, o->scheme, mm, o->xbc, o->zbc, "fluidmech!");
    ctx->flags = saved;
    if (st) return st;
    RRArgs a{};
    a.Drho = fl->Drho; a.rhoWo = fl->rhoWo; a.rhoWoo = fl->rhoWoo; a.rhoUo = fl->rhoUo; a.rhoUoo = fl->rhoUoo;
    a.MFS = fl->MFS; a.rhow = fl->rhow;
    a.advz = az; a.advx = ax; a.mfsmean = s->dots + 100;
    if (bn) { a.srcW = bn->src_W; a.srcU = bn->src_U; a.srcP = bn->src_P; }
    a.a2 = o->a2; a.a3 = o->a3; a.dt = o->dt; a.g0 = par->g0; a.D = g->D; a.h = g->h;
    a.oc = o->open_cnv ? 1.0 : 0.0;
    a.bnchm = bn != nullptr;
    if (x0) { a.gW = fl->W; a.gU = fl->U; a.gP = fl->P; a.gx = x0; }
    const int64_t n = (int64_t)(Nz + 1) * (Nx + 2) + (int64_t)(Nz + 2) * (Nx + 1) + (int64_t)(Nz + 2) * (Nx + 2);
    hipLaunchKernelGGL(k_rr, dim3(nblk(n)), dim3(NT), 0, ctx->stream, a, Nz, Nx, RR);
    if (bn) {
        const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1);
        const int64_t kp = nW + nU + F.pr + (int64_t)F.pc * (Nz + 2);
        const int64_t q = F.nu0 - nW;
        hipLaunchKernelGGL(k_rr_pins, dim3(1), dim3(1), 0, ctx->stream, RR, kp, bn->P_exact, F.pr, F.pc, F.nu0,
                           bn->U_exact, ((int)q % (Nz + 2)), ((int)q / (Nz + 2)));
    }
    return XC_OK;
}

// pshift: the pressure level a converged solve with the pin fix sets (x_P += pshift), left to the caller
// (xc_fluidmech adds it as it scatters the solution; has_pshift: there is one)
struct SolveOut { int iters; bool conv; double rel, rel0; int learned_steps; bool dropped; bool has_pshift; double pshift; };
// cost of a learned preconditioner step in plain steps (restriction, network, prolongation besides the V-cycle)
constexpr double LC_COST = 1.25;
// a learned correction dropped at its first step in this many solves in a row is left out of the next LC_SKIP
// solves, twice as many each time the retry after such a period is dropped again (up to LC_SKIP << LC_BACKOFF)
constexpr int LC_STREAK = 3, LC_SKIP = 16, LC_BACKOFF = 6;

struct LcCfg {
    bool on;
    int mode;  // 1: z = e + M (v - LL e) with e = P N(R v); 2: z = M v + P N(R v)
    int level, f, Hc, Wc;
    float is[4], os[3];
    double eta_ref;
};

// learned correction settings for this solve (off unless requested, loaded and the grid divides)
int lc_config(xc_ctx* ctx, StokesCtx* s, const xc_fm_opts* o, const xc_grid* g, const xc_par* par, LcCfg* lc) {
    *lc = LcCfg{};
    if (!o->learned || !xc_unet_has_params(ctx)) return XC_OK;
    const int lev = o->lc_level > 0 ? o->lc_level : 2;
    const int f = 1 << lev;
    if (g->Nz % f || g->Nx % f) return xc_fail(ctx, XC_EINVAL_SHAPE, "fluidmech!: grid not divisible by 2^%d", lev);
    const int Hc = (int)g->Nz / f, Wc = (int)g->Nx / f;
    if (Hc % 8 || Wc % 8 || Hc < 8 || Wc < 8 || (int64_t)Hc * Wc > s->lc_cap)
        return xc_fail(ctx, XC_EINVAL_SHAPE, "fluidmech!: learned correction grid %d x %d (multiples of 8 needed)", Hc, Wc);
    lc->on = true;
    lc->mode = o->learned == 2 ? 2 : 1;
    lc->level = lev;
    lc->f = f;
    lc->Hc = Hc;
    lc->Wc = Wc;
    xc_unet_io_scales(ctx, lc->is, lc->os);
    lc->eta_ref = par->etam0;
    return XC_OK;
}

void lc_restrict(xc_ctx* ctx, StokesCtx* s, const LevD& d, const LcCfg& lc, const double* v, float* out, float* sigma,
                 int compute) {
    LcArgs a{};
    a.L = d;
    a.f = lc.f;
    a.Hc = lc.Hc;
    a.Wc = lc.Wc;
    a.v = v;
    a.out = out;
    a.part = s->lc_part;
    a.eta_ref = lc.eta_ref;
    for (int k = 0; k < 4; ++k) a.s[k] = lc.is[k];
    const int64_t nc = (int64_t)lc.Hc * lc.Wc;
    const unsigned nbk = nblk(nc);
    hipLaunchKernelGGL(k_lc_restrict, dim3(nbk), dim3(NT), 0, ctx->stream, a);
    hipLaunchKernelGGL(k_lc_norm, dim3(1), dim3(1024), 0, ctx->stream, out, s->lc_part, (int)nbk, nc, sigma, compute);
}

// z = M^-1 v.  Plain: the block preconditioner.  Learned, mode 1: the network's coarse estimate
// e = P N(R v) of LL^-1 v first, then the block preconditioner on what it leaves,
// z = e + M_plain (v - LL e); the multigrid smoothing removes the block-wise roughness
// of the prolonged estimate.  Mode 2: the network estimates what the block preconditioner misses,
// z = M_plain v + P N(R v) (trained on R (LL^-1 v - M_plain v)); no extra operator application.
int precond_full(xc_ctx* ctx, StokesCtx* s, const Flags& F, const MGOpts& mo, int64_t n, const LcCfg& lc,
                 const double* v, double* z) {
    if (!lc.on) {
        precond(ctx, s, F, v, z, mo);
        return XC_OK;
    }
    const LevD dd = s->lev[0].d();
    const unsigned nb = nblk(n);
    lc_restrict(ctx, s, dd, lc, v, s->lc_in, s->lc_sigma, 1);
    const float* e;
    int st;
    if ((st = xc_unet_infer(ctx, s->lc_in, 1, lc.Hc, lc.Wc, &e))) return st;
    if (lc.mode == 2) precond(ctx, s, F, v, z, mo);
    else hipLaunchKernelGGL(k_zero, dim3(nb), dim3(NT), 0, ctx->stream, z, n);
    LcProArgs pa{};
    pa.L = dd;
    pa.f = lc.f;
    pa.Hc = lc.Hc;
    pa.Wc = lc.Wc;
    pa.e = e;
    pa.sigma = s->lc_sigma;
    for (int k = 0; k < 3; ++k) pa.os[k] = lc.os[k];
    pa.z = z;
    pa.n = n;
    hipLaunchKernelGGL(k_lc_prolong, dim3(nb), dim3(NT), 0, ctx->stream, pa);
    if (lc.mode == 2) return XC_OK;
    hipLaunchKernelGGL(k_ll_cells<true>, dim3(ll_cells_blocks(dd.Nz, dd.Nx)), dim3(NT), 0, ctx->stream, dd, F, z, v,
                       s->tmp);
    precond(ctx, s, F, s->tmp, s->rs, mo);  // rs is free inside an Arnoldi cycle
    hipLaunchKernelGGL(k_axpy1, dim3(nb), dim3(NT), 0, ctx->stream, s->rs, z, n);
    return XC_OK;
}

// right-preconditioned FGMRES(m) on LL x = b from the given x; stops at ||b - LL x|| <= target
// ---- horizontal-mean correction of a cycle's starting guess
// The Krylov solve converges slowly in one subspace: the fields that do not vary in x (k_x = 0 modes --
// the horizontally averaged mass flux through each row of faces and the pressure profile that carries it),
// which the preconditioner's Schur approximation treats least accurately; FGMRES then stalls for ~8
// iterations at a residual of ~1e-7 of ||RR|| before it resolves them.  Before every cycle the starting guess
// is corrected on that subspace exactly (a Galerkin, two-level step): with R the map from profiles (one value
// per row of W, U and P) to x-uniform fields (every column, ghost columns included, equal to the profile),
// x += R L0^-1 R^T (b - LL x) with L0 = R^T LL R, so that the corrected residual has no x-uniform component.
// L0 couples rows z - 1 .. z + 1 only; it is probed with 9 x-uniform comb vectors (one field, rows of one
// residue mod 3), each response row decoded to the comb row within one of it, and factorized on the host
// (LU with partial pivoting, band 5 with rows ordered W, U, P per z).  Profiles are [W rows 0..Nz][U rows
// 0..Nz+1][P rows 0..Nz+1].
__device__ __forceinline__ void xm_locate(int64_t i, int Nz, int Nx, int& f, int& row) {
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1);
    if (i < nW) { f = 0; row = (int)(i % (Nz + 1)); }
    else if (i < nW + nU) { f = 1; row = (int)((i - nW) % (Nz + 2)); }
    else { f = 2; row = (int)((i - nW - nU) % (Nz + 2)); }
}
// comb_f < 0: x += R p; otherwise x = R c with c the comb (field comb_f, rows = comb_k mod 3)
__global__ __launch_bounds__(NT) void k_xm_expand(const double* __restrict__ p, double* __restrict__ x, int64_t n,
                                                  int Nz, int Nx, int comb_f, int comb_k) {
    const int64_t i = gidx();
    if (i >= n) return;
    int f, row;
    xm_locate(i, Nz, Nx, f, row);
    if (comb_f >= 0) x[i] = (f == comb_f && row % 3 == comb_k) ? 1.0 : 0.0;
    else x[i] += p[row + (f == 0 ? 0 : f == 1 ? Nz + 1 : 2 * Nz + 3)];
}
// part[p][c][q] = the sum of profile row q's entries in columns c XM_CH .. c XM_CH + XM_CH - 1 of its field, for
// vector p = blockIdx.z (v + p vstride)
// sq (optional): the launch has one more z plane, whose first workgroup sums the nsq partials sq into sq_out
// (k_sum_final1's work, in the same launch), and bq (optional) into bq_out
__global__ __launch_bounds__(NT) void k_xm_part(const double* __restrict__ v, int64_t vstride,
                                                double* __restrict__ part, int Nz, int Nx, int n0,
                                                const double* __restrict__ sq = nullptr, int nsq = 0,
                                                double* sq_out = nullptr, const double* __restrict__ bq = nullptr,
                                                double* bq_out = nullptr) {
    if (sq && blockIdx.z == gridDim.z - 1) {
        if (blockIdx.x == 0 && blockIdx.y == 0) {
            block_sum1(sq, nsq, sq_out);
            if (bq) {
                __syncthreads();  // block_sum1's LDS is read by thread 0 before it is reused
                block_sum1(bq, nsq, bq_out);
            }
        }
        return;
    }
    const int q = blockIdx.x * NT + threadIdx.x, c = blockIdx.y;
    if (q >= n0) return;
    v += (int64_t)blockIdx.z * vstride;
    part += (int64_t)blockIdx.z * gridDim.y * n0;
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1);
    int rows, cols, row;
    int64_t base;
    if (q <= Nz) { rows = Nz + 1; cols = Nx + 2; row = q; base = 0; }
    else if (q <= 2 * Nz + 2) { rows = Nz + 2; cols = Nx + 1; row = q - (Nz + 1); base = nW; }
    else { rows = Nz + 2; cols = Nx + 2; row = q - (2 * Nz + 3); base = nW + nU; }
    const int c0 = c * XM_CH, c1 = min(cols, c0 + XM_CH);
    const double* col = v + base + row;
    double sum = 0.0;
    int cc = c0;
    for (; cc + 8 <= c1; cc += 8) {  // eight loads in flight, summed in column order
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = col[(int64_t)(cc + u) * rows];
#pragma unroll
        for (int u = 0; u < 8; ++u) sum += t[u];
    }
    for (; cc < c1; ++cc) sum += col[(int64_t)cc * rows];
    part[(int64_t)c * n0 + q] = sum;
}
// out[p][q] = sum over the nch partials of vector p (fixed order)
__global__ __launch_bounds__(NT) void k_xm_final(const double* __restrict__ part, int nch, int n0, double* out) {
    const int q = blockIdx.x * NT + threadIdx.x;
    if (q >= n0) return;
    part += (int64_t)blockIdx.y * nch * n0;
    double sum = 0.0;
    for (int c = 0; c < nch; ++c) sum += part[(int64_t)c * n0 + q];
    out[(int64_t)blockIdx.y * n0 + q] = sum;
}
// out[p] (n0 each, device or mapped host) = R^T (v + p vstride) for p < nv
// sq, nsq, sq_out (and bq, bq_out): also the sums of nsq partials, in the same launch (k_xm_part)
void xm_sum(xc_ctx* ctx, StokesCtx* s, const double* v, double* out, int nv = 1, int64_t vstride = 0,
            const double* sq = nullptr, int nsq = 0, double* sq_out = nullptr, const double* bq = nullptr,
            double* bq_out = nullptr) {
    const int n0 = s->xm_n0;
    hipLaunchKernelGGL(k_xm_part, dim3((n0 + NT - 1) / NT, s->xm_nch, nv + (sq ? 1 : 0)), dim3(NT), 0, ctx->stream,
                       v, vstride, s->xm_part, s->Nz, s->Nx, n0, sq, nsq, sq_out, bq, bq_out);
    hipLaunchKernelGGL(k_xm_final, dim3((n0 + NT - 1) / NT, nv), dim3(NT), 0, ctx->stream, s->xm_part, s->xm_nch,
                       n0, out);
}

// LU with partial pivoting of a band matrix (kl sub-, ku super-diagonals); row i of a holds columns
// i - kl .. i + kl + ku (the room the row interchanges fill), W = 2 kl + ku + 1 entries a row
constexpr int XM_KL = 5, XM_KU = 5, XM_W = 2 * XM_KL + XM_KU + 1;
inline double& xm_at(std::vector<double>& a, int i, int j) { return a[(size_t)i * XM_W + (j - i + XM_KL)]; }
bool xm_band_lu(int n, std::vector<double>& a, std::vector<int>& piv) {
    piv.assign(n, 0);
    for (int k = 0; k < n; ++k) {
        const int last = std::min(n - 1, k + XM_KL), right = std::min(n - 1, k + XM_KL + XM_KU);
        int p = k;
        for (int i = k + 1; i <= last; ++i)
            if (fabs(xm_at(a, i, k)) > fabs(xm_at(a, p, k))) p = i;
        const double d = xm_at(a, p, k);
        if (!(d != 0.0) || !std::isfinite(d)) return false;
        piv[k] = p;
        if (p != k)
            for (int j = k; j <= right; ++j) std::swap(xm_at(a, k, j), xm_at(a, p, j));
        for (int i = k + 1; i <= last; ++i) {
            const double l = xm_at(a, i, k) / xm_at(a, k, k);
            xm_at(a, i, k) = l;
            if (l != 0.0)
                for (int j = k + 1; j <= right; ++j) xm_at(a, i, j) -= l * xm_at(a, k, j);
        }
    }
    return true;
}
void xm_band_solve(int n, std::vector<double>& a, const std::vector<int>& piv, double* b) {
    for (int k = 0; k < n; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int i = k + 1; i <= std::min(n - 1, k + XM_KL); ++i) b[i] -= xm_at(a, i, k) * b[k];
    }
    for (int k = n - 1; k >= 0; --k) {
        double v = b[k];
        for (int j = k + 1; j <= std::min(n - 1, k + XM_KL + XM_KU); ++j) v -= xm_at(a, k, j) * b[j];
        b[k] = v / xm_at(a, k, k);
    }
}
// profile index q of (field f, row z) and its position in the banded ordering (3 z + f; W has no row Nz + 1,
// whose position holds an identity row)
inline int xm_prof(int f, int z, int Nz) { return z + (f == 0 ? 0 : f == 1 ? Nz + 1 : 2 * Nz + 3); }

// Probe and factorize L0 for the current coefficients and flags (uses s->ws and s->tmp; synchronizes).
int xm_build(xc_ctx* ctx, StokesCtx* s, const Flags& F) {
    const int Nz = s->Nz, Nx = s->Nx, n0 = s->xm_n0;
    const LevD dd = s->lev[0].d();
    const int64_t n = s->lev[0].nv + (int64_t)(Nz + 2) * (Nx + 2);
    s->xm_ok = false;
    const bool batch = s->m >= XM_PROBES;  // the responses side by side in the (idle) basis Z, summed together
    for (int pf = 0; pf < 2; ++pf)
        for (int k = 0; k < 3; ++k) {
            const int f = 2 * pf, pr = 3 * pf + k;  // fields W and P
            double* resp = batch ? s->Z + (int64_t)pr * n : s->tmp;
            hipLaunchKernelGGL(k_xm_expand, dim3(nblk(n)), dim3(NT), 0, ctx->stream, (const double*)nullptr, s->ws, n,
                               Nz, Nx, f, k);
            hipLaunchKernelGGL(k_ll_cells<false>, dim3(ll_cells_blocks(dd.Nz, dd.Nx)), dim3(NT), 0, ctx->stream, dd,
                               F, s->ws, s->ws, resp);
            if (!batch) xm_sum(ctx, s, resp, s->xm_hdev + (size_t)(2 + pr) * n0);
        }
    if (batch) xm_sum(ctx, s, s->Z, s->xm_hdev + 2 * (size_t)n0, XM_PROBES, n);
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int nb = 3 * (Nz + 2);
    const int rows[3] = {Nz + 1, Nz + 2, Nz + 2};
    std::vector<double>& a = s->xm_lu;
    a.assign((size_t)nb * XM_W, 0.0);
    xm_at(a, 3 * (Nz + 1), 3 * (Nz + 1)) = 1.0;  // W has no row Nz + 1
    for (int z = 0; z < Nz + 2; ++z) xm_at(a, 3 * z + 1, 3 * z + 1) = 1.0;  // U is not corrected
    for (int pf = 0; pf < 2; ++pf)
        for (int k = 0; k < 3; ++k) {
            const int f = 2 * pf;
            const double* resp = s->xm_host + (size_t)(2 + 3 * pf + k) * n0;
            for (int fr = 0; fr < 3; fr += 2)
                for (int zr = 0; zr < rows[fr]; ++zr) {
                    const double v = resp[xm_prof(fr, zr, Nz)];
                    if (v == 0.0) continue;
                    const int zt = zr - 1 + (((k - (zr - 1)) % 3) + 3) % 3;  // the comb row within one of zr
                    const int i = 3 * zr + fr, j = 3 * zt + f;
                    if (!std::isfinite(v) || zt < 0 || zt >= rows[f] || j - i > XM_KU || i - j > XM_KL) return XC_OK;
                    xm_at(a, i, j) += v;
                }
        }
    s->xm_ok = xm_band_lu(nb, a, s->xm_piv);
    return XC_OK;
}

// Decide for this solve (fluidmech!) from o->xmean: -1 or 0 off; 1 probe the operator now; 2 reuse the operator
// the last probing solve of this grid left (probe if there is none).  Reuse keeps the results of a solve a
// function of the calls of the current time step only (run! probes in each step's first Picard sweep), so a
// resumed run repeats the uninterrupted one bit for bit.
int xm_prepare(xc_ctx* ctx, StokesCtx* s, const Flags& F, const xc_fm_opts* o) {
    s->xm_active = false;
    if (F.bnchm || o->xmean <= 0) return XC_OK;
    if (o->xmean == 1 || !s->xm_built) {
        int st;
        if ((st = xm_build(ctx, s, F))) return st;
        s->xm_built = true;
    }
    s->xm_active = s->xm_ok;
    return XC_OK;
}

// x += R L0^-1 q with q = R^T r of the residual of x (in s->xm_host, reduced with that residual); false if the
// correction is not finite (x unchanged)
bool xm_correct(xc_ctx* ctx, StokesCtx* s, double* x, int64_t n) {
    const int Nz = s->Nz, n0 = s->xm_n0, nb = 3 * (Nz + 2);
    const int rows[3] = {Nz + 1, Nz + 2, Nz + 2};
    std::vector<double> b((size_t)nb, 0.0);
    for (int f = 0; f < 3; f += 2)  // the U rows stay 0 (identity rows): U is not corrected
        for (int z = 0; z < rows[f]; ++z) b[3 * z + f] = s->xm_host[xm_prof(f, z, Nz)];
    xm_band_solve(nb, s->xm_lu, s->xm_piv, b.data());
    double* d = s->xm_host + n0;
    for (int f = 0; f < 3; ++f)
        for (int z = 0; z < rows[f]; ++z) {
            const double v = b[3 * z + f];
            if (!std::isfinite(v)) return false;
            d[xm_prof(f, z, Nz)] = v;
        }
    if (hipMemcpyAsync(s->xm_delta, d, (size_t)n0 * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
        return false;
    hipLaunchKernelGGL(k_xm_expand, dim3(nblk(n)), dim3(NT), 0, ctx->stream, (const double*)s->xm_delta, x, n, Nz,
                       s->Nx, -1, 0);
    return true;
}

// The residual of the starting x: r0_ready: already in s->rs with its squared norm in s->host[R0_SLOT] (enqueued by
// the caller before a synchronization it needed anyway, enqueue_r0), else computed here.
constexpr int R0_SLOT = 4;
// b2: where ||r||^2 goes (default: the host's slot); bb: where ||b||^2 goes (none by default)
void enqueue_r0(xc_ctx* ctx, StokesCtx* s, const Flags& F, const double* x, const double* b, bool with_q = false,
                double* b2 = nullptr, double* bb = nullptr) {
    const LevD dd = s->lev[0].d();
    const unsigned nbl = ll_cells_blocks(dd.Nz, dd.Nx);
    const int nw = (int)(nbl * (NT / 64));
    double* part = s->part + DOT_BLOCKS;
    double* bpart = bb ? part + nw : nullptr;
    hipLaunchKernelGGL(k_ll_cells<true>, dim3(nbl), dim3(NT), 0, ctx->stream, dd, F, x, b, s->rs, part,
                       (double*)nullptr, bpart);
    double* b2out = b2 ? b2 : s->host_dev + R0_SLOT;
    if (s->xm_active && with_q)  // R^T r for the cycle start's correction, the norms in the same launch
        xm_sum(ctx, s, s->rs, s->xm_hdev, 1, 0, part, nw, b2out, bpart, bb);
    else
        hipLaunchKernelGGL(k_sum_final1, dim3(1), dim3(NT), 0, ctx->stream, (const double*)part, nw, b2out,
                           (const double*)bpart, bb);
}
int fgmres_run(xc_ctx* ctx, StokesCtx* s, const Flags& F, const MGOpts& mo, int m, int64_t n, const double* b,
               double* x, double target, double bnorm, int maxit, const LcCfg& lc, SolveOut* out,
               bool r0_ready = false) {
    int st;
    const unsigned nb = nblk(n);
    const LevD dd = s->lev[0].d();
    int iters = 0;
    double rel = 0.0, rel0 = -1.0;
    bool conv = false;
    // The Arnoldi steps run on the device without host round trips: the Gram-Schmidt
    // coefficients, the Hessenberg column and the Givens rotations never leave the GPU.
    // The host learns the residual estimate |g[j+1]| of step j only after it has
    // enqueued step j+1, so the device never waits for the host.  Near convergence
    // (step j's estimate extrapolated from the last two below NEAR_CONV times the bar)
    // the host waits for step j before enqueuing step j+1 instead: if step j has
    // converged, its column is closed with its provisional values and the preconditioner,
    // operator and reduction of a step the solution would not use are never run (the
    // device idles for one host round trip instead).
    Kry K{};
    K.H = s->kry;
    K.cs = K.H + (size_t)(m + 1) * m;
    K.sn = K.cs + m;
    K.g = K.sn + m;
    K.y = K.g + m + 1;
    double* dd1 = s->d3;                  // step reductions (2m + 3)
    double* cf = s->d3 + 3 * (m + 2);     // update coefficients and nu
    volatile double* hout = s->host + 8;  // 2 per step: provisional |g[j+1]|, nu_{j+1}
    // A learned step is kept only while it cuts the residual at least as much as the plain preconditioner does
    // per iteration on average (its measured rate), raised to the learned step's relative cost; the first step
    // that does not switches the rest of the solve to the plain preconditioner.  The Krylov basis is kept:
    // FGMRES allows a different preconditioner in every step.
    LcCfg cur = lc;
    const double lc_bar = pow(s->plain_rate > 0 ? s->plain_rate : 0.2, LC_COST);
    // With the horizontal-mean correction every step also fixes its constant-pressure component exactly (the single
    // interior pin: LL 1_P = e_pin): the preconditioner leaves errors in the pressure level that the residual hardly
    // sees, and without the fix they differ from one Picard sweep to the next, which the Picard test reads as a
    // change of P (config 1: 210 sweeps in 20 steps instead of the 160 of the sparse-LU oracle, with it 160).
    const bool pinfix = s->xm_active && F.pin_kind == 1 && !F.bnchm;
    bool has_pshift = false;
    double pshift = 0.0;
    int lsteps = 0;
    bool dropped = false;
    constexpr double NEAR_CONV = 3.0;
    bool first = true;
    while (true) {
        if (!(first && r0_ready)) {  // r = b - LL x and ||r||^2 (per-wave partials, summed into the mapped host buffer)
            enqueue_r0(ctx, s, F, x, b, s->xm_active);  // with R^T r (the first residual has it: xc_fluidmech)
            XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
        }
        first = false;
        // a cycle that will run: its starting guess corrected on the x-uniform fields first.  The corrected
        // residual's norm stays on the device (the cycle start reads it); the host reads it once step 0 is done,
        // which it waits for anyway.  (Should the correction have converged the solve, the cycle's first step
        // finds that out.)
        const bool dev_beta = s->xm_active && iters < maxit && isfinite(s->host[R0_SLOT]) &&
                              sqrt(s->host[R0_SLOT]) > target && xm_correct(ctx, s, x, n);
        if (dev_beta) enqueue_r0(ctx, s, F, x, b, false, s->xm_b2);
        double beta = sqrt(s->host[R0_SLOT]);
        bool beta_known = !dev_beta;
        auto know_beta = [&]() {  // once step 0 has completed (its event waited for)
            if (beta_known) return;
            beta_known = true;
            beta = sqrt(s->host[R0_SLOT]);
            rel = bnorm > 0 ? beta / bnorm : beta;
            if (rel0 < 0) rel0 = rel;
        };
        if (beta_known) {
            rel = bnorm > 0 ? beta / bnorm : beta;
            if (rel0 < 0) rel0 = rel;
        }
        if (beta_known && !isfinite(beta)) break;  // not converged; a learned solve falls back to the plain one
        if (beta_known && beta <= target) {
            conv = true;
            if (pinfix) {
                // The steps fixed the pressure level through the pin row, where it rests on the divergence of one
                // cell; set it instead from the global mass balance, the sum of every divergence row, which the
                // boundary fluxes fix: x_P += c with c the sum of the final residual over the interior P rows (its
                // row sums R^T r came with it)
                const int Nz = dd.Nz;
                double c = 0.0;
                for (int z = 1; z <= Nz; ++z) c += s->xm_host[xm_prof(2, z, Nz)];
                if (isfinite(c)) {
                    has_pshift = true;
                    pshift = c;
                }
            }
            break;
        }
        if (iters >= maxit) break;
        if (dev_beta)
            hipLaunchKernelGGL(k_cycle_start_dev, dim3(nb), dim3(NT), 0, ctx->stream, s->rs, (const double*)s->xm_b2,
                               s->V, n, K, m, s->host_dev + R0_SLOT);
        else
            hipLaunchKernelGGL(k_cycle_start, dim3(nb), dim3(NT), 0, ctx->stream, s->rs, 1.0 / beta, s->V, n, K, m,
                               beta);
        int used = 0;  // basis columns whose results are final
        int j = 0;
        bool waited = false;  // step j - 1 was waited for before step j was enqueued
        bool early = false;   // converged at step j - 1 with step j not enqueued: close column j - 1
        for (; j < m; ++j) {
            double* vj = s->V + (int64_t)j * n;  // p_j, provisional until this step's update
            double* zj = s->Z + (int64_t)j * n;
            waited = false;
            if (j >= 2 && !cur.on) {
                const double r1 = hout[2 * (j - 2)], r0 = j >= 3 ? hout[2 * (j - 3)] : beta;
                const double pred = r0 > 0.0 ? r1 * (r1 / r0) : r1;
                if (pred <= 0.5 * target * NEAR_CONV) {
                    XC_TRY_HIP(ctx, hipEventSynchronize(s->ev[(j - 1) & 1]));
                    waited = true;
                    used = j;
                    ++iters;
                    if (hout[2 * (j - 1)] <= 0.5 * target || hout[2 * (j - 1) + 1] == 0.0 || iters >= maxit) {
                        early = true;
                        break;
                    }
                }
            }
            if ((st = precond_full(ctx, s, F, mo, n, cur, vj, zj))) return st;
            hipLaunchKernelGGL(k_ll_cells<false>, dim3(ll_cells_blocks(dd.Nz, dd.Nx)), dim3(NT), 0, ctx->stream, dd, F,
                               zj, pinfix ? vj : zj, s->ws, (double*)nullptr, pinfix ? s->xm_pinc + j : (double*)nullptr);
            {
                const dim3 g((unsigned)md_blocks(n)), b(NT), g1(1);
                switch (md_ept(n)) {
                    case 1: hipLaunchKernelGGL(k_multidot2<1>, g, b, 0, ctx->stream, s->V, n, j, vj, s->ws, s->part); break;
                    case 2: hipLaunchKernelGGL(k_multidot2<2>, g, b, 0, ctx->stream, s->V, n, j, vj, s->ws, s->part); break;
                    case 4: hipLaunchKernelGGL(k_multidot2<4>, g, b, 0, ctx->stream, s->V, n, j, vj, s->ws, s->part); break;
                    default: hipLaunchKernelGGL(k_multidot2<8>, g, b, 0, ctx->stream, s->V, n, j, vj, s->ws, s->part); break;
                }
                hipLaunchKernelGGL(k_arnoldi2, g1, dim3(ARN_NT), 0, ctx->stream, K, m, j, s->part, (int)g.x, dd1, cf,
                                   s->host_dev + 8);
            }
            hipLaunchKernelGGL(k_dcgs_update, dim3(nb), dim3(NT), 0, ctx->stream, s->V, n, j, dd1, cf, vj, s->ws,
                               s->V + (int64_t)(j + 1) * n);
            XC_TRY_HIP(ctx, hipEventRecord(s->ev[j & 1], ctx->stream));
            if (cur.on) {  // judge the learned step now (the pipeline waits for it)
                XC_TRY_HIP(ctx, hipEventSynchronize(s->ev[j & 1]));
                know_beta();
                if (!isfinite(beta)) break;
                const double before = j == 0 ? beta : hout[2 * (j - 1)];
                ++lsteps;
                if (!(hout[2 * j] <= before * lc_bar)) {
                    cur.on = false;
                    dropped = true;
                }
            }
            if (j == 0 || waited) continue;
            // step j-1 is complete once its event has fired; step j (enqueued) finalises its column
            XC_TRY_HIP(ctx, hipEventSynchronize(s->ev[(j - 1) & 1]));
            know_beta();
            if (!isfinite(beta)) break;
            used = j;
            ++iters;
            if (hout[2 * (j - 1)] <= 0.5 * target || hout[2 * (j - 1) + 1] == 0.0 || iters >= maxit) break;
        }
        // column used - 1 keeps its provisional values when the cycle converged before its next step was
        // enqueued (early) or when the cycle is full (its last step was not tested yet)
        if (!beta_known) {  // the cycle ended before step 0 was waited for (m = 1)
            XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            know_beta();
        }
        if (!isfinite(beta)) break;  // a non-finite corrected residual: not converged
        bool close = early;
        if (!early && j == m) {
            XC_TRY_HIP(ctx, hipEventSynchronize(s->ev[(m - 1) & 1]));
            used = m;
            ++iters;
            close = true;
        }
        if (used > 0) {
            const unsigned gu = nb < 1024u ? nb : 1024u;
            const size_t lds = ((size_t)used * (used + 1) / 2 + (size_t)used) * sizeof(double);
            hipLaunchKernelGGL(k_update_x_solve, dim3(gu), dim3(NT), lds, ctx->stream, K, m, used, close ? 1 : 0,
                               s->Z, n, x, pinfix ? (const double*)s->xm_pinc : (const double*)nullptr,
                               n - (int64_t)(dd.Nz + 2) * (dd.Nx + 2));
        }
    }
    out->iters = iters;
    out->conv = conv;
    out->rel = rel;
    out->rel0 = rel0;
    out->learned_steps = lsteps;
    out->dropped = dropped;
    out->has_pshift = has_pshift;
    out->pshift = pshift;
    return XC_OK;
}

int check_fm(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_fm_opts* o) {
    if (!fl || !g || !o) return xc_fail(ctx, XC_EINVAL_ARG, "fluidmech!: NULL argument");
    const int64_t Nz = g->Nz, Nx = g->Nx;
    if (Nz < 4 || Nx < 4) return xc_fail(ctx, XC_EINVAL_SHAPE, "fluidmech!: grid must be at least 4x4");
    struct { const xc_mat* m; int64_t r, c; const char* n; } t[] = {
        {&fl->W, Nz + 1, Nx + 2, "W"}, {&fl->U, Nz + 2, Nx + 1, "U"}, {&fl->P, Nz + 2, Nx + 2, "P"},
        {&fl->eta, Nz, Nx, "eta"}, {&fl->etaco, Nz + 1, Nx + 1, "etaco"}, {&fl->rho, Nz, Nx, "rho"},
        {&fl->rhow, Nz + 1, Nx, "rhow"}, {&fl->rhou, Nz, Nx + 1, "rhou"}, {&fl->Drho, Nz + 1, Nx, "Drho"},
        {&fl->MFS, Nz, Nx, "MFS"}, {&fl->rhoWo, Nz + 1, Nx, "rhoWo"}, {&fl->rhoWoo, Nz + 1, Nx, "rhoWoo"},
        {&fl->rhoUo, Nz, Nx + 1, "rhoUo"}, {&fl->rhoUoo, Nz, Nx + 1, "rhoUoo"}};
    for (auto& e : t)
        if (!xc_shape(*e.m, e.r, e.c)) return xc_fail(ctx, XC_EINVAL_SHAPE, "FluidState.%s has the wrong shape", e.n);
    if (o->sds != 1 && o->sds != -1) return xc_fail(ctx, XC_EINVAL_ARG, "fluidmech!: sds must be +-1");
    if (o->xbc != XC_PERIODIC && o->xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "fluidmech!: xBC");
    if (o->zbc != XC_PERIODIC && o->zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "fluidmech!: zBC");
    if (xc_scheme_halo(o->scheme) < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "fluidmech!: scheme");
    if (o->mg_levels < 0) return xc_fail(ctx, XC_EINVAL_ARG, "fluidmech!: mg_levels must be >= 0");
    return XC_OK;
}

}  // namespace

extern "C" {

int xc_stokes_apply(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_fm_opts* o,
                    int bnchm, const double* x, double* y) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_fm(ctx, fl, g, o))) return st;
    StokesCtx* s;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, 1, &s, o->mg_levels))) return st;
    build_levels(ctx, s, fl, g->h);
    const Flags F = make_flags(o, par, (int)g->Nz, (int)g->Nx, bnchm != 0);
    const int64_t n = s->lev[0].nv + (g->Nz + 2) * (g->Nx + 2);
    hipLaunchKernelGGL(k_ll_cells<false>, dim3(ll_cells_blocks(g->Nz, g->Nx)), dim3(NT), 0, ctx->stream, s->lev[0].d(), F,
                       x, x, y);
    return xc_finish(ctx, "stokes_apply");
}

int xc_stokes_rhs(xc_ctx* ctx, xc_fluid* fl, const xc_phase* ph, const xc_grid* g, const xc_par* par,
                  const xc_fm_opts* o, const xc_bnchm* bn, double* rr) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_fm(ctx, fl, g, o))) return st;
    if (!bn && !ph) return xc_fail(ctx, XC_EINVAL_ARG, "fluidmech!: phase required outside the benchmark mode");
    StokesCtx* s;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, 1, &s, o->mg_levels))) return st;
    const Flags F = make_flags(o, par, (int)g->Nz, (int)g->Nx, bn != nullptr);
    if ((st = build_rhs(ctx, s, fl, ph, g, par, o, bn, F, rr))) return st;
    return xc_finish(ctx, "stokes_rhs");
}

int xc_fluidmech(xc_ctx* ctx, xc_fluid* fl, const xc_phase* ph, const xc_grid* g, const xc_par* par,
                 const xc_fm_opts* o, const xc_bnchm* bn, xc_solve_stats* stats) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_fm(ctx, fl, g, o))) return st;
    if (!bn && !ph) return xc_fail(ctx, XC_EINVAL_ARG, "fluidmech!: phase required outside the benchmark mode");
    const int m = o->restart > 0 ? (o->restart < 60 ? o->restart : 60) : 30;
    StokesCtx* s;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, m, &s, o->mg_levels))) return st;
    const int Nz = (int)g->Nz, Nx = (int)g->Nx;
    const Flags F = make_flags(o, par, Nz, Nx, bn != nullptr);
    const int64_t n = s->lev[0].nv + (int64_t)(Nz + 2) * (Nx + 2);
    const unsigned nb = nblk(n);
    // system: coefficients, RR, initial guess
    build_levels(ctx, s, fl, g->h, false);
    const bool depth_pending = depth_launch(ctx, s, F, o);
    if ((st = pois_table(ctx, s, F, Nz, Nx, g->h))) return st;
    if ((st = xm_prepare(ctx, s, F, o))) return st;
    if ((st = build_rhs(ctx, s, fl, ph, g, par, o, bn, F, s->bs, s->xs))) return st;  // and the starting guess
    // the starting residual, its x-uniform part R^T r and ||RR|| (summed beside ||r||): one synchronization
    enqueue_r0(ctx, s, F, s->xs, s->bs, true, nullptr, s->host_dev);
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double bnorm = sqrt(s->host[0]);
    if ((st = choose_depth(ctx, s, F, o, g->h, depth_pending, true))) return st;  // synchronized above
    const MGOpts mo = mg_opts(o, s);
    build_coarse(ctx, s, s->leff);
    const double target = o->tol * (bnorm > 0 ? bnorm : 1.0);
    const int maxit = o->maxit > 0 ? o->maxit : 500;
    int iters = 0;
    double rel = 0.0, rel0 = -1.0;
    bool conv = false;
    LcCfg lc{};
    if ((st = lc_config(ctx, s, o, g, par, &lc))) return st;
    int fallback = 0;
    if (lc.on && s->lc_skip > 0) {  // recently useless on these systems: plain this time
        --s->lc_skip;
        lc.on = false;
        fallback = 1;
    }
    SolveOut so{};
    if ((st = fgmres_run(ctx, s, F, mo, m, n, s->bs, s->xs, target, bnorm, maxit, lc, &so, true))) return st;
    if (lc.on && !so.conv && !(so.rel < 1e30)) {  // non-finite iterate: once more, plain, from the initial guess
        const int used = so.iters;
        const double rel0_l = so.rel0;
        lc.on = false;
        so.dropped = true;
        hipLaunchKernelGGL(k_gather, dim3(nb), dim3(NT), 0, ctx->stream, fl->W, fl->U, fl->P, s->xs, Nz, Nx);
        const int ls = so.learned_steps;
        if ((st = fgmres_run(ctx, s, F, mo, m, n, s->bs, s->xs, target, bnorm, maxit - used, lc, &so))) return st;
        so.iters += used;
        so.rel0 = rel0_l;
        so.learned_steps = ls;
        so.dropped = true;
    }
    if (lc.on) {
        fallback |= so.dropped ? 1 : 0;
        if (so.dropped && so.learned_steps <= 1) {
            if (++s->lc_fail_streak >= LC_STREAK) {
                s->lc_skip = LC_SKIP << (s->lc_backoff < LC_BACKOFF ? s->lc_backoff : LC_BACKOFF);
                ++s->lc_backoff;
                s->lc_fail_streak = LC_STREAK - 1;
            }
        } else {
            s->lc_fail_streak = 0;
            s->lc_backoff = 0;
        }
    } else if (so.conv && so.iters > 0 && so.rel > 0 && so.rel0 > so.rel) {
        // the plain preconditioner's measured reduction per iteration (exponential average over solves)
        const double r = pow(so.rel / so.rel0, 1.0 / so.iters);
        s->plain_rate = s->plain_rate > 0 ? 0.7 * s->plain_rate + 0.3 * r : r;
    }
    iters = so.iters;
    conv = so.conv;
    rel = so.rel;
    rel0 = so.rel0;
    s->xm_active = false;
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(NT), 0, ctx->stream, s->xs, fl->W, fl->U, fl->P, Nz, Nx,
                       so.has_pshift ? 1 : 0, so.pshift);
    if (stats) {
        stats->iters = iters;
        stats->converged = conv;
        stats->rel_resid = rel;
        stats->init_rel_resid = rel0;
        stats->mg_levels = s->leff;
        stats->inertia_ratio = s->inertia;
        stats->fallback_used = fallback;
        stats->learned_steps = so.learned_steps;
    }
    if ((st = xc_finish(ctx, "fluidmech!"))) return st;
    if (!conv) return xc_fail(ctx, XC_ENOCONV, "fluidmech!: FGMRES reached %d iterations at relative residual %.3e",
                              iters, rel);
    return XC_OK;
}


// shared set-up of the solver-piece entry points: coefficients, flags, Poisson table
static int piece_setup(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_fm_opts* o,
                       int bnchm, int m, StokesCtx** sp, Flags* F, MGOpts* mo, int64_t* n) {
    int st;
    if ((st = check_fm(ctx, fl, g, o))) return st;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, m, sp, o->mg_levels))) return st;
    StokesCtx* s = *sp;
    s->xm_active = false;  // the solver pieces run the plain method
    build_levels(ctx, s, fl, g->h, false);
    *F = make_flags(o, par, (int)g->Nz, (int)g->Nx, bnchm != 0);
    if ((st = choose_depth(ctx, s, *F, o, g->h, depth_launch(ctx, s, *F, o), false))) return st;
    build_coarse(ctx, s, s->leff);
    if ((st = pois_table(ctx, s, *F, (int)g->Nz, (int)g->Nx, g->h))) return st;
    *mo = mg_opts(o, s);
    *n = s->lev[0].nv + (g->Nz + 2) * (g->Nx + 2);
    return XC_OK;
}

int xc_stokes_precond(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_fm_opts* o,
                      int bnchm, const double* v, double* z) {
    XC_CHECK_CTX(ctx);
    if (!v || !z) return xc_fail(ctx, XC_EINVAL_ARG, "stokes_precond: NULL vector");
    StokesCtx* s;
    Flags F;
    MGOpts mo;
    int64_t n;
    int st;
    if ((st = piece_setup(ctx, fl, g, par, o, bnchm, 1, &s, &F, &mo, &n))) return st;
    precond(ctx, s, F, v, z, mo);
    return xc_finish(ctx, "stokes_precond");
}

int xc_stokes_solve(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_fm_opts* o,
                    int bnchm, const double* rhs, double* x, xc_solve_stats* stats) {
    XC_CHECK_CTX(ctx);
    if (!rhs || !x) return xc_fail(ctx, XC_EINVAL_ARG, "stokes_solve: NULL vector");
    const int m = o && o->restart > 0 ? (o->restart < 60 ? o->restart : 60) : 30;
    StokesCtx* s;
    Flags F;
    MGOpts mo;
    int64_t n;
    int st;
    if ((st = piece_setup(ctx, fl, g, par, o, bnchm, m, &s, &F, &mo, &n))) return st;
    if ((st = multidot(ctx, s, rhs, 0, rhs, n))) return st;
    const double bnorm = sqrt(s->host[0]);
    const double target = o->tol * (bnorm > 0 ? bnorm : 1.0);
    const int maxit = o->maxit > 0 ? o->maxit : 500;
    LcCfg lc{};
    if ((st = lc_config(ctx, s, o, g, par, &lc))) return st;
    SolveOut so{};
    if ((st = fgmres_run(ctx, s, F, mo, m, n, rhs, x, target, bnorm, maxit, lc, &so))) return st;
    if (so.has_pshift) {  // (not with the solver pieces, which run without the horizontal-mean correction)
        const int64_t nP = (int64_t)(g->Nz + 2) * (g->Nx + 2);
        hipLaunchKernelGGL(k_p_shift, dim3(nblk(nP)), dim3(NT), 0, ctx->stream, x + (n - nP), nP, so.pshift);
    }
    if (stats) {
        stats->iters = so.iters;
        stats->converged = so.conv;
        stats->rel_resid = so.rel;
        stats->init_rel_resid = so.rel0;
        stats->mg_levels = s->leff;
        stats->inertia_ratio = s->inertia;
        stats->fallback_used = so.dropped ? 1 : 0;
        stats->learned_steps = so.learned_steps;
    }
    if ((st = xc_finish(ctx, "stokes_solve"))) return st;
    if (!so.conv) return xc_fail(ctx, XC_ENOCONV, "stokes_solve: FGMRES reached %d iterations at relative residual %.3e",
                                 so.iters, so.rel);
    return XC_OK;
}

int xc_stokes_lc_restrict(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, const xc_par* par, int level,
                          const double* vec, float* out4, float* sigma, int compute_sigma) {
    XC_CHECK_CTX(ctx);
    if (!vec || !out4 || !sigma) return xc_fail(ctx, XC_EINVAL_ARG, "stokes_lc_restrict: NULL buffer");
    xc_fm_opts o{};
    o.sds = -1;
    o.top_cnv = 1;
    o.bot_cnv = 1;
    o.learned = 1;
    o.lc_level = level;
    int st;
    if ((st = check_fm(ctx, fl, g, &o))) return st;
    StokesCtx* s;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, 1, &s))) return st;
    build_levels(ctx, s, fl, g->h);
    const int f = 1 << (level > 0 ? level : 2);
    if (level < 1 || g->Nz % f || g->Nx % f) return xc_fail(ctx, XC_EINVAL_SHAPE, "stokes_lc_restrict: level %d", level);
    LcCfg lc{};
    lc.on = true;
    lc.level = level;
    lc.f = f;
    lc.Hc = (int)g->Nz / f;
    lc.Wc = (int)g->Nx / f;
    xc_unet_io_scales(ctx, lc.is, lc.os);
    lc.eta_ref = par->etam0;
    lc_restrict(ctx, s, s->lev[0].d(), lc, vec, out4, sigma, compute_sigma);
    return xc_finish(ctx, "stokes_lc_restrict");
}

int xc_stokes_krylov_vector(xc_ctx* ctx, int j, double* out) {
    XC_CHECK_CTX(ctx);
    StokesCtx* s = (StokesCtx*)ctx->ext[XC_EXT_STOKES];
    if (!s || !out || j < 0 || j > s->m) return xc_fail(ctx, XC_EINVAL_ARG, "stokes_krylov_vector: no basis vector %d", j);
    const int64_t n = s->lev[0].nv + (int64_t)(s->Nz + 2) * (s->Nx + 2);
    XC_TRY_HIP(ctx, hipMemcpyAsync(out, s->V + (int64_t)j * n, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    return xc_finish(ctx, "stokes_krylov_vector");
}

int xc_stokes_lc_prolong(xc_ctx* ctx, const xc_fluid* fl, const xc_grid* g, int level, const float* e3,
                         const float* sigma, double* z) {
    XC_CHECK_CTX(ctx);
    if (!e3 || !sigma || !z) return xc_fail(ctx, XC_EINVAL_ARG, "stokes_lc_prolong: NULL buffer");
    xc_fm_opts o{};
    o.sds = -1;
    o.top_cnv = 1;
    o.bot_cnv = 1;
    int st;
    if ((st = check_fm(ctx, fl, g, &o))) return st;
    StokesCtx* s;
    if ((st = setup(ctx, (int)g->Nz, (int)g->Nx, 1, &s))) return st;
    build_levels(ctx, s, fl, g->h);
    const int f = 1 << level;
    if (level < 1 || g->Nz % f || g->Nx % f) return xc_fail(ctx, XC_EINVAL_SHAPE, "stokes_lc_prolong: level %d", level);
    LcProArgs pa{};
    pa.L = s->lev[0].d();
    pa.f = f;
    pa.Hc = (int)g->Nz / f;
    pa.Wc = (int)g->Nx / f;
    pa.e = e3;
    pa.sigma = sigma;
    xc_unet_io_scales(ctx, nullptr, pa.os);
    pa.z = z;
    pa.n = s->lev[0].nv + (g->Nz + 2) * (g->Nx + 2);
    hipLaunchKernelGGL(k_lc_prolong, dim3(nblk(pa.n)), dim3(NT), 0, ctx->stream, pa);
    return xc_finish(ctx, "stokes_lc_prolong");
}
}  // extern "C"
