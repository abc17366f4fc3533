// xc_fields.hip -- constitutive update (update!), phase evolution (phsevo!),
// history shift (store_previous!), phase velocities and the CFL time step.
//
// Reference: src/update.jl, src/store.jl, src/phsevo.jl:164-214, src/timing.jl.
//
// MI355X layout of the work: the ~40 reference broadcasts/kernels of one
// update! become 8 launches.  Values the reference stores and then re-reads at
// neighbours (rho, chi, mu, exz, txz) are recomputed from their inputs where
// that saves a pass -- the recomputation is the same expression on the same
// operands, hence the same bits.  Global minima (update.jl:331,396) are
// per-workgroup partials plus one finalising workgroup; the row means of rhow
// (update.jl:86,122) are chunked partial sums combined in a fixed order, and
// the lithostatic z-scan (update.jl:125-131) stays sequential like the
// reference's loop.  FP contraction is off to keep the reference's rounding.
#include <math.h>

#include "xc_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr double EPS = 2.220446049250313e-16;
constexpr int NT = 256;

__device__ __forceinline__ int gidx(int i, int n, int bc) {
    if (i < 0) return bc == XC_PERIODIC ? n - 1 : 0;
    if (i >= n) return bc == XC_PERIODIC ? 0 : n - 1;
    return i;
}
__device__ __forceinline__ double jmax(double a, double b) {
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}
__device__ __forceinline__ double jclamp(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

struct DM {  // device view of one matrix
    double* p;
    int64_t ld;
    __device__ __forceinline__ double& operator()(int i, int j) const { return p[i + (int64_t)j * ld]; }
};
struct DB {
    uint8_t* p;
    int64_t ld;
    __device__ __forceinline__ uint8_t& operator()(int i, int j) const { return p[i + (int64_t)j * ld]; }
};
static inline DM dm(const xc_mat& m) { return DM{m.p, m.ld}; }
static inline DB db(const xc_bmat& m) { return DB{m.p, m.ld}; }

struct Par {
    double rhom0, rhox0, d0, g0, etam0, etax0;
    double AA[4], BB[4], CC[4];
    double Ptop, etacntr, L0, l0, D0, h, invh;
    int Nz, Nx, xbc, zbc;
};

// --- update_volume_fractions! (update.jl:18-35), one cell ------------------
struct VF { double rho, chi, mu; };
__device__ __forceinline__ VF volfrac(double x, double m, const Par& p) {
    VF r;
    r.rho = 1.0 / (m / p.rhom0 + x / p.rhox0);
    r.chi = jclamp(x * r.rho / p.rhox0, EPS, 1.0 - EPS);
    r.mu = jclamp(m * r.rho / p.rhom0, EPS, 1.0 - EPS);
    return r;
}

// K1: volume fractions + face interpolations (update.jl:18-104).  Domain
// (Nz+2) x (Nx+2); each lane writes every array entry that lives at its index.
struct InterpArgs {
    DM x, m, rho, chi, mu, rhow, rhou, Drhom, Drhox, chiw, muw, chiu, muu, x_w, m_w, x_u, m_u;
    DB hasx, hasm;
    Par p;
    xc_dom d;
};

__global__ __launch_bounds__(NT) void k_interp(InterpArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    const Par& p = a.p;
    const int Nz = p.Nz, Nx = p.Nx;
    if (j >= Nz + 2 || i >= Nx + 2) return;
    auto vf = [&](int z, int x) { return volfrac(a.x(z, x), a.m(z, x), p); };
    const int z0 = gidx(j - 1, Nz, p.zbc), z1 = gidx(j, Nz, p.zbc);
    const int x0 = gidx(i - 1, Nx, p.xbc), x1 = gidx(i, Nx, p.xbc);
    // the four cells every output of this lane reads, each evaluated once (gidx is the identity inside
    // the grid, so the cell (j, i) is (z1, x1) where it exists, (z0, i) is (z0, x1) for i < Nx, and
    // (j, x0) is (z1, x0) for j < Nz)
    const VF A = vf(z0, x0), B = vf(z0, x1), C = vf(z1, x0), D = vf(z1, x1);
    if (j < Nz && i < Nx) {
        const double x = a.x(j, i), m = a.m(j, i);
        a.rho(j, i) = D.rho;
        a.chi(j, i) = D.chi;
        a.mu(j, i) = D.mu;
        const double sq = sqrt(EPS);
        a.hasx(j, i) = x >= sq;
        a.hasm(j, i) = m >= sq;
    }
    if (j <= Nz && i < Nx) {  // z-faces (Nz+1, Nx)
        const double rw = (B.rho + D.rho) / 2.0;
        a.rhow(j, i) = rw;
        a.Drhom(j, i) = p.rhom0 - rw;
        a.Drhox(j, i) = p.rhox0 - rw;
        a.x_w(j, i) = (a.x(z0, i) + a.x(z1, i)) / 2.0;
        a.m_w(j, i) = (a.m(z0, i) + a.m(z1, i)) / 2.0;
    }
    if (j < Nz && i <= Nx) {  // x-faces (Nz, Nx+1)
        a.rhou(j, i) = (C.rho + D.rho) / 2.0;
        a.x_u(j, i) = (a.x(j, x0) + a.x(j, x1)) / 2.0;
        a.m_u(j, i) = (a.m(j, x0) + a.m(j, x1)) / 2.0;
    }
    if (j <= Nz) {  // W-type (Nz+1, Nx+2): icx column i
        a.chiw(j, i) = (A.chi + C.chi) / 2.0;
        a.muw(j, i) = (A.mu + C.mu) / 2.0;
    }
    if (i <= Nx) {  // U-type (Nz+2, Nx+1): icz row j
        a.chiu(j, i) = (A.chi + B.chi) / 2.0;
        a.muu(j, i) = (A.mu + B.mu) / 2.0;
    }
}

// K2: chunked row sums of rhow (update.jl:86,122): partial[c][j] over columns of chunk c
constexpr int ROW_CHUNK = 32;
__global__ __launch_bounds__(64) void k_rowsum(DM rhow, int rows, int cols, double* partial) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= rows) return;
    const int i0 = c * ROW_CHUNK;
    double v[ROW_CHUNK];
#pragma unroll
    for (int u = 0; u < ROW_CHUNK; ++u) v[u] = i0 + u < cols ? rhow(j, i0 + u) : 0.0;
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < ROW_CHUNK; ++u) acc += v[u];
    partial[(int64_t)c * rows + j] = acc;
}

// K3: row means + sequential lithostatic scan (update.jl:122-131) in one workgroup of RM_NT lanes: the
// means in one pass, then one lane adds them in row order (operands read ahead from LDS in blocks of 16,
// the row count padded so that the blocks need no bounds tests), the results stored by all lanes
constexpr int RM_NT = 1024;
__global__ __launch_bounds__(RM_NT) void k_rowmean_scan(const double* partial, int rows, int nchunk, int cols,
                                                         double g0, double h, double Ptop, double* ref, double* plcol) {
    extern __shared__ double sref[];  // rows + 16 means, then rows + 16 scan results
    const int pad = ((rows + 15) & ~15) + 16;
    double* spl = sref + pad;
    const double s = 1.0 / (double)cols;
    for (int j = threadIdx.x; j < pad; j += RM_NT) {
        double acc = 0.0;
        if (j < rows) {
            for (int c0 = 0; c0 < nchunk; c0 += 16) {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) v[u] = c0 + u < nchunk ? partial[(int64_t)(c0 + u) * rows + j] : 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u) acc += v[u];
            }
            const double r = acc * s;
            ref[j] = r;
            acc = r;
        }
        sref[j] = acc;  // 0 in the padding
    }
    __syncthreads();
    const int Nz = rows - 1;
    if (threadIdx.x == 0) {
        double pl = sref[0] * g0 * h / 2.0 + Ptop;
        spl[0] = pl;
        for (int j0 = 1; j0 < Nz; j0 += 16) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = sref[j0 + u];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                pl = pl + v[u] * g0 * h;
                spl[j0 + u] = pl;
            }
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < Nz; j += RM_NT) plcol[j] = spl[j];
}

// K4: Drho, Pl, Pt (update.jl:87, 132-134).  Domain (Nz+1) x Nx.
__global__ __launch_bounds__(NT) void k_drho_pt(DM rhow, DM Drho, DM Pl, DM Pt, DM P, const double* ref,
                                                const double* plcol, double floor_p, int Nz, int Nx, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j > Nz || i >= Nx) return;
    Drho(j, i) = rhow(j, i) - ref[j];
    if (j < Nz) {
        const double pl = plcol[j];
        Pl(j, i) = pl;
        Pt(j, i) = jmax(floor_p, pl + P(j + 1, i + 1));
    }
}

// x^y for x > 0 (and x = 0 with y > 0) as exp(y log x): within ~|y log x| + 2 ulp of pow(x, y), at a third of
// pow's instructions (pow carries log and exp in double-double); the rheology's bases are volume fractions over
// positive constants and viscosity ratios: |y log x| stays below ~400 (the default CC[1] = 0.09 on a volume
// fraction of 2.2e-16, a term of ~1e-169), a relative error below 1e-13 on each term
__device__ __forceinline__ double pos_pow(double x, double y) { return exp(y * log(x)); }

// permission-weight rheology of one cell (update.jl:144-177)
__device__ double etamix_cell(double f1, double f2, const Par& p) {
    double s11 = pos_pow(f1 / p.BB[0], 1.0 / p.CC[0]), s12 = pos_pow(f2 / p.BB[1], 1.0 / p.CC[1]);
    double s21 = pos_pow(f1 / p.BB[2], 1.0 / p.CC[2]), s22 = pos_pow(f2 / p.BB[3], 1.0 / p.CC[3]);
    double n1 = s11 + s12, n2 = s21 + s22;
    double Sf11 = s11 / n1, Sf12 = s12 / n1, Sf21 = s21 / n2, Sf22 = s22 / n2;
    double a1 = p.AA[0] * Sf11 + p.AA[1] * Sf12;
    double a2 = p.AA[2] * Sf21 + p.AA[3] * Sf22;
    double Xf11 = a1 * f1 + (1.0 - a1) * Sf11;
    double Xf12 = a1 * f2 + (1.0 - a1) * Sf12;
    double Xf21 = a2 * f1 + (1.0 - a2) * Sf21;
    double Xf22 = a2 * f2 + (1.0 - a2) * Sf22;
    const double ex = p.etax0, em = p.etam0;
    // pow(ex / ex, Xf11) and pow(em / em, Xf22) of the reference are pow(1, y) = 1 exactly (IEEE 754 pow)
    double thtv1 = 1.0 * pos_pow(em / ex, Xf12);
    double thtv2 = pos_pow(ex / em, Xf21) * 1.0;
    return f1 * (ex * thtv1) + f2 * (em * thtv2);
}

// exz at corner (j, i) of (Nz+1, Nx+1) (update.jl:215-223)
__device__ __forceinline__ double exz_at(const DM& W, const DM& U, int j, int i, double invh) {
    double dUdz = (U(j + 1, i) - U(j, i)) * invh;
    double dWdx = (W(j, i + 1) - W(j, i)) * invh;
    return (dUdz + dWdx) / 2.0;
}

// K5: rheology, kinematics, both viscosity kernels up to their global caps
// (update.jl:144-308, 345-370).  Domain (Nz+1) x (Nx+1): corners write exz,
// cells write everything else and the uncapped etai / etasi to scratch, plus
// per-workgroup minima of both.
struct CellArgs {
    DM chi, mu, etamix, W, U, exx, ezz, Div_V, exz, eII, V, wx, wm, vx, vm, rho;
    DM etae, ke, kx, ReL, fReL, Rel, fRel, ks, etat, Rc;
    double* etai;   // (Nz*Nx) scratch, uncapped eta
    double* etasi;  // (Nz*Nx) scratch, uncapped etas
    double* bmin;   // 2 * nb partial minima
    Par p;
    xc_dom d;
};

__global__ __launch_bounds__(NT) void k_cells(CellArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    const Par& p = a.p;
    const int Nz = p.Nz, Nx = p.Nx;
    double mn_eta = INFINITY, mn_etas = INFINITY;
    if (j <= Nz && i <= Nx) a.exz(j, i) = exz_at(a.W, a.U, j, i, p.invh);
    if (j < Nz && i < Nx) {
        const double f1 = a.chi(j, i), f2 = a.mu(j, i);
        const double em = etamix_cell(f1, f2, p);
        a.etamix(j, i) = em;
        // update.jl:202-213
        const double Uxc = (a.U(j + 1, i + 1) - a.U(j + 1, i)) * p.invh;
        const double Wzc = (a.W(j + 1, i + 1) - a.W(j, i + 1)) * p.invh;
        const double div = Uxc + Wzc;
        const double exx = Uxc - div / 3.0, ezz = Wzc - div / 3.0;
        a.Div_V(j, i) = div;
        a.exx(j, i) = exx;
        a.ezz(j, i) = ezz;
        // update.jl:225-241
        const double e1 = exz_at(a.W, a.U, j, i, p.invh), e2 = exz_at(a.W, a.U, j + 1, i, p.invh);
        const double e3 = exz_at(a.W, a.U, j, i + 1, p.invh), e4 = exz_at(a.W, a.U, j + 1, i + 1, p.invh);
        const double sq = (e1 * e1 + e2 * e2 + e3 * e3 + e4 * e4) / 4.0;
        const double eII = sqrt(EPS + (exx * exx + ezz * ezz) / 2.0 + sq);
        a.eII(j, i) = eII;
        const double wc = (a.W(j, i + 1) + a.W(j + 1, i + 1)) / 2.0;
        const double uc = (a.U(j + 1, i) + a.U(j + 1, i + 1)) / 2.0;
        const double V = sqrt(wc * wc + uc * uc);
        a.V(j, i) = V;
        // update.jl:243-253
        const double wa = (a.wx(j, i + 1) + a.wx(j + 1, i + 1)) / 2.0;
        const double vx = sqrt(wa * wa + EPS);
        a.vx(j, i) = vx;
        const double wb = (a.wm(j, i + 1) + a.wm(j + 1, i + 1)) / 2.0;
        a.vm(j, i) = sqrt(wb * wb + EPS);
        // update.jl:286-308
        const double rho = a.rho(j, i);
        const double ke = eII * p.L0 * p.L0;
        const double ReL = V * p.L0 / (em / rho);
        const double fReL = 1.0 - exp(-ReL);
        const double etae = fReL * ke * rho;
        const double etai = em + etae;
        a.etae(j, i) = etae;
        a.ke(j, i) = ke;
        a.ReL(j, i) = ReL;
        a.fReL(j, i) = fReL;
        // update.jl:345-370 (its kx supersedes the kx = ke of update.jl:304)
        const double Rel = vx * p.l0 * rho / em;
        const double fRel = 1.0 - exp(-Rel);
        const double ks = vx * p.l0;
        const double etat = fRel * ks * rho;
        const double etasi = em + etat;
        a.Rel(j, i) = Rel;
        a.fRel(j, i) = fRel;
        a.ks(j, i) = ks;
        a.etat(j, i) = etat;
        a.kx(j, i) = ks + fReL * ke;
        a.Rc(j, i) = V / jmax(vx, EPS);
        const int64_t o = j + (int64_t)i * Nz;
        a.etai[o] = etai;
        a.etasi[o] = etasi;
        mn_eta = etai;
        mn_etas = etasi;
    }
    // workgroup minima (exact: min is order independent): wave minima by DPP, then the NT / 64 of them
    __shared__ double s1[NT / 64], s2[NT / 64];
    mn_eta = wave_min(mn_eta);
    mn_etas = wave_min(mn_etas);
    if ((threadIdx.x & 63) == 0) {
        s1[threadIdx.x >> 6] = mn_eta;
        s2[threadIdx.x >> 6] = mn_etas;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double m1 = s1[0], m2 = s2[0];
        for (int w = 1; w < NT / 64; ++w) {
            m1 = fmin(m1, s1[w]);
            m2 = fmin(m2, s2[w]);
        }
        a.bmin[blockIdx.x] = m1;
        a.bmin[a.d.nb + blockIdx.x] = m2;
    }
}

// the minima of both viscosity kernels over the workgroup minima, one wave each, then the caps
__global__ __launch_bounds__(128) void k_min_final(const double* bmin, int nb, double etacntr, double* emax) {
    const int q = threadIdx.x >> 6;
    const double m = xc_rows_reduce<true>(bmin, nb, q);
    if ((threadIdx.x & 63) == 0) emax[q] = m * etacntr;  // update.jl:331, 396
}

// K6: harmonic cap + 50/50 Picard blend (update.jl:332-337, 397-402).  Cells.
__global__ __launch_bounds__(NT) void k_blend(DM eta, DM etas, DM ReD, DM Red, DM V, DM vx, DM rho,
                                              const double* etai, const double* etasi, const double* emax,
                                              double D0, double d0, int Nz, int Nx, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j >= Nz || i >= Nx) return;
    const int64_t o = j + (int64_t)i * Nz;
    const double r = rho(j, i);
    double e = 1.0 / (1.0 / emax[0] + 1.0 / etai[o]);
    e = 0.5 * e + (1.0 - 0.5) * eta(j, i);
    eta(j, i) = e;
    ReD(j, i) = V(j, i) * D0 / (e / r);
    double s = 1.0 / (1.0 / emax[1] + 1.0 / etasi[o]);
    s = 0.5 * s + (1.0 - 0.5) * etas(j, i);
    etas(j, i) = s;
    Red(j, i) = vx(j, i) * d0 / (s / r);
}

// K7: corner eta (update.jl:411-426, always periodic-x / closed-z as called at
// update.jl:341) and etasw (update.jl:405-407, zBC).  Domain (Nz+1) x (Nx+1).
__global__ __launch_bounds__(NT) void k_corner(DM eta, DM etaco, DM etas, DM etasw, int Nz, int Nx, int zbc, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j > Nz || i > Nx) return;
    const int z0 = gidx(j - 1, Nz, XC_CLOSED), z1 = gidx(j, Nz, XC_CLOSED);
    const int x0 = gidx(i - 1, Nx, XC_PERIODIC), x1 = gidx(i, Nx, XC_PERIODIC);
    const double e00 = eta(z0, x0), e10 = eta(z1, x0), e01 = eta(z0, x1), e11 = eta(z1, x1);
    etaco(j, i) = pow(e00 * e10 * e01 * e11, 0.25);
    if (i < Nx) etasw(j, i) = sqrt(etas(gidx(j - 1, Nz, zbc), i) * etas(gidx(j, Nz, zbc), i));
}

// K8: stresses + Ra (update.jl:432-455).  Domain (Nz+1) x (Nx+1).
__global__ __launch_bounds__(NT) void k_stress(DM etaco, DM exz, DM txz, DM eta, DM exx, DM ezz, DM txx, DM tzz,
                                               DM tII, DM Ra, DM V, DM kx, double D0, int Nz, int Nx, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j > Nz || i > Nx) return;
    txz(j, i) = etaco(j, i) * exz(j, i);
    if (j < Nz && i < Nx) {
        const double e = eta(j, i);
        const double txx_c = e * exx(j, i), tzz_c = e * ezz(j, i);
        txx(j, i) = txx_c;
        tzz(j, i) = tzz_c;
        const double t1 = etaco(j, i) * exz(j, i), t2 = etaco(j + 1, i) * exz(j + 1, i);
        const double t3 = etaco(j, i + 1) * exz(j, i + 1), t4 = etaco(j + 1, i + 1) * exz(j + 1, i + 1);
        const double sq = (t1 * t1 + t2 * t2 + t3 * t3 + t4 * t4) / 4.0;
        tII(j, i) = sqrt(EPS + (txx_c * txx_c + tzz_c * tzz_c) / 2.0 + sq);
        Ra(j, i) = V(j, i) * D0 / kx(j, i);
    }
}

// ---- phsevo! pointwise tail (phsevo.jl:189, 197-211).  Cells.
struct PhsArgs {
    DM advn_X, advn_M, advn_rho, dffn_X, Gx, dXdt, dXdto, dXdtoo, X, Xo, Xoo, M, x, m, rho, bndshape;
    double a1, a2, a3, b1, b2, b3, dt, alpha, G0;
    int Nz, Nx;
    xc_dom d;
};
__global__ __launch_bounds__(NT) void k_phsevo_tail(PhsArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    if (j >= a.Nz || i >= a.Nx) return;
    const double aX = a.advn_X(j, i);
    a.advn_rho(j, i) = aX + a.advn_M(j, i);
    const double Gx = a.G0 * (1.0 - a.x(j, i)) * a.bndshape(j, i);
    a.Gx(j, i) = Gx;
    const double dXdt = -aX + a.dffn_X(j, i) + Gx;
    a.dXdt(j, i) = dXdt;
    double X = a.X(j, i);
    const double res = (a.a1 * X - a.a2 * a.Xo(j, i) - a.a3 * a.Xoo(j, i)) -
                       (a.b1 * dXdt + a.b2 * a.dXdto(j, i) + a.b3 * a.dXdtoo(j, i)) * a.dt;
    const double upd = -a.alpha * res / a.a1;
    X = X + upd;
    const double r = a.rho(j, i);
    X = jclamp(X, r * EPS, r * (1.0 - EPS));
    a.X(j, i) = X;
    const double M = r - X;
    a.M(j, i) = M;
    a.x(j, i) = X / r;
    a.m(j, i) = M / r;
}

// ---- store_previous! (store.jl:13-30).  Domain (Nz+1) x (Nx+1).
struct StoreArgs {
    DM Xoo, Xo, X, Moo, Mo, M, dXdtoo, dXdto, dXdt, rhooo, rhoo, rho, drhodtoo, drhodto, drhodt;
    DM rhoWoo, rhoWo, rhow, W, rhoUoo, rhoUo, rhou, U;
    int Nz, Nx;
    xc_dom d;
};
__global__ __launch_bounds__(NT) void k_store(StoreArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    if (j > a.Nz || i > a.Nx) return;
    if (j < a.Nz && i < a.Nx) {
        a.Xoo(j, i) = a.Xo(j, i); a.Xo(j, i) = a.X(j, i);
        a.Moo(j, i) = a.Mo(j, i); a.Mo(j, i) = a.M(j, i);
        a.dXdtoo(j, i) = a.dXdto(j, i); a.dXdto(j, i) = a.dXdt(j, i);
        a.rhooo(j, i) = a.rhoo(j, i); a.rhoo(j, i) = a.rho(j, i);
        a.drhodtoo(j, i) = a.drhodto(j, i); a.drhodto(j, i) = a.drhodt(j, i);
    }
    if (i < a.Nx) {  // (Nz+1, Nx)
        a.rhoWoo(j, i) = a.rhoWo(j, i);
        a.rhoWo(j, i) = a.rhow(j, i) * a.W(j, i + 1);
    }
    if (j < a.Nz) {  // (Nz, Nx+1)
        a.rhoUoo(j, i) = a.rhoUo(j, i);
        a.rhoUo(j, i) = a.rhou(j, i) * a.U(j + 1, i);
    }
}

// ---- update_segregation_speed! + update_phase_velocities! (store.jl:47-105).
// Domain (Nz+2) x (Nx+2).  The ghost columns of wx are the tapered values of
// the opposite interior columns (store.jl:54-59), recomputed here.
struct PVArgs {
    DM wx, wm, etasw, Drhox, bndtaperw, x_w, m_w, x_u, m_u, W, U, Wx, Wm, Ux, Um;
    DM xisw, xixw, xiew, xisu, xixu, xieu;
    double d02, g0;
    int Nz, Nx, xbc, zbc;
    xc_dom d;
};
__global__ __launch_bounds__(NT) void k_phase_vel(PVArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    const int Nz = a.Nz, Nx = a.Nx;
    if (j >= Nz + 2 || i >= Nx + 2) return;
    if (j <= Nz) {
        const int s = i == 0 ? Nx : (i == Nx + 1 ? 1 : i);  // source column of wx (store.jl:58-59)
        const double wxv = a.d02 / a.etasw(j, s - 1) * a.Drhox(j, s - 1) * a.g0 * a.bndtaperw(j, s);
        a.wx(j, i) = wxv;
        const int ci = gidx(i - 1, Nx, a.xbc);
        const double xw = a.x_w(j, ci), mwm = jmax(a.m_w(j, ci), EPS);
        const double wmv = -xw / mwm * wxv;
        a.wm(j, i) = wmv;
        const double sx = a.xisw(j, i) + a.xixw(j, i);
        a.Wx(j, i) = a.W(j, i) + wxv + sx + a.xiew(j, i);
        a.Wm(j, i) = a.W(j, i) + wmv - xw / mwm * sx + a.xiew(j, i);
    }
    if (i <= Nx) {
        const int zj = gidx(j - 1, Nz, a.zbc);
        const double xu = a.x_u(zj, i), mu = jmax(a.m_u(zj, i), EPS);
        const double sx = a.xisu(j, i) + a.xixu(j, i);
        a.Ux(j, i) = a.U(j, i) + sx + a.xieu(j, i);
        a.Um(j, i) = a.U(j, i) + (-xu) / mu * sx + a.xieu(j, i);
    }
}

// ---- update_dt reductions (update.jl:494-502): max(kx) and max |Wx, Wm, Ux, Um|
__global__ __launch_bounds__(NT) void k_dt_partial(DM kx, DM Wx, DM Wm, DM Ux, DM Um, int Nz, int Nx, xc_dom d,
                                                   double* part) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    double km = -INFINITY, vm = 0.0;
    if (j < Nz && i < Nx) km = kx(j, i);
    if (j <= Nz && i < Nx + 2) vm = fmax(fabs(Wx(j, i)), fabs(Wm(j, i)));
    if (j < Nz + 2 && i <= Nx) vm = fmax(vm, fmax(fabs(Ux(j, i)), fabs(Um(j, i))));
    __shared__ double s1[NT], s2[NT];
    s1[threadIdx.x] = km;
    s2[threadIdx.x] = vm;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + st]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = s1[0];
        part[d.nb + blockIdx.x] = s2[0];
    }
}
__global__ __launch_bounds__(NT) void k_dt_final(const double* part, int nb, double* out) {
    __shared__ double s1[NT], s2[NT];
    double a = -INFINITY, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += NT) {
        a = fmax(a, part[k]);
        b = fmax(b, part[nb + k]);
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + st]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + st]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

// ---------------------------------------------------------------------------
// host side

Par make_par(const xc_grid* g, const xc_par* par, const xc_scales* sc, int xbc, int zbc) {
    Par p{};
    p.rhom0 = par->rhom0; p.rhox0 = par->rhox0; p.d0 = par->d0; p.g0 = par->g0;
    p.etam0 = par->etam0; p.etax0 = par->etax0;
    for (int k = 0; k < 4; ++k) { p.AA[k] = par->AA[k]; p.BB[k] = par->BB[k]; p.CC[k] = par->CC[k]; }
    p.Ptop = par->Ptop; p.etacntr = par->etacntr;
    if (sc) { p.L0 = sc->L0; p.l0 = sc->l0; p.D0 = sc->D0; }
    p.h = g->h; p.invh = 1.0 / g->h;
    p.Nz = (int)g->Nz; p.Nx = (int)g->Nx; p.xbc = xbc; p.zbc = zbc;
    return p;
}

bool shp(const xc_mat& m, int64_t r, int64_t c) { return xc_shape(m, r, c); }
bool shpb(const xc_bmat& m, int64_t r, int64_t c) {
    return m.p && m.rows == r && m.cols == c && (c <= 1 || m.ld >= r);
}

// Shape checks of every PhaseState / FluidState field against state.jl:101-122, 218-259.
int check_fluid(xc_ctx* ctx, const xc_fluid* f, int64_t Nz, int64_t Nx) {
    struct { const xc_mat* m; int64_t r, c; const char* n; } t[] = {
        {&f->W, Nz + 1, Nx + 2, "W"}, {&f->U, Nz + 2, Nx + 1, "U"}, {&f->P, Nz + 2, Nx + 2, "P"},
        {&f->eta, Nz, Nx, "eta"}, {&f->etaco, Nz + 1, Nx + 1, "etaco"}, {&f->rho, Nz, Nx, "rho"},
        {&f->rhow, Nz + 1, Nx, "rhow"}, {&f->rhou, Nz, Nx + 1, "rhou"}, {&f->Drho, Nz + 1, Nx, "Drho"},
        {&f->rhoo, Nz, Nx, "rhoo"}, {&f->rhooo, Nz, Nx, "rhooo"}, {&f->drhodt, Nz, Nx, "drhodt"},
        {&f->drhodto, Nz, Nx, "drhodto"}, {&f->drhodtoo, Nz, Nx, "drhodtoo"}, {&f->MFS, Nz, Nx, "MFS"},
        {&f->rhoWo, Nz + 1, Nx, "rhoWo"}, {&f->rhoWoo, Nz + 1, Nx, "rhoWoo"},
        {&f->rhoUo, Nz, Nx + 1, "rhoUo"}, {&f->rhoUoo, Nz, Nx + 1, "rhoUoo"}};
    for (auto& e : t)
        if (!shp(*e.m, e.r, e.c))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "FluidState.%s must be (%lld, %lld), got (%lld, %lld)", e.n,
                           (long long)e.r, (long long)e.c, (long long)e.m->rows, (long long)e.m->cols);
    return XC_OK;
}

int check_phase(xc_ctx* ctx, const xc_phase* p, int64_t Nz, int64_t Nx) {
    const int64_t C[2] = {Nz, Nx}, Wt[2] = {Nz + 1, Nx + 2}, Ut[2] = {Nz + 2, Nx + 1}, K[2] = {Nz + 1, Nx + 1},
                  Zf[2] = {Nz + 1, Nx}, Xf[2] = {Nz, Nx + 1};
    struct { const xc_mat* m; const int64_t* s; const char* n; } t[] = {
        {&p->x, C, "x"}, {&p->m, C, "m"}, {&p->X, C, "X"}, {&p->M, C, "M"}, {&p->Xo, C, "Xo"}, {&p->Xoo, C, "Xoo"},
        {&p->Mo, C, "Mo"}, {&p->Moo, C, "Moo"}, {&p->chi, C, "chi"}, {&p->mu, C, "mu"},
        {&p->chiw, Wt, "chiw"}, {&p->muw, Wt, "muw"}, {&p->chiu, Ut, "chiu"}, {&p->muu, Ut, "muu"},
        {&p->x_w, Zf, "x_w"}, {&p->m_w, Zf, "m_w"}, {&p->x_u, Xf, "x_u"}, {&p->m_u, Xf, "m_u"},
        {&p->Drhom, Zf, "Drhom"}, {&p->Drhox, Zf, "Drhox"}, {&p->Pl, C, "Pl"}, {&p->Pt, C, "Pt"},
        {&p->etamix, C, "etamix"}, {&p->etae, C, "etae"}, {&p->ke, C, "ke"}, {&p->kx, C, "kx"},
        {&p->etat, C, "etat"}, {&p->etas, C, "etas"}, {&p->etasw, Zf, "etasw"}, {&p->ks, C, "ks"},
        {&p->V, C, "V"}, {&p->Div_V, C, "Div_V"}, {&p->exx, C, "exx"}, {&p->ezz, C, "ezz"}, {&p->exz, K, "exz"},
        {&p->eII, C, "eII"}, {&p->vx, C, "vx"}, {&p->vm, C, "vm"}, {&p->txx, C, "txx"}, {&p->tzz, C, "tzz"},
        {&p->txz, K, "txz"}, {&p->tII, C, "tII"}, {&p->fReL, C, "fReL"}, {&p->ReL, C, "ReL"}, {&p->ReD, C, "ReD"},
        {&p->Ra, C, "Ra"}, {&p->fRel, C, "fRel"}, {&p->Rel, C, "Rel"}, {&p->Red, C, "Red"}, {&p->Rc, C, "Rc"},
        {&p->Wx, Wt, "Wx"}, {&p->Wm, Wt, "Wm"}, {&p->Ux, Ut, "Ux"}, {&p->Um, Ut, "Um"}, {&p->wx, Wt, "wx"},
        {&p->wm, Wt, "wm"}, {&p->bndtaperw, Wt, "bndtaperw"}, {&p->bndshape, C, "bndshape"},
        {&p->advn_X, C, "advn_X"}, {&p->advn_M, C, "advn_M"}, {&p->advn_rho, C, "advn_rho"},
        {&p->qx_advn_X, Ut, "qx_advn_X"}, {&p->qz_advn_X, Wt, "qz_advn_X"}, {&p->qx_advn_M, Ut, "qx_advn_M"},
        {&p->qz_advn_M, Wt, "qz_advn_M"}, {&p->qx_dffn_X, Ut, "qx_dffn_X"}, {&p->qz_dffn_X, Wt, "qz_dffn_X"},
        {&p->qx_dffn_M, Ut, "qx_dffn_M"}, {&p->qz_dffn_M, Wt, "qz_dffn_M"}, {&p->dffn_X, C, "dffn_X"},
        {&p->Gx, C, "Gx"}, {&p->dXdt, C, "dXdt"}, {&p->dXdto, C, "dXdto"}, {&p->dXdtoo, C, "dXdtoo"}};
    for (auto& e : t)
        if (!shp(*e.m, e.s[0], e.s[1]))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "PhaseState.%s must be (%lld, %lld), got (%lld, %lld)", e.n,
                           (long long)e.s[0], (long long)e.s[1], (long long)e.m->rows, (long long)e.m->cols);
    if (!shpb(p->hasx, Nz, Nx) || !shpb(p->hasm, Nz, Nx))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "PhaseState.hasx/hasm must be (Nz, Nx)");
    return XC_OK;
}

int check_grid(xc_ctx* ctx, const xc_grid* g) {
    if (!g || g->Nz < 1 || g->Nx < 1 || g->Nz > (1 << 28) || g->Nx > (1 << 28) || !(g->h > 0))
        return xc_fail(ctx, XC_EINVAL_ARG, "bad grid");
    return XC_OK;
}

int check_bc(xc_ctx* ctx, int xbc, int zbc) {
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "unsupported xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "unsupported zBC = %d", zbc);
    return XC_OK;
}

#define LAUNCH(kern, dom, ...) hipLaunchKernelGGL(kern, dim3((dom).nb), dim3(NT), 0, ctx->stream, __VA_ARGS__)

}  // namespace

extern "C" {

int xc_time_coefs(int tint, int64_t step, double c[6]) {
    static const double be[6] = {1, 1, 0, 1, 0, 0};
    static const double bd2[6] = {1.5, 2.0, -0.5, 1, 0, 0};
    static const double cn[6] = {1, 1, 0, 0.5, 0.5, 0};
    static const double bd2si[6] = {1.5, 2.0, -0.5, 0.75, 0.5, -0.25};
    const double* src = nullptr;
    if (tint == 0 || tint == 1) src = (tint == 0 || step <= 2) ? be : bd2;
    else if (tint == 2 || tint == 3) src = step == 1 ? be : ((tint == 2 || step == 2) ? cn : bd2si);
    else return XC_EINVAL_ARG;
    for (int k = 0; k < 6; ++k) c[k] = src[k];
    return XC_OK;
}

// update! with the lithostatic chain (row sums of rhow, their scan, Drho / Pl / Pt), which nothing else in
// update! reads, on the stream `side` beside the cell kernels (ev_a: its start after k_interp, ev_b: its end,
// which the context's stream waits for before returning); side == nullptr: all on the context's stream
int xc_update_side(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, const xc_grid* g, const xc_par* par,
                   const xc_scales* sc, int xbc, int zbc, hipStream_t side, hipEvent_t ev_a, hipEvent_t ev_b) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_grid(ctx, g)) || (st = check_bc(ctx, xbc, zbc))) return st;
    const int64_t Nz = g->Nz, Nx = g->Nx;
    if ((st = check_fluid(ctx, fl, Nz, Nx)) || (st = check_phase(ctx, ph, Nz, Nx))) return st;
    if (!par || !sc) return xc_fail(ctx, XC_EINVAL_ARG, "update!: par and scales required");
    Par p = make_par(g, par, sc, xbc, zbc);

    // scratch: row partials, row means, Pl column, etai, etasi, block minima, emax
    const int rows = (int)Nz + 1;
    const int nchunk = (int)((Nx + ROW_CHUNK - 1) / ROW_CHUNK);
    xc_dom dK = xc_make_dom(Nz + 1, Nx + 1);
    const size_t n_part = (size_t)nchunk * rows, n_cell = (size_t)(Nz * Nx);
    const size_t total = n_part + rows + Nz + 2 * n_cell + 2 * (size_t)dK.nb + 2;
    double* sc0;
    if ((st = xc_scratch(ctx, XC_SLOT_FIELDS, total * sizeof(double), (void**)&sc0))) return st;
    double* part = sc0;
    double* ref = part + n_part;
    double* plcol = ref + rows;
    double* etai = plcol + Nz;
    double* etasi = etai + n_cell;
    double* bmin = etasi + n_cell;
    double* emax = bmin + 2 * dK.nb;

    InterpArgs ia{dm(ph->x), dm(ph->m), dm(fl->rho), dm(ph->chi), dm(ph->mu), dm(fl->rhow), dm(fl->rhou),
                  dm(ph->Drhom), dm(ph->Drhox), dm(ph->chiw), dm(ph->muw), dm(ph->chiu), dm(ph->muu),
                  dm(ph->x_w), dm(ph->m_w), dm(ph->x_u), dm(ph->m_u), db(ph->hasx), db(ph->hasm), p,
                  xc_make_dom(Nz + 2, Nx + 2)};
    LAUNCH(k_interp, ia.d, ia);
    const hipStream_t ls = side ? side : ctx->stream;
    if (side) {
        XC_TRY_HIP(ctx, hipEventRecord(ev_a, ctx->stream));
        XC_TRY_HIP(ctx, hipStreamWaitEvent(side, ev_a, 0));
    }
    hipLaunchKernelGGL(k_rowsum, dim3((rows + 63) / 64, nchunk), dim3(64), 0, ls, dm(fl->rhow), rows, (int)Nx, part);
    hipLaunchKernelGGL(k_rowmean_scan, dim3(1), dim3(RM_NT), 2 * (((rows + 15) & ~15) + 16) * sizeof(double), ls,
                       part, rows, nchunk, (int)Nx, par->g0, g->h, par->Ptop, ref, plcol);
    xc_dom dZ = xc_make_dom(Nz + 1, Nx);
    hipLaunchKernelGGL(k_drho_pt, dim3(dZ.nb), dim3(NT), 0, ls, dm(fl->rhow), dm(fl->Drho), dm(ph->Pl), dm(ph->Pt),
                       dm(fl->P), (const double*)ref, (const double*)plcol, par->Ptop / 100.0, (int)Nz, (int)Nx, dZ);
    if (side) XC_TRY_HIP(ctx, hipEventRecord(ev_b, side));
    CellArgs ca{dm(ph->chi), dm(ph->mu), dm(ph->etamix), dm(fl->W), dm(fl->U), dm(ph->exx), dm(ph->ezz),
                dm(ph->Div_V), dm(ph->exz), dm(ph->eII), dm(ph->V), dm(ph->wx), dm(ph->wm), dm(ph->vx),
                dm(ph->vm), dm(fl->rho), dm(ph->etae), dm(ph->ke), dm(ph->kx), dm(ph->ReL), dm(ph->fReL),
                dm(ph->Rel), dm(ph->fRel), dm(ph->ks), dm(ph->etat), dm(ph->Rc), etai, etasi, bmin, p, dK};
    LAUNCH(k_cells, dK, ca);
    hipLaunchKernelGGL(k_min_final, dim3(1), dim3(128), 0, ctx->stream, bmin, dK.nb, par->etacntr, emax);
    xc_dom dC = xc_make_dom(Nz, Nx);
    LAUNCH(k_blend, dC, dm(fl->eta), dm(ph->etas), dm(ph->ReD), dm(ph->Red), dm(ph->V), dm(ph->vx), dm(fl->rho),
           etai, etasi, emax, sc->D0, par->d0, (int)Nz, (int)Nx, dC);
    LAUNCH(k_corner, dK, dm(fl->eta), dm(fl->etaco), dm(ph->etas), dm(ph->etasw), (int)Nz, (int)Nx, zbc, dK);
    LAUNCH(k_stress, dK, dm(fl->etaco), dm(ph->exz), dm(ph->txz), dm(fl->eta), dm(ph->exx), dm(ph->ezz),
           dm(ph->txx), dm(ph->tzz), dm(ph->tII), dm(ph->Ra), dm(ph->V), dm(ph->kx), sc->D0, (int)Nz, (int)Nx, dK);
    if (side) XC_TRY_HIP(ctx, hipStreamWaitEvent(ctx->stream, ev_b, 0));
    return xc_finish(ctx, "update!");
}
int xc_update(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_scales* sc,
              int xbc, int zbc) {
    return xc_update_side(ctx, ph, fl, g, par, sc, xbc, zbc, nullptr, nullptr, nullptr);
}

int xc_phsevo(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, const xc_grid* g, const xc_par* par, const xc_scales* sc,
              int scheme, int xbc, int zbc, const double coef[6], double dt, double alpha) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_grid(ctx, g)) || (st = check_bc(ctx, xbc, zbc))) return st;
    const int64_t Nz = g->Nz, Nx = g->Nx;
    if ((st = check_fluid(ctx, fl, Nz, Nx)) || (st = check_phase(ctx, ph, Nz, Nx))) return st;
    if (!coef || !sc) return xc_fail(ctx, XC_EINVAL_ARG, "phsevo!: coef and scales required");
    if (xc_scheme_halo(scheme) < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "phsevo!: unsupported scheme = %d", scheme);
    const unsigned saved = ctx->flags;
    ctx->flags &= ~XC_CTX_SYNC;  // one synchronization at the end, not per stage
    auto sub = [](const xc_mat& a, int64_t r0, int64_t c0, int64_t nr, int64_t nc) {
        xc_mat v{a.p + r0 + c0 * a.ld, nr, nc, a.ld};
        return v;
    };
    // phsevo.jl:177-188: face views of the phase velocities, two WENO advections with fluxes
    xc_mat Wx = sub(ph->Wx, 0, 1, Nz + 1, Nx), Wm = sub(ph->Wm, 0, 1, Nz + 1, Nx);
    xc_mat Ux = sub(ph->Ux, 1, 0, Nz, Nx + 1), Um = sub(ph->Um, 1, 0, Nz, Nx + 1);
    {
        const xc_adv_member pm[2] = {{ph->advn_X, ph->qx_advn_X, ph->qz_advn_X, ph->X, Ux, Wx, g->h},
                                     {ph->advn_M, ph->qx_advn_M, ph->qz_advn_M, ph->M, Um, Wm, g->h}};
        st = xc_advect_centered_pair(ctx, scheme, pm, xbc, zbc, "phsevo!");
    }
    // phsevo.jl:192-194: diffusion of chi with k = X .* kx
    if (!st) st = xc_diffus_centered_kprod(ctx, ph->dffn_X, &ph->qx_dffn_X, &ph->qz_dffn_X, ph->chi, ph->X, ph->kx, g->h, xbc, zbc);
    ctx->flags = saved;
    if (st) return st;
    xc_dom dC = xc_make_dom(Nz, Nx);
    PhsArgs a{dm(ph->advn_X), dm(ph->advn_M), dm(ph->advn_rho), dm(ph->dffn_X), dm(ph->Gx), dm(ph->dXdt),
              dm(ph->dXdto), dm(ph->dXdtoo), dm(ph->X), dm(ph->Xo), dm(ph->Xoo), dm(ph->M), dm(ph->x), dm(ph->m),
              dm(fl->rho), dm(ph->bndshape), coef[0], coef[1], coef[2], coef[3], coef[4], coef[5], dt, alpha,
              sc->G0, (int)Nz, (int)Nx, dC};
    LAUNCH(k_phsevo_tail, dC, a);
    return xc_finish(ctx, "phsevo!");
}

int xc_store_previous(xc_ctx* ctx, xc_fluid* fl, xc_phase* ph) {
    XC_CHECK_CTX(ctx);
    const int64_t Nz = ph->X.rows, Nx = ph->X.cols;
    int st;
    if ((st = check_fluid(ctx, fl, Nz, Nx)) || (st = check_phase(ctx, ph, Nz, Nx))) return st;
    xc_dom d = xc_make_dom(Nz + 1, Nx + 1);
    StoreArgs a{dm(ph->Xoo), dm(ph->Xo), dm(ph->X), dm(ph->Moo), dm(ph->Mo), dm(ph->M), dm(ph->dXdtoo),
                dm(ph->dXdto), dm(ph->dXdt), dm(fl->rhooo), dm(fl->rhoo), dm(fl->rho), dm(fl->drhodtoo),
                dm(fl->drhodto), dm(fl->drhodt), dm(fl->rhoWoo), dm(fl->rhoWo), dm(fl->rhow), dm(fl->W),
                dm(fl->rhoUoo), dm(fl->rhoUo), dm(fl->rhou), dm(fl->U), (int)Nz, (int)Nx, d};
    LAUNCH(k_store, d, a);
    return xc_finish(ctx, "store_previous!");
}

int xc_update_phase_velocities(xc_ctx* ctx, xc_phase* ph, const xc_fluid* fl, const xc_noise* ns,
                               const xc_grid* g, const xc_par* par, int xbc, int zbc) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_grid(ctx, g)) || (st = check_bc(ctx, xbc, zbc))) return st;
    const int64_t Nz = g->Nz, Nx = g->Nx;
    if ((st = check_fluid(ctx, fl, Nz, Nx)) || (st = check_phase(ctx, ph, Nz, Nx))) return st;
    if (!ns || !shp(ns->xisw, Nz + 1, Nx + 2) || !shp(ns->xixw, Nz + 1, Nx + 2) || !shp(ns->xiew, Nz + 1, Nx + 2) ||
        !shp(ns->xisu, Nz + 2, Nx + 1) || !shp(ns->xixu, Nz + 2, Nx + 1) || !shp(ns->xieu, Nz + 2, Nx + 1))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "update_phase_velocities!: NoiseState flux shapes");
    xc_dom d = xc_make_dom(Nz + 2, Nx + 2);
    PVArgs a{dm(ph->wx), dm(ph->wm), dm(ph->etasw), dm(ph->Drhox), dm(ph->bndtaperw), dm(ph->x_w), dm(ph->m_w),
             dm(ph->x_u), dm(ph->m_u), dm(fl->W), dm(fl->U), dm(ph->Wx), dm(ph->Wm), dm(ph->Ux), dm(ph->Um),
             dm(ns->xisw), dm(ns->xixw), dm(ns->xiew), dm(ns->xisu), dm(ns->xixu), dm(ns->xieu),
             par->d0 * par->d0, par->g0, (int)Nz, (int)Nx, xbc, zbc, d};
    LAUNCH(k_phase_vel, d, a);
    return xc_finish(ctx, "update_phase_velocities!");
}

int xc_update_dt(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, const xc_grid* g, const xc_par* par,
                 double dt_prev, double* dt_out) {
    XC_CHECK_CTX(ctx);
    int st;
    if ((st = check_grid(ctx, g))) return st;
    const int64_t Nz = g->Nz, Nx = g->Nx;
    if ((st = check_phase(ctx, ph, Nz, Nx))) return st;
    if (!dt_out) return xc_fail(ctx, XC_EINVAL_ARG, "update_dt: dt_out is NULL");
    xc_dom d = xc_make_dom(Nz + 2, Nx + 2);
    double* part;
    if ((st = xc_scratch(ctx, XC_SLOT_REDUCE, (2 * (size_t)d.nb + 2) * sizeof(double), (void**)&part))) return st;
    LAUNCH(k_dt_partial, d, dm(ph->kx), dm(ph->Wx), dm(ph->Wm), dm(ph->Ux), dm(ph->Um), (int)Nz, (int)Nx, d, part);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(NT), 0, ctx->stream, part, d.nb, part + 2 * d.nb);
    double mx[2];
    XC_TRY_HIP(ctx, hipMemcpyAsync(mx, part + 2 * d.nb, sizeof(mx), hipMemcpyDeviceToHost, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // update.jl:497-501
    const double hh = g->h / 2.0;
    const double dtk = hh * hh / mx[0];
    const double vmax = mx[1] + EPS;
    const double dta = hh / vmax;
    const double a = 1.5 * dt_prev, b = dtk < par->CFL * dta ? dtk : par->CFL * dta;
    const double r = a < b ? a : b;
    *dt_out = r < par->dtmax ? r : par->dtmax;
    return XC_OK;
}

}  // extern "C"
