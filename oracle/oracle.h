/*
 * oracle.h -- CPU restatement of xcore.jl's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing that is measured or shipped.  The product
 * path (xcore.jl_amd/) never calls into it and fails loudly when its HIP
 * library is missing.
 *
 * Every function restates one reference routine with the same operation
 * order (compiled with -ffp-contract=off, so a*b+c is never fused, as in
 * Julia without @fastmath) and cites the file:line it follows.
 *
 * Storage: column-major, z (rows) fastest, element (i,j) at p[i + j*ld],
 * 0-based -- the memory layout of a Julia Matrix{Float64}.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   advection / diffusion   convergence orders of test_advect.jl/test_diffus.jl,
 *                           exactness on linear data, telescoping conservation,
 *                           bitwise face identity, fill_ghosts! 4x4 fixture
 *   time_coefs              exact table of test_run.jl:5-25
 *   rheology                naive restatement of test_update.jl:8-28
 *   kinematics              test_update.jl:49-94 orders and Div V
 *   FFT filter              numpy pocketfft goldens, cosine eigen-mode, constants
 *   Stokes                  MMS order >1.8 (test_mms.jl), scipy splu solves
 */
#ifndef XCORE_ORACLE_H
#define XCORE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double* p; int64_t rows, cols, ld; } omat;

enum { OR_OK = 0, OR_ESHAPE = 1, OR_ESCHEME = 2, OR_EBC = 3, OR_EARG = 6 };
enum { OR_CENTR = 0, OR_UPWD1, OR_QUICK, OR_FROMM, OR_WENO3, OR_WENO5, OR_TVDIM };
enum { OR_PERIODIC = 0, OR_CLOSED = 1 };

int  or_scheme_halo(int scheme);
int  or_fill_ghosts(omat f, int halo, int xbc, int zbc);
int  or_embed_interior(omat fh, omat f, int halo);
int  or_advect(omat adv, omat fh, omat u, omat w, double h, int scheme);
int  or_advect_with_flux(omat adv, omat qx, omat qz, omat fh, omat u, omat w,
                         double h, int scheme, int xbc, int zbc);
int  or_advect_centered(omat adv, const omat* qx, const omat* qz, omat f, omat u, omat w,
                        double h, int scheme, int xbc, int zbc);
int  or_diffus(omat dff, omat fh, omat kh, double h, int halo);
int  or_diffus_with_flux(omat dff, omat qx, omat qz, omat fh, omat kh, double h,
                         int halo, int xbc, int zbc);
int  or_diffus_centered(omat dff, const omat* qx, const omat* qz, omat f, omat k,
                        double h, int xbc, int zbc);


/* ---- state containers: field-by-field mirrors of src/state.jl ------------
 * Shapes (N = Nz x Nx): cell (Nz,Nx); W-type (Nz+1,Nx+2); U-type (Nz+2,Nx+1);
 * P (Nz+2,Nx+2); corner (Nz+1,Nx+1); z-face (Nz+1,Nx); x-face (Nz,Nx+1). */
typedef struct { uint8_t* p; int64_t rows, cols, ld; } obmat;

typedef struct {  /* state.jl:74-93 FluidState (solver cache excluded) */
    omat W, U, P, eta, etaco, rho, rhow, rhou, Drho, rhoo, rhooo,
         drhodt, drhodto, drhodtoo, MFS, rhoWo, rhoWoo, rhoUo, rhoUoo;
} or_fluid;

typedef struct {  /* state.jl:165-209 PhaseState */
    omat x, m, X, M, Xo, Xoo, Mo, Moo, chi, mu;
    obmat hasx, hasm;
    omat chiw, muw, chiu, muu, x_w, m_w, x_u, m_u, Drhom, Drhox, Pl, Pt,
         etamix, etae, ke, kx, etat, etas, etasw, ks, V, Div_V, exx, ezz, exz, eII,
         vx, vm, txx, tzz, txz, tII, fReL, ReL, ReD, Ra, fRel, Rel, Red, Rc,
         Wx, Wm, Ux, Um, wx, wm, bndtaperw, bndshape, advn_X, advn_M, advn_rho,
         qx_advn_X, qz_advn_X, qx_advn_M, qz_advn_M, qx_dffn_X, qz_dffn_X, qx_dffn_M, qz_dffn_M,
         dffn_X, Gx, dXdt, dXdto, dXdtoo;
} or_phase;

typedef struct {  /* noise.jl:38-53 NoiseState (filters are rebuilt from padL0/padl0) */
    omat psie, psix, psis, psieo, psixo, psiso, re, rs;
    int64_t padL0, padl0;
    double fL, fl;
    omat xiew, xieu, xixw, xixu, xisw, xisu, xie, xix, xis;
} or_noise;

typedef struct { int64_t Nz, Nx; double h, D, L; } or_grid;   /* grid.jl:24-43 scalars */

typedef struct {  /* parameters.jl:87-164, the knobs the hot path reads */
    double rhom0, rhox0, d0, g0, etam0, etax0;
    double AA[4], BB[4], CC[4];  /* row-major 2x2: [11, 12, 21, 22] */
    double Xi, Ptop, alpha, gamma, CFL, dtmax, etacntr;
} or_par;

typedef struct { double L0, L0h, l0, l0h, D0, G0, txi0; } or_scales;  /* scales.jl:290-334 subset */

/* update.jl */
int or_update_volume_fractions(or_phase* ph, or_fluid* fl, const or_par* par);
int or_update_interpolations(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, int xbc, int zbc);
int or_update_pressure(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par);
int or_update_rheology(or_phase* ph, const or_par* par);
int or_update_kinematics(or_phase* ph, or_fluid* fl, const or_grid* g);
int or_update_viscosity(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc);
int or_update_segregation_viscosity(or_phase* ph, or_fluid* fl, const or_par* par, const or_scales* sc, int zbc);
int or_update_stresses(or_phase* ph, or_fluid* fl, const or_grid* g, const or_scales* sc);
int or_update(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc, int xbc, int zbc);
int or_compute_corner_eta(omat etaco, omat eta, int xbc, int zbc);
double or_update_dt(const or_phase* ph, const or_fluid* fl, const or_grid* g, const or_par* par, double dt_prev);
/* store.jl */
int or_store_previous(or_fluid* fl, or_phase* ph);
int or_update_segregation_speed(or_phase* ph, const or_grid* g, const or_par* par, int xbc);
int or_update_phase_velocities(or_phase* ph, or_fluid* fl, const or_noise* ns, const or_grid* g,
                               const or_par* par, int xbc, int zbc);
/* phsevo.jl:164-214; coef = {a1, a2, a3, b1, b2, b3} */
int or_phsevo(or_phase* ph, or_fluid* fl, const or_grid* g, const or_par* par, const or_scales* sc,
              int scheme, int xbc, int zbc, const double coef[6], double dt, double alpha);
/* timing.jl:18-36; tint: 0 be1im, 1 bd2im, 2 cn2si, 3 bd2si */
int or_time_coefs(int tint, int64_t step, double coef[6]);

/* noise.jl + the counter-based normal generator that replaces MersenneTwister */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
int  or_normal_fill(omat a, uint64_t seed, uint32_t stream, uint32_t draw);
int  or_fft_filter(omat psi, omat out, int64_t pad, double h, double sigma);
int  or_noise_step(or_noise* ns, const or_phase* ph, const or_grid* g, const or_par* par, const or_scales* sc,
              double dt, int first_iter, uint64_t seed, uint32_t draw);
int  or_store_noise(or_noise* ns);

/* fluidmech.jl:71-534 assembly: options mirror the fluidmech! keywords */
typedef struct {
    int sds, top_cnv, bot_cnv, open_cnv, xbc, zbc, scheme;
    double dt, a1, a2, a3, b1, b2, b3, alpha, gamma, g0;
} or_fm_opts;

/* LL (after pins) as 0-based COO triplets with duplicates, and RR.  bn = NULL, or the six
 * BnchmData arrays {src_W, src_U, src_P, W_exact, U_exact, P_exact} (state.jl:131-134). */
int or_fluidmech_assemble(or_fluid* fl, const or_phase* ph, const or_grid* g, const or_fm_opts* o,
                          const omat* bn, int64_t* rows, int64_t* cols, double* vals, int64_t cap, int64_t* nnz_out,
                          double* RR);

/* threads used by the OpenMP loops (1 = scalar restatement) */
void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
