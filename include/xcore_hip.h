/*
 * xcore_hip.h -- C ABI of libxcore_hip.so, the MI355X (gfx950) implementation
 * of xcore.jl's per-timestep hot path.
 *
 * Every entry point takes plain device pointers wrapped in xc_mat (column-major,
 * z fastest, element (i,j) at p[i + j*ld], 0-based), which is the memory layout
 * of a Julia Matrix{Float64} and of any strided view of one whose first
 * dimension is contiguous (view(qx, 2:Nz+1, :) etc.).  No torch types appear.
 *
 * Ownership: the caller owns every array passed in; the library never frees
 * caller memory.  Scratch buffers, FFT tables and solver hierarchies live in
 * the opaque xc_ctx.  Work is enqueued on the context's HIP stream; with the
 * XC_CTX_SYNC flag every entry point synchronizes before returning, matching
 * the reference's KernelAbstractions.synchronize after each launch.
 *
 * Errors: every function returns an xc_status.  Size mismatches return
 * XC_EINVAL_SHAPE (the reference's AssertionError), unknown schemes and BCs
 * XC_EINVAL_SCHEME / XC_EINVAL_BC (the reference's ArgumentError); the message
 * is available from xc_last_error(ctx).
 */
#ifndef XCORE_HIP_H
#define XCORE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double* p; int64_t rows, cols, ld; } xc_mat;
typedef struct xc_ctx xc_ctx;

typedef enum {
    XC_OK = 0,
    XC_EINVAL_SHAPE = 1,   /* AssertionError in the reference */
    XC_EINVAL_SCHEME = 2,  /* ArgumentError: unsupported scheme */
    XC_EINVAL_BC = 3,      /* ArgumentError: unsupported BC */
    XC_EHIP = 4,           /* HIP runtime error */
    XC_ENOCONV = 5,        /* iterative solve did not reach its tolerance */
    XC_EINVAL_ARG = 6,     /* any other invalid argument */
    XC_ENOMEM = 7
} xc_status;

/* advect.jl:346-355 scheme symbols */
enum { XC_CENTR = 0, XC_UPWD1 = 1, XC_QUICK = 2, XC_FROMM = 3, XC_WENO3 = 4, XC_WENO5 = 5, XC_TVDIM = 6 };
/* ghosts.jl:18-51 BC symbols (:periodic, :closed) */
enum { XC_PERIODIC = 0, XC_CLOSED = 1 };

#define XC_CTX_SYNC 1u   /* synchronize the stream before every entry point returns */
#define XC_CTX_EXACT 2u  /* reference operation order everywhere (bit-identical to a scalar
                            restatement); default: WENO5 weights with one division + FMA (<= 1e-12) */
#define XC_CTX_AUTOTUNE 4u  /* the first advection call at a new (scheme, fluxes, grid) times every tile
                               shape and keeps the fastest (synchronizes once); later calls reuse it */

/* ---- context (no reference counterpart: replaces the KA backend object,
 *      backend.jl:12-13, as the owner of device state) ------------------- */
int         xc_create(xc_ctx** ctx, int device, unsigned flags);
int         xc_destroy(xc_ctx* ctx);
int         xc_set_stream(xc_ctx* ctx, void* hip_stream);  /* enqueue on this stream (NULL = default stream) */
int         xc_use_own_stream(xc_ctx* ctx);                /* back to the stream xc_create made */
void*       xc_get_stream(xc_ctx* ctx);
int         xc_set_flags(xc_ctx* ctx, unsigned flags);
int         xc_sync(xc_ctx* ctx);
const char* xc_last_error(const xc_ctx* ctx);
/* backend.jl:12-13 xcore_zeros: zero-filled device allocation */
int         xc_malloc(xc_ctx* ctx, size_t bytes, void** out);
int         xc_free(xc_ctx* ctx, void* p);

/* ---- halo handling ------------------------------------------------------- */
/* advect.jl:346-355 scheme_halo; returns the halo width or -1 */
int xc_scheme_halo(int scheme);
/* ghosts.jl:18-51 fill_ghosts!(f, halo; zBC, xBC) */
int xc_fill_ghosts(xc_ctx* ctx, xc_mat f_halo, int halo, int xbc, int zbc);
/* ghosts.jl:61-64 embed_interior!(f_halo, f, halo) */
int xc_embed_interior(xc_ctx* ctx, xc_mat f_halo, xc_mat f, int halo);

/* ---- advection (src/kernels/advect.jl) ----------------------------------- */
/* advect.jl:374-425 advect!(adv, f_halo, u, w, h, scheme).  For XC_TVDIM u is
 * (Nz, Nx+3) and w (Nz+3, Nx) as in the reference. */
int xc_advect(xc_ctx* ctx, int scheme, xc_mat adv, xc_mat f_halo, xc_mat u, xc_mat w, double h);
/* advect.jl:717-798 advect_with_flux!(adv, qx, qz, f_halo, u, w, h, scheme; xBC, zBC);
 * qx (Nz+2, Nx+1), qz (Nz+1, Nx+2) with the MATLAB boundary padding. */
int xc_advect_with_flux(xc_ctx* ctx, int scheme, xc_mat adv, xc_mat qx, xc_mat qz,
                        xc_mat f_halo, xc_mat u, xc_mat w, double h, int xbc, int zbc);
/* phsevo.jl:21-59 advect_centered!(adv, f, u, w, h, scheme; xBC, zBC) when qx == qz == NULL,
 * phsevo.jl:90-128 advect_centered_with_flux! otherwise.  No halo buffer is
 * built: the BCs become an index map inside the kernel (needs Nz, Nx >= halo). */
int xc_advect_centered(xc_ctx* ctx, int scheme, xc_mat adv, const xc_mat* qx, const xc_mat* qz,
                       xc_mat f, xc_mat u, xc_mat w, double h, int xbc, int zbc);

/* ---- diffusion (src/kernels/diffus.jl) ----------------------------------- */
/* diffus.jl:45-57 diffus!(dff, f_halo, k_halo, h; halo) */
int xc_diffus(xc_ctx* ctx, xc_mat dff, xc_mat f_halo, xc_mat k_halo, double h, int halo);
/* diffus.jl:105-147 diffus_with_flux!(dff, qx, qz, f_halo, k_halo, h; halo, xBC, zBC) */
int xc_diffus_with_flux(xc_ctx* ctx, xc_mat dff, xc_mat qx, xc_mat qz, xc_mat f_halo, xc_mat k_halo,
                        double h, int halo, int xbc, int zbc);
/* phsevo.jl:67-79 diffus_centered! (qx == qz == NULL) and 135-149 diffus_centered_with_flux! */
int xc_diffus_centered(xc_ctx* ctx, xc_mat dff, const xc_mat* qx, const xc_mat* qz,
                       xc_mat f, xc_mat k, double h, int xbc, int zbc);


/* ---- state containers: field-by-field mirrors of src/state.jl -------------
 * Shapes (Nz x Nx grid): cell (Nz,Nx); W-type (Nz+1,Nx+2); U-type (Nz+2,Nx+1);
 * P (Nz+2,Nx+2); corner (Nz+1,Nx+1); z-face (Nz+1,Nx); x-face (Nz,Nx+1). */
typedef struct { uint8_t* p; int64_t rows, cols, ld; } xc_bmat;   /* Matrix{Bool} */

typedef struct {  /* state.jl:74-93 FluidState (the host-side solver cache is not part of it) */
    xc_mat W, U, P, eta, etaco, rho, rhow, rhou, Drho, rhoo, rhooo,
           drhodt, drhodto, drhodtoo, MFS, rhoWo, rhoWoo, rhoUo, rhoUoo;
} xc_fluid;

typedef struct {  /* state.jl:165-209 PhaseState */
    xc_mat x, m, X, M, Xo, Xoo, Mo, Moo, chi, mu;
    xc_bmat hasx, hasm;
    xc_mat chiw, muw, chiu, muu, x_w, m_w, x_u, m_u, Drhom, Drhox, Pl, Pt,
           etamix, etae, ke, kx, etat, etas, etasw, ks, V, Div_V, exx, ezz, exz, eII,
           vx, vm, txx, tzz, txz, tII, fReL, ReL, ReD, Ra, fRel, Rel, Red, Rc,
           Wx, Wm, Ux, Um, wx, wm, bndtaperw, bndshape, advn_X, advn_M, advn_rho,
           qx_advn_X, qz_advn_X, qx_advn_M, qz_advn_M, qx_dffn_X, qz_dffn_X, qx_dffn_M, qz_dffn_M,
           dffn_X, Gx, dXdt, dXdto, dXdtoo;
} xc_phase;

typedef struct {  /* noise.jl:38-53 NoiseState; the Gaussian filters and the RNG live in xc_ctx */
    xc_mat psie, psix, psis, psieo, psixo, psiso, re, rs;
    int64_t padL0, padl0;
    double fL, fl;
    xc_mat xiew, xieu, xixw, xixu, xisw, xisu, xie, xix, xis;
    uint64_t seed;   /* the NoiseState's RNG (noise.jl:91): Philox4x32-10 key ... */
    uint64_t draws;  /* ... and its counter, advanced by one per white-noise draw */
} xc_noise;

typedef struct { int64_t Nz, Nx; double h, D, L; } xc_grid;  /* grid.jl:24-43 (scalars) */

typedef struct {  /* parameters.jl:87-164: the knobs the hot path reads */
    double rhom0, rhox0, d0, g0, etam0, etax0;
    double AA[4], BB[4], CC[4];  /* row-major 2x2: [11, 12, 21, 22] */
    double Xi, Ptop, alpha, gamma, CFL, dtmax, etacntr;
    /* read by initialize! only (run.jl:71-140) */
    double x0, Da, dxr, dxg;
    uint64_t seed;
    int open_sgr;
} xc_par;

typedef struct {  /* scales.jl:290-334 subset */
    double L0, L0h, l0, l0h, D0, G0, txi0;
    double bnd_w, dt0;  /* read by initialize! only */
} xc_scales;

/* ---- characteristic scales (src/scales.jl) --------------------------------- */
typedef struct {  /* the scalars scales.jl:88's _compute_scales_inner reads: Parameters + the grid spacing */
    double D, d0, L0, l0, x0, Da, rhom0, rhox0, etam0, g0, Xi, h;
} xc_scales_in;
typedef struct {  /* scales.jl:22-62 Scales{Float64}: every field, in order */
    double h0, D0, d0, L0, L0h, l0, l0h, bnd_w, rho0, Drho0, chi0, Dchi0, eta0, W0, w0, W0l, w0l, W0t, w0t, W0i,
           Ri0, ReL0, Rel0, fReL0, fRel0, eII0, ke0, ks0, kx0, tW0, tw0, tk0, ti0, txi0, t0, dt0, taue0, taus0,
           St0, xie0, xix0, xis0, G0, etae0, etat0, p0, Noe0, Nox0, Nos0, Rc0, Ra0, ReD0, Red0;
} xc_scales_full;
/* scales.jl:85-185 compute_scales(par, grid; tol, maxiter) for n parameter sets (ensemble members) at once;
 * host only (no context, no device work) */
int xc_compute_scales(int n, const xc_scales_in* in, double tol, int maxiter, xc_scales_full* out);

/* ---- ensembles: many grids of any sizes advance in one launch (no counterpart in the reference;
 * each member is exactly the single-grid advect_centered[_with_flux]! / diffus_centered[_with_flux]!
 * of phsevo.jl:21-149 on its own arrays).  qx.p == NULL on every member: no fluxes. */
typedef struct { xc_mat adv, qx, qz, f, u, w; double h; } xc_adv_member;
typedef struct { xc_mat dff, qx, qz, f, k; double h; } xc_dif_member;
int xc_advect_centered_batch(xc_ctx* ctx, int scheme, int n, const xc_adv_member* members, int xbc, int zbc);
int xc_diffus_centered_batch(xc_ctx* ctx, int n, const xc_dif_member* members, int xbc, int zbc);
/* Reusable form: the plan keeps the members' argument table on the device, so xc_batch_run only
 * launches one kernel -- the form to record in a hipGraph and replay every step (the members'
 * arrays must stay where they are).  The arithmetic mode (XC_CTX_EXACT) is read at run time. */
typedef struct xc_batch_plan xc_batch_plan;
int xc_advect_batch_plan(xc_ctx* ctx, int scheme, int n, const xc_adv_member* members, int xbc, int zbc,
                         xc_batch_plan** plan);
int xc_diffus_batch_plan(xc_ctx* ctx, int n, const xc_dif_member* members, int xbc, int zbc, xc_batch_plan** plan);
int xc_batch_run(xc_ctx* ctx, const xc_batch_plan* plan);
int xc_batch_destroy(xc_ctx* ctx, xc_batch_plan* plan);  /* synchronizes the context's stream */

/* Tile shape the advection kernels use at this (scheme, fluxes, grid): height tz (z) x width tx (tx < 0:
 * the persistent, software-pipelined 64 x |tx| kernel that streams each next halo tile into LDS by DMA),
 * and the measured time per launch of every candidate -- 64x16, 128x8, 64x8, 32x32, 64x16 pipelined,
 * 64x8 pipelined -- in us[6] (if non-NULL).  No counterpart in the reference (the autotuner of the tile
 * shapes).  XC_ENOCONV: not tuned yet (the default shape). */
int xc_advect_tile(xc_ctx* ctx, int scheme, int with_flux, int64_t Nz, int64_t Nx, int* tz, int* tx, float* us);
/* pin the tile shape for (scheme, fluxes, grid), e.g. a choice tuned in an earlier run: 64x16, 128x8, 64x8,
 * 32x32, or 64x-16 / 64x-8 for the pipelined kernels (not for tvdim) */
int xc_advect_set_tile(xc_ctx* ctx, int scheme, int with_flux, int64_t Nz, int64_t Nx, int tz, int tx);

/* Diagnostic, no counterpart in the reference: the memory traffic of a stencil launch without its
 * arithmetic.  out[k](i + (k > 0), j) = sum_r in[r](i, j) for every cell (i, j) of out[0]'s shape
 * (0..3 inputs, 1..3 outputs; outputs k >= 1 one row down, as the padded face fluxes), written with
 * non-temporal stores when nts bit 0 is set -- the bandwidth HBM sustains for that mix of streams, the
 * achievable bound a stencil kernel with the same traffic is measured against.  nts bit 1: the best case of
 * the same traffic, every stream aligned and not shifted, 16-byte accesses (needs even rows and 16-byte
 * aligned columns).  nts bit 2: non-temporal loads. */
int xc_diag_stream(xc_ctx* ctx, int nr, const xc_mat* in, int nw, const xc_mat* out, int nts);

/* ---- field updates -------------------------------------------------------- */
/* phsevo.jl:164-214 phsevo!(phase, fluid, grid, par, scales; ADVN, xBC, zBC, a1..b3, dt, alpha);
 * coef = {a1, a2, a3, b1, b2, b3} from time_coefs (timing.jl:18-36) */
int xc_phsevo(xc_ctx* ctx, xc_phase* phase, xc_fluid* fluid, const xc_grid* grid, const xc_par* par,
              const xc_scales* scales, int scheme, int xbc, int zbc, const double coef[6], double dt,
              double alpha);
/* update.jl:525-538 update!(phase, fluid, grid, par, scales; xBC, zBC) */
int xc_update(xc_ctx* ctx, xc_phase* phase, xc_fluid* fluid, const xc_grid* grid, const xc_par* par,
              const xc_scales* scales, int xbc, int zbc);
/* update.jl:494-502 update_dt(phase, fluid, grid, par, dt_prev): synchronizes, writes *dt_out */
int xc_update_dt(xc_ctx* ctx, const xc_phase* phase, const xc_fluid* fluid, const xc_grid* grid,
                 const xc_par* par, double dt_prev, double* dt_out);
/* store.jl:13-30 store_previous!(fluid, phase) */
int xc_store_previous(xc_ctx* ctx, xc_fluid* fluid, xc_phase* phase);
/* store.jl:80-105 update_phase_velocities!(phase, fluid, ns, grid, par; xBC, zBC),
 * including update_segregation_speed! (store.jl:47-68) */
int xc_update_phase_velocities(xc_ctx* ctx, xc_phase* phase, const xc_fluid* fluid, const xc_noise* ns,
                               const xc_grid* grid, const xc_par* par, int xbc, int zbc);
/* noise.jl:138-239 noise!(ns, phase, grid, par, scales, dt; first_iter).  When first_iter
 * is nonzero fresh white noise re, rs is drawn (Philox4x32-10 + Box-Muller, keyed by
 * ns->seed, counter ns->draws, which is then incremented).  The Gaussian filter runs as
 * separable batched mixed-radix FFTs in LDS (line lengths Nx and Nz+pad up to 2048). */
int xc_noise_step(xc_ctx* ctx, xc_noise* ns, const xc_phase* phase, const xc_grid* grid, const xc_par* par,
                  const xc_scales* scales, double dt, int first_iter);
/* noise.jl:115-120 store_noise!(ns) */
int xc_store_noise(xc_ctx* ctx, xc_noise* ns);
/* the white-noise generator alone: standard normals into a (column-major) matrix */
int xc_normal_fill(xc_ctx* ctx, xc_mat a, uint64_t seed, uint32_t stream, uint32_t draw);
/* noise.jl:244-270 compute_fft_filter of one field (psi -> out, out != psi) */
int xc_fft_filter(xc_ctx* ctx, xc_mat psi, xc_mat out, int64_t pad, double h, double sigma);

/* ---- Stokes (src/fluidmech.jl) ------------------------------------------- */
typedef struct {  /* fluidmech! keywords (fluidmech.jl:71-80) and the iterative-solver controls */
    int sds, top_cnv, bot_cnv, open_cnv, xbc, zbc;
    int scheme;       /* par.ADVN, used for the momentum advection on the right-hand side */
    double dt, a1, a2, a3, b1, b2, b3;
    double tol;       /* stop at ||RR - LL x||_2 <= tol ||RR||_2 (1e-8) */
    int maxit;        /* FGMRES iteration cap (all restarts) */
    int restart;      /* FGMRES restart length */
    int mg_pre, mg_post, mg_coarse;  /* multigrid smoothing sweeps on the finest level (pre, post) and on
                                        the coarsest; 0 selects the defaults 3, 3 and, on the coarsest,
                                        20 at the 4 x 4 bottom or 2 where the cycle stopped early */
    double mg_omega;  /* damped-Jacobi weight of the multigrid smoother (0: per level from its inertia ratio
                         R = a1 rho h^2 / (dt eta): 0.95 where R >= 1, 0.8 elsewhere and without inertia); the
                         finest level's Chebyshev passes (mg_jacobi) set their own weights */
    int learned;      /* the learned correction (needs xc_unet_set_params): 1 the network's coarse estimate e of
                         LL^-1 v first, z = e + M (v - LL e); 2 the network's estimate of what the plain
                         preconditioner M misses, z = M v + P N(R v).  A learned step is kept while it cuts the
                         residual at least as much as plain steps do per unit of cost; otherwise the rest of the
                         solve is plain (same Krylov basis; fallback_used = 1), and after 3 solves in a row
                         dropped at their first step the next 16 solves are plain (32, 64, ... up to 1024 each
                         time the solve that retries it is dropped again) */
    int lc_level;     /* grid of the learned correction: Nz / 2^lc_level x Nx / 2^lc_level cells */
    int mg_sub;       /* pre- and post-smoothing sweeps on the levels between (0: default 1; < 0: as the
                         finest level) */
    int mg_levels;    /* most multigrid levels, the finest included (0: chosen per solve from the inertia ratio,
                         see xc_solve_stats) */
    int mg_fp64;      /* 1: the finest level's smoother in double precision; 0 (default): in single precision
                         (its vectors and a copy of its coefficients; FGMRES, the operator and the coarser
                         levels stay in double precision, and so does the convergence test) */
    int xmean;        /* horizontal-mean correction (outside the benchmark mode): every FGMRES cycle's starting
                         guess is first corrected exactly on the fields that do not vary in x,
                         x += R (R^T LL R)^-1 R^T (RR - LL x) with R the x-uniform extension of one value per row
                         of W, U and P (the Krylov method converges slowest on that subspace).  -1: off; 1: probe
                         R^T LL R in this solve (9 operator applications); 2: reuse the operator the last probing
                         solve on this grid left (probe if none); 0 (default): off in xc_fluidmech, and in
                         xc_run_step probed in each time step's first Picard sweep and reused in the others */
    int mg_jacobi;    /* 1: the finest level's pre- and post-smoothing are damped Jacobi sweeps (mg_omega) in every
                         case; 0 (default): with the single-precision fine smoother (mg_fp64 = 0), Chebyshev passes
                         (the same sweeps, each adding a multiple of the last update) where the finest level's
                         inertia ratio is >= 1 and it is not the only level, damped Jacobi elsewhere */
} xc_fm_opts;

typedef struct {  /* state.jl:131-134 BnchmData */
    xc_mat src_W, src_U, src_P, W_exact, U_exact, P_exact;
} xc_bnchm;

typedef struct {
    int iters;             /* FGMRES iterations used */
    int converged;         /* 1 when the residual target was met */
    double rel_resid;      /* final ||RR - LL x|| / ||RR|| */
    double init_rel_resid; /* the same for the initial guess (previous W, U, P) */
    int mg_levels;         /* multigrid levels the velocity preconditioner's V-cycle used */
    int fallback_used;     /* 1 if the learned correction was dropped during the solve or skipped */
    double inertia_ratio;  /* a1 min(rho) h^2 / (dt max(eta)) on the finest level (0 without inertia or when
                              mg_levels is set): the
                              V-cycle stops at the first level l >= 1 with 4^l times it >= 6 (mg_levels == 0),
                              where two damped-Jacobi sweeps cut the smoothest error mode 4x */
    int learned_steps;     /* FGMRES steps preconditioned with the learned correction */
} xc_solve_stats;

/* fluidmech.jl:71-534 fluidmech!(state, grid, par; phase, bnchm_data, sds, top_cnv, bot_cnv, open_cnv,
 * xBC, zBC, dt, a1..b3).  Builds the same saddle-point system LL x = RR (MFS update, momentum
 * advection, BC rows, pins) matrix-free and solves it by right-preconditioned FGMRES with a
 * block-triangular preconditioner (geometric multigrid on the velocity block, rho/eta Schur
 * diagonal), starting from the current (W, U, P).  phase == NULL with bnchm != NULL is the MMS path.
 * Returns XC_ENOCONV (state holds the best iterate) if tol is not met within maxit. */
int xc_fluidmech(xc_ctx* ctx, xc_fluid* fluid, const xc_phase* phase, const xc_grid* grid, const xc_par* par,
                 const xc_fm_opts* opts, const xc_bnchm* bnchm, xc_solve_stats* stats);
/* y = LL x for stacked device vectors [vec(W); vec(U); vec(P)] (no reference counterpart: exposes
 * the matrix-free operator of fluidmech.jl:145-488 for checking against an assembled matrix) */
int xc_stokes_apply(xc_ctx* ctx, const xc_fluid* fluid, const xc_grid* grid, const xc_par* par,
                    const xc_fm_opts* opts, int bnchm, const double* x, double* y);
/* RR of fluidmech.jl (after its MFS update when phase != NULL) into a stacked device vector */
int xc_stokes_rhs(xc_ctx* ctx, xc_fluid* fluid, const xc_phase* phase, const xc_grid* grid, const xc_par* par,
                  const xc_fm_opts* opts, const xc_bnchm* bnchm, double* rr);
/* No reference counterpart (the solver pieces behind fluidmech!, exposed to generate the learned
 * correction's training data and to test it): z = M^-1 v with the plain block preconditioner;
 * FGMRES on LL x = rhs from the given x (opts->learned selects the preconditioner); and the
 * learned correction's input map: block averages over 2^level x 2^level cells of the cell-centred
 * (W, U, P) of vec plus log(eta / etam0) into out4 ([p][4] FP32, p = x * Hc + z), channels 1..3
 * divided by sigma = their RMS (computed when compute_sigma, else read from *sigma; device FP32). */
int xc_stokes_precond(xc_ctx* ctx, const xc_fluid* fluid, const xc_grid* grid, const xc_par* par,
                      const xc_fm_opts* opts, int bnchm, const double* v, double* z);
int xc_stokes_solve(xc_ctx* ctx, const xc_fluid* fluid, const xc_grid* grid, const xc_par* par,
                    const xc_fm_opts* opts, int bnchm, const double* rhs, double* x, xc_solve_stats* stats);
int xc_stokes_lc_restrict(xc_ctx* ctx, const xc_fluid* fluid, const xc_grid* grid, const xc_par* par, int level,
                          const double* vec, float* out4, float* sigma, int compute_sigma);
/* basis vector j of the last FGMRES cycle (unit norm) into out (the training data of the learned
 * correction are the directions FGMRES actually feeds its preconditioner) */
int xc_stokes_krylov_vector(xc_ctx* ctx, int j, double* out);
/* z += P e: the learned correction's prolongation of coarse cell values e3 ([p][3], times *sigma and
 * the output scales) to the stacked vector z (bilinear to fine cells, averaged onto interior faces) */
int xc_stokes_lc_prolong(xc_ctx* ctx, const xc_fluid* fluid, const xc_grid* grid, int level, const float* e3,
                         const float* sigma, double* z);

/* ---- time loop (src/run.jl, src/diagnose.jl, src/history.jl) ------------ */
typedef struct {  /* run! keywords (run.jl:167-181) and Picard controls (parameters.jl:150-154) */
    int tint;          /* TINT: 0 :be1im, 1 :bd2im, 2 :cn2si, 3 :bd2si */
    int scheme;        /* ADVN */
    int xbc, zbc;
    double rtol, atol; /* Picard stopping (par.rtol, par.atol) */
    int maxit;         /* par.maxit */
    int fixed_sweeps;  /* > 0: exactly this many Picard sweeps (for reproducible comparisons) */
    xc_fm_opts fm;     /* sds, top_cnv, bot_cnv, open_cnv + Stokes solver controls (dt, a*, b* are set per step) */
} xc_step_opts;

typedef struct { double time, dt; int64_t step; } xc_step_state;

typedef struct {
    int sweeps;                     /* Picard sweeps taken */
    double resnorm;                 /* last Picard residual */
    int stokes_iters;               /* FGMRES iterations over all sweeps */
    double stokes_worst_rel_resid;  /* largest final Stokes relative residual */
    int stokes_fallbacks;           /* sweeps where the learned correction was rejected */
    int stokes_mg_levels;           /* V-cycle depth of the last Stokes solve */
    double stokes_inertia_ratio;    /* its inertia ratio (xc_solve_stats.inertia_ratio) */
    int stokes_iters_first;         /* FGMRES iterations of the first sweep's solve */
    double stokes_init_rel_first;   /* its initial relative residual (the previous step's solution) */
    double stokes_init_rel_later;   /* the largest initial relative residual of the later sweeps */
    int stokes_learned_steps;       /* FGMRES steps preconditioned with the learned correction */
} xc_step_stats;

/* run.jl:204-275: one outer step of run! -- time_coefs, store_previous!, store_noise!, then the
 * Picard loop (>= 3 sweeps of snapshot!, phsevo!, fluidmech!, noise!, update_phase_velocities!,
 * update!, compute_resnorm), time += dt, dt = update_dt(...).  noise! and update!'s lithostatic chain run on
 * a second stream the context owns (default priority), ordered after the work queued before the call and before
 * all work queued after it on the context's stream; the results are those of the entry points called one after
 * another.  state->step/time/dt advance only when
 * the step succeeds.  A failed step (XC_ENOCONV from a Stokes solve or a NaN Picard residual, or any
 * other error) returns with state unchanged but the fields part-way through the step: they are not a
 * state of the run, which has to resume from a checkpoint (the reference raises at the same point,
 * diagnose.jl:80). */
int xc_run_step(xc_ctx* ctx, xc_phase* phase, xc_fluid* fluid, xc_noise* ns, const xc_grid* grid,
                const xc_par* par, const xc_scales* scales, const xc_step_opts* opts, xc_step_state* state,
                xc_step_stats* stats);
/* Ensembles (no reference counterpart): n independent runs, each with its own context (stream, solver
 * hierarchy), grid, parameters, scales, noise seed and state, advance one step of run! each (as
 * xc_run_step) with their work running concurrently on the device: nthreads host threads take members in
 * turn (nthreads <= 0: 4, one per hardware queue HIP uses by default; 64 members of 128^2-384^2 took 210-215 ms
 * per step with 4, 8 or 16 threads and 297 ms with 32).  Each member's results equal its step run alone, bit for bit.  status: the member's xc_status; the
 * function returns the first member's error (its message in that member's context), else XC_OK. */
typedef struct {
    xc_ctx* ctx;
    xc_phase* phase;
    xc_fluid* fluid;
    xc_noise* ns;
    const xc_grid* grid;
    const xc_par* par;
    const xc_scales* scales;
    const xc_step_opts* opts;
    xc_step_state* state;
    xc_step_stats* stats;
    int status;
} xc_ensemble_member;
int xc_ensemble_run_step(int n, xc_ensemble_member* members, int nthreads);
/* diagnose.jl:42-49 snapshot!(res, phase, fluid): the context holds one snapshot (of the grid last
 * passed; the reference's StepResidual, diagnose.jl:13-28) */
int xc_snapshot(xc_ctx* ctx, const xc_phase* phase, const xc_fluid* fluid);
/* diagnose.jl:57-71 compute_resnorm(res, phase, fluid, dt) -> out = {resnorm, res_mass, res_mmnt}
 * against the context's snapshot; XC_EINVAL_ARG when no snapshot was taken or it is of another grid
 * size (synchronizes) */
int xc_compute_resnorm(xc_ctx* ctx, const xc_phase* phase, const xc_fluid* fluid, double dt, double out[3]);
/* history.jl:122-151 raw sums: {sum rho, sum M, sum X, sum Gx, top/bottom line sums of qz_advn_X,
 * qz_advn_M, qz_dffn_X, qz_dffn_M over the interior columns} (12 values, synchronizes) */
int xc_history_sums(xc_ctx* ctx, const xc_phase* phase, const xc_fluid* fluid, double out[12]);
/* history.jl:101-222 record_history!'s reductions (synchronizes): out[0..11] as xc_history_sums, then
 * for each field f of {x, V, vx, vm, ns.xie, ns.xix, ns.xis, Ra, Rc, ReD, Red, ks, ke, kx, eta, etamix,
 * etae, etas, etat} (history.jl:184-221) the 7 values out[12 + 7 f + q]: q = 0 min, 1 max, 2 sum,
 * 3 sum of squares, 4 sum of (a - a[0]), 5 sum of (a - a[0])^2 (mean and n-1 std without
 * cancellation), 6 sum of log(max(a, eps)) (the geometric mean).  The host turns them into the
 * History record (min/mean/max/std, min/rms/max, min/geomean/max). */
#define XC_HIST_NSTATS 145
int xc_history_stats(xc_ctx* ctx, const xc_phase* phase, const xc_fluid* fluid, const xc_noise* ns,
                     double out[XC_HIST_NSTATS]);
/* No counterpart in the reference: xc_history_stats for callers that cannot allocate (code under HIP graph
 * capture, tight loops, harnesses that forbid hipMalloc).  Enqueues on `stream` (a hipStream_t) the same
 * reductions into out_dev (XC_HIST_NSTATS doubles of device memory) and returns without synchronizing; no
 * context, no allocation: the partial sums live in a fixed 114.5 KiB device array of the library (one per
 * device, held until the process ends), whatever the grid.  Grids of more than 536,870,911 cells are
 * refused (XC_EINVAL_SHAPE).  Calls that overlap on one device race on that array. */
int xc_history_stats_noalloc(void* stream, const xc_phase* phase, const xc_fluid* fluid, const xc_noise* ns,
                             double* out_dev);
/* run.jl:71-140 initialize!(phase, fluid, ns, hst, grid, par, scales; xBC, zBC), fresh start: the
 * boundary taper and shape (run.jl:24-55), x = xin (1 + dxr rp + dxg gp) with rp the normalised white
 * noise (the library's Philox4x32-10 generator, seed par->seed, stream 2, in place of Julia's
 * MersenneTwister) -- or x = *x_init when x_init != NULL --, m = 1 - x, eta = etaco = etas = etam0,
 * update!, X = rho x, M = rho m, noise!(dt0, first_iter), update_phase_velocities!, Xo = Xoo = X,
 * Mo = Moo = M, rhoo = rhooo = rho, store_noise!.  The t = 0 history record is the caller's
 * (xc_history_stats), as is a restart (load the checkpoint instead). */
int xc_initialize(xc_ctx* ctx, xc_phase* phase, xc_fluid* fluid, xc_noise* ns, const xc_grid* grid,
                  const xc_par* par, const xc_scales* scales, int xbc, int zbc, const xc_mat* x_init);

/* ---- learned coarse correction (no counterpart in the reference: the learned Stokes
 * preconditioner).  U-Net, 4 levels, channels 16/32/48/64, BF16 weights and activations with
 * FP32 accumulation.  Tensors are device FP32, pixel-major, pixel p = (b*W + x)*H + z:
 * inputs [p][4] = (log-viscosity, r_W, r_U, r_P), outputs [p][3] = (e_W, e_U, e_P); H, W
 * multiples of 8.  Parameters: one flat FP32 vector (layer order and layouts in xc_unet.hip). */
int xc_unet_num_params(int64_t* n);
int xc_unet_set_params(xc_ctx* ctx, const float* host_params, int64_t n);  /* also resets Adam */
int xc_unet_get_params(xc_ctx* ctx, float* host_params, int64_t n);
int xc_unet_forward(xc_ctx* ctx, const float* in, int batch, int H, int W, float* out);
/* fixed per-channel scales of the trained network: inputs are multiplied by in_scale, outputs by out_scale */
int xc_unet_set_io_scales(xc_ctx* ctx, const float in_scale[4], const float out_scale[3]);
/* loss = 0.5 mean_pixels sum_channels (out - tgt)^2 (to *loss on the host if non-NULL, which
 * synchronizes) and its gradient (to the device vector grad if non-NULL) */
int xc_unet_loss_grad(xc_ctx* ctx, const float* in, const float* tgt, int batch, int H, int W,
                      double* loss, float* grad);
/* loss_grad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8) with learning rate lr */
int xc_unet_train_step(xc_ctx* ctx, const float* in, const float* tgt, int batch, int H, int W,
                       float lr, double* loss);

/* timing.jl:18-36 time_coefs(T, TINT, step); tint 0 :be1im, 1 :bd2im, 2 :cn2si, 3 :bd2si */
int xc_time_coefs(int tint, int64_t step, double coef_out[6]);

#ifdef __cplusplus
}
#endif
#endif
