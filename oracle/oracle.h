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

/* threads used by the OpenMP loops (1 = scalar restatement) */
void or_set_threads(int n);
int  or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
