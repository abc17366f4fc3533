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

/* ---- context (no reference counterpart: replaces the KA backend object,
 *      backend.jl:12-13, as the owner of device state) ------------------- */
int         xc_create(xc_ctx** ctx, int device, unsigned flags);
int         xc_destroy(xc_ctx* ctx);
int         xc_set_stream(xc_ctx* ctx, void* hip_stream);  /* NULL = the context's own stream */
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

#ifdef __cplusplus
}
#endif
#endif
