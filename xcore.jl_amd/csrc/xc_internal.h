// xc_internal.h -- shared definitions of libxcore_hip (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "xcore_hip.h"

#define XC_MAX_SCRATCH 32

struct xc_ctx {
    int device;
    unsigned flags;
    hipStream_t own_stream;   // created by xc_create
    hipStream_t stream;       // where work is enqueued (own_stream or the caller's)
    char err[1024];
    void* scratch[XC_MAX_SCRATCH];
    size_t scratch_bytes[XC_MAX_SCRATCH];
    void* ext[8];             // per-subsystem state (FFT plans, solver hierarchies, ...)
    void (*ext_free[8])(xc_ctx*, void*);  // releases ext[i] in xc_destroy
};

enum { XC_EXT_NOISE = 0, XC_EXT_STOKES = 1, XC_EXT_UNET = 2, XC_EXT_STEP = 3, XC_EXT_TUNE = 4, XC_EXT_BATCH = 5 };

// Scratch slots: one growing device buffer per slot, owned by the context.
enum {
    XC_SLOT_REDUCE = 0,
    XC_SLOT_HOSTRED = 1,
    XC_SLOT_FIELDS = 2,
    XC_SLOT_NOISE = 8,
    XC_SLOT_STOKES = 16,
};

int xc_fail(xc_ctx* ctx, int code, const char* fmt, ...);
// After enqueueing kernels: report launch errors; synchronize in XC_CTX_SYNC mode.
int xc_finish(xc_ctx* ctx, const char* what);
// two centred advections of one scheme in one launch (both with fluxes or both without), by value: the results
// of two xc_advect_centered calls; name: the caller, for error messages
extern "C" int xc_advect_centered_pair(xc_ctx* ctx, int scheme, const xc_adv_member* m, int xbc, int zbc,
                                       const char* name);
// update! with its lithostatic chain on a second stream (xc_fields.hip; side == nullptr: xc_update)
extern "C" int xc_update_side(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, const xc_grid* g, const xc_par* par,
                              const xc_scales* sc, int xbc, int zbc, hipStream_t side, hipEvent_t ev_a,
                              hipEvent_t ev_b);
// Device scratch of at least `bytes` in `slot` (contents undefined).
int xc_scratch(xc_ctx* ctx, int slot, size_t bytes, void** out);

#define XC_TRY_HIP(ctx, call)                                                              \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) return xc_fail((ctx), XC_EHIP, "%s: %s", #call, hipGetErrorString(e_)); \
    } while (0)

#define XC_CHECK_CTX(ctx)                                                                  \
    do {                                                                                   \
        if (!(ctx)) return XC_EINVAL_ARG;                                                  \
    } while (0)

// Diffusion with diffusivity k1 .* k2 (phsevo.jl:192-194, Xkx = X .* kx) -- product formed per cell.
int xc_diffus_centered_kprod(xc_ctx* ctx, xc_mat dff, const xc_mat* qx, const xc_mat* qz, xc_mat f, xc_mat k1,
                             xc_mat k2, double h, int xbc, int zbc);

// record_history!'s reductions (xc_step.hip) enqueued on stream st: part holds XC_HIST_PART_N doubles of
// partial sums (12 x 512 for the sums, then 19 x 7 x 64 for the field statistics), out XC_HIST_NSTATS results
#define XC_HIST_PART_N (12 * 512 + 19 * 7 * 64)
int xc_history_enqueue(xc_ctx* ctx, hipStream_t st, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns,
                       double* part, double* out, const char* who);

// looks up the no-allocation scratch array's address on the current device (xc_noalloc.hip; called by xc_create)
int xc_noalloc_init_device();

// Learned coarse correction (xc_unet.hip), used by the Stokes preconditioner:
// network forward into the context's own output buffer ([p][3] FP32, valid until the next call),
// whether parameters were loaded, and the per-channel input/output scales of the trained network.
int xc_unet_infer(xc_ctx* ctx, const float* in, int B, int H, int W, const float** out);
bool xc_unet_has_params(xc_ctx* ctx);
void xc_unet_io_scales(xc_ctx* ctx, float in[4], float out[3]);

// 2-D launch domain of R x C elements, 64 (z) x 4 (x) lanes per workgroup, XCD-aware order
struct xc_dom { int R, C, nbz, nb; };
static inline xc_dom xc_make_dom(int64_t R, int64_t C) {
    xc_dom d; d.R = (int)R; d.C = (int)C; d.nbz = (int)((R + 63) / 64);
    d.nb = d.nbz * (int)((C + 3) / 4); if (d.nb < 1) d.nb = 1; return d;
}

// Column-major accessor
#define XM(m, i, j) ((m).p[(int64_t)(i) + (int64_t)(j) * (m).ld])

static inline bool xc_mat_ok(const xc_mat& m) {
    return m.p != nullptr && m.rows >= 0 && m.cols >= 0 && (m.cols <= 1 || m.ld >= m.rows);
}
static inline bool xc_shape(const xc_mat& m, int64_t r, int64_t c) {
    return m.rows == r && m.cols == c && xc_mat_ok(m);
}

// element q (column-major order) of a matrix: a contiguous matrix (ld == rows) is indexed directly, any
// other through one 32-bit division (the matrices here hold fewer than 2^31 elements)
__device__ __forceinline__ double xc_mat_at(const xc_mat& m, int64_t q) {
    if (m.ld == m.rows) return m.p[q];
    const unsigned uq = (unsigned)q, ur = (unsigned)m.rows, c = uq / ur;
    return m.p[(uq - c * ur) + (int64_t)c * m.ld];
}
// Hardware dispatches workgroup b of a 1-D grid to XCD (b % 8).  Remap so each
// XCD walks a contiguous run of logical tiles: neighbouring tiles then share an
// L2 and their halos are re-read from cache rather than from HBM.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int xcd = b & 7, slot = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + slot;
}
// Wavefront reductions by DPP: row shifts within the rows of 16 lanes (row_shr 1, 2, 4, 8: lane 15 of
// each row then holds the row's result), then the row broadcasts (row_bcast 15 into rows 1 and 3,
// row_bcast 31 into rows 2 and 3: lane 63 holds the wave's), read back to every lane.  A lane whose source
// lies outside its row, or whose row the mask excludes, takes OLD (the operation's identity).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_shift(double v, double old = 0.0) {
    const int lo = __builtin_amdgcn_update_dpp((int)__double2loint(old), (int)__double2loint(v), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)__double2hiint(old), (int)__double2hiint(v), CTRL, ROWS, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63(double v) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_shift<0x111, 0xf>(v);
    v += dpp_shift<0x112, 0xf>(v);
    v += dpp_shift<0x114, 0xf>(v);
    v += dpp_shift<0x118, 0xf>(v);
    v += dpp_shift<0x142, 0xa>(v);
    v += dpp_shift<0x143, 0xc>(v);
    return lane63(v);
}
__device__ __forceinline__ double wave_min(double v) {
    const double I = INFINITY;
    v = fmin(v, dpp_shift<0x111, 0xf>(v, I));
    v = fmin(v, dpp_shift<0x112, 0xf>(v, I));
    v = fmin(v, dpp_shift<0x114, 0xf>(v, I));
    v = fmin(v, dpp_shift<0x118, 0xf>(v, I));
    v = fmin(v, dpp_shift<0x142, 0xa>(v, I));
    v = fmin(v, dpp_shift<0x143, 0xc>(v, I));
    return lane63(v);
}
// out[q] = the sum (or, MIN, the minimum) of the nb partials part[q * nb + b] for q < nq: one wave per q
// (blockDim.x = 64 * waves), every lane's loads in flight together, a fixed order of operations
template <bool MIN>
__device__ __forceinline__ double xc_rows_reduce(const double* __restrict__ part, int nb, int q) {
    const int lane = threadIdx.x & 63;
    constexpr int U = 8;
    double acc = MIN ? INFINITY : 0.0;
    for (int b0 = lane; b0 < nb; b0 += 64 * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int b = b0 + 64 * u;
            v[u] = part[(int64_t)q * nb + (b < nb ? b : nb - 1)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (b0 + 64 * u < nb) acc = MIN ? fmin(acc, v[u]) : acc + v[u];
    }
    return MIN ? wave_min(acc) : wave_total(acc);
}

__device__ __forceinline__ void xc_dom_ij(const xc_dom& d, int* j, int* i) {
    const int t = xcd_tile(blockIdx.x, d.nb);
    *j = (t % d.nbz) * 64 + (threadIdx.x & 63);
    *i = (t / d.nbz) * 4 + (threadIdx.x >> 6);
}
