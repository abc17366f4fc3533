// xc_stencil.hip -- advection and diffusion stencils for gfx950 (FP64).
//
// Reference: src/kernels/advect.jl, diffus.jl, ghosts.jl and the
// halo-building wrappers of src/phsevo.jl:21-149.
//
// Design (MI355X):
//  * One workgroup (256 lanes = 4 wave64) owns a 64 (z) x 16 (x) tile of cells.
//    z is the contiguous axis, so one wavefront reads one tile column with a
//    single coalesced 512 B access.
//  * The tile plus its scheme halo is staged once in LDS (BCs applied as an
//    index map while staging -- no halo buffer is ever built in HBM).
//  * Every face flux is reconstructed exactly once and kept in LDS; the cell
//    divergence then reads its four faces.  The reference reconstructs 8 face
//    states per cell in the divergence kernel and 4 more per cell again in the
//    flux kernels; here it is 4 per cell in total.
//  * The flux arrays qx/qz (with the MATLAB boundary padding) are written from
//    the same face values, so adv == div(q)/h holds bitwise as in the reference.
//  * Workgroups are remapped XCD-contiguously (xcd_tile) so neighbouring tiles
//    share one of the 8 L2s.
//  * FP contraction is off in this file: every face value is bit-identical to
//    the reference's operation order (Julia does not fuse a*b+c).
#include "xc_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int TZ = 64;   // tile height (z, contiguous)
constexpr int TX = 16;   // tile width (x)
constexpr int NT = 256;  // lanes per workgroup

__host__ __device__ constexpr int halo_of(int s) {
    return s == XC_CENTR ? 1 : s == XC_UPWD1 ? 1 : s == XC_WENO5 ? 3 : 2;
}

__device__ __forceinline__ double vpos(double v) { return (v + fabs(v)) * 0.5; }
__device__ __forceinline__ double vneg(double v) { return (v - fabs(v)) * 0.5; }

// Julia max/min on non-NaN doubles (-0.0 < +0.0)
__device__ __forceinline__ double jmax(double a, double b) {
    return a > b ? a : (b > a ? b : (signbit(a) ? b : a));
}
__device__ __forceinline__ double jmin(double a, double b) {
    return a < b ? a : (b < a ? b : (signbit(a) ? a : b));
}

// advect.jl:35-53, Jiang-Shu WENO5 with eps = eps(Float64)
__device__ __forceinline__ double weno5_poly(double fmm, double fm, double fc, double fp, double fpp) {
    const double e = 2.220446049250313e-16;
    const double sixth = 1.0 / 6.0, quarter = 0.25, tt = 13.0 / 12.0;
    double p1 = sixth * (2.0 * fmm - 7.0 * fm + 11.0 * fc);
    double p2 = sixth * (-fm + 5.0 * fc + 2.0 * fp);
    double p3 = sixth * (2.0 * fc + 5.0 * fp - fpp);
    double s1 = fmm - 2.0 * fm + fc, t1 = fmm - 4.0 * fm + 3.0 * fc;
    double s2 = fm - 2.0 * fc + fp, t2 = fm - fp;
    double s3 = fc - 2.0 * fp + fpp, t3 = 3.0 * fc - 4.0 * fp + fpp;
    double a1 = tt * (s1 * s1) + quarter * (t1 * t1) + e;
    double a2 = tt * (s2 * s2) + quarter * (t2 * t2) + e;
    double a3 = tt * (s3 * s3) + quarter * (t3 * t3) + e;
    double w1 = 0.1 / (a1 * a1);
    double w2 = 0.6 / (a2 * a2);
    double w3 = 0.3 / (a3 * a3);
    return (w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3);
}

// Same reconstruction with one division: the weights d_k/(b_k+eps)^2 are scaled
// by prod_k (b_k+eps)^2 before normalising (mathematically identical; differs
// from the reference's rounding in the last bits).  FMA contraction allowed.
__device__ __forceinline__ double weno5_poly_fast(double fmm, double fm, double fc, double fp, double fpp) {
#pragma clang fp contract(fast)
    const double e = 2.220446049250313e-16;
    const double sixth = 1.0 / 6.0, quarter = 0.25, tt = 13.0 / 12.0;
    const double p1 = sixth * (2.0 * fmm - 7.0 * fm + 11.0 * fc);
    const double p2 = sixth * (-fm + 5.0 * fc + 2.0 * fp);
    const double p3 = sixth * (2.0 * fc + 5.0 * fp - fpp);
    const double s1 = fmm - 2.0 * fm + fc, t1 = fmm - 4.0 * fm + 3.0 * fc;
    const double s2 = fm - 2.0 * fc + fp, t2 = fm - fp;
    const double s3 = fc - 2.0 * fp + fpp, t3 = 3.0 * fc - 4.0 * fp + fpp;
    const double a1 = tt * (s1 * s1) + quarter * (t1 * t1) + e;
    const double a2 = tt * (s2 * s2) + quarter * (t2 * t2) + e;
    const double a3 = tt * (s3 * s3) + quarter * (t3 * t3) + e;
    const double q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3;
    const double w1 = 0.1 * (q2 * q3), w2 = 0.6 * (q1 * q3), w3 = 0.3 * (q1 * q2);
    return (w1 * p1 + w2 * p2 + w3 * p3) / (w1 + w2 + w3);
}

// advect.jl:216-229 (eps fixed at 1e-6)
__device__ __forceinline__ double weno3_poly(double fm, double fc, double fp) {
    const double e = 1e-6;
    double p1 = 0.5 * (fc + fp);
    double p2 = 0.5 * (3.0 * fc - fm);
    double b1 = (fp - fc) * (fp - fc);
    double b2 = (fc - fm) * (fc - fm);
    double w1 = (1.0 / 3.0) / (b1 + e);
    double w2 = (2.0 / 3.0) / (b2 + e);
    return (w1 * p1 + w2 * p2) / (w1 + w2);
}

// advect.jl:275-283, superbee limiter
__device__ __forceinline__ double tvd_flux(double fm, double fc, double fp, double vm, double vp) {
    const double e = 2.220446049250313e-16;
    double num = vm * (fc - fm);
    double den = vp * (fp - fc);
    double R = num / (den + copysign(e, den));
    double l = jmax(0.0, jmax(jmin(1.0, 2.0 * R), jmin(2.0, R)));
    return fc + 0.5 * l * (fp - fc);
}

// Flux through one face.  c[k*st] is the cell k places downstream of the face
// (k = 0: the cell after the face, k = -1: the cell before it).  vl/vr are the
// neighbouring face velocities (TVD only).  Same algebra as the reference's
// per-face flux kernels, advect.jl:444-695.
template <int S, bool FAST>
__device__ __forceinline__ double face_flux(const double* c, int st, double vl, double v, double vr) {
#define F(k) c[(k) * st]
    if constexpr (S == XC_CENTR) {
        return v * (F(-1) + F(0)) * 0.5;
    } else if constexpr (S == XC_TVDIM) {
        double fppos = tvd_flux(F(-2), F(-1), F(0), vpos(vl), vpos(v));
        double fpneg = tvd_flux(F(1), F(0), F(-1), vneg(vr), vneg(v));
        return vpos(v) * fppos + vneg(v) * fpneg;
    } else {
        double fppos, fpneg;
        if constexpr (S == XC_UPWD1) {
            fppos = F(-1);
            fpneg = F(0);
        } else if constexpr (S == XC_QUICK) {
            const double sixth = 1.0 / 6.0;
            fppos = sixth * (2 * F(0) + 5 * F(-1) - F(-2));
            fpneg = sixth * (2 * F(-1) + 5 * F(0) - F(1));
        } else if constexpr (S == XC_FROMM) {
            fppos = F(-1) + 0.25 * (F(0) - F(-2));
            fpneg = F(0) + 0.25 * (F(-1) - F(1));
        } else if constexpr (S == XC_WENO3) {
            fppos = weno3_poly(F(-2), F(-1), F(0));
            fpneg = weno3_poly(F(1), F(0), F(-1));
        } else {
            if constexpr (FAST) {
                fppos = weno5_poly_fast(F(-3), F(-2), F(-1), F(0), F(1));
                fpneg = weno5_poly_fast(F(2), F(1), F(0), F(-1), F(-2));
            } else {
                fppos = weno5_poly(F(-3), F(-2), F(-1), F(0), F(1));
                fpneg = weno5_poly(F(2), F(1), F(0), F(-1), F(-2));
            }
        }
        return vpos(v) * fppos + vneg(v) * fpneg;
    }
#undef F
}

// cell index map for the BCs of fill_ghosts! (ghosts.jl:18-51), valid for |i| within one halo
__device__ __forceinline__ int map_cell(int i, int n, int bc) {
    if (bc == XC_PERIODIC) return i < 0 ? i + n : (i >= n ? i - n : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
// face index map for the padded TVD velocities (phsevo.jl:34-53): faces -1 and n+1
__device__ __forceinline__ int map_face(int f, int n, int bc) {
    if (f < 0) return bc == XC_PERIODIC ? n - 1 : 0;
    if (f > n) return bc == XC_PERIODIC ? 1 : n;
    return f;
}

struct AdvArgs {
    double* adv; int64_t ld_adv;
    double* qx; int64_t ld_qx;
    double* qz; int64_t ld_qz;
    const double* f; int64_t ld_f;
    const double* u; int64_t ld_u;
    const double* w; int64_t ld_w;
    int Nz, Nx, nbz, nb;
    double invh;
    int xbc, zbc;
};

// HIN: f is a halo'd buffer (advect!); otherwise f is (Nz, Nx) and the BCs are
// an index map (advect_centered!).  FLUX: also write qx/qz with padding.
template <int S, bool HIN, bool FLUX, bool FAST>
__global__ __launch_bounds__(NT) void k_advect(AdvArgs a) {
    constexpr int G = halo_of(S);
    constexpr int PZ = TZ + 2 * G;
    constexpr int PX = TX + 2 * G;
    __shared__ double fs[PX * PZ];
    __shared__ double qxs[(TX + 1) * TZ];
    __shared__ double qzs[TX * (TZ + 1)];

    const int t = xcd_tile(blockIdx.x, a.nb);
    const int z0 = (t % a.nbz) * TZ;
    const int x0 = (t / a.nbz) * TX;
    const int tid = threadIdx.x;
    const int Nz = a.Nz, Nx = a.Nx;

    // ---- stage the tile + halo (cross-shaped use; corners loaded for simplicity)
    for (int k = tid; k < PX * PZ; k += NT) {
        const int lz = k % PZ, lx = k / PZ;
        int gz = z0 + lz - G, gx = x0 + lx - G;
        gz = gz < Nz + G - 1 ? gz : Nz + G - 1;
        gx = gx < Nx + G - 1 ? gx : Nx + G - 1;
        int hz, hx;
        if constexpr (HIN) {
            hz = gz + G;
            hx = gx + G;
        } else {
            hz = map_cell(gz, Nz, a.zbc);
            hx = map_cell(gx, Nx, a.xbc);
        }
        fs[k] = a.f[hz + (int64_t)hx * a.ld_f];
    }
    __syncthreads();

    const int qx_top = (a.zbc == XC_PERIODIC && Nz > 1) ? Nz - 1 : 0;  // source row of qx[0, :]
    const int qx_bot = (a.zbc == XC_PERIODIC && Nz > 1) ? 0 : Nz - 1;  // source row of qx[Nz+1, :]
    const int qz_lft = (a.xbc == XC_PERIODIC && Nx > 1) ? Nx - 1 : 0;
    const int qz_rgt = (a.xbc == XC_PERIODIC && Nx > 1) ? 0 : Nx - 1;

    // ---- x-faces: TZ rows x (TX+1) faces
    for (int k = tid; k < TZ * (TX + 1); k += NT) {
        const int lz = k % TZ, fx = k / TZ;
        const int iz = z0 + lz, F = x0 + fx;
        double q = 0.0;
        if (iz < Nz && F <= Nx) {
            double vl = 0.0, v, vr = 0.0;
            if constexpr (S == XC_TVDIM) {
                if constexpr (HIN) {
                    vl = a.u[iz + (int64_t)F * a.ld_u];
                    v = a.u[iz + (int64_t)(F + 1) * a.ld_u];
                    vr = a.u[iz + (int64_t)(F + 2) * a.ld_u];
                } else {
                    vl = a.u[iz + (int64_t)map_face(F - 1, Nx, a.xbc) * a.ld_u];
                    v = a.u[iz + (int64_t)F * a.ld_u];
                    vr = a.u[iz + (int64_t)map_face(F + 1, Nx, a.xbc) * a.ld_u];
                }
            } else {
                v = a.u[iz + (int64_t)F * a.ld_u];
            }
            q = face_flux<S, FAST>(&fs[(fx + G) * PZ + lz + G], PZ, vl, v, vr);
            if constexpr (FLUX) {
                if (fx < TX || F == Nx) {
                    double* col = a.qx + (int64_t)F * a.ld_qx;
                    col[iz + 1] = q;
                    if (iz == qx_top) col[0] = q;
                    if (iz == qx_bot) col[Nz + 1] = q;
                }
            }
        }
        qxs[fx * TZ + lz] = q;
    }
    // ---- z-faces: (TZ+1) faces x TX columns
    for (int k = tid; k < (TZ + 1) * TX; k += NT) {
        const int fz = k % (TZ + 1), lx = k / (TZ + 1);
        const int F = z0 + fz, ix = x0 + lx;
        double q = 0.0;
        if (ix < Nx && F <= Nz) {
            double vl = 0.0, v, vr = 0.0;
            if constexpr (S == XC_TVDIM) {
                if constexpr (HIN) {
                    vl = a.w[F + (int64_t)ix * a.ld_w];
                    v = a.w[F + 1 + (int64_t)ix * a.ld_w];
                    vr = a.w[F + 2 + (int64_t)ix * a.ld_w];
                } else {
                    vl = a.w[map_face(F - 1, Nz, a.zbc) + (int64_t)ix * a.ld_w];
                    v = a.w[F + (int64_t)ix * a.ld_w];
                    vr = a.w[map_face(F + 1, Nz, a.zbc) + (int64_t)ix * a.ld_w];
                }
            } else {
                v = a.w[F + (int64_t)ix * a.ld_w];
            }
            q = face_flux<S, FAST>(&fs[(lx + G) * PZ + fz + G], 1, vl, v, vr);
            if constexpr (FLUX) {
                if (fz < TZ || F == Nz) {
                    a.qz[F + (int64_t)(ix + 1) * a.ld_qz] = q;
                    if (ix == qz_lft) a.qz[F] = q;
                    if (ix == qz_rgt) a.qz[F + (int64_t)(Nx + 1) * a.ld_qz] = q;
                }
            }
        }
        qzs[lx * (TZ + 1) + fz] = q;
    }
    __syncthreads();

    // ---- divergence, advect.jl:94: (qxr - qxl + qzb - qzt) * invh
    for (int k = tid; k < TZ * TX; k += NT) {
        const int lz = k % TZ, lx = k / TZ;
        const int iz = z0 + lz, ix = x0 + lx;
        if (iz < Nz && ix < Nx) {
            const double qxr = qxs[(lx + 1) * TZ + lz], qxl = qxs[lx * TZ + lz];
            const double qzb = qzs[lx * (TZ + 1) + lz + 1], qzt = qzs[lx * (TZ + 1) + lz];
            a.adv[iz + (int64_t)ix * a.ld_adv] = (qxr - qxl + qzb - qzt) * a.invh;
        }
    }
}

template <int S, bool HIN, bool FLUX>
hipError_t launch_advect_t(const AdvArgs& a0, hipStream_t s, bool exact) {
    AdvArgs a = a0;
    // only WENO5 has a distinct fast form; every other scheme is the reference arithmetic
    if (S == XC_WENO5 && !exact) hipLaunchKernelGGL((k_advect<S, HIN, FLUX, true>), dim3(a.nb), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((k_advect<S, HIN, FLUX, false>), dim3(a.nb), dim3(NT), 0, s, a);
    return hipGetLastError();
}

template <bool HIN, bool FLUX>
hipError_t launch_advect_s(int scheme, const AdvArgs& a, hipStream_t s, bool exact) {
    switch (scheme) {
        case XC_CENTR: return launch_advect_t<XC_CENTR, HIN, FLUX>(a, s, exact);
        case XC_UPWD1: return launch_advect_t<XC_UPWD1, HIN, FLUX>(a, s, exact);
        case XC_QUICK: return launch_advect_t<XC_QUICK, HIN, FLUX>(a, s, exact);
        case XC_FROMM: return launch_advect_t<XC_FROMM, HIN, FLUX>(a, s, exact);
        case XC_WENO3: return launch_advect_t<XC_WENO3, HIN, FLUX>(a, s, exact);
        case XC_WENO5: return launch_advect_t<XC_WENO5, HIN, FLUX>(a, s, exact);
        default: return launch_advect_t<XC_TVDIM, HIN, FLUX>(a, s, exact);
    }
}

int advect_common(xc_ctx* ctx, int scheme, bool hin, xc_mat adv, const xc_mat* qx, const xc_mat* qz,
                  xc_mat f, xc_mat u, xc_mat w, double h, int xbc, int zbc, const char* name) {
    XC_CHECK_CTX(ctx);
    const int g = xc_scheme_halo(scheme);
    if (g < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "%s: unsupported scheme = %d", name, scheme);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported xBC = %d", name, xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported zBC = %d", name, zbc);
    const int64_t Nz = adv.rows, Nx = adv.cols;
    if (!xc_mat_ok(adv) || Nz < 1 || Nx < 1) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: adv must be a non-empty matrix", name);
    if (Nz > (1 << 30) || Nx > (1 << 30)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: grid too large", name);
    if (hin) {
        if (!xc_shape(f, Nz + 2 * g, Nx + 2 * g))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "advect!: f_halo size doesn't match adv + 2*halo");
        if (scheme == XC_TVDIM) {
            if (!xc_shape(u, Nz, Nx + 3)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect!: tvdim requires halo'd u of size (Nz, Nx+3)");
            if (!xc_shape(w, Nz + 3, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect!: tvdim requires halo'd w of size (Nz+3, Nx)");
        } else {
            if (!xc_shape(u, Nz, Nx + 1)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect!: u must be (Nz, Nx+1)");
            if (!xc_shape(w, Nz + 1, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect!: w must be (Nz+1, Nx)");
        }
    } else {
        if (!xc_shape(f, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: f must be (Nz, Nx)", name);
        if (!xc_shape(u, Nz, Nx + 1)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: u must be (Nz, Nx+1)", name);
        if (!xc_shape(w, Nz + 1, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: w must be (Nz+1, Nx)", name);
        if (Nz < g || Nx < g)
            return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: grid %lldx%lld smaller than the scheme halo %d", name,
                           (long long)Nz, (long long)Nx, g);
    }
    const bool flux = qx && qz;
    if ((qx != nullptr) != (qz != nullptr)) return xc_fail(ctx, XC_EINVAL_ARG, "%s: pass both qx and qz or neither", name);
    if (flux) {
        if (!xc_shape(*qx, Nz + 2, Nx + 1)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_with_flux!: qx must be (Nz+2, Nx+1)");
        if (!xc_shape(*qz, Nz + 1, Nx + 2)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_with_flux!: qz must be (Nz+1, Nx+2)");
    }
    AdvArgs a{};
    a.adv = adv.p; a.ld_adv = adv.ld;
    if (flux) { a.qx = qx->p; a.ld_qx = qx->ld; a.qz = qz->p; a.ld_qz = qz->ld; }
    a.f = f.p; a.ld_f = f.ld;
    a.u = u.p; a.ld_u = u.ld;
    a.w = w.p; a.ld_w = w.ld;
    a.Nz = (int)Nz; a.Nx = (int)Nx;
    a.nbz = (int)((Nz + TZ - 1) / TZ);
    const int nbx = (int)((Nx + TX - 1) / TX);
    a.nb = a.nbz * nbx;
    a.invh = 1.0 / h;
    a.xbc = xbc; a.zbc = zbc;
    hipError_t e;
    const bool ex = (ctx->flags & XC_CTX_EXACT) != 0;
    if (hin) e = flux ? launch_advect_s<true, true>(scheme, a, ctx->stream, ex) : launch_advect_s<true, false>(scheme, a, ctx->stream, ex);
    else     e = flux ? launch_advect_s<false, true>(scheme, a, ctx->stream, ex) : launch_advect_s<false, false>(scheme, a, ctx->stream, ex);
    if (e != hipSuccess) return xc_fail(ctx, XC_EHIP, "%s: launch: %s", name, hipGetErrorString(e));
    return xc_finish(ctx, name);
}

// ---------------------------------------------------------------------------
// diffusion (diffus.jl) -- memory-bound 5-point stencil, one cell per lane;
// neighbour reads hit L1/L2.  Each lane also writes the x-face left of its
// cell and the z-face above it (and the closing faces on the last row/column).

struct DifArgs {
    double* dff; int64_t ld_d;
    double* qx; int64_t ld_qx;
    double* qz; int64_t ld_qz;
    const double* f; int64_t ld_f;
    const double* k; int64_t ld_k;
    const double* k2; int64_t ld_k2;   // optional: diffusivity is k .* k2 (phsevo.jl:192)
    int Nz, Nx, halo, nbz, nb;
    double invh, invh2;
    int xbc, zbc;
};

template <bool HIN, bool FLUX>
__global__ __launch_bounds__(NT) void k_diffus(DifArgs a) {
    const int t = xcd_tile(blockIdx.x, a.nb);
    const int iz = (t % a.nbz) * 64 + (threadIdx.x & 63);
    const int ix = (t / a.nbz) * 4 + (threadIdx.x >> 6);
    const int Nz = a.Nz, Nx = a.Nx;
    if (iz >= Nz || ix >= Nx) return;
    // index of cell (z, x) in the input buffers
    auto at = [&](const double* p, int64_t ld, int z, int x) -> double {
        if constexpr (HIN) return p[(z + a.halo) + (int64_t)(x + a.halo) * ld];
        else return p[map_cell(z, Nz, a.zbc) + (int64_t)map_cell(x, Nx, a.xbc) * ld];
    };
    auto kat = [&](int z, int x) -> double {
        if constexpr (HIN) return at(a.k, a.ld_k, z, x);
        else {
            const int64_t o = map_cell(z, Nz, a.zbc), c = map_cell(x, Nx, a.xbc);
            return a.k2 ? a.k[o + c * a.ld_k] * a.k2[o + c * a.ld_k2] : a.k[o + c * a.ld_k];
        }
    };
    const double fc = at(a.f, a.ld_f, iz, ix);
    const double fmx = at(a.f, a.ld_f, iz, ix - 1), fpx = at(a.f, a.ld_f, iz, ix + 1);
    const double fmz = at(a.f, a.ld_f, iz - 1, ix), fpz = at(a.f, a.ld_f, iz + 1, ix);
    const double kc = kat(iz, ix);
    const double kmx = kat(iz, ix - 1), kpx = kat(iz, ix + 1);
    const double kmz = kat(iz - 1, ix), kpz = kat(iz + 1, ix);
    // diffus.jl:23-29
    const double flux_xr = (kpx + kc) * (fpx - fc);
    const double flux_xl = (kc + kmx) * (fc - fmx);
    const double flux_zb = (kpz + kc) * (fpz - fc);
    const double flux_zt = (kc + kmz) * (fc - fmz);
    a.dff[iz + (int64_t)ix * a.ld_d] = 0.5 * a.invh2 * (flux_xr - flux_xl + flux_zb - flux_zt);
    if constexpr (FLUX) {
        // diffus.jl:77 / 89: q = -(kL + kR) * 1/2 * (fR - fL) / h
        const int qx_top = (a.zbc == XC_PERIODIC && Nz > 1) ? Nz - 1 : 0;
        const int qx_bot = (a.zbc == XC_PERIODIC && Nz > 1) ? 0 : Nz - 1;
        const int qz_lft = (a.xbc == XC_PERIODIC && Nx > 1) ? Nx - 1 : 0;
        const int qz_rgt = (a.xbc == XC_PERIODIC && Nx > 1) ? 0 : Nx - 1;
        auto put_qx = [&](int F, double q) {
            double* col = a.qx + (int64_t)F * a.ld_qx;
            col[iz + 1] = q;
            if (iz == qx_top) col[0] = q;
            if (iz == qx_bot) col[Nz + 1] = q;
        };
        auto put_qz = [&](int F, double q) {
            a.qz[F + (int64_t)(ix + 1) * a.ld_qz] = q;
            if (ix == qz_lft) a.qz[F] = q;
            if (ix == qz_rgt) a.qz[F + (int64_t)(Nx + 1) * a.ld_qz] = q;
        };
        put_qx(ix, -(kmx + kc) * 0.5 * (fc - fmx) * a.invh);
        if (ix == Nx - 1) put_qx(Nx, -(kc + kpx) * 0.5 * (fpx - fc) * a.invh);
        put_qz(iz, -(kmz + kc) * 0.5 * (fc - fmz) * a.invh);
        if (iz == Nz - 1) put_qz(Nz, -(kc + kpz) * 0.5 * (fpz - fc) * a.invh);
    }
}

int diffus_common(xc_ctx* ctx, bool hin, int halo, xc_mat dff, const xc_mat* qx, const xc_mat* qz,
                  xc_mat f, xc_mat k, double h, int xbc, int zbc, const char* name, const xc_mat* k2 = nullptr) {
    XC_CHECK_CTX(ctx);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported xBC = %d", name, xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported zBC = %d", name, zbc);
    const int64_t Nz = dff.rows, Nx = dff.cols;
    if (!xc_mat_ok(dff) || Nz < 1 || Nx < 1) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: dff must be non-empty", name);
    if (hin) {
        if (halo < 1) return xc_fail(ctx, XC_EINVAL_ARG, "diffus!: halo must be >= 1");
        if (f.rows != k.rows || f.cols != k.cols || !xc_mat_ok(f) || !xc_mat_ok(k))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus!: f and k must have matching halo sizes");
        if (Nz != f.rows - 2 * halo) return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus!: output row count must match f interior");
        if (Nx != f.cols - 2 * halo) return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus!: output col count must match f interior");
    } else {
        halo = 1;
        if (!xc_shape(f, Nz, Nx) || !xc_shape(k, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: f and k must be (Nz, Nx)", name);
        if (k2 && !xc_shape(*k2, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: k2 must be (Nz, Nx)", name);
    }
    if ((qx != nullptr) != (qz != nullptr)) return xc_fail(ctx, XC_EINVAL_ARG, "%s: pass both qx and qz or neither", name);
    const bool flux = qx && qz;
    if (flux) {
        if (!xc_shape(*qx, Nz + 2, Nx + 1)) return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_with_flux!: qx must be (Nz+2, Nx+1)");
        if (!xc_shape(*qz, Nz + 1, Nx + 2)) return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_with_flux!: qz must be (Nz+1, Nx+2)");
    }
    DifArgs a{};
    a.dff = dff.p; a.ld_d = dff.ld;
    if (flux) { a.qx = qx->p; a.ld_qx = qx->ld; a.qz = qz->p; a.ld_qz = qz->ld; }
    a.f = f.p; a.ld_f = f.ld; a.k = k.p; a.ld_k = k.ld;
    if (k2) { a.k2 = k2->p; a.ld_k2 = k2->ld; }
    a.Nz = (int)Nz; a.Nx = (int)Nx; a.halo = halo;
    a.nbz = (int)((Nz + 63) / 64);
    a.nb = a.nbz * (int)((Nx + 3) / 4);
    a.invh = 1.0 / h;
    a.invh2 = a.invh * a.invh;  // diffus.jl:53 T(inv(h)^2)
    a.xbc = xbc; a.zbc = zbc;
    if (hin) {
        if (flux) hipLaunchKernelGGL((k_diffus<true, true>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else hipLaunchKernelGGL((k_diffus<true, false>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
    } else {
        if (flux) hipLaunchKernelGGL((k_diffus<false, true>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else hipLaunchKernelGGL((k_diffus<false, false>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
    }
    return xc_finish(ctx, name);
}

// ---------------------------------------------------------------------------
// fill_ghosts! / embed_interior! (ghosts.jl)

__global__ void k_ghosts_x(double* f, int64_t ld, int Nzt, int Nx, int halo, int bc) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;  // 0..2*halo-1
    if (j >= Nzt) return;
    const bool left = k < halo;
    const int c = left ? k : Nx + halo + (k - halo);
    int src;
    if (bc == XC_PERIODIC) src = left ? c + Nx : c - Nx;
    else src = left ? halo : Nx + halo - 1;
    f[j + (int64_t)c * ld] = f[j + (int64_t)src * ld];
}

__global__ void k_ghosts_z(double* f, int64_t ld, int Nxt, int Nz, int halo, int bc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= Nxt) return;
    const bool top = k < halo;
    const int r = top ? k : Nz + halo + (k - halo);
    int src;
    if (bc == XC_PERIODIC) src = top ? r + Nz : r - Nz;
    else src = top ? halo : Nz + halo - 1;
    f[r + (int64_t)i * ld] = f[src + (int64_t)i * ld];
}

__global__ void k_embed(double* fh, int64_t ldh, const double* f, int64_t ldf, int Nz, int Nx, int halo) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (j >= Nz) return;
    fh[(j + halo) + (int64_t)(i + halo) * ldh] = f[j + (int64_t)i * ldf];
}

}  // namespace

int xc_diffus_centered_kprod(xc_ctx* ctx, xc_mat dff, const xc_mat* qx, const xc_mat* qz, xc_mat f, xc_mat k1,
                             xc_mat k2, double h, int xbc, int zbc) {
    return diffus_common(ctx, false, 1, dff, qx, qz, f, k1, h, xbc, zbc, "diffus_centered_with_flux!", &k2);
}

extern "C" {

int xc_scheme_halo(int scheme) {
    if (scheme < XC_CENTR || scheme > XC_TVDIM) return -1;
    return halo_of(scheme);
}

int xc_fill_ghosts(xc_ctx* ctx, xc_mat f, int halo, int xbc, int zbc) {
    XC_CHECK_CTX(ctx);
    if (!xc_mat_ok(f) || halo < 0) return xc_fail(ctx, XC_EINVAL_ARG, "fill_ghosts!: bad array");
    const int64_t Nz = f.rows - 2 * halo, Nx = f.cols - 2 * halo;
    if (Nz < 1 || Nx < 1) return xc_fail(ctx, XC_EINVAL_SHAPE, "fill_ghosts!: interior must be non-empty (halo too large?)");
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "fill_ghosts!: unsupported xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "fill_ghosts!: unsupported zBC = %d", zbc);
    if ((xbc == XC_PERIODIC && Nx < halo) || (zbc == XC_PERIODIC && Nz < halo))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "fill_ghosts!: periodic fill needs an interior at least `halo` wide");
    if (halo == 0) return XC_OK;
    hipLaunchKernelGGL(k_ghosts_x, dim3((unsigned)((f.rows + 255) / 256), 2 * halo), dim3(256), 0, ctx->stream,
                       f.p, f.ld, (int)f.rows, (int)Nx, halo, xbc);
    hipLaunchKernelGGL(k_ghosts_z, dim3((unsigned)((f.cols + 255) / 256), 2 * halo), dim3(256), 0, ctx->stream,
                       f.p, f.ld, (int)f.cols, (int)Nz, halo, zbc);
    return xc_finish(ctx, "fill_ghosts!");
}

int xc_embed_interior(xc_ctx* ctx, xc_mat fh, xc_mat f, int halo) {
    XC_CHECK_CTX(ctx);
    if (!xc_mat_ok(fh) || !xc_mat_ok(f) || fh.rows != f.rows + 2 * halo || fh.cols != f.cols + 2 * halo)
        return xc_fail(ctx, XC_EINVAL_SHAPE, "embed_interior!: f_halo must be f + 2*halo");
    if (f.rows == 0 || f.cols == 0) return XC_OK;
    hipLaunchKernelGGL(k_embed, dim3((unsigned)((f.rows + 255) / 256), (unsigned)f.cols), dim3(256), 0, ctx->stream,
                       fh.p, fh.ld, f.p, f.ld, (int)f.rows, (int)f.cols, halo);
    return xc_finish(ctx, "embed_interior!");
}

int xc_advect(xc_ctx* ctx, int scheme, xc_mat adv, xc_mat f_halo, xc_mat u, xc_mat w, double h) {
    return advect_common(ctx, scheme, true, adv, nullptr, nullptr, f_halo, u, w, h, XC_PERIODIC, XC_CLOSED, "advect!");
}

int xc_advect_with_flux(xc_ctx* ctx, int scheme, xc_mat adv, xc_mat qx, xc_mat qz, xc_mat f_halo, xc_mat u,
                        xc_mat w, double h, int xbc, int zbc) {
    return advect_common(ctx, scheme, true, adv, &qx, &qz, f_halo, u, w, h, xbc, zbc, "advect_with_flux!");
}

int xc_advect_centered(xc_ctx* ctx, int scheme, xc_mat adv, const xc_mat* qx, const xc_mat* qz, xc_mat f, xc_mat u,
                       xc_mat w, double h, int xbc, int zbc) {
    return advect_common(ctx, scheme, false, adv, qx, qz, f, u, w, h, xbc, zbc,
                         qx ? "advect_centered_with_flux!" : "advect_centered!");
}

int xc_diffus(xc_ctx* ctx, xc_mat dff, xc_mat f_halo, xc_mat k_halo, double h, int halo) {
    return diffus_common(ctx, true, halo, dff, nullptr, nullptr, f_halo, k_halo, h, XC_PERIODIC, XC_CLOSED, "diffus!");
}

int xc_diffus_with_flux(xc_ctx* ctx, xc_mat dff, xc_mat qx, xc_mat qz, xc_mat f_halo, xc_mat k_halo, double h,
                        int halo, int xbc, int zbc) {
    return diffus_common(ctx, true, halo, dff, &qx, &qz, f_halo, k_halo, h, xbc, zbc, "diffus_with_flux!");
}

int xc_diffus_centered(xc_ctx* ctx, xc_mat dff, const xc_mat* qx, const xc_mat* qz, xc_mat f, xc_mat k, double h,
                       int xbc, int zbc) {
    return diffus_common(ctx, false, 1, dff, qx, qz, f, k, h, xbc, zbc,
                         qx ? "diffus_centered_with_flux!" : "diffus_centered!");
}

}  // extern "C"
