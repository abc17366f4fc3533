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

constexpr int TZ = 64;   // default tile height (z, contiguous)
constexpr int TX = 16;   // default tile width (x)
constexpr int NT = 256;  // lanes per workgroup
// advection tile shapes the autotuner chooses from (variant 0 is the default)
constexpr int NVAR = 6;
constexpr int VAR_TZ[NVAR] = {64, 128, 64, 32, 64, 64};
constexpr int VAR_TX[NVAR] = {16, 8, 8, 32, 16, 8};   // variants 4, 5: persistent, LDS-DMA pipelined
constexpr int VAR_PIPE = 4;                           // the first persistent variant

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

// Same reconstruction, fewer operations (the fast path; differs from the reference's rounding in the
// last bits): the 1/6 of the candidate polynomials and the linear weights d_k = (0.1, 0.6, 0.3) are
// folded into the polynomial coefficients, the nonlinear weights d_k/(b_k+eps)^2 are scaled by
// prod_k (b_k+eps)^2 so that one quotient remains, and that quotient is a reciprocal refined by two
// Newton-Raphson steps plus one residual correction of the product (the denominator is a positive
// normal number: each b_k + eps >= eps).  FMA contraction allowed.  When the largest (b_k + eps) exceeds
// 2^200 they are first scaled by the power of two that brings the largest into [0.5, 1): the quotient is
// homogeneous of degree 0 in them and a power-of-two scaling is exact (the hardware reciprocal depends
// on the mantissa only), so the result has the same bits as unscaled wherever the unscaled products stay
// finite -- and the products of four squares (~ |df|^8) no longer overflow for fields of magnitude ~1e38
// and above, up to the ~1e77 where the reference's own weights overflow.  Below 2^200 the products stay
// under 2^800 and the scaling is skipped (the same bits).
__device__ __forceinline__ double weno5_poly_fast(double fmm, double fm, double fc, double fp, double fpp) {
#pragma clang fp contract(fast)
    const double e = 2.220446049250313e-16;
    const double tt = 13.0 / 12.0;
    const double P1 = fmm * (0.1 * 2.0 / 6.0) + fm * (-0.1 * 7.0 / 6.0) + fc * (0.1 * 11.0 / 6.0);
    const double P2 = fm * (-0.6 / 6.0) + fc * (0.6 * 5.0 / 6.0) + fp * (0.6 * 2.0 / 6.0);
    const double P3 = fc * (0.3 * 2.0 / 6.0) + fp * (0.3 * 5.0 / 6.0) + fpp * (-0.3 / 6.0);
    const double s1 = fmm - 2.0 * fm + fc, t1 = fmm - 4.0 * fm + 3.0 * fc;
    const double s2 = fm - 2.0 * fc + fp, t2 = fm - fp;
    const double s3 = fc - 2.0 * fp + fpp, t3 = 3.0 * fc - 4.0 * fp + fpp;
    const double b1 = (tt * s1) * s1 + ((0.25 * t1) * t1 + e);
    const double b2 = (tt * s2) * s2 + ((0.25 * t2) * t2 + e);
    const double b3 = (tt * s3) * s3 + ((0.25 * t3) * t3 + e);
    double a1 = b1, a2 = b2, a3 = b3;
    const double bm = fmax(b1, fmax(b2, b3));
    if (bm > 0x1p200) {  // products of four squares would pass 2^800: rescale (rare; exact either way)
        const int ex = __builtin_amdgcn_frexp_exp(bm);
        a1 = __builtin_amdgcn_ldexp(b1, -ex);
        a2 = __builtin_amdgcn_ldexp(b2, -ex);
        a3 = __builtin_amdgcn_ldexp(b3, -ex);
    }
    const double q1 = a1 * a1, q2 = a2 * a2, q3 = a3 * a3;
    const double q23 = q2 * q3, q13 = q1 * q3, q12 = q1 * q2;
    const double num = q23 * P1 + q13 * P2 + q12 * P3;
    const double den = q23 * 0.1 + q13 * 0.6 + q12 * 0.3;
    double r = __builtin_amdgcn_rcp(den);
    r = fma(r, fma(-den, r, 1.0), r);
    r = fma(r, fma(-den, r, 1.0), r);
    const double y = num * r;
    return fma(r, fma(-den, y, num), y);
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
        // Only the upwind reconstruction can contribute: for v > 0, vpos(v) = v and vneg(v) = 0
        // exactly (and the reverse for v < 0), so v * f_upwind has the bits of the reference's
        // vpos(v) * fppos + vneg(v) * fpneg (advect.jl:613-617) for finite values.  The stencil is
        // mirrored by selecting its points, not by branching.
        const bool p = v > 0.0;
        double fp;
        if constexpr (S == XC_UPWD1) {
            fp = p ? F(-1) : F(0);
        } else if constexpr (S == XC_QUICK) {
            const double sixth = 1.0 / 6.0;
            const double s0 = p ? F(-2) : F(1), s1 = p ? F(-1) : F(0), s2 = p ? F(0) : F(-1);
            fp = sixth * (2 * s2 + 5 * s1 - s0);
        } else if constexpr (S == XC_FROMM) {
            const double s0 = p ? F(-2) : F(1), s1 = p ? F(-1) : F(0), s2 = p ? F(0) : F(-1);
            fp = s1 + 0.25 * (s2 - s0);
        } else if constexpr (S == XC_WENO3) {
            fp = p ? weno3_poly(F(-2), F(-1), F(0)) : weno3_poly(F(1), F(0), F(-1));
        } else {
            const double a0 = p ? F(-3) : F(2), a1 = p ? F(-2) : F(1), a2 = p ? F(-1) : F(0);
            const double a3 = p ? F(0) : F(-1), a4 = p ? F(1) : F(-2);
            if constexpr (FAST) fp = weno5_poly_fast(a0, a1, a2, a3, a4);
            else fp = weno5_poly(a0, a1, a2, a3, a4);
        }
        return v * fp;
    }
#undef F
}

// cell index map for the BCs of fill_ghosts! (ghosts.jl:18-51), valid for |i| within one halo
// Workgroup barrier for LDS traffic only: waits for this wave's LDS operations (lgkmcnt(0)) and
// meets the other waves, but leaves global stores in flight (a full __syncthreads() would also wait
// for every outstanding global memory operation).  Used where no thread reads global memory
// another thread of the workgroup wrote.  The empty asm keeps the compiler from moving memory
// accesses across it.
__device__ __forceinline__ void lds_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt(0xc07f);  // vmcnt(63) expcnt(7) lgkmcnt(0)
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

__device__ __forceinline__ int map_cell(int i, int n, int bc) {
    if (bc == XC_PERIODIC) return i < 0 ? i + n : (i >= n ? i - n : i);
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}
// x of lane l + 1 in lane l (lane 63 receives 0): DPP wavefront shift by one lane.  Every lane of
// the wave must execute it.
__device__ __forceinline__ double lane_next(double x) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// Output store; NTS: non-temporal (the line is streamed past the caches instead of being kept for a
// later reader).  Used when a launch's outputs are much larger than the last-level cache.
template <bool NTS>
__device__ __forceinline__ void put(double* p, double v) {
    if constexpr (NTS) __builtin_nontemporal_store(v, p);
    else *p = v;
}
// Outputs of one launch above this many bytes are written non-temporally: a quarter of the 256 MB
// last-level cache, beyond which a later kernel would find little of them there anyway.
constexpr int64_t NT_STORE_BYTES = 64ll << 20;
// bytes a stencil launch writes: the cell output, and the padded face fluxes when requested
inline int64_t out_bytes(int64_t Nz, int64_t Nx, bool flux) {
    return 8 * (Nz * Nx + (flux ? (Nz + 2) * (Nx + 1) + (Nz + 1) * (Nx + 2) : 0));
}
// map_cell for a wave-uniform periodic flag, as selects (no branch between a tile's loads)
__device__ __forceinline__ int map_cell_sel(int i, int n, bool per) {
    const int lo = per ? i + n : 0, hi = per ? i - n : n - 1;
    return i < 0 ? lo : (i >= n ? hi : i);
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
    int nts;  // host side: launch the non-temporal-store form (outputs larger than NT_STORE_BYTES)
};

// HIN: f is a halo'd buffer (advect!); otherwise f is (Nz, Nx) and the BCs are
// an index map (advect_centered!).  FLUX: also write qx/qz with padding.
// One tile t of one grid; the body of k_advect and of the ensemble kernel k_advect_batch.
// Every global load of the tile (the f halo tile and the face velocities) is issued back to back
// into registers before the first LDS store, so a wave keeps all of them in flight at once; indexes
// past the grid are clamped to a valid element whose value is never used.
template <int S, bool HIN, bool FLUX, bool FAST, int BZ, int BX, bool NTS = false>
__device__ __forceinline__ void advect_tile(const AdvArgs& a, int t) {
    constexpr int G = halo_of(S);
    constexpr int PZ = BZ + 2 * G;
    constexpr int PX = BX + 2 * G;
    constexpr int NF = (PX * PZ + NT - 1) / NT;        // f halo-tile elements per lane
    constexpr int NU = (BZ * (BX + 1) + NT - 1) / NT;  // x-faces per lane
    constexpr int NW = ((BZ + 1) * BX + NT - 1) / NT;  // z-faces per lane
    constexpr int NV = S == XC_TVDIM ? 3 : 1;          // velocities per face (TVD: left, own, right)
    __shared__ double fs[PX * PZ];
    __shared__ double qxs[(BX + 1) * BZ];
    __shared__ double qzs[BX * (BZ + 1)];

    const int z0 = (t % a.nbz) * BZ;
    const int x0 = (t / a.nbz) * BX;
    const int tid = threadIdx.x;
    const int Nz = a.Nz, Nx = a.Nx;
    const bool zper = a.zbc == XC_PERIODIC, xper = a.xbc == XC_PERIODIC;

    // ---- issue every load: f tile + halo (cross-shaped use; corners loaded for simplicity)
    double rf[NF];
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const int k = min(tid + i * NT, PX * PZ - 1);
        const int lz = k % PZ, lx = k / PZ;
        const int gz = min(z0 + lz - G, Nz + G - 1), gx = min(x0 + lx - G, Nx + G - 1);
        int hz, hx;
        if constexpr (HIN) {
            hz = gz + G;
            hx = gx + G;
        } else {
            hz = map_cell_sel(gz, Nz, zper);
            hx = map_cell_sel(gx, Nx, xper);
        }
        rf[i] = a.f[hz + (int64_t)hx * a.ld_f];
    }
    double ru[NU][NV], rw[NW][NV];
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int k = min(tid + i * NT, BZ * (BX + 1) - 1);
        const int iz = min(z0 + k % BZ, Nz - 1), F = min(x0 + k / BZ, Nx);
        if constexpr (S == XC_TVDIM) {
            if constexpr (HIN) {
                ru[i][0] = a.u[iz + (int64_t)F * a.ld_u];
                ru[i][1] = a.u[iz + (int64_t)(F + 1) * a.ld_u];
                ru[i][2] = a.u[iz + (int64_t)(F + 2) * a.ld_u];
            } else {
                ru[i][0] = a.u[iz + (int64_t)map_face(F - 1, Nx, a.xbc) * a.ld_u];
                ru[i][1] = a.u[iz + (int64_t)F * a.ld_u];
                ru[i][2] = a.u[iz + (int64_t)map_face(F + 1, Nx, a.xbc) * a.ld_u];
            }
        } else {
            ru[i][0] = a.u[iz + (int64_t)F * a.ld_u];
        }
    }
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int k = min(tid + i * NT, (BZ + 1) * BX - 1);
        const int F = min(z0 + k % (BZ + 1), Nz), ix = min(x0 + k / (BZ + 1), Nx - 1);
        if constexpr (S == XC_TVDIM) {
            if constexpr (HIN) {
                rw[i][0] = a.w[F + (int64_t)ix * a.ld_w];
                rw[i][1] = a.w[F + 1 + (int64_t)ix * a.ld_w];
                rw[i][2] = a.w[F + 2 + (int64_t)ix * a.ld_w];
            } else {
                rw[i][0] = a.w[map_face(F - 1, Nz, a.zbc) + (int64_t)ix * a.ld_w];
                rw[i][1] = a.w[F + (int64_t)ix * a.ld_w];
                rw[i][2] = a.w[map_face(F + 1, Nz, a.zbc) + (int64_t)ix * a.ld_w];
            }
        } else {
            rw[i][0] = a.w[F + (int64_t)ix * a.ld_w];
        }
    }
#pragma unroll
    for (int i = 0; i < NF; ++i)
        if (tid + i * NT < PX * PZ) fs[tid + i * NT] = rf[i];
    lds_barrier();

    const int qx_top = (zper && Nz > 1) ? Nz - 1 : 0;  // source row of qx[0, :]
    const int qx_bot = (zper && Nz > 1) ? 0 : Nz - 1;  // source row of qx[Nz+1, :]
    const int qz_lft = (xper && Nx > 1) ? Nx - 1 : 0;
    const int qz_rgt = (xper && Nx > 1) ? 0 : Nx - 1;
    constexpr int VI = S == XC_TVDIM ? 1 : 0;  // index of the face's own velocity

    // ---- x-faces: BZ rows x (BX+1) faces
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        const int k = tid + i * NT;
        if (k < BZ * (BX + 1)) {
            const int lz = k % BZ, fx = k / BZ;
            const int iz = z0 + lz, F = x0 + fx;
            double q = 0.0;
            if (iz < Nz && F <= Nx) {
                q = face_flux<S, FAST>(&fs[(fx + G) * PZ + lz + G], PZ, ru[i][0], ru[i][VI], ru[i][NV - 1]);
                if constexpr (FLUX) {
                    if (fx < BX || F == Nx) {
                        double* col = a.qx + (int64_t)F * a.ld_qx;
                        put<NTS>(&col[iz + 1], q);
                        if (iz == qx_top) put<NTS>(&col[0], q);
                        if (iz == qx_bot) put<NTS>(&col[Nz + 1], q);
                    }
                }
            }
            qxs[fx * BZ + lz] = q;
        }
    }
    // ---- z-faces: (BZ+1) faces x BX columns
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const int k = tid + i * NT;
        if (k < (BZ + 1) * BX) {
            const int fz = k % (BZ + 1), lx = k / (BZ + 1);
            const int F = z0 + fz, ix = x0 + lx;
            double q = 0.0;
            if (ix < Nx && F <= Nz) {
                q = face_flux<S, FAST>(&fs[(lx + G) * PZ + fz + G], 1, rw[i][0], rw[i][VI], rw[i][NV - 1]);
                if constexpr (FLUX) {
                    if (fz < BZ || F == Nz) {
                        put<NTS>(&a.qz[F + (int64_t)(ix + 1) * a.ld_qz], q);
                        if (ix == qz_lft) put<NTS>(&a.qz[F], q);
                        if (ix == qz_rgt) put<NTS>(&a.qz[F + (int64_t)(Nx + 1) * a.ld_qz], q);
                    }
                }
            }
            qzs[lx * (BZ + 1) + fz] = q;
        }
    }
    lds_barrier();

    // ---- divergence, advect.jl:94: (qxr - qxl + qzb - qzt) * invh
#pragma unroll
    for (int k = tid; k < BZ * BX; k += NT) {
        const int lz = k % BZ, lx = k / BZ;
        const int iz = z0 + lz, ix = x0 + lx;
        if (iz < Nz && ix < Nx) {
            const double qxr = qxs[(lx + 1) * BZ + lz], qxl = qxs[lx * BZ + lz];
            const double qzb = qzs[lx * (BZ + 1) + lz + 1], qzt = qzs[lx * (BZ + 1) + lz];
            put<NTS>(&a.adv[iz + (int64_t)ix * a.ld_adv], (qxr - qxl + qzb - qzt) * a.invh);
        }
    }
}

// Lane-row form of advect_tile for 64-row tiles: lane l of every wave owns tile row l and each wave
// takes whole columns (of the f halo tile, of x-faces, of z-faces, of cells), so column indexes and
// column addresses are wave-uniform (scalar registers), only the row offset is per lane, and every
// LDS access of a wave touches 64 consecutive doubles.  Same loads, face arithmetic and stores as
// advect_tile, hence bitwise the same results.  The tile is split in two parts shared by the one-tile
// kernel and the persistent one: LrVel holds the face velocities a wave needs (loaded into registers),
// lr_faces computes and stores the faces and the divergence from an f halo tile already in LDS.
template <int S, int BX>
struct LrShape {
    static constexpr int BZ = 64;
    static constexpr int G = halo_of(S);
    static constexpr int PZ = BZ + 2 * G;
    static constexpr int PX = BX + 2 * G;
    static constexpr int NWV = NT / 64;                  // waves per workgroup
    static constexpr int CF = (PX + NWV - 1) / NWV;      // halo-tile columns per wave
    static constexpr int CU = (BX + 1 + NWV - 1) / NWV;  // x-face columns per wave
    static constexpr int CW = (BX + NWV - 1) / NWV;      // z-face (and cell) columns per wave
    static constexpr int NV = S == XC_TVDIM ? 3 : 1;     // velocities per face (TVD: left, own, right)
    static constexpr int VI = S == XC_TVDIM ? 1 : 0;
    static constexpr int WC = 1 % NWV;                   // the wave that also takes the closing z-faces
    static_assert(BX <= 64 && NT % 64 == 0, "lane-row tiles");
};

template <int S, int BX>
struct LrVel {
    using L = LrShape<S, BX>;
    double ru[L::CU][L::NV];  // x-face velocities of the wave's x-face columns, row of the lane
    double rw[L::CW][L::NV];  // z-face velocities of the wave's cell columns, face row of the lane
    double rc[L::NV];         // closing z-faces (tile row BZ) of the columns x0 + l, lanes l < BX
};

// Velocity load; NTL: non-temporal (each velocity is read by one tile only, so the line need not stay cached)
template <bool NTL>
__device__ __forceinline__ double get(const double* p) {
    if constexpr (NTL) return __builtin_nontemporal_load(p);
    else return *p;
}

template <int S, bool HIN, int BX, bool NTL = false>
__device__ __forceinline__ void lr_load_vel(const AdvArgs& a, int z0, int x0, int l, int wv, LrVel<S, BX>& v) {
    using L = LrShape<S, BX>;
    constexpr int BZ = L::BZ, NWV = L::NWV;
    const int Nz = a.Nz, Nx = a.Nx;
    const int izc = min(z0 + l, Nz - 1);
#pragma unroll
    for (int j = 0; j < L::CU; ++j) {
        const int F = min(x0 + min(wv + j * NWV, BX), Nx);
        if constexpr (S == XC_TVDIM) {
            int c0 = F, c2 = F + 2, c1 = F + 1;
            if constexpr (!HIN) {
                c0 = map_face(F - 1, Nx, a.xbc);
                c1 = F;
                c2 = map_face(F + 1, Nx, a.xbc);
            }
            v.ru[j][0] = a.u[izc + (int64_t)c0 * a.ld_u];
            v.ru[j][1] = a.u[izc + (int64_t)c1 * a.ld_u];
            v.ru[j][2] = a.u[izc + (int64_t)c2 * a.ld_u];
        } else {
            v.ru[j][0] = get<NTL>(&a.u[izc + (int64_t)F * a.ld_u]);
        }
    }
    auto wload = [&](double* r, int F, int ix) {  // F: face row, 0..Nz; ix: cell column, 0..Nx-1
        const double* col = a.w + (int64_t)ix * a.ld_w;
        if constexpr (S == XC_TVDIM) {
            if constexpr (HIN) {
                r[0] = col[F];
                r[1] = col[F + 1];
                r[2] = col[F + 2];
            } else {
                r[0] = col[map_face(F - 1, Nz, a.zbc)];
                r[1] = col[F];
                r[2] = col[map_face(F + 1, Nz, a.zbc)];
            }
        } else {
            r[0] = get<NTL>(&col[F]);
        }
    };
#pragma unroll
    for (int j = 0; j < L::CW; ++j)
        wload(v.rw[j], min(z0 + l, Nz), min(x0 + min(wv + j * NWV, BX - 1), Nx - 1));
    if (wv == L::WC) wload(v.rc, min(z0 + BZ, Nz), min(x0 + min(l, BX - 1), Nx - 1));
}

// Faces and divergence of the tile at (z0, x0) from the f halo tile fs (PX columns of PZ rows, the
// BCs already applied) and the velocities v; qxs (BX + 1 columns of BZ) and qzc (BX) are workgroup
// LDS scratch.  Ends after the last LDS read of fs, qxs and qzc but before any barrier.
template <int S, bool FLUX, bool FAST, int BX, bool NTS>
__device__ __forceinline__ void lr_faces(const AdvArgs& a, int z0, int x0, int l, int wv, const double* fs,
                                         double* qxs, double* qzc, const LrVel<S, BX>& v) {
    using L = LrShape<S, BX>;
    constexpr int BZ = L::BZ, G = L::G, PZ = L::PZ, NWV = L::NWV, NV = L::NV, VI = L::VI;
    const int Nz = a.Nz, Nx = a.Nx;
    const bool zper = a.zbc == XC_PERIODIC, xper = a.xbc == XC_PERIODIC;
    const int qx_top = (zper && Nz > 1) ? Nz - 1 : 0;  // source row of qx[0, :]
    const int qx_bot = (zper && Nz > 1) ? 0 : Nz - 1;  // source row of qx[Nz+1, :]
    const int qz_lft = (xper && Nx > 1) ? Nx - 1 : 0;
    const int qz_rgt = (xper && Nx > 1) ? 0 : Nx - 1;
    const int iz = z0 + l;

    // ---- x-faces: column F = x0 + fx of the wave, row iz of the lane
#pragma unroll
    for (int j = 0; j < L::CU; ++j) {
        const int fx = wv + j * NWV;
        if (fx <= BX) {
            const int F = x0 + fx;
            double q = 0.0;
            if (F <= Nx && iz < Nz) {
                q = face_flux<S, FAST>(&fs[(fx + G) * PZ + l + G], PZ, v.ru[j][0], v.ru[j][VI], v.ru[j][NV - 1]);
                if constexpr (FLUX) {
                    if (fx < BX || F == Nx) {
                        double* col = a.qx + (int64_t)F * a.ld_qx;
                        put<NTS>(&col[iz + 1], q);
                        if (iz == qx_top) put<NTS>(&col[0], q);
                        if (iz == qx_bot) put<NTS>(&col[Nz + 1], q);
                    }
                }
            }
            qxs[fx * BZ + l] = q;
        }
    }
    // ---- z-faces: face row F = z0 + l of the wave's column ix
    auto zface = [&](const double* r, int lx, int fz) -> double {
        const int F = z0 + fz, ix = x0 + lx;
        double q = 0.0;
        if (ix < Nx && F <= Nz) {
            q = face_flux<S, FAST>(&fs[(lx + G) * PZ + fz + G], 1, r[0], r[VI], r[NV - 1]);
            if constexpr (FLUX) {
                if (fz < BZ || F == Nz) {
                    put<NTS>(&a.qz[F + (int64_t)(ix + 1) * a.ld_qz], q);
                    if (ix == qz_lft) put<NTS>(&a.qz[F], q);
                    if (ix == qz_rgt) put<NTS>(&a.qz[F + (int64_t)(Nx + 1) * a.ld_qz], q);
                }
            }
        }
        return q;
    };
    double zq[L::CW];
#pragma unroll
    for (int j = 0; j < L::CW; ++j) {
        const int lx = wv + j * NWV;
        zq[j] = lx < BX ? zface(v.rw[j], lx, l) : 0.0;
    }
    if (wv == L::WC && l < BX) qzc[l] = zface(v.rc, l, BZ);
    lds_barrier();

    // ---- divergence, advect.jl:94: (qxr - qxl + qzb - qzt) * invh; the face below row l is lane
    // l + 1's z-face (a wave shift), below row 63 the closing face
#pragma unroll
    for (int j = 0; j < L::CW; ++j) {
        const int lx = wv + j * NWV;
        const int ix = x0 + lx;
        const double below = lane_next(zq[j]);
        if (lx < BX && ix < Nx && iz < Nz) {
            const double qxr = qxs[(lx + 1) * BZ + l], qxl = qxs[lx * BZ + l];
            const double qzb = l == 63 ? qzc[lx] : below, qzt = zq[j];
            put<NTS>(&a.adv[iz + (int64_t)ix * a.ld_adv], (qxr - qxl + qzb - qzt) * a.invh);
        }
    }
}

template <int S, bool HIN, bool FLUX, bool FAST, int BX, bool NTS = false>
__device__ __forceinline__ void advect_tile_lr(const AdvArgs& a, int t) {
    using L = LrShape<S, BX>;
    constexpr int BZ = L::BZ, G = L::G, PZ = L::PZ, PX = L::PX, NWV = L::NWV, CF = L::CF;
    __shared__ double fs[PX * PZ];
    __shared__ double qxs[(BX + 1) * BZ];
    __shared__ double qzc[BX];  // closing z-faces; the other z-faces stay in registers

    const int l = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int z0 = (t % a.nbz) * BZ, x0 = (t / a.nbz) * BX;
    const int Nz = a.Nz, Nx = a.Nx;
    const bool zper = a.zbc == XC_PERIODIC, xper = a.xbc == XC_PERIODIC;

    // ---- issue every load of the tile before the first LDS store
    // halo-tile rows of this lane: l - G, and (lanes l < 2G) BZ - G + l
    const int gzA = min(z0 + l - G, Nz + G - 1);
    const int gzB = min(z0 + BZ - G + min(l, 2 * G - 1), Nz + G - 1);
    int hzA, hzB;
    if constexpr (HIN) {
        hzA = gzA + G;
        hzB = gzB + G;
    } else {
        hzA = map_cell_sel(gzA, Nz, zper);
        hzB = map_cell_sel(gzB, Nz, zper);
    }
    double rf[CF][2];
#pragma unroll
    for (int j = 0; j < CF; ++j) {
        const int c = min(wv + j * NWV, PX - 1);  // past PX: a duplicate, never stored
        const int gx = min(x0 + c - G, Nx + G - 1);
        int hx;
        if constexpr (HIN) hx = gx + G;
        else hx = map_cell_sel(gx, Nx, xper);
        const double* col = a.f + (int64_t)hx * a.ld_f;
        // NTS launches: the columns no neighbouring tile reads (more than G from either edge of the tile)
        // are loaded non-temporally, so they do not displace the shared halo columns from the cache
        if (NTS && c >= 2 * G && c < PX - 2 * G) {
            rf[j][0] = get<true>(&col[hzA]);
            rf[j][1] = get<true>(&col[hzB]);
        } else {
            rf[j][0] = col[hzA];
            rf[j][1] = col[hzB];
        }
    }
    LrVel<S, BX> v;
    lr_load_vel<S, HIN, BX, NTS>(a, z0, x0, l, wv, v);

#pragma unroll
    for (int j = 0; j < CF; ++j) {
        const int c = wv + j * NWV;
        if (c < PX) {
            fs[c * PZ + l] = rf[j][0];
            if (l < 2 * G) fs[c * PZ + BZ + l] = rf[j][1];
        }
    }
    lds_barrier();
    lr_faces<S, FLUX, FAST, BX, NTS>(a, z0, x0, l, wv, fs, qxs, qzc, v);
}

// the tile body a kernel uses: lane-row form for 64-row tiles, the generic form otherwise
template <int S, bool HIN, bool FLUX, bool FAST, int BZ, int BX, bool NTS = false>
__device__ __forceinline__ void advect_tile_any(const AdvArgs& a, int t) {
    if constexpr (BZ == 64) advect_tile_lr<S, HIN, FLUX, FAST, BX, NTS>(a, t);
    else advect_tile<S, HIN, FLUX, FAST, BZ, BX, NTS>(a, t);
}

typedef __attribute__((address_space(3))) char lds_char_t;

// LDS DMA of one dword a lane: lane k of the wave copies the dword at gbase + voff (its own offset) to
// LDS byte lds + 4 k.  Written as inline asm so that the compiler's wait-count insertion does not treat
// it as a pending write to every LDS array (with the builtin it waits for the DMA before each LDS read,
// including reads of the other buffer); the caller waits for it explicitly (vmcnt) before reading the
// destination.  M0 (the LDS base) is saved and restored around it.
__device__ __forceinline__ void dma_dword(const char* gbase, unsigned voff, unsigned lds) {
    unsigned saved;
    asm volatile("s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\tglobal_load_lds_dword %2, %3\n\ts_mov_b32 m0, %0"
                 : "=&s"(saved) : "s"(lds), "v"(voff), "s"(gbase) : "memory");
}
__device__ __forceinline__ unsigned lds_offset(void* p) {
    return (unsigned)(uintptr_t)(lds_char_t*)p;
}

// Persistent, software-pipelined lane-row form (centred input, schemes without padded velocities).
// Each workgroup walks its XCD's tiles.  The f halo tile of the next tile is copied from global memory
// straight into the other half of a double-buffered LDS tile by the LDS DMA path (global_load_lds:
// lane k of a wave writes dword k after the wave-uniform LDS base, no VGPRs involved), and the next
// tile's face velocities are loaded into a second register set, both issued before the current tile's
// reconstructions: the memory system streams the next tile while the arithmetic of this one runs.
// A halo column of PZ doubles is 2 PZ dwords, ND dword-DMA instructions of 64 lanes; each lane's
// address is its row's mapped source (the BCs as an index map, at 4-byte granularity), so clamped and
// wrapped halo rows need no fix-up.  Same face arithmetic (lr_faces) as k_advect: bitwise the same.
template <int S, bool FLUX, bool FAST, int BX, bool NTS>
__global__ __launch_bounds__(NT) void k_advect_dma(AdvArgs a) {
    static_assert(S != XC_TVDIM, "tvdim needs its padded velocities");
    using L = LrShape<S, BX>;
    constexpr int BZ = L::BZ, G = L::G, PZ = L::PZ, PX = L::PX, NWV = L::NWV, CF = L::CF;
    constexpr int ND = (2 * PZ + 63) / 64;
    __shared__ double fs[2][PX * PZ];
    __shared__ double qxs[(BX + 1) * BZ];
    __shared__ double qzc[BX];
    const int l = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int Nz = a.Nz, Nx = a.Nx;
    const bool zper = a.zbc == XC_PERIODIC, xper = a.xbc == XC_PERIODIC;

    auto dma_f = [&](int t, int b) {
        const int z0 = (t % a.nbz) * BZ, x0 = (t / a.nbz) * BX;
        int off[ND];  // byte offset in a column of the dword this lane copies, per DMA instruction
#pragma unroll
        for (int k = 0; k < ND; ++k) {
            const int d = 64 * k + l;
            const int r = min(d >> 1, PZ - 1);
            const int gz = min(z0 + r - G, Nz + G - 1);
            off[k] = map_cell_sel(gz, Nz, zper) * 8 + (d & 1) * 4;
        }
#pragma unroll
        for (int j = 0; j < CF; ++j) {
            const int c = wv + j * NWV;
            if (c < PX) {
                const int gx = min(x0 + c - G, Nx + G - 1);
                const char* col = reinterpret_cast<const char*>(a.f + (int64_t)map_cell_sel(gx, Nx, xper) * a.ld_f);
                const unsigned dst = lds_offset(&fs[b][c * PZ]);
#pragma unroll
                for (int k = 0; k < ND; ++k)
                    if (64 * k + l < 2 * PZ)  // lanes past the column's last dword stay off
                        dma_dword(col, (unsigned)off[k], dst + 256 * k);
            }
        }
    };
    auto tile_of = [&](int it) { return xcd_tile(it, a.nb); };

    int it = blockIdx.x;
    if (it >= a.nb) return;
    LrVel<S, BX> v;
    {
        const int t = tile_of(it);
        dma_f(t, 0);
        lr_load_vel<S, false, BX, NTS>(a, (t % a.nbz) * BZ, (t / a.nbz) * BX, l, wv, v);
    }
    int b = 0;
    for (; it < a.nb; it += gridDim.x) {
        const int t = tile_of(it);
        const int z0 = (t % a.nbz) * BZ, x0 = (t / a.nbz) * BX;
        // this tile's halo has landed in fs[b] (every wave's share: wait, then meet), and every wave is
        // done with fs[b ^ 1], qxs and qzc of the previous tile
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) expcnt(7) lgkmcnt(0)
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        LrVel<S, BX> vn;
        const int itn = it + (int)gridDim.x;
        if (itn < a.nb) {
            const int tn = tile_of(itn);
            dma_f(tn, b ^ 1);
            lr_load_vel<S, false, BX, NTS>(a, (tn % a.nbz) * BZ, (tn / a.nbz) * BX, l, wv, vn);
        }
        lr_faces<S, FLUX, FAST, BX, NTS>(a, z0, x0, l, wv, fs[b], qxs, qzc, v);
        v = vn;
        b ^= 1;
    }
}

template <int S, bool HIN, bool FLUX, bool FAST, int BZ, int BX, bool NTS>
__global__ __launch_bounds__(NT) void k_advect(AdvArgs a) {
    advect_tile_any<S, HIN, FLUX, FAST, BZ, BX, NTS>(a, xcd_tile(blockIdx.x, a.nb));
}

// Ensemble form: the tiles of n grids of any sizes in one launch.  Workgroup b (XCD-remapped)
// belongs to the grid whose tile range [first[i], first[i+1]) holds it.
template <int S, bool FLUX, bool FAST, int BZ, int BX>
__global__ __launch_bounds__(NT) void k_advect_batch(const AdvArgs* __restrict__ items,
                                                     const int* __restrict__ first, int n, int total) {
    const int b = xcd_tile(blockIdx.x, total);
    int lo = 0, hi = n - 1;
    while (lo < hi) {  // last i with first[i] <= b
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const AdvArgs a = items[lo];
    advect_tile_any<S, false, FLUX, FAST, BZ, BX>(a, b - first[lo]);
}

// Two grids' tiles in one launch, their arguments by value (phsevo!'s advections of X and M, fluidmech!'s
// momentum advections of rho W and rho U): workgroup b takes tile b of the first grid or tile b - a0.nb of the
// second; the same tile body as k_advect, hence the same results.
template <int S, bool FLUX, bool FAST, int BZ, int BX>
__global__ __launch_bounds__(NT) void k_advect_pair(AdvArgs a0, AdvArgs a1) {
    const int b = xcd_tile(blockIdx.x, a0.nb + a1.nb);
    if (b < a0.nb) advect_tile_any<S, false, FLUX, FAST, BZ, BX>(a0, b);
    else advect_tile_any<S, false, FLUX, FAST, BZ, BX>(a1, b - a0.nb);
}
template <int S, bool FLUX>
void launch_advect_pair_t(const AdvArgs& a0, const AdvArgs& a1, bool exact, hipStream_t st) {
    constexpr int BZ = VAR_TZ[2], BX = VAR_TX[2];
    const dim3 grid((unsigned)(a0.nb + a1.nb));
    if constexpr (S == XC_WENO5) {
        if (!exact) {
            hipLaunchKernelGGL((k_advect_pair<S, FLUX, true, BZ, BX>), grid, dim3(NT), 0, st, a0, a1);
            return;
        }
    }
    hipLaunchKernelGGL((k_advect_pair<S, FLUX, false, BZ, BX>), grid, dim3(NT), 0, st, a0, a1);
}
template <bool FLUX>
void launch_advect_pair_s(int scheme, const AdvArgs& a0, const AdvArgs& a1, bool exact, hipStream_t st) {
    switch (scheme) {
        case XC_CENTR: launch_advect_pair_t<XC_CENTR, FLUX>(a0, a1, exact, st); break;
        case XC_UPWD1: launch_advect_pair_t<XC_UPWD1, FLUX>(a0, a1, exact, st); break;
        case XC_QUICK: launch_advect_pair_t<XC_QUICK, FLUX>(a0, a1, exact, st); break;
        case XC_FROMM: launch_advect_pair_t<XC_FROMM, FLUX>(a0, a1, exact, st); break;
        case XC_WENO3: launch_advect_pair_t<XC_WENO3, FLUX>(a0, a1, exact, st); break;
        case XC_WENO5: launch_advect_pair_t<XC_WENO5, FLUX>(a0, a1, exact, st); break;
        default: launch_advect_pair_t<XC_TVDIM, FLUX>(a0, a1, exact, st); break;
    }
}

template <int S, bool HIN, bool FLUX, bool FAST, int V>
hipError_t launch_advect_v(const AdvArgs& a0, hipStream_t s) {
    constexpr int BZ = VAR_TZ[V], BX = VAR_TX[V];
    AdvArgs a = a0;
    a.nbz = (a.Nz + BZ - 1) / BZ;
    a.nb = a.nbz * ((a.Nx + BX - 1) / BX);
    if (!HIN && a.nts) hipLaunchKernelGGL((k_advect<S, HIN, FLUX, FAST, BZ, BX, !HIN>), dim3(a.nb), dim3(NT), 0, s, a);
    else hipLaunchKernelGGL((k_advect<S, HIN, FLUX, FAST, BZ, BX, false>), dim3(a.nb), dim3(NT), 0, s, a);
    return hipGetLastError();
}

// persistent grid: every resident workgroup slot of the device once (a multiple of the 8 XCDs)
template <typename K>
int pipe_grid(K kernel, int nb) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 8)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, NT, 0) != hipSuccess || per < 1) per = 4;
    const int g = ((per * cus) / 8) * 8;
    return nb < g ? nb : (g > 0 ? g : nb);
}

template <int S, bool FLUX, bool FAST, int BX>
hipError_t launch_advect_dma(const AdvArgs& a0, hipStream_t s) {
    AdvArgs a = a0;
    a.nbz = (a.Nz + 63) / 64;
    a.nb = a.nbz * ((a.Nx + BX - 1) / BX);
    if (a.nts) {
        auto kern = k_advect_dma<S, FLUX, FAST, BX, true>;
        hipLaunchKernelGGL(kern, dim3(pipe_grid(kern, a.nb)), dim3(NT), 0, s, a);
    } else {
        auto kern = k_advect_dma<S, FLUX, FAST, BX, false>;
        hipLaunchKernelGGL(kern, dim3(pipe_grid(kern, a.nb)), dim3(NT), 0, s, a);
    }
    return hipGetLastError();
}

template <int S, bool HIN, bool FLUX, bool FAST>
hipError_t launch_advect_f(const AdvArgs& a, hipStream_t s, int v) {
    if constexpr (HIN) {  // halo'd-input entry points keep the default tile
        return launch_advect_v<S, HIN, FLUX, FAST, 0>(a, s);
    } else {
        if constexpr (S != XC_TVDIM) {
            if (v >= VAR_PIPE) return v == VAR_PIPE ? launch_advect_dma<S, FLUX, FAST, VAR_TX[VAR_PIPE]>(a, s)
                                                    : launch_advect_dma<S, FLUX, FAST, VAR_TX[VAR_PIPE + 1]>(a, s);
        }
        switch (v) {
            case 1: return launch_advect_v<S, HIN, FLUX, FAST, 1>(a, s);
            case 2: return launch_advect_v<S, HIN, FLUX, FAST, 2>(a, s);
            case 3: return launch_advect_v<S, HIN, FLUX, FAST, 3>(a, s);
            default: return launch_advect_v<S, HIN, FLUX, FAST, 0>(a, s);
        }
    }
}

template <int S, bool HIN, bool FLUX>
hipError_t launch_advect_t(const AdvArgs& a, hipStream_t s, bool exact, int v) {
    // only WENO5 has a distinct fast form; every other scheme is the reference arithmetic
    if constexpr (S == XC_WENO5) {
        if (!exact) return launch_advect_f<S, HIN, FLUX, true>(a, s, v);
    }
    return launch_advect_f<S, HIN, FLUX, false>(a, s, v);
}

template <bool HIN, bool FLUX>
hipError_t launch_advect_s(int scheme, const AdvArgs& a, hipStream_t s, bool exact, int v) {
    switch (scheme) {
        case XC_CENTR: return launch_advect_t<XC_CENTR, HIN, FLUX>(a, s, exact, v);
        case XC_UPWD1: return launch_advect_t<XC_UPWD1, HIN, FLUX>(a, s, exact, v);
        case XC_QUICK: return launch_advect_t<XC_QUICK, HIN, FLUX>(a, s, exact, v);
        case XC_FROMM: return launch_advect_t<XC_FROMM, HIN, FLUX>(a, s, exact, v);
        case XC_WENO3: return launch_advect_t<XC_WENO3, HIN, FLUX>(a, s, exact, v);
        case XC_WENO5: return launch_advect_t<XC_WENO5, HIN, FLUX>(a, s, exact, v);
        default: return launch_advect_t<XC_TVDIM, HIN, FLUX>(a, s, exact, v);
    }
}

// ---- tile autotuner: per (scheme, fluxes, arithmetic, grid) the fastest tile shape, measured once
struct TuneEntry { int scheme, flux, exact, Nz, Nx, v; float us[NVAR]; };
struct TuneCtx { std::vector<TuneEntry> e; };
void tune_free(xc_ctx*, void* p) { delete (TuneCtx*)p; }
TuneCtx* tune_ctx(xc_ctx* ctx) {
    TuneCtx* t = (TuneCtx*)ctx->ext[XC_EXT_TUNE];
    if (!t) {
        t = new TuneCtx();
        ctx->ext[XC_EXT_TUNE] = t;
        ctx->ext_free[XC_EXT_TUNE] = tune_free;
    }
    return t;
}
TuneEntry* tune_find(xc_ctx* ctx, int scheme, int flux, int exact, int Nz, int Nx) {
    for (auto& e : tune_ctx(ctx)->e)
        if (e.scheme == scheme && e.flux == flux && e.exact == exact && e.Nz == Nz && e.Nx == Nx) return &e;
    return nullptr;
}
// time every tile variant (3 launches each, HIP events on the context's stream) and keep the fastest;
// every variant computes the same values, so the outputs are those of the last launch either way
int tune_advect(xc_ctx* ctx, int scheme, bool flux, bool ex, const AdvArgs& a, int* out) {
    TuneEntry t{scheme, flux ? 1 : 0, ex ? 1 : 0, a.Nz, a.Nx, 0, {}};
    hipEvent_t e0, e1;
    XC_TRY_HIP(ctx, hipEventCreate(&e0));
    XC_TRY_HIP(ctx, hipEventCreate(&e1));
    float best = 1e30f;
    for (int v = 0; v < NVAR; ++v) {
        if (v >= VAR_PIPE && scheme == XC_TVDIM) { t.us[v] = 1e30f; continue; }
        hipError_t e = hipSuccess;
        for (int r = 0; r < 4 && e == hipSuccess; ++r) {  // first launch warms up
            if (r == 1) (void)hipEventRecord(e0, ctx->stream);
            e = flux ? launch_advect_s<false, true>(scheme, a, ctx->stream, ex, v)
                     : launch_advect_s<false, false>(scheme, a, ctx->stream, ex, v);
        }
        (void)hipEventRecord(e1, ctx->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.us[v] = e == hipSuccess ? 1e3f * ms / 3.f : 1e30f;
        if (t.us[v] < best) { best = t.us[v]; t.v = v; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    tune_ctx(ctx)->e.push_back(t);
    *out = t.v;
    return XC_OK;
}

// ensemble launches use the 64 x 8 tile (variant 2)
template <int S, bool FLUX, bool FAST>
void launch_advect_batch_t(const AdvArgs* items, const int* first, int n, int total, hipStream_t s) {
    hipLaunchKernelGGL((k_advect_batch<S, FLUX, FAST, VAR_TZ[2], VAR_TX[2]>), dim3(total), dim3(NT), 0, s, items,
                       first, n, total);
}
template <bool FLUX>
void launch_advect_batch_s(int scheme, bool exact, const AdvArgs* items, const int* first, int n, int total,
                           hipStream_t s) {
    switch (scheme) {
        case XC_CENTR: launch_advect_batch_t<XC_CENTR, FLUX, false>(items, first, n, total, s); break;
        case XC_UPWD1: launch_advect_batch_t<XC_UPWD1, FLUX, false>(items, first, n, total, s); break;
        case XC_QUICK: launch_advect_batch_t<XC_QUICK, FLUX, false>(items, first, n, total, s); break;
        case XC_FROMM: launch_advect_batch_t<XC_FROMM, FLUX, false>(items, first, n, total, s); break;
        case XC_WENO3: launch_advect_batch_t<XC_WENO3, FLUX, false>(items, first, n, total, s); break;
        case XC_WENO5:
            if (exact) launch_advect_batch_t<XC_WENO5, FLUX, false>(items, first, n, total, s);
            else launch_advect_batch_t<XC_WENO5, FLUX, true>(items, first, n, total, s);
            break;
        default: launch_advect_batch_t<XC_TVDIM, FLUX, false>(items, first, n, total, s); break;
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
    a.invh = 1.0 / h;
    a.xbc = xbc; a.zbc = zbc;
    a.nts = out_bytes(a.Nz, a.Nx, flux) > NT_STORE_BYTES;
    hipError_t e;
    const bool ex = (ctx->flags & XC_CTX_EXACT) != 0;
    int v = 0;
    if (!hin) {
        const TuneEntry* te = tune_find(ctx, scheme, flux, ex, a.Nz, a.Nx);
        if (te) v = te->v;
        else if (ctx->flags & XC_CTX_AUTOTUNE) {
            int st;
            if ((st = tune_advect(ctx, scheme, flux, ex, a, &v))) return st;
        }
    }
    if (hin) e = flux ? launch_advect_s<true, true>(scheme, a, ctx->stream, ex, 0) : launch_advect_s<true, false>(scheme, a, ctx->stream, ex, 0);
    else     e = flux ? launch_advect_s<false, true>(scheme, a, ctx->stream, ex, v) : launch_advect_s<false, false>(scheme, a, ctx->stream, ex, v);
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
    int nts;  // host side: launch the non-temporal-store form (outputs larger than NT_STORE_BYTES)
};

template <bool HIN, bool FLUX>
__global__ __launch_bounds__(NT) void k_diffus(DifArgs a) {
    constexpr bool NTS = false;
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
    put<NTS>(&a.dff[iz + (int64_t)ix * a.ld_d], 0.5 * a.invh2 * (flux_xr - flux_xl + flux_zb - flux_zt));
    if constexpr (FLUX) {
        // diffus.jl:77 / 89: q = -(kL + kR) * 1/2 * (fR - fL) / h
        const int qx_top = (a.zbc == XC_PERIODIC && Nz > 1) ? Nz - 1 : 0;
        const int qx_bot = (a.zbc == XC_PERIODIC && Nz > 1) ? 0 : Nz - 1;
        const int qz_lft = (a.xbc == XC_PERIODIC && Nx > 1) ? Nx - 1 : 0;
        const int qz_rgt = (a.xbc == XC_PERIODIC && Nx > 1) ? 0 : Nx - 1;
        auto put_qx = [&](int F, double q) {
            double* col = a.qx + (int64_t)F * a.ld_qx;
            put<NTS>(&col[iz + 1], q);
            if (iz == qx_top) put<NTS>(&col[0], q);
            if (iz == qx_bot) put<NTS>(&col[Nz + 1], q);
        };
        auto put_qz = [&](int F, double q) {
            put<NTS>(&a.qz[F + (int64_t)(ix + 1) * a.ld_qz], q);
            if (ix == qz_lft) put<NTS>(&a.qz[F], q);
            if (ix == qz_rgt) put<NTS>(&a.qz[F + (int64_t)(Nx + 1) * a.ld_qz], q);
        };
        put_qx(ix, -(kmx + kc) * 0.5 * (fc - fmx) * a.invh);
        if (ix == Nx - 1) put_qx(Nx, -(kc + kpx) * 0.5 * (fpx - fc) * a.invh);
        put_qz(iz, -(kmz + kc) * 0.5 * (fc - fmz) * a.invh);
        if (iz == Nz - 1) put_qz(Nz, -(kc + kpz) * 0.5 * (fpz - fc) * a.invh);
    }
}

// LDS-tiled form of k_diffus for the centred entry points (f, k are (Nz, Nx) and the BCs are an index
// map): the 64 x 16 tile of f and of the diffusivity (k, or k .* k2 formed once per cell) plus a
// one-cell halo is staged once, so each cell costs two global loads instead of ten.  Same arithmetic
// and operation order as k_diffus.
template <bool FLUX, bool NTS = false>
__device__ __forceinline__ void diffus_tile(const DifArgs& a, int t) {
    constexpr int PZ = TZ + 2, PX = TX + 2;
    __shared__ double fs[PX * PZ];
    __shared__ double ks[PX * PZ];
    constexpr int NF = (PX * PZ + NT - 1) / NT;
    const int z0 = (t % a.nbz) * TZ;
    const int x0 = (t / a.nbz) * TX;
    const int Nz = a.Nz, Nx = a.Nx;
    const bool zper = a.zbc == XC_PERIODIC, xper = a.xbc == XC_PERIODIC;
    // every load of the tile is issued before the first LDS store (see advect_tile)
    double rf[NF], rk[NF], rk2[NF];
    const bool has_k2 = a.k2 != nullptr;
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const int k = min((int)threadIdx.x + i * NT, PX * PZ - 1);
        const int lz = k % PZ, lx = k / PZ;
        // rows/columns past the grid end map like the one after the last
        const int gz = min(z0 + lz - 1, Nz), gx = min(x0 + lx - 1, Nx);
        const int64_t hz = map_cell_sel(gz, Nz, zper), hx = map_cell_sel(gx, Nx, xper);
        if (NTS && lx >= 2 && lx < PX - 2) {  // columns no neighbouring tile reads: non-temporal
            rf[i] = get<true>(&a.f[hz + hx * a.ld_f]);
            rk[i] = get<true>(&a.k[hz + hx * a.ld_k]);
            if (has_k2) rk2[i] = get<true>(&a.k2[hz + hx * a.ld_k2]);
        } else {
            rf[i] = a.f[hz + hx * a.ld_f];
            rk[i] = a.k[hz + hx * a.ld_k];
            if (has_k2) rk2[i] = a.k2[hz + hx * a.ld_k2];
        }
    }
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const int k = threadIdx.x + i * NT;
        if (k < PX * PZ) {
            fs[k] = rf[i];
            ks[k] = has_k2 ? rk[i] * rk2[i] : rk[i];
        }
    }
    lds_barrier();
    const int qx_top = (a.zbc == XC_PERIODIC && Nz > 1) ? Nz - 1 : 0;
    const int qx_bot = (a.zbc == XC_PERIODIC && Nz > 1) ? 0 : Nz - 1;
    const int qz_lft = (a.xbc == XC_PERIODIC && Nx > 1) ? Nx - 1 : 0;
    const int qz_rgt = (a.xbc == XC_PERIODIC && Nx > 1) ? 0 : Nx - 1;
#pragma unroll
    for (int k = threadIdx.x; k < TZ * TX; k += NT) {
        const int lz = k % TZ, lx = k / TZ;
        const int iz = z0 + lz, ix = x0 + lx;
        if (iz >= Nz || ix >= Nx) continue;
        const int c = (lx + 1) * PZ + lz + 1;
        const double fc = fs[c], fmx = fs[c - PZ], fpx = fs[c + PZ], fmz = fs[c - 1], fpz = fs[c + 1];
        const double kc = ks[c], kmx = ks[c - PZ], kpx = ks[c + PZ], kmz = ks[c - 1], kpz = ks[c + 1];
        const double flux_xr = (kpx + kc) * (fpx - fc);
        const double flux_xl = (kc + kmx) * (fc - fmx);
        const double flux_zb = (kpz + kc) * (fpz - fc);
        const double flux_zt = (kc + kmz) * (fc - fmz);
        put<NTS>(&a.dff[iz + (int64_t)ix * a.ld_d], 0.5 * a.invh2 * (flux_xr - flux_xl + flux_zb - flux_zt));
        if constexpr (FLUX) {
            auto put_qx = [&](int F, double q) {
                double* col = a.qx + (int64_t)F * a.ld_qx;
                put<NTS>(&col[iz + 1], q);
                if (iz == qx_top) put<NTS>(&col[0], q);
                if (iz == qx_bot) put<NTS>(&col[Nz + 1], q);
            };
            auto put_qz = [&](int F, double q) {
                put<NTS>(&a.qz[F + (int64_t)(ix + 1) * a.ld_qz], q);
                if (ix == qz_lft) put<NTS>(&a.qz[F], q);
                if (ix == qz_rgt) put<NTS>(&a.qz[F + (int64_t)(Nx + 1) * a.ld_qz], q);
            };
            put_qx(ix, -(kmx + kc) * 0.5 * (fc - fmx) * a.invh);
            if (ix == Nx - 1) put_qx(Nx, -(kc + kpx) * 0.5 * (fpx - fc) * a.invh);
            put_qz(iz, -(kmz + kc) * 0.5 * (fc - fmz) * a.invh);
            if (iz == Nz - 1) put_qz(Nz, -(kc + kpz) * 0.5 * (fpz - fc) * a.invh);
        }
    }
}
template <bool FLUX, bool NTS>
__global__ __launch_bounds__(NT) void k_diffus_lds(DifArgs a) {
    diffus_tile<FLUX, NTS>(a, xcd_tile(blockIdx.x, a.nb));
}
template <bool FLUX>
__global__ __launch_bounds__(NT) void k_diffus_batch(const DifArgs* __restrict__ items, const int* __restrict__ first,
                                                     int n, int total) {
    const int b = xcd_tile(blockIdx.x, total);
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    const DifArgs a = items[lo];
    diffus_tile<FLUX>(a, b - first[lo]);
}

// ---- ensemble tables: per-grid argument blocks staged in pinned memory and copied to the device
struct BatchCtx {
    char* host = nullptr;   // pinned
    char* dev = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;  // the last copy out of `host` has completed
};
void batch_free(xc_ctx*, void* p) {
    BatchCtx* b = (BatchCtx*)p;
    if (b->done) (void)hipEventSynchronize(b->done), (void)hipEventDestroy(b->done);
    if (b->host) (void)hipHostFree(b->host);
    if (b->dev) (void)hipFree(b->dev);
    delete b;
}
// room for `bytes` of table; returns host staging and device pointers (the previous copy has finished)
int batch_table(xc_ctx* ctx, size_t bytes, char** host, char** dev) {
    BatchCtx* b = (BatchCtx*)ctx->ext[XC_EXT_BATCH];
    if (!b) {
        b = new BatchCtx();
        ctx->ext[XC_EXT_BATCH] = b;
        ctx->ext_free[XC_EXT_BATCH] = batch_free;
        XC_TRY_HIP(ctx, hipEventCreateWithFlags(&b->done, hipEventDisableTiming));
    }
    XC_TRY_HIP(ctx, hipEventSynchronize(b->done));
    if (bytes > b->cap) {
        if (b->host) (void)hipHostFree(b->host);
        if (b->dev) (void)hipFree(b->dev);
        b->host = nullptr;
        b->dev = nullptr;
        b->cap = 0;
        XC_TRY_HIP(ctx, hipHostMalloc((void**)&b->host, bytes, hipHostMallocDefault));
        XC_TRY_HIP(ctx, hipMalloc((void**)&b->dev, bytes));
        b->cap = bytes;
    }
    *host = b->host;
    *dev = b->dev;
    return XC_OK;
}
int batch_upload(xc_ctx* ctx, size_t bytes) {
    BatchCtx* b = (BatchCtx*)ctx->ext[XC_EXT_BATCH];
    XC_TRY_HIP(ctx, hipMemcpyAsync(b->dev, b->host, bytes, hipMemcpyHostToDevice, ctx->stream));
    XC_TRY_HIP(ctx, hipEventRecord(b->done, ctx->stream));
    return XC_OK;
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
    a.nts = out_bytes(a.Nz, a.Nx, flux) > NT_STORE_BYTES;
    if (hin) {
        if (flux) hipLaunchKernelGGL((k_diffus<true, true>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else hipLaunchKernelGGL((k_diffus<true, false>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
    } else {
        a.nbz = (int)((Nz + TZ - 1) / TZ);
        a.nb = a.nbz * (int)((Nx + TX - 1) / TX);
        if (flux && a.nts) hipLaunchKernelGGL((k_diffus_lds<true, true>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else if (flux) hipLaunchKernelGGL((k_diffus_lds<true, false>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else if (a.nts) hipLaunchKernelGGL((k_diffus_lds<false, true>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
        else hipLaunchKernelGGL((k_diffus_lds<false, false>), dim3(a.nb), dim3(NT), 0, ctx->stream, a);
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

int xc_advect_tile(xc_ctx* ctx, int scheme, int with_flux, int64_t Nz, int64_t Nx, int* tz, int* tx, float* us) {
    XC_CHECK_CTX(ctx);
    if (!tz || !tx) return xc_fail(ctx, XC_EINVAL_ARG, "advect_tile: NULL output");
    const TuneEntry* te = tune_find(ctx, scheme, with_flux ? 1 : 0, (ctx->flags & XC_CTX_EXACT) ? 1 : 0, (int)Nz, (int)Nx);
    const int v = te ? te->v : 0;
    *tz = VAR_TZ[v];
    *tx = v >= VAR_PIPE ? -VAR_TX[v] : VAR_TX[v];
    if (us)
        for (int k = 0; k < NVAR; ++k) us[k] = te ? te->us[k] : 0.f;
    return te ? XC_OK : XC_ENOCONV;
}

int xc_advect_set_tile(xc_ctx* ctx, int scheme, int with_flux, int64_t Nz, int64_t Nx, int tz, int tx) {
    XC_CHECK_CTX(ctx);
    int v = -1;
    for (int k = 0; k < NVAR; ++k)
        if (VAR_TZ[k] == tz && (k < VAR_PIPE ? VAR_TX[k] : -VAR_TX[k]) == tx && (k < VAR_PIPE || scheme != XC_TVDIM)) v = k;
    if (v < 0) return xc_fail(ctx, XC_EINVAL_ARG, "advect_set_tile: no %d x %d tile", tz, tx);
    const int ex = (ctx->flags & XC_CTX_EXACT) ? 1 : 0;
    TuneEntry* te = tune_find(ctx, scheme, with_flux ? 1 : 0, ex, (int)Nz, (int)Nx);
    if (te) te->v = v;
    else tune_ctx(ctx)->e.push_back(TuneEntry{scheme, with_flux ? 1 : 0, ex, (int)Nz, (int)Nx, v, {}});
    return XC_OK;
}

int xc_advect_centered_pair(xc_ctx* ctx, int scheme, const xc_adv_member* m, int xbc, int zbc, const char* name) {
    XC_CHECK_CTX(ctx);
    const int g = xc_scheme_halo(scheme);
    if (g < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "%s: unsupported scheme = %d", name, scheme);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported xBC = %d", name, xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "%s: unsupported zBC = %d", name, zbc);
    const bool flux = m[0].qx.p != nullptr;
    AdvArgs a[2];
    for (int i = 0; i < 2; ++i) {
        const xc_adv_member& q = m[i];
        const int64_t Nz = q.adv.rows, Nx = q.adv.cols;
        if (!xc_mat_ok(q.adv) || Nz < g || Nx < g || Nz > (1 << 30) || Nx > (1 << 30) || !xc_shape(q.f, Nz, Nx) ||
            !xc_shape(q.u, Nz, Nx + 1) || !xc_shape(q.w, Nz + 1, Nx))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: advection %d has inconsistent shapes", name, i);
        if ((q.qx.p != nullptr) != flux || (q.qz.p != nullptr) != flux ||
            (flux && (!xc_shape(q.qx, Nz + 2, Nx + 1) || !xc_shape(q.qz, Nz + 1, Nx + 2))))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: advection %d flux shapes", name, i);
        AdvArgs& b = a[i];
        b = AdvArgs{};
        b.adv = q.adv.p; b.ld_adv = q.adv.ld;
        if (flux) { b.qx = q.qx.p; b.ld_qx = q.qx.ld; b.qz = q.qz.p; b.ld_qz = q.qz.ld; }
        b.f = q.f.p; b.ld_f = q.f.ld;
        b.u = q.u.p; b.ld_u = q.u.ld;
        b.w = q.w.p; b.ld_w = q.w.ld;
        b.Nz = (int)Nz; b.Nx = (int)Nx;
        b.nbz = (b.Nz + VAR_TZ[2] - 1) / VAR_TZ[2];
        b.nb = b.nbz * ((b.Nx + VAR_TX[2] - 1) / VAR_TX[2]);
        b.invh = 1.0 / q.h;
        b.xbc = xbc; b.zbc = zbc;
    }
    if ((int64_t)a[0].nb + a[1].nb > (1 << 30)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: too many tiles", name);
    const bool ex = (ctx->flags & XC_CTX_EXACT) != 0;
    if (flux) launch_advect_pair_s<true>(scheme, a[0], a[1], ex, ctx->stream);
    else launch_advect_pair_s<false>(scheme, a[0], a[1], ex, ctx->stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xc_fail(ctx, XC_EHIP, "%s: launch: %s", name, hipGetErrorString(e));
    return xc_finish(ctx, name);
}

int xc_advect_centered_batch(xc_ctx* ctx, int scheme, int n, const xc_adv_member* m, int xbc, int zbc) {
    XC_CHECK_CTX(ctx);
    const int g = xc_scheme_halo(scheme);
    if (g < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "advect_centered_batch: unsupported scheme = %d", scheme);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "advect_centered_batch: xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "advect_centered_batch: zBC = %d", zbc);
    if (n < 1 || !m) return xc_fail(ctx, XC_EINVAL_ARG, "advect_centered_batch: no members");
    const bool flux = m[0].qx.p != nullptr;
    const size_t bytes = (size_t)n * sizeof(AdvArgs) + (size_t)(n + 1) * sizeof(int);
    char *host, *dev;
    int st;
    if ((st = batch_table(ctx, bytes, &host, &dev))) return st;
    AdvArgs* A = (AdvArgs*)host;
    int* first = (int*)(host + (size_t)n * sizeof(AdvArgs));
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const xc_adv_member& q = m[i];
        const int64_t Nz = q.adv.rows, Nx = q.adv.cols;
        if (!xc_mat_ok(q.adv) || Nz < g || Nx < g || !xc_shape(q.f, Nz, Nx) || !xc_shape(q.u, Nz, Nx + 1) ||
            !xc_shape(q.w, Nz + 1, Nx))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_centered_batch: member %d has inconsistent shapes", i);
        if ((q.qx.p != nullptr) != flux || (q.qz.p != nullptr) != flux)
            return xc_fail(ctx, XC_EINVAL_ARG, "advect_centered_batch: fluxes for every member or for none");
        if (flux && (!xc_shape(q.qx, Nz + 2, Nx + 1) || !xc_shape(q.qz, Nz + 1, Nx + 2)))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_centered_batch: member %d flux shapes", i);
        AdvArgs a{};
        a.adv = q.adv.p; a.ld_adv = q.adv.ld;
        if (flux) { a.qx = q.qx.p; a.ld_qx = q.qx.ld; a.qz = q.qz.p; a.ld_qz = q.qz.ld; }
        a.f = q.f.p; a.ld_f = q.f.ld;
        a.u = q.u.p; a.ld_u = q.u.ld;
        a.w = q.w.p; a.ld_w = q.w.ld;
        a.Nz = (int)Nz; a.Nx = (int)Nx;
        a.nbz = (a.Nz + VAR_TZ[2] - 1) / VAR_TZ[2];
        a.nb = a.nbz * ((a.Nx + VAR_TX[2] - 1) / VAR_TX[2]);
        a.invh = 1.0 / q.h;
        a.xbc = xbc; a.zbc = zbc;
        A[i] = a;
        first[i] = (int)total;
        total += a.nb;
        if (total > (1 << 30)) return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_centered_batch: too many tiles");
    }
    first[n] = (int)total;
    if ((st = batch_upload(ctx, bytes))) return st;
    const AdvArgs* dA = (const AdvArgs*)dev;
    const int* dfirst = (const int*)(dev + (size_t)n * sizeof(AdvArgs));
    const bool ex = (ctx->flags & XC_CTX_EXACT) != 0;
    if (flux) launch_advect_batch_s<true>(scheme, ex, dA, dfirst, n, (int)total, ctx->stream);
    else launch_advect_batch_s<false>(scheme, ex, dA, dfirst, n, (int)total, ctx->stream);
    return xc_finish(ctx, "advect_centered_batch");
}

int xc_diffus_centered_batch(xc_ctx* ctx, int n, const xc_dif_member* m, int xbc, int zbc) {
    XC_CHECK_CTX(ctx);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "diffus_centered_batch: xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "diffus_centered_batch: zBC = %d", zbc);
    if (n < 1 || !m) return xc_fail(ctx, XC_EINVAL_ARG, "diffus_centered_batch: no members");
    const bool flux = m[0].qx.p != nullptr;
    const size_t bytes = (size_t)n * sizeof(DifArgs) + (size_t)(n + 1) * sizeof(int);
    char *host, *dev;
    int st;
    if ((st = batch_table(ctx, bytes, &host, &dev))) return st;
    DifArgs* A = (DifArgs*)host;
    int* first = (int*)(host + (size_t)n * sizeof(DifArgs));
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const xc_dif_member& q = m[i];
        const int64_t Nz = q.dff.rows, Nx = q.dff.cols;
        if (!xc_mat_ok(q.dff) || Nz < 1 || Nx < 1 || !xc_shape(q.f, Nz, Nx) || !xc_shape(q.k, Nz, Nx))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_centered_batch: member %d has inconsistent shapes", i);
        if ((q.qx.p != nullptr) != flux || (q.qz.p != nullptr) != flux)
            return xc_fail(ctx, XC_EINVAL_ARG, "diffus_centered_batch: fluxes for every member or for none");
        if (flux && (!xc_shape(q.qx, Nz + 2, Nx + 1) || !xc_shape(q.qz, Nz + 1, Nx + 2)))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_centered_batch: member %d flux shapes", i);
        DifArgs a{};
        a.dff = q.dff.p; a.ld_d = q.dff.ld;
        if (flux) { a.qx = q.qx.p; a.ld_qx = q.qx.ld; a.qz = q.qz.p; a.ld_qz = q.qz.ld; }
        a.f = q.f.p; a.ld_f = q.f.ld; a.k = q.k.p; a.ld_k = q.k.ld;
        a.Nz = (int)Nz; a.Nx = (int)Nx; a.halo = 1;
        a.nbz = (int)((Nz + TZ - 1) / TZ);
        a.nb = a.nbz * (int)((Nx + TX - 1) / TX);
        a.invh = 1.0 / q.h;
        a.invh2 = a.invh * a.invh;
        a.xbc = xbc; a.zbc = zbc;
        A[i] = a;
        first[i] = (int)total;
        total += a.nb;
    }
    first[n] = (int)total;
    if ((st = batch_upload(ctx, bytes))) return st;
    const DifArgs* dA = (const DifArgs*)dev;
    const int* dfirst = (const int*)(dev + (size_t)n * sizeof(DifArgs));
    if (flux) hipLaunchKernelGGL((k_diffus_batch<true>), dim3((unsigned)total), dim3(NT), 0, ctx->stream, dA, dfirst, n, (int)total);
    else hipLaunchKernelGGL((k_diffus_batch<false>), dim3((unsigned)total), dim3(NT), 0, ctx->stream, dA, dfirst, n, (int)total);
    return xc_finish(ctx, "diffus_centered_batch");
}

struct xc_batch_plan {
    int kind;      // 0: advection, 1: diffusion
    int scheme;
    int flux;
    int n, total;
    char* dev;     // argument blocks + first-tile table
};

static int plan_make(xc_ctx* ctx, int kind, int scheme, int n, bool flux, int64_t total, const char* host,
                     size_t bytes, xc_batch_plan** out) {
    xc_batch_plan* p = new xc_batch_plan{kind, scheme, flux ? 1 : 0, n, (int)total, nullptr};
    hipError_t e = hipMalloc((void**)&p->dev, bytes);
    if (e == hipSuccess) e = hipMemcpy(p->dev, host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->dev) (void)hipFree(p->dev);
        delete p;
        return xc_fail(ctx, XC_EHIP, "batch plan: %s", hipGetErrorString(e));
    }
    *out = p;
    return XC_OK;
}

int xc_advect_batch_plan(xc_ctx* ctx, int scheme, int n, const xc_adv_member* m, int xbc, int zbc,
                         xc_batch_plan** out) {
    XC_CHECK_CTX(ctx);
    if (!out) return xc_fail(ctx, XC_EINVAL_ARG, "advect_batch_plan: NULL output");
    // validate and lay out exactly as the immediate form, then keep the table in its own buffer
    const int g = xc_scheme_halo(scheme);
    if (g < 0) return xc_fail(ctx, XC_EINVAL_SCHEME, "advect_batch_plan: unsupported scheme = %d", scheme);
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "advect_batch_plan: xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "advect_batch_plan: zBC = %d", zbc);
    if (n < 1 || !m) return xc_fail(ctx, XC_EINVAL_ARG, "advect_batch_plan: no members");
    const bool flux = m[0].qx.p != nullptr;
    const size_t bytes = (size_t)n * sizeof(AdvArgs) + (size_t)(n + 1) * sizeof(int);
    std::vector<char> host(bytes);
    AdvArgs* A = (AdvArgs*)host.data();
    int* first = (int*)(host.data() + (size_t)n * sizeof(AdvArgs));
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const xc_adv_member& q = m[i];
        const int64_t Nz = q.adv.rows, Nx = q.adv.cols;
        if (!xc_mat_ok(q.adv) || Nz < g || Nx < g || !xc_shape(q.f, Nz, Nx) || !xc_shape(q.u, Nz, Nx + 1) ||
            !xc_shape(q.w, Nz + 1, Nx))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_batch_plan: member %d has inconsistent shapes", i);
        if ((q.qx.p != nullptr) != flux || (q.qz.p != nullptr) != flux)
            return xc_fail(ctx, XC_EINVAL_ARG, "advect_batch_plan: fluxes for every member or for none");
        if (flux && (!xc_shape(q.qx, Nz + 2, Nx + 1) || !xc_shape(q.qz, Nz + 1, Nx + 2)))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "advect_batch_plan: member %d flux shapes", i);
        AdvArgs a{};
        a.adv = q.adv.p; a.ld_adv = q.adv.ld;
        if (flux) { a.qx = q.qx.p; a.ld_qx = q.qx.ld; a.qz = q.qz.p; a.ld_qz = q.qz.ld; }
        a.f = q.f.p; a.ld_f = q.f.ld;
        a.u = q.u.p; a.ld_u = q.u.ld;
        a.w = q.w.p; a.ld_w = q.w.ld;
        a.Nz = (int)Nz; a.Nx = (int)Nx;
        a.nbz = (a.Nz + VAR_TZ[2] - 1) / VAR_TZ[2];
        a.nb = a.nbz * ((a.Nx + VAR_TX[2] - 1) / VAR_TX[2]);
        a.invh = 1.0 / q.h;
        a.xbc = xbc; a.zbc = zbc;
        A[i] = a;
        first[i] = (int)total;
        total += a.nb;
    }
    first[n] = (int)total;
    return plan_make(ctx, 0, scheme, n, flux, total, host.data(), bytes, out);
}

int xc_diffus_batch_plan(xc_ctx* ctx, int n, const xc_dif_member* m, int xbc, int zbc, xc_batch_plan** out) {
    XC_CHECK_CTX(ctx);
    if (!out) return xc_fail(ctx, XC_EINVAL_ARG, "diffus_batch_plan: NULL output");
    if (xbc != XC_PERIODIC && xbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "diffus_batch_plan: xBC = %d", xbc);
    if (zbc != XC_PERIODIC && zbc != XC_CLOSED) return xc_fail(ctx, XC_EINVAL_BC, "diffus_batch_plan: zBC = %d", zbc);
    if (n < 1 || !m) return xc_fail(ctx, XC_EINVAL_ARG, "diffus_batch_plan: no members");
    const bool flux = m[0].qx.p != nullptr;
    const size_t bytes = (size_t)n * sizeof(DifArgs) + (size_t)(n + 1) * sizeof(int);
    std::vector<char> host(bytes);
    DifArgs* A = (DifArgs*)host.data();
    int* first = (int*)(host.data() + (size_t)n * sizeof(DifArgs));
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const xc_dif_member& q = m[i];
        const int64_t Nz = q.dff.rows, Nx = q.dff.cols;
        if (!xc_mat_ok(q.dff) || Nz < 1 || Nx < 1 || !xc_shape(q.f, Nz, Nx) || !xc_shape(q.k, Nz, Nx))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_batch_plan: member %d has inconsistent shapes", i);
        if ((q.qx.p != nullptr) != flux || (q.qz.p != nullptr) != flux)
            return xc_fail(ctx, XC_EINVAL_ARG, "diffus_batch_plan: fluxes for every member or for none");
        if (flux && (!xc_shape(q.qx, Nz + 2, Nx + 1) || !xc_shape(q.qz, Nz + 1, Nx + 2)))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diffus_batch_plan: member %d flux shapes", i);
        DifArgs a{};
        a.dff = q.dff.p; a.ld_d = q.dff.ld;
        if (flux) { a.qx = q.qx.p; a.ld_qx = q.qx.ld; a.qz = q.qz.p; a.ld_qz = q.qz.ld; }
        a.f = q.f.p; a.ld_f = q.f.ld; a.k = q.k.p; a.ld_k = q.k.ld;
        a.Nz = (int)Nz; a.Nx = (int)Nx; a.halo = 1;
        a.nbz = (int)((Nz + TZ - 1) / TZ);
        a.nb = a.nbz * (int)((Nx + TX - 1) / TX);
        a.invh = 1.0 / q.h;
        a.invh2 = a.invh * a.invh;
        a.xbc = xbc; a.zbc = zbc;
        A[i] = a;
        first[i] = (int)total;
        total += a.nb;
    }
    first[n] = (int)total;
    return plan_make(ctx, 1, 0, n, flux, total, host.data(), bytes, out);
}

int xc_batch_run(xc_ctx* ctx, const xc_batch_plan* p) {
    XC_CHECK_CTX(ctx);
    if (!p) return xc_fail(ctx, XC_EINVAL_ARG, "batch_run: NULL plan");
    const bool ex = (ctx->flags & XC_CTX_EXACT) != 0;
    if (p->kind == 0) {
        const AdvArgs* dA = (const AdvArgs*)p->dev;
        const int* dfirst = (const int*)(p->dev + (size_t)p->n * sizeof(AdvArgs));
        if (p->flux) launch_advect_batch_s<true>(p->scheme, ex, dA, dfirst, p->n, p->total, ctx->stream);
        else launch_advect_batch_s<false>(p->scheme, ex, dA, dfirst, p->n, p->total, ctx->stream);
    } else {
        const DifArgs* dA = (const DifArgs*)p->dev;
        const int* dfirst = (const int*)(p->dev + (size_t)p->n * sizeof(DifArgs));
        if (p->flux) hipLaunchKernelGGL((k_diffus_batch<true>), dim3((unsigned)p->total), dim3(NT), 0, ctx->stream, dA, dfirst, p->n, p->total);
        else hipLaunchKernelGGL((k_diffus_batch<false>), dim3((unsigned)p->total), dim3(NT), 0, ctx->stream, dA, dfirst, p->n, p->total);
    }
    return xc_finish(ctx, "batch_run");
}

int xc_batch_destroy(xc_ctx* ctx, xc_batch_plan* p) {
    XC_CHECK_CTX(ctx);
    if (!p) return XC_OK;
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));  // no launch may still read the table
    (void)hipFree(p->dev);
    delete p;
    return XC_OK;
}

}  // extern "C"
