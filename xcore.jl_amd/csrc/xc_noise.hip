// xc_noise.hip -- Ornstein-Uhlenbeck noise with the zero-padded Gaussian FFT
// filter (src/noise.jl) for gfx950, FP64.
//
// * White noise: counter-based Philox4x32-10 + Box-Muller, one lane per pair
//   of normals; any (seed, draw, element) is reproducible without state.
// * Filter: exp(-s^2 (kz^2 + kx^2)) = exp(-s^2 kz^2) exp(-s^2 kx^2), so the 2-D
//   FFT -> multiply -> inverse FFT of noise.jl:254-256 runs as two batched
//   1-D passes (x periodic over Nx; z over the zero-padded Nz+pad).  Each
//   workgroup keeps its lines in LDS for the forward transform, the filter
//   multiply and the inverse transform (Stockham autosort, mixed radix
//   2/3/4/5/7/11/13, twiddles from a table computed on the host in long
//   double).  Two real lines ride in one complex transform: the multiplier is
//   real and even, so real and imaginary parts are filtered independently.
//   Filters several cells wide, whose taps beyond a half-width R carry at most
//   2^-54 of the kernel's mass, run instead as direct convolutions with the
//   2R+1 taps of the same circular convolution (k_gconv_x for R <= 43,
//   k_gconv_z for R <= 128; lines of at least 64 points).
// * Mean/std rescale (noise.jl:261-267): one pass over all six raw/filtered
//   fields summing x - x[0] and its square (the shift keeps the one-pass
//   variance free of cancellation), then one wave per field.
// * Flux differencing, tapers and magnitudes (noise.jl:195-236) as two
//   elementwise kernels that apply the rescale on the fly.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "xc_fft.h"
#include "xc_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int NT = 256;
constexpr double EPS = 2.220446049250313e-16;

// ---------------------------------------------------------------- RNG
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// element pair p of a column-major matrix -> two standard normals (same mapping as the oracle)
__global__ __launch_bounds__(NT) void k_normals(double* a0, int64_t ld0, double* a1, int64_t ld1, int rows,
                                                int64_t n, uint64_t seed, uint32_t draw, uint32_t stream0) {
    const int64_t p = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (2 * p >= n) return;
    const int s = blockIdx.y;
    double* a = s == 0 ? a0 : a1;
    const int64_t ld = s == 0 ? ld0 : ld1;
    uint32_t c[4] = {(uint32_t)p, (uint32_t)((uint64_t)p >> 32), draw, stream0 + (uint32_t)s};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const uint64_t A = ((uint64_t)c[1] << 32) | c[0], B = ((uint64_t)c[3] << 32) | c[2];
    const double u1 = ((double)(A >> 11) + 1.0) * 0x1.0p-53;
    const double u2 = (double)(B >> 11) * 0x1.0p-53;
    const double r = sqrt(-2.0 * log(u1));
    const double th = 2.0 * 3.14159265358979323846 * u2;
    const int64_t e0 = 2 * p, e1 = e0 + 1;
    a[(e0 % rows) + (e0 / rows) * ld] = r * cos(th);
    if (e1 < n) a[(e1 % rows) + (e1 / rows) * ld] = r * sin(th);
}

// ---------------------------------------------------------------- OU update (noise.jl:159-183)
struct OUArgs {
    const double *V, *vx, *fReL, *ke, *chi, *ks;
    int64_t ldV, ldvx, ldf, ldk, ldc, ldks;
    double *psie, *psix, *psis;
    int64_t lde, ldx, lds;
    const double *psieo, *psixo, *psiso, *re, *rs;
    int64_t ldeo, ldxo, ldso, ldre, ldrs;
    double Xi, L0h, l0h, txi0, dt, fL, fl;
    int Nz, Nx;
    xc_dom d;
};
__global__ __launch_bounds__(NT) void k_ou(OUArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    if (j >= a.Nz || i >= a.Nx) return;
    double taue = a.L0h / (2.0 * (a.V[j + i * a.ldV] + EPS));
    double taus = a.l0h / (2.0 * (a.vx[j + i * a.ldvx] + EPS));
    const double St = a.txi0 / (taue + EPS);
    const double fReL = a.fReL[j + i * a.ldf], ke = a.ke[j + i * a.ldk], chi = a.chi[j + i * a.ldc];
    const double q1 = fReL * ke / taue;
    const double sge = a.Xi * sqrt(q1 > 0.0 ? q1 : 0.0);
    const double q2 = chi * fReL * ke / taue * St / (1.0 + St * St);
    const double sgx = a.Xi * sqrt(q2 > 0.0 ? q2 : 0.0);
    const double q3 = chi * a.ks[j + i * a.ldks] / taus;
    const double sgs = a.Xi * sqrt(q3 > 0.0 ? q3 : 0.0);
    taue = 1.0 / (1.0 / taue + 1.0 / (1e2 * a.dt)) + a.dt;
    taus = 1.0 / (1.0 / taus + 1.0 / (1e2 * a.dt)) + a.dt;
    const double Fte = exp(-a.dt / taue), Fts = exp(-a.dt / taus);
    double ge = 1.0 - Fte * Fte, gs = 1.0 - Fts * Fts;
    ge = ge > 0.0 ? ge : 0.0;
    gs = gs > 0.0 ? gs : 0.0;
    const double re = a.re[j + i * a.ldre], rs = a.rs[j + i * a.ldrs];
    a.psie[j + i * a.lde] = Fte * a.psieo[j + i * a.ldeo] + sqrt(ge) * sge * a.fL * re;
    a.psix[j + i * a.ldx] = Fte * a.psixo[j + i * a.ldxo] + sqrt(ge) * sgx * a.fL * re;
    a.psis[j + i * a.lds] = Fts * a.psiso[j + i * a.ldso] + sqrt(gs) * sgs * a.fl * rs;
}

struct FieldIO {
    const double* in;
    int64_t ld_in;
    double* out;
    int64_t ld_out;
};

// x pass: rows of a (Nz x Nx) field, periodic over Nx, 2*nl rows per workgroup.
// blockIdx.y = field.
struct XPassArgs {
    FieldIO io[3];
    Plan plan[3];
    int Nz, Nx, nl, nblk;
};
template <int LG>  // 8..10: the power-of-two path with prefetched twiddles (P2Fast); 0: lds_fft
__global__ __launch_bounds__(NT) void k_filter_x(XPassArgs a) {
    extern __shared__ double2 sm[];
    const int f = blockIdx.y;
    const Plan& p = a.plan[f];
    const FieldIO io = a.io[f];
    const int n = a.Nx, nl = a.nl, rows = 2 * nl;
    double2 w[P2Fast<LG>::NST][3];
    if constexpr (LG > 0) fft_p2_twiddles<LG>(nl, p.tw, w);
    const int r0 = xcd_tile(blockIdx.x, gridDim.x) * rows;  // row blocks sharing cache lines on one XCD
    double2* A = sm;
    double2* B = sm + nl * n;
    for (int k = threadIdx.x; k < rows * n; k += NT) {
        const int rr = k % rows, c = k / rows;
        const int r = r0 + rr;
        const double v = r < a.Nz ? io.in[r + (int64_t)c * io.ld_in] : 0.0;
        double* slot = reinterpret_cast<double*>(&A[(rr >> 1) * n + c]);
        slot[rr & 1] = v;
    }
    __syncthreads();
    double2* X;
    if constexpr (LG > 0) X = lds_fft_p2<false, LG>(A, B, nl, w);
    else X = lds_fft<false>(A, B, nl, p);
    for (int k = threadIdx.x; k < nl * n; k += NT) {
        const double g = p.g[k % n];
        X[k].x *= g;
        X[k].y *= g;
    }
    __syncthreads();
    double2* Y;
    if constexpr (LG > 0) Y = lds_fft_p2<true, LG>(X, X == A ? B : A, nl, w);
    else Y = lds_fft<true>(X, X == A ? B : A, nl, p);
    for (int k = threadIdx.x; k < rows * n; k += NT) {
        const int rr = k % rows, c = k / rows;
        const int r = r0 + rr;
        if (r < a.Nz) {
            const double* slot = reinterpret_cast<const double*>(&Y[(rr >> 1) * n + c]);
            io.out[r + (int64_t)c * io.ld_out] = slot[rr & 1];
        }
    }
}

// z pass: columns zero-padded to Nz+pad (hp = pad/2 zeros on top), 2*nl columns per workgroup.
struct ZPassArgs {
    FieldIO io[3];
    Plan plan[3];
    int hp[3];
    int nl[3];
    int Nz, Nx;
};
__global__ __launch_bounds__(NT) void k_filter_z(ZPassArgs a) {
    extern __shared__ double2 sm[];
    const int f = blockIdx.y;
    const Plan& p = a.plan[f];
    const FieldIO io = a.io[f];
    const int n = p.n, nl = a.nl[f], cols = 2 * nl, hp = a.hp[f];
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * cols;
    if (c0 >= a.Nx) return;  // whole workgroup (uniform)
    double2* A = sm;
    double2* B = sm + nl * n;
    for (int k = threadIdx.x; k < cols * n; k += NT) {
        const int z = k % n, cc = k / n;
        const int c = c0 + cc;
        const int r = z - hp;
        const double v = (c < a.Nx && r >= 0 && r < a.Nz) ? io.in[r + (int64_t)c * io.ld_in] : 0.0;
        double* slot = reinterpret_cast<double*>(&A[(cc >> 1) * n + z]);
        slot[cc & 1] = v;
    }
    __syncthreads();
    double2* X = lds_fft<false>(A, B, nl, p);
    for (int k = threadIdx.x; k < nl * n; k += NT) {
        const double g = p.g[k % n];
        X[k].x *= g;
        X[k].y *= g;
    }
    __syncthreads();
    double2* Y = lds_fft<true>(X, X == A ? B : A, nl, p);
    for (int k = threadIdx.x; k < cols * a.Nz; k += NT) {
        const int r = k % a.Nz, cc = k / a.Nz;
        const int c = c0 + cc;
        if (c < a.Nx) {
            const double* slot = reinterpret_cast<const double*>(&Y[(cc >> 1) * n + r + hp]);
            io.out[r + (int64_t)c * io.ld_out] = slot[cc & 1];
        }
    }
}

// Narrow filters as a direct convolution.  The FFT filter of a line of period n is the circular
// convolution with c[m] = sum_t g_t cos(2 pi t m / n) (g real and even in t).  When the taps beyond
// |m| = R carry at most 2^-54 of the total mass (a filter several cells wide), the 2R+1 taps are applied
// directly: each lane computes CV_P consecutive outputs from a window of CV_P + 2R (rounded up) inputs,
// the taps CV_P at a time from two register blocks of the window (ping-pong, no copies), every tap a
// uniform (broadcast) LDS read.  taps[u] = c[|R - u|] for u <= 2R and 0 up to nblk CV_P, so output j of a lane
// is sum_u taps[u] window[j + u], window[i] = input (first output - R + i).  The work is FMA-bound: CV_P
// is small enough that a 512^2 field gives every SIMD several waves.
constexpr int CV_P = 4;
constexpr int CV_MAXTAP = 257;  // 2R+1 at most: beyond it the FFT passes are used
constexpr int CV_MINLINE = 64;  // shorter lines: the FFT passes (the window indexing wraps once per step)
struct ConvField {
    const double* in;
    int64_t ld_in;
    double* out;
    int64_t ld_out;
    const double* taps;
    int R, nblk;
};
// acc[s & 1][j] += sum_s tp[s] win[j + s], win = lo | hi (even and odd taps in separate sums: twice the
// independent FMA chains)
__device__ __forceinline__ void conv_block(const double* __restrict__ tp, const double (&lo)[CV_P],
                                           const double (&hi)[CV_P], double (&acc)[2][CV_P]) {
#pragma unroll
    for (int s = 0; s < CV_P; ++s) {
        const double w = tp[s];
#pragma unroll
        for (int j = 0; j < CV_P; ++j)
            acc[s & 1][j] = fma(w, j + s < CV_P ? lo[j + s] : hi[j + s - CV_P], acc[s & 1][j]);
    }
}
// taps to LDS (every wait in the convolution loop is then an in-order LDS wait, so reads of the next block
// overlap the FMAs of this one; scalar loads would force waits for every outstanding LDS read)
__device__ __forceinline__ void stage_taps(const ConvField& F, double* stap) {
    for (int u = threadIdx.x; u < F.nblk * CV_P; u += blockDim.x) stap[u] = F.taps[u];
}
// the blocked convolution of a lane: get(i, v) fills v with window elements i .. i + CV_P - 1
template <class Get>
__device__ __forceinline__ void conv_lane(const double* __restrict__ taps, int nblk, Get get, double (&out)[CV_P]) {
    double a[CV_P], b[CV_P], acc[2][CV_P];
#pragma unroll
    for (int j = 0; j < CV_P; ++j) acc[0][j] = acc[1][j] = 0.0;
    get(0, a);
    int k = 0;
    for (; k + 1 < nblk; k += 2) {
        get((k + 1) * CV_P, b);
        conv_block(taps + k * CV_P, a, b, acc);
        get((k + 2) * CV_P, a);
        conv_block(taps + (k + 1) * CV_P, b, a, acc);
    }
    if (k < nblk) {
        get((k + 1) * CV_P, b);
        conv_block(taps + k * CV_P, a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < CV_P; ++j) out[j] = acc[0][j] + acc[1][j];
}
// x pass, periodic over Nx.  A workgroup of CVX_W waves takes 64 rows and CVX_C = CVX_W CV_P output
// columns: the CVX_C + 2R (rounded up) window columns go to LDS ([column][row], loads in flight U at a
// time, each a coalesced 64-row column), then lane = row and wave w computes columns w CV_P ..
// w CV_P + CV_P - 1.  blockIdx.y = field.
constexpr int CVX_W = 8, CVX_C = CVX_W * CV_P;
// window tile and taps within 64 KiB of LDS (R <= 43)
constexpr int CVX_MAXBLK = ((8192 - CV_MAXTAP - CV_P) / 64 - CVX_C) / CV_P;
struct ConvXArgs {
    ConvField f[3];
    int Nz, Nx, nrc, ncc;  // row chunks of 64, column chunks of CVX_C
};
__global__ __launch_bounds__(64 * CVX_W) void k_gconv_x(ConvXArgs a) {
    extern __shared__ double tile[];
    __shared__ double stap[CV_MAXTAP + CV_P];
    const ConvField F = a.f[blockIdx.y];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t = xcd_tile(blockIdx.x, gridDim.x);  // neighbouring column chunks (shared halos) on one XCD
    const int rc = t % a.nrc, cc = t / a.nrc;
    const int r = rc * 64 + lane, Nx = a.Nx;
    const bool ok = r < a.Nz;
    const double* x = F.in + (ok ? r : a.Nz - 1);
    const int c0 = cc * CVX_C;
    const int wc = CVX_C + F.nblk * CV_P;  // window columns: column (c0 - R + i) mod Nx
    // this wave's window columns i = wave + CVX_W k, their columns advanced by CVX_W < Nx with one wrap
    int col = c0 - F.R + wave;
    col += col < 0 ? Nx : 0;
    col -= col >= Nx ? Nx : 0;
    constexpr int U = 8;
    for (int i0 = wave; i0 < wc; i0 += CVX_W * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {  // in bounds past the window too (the column wraps), so unconditional
            v[u] = x[(int64_t)col * F.ld_in];
            col += CVX_W;
            col -= col >= Nx ? Nx : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int i = i0 + u * CVX_W;
            if (i < wc) tile[i * 64 + lane] = v[u];
        }
    }
    stage_taps(F, stap);
    __syncthreads();
    const double* tw = tile + wave * CV_P * 64 + lane;
    double acc[CV_P];
    conv_lane(stap, F.nblk, [&](int i, double (&v)[CV_P]) {
#pragma unroll
        for (int s = 0; s < CV_P; ++s) v[s] = tw[(i + s) * 64];
    }, acc);
    // stores without per-store branches, so that none waits on the one before
    const int cw = c0 + wave * CV_P;
    double* y = F.out + r + (int64_t)cw * F.ld_out;
    if (ok && cw + CV_P <= Nx) {
#pragma unroll
        for (int j = 0; j < CV_P; ++j) y[(int64_t)j * F.ld_out] = acc[j];
    } else if (ok) {
        for (int j = 0; j < CV_P && cw + j < Nx; ++j) y[(int64_t)j * F.ld_out] = acc[j];
    }
}
// z pass over the column zero-padded to period n = Nz + pad (the circular convolution does not depend on
// where the zeros sit, so the column is taken as rows 0..Nz-1 then n - Nz zeros, repeated).  One wave a
// segment of 64 CV_P rows of one column: the window goes to LDS with element q at (q % CV_P) S + q / CV_P,
// so that lane l's element CV_P l + i sits at (i % CV_P) S + l + i / CV_P: consecutive lanes read
// consecutive words.  blockIdx.y = field.
struct ConvZArgs {
    ConvField f[3];
    int n[3];
    int Nz, Nx, nseg;
};
__global__ __launch_bounds__(64) void k_gconv_z(ConvZArgs a) {
    extern __shared__ double win[];
    __shared__ double stap[CV_MAXTAP + CV_P];
    const ConvField F = a.f[blockIdx.y];
    const int n = a.n[blockIdx.y], Nz = a.Nz;
    const int lane = threadIdx.x;
    const int c = blockIdx.x / a.nseg, seg = blockIdx.x - c * a.nseg;
    const int r0 = seg * 64 * CV_P;
    const int S = 64 + F.nblk;  // words per residue class
    const int len = CV_P * S;   // window elements: row (r0 - R + q) mod n
    const double* x = F.in + (int64_t)c * F.ld_in;
    // lane's elements q = lane + 64 k, their rows advanced by 64 <= n with one wrap
    int j = r0 - F.R + lane;
    j += j < 0 ? n : 0;
    j -= j >= n ? n : 0;
    constexpr int U = 8;
    for (int q0 = lane; q0 < len; q0 += 64 * U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {  // unconditional loads (clamped row): all in flight at once
            const double t = x[j < Nz ? j : Nz - 1];
            v[u] = j < Nz ? t : 0.0;
            j += 64;
            j -= j >= n ? n : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = q0 + 64 * u;
            if (q < len) win[(q % CV_P) * S + q / CV_P] = v[u];
        }
    }
    stage_taps(F, stap);
    __syncthreads();
    double acc[CV_P];
    conv_lane(stap, F.nblk, [&](int i, double (&v)[CV_P]) {  // i is a multiple of CV_P
#pragma unroll
        for (int s = 0; s < CV_P; ++s) v[s] = win[s * S + lane + i / CV_P];
    }, acc);
    // stores without per-store branches, so that none waits on the one before
    double* y = F.out + (int64_t)c * F.ld_out + r0 + CV_P * lane;
    const int rem = Nz - (r0 + CV_P * lane);
    if (rem >= CV_P) {
#pragma unroll
        for (int q = 0; q < CV_P; ++q) y[q] = acc[q];
    } else {
        for (int q = 0; q < rem; ++q) y[q] = acc[q];
    }
}

// ---------------------------------------------------------------- statistics
// six (Nz x Nx) arrays: raw e, x, s and filtered e, x, s
struct StatArgs {
    const double* a[6];
    int64_t ld[6];
    int Nz, Nx;
    xc_dom d;
    double* part;   // [12][nb]: sums of d = x - x[0], then of d^2
    double* stats;  // [6] means, [6] stds (n - 1)
};
// One pass over the six arrays: per workgroup, the sums of d = x - x[0] and of d^2.  The shift by the array's
// first element keeps the one-pass variance below free of cancellation where the mean is large against the
// spread (the two-pass form read every array twice).
__global__ __launch_bounds__(NT) void k_stat_partial(StatArgs s) {
    int j, i;
    xc_dom_ij(s.d, &j, &i);
    const bool in = j < s.Nz && i < s.Nx;
    __shared__ double red[12][NT / 64];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double K = s.a[q][0];
        const double d = in ? s.a[q][j + (int64_t)i * s.ld[q]] - K : 0.0;
        const double d1 = wave_total(d);  // DPP wave sums, then the NT / 64 wave sums
        const double d2 = wave_total(d * d);
        if ((threadIdx.x & 63) == 0) {
            red[q][threadIdx.x >> 6] = d1;
            red[6 + q][threadIdx.x >> 6] = d2;
        }
    }
    __syncthreads();
    if (threadIdx.x < 12) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) t += red[threadIdx.x][w];
        s.part[(int64_t)threadIdx.x * s.d.nb + blockIdx.x] = t;
    }
}
// stats[q] = mean = x[0] + sum d / n, stats[6 + q] = std (n - 1) = sqrt((sum d^2 - (sum d)^2 / n) / (n - 1))
// (0 where that is not positive; noise.jl:261-267 rescales only where the std exceeds eps).  One wave per row of
// partials.
__global__ __launch_bounds__(12 * 64) void k_stat_final(StatArgs s, double n) {
    __shared__ double t[12];
    const int q = threadIdx.x >> 6;
    const double v = xc_rows_reduce<false>(s.part, s.d.nb, q);
    if ((threadIdx.x & 63) == 0) t[q] = v;
    __syncthreads();
    if (threadIdx.x < 6) {
        const int f = threadIdx.x;
        const double t1 = t[f], t2 = t[6 + f];
        s.stats[f] = s.a[f][0] + t1 / n;
        const double var = (t2 - t1 * t1 / n) / (n - 1.0);
        s.stats[6 + f] = sqrt(var > 0.0 ? var : 0.0);
    }
}

// rescaled filtered value (noise.jl:265-267); st: [0..5] means, [6..11] stds
__device__ __forceinline__ double resc(const double* flt, int64_t ld, int j, int i, const double* st, int f) {
    const double v = flt[j + (int64_t)i * ld];
    const double rm = st[f], fm = st[3 + f], rs = st[6 + f], fs = st[9 + f];
    if (fs > EPS) return (v - fm) * rs / fs + rm;
    return v;
}

// ---------------------------------------------------------------- fluxes (noise.jl:190-220)
struct FluxArgs {
    const double* flt[3];
    int64_t ldf[3];
    const double* st;
    const double *chiw, *muw, *chiu, *muu;
    int64_t ldchiw, ldmuw, ldchiu, ldmuu;
    double *xiew, *xieu, *xixw, *xixu, *xisw, *xisu;
    int64_t ldew, ldeu, ldxw, ldxu, ldsw, ldsu;
    int Nz, Nx;
    xc_dom d;
};
__global__ __launch_bounds__(NT) void k_noise_flux(FluxArgs a) {
    int j, i;
    xc_dom_ij(a.d, &j, &i);
    const int Nz = a.Nz, Nx = a.Nx;
    if (j >= Nz + 2 || i >= Nx + 2) return;
    // the rescale of field f (noise.jl:265-267), in the reference's operation order ((v - fm) rs) / fs + rm
    // (identity when fs <= eps), as resc() and k_resc_inplace compute it
    double rm[3], fm[3], rs[3], fs[3];
    bool on[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
        rm[f] = a.st[f];
        fm[f] = a.st[3 + f];
        rs[f] = a.st[6 + f];
        fs[f] = a.st[9 + f];
        on[f] = fs[f] > EPS;
    }
    // ext(f, j, i) = flt_f[icz[j], icx[i]] with grid.icz closed, grid.icx periodic (grid.jl:68-69)
    auto ext = [&](int f, int jj, int ii) -> double {
        const int z = jj - 1 < 0 ? 0 : (jj - 1 >= Nz ? Nz - 1 : jj - 1);
        const int x = ii - 1 < 0 ? Nx - 1 : (ii - 1 >= Nx ? 0 : ii - 1);
        const double v = a.flt[f][z + (int64_t)x * a.ldf[f]];
        return on[f] ? (v - fm[f]) * rs[f] / fs[f] + rm[f] : v;
    };
    auto psiec = [&](int jj, int ii) -> double {
        return (ext(0, jj, ii) + ext(0, jj, ii + 1) + ext(0, jj + 1, ii) + ext(0, jj + 1, ii + 1)) / 4.0;
    };
    auto taper = [](double chi, double mu) {
        const double m = mu - 0.4;
        return (1.0 - exp(-chi / 1e-4)) * (1.0 - exp(-(m > 0.0 ? m : 0.0) / 0.05));
    };
    if (j >= 2 && j <= Nz - 1 && i >= 2 && i <= Nx - 1) {
        // away from the ghost rows and columns every ext(f, jj, ii) above is the rescaled value at
        // (jj - 1, ii - 1) itself: the same terms in the same order, each distinct one evaluated once (8 of
        // psi's 16, 3 of the 4 of each other field), psiec(j, i) shared by both face types
        auto g = [&](int f, int z, int x) -> double {
            const double v = a.flt[f][z + (int64_t)x * a.ldf[f]];
            return on[f] ? (v - fm[f]) * rs[f] / fs[f] + rm[f] : v;
        };
        const double p11 = g(0, j - 1, i - 1), p12 = g(0, j - 1, i), p21 = g(0, j, i - 1), p22 = g(0, j, i);
        const double p01 = g(0, j - 2, i - 1), p02 = g(0, j - 2, i), p10 = g(0, j - 1, i - 2), p20 = g(0, j, i - 2);
        const double ps = (p11 + p12 + p21 + p22) / 4.0;           // psiec(j, i)
        const double pu = (p01 + p02 + p11 + p12) / 4.0;           // psiec(j - 1, i)
        const double pw = (p10 + p11 + p20 + p21) / 4.0;           // psiec(j, i - 1)
        const double x11 = g(1, j - 1, i - 1), x12 = g(1, j - 1, i), x21 = g(1, j, i - 1);
        const double s11 = g(2, j - 1, i - 1), s12 = g(2, j - 1, i), s21 = g(2, j, i - 1);
        const double tu = taper(a.chiu[j + (int64_t)i * a.ldchiu], a.muu[j + (int64_t)i * a.ldmuu]);
        const double tw = taper(a.chiw[j + (int64_t)i * a.ldchiw], a.muw[j + (int64_t)i * a.ldmuw]);
        a.xieu[j + (int64_t)i * a.ldeu] = ps - pu;
        a.xixu[j + (int64_t)i * a.ldxu] = (-(x12 - x11)) * tu;
        a.xisu[j + (int64_t)i * a.ldsu] = (-(s12 - s11)) * tu;
        a.xiew[j + (int64_t)i * a.ldew] = -(ps - pw);
        a.xixw[j + (int64_t)i * a.ldxw] = (-(x21 - x11)) * tw;
        a.xisw[j + (int64_t)i * a.ldsw] = (-(s21 - s11)) * tw;
        return;
    }
    if (i <= Nx) {  // U-type (Nz+2, Nx+1)
        const int z = j - 1 < 0 ? 0 : (j - 1 >= Nz ? Nz - 1 : j - 1);
        a.xieu[j + (int64_t)i * a.ldeu] = psiec(z + 1, i) - psiec(z, i);
        const double tap = taper(a.chiu[j + (int64_t)i * a.ldchiu], a.muu[j + (int64_t)i * a.ldmuu]);
        a.xixu[j + (int64_t)i * a.ldxu] = (-(ext(1, j, i + 1) - ext(1, j, i))) * tap;
        a.xisu[j + (int64_t)i * a.ldsu] = (-(ext(2, j, i + 1) - ext(2, j, i))) * tap;
    }
    if (j <= Nz) {  // W-type (Nz+1, Nx+2)
        const int x = i - 1 < 0 ? Nx - 1 : (i - 1 >= Nx ? 0 : i - 1);
        a.xiew[j + (int64_t)i * a.ldew] = -(psiec(j, x + 1) - psiec(j, x));
        const double tap = taper(a.chiw[j + (int64_t)i * a.ldchiw], a.muw[j + (int64_t)i * a.ldmuw]);
        a.xixw[j + (int64_t)i * a.ldxw] = (-(ext(1, j + 1, i) - ext(1, j, i))) * tap;
        a.xisw[j + (int64_t)i * a.ldsw] = (-(ext(2, j + 1, i) - ext(2, j, i))) * tap;
    }
}

// noise.jl:223-236
__global__ __launch_bounds__(NT) void k_noise_mag(const double* sw, int64_t ldsw, const double* su, int64_t ldsu,
                                                  const double* xw, int64_t ldxw, const double* xu, int64_t ldxu,
                                                  const double* ew, int64_t ldew, const double* eu, int64_t ldeu,
                                                  double* xis, int64_t lds, double* xix, int64_t ldx, double* xie,
                                                  int64_t lde, int Nz, int Nx, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j >= Nz || i >= Nx) return;
    auto mag = [&](const double* w, int64_t lw, const double* u, int64_t lu) {
        const double wz = (w[j + (int64_t)(i + 1) * lw] + w[j + 1 + (int64_t)(i + 1) * lw]) / 2.0;
        const double wx = (u[j + 1 + (int64_t)i * lu] + u[j + 1 + (int64_t)(i + 1) * lu]) / 2.0;
        return sqrt(wz * wz + wx * wx);
    };
    xis[j + (int64_t)i * lds] = mag(sw, ldsw, su, ldsu);
    xix[j + (int64_t)i * ldx] = mag(xw, ldxw, xu, ldxu);
    xie[j + (int64_t)i * lde] = mag(ew, ldew, eu, ldeu);
}

__global__ __launch_bounds__(NT) void k_resc_inplace(double* o, int64_t ldo, const double* st, int Nz, int Nx,
                                                     xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j >= Nz || i >= Nx) return;
    o[j + (int64_t)i * ldo] = resc(o, ldo, j, i, st, 0);
}

__global__ __launch_bounds__(NT) void k_copy3(const double* a, int64_t lda, double* ao, int64_t ldao, const double* b,
                                              int64_t ldb, double* bo, int64_t ldbo, const double* c, int64_t ldc,
                                              double* co, int64_t ldco, int Nz, int Nx, xc_dom d) {
    int j, i;
    xc_dom_ij(d, &j, &i);
    if (j >= Nz || i >= Nx) return;
    ao[j + (int64_t)i * ldao] = a[j + (int64_t)i * lda];
    bo[j + (int64_t)i * ldbo] = b[j + (int64_t)i * ldb];
    co[j + (int64_t)i * ldco] = c[j + (int64_t)i * ldc];
}

// ---------------------------------------------------------------- host: plans
struct PlanEntry {
    int n;
    double sigma, h;
    std::vector<int> rad;
    double2* tw;  // device
    double* g;    // device
    double* taps = nullptr;  // device, direct-convolution taps (ConvField layout) when R >= 0
    int R = -1, nblk = 0;
};
struct NoiseCtx {
    std::vector<PlanEntry> plans;
};

void noise_free(xc_ctx*, void* p) {
    NoiseCtx* nc = (NoiseCtx*)p;
    for (auto& e : nc->plans) {
        (void)hipFree(e.tw);
        (void)hipFree(e.g);
        if (e.taps) (void)hipFree(e.taps);
    }
    delete nc;
}

// fftfreq(n, 1/h)[t] * 2 pi (AbstractFFTs ordering)
double kfreq(int t, int n, double h) {
    const int nn = (n + 1) / 2;
    const double mult = (1.0 / h) / (double)n;
    return 2.0 * 3.14159265358979323846 * ((double)(t < nn ? t : t - n) * mult);
}

// The taps of the circular convolution the FFT filter of period n applies, in long double: c[m] =
// sum_t g_t cos(2 pi ((t m) mod n) / n), g_t = exp(-sigma^2 k_t^2) / n.  R = the least half-width whose
// dropped taps (R < m < n - R) carry at most 2^-54 of sum |c|; -1 when 2R+1 would reach n or CV_MAXTAP.
int conv_taps(int n, double sigma, double h, std::vector<double>& taps, int& nblk) {
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> cs(n), g(n), c(n);
    for (int j = 0; j < n; ++j) cs[j] = cosl(2.0L * pi * (long double)j / (long double)n);
    const int nn = (n + 1) / 2;
    for (int t = 0; t < n; ++t) {
        const long double k = 2.0L * pi * (long double)(t < nn ? t : t - n) / ((long double)n * (long double)h);
        g[t] = expl(-(long double)sigma * (long double)sigma * k * k) / (long double)n;
    }
    for (int m = 0; m <= n / 2; ++m) {
        long double acc = 0.0L;
        for (int t = 0; t < n; ++t) acc += g[t] * cs[(int)(((int64_t)t * m) % n)];
        c[m] = acc;
        c[(n - m) % n] = acc;
    }
    long double tot = 0.0L;
    for (int m = 0; m < n; ++m) tot += fabsl(c[m]);
    const long double bar = ldexpl(tot, -54);
    // tail[m] = sum of |c| over R < m' < n - R for R = m, from the middle outwards
    for (int R = 0; 2 * R + 1 < n && 2 * R + 1 <= CV_MAXTAP; ++R) {
        long double drop = 0.0L;
        for (int m = R + 1; m < n - R; ++m) drop += fabsl(c[m]);
        if (drop <= bar) {
            nblk = (2 * R + 1 + CV_P - 1) / CV_P;
            taps.assign((size_t)nblk * CV_P, 0.0);
            for (int u = 0; u <= 2 * R; ++u) taps[u] = (double)c[R > u ? R - u : u - R];
            return R;
        }
    }
    return -1;
}

int get_plan(xc_ctx* ctx, int n, double sigma, double h, Plan* out, ConvField* cf = nullptr) {
    NoiseCtx* nc = (NoiseCtx*)ctx->ext[XC_EXT_NOISE];
    if (!nc) {
        nc = new NoiseCtx();
        ctx->ext[XC_EXT_NOISE] = nc;
        ctx->ext_free[XC_EXT_NOISE] = noise_free;
    }
    for (auto& e : nc->plans)
        if (e.n == n && e.sigma == sigma && e.h == h) {
            out->n = n;
            out->nst = (int)e.rad.size();
            for (int s = 0; s < out->nst; ++s) out->rad[s] = e.rad[s];
            out->tw = e.tw;
            out->g = e.g;
            plan_mark_pow2(*out);
            if (cf) {
                cf->taps = e.taps;
                cf->R = e.R;
                cf->nblk = e.nblk;
            }
            return XC_OK;
        }
    PlanEntry e;
    e.n = n;
    e.sigma = sigma;
    e.h = h;
    if (n > MAXLINE || !fft_factorize(n, e.rad))
        return xc_fail(ctx, XC_EINVAL_SHAPE,
                       "noise filter: line length %d must be <= %d", n, MAXLINE);
    std::vector<double2> tw;
    fft_twiddles(n, tw);
    std::vector<double> g(n);
    for (int t = 0; t < n; ++t) {
        const double k = kfreq(t, n, h);
        g[t] = exp(-(sigma * sigma) * (k * k)) / (double)n;  // inverse-transform normalisation folded in
    }
    XC_TRY_HIP(ctx, hipMalloc(&e.tw, n * sizeof(double2)));
    XC_TRY_HIP(ctx, hipMalloc(&e.g, n * sizeof(double)));
    XC_TRY_HIP(ctx, hipMemcpy(e.tw, tw.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    XC_TRY_HIP(ctx, hipMemcpy(e.g, g.data(), n * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> taps;
    e.R = conv_taps(n, sigma, h, taps, e.nblk);
    if (e.R >= 0) {
        XC_TRY_HIP(ctx, hipMalloc(&e.taps, taps.size() * sizeof(double)));
        XC_TRY_HIP(ctx, hipMemcpy(e.taps, taps.data(), taps.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    nc->plans.push_back(e);
    return get_plan(ctx, n, sigma, h, out, cf);
}

// Filter three (or fewer) fields: in[f] -> out[f] using scratch tmp[f] between the passes.
int filter_fields(xc_ctx* ctx, int nf, const xc_mat* in, xc_mat* out, xc_mat* tmp, const int64_t* pad, double h,
                  const double* sigma) {
    const int Nz = (int)in[0].rows, Nx = (int)in[0].cols;
    int st;
    XPassArgs xa{};
    ZPassArgs za{};
    ConvXArgs cx{};
    ConvZArgs cz{};
    bool dx = true, dz = true;  // every field's pass has direct-convolution taps
    for (int f = 0; f < nf; ++f) {
        if ((st = get_plan(ctx, Nx, sigma[f], h, &xa.plan[f], &cx.f[f]))) return st;
        if ((st = get_plan(ctx, Nz + (int)pad[f], sigma[f], h, &za.plan[f], &cz.f[f]))) return st;
        dx = dx && cx.f[f].R >= 0 && cx.f[f].nblk <= CVX_MAXBLK && Nx >= CV_MINLINE;
        dz = dz && cz.f[f].R >= 0 && Nz + pad[f] >= CV_MINLINE;
        cx.f[f].in = in[f].p; cx.f[f].ld_in = in[f].ld; cx.f[f].out = tmp[f].p; cx.f[f].ld_out = tmp[f].ld;
        cz.f[f].in = tmp[f].p; cz.f[f].ld_in = tmp[f].ld; cz.f[f].out = out[f].p; cz.f[f].ld_out = out[f].ld;
        cz.n[f] = Nz + (int)pad[f];
        xa.io[f] = FieldIO{in[f].p, in[f].ld, tmp[f].p, tmp[f].ld};
        za.io[f] = FieldIO{tmp[f].p, tmp[f].ld, out[f].p, out[f].ld};
        za.hp[f] = (int)(pad[f] / 2);
        // one complex line (two columns) per workgroup: these transforms are latency-bound, and more
        // workgroups keep more CUs busy (the per-line arithmetic does not depend on the grouping)
        za.nl[f] = 1;
    }
    xa.Nz = za.Nz = Nz;
    xa.Nx = za.Nx = Nx;
    // the x plans of the fields are the same length; the power-of-two path when every field's plan has it
    int lgx = p2_fast_lg(xa.plan[0], 1);
    for (int f = 1; f < nf; ++f)
        if (p2_fast_lg(xa.plan[f], 1) != lgx) lgx = 0;
    xa.nl = 1;
    const int xblocks = (Nz + 2 * xa.nl - 1) / (2 * xa.nl);
    const dim3 gx(xblocks, nf);
    const size_t ldx = 2 * xa.nl * Nx * sizeof(double2);
    if (dx) {
        cx.Nz = Nz;
        cx.Nx = Nx;
        cx.nrc = (Nz + 63) / 64;
        cx.ncc = (Nx + CVX_C - 1) / CVX_C;
        int nbmax = 0;
        for (int f = 0; f < nf; ++f) nbmax = cx.f[f].nblk > nbmax ? cx.f[f].nblk : nbmax;
        const size_t lx = (size_t)(CVX_C + nbmax * CV_P) * 64 * sizeof(double);
        hipLaunchKernelGGL(k_gconv_x, dim3(cx.nrc * cx.ncc, nf), dim3(64 * CVX_W), lx, ctx->stream, cx);
    } else switch (lgx) {
        case 8: hipLaunchKernelGGL(k_filter_x<8>, gx, dim3(NT), ldx, ctx->stream, xa); break;
        case 9: hipLaunchKernelGGL(k_filter_x<9>, gx, dim3(NT), ldx, ctx->stream, xa); break;
        case 10: hipLaunchKernelGGL(k_filter_x<10>, gx, dim3(NT), ldx, ctx->stream, xa); break;
        default: hipLaunchKernelGGL(k_filter_x<0>, gx, dim3(NT), ldx, ctx->stream, xa); break;
    }
    int nlmin = za.nl[0], nmax = za.plan[0].n * za.nl[0];
    for (int f = 1; f < nf; ++f) {
        nlmin = za.nl[f] < nlmin ? za.nl[f] : nlmin;
        nmax = za.plan[f].n * za.nl[f] > nmax ? za.plan[f].n * za.nl[f] : nmax;
    }
    if (dz) {
        cz.Nz = Nz;
        cz.Nx = Nx;
        cz.nseg = (Nz + 64 * CV_P - 1) / (64 * CV_P);
        int nbmax = 0;
        for (int f = 0; f < nf; ++f) nbmax = cz.f[f].nblk > nbmax ? cz.f[f].nblk : nbmax;
        const size_t lz = (size_t)CV_P * (64 + nbmax) * sizeof(double);
        hipLaunchKernelGGL(k_gconv_z, dim3(Nx * cz.nseg, nf), dim3(64), lz, ctx->stream, cz);
        return XC_OK;
    }
    const int zblocks = (Nx + 2 * nlmin - 1) / (2 * nlmin);
    hipLaunchKernelGGL(k_filter_z, dim3(zblocks, nf), dim3(NT), 2 * nmax * sizeof(double2), ctx->stream, za);
    return XC_OK;
}

bool shp(const xc_mat& m, int64_t r, int64_t c) { return xc_shape(m, r, c); }

}  // namespace

extern "C" {

int xc_normal_fill(xc_ctx* ctx, xc_mat a, uint64_t seed, uint32_t stream, uint32_t draw) {
    XC_CHECK_CTX(ctx);
    if (!xc_mat_ok(a)) return xc_fail(ctx, XC_EINVAL_ARG, "normal_fill: bad matrix");
    const int64_t n = a.rows * a.cols;
    if (n == 0) return XC_OK;
    const int64_t pairs = (n + 1) / 2;
    hipLaunchKernelGGL(k_normals, dim3((unsigned)((pairs + NT - 1) / NT), 1), dim3(NT), 0, ctx->stream, a.p, a.ld,
                       a.p, a.ld, (int)a.rows, n, seed, draw, stream);
    return xc_finish(ctx, "normal_fill");
}

int xc_fft_filter(xc_ctx* ctx, xc_mat psi, xc_mat out, int64_t pad, double h, double sigma) {
    XC_CHECK_CTX(ctx);
    if (!xc_mat_ok(psi) || !xc_shape(out, psi.rows, psi.cols) || psi.rows < 2 || psi.cols < 1 || pad < 0)
        return xc_fail(ctx, XC_EINVAL_SHAPE, "compute_fft_filter: shapes");
    const int Nz = (int)psi.rows, Nx = (int)psi.cols;
    double* sc;
    const xc_dom d = xc_make_dom(Nz, Nx);
    const size_t ncell = (size_t)Nz * Nx;
    int st;
    if ((st = xc_scratch(ctx, XC_SLOT_NOISE, (ncell + 12 * (size_t)d.nb + 12) * sizeof(double), (void**)&sc))) return st;
    xc_mat tmp{sc, Nz, Nx, Nz};
    if ((st = filter_fields(ctx, 1, &psi, &out, &tmp, &pad, h, &sigma))) return st;
    // statistics over (raw, filtered) pairs; unused slots point at the same arrays
    StatArgs sa{};
    for (int q = 0; q < 3; ++q) {
        sa.a[q] = psi.p; sa.ld[q] = psi.ld;
        sa.a[3 + q] = out.p; sa.ld[3 + q] = out.ld;
    }
    sa.Nz = Nz; sa.Nx = Nx; sa.d = d;
    sa.part = sc + ncell;
    sa.stats = sa.part + 12 * d.nb;
    hipLaunchKernelGGL(k_stat_partial, dim3(d.nb), dim3(NT), 0, ctx->stream, sa);
    hipLaunchKernelGGL(k_stat_final, dim3(1), dim3(12 * 64), 0, ctx->stream, sa, (double)ncell);
    // rescale to the raw mean/std (field 0's statistics), elementwise in place
    hipLaunchKernelGGL(k_resc_inplace, dim3(d.nb), dim3(NT), 0, ctx->stream, out.p, out.ld, sa.stats, Nz, Nx, d);
    return xc_finish(ctx, "compute_fft_filter");
}

int xc_store_noise(xc_ctx* ctx, xc_noise* ns) {
    XC_CHECK_CTX(ctx);
    const int64_t Nz = ns->psie.rows, Nx = ns->psie.cols;
    if (!shp(ns->psieo, Nz, Nx) || !shp(ns->psix, Nz, Nx) || !shp(ns->psixo, Nz, Nx) || !shp(ns->psis, Nz, Nx) ||
        !shp(ns->psiso, Nz, Nx))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "store_noise!: NoiseState shapes");
    const xc_dom d = xc_make_dom(Nz, Nx);
    hipLaunchKernelGGL(k_copy3, dim3(d.nb), dim3(NT), 0, ctx->stream, ns->psie.p, ns->psie.ld, ns->psieo.p,
                       ns->psieo.ld, ns->psix.p, ns->psix.ld, ns->psixo.p, ns->psixo.ld, ns->psis.p, ns->psis.ld,
                       ns->psiso.p, ns->psiso.ld, (int)Nz, (int)Nx, d);
    return xc_finish(ctx, "store_noise!");
}

int xc_noise_step(xc_ctx* ctx, xc_noise* ns, const xc_phase* ph, const xc_grid* g, const xc_par* par,
                  const xc_scales* sc, double dt, int first_iter) {
    XC_CHECK_CTX(ctx);
    if (!ns || !ph || !g || !par || !sc) return xc_fail(ctx, XC_EINVAL_ARG, "noise!: NULL argument");
    const int64_t Nz = g->Nz, Nx = g->Nx;
    const xc_mat* cellm[] = {&ns->psie, &ns->psix, &ns->psis, &ns->psieo, &ns->psixo, &ns->psiso, &ns->re, &ns->rs,
                             &ns->xie, &ns->xix, &ns->xis, &ph->V, &ph->vx, &ph->fReL, &ph->ke, &ph->chi, &ph->ks};
    for (auto m : cellm)
        if (!shp(*m, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "noise!: cell-centred fields must be (Nz, Nx)");
    if (!shp(ns->xiew, Nz + 1, Nx + 2) || !shp(ns->xixw, Nz + 1, Nx + 2) || !shp(ns->xisw, Nz + 1, Nx + 2) ||
        !shp(ns->xieu, Nz + 2, Nx + 1) || !shp(ns->xixu, Nz + 2, Nx + 1) || !shp(ns->xisu, Nz + 2, Nx + 1) ||
        !shp(ph->chiw, Nz + 1, Nx + 2) || !shp(ph->muw, Nz + 1, Nx + 2) || !shp(ph->chiu, Nz + 2, Nx + 1) ||
        !shp(ph->muu, Nz + 2, Nx + 1))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "noise!: staggered noise/volume-fraction fields");
    if (Nz < 2 || ns->padL0 < 0 || ns->padl0 < 0) return xc_fail(ctx, XC_EINVAL_ARG, "noise!: bad grid/padding");
    const xc_dom d = xc_make_dom(Nz, Nx);
    // 1. white noise (noise.jl:154-157)
    if (first_iter) {
        const int64_t n = Nz * Nx, pairs = (n + 1) / 2;
        if (ns->re.ld != ns->rs.ld) {
            hipLaunchKernelGGL(k_normals, dim3((unsigned)((pairs + NT - 1) / NT), 1), dim3(NT), 0, ctx->stream,
                               ns->re.p, ns->re.ld, ns->re.p, ns->re.ld, (int)Nz, n, ns->seed, (uint32_t)ns->draws, 0u);
            hipLaunchKernelGGL(k_normals, dim3((unsigned)((pairs + NT - 1) / NT), 1), dim3(NT), 0, ctx->stream,
                               ns->rs.p, ns->rs.ld, ns->rs.p, ns->rs.ld, (int)Nz, n, ns->seed, (uint32_t)ns->draws, 1u);
        } else {
            hipLaunchKernelGGL(k_normals, dim3((unsigned)((pairs + NT - 1) / NT), 2), dim3(NT), 0, ctx->stream,
                               ns->re.p, ns->re.ld, ns->rs.p, ns->rs.ld, (int)Nz, n, ns->seed, (uint32_t)ns->draws, 0u);
        }
        ns->draws += 1;
    }
    // 2. OU update (noise.jl:159-183)
    const double h = g->h, L0h = sc->L0h, l0h = sc->l0h;
    OUArgs oa{};
    oa.V = ph->V.p; oa.ldV = ph->V.ld; oa.vx = ph->vx.p; oa.ldvx = ph->vx.ld;
    oa.fReL = ph->fReL.p; oa.ldf = ph->fReL.ld; oa.ke = ph->ke.p; oa.ldk = ph->ke.ld;
    oa.chi = ph->chi.p; oa.ldc = ph->chi.ld; oa.ks = ph->ks.p; oa.ldks = ph->ks.ld;
    oa.psie = ns->psie.p; oa.lde = ns->psie.ld; oa.psix = ns->psix.p; oa.ldx = ns->psix.ld;
    oa.psis = ns->psis.p; oa.lds = ns->psis.ld;
    oa.psieo = ns->psieo.p; oa.ldeo = ns->psieo.ld; oa.psixo = ns->psixo.p; oa.ldxo = ns->psixo.ld;
    oa.psiso = ns->psiso.p; oa.ldso = ns->psiso.ld; oa.re = ns->re.p; oa.ldre = ns->re.ld;
    oa.rs = ns->rs.p; oa.ldrs = ns->rs.ld;
    oa.Xi = par->Xi; oa.L0h = L0h; oa.l0h = l0h; oa.txi0 = sc->txi0; oa.dt = dt;
    oa.fL = 2.0 / sqrt(1.0 - exp(-(h * h) / (2 * (L0h * L0h))));
    oa.fl = 2.0 / sqrt(1.0 - exp(-(h * h) / (2 * (l0h * l0h))));
    oa.Nz = (int)Nz; oa.Nx = (int)Nx; oa.d = d;
    hipLaunchKernelGGL(k_ou, dim3(d.nb), dim3(NT), 0, ctx->stream, oa);
    // 3. filters (noise.jl:186-188)
    const size_t ncell = (size_t)Nz * Nx;
    double* sb;
    int st;
    if ((st = xc_scratch(ctx, XC_SLOT_NOISE, (6 * ncell + 12 * (size_t)d.nb + 12) * sizeof(double), (void**)&sb)))
        return st;
    xc_mat in[3] = {ns->psie, ns->psix, ns->psis};
    xc_mat tmp[3], flt[3];
    for (int f = 0; f < 3; ++f) {
        tmp[f] = xc_mat{sb + f * ncell, Nz, Nx, Nz};
        flt[f] = xc_mat{sb + (3 + f) * ncell, Nz, Nx, Nz};
    }
    const int64_t pads[3] = {ns->padL0, ns->padL0, ns->padl0};
    const double sig[3] = {L0h, L0h, l0h};
    if ((st = filter_fields(ctx, 3, in, flt, tmp, pads, h, sig))) return st;
    StatArgs sa{};
    for (int f = 0; f < 3; ++f) {
        sa.a[f] = in[f].p; sa.ld[f] = in[f].ld;
        sa.a[3 + f] = flt[f].p; sa.ld[3 + f] = flt[f].ld;
    }
    sa.Nz = (int)Nz; sa.Nx = (int)Nx; sa.d = d;
    sa.part = sb + 6 * ncell;
    sa.stats = sa.part + 12 * d.nb;
    hipLaunchKernelGGL(k_stat_partial, dim3(d.nb), dim3(NT), 0, ctx->stream, sa);
    hipLaunchKernelGGL(k_stat_final, dim3(1), dim3(12 * 64), 0, ctx->stream, sa, (double)ncell);
    // 4. fluxes and magnitudes (noise.jl:190-236)
    FluxArgs fa{};
    for (int f = 0; f < 3; ++f) { fa.flt[f] = flt[f].p; fa.ldf[f] = flt[f].ld; }
    fa.st = sa.stats;
    fa.chiw = ph->chiw.p; fa.ldchiw = ph->chiw.ld; fa.muw = ph->muw.p; fa.ldmuw = ph->muw.ld;
    fa.chiu = ph->chiu.p; fa.ldchiu = ph->chiu.ld; fa.muu = ph->muu.p; fa.ldmuu = ph->muu.ld;
    fa.xiew = ns->xiew.p; fa.ldew = ns->xiew.ld; fa.xieu = ns->xieu.p; fa.ldeu = ns->xieu.ld;
    fa.xixw = ns->xixw.p; fa.ldxw = ns->xixw.ld; fa.xixu = ns->xixu.p; fa.ldxu = ns->xixu.ld;
    fa.xisw = ns->xisw.p; fa.ldsw = ns->xisw.ld; fa.xisu = ns->xisu.p; fa.ldsu = ns->xisu.ld;
    fa.Nz = (int)Nz; fa.Nx = (int)Nx; fa.d = xc_make_dom(Nz + 2, Nx + 2);
    hipLaunchKernelGGL(k_noise_flux, dim3(fa.d.nb), dim3(NT), 0, ctx->stream, fa);
    hipLaunchKernelGGL(k_noise_mag, dim3(d.nb), dim3(NT), 0, ctx->stream, ns->xisw.p, ns->xisw.ld, ns->xisu.p,
                       ns->xisu.ld, ns->xixw.p, ns->xixw.ld, ns->xixu.p, ns->xixu.ld, ns->xiew.p, ns->xiew.ld,
                       ns->xieu.p, ns->xieu.ld, ns->xis.p, ns->xis.ld, ns->xix.p, ns->xix.ld, ns->xie.p, ns->xie.ld,
                       (int)Nz, (int)Nx, d);
    return xc_finish(ctx, "noise!");
}

}  // extern "C"
