// xc_fft.h -- batched 1-D complex FFTs held in LDS (device code shared by the
// noise filter and the pressure-Poisson part of the Stokes preconditioner).
//
// Stockham autosort, mixed radix 2/3/4/5/7/11/13 plus a generic stage for any
// remaining prime factor; `nl` lines of length n sit contiguously in LDS and
// the workgroup (FFT_NT lanes) cooperates on every stage.  Twiddles
// W_n^t = exp(-2 pi i t / n) come from a table computed on the host in long
// double.  Forward transforms are unnormalised, inverse ones use conj(W).
#ifndef XC_FFT_H
#define XC_FFT_H
#include <hip/hip_runtime.h>
#include <math.h>

#include <vector>

namespace {

constexpr int FFT_NT = 256;   // workgroup size of every kernel that calls lds_fft
constexpr int MAXLINE = 2048; // complex points per LDS buffer (64 KiB ping-pong)
constexpr int MAXST = 24;

struct Plan {
    int n, nst;
    int rad[MAXST];
    const double2* tw;  // W_n^t = exp(-2 pi i t / n), t in [0, n)
    const double* g;    // filter multiplier exp(-s^2 k_t^2) / n, t in [0, n)
    int lg1;            // log2(n) + 1 when n is a power of two factored into radices 4 and 2, else 0
};

// host: mark a completed plan whose length is a power of two (index arithmetic by shifts and masks)
inline void plan_mark_pow2(Plan& p) {
    p.lg1 = 0;
    if (p.n < 1 || (p.n & (p.n - 1))) return;
    for (int s = 0; s < p.nst; ++s)
        if (p.rad[s] != 2 && p.rad[s] != 4) return;
    int lg = 0;
    while ((1 << lg) < p.n) ++lg;
    p.lg1 = lg + 1;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// One Stockham stage of radix R over nl lines of length n (src -> dst).  P2: n and Ns are powers of two
// (lg = log2 n, lgNs = log2 Ns) and the index arithmetic is shifts and masks instead of divisions.
template <int R, bool INV, bool P2 = false>
__device__ __forceinline__ void stage(const double2* __restrict__ src, double2* __restrict__ dst, int n, int nl,
                                      int Ns, const double2* __restrict__ tw, int lg = 0, int lgNs = 0) {
    constexpr int LR = R == 2 ? 1 : (R == 4 ? 2 : 0);
    static_assert(!P2 || LR > 0, "power-of-two stages are radix 2 or 4");
    const int nR = P2 ? n >> LR : n / R;
    const int strideT = P2 ? 1 << (lg - lgNs - LR) : n / (Ns * R);
    double2 wr[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
        double2 w = tw[m * nR];  // W_R^m = W_n^{m n / R}
        wr[m] = INV ? make_double2(w.x, -w.y) : w;
    }
    for (int k = threadIdx.x; k < nl * nR; k += FFT_NT) {
        const int l = P2 ? k >> (lg - LR) : k / nR, j = k - l * nR;
        const int base = l * n;
        const int jm = P2 ? j & (Ns - 1) : j % Ns;
        const int t0 = jm * strideT;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double2 x = src[base + j + r * nR];
            if (r > 0) {
                double2 w = tw[P2 ? (t0 * r) & (n - 1) : (t0 * r) % n];
                if (INV) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        double2 o[R];
        if constexpr (R == 2) {
            o[0] = cadd(v[0], v[1]);
            o[1] = csub(v[0], v[1]);
        } else if constexpr (R == 4) {
            const double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
            const double2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
            // forward: W_4 = -i
            const double2 jd13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
            o[0] = cadd(s02, s13);
            o[2] = csub(s02, s13);
            o[1] = cadd(d02, jd13);
            o[3] = csub(d02, jd13);
        } else {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                double2 s = v[0];
#pragma unroll
                for (int r = 1; r < R; ++r) s = cadd(s, cmul(v[r], wr[(r * q) % R]));
                o[q] = s;
            }
        }
        const int idxD = P2 ? ((j >> lgNs) << (lgNs + LR)) + jm : (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + idxD + r * Ns] = o[r];
    }
}

// Stage of any radix R (remaining prime factors): one output per lane,
// out[q] = sum_r x[j + r n/R] W_n^{r (t0 + q n/R)}.
template <bool INV>
__device__ __forceinline__ void stage_generic(const double2* __restrict__ src, double2* __restrict__ dst, int n,
                                              int nl, int Ns, int R, const double2* __restrict__ tw) {
    const int nR = n / R;
    const int strideT = n / (Ns * R);
    for (int k = threadIdx.x; k < nl * n; k += FFT_NT) {
        const int l = k / n, rem = k - l * n;
        const int j = rem / R, q = rem - j * R;
        const int base = l * n;
        const int jm = j % Ns;
        const int e = jm * strideT + q * nR;
        double2 s = make_double2(0.0, 0.0);
        int t = 0;
        for (int r = 0; r < R; ++r) {
            double2 w = tw[t];
            if (INV) w.y = -w.y;
            s = cadd(s, cmul(src[base + j + r * nR], w));
            t += e;
            if (t >= n) t -= n * (t / n);
        }
        dst[base + (j / Ns) * Ns * R + jm + q * Ns] = s;
    }
}

template <bool INV>
__device__ double2* lds_fft(double2* a, double2* b, int nl, const Plan& p) {
    double2* src = a;
    double2* dst = b;
    int Ns = 1;
    if (p.lg1 > 0) {  // power of two: radix-4 and radix-2 stages with shift/mask indexing
        const int lg = p.lg1 - 1;
        int lgNs = 0;
        for (int s = 0; s < p.nst; ++s) {
            if (p.rad[s] == 4) stage<4, INV, true>(src, dst, p.n, nl, Ns, p.tw, lg, lgNs);
            else stage<2, INV, true>(src, dst, p.n, nl, Ns, p.tw, lg, lgNs);
            __syncthreads();
            double2* t = src;
            src = dst;
            dst = t;
            const int lr = p.rad[s] == 4 ? 2 : 1;
            Ns <<= lr;
            lgNs += lr;
        }
        return src;
    }
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.rad[s];
        switch (R) {
            case 2: stage<2, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 3: stage<3, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 4: stage<4, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 5: stage<5, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 7: stage<7, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 11: stage<11, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            case 13: stage<13, INV>(src, dst, p.n, nl, Ns, p.tw); break;
            default: stage_generic<INV>(src, dst, p.n, nl, Ns, R, p.tw); break;
        }
        __syncthreads();
        double2* t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

// Power-of-two transforms of length 2^LG in which every stage has at most FFT_NT butterflies over the nl
// lines (one per lane): the lane's twiddles of every stage are fetched once, up front (fft_p2_twiddles,
// issued with the caller's data loads), instead of at each stage.  The stages are those fft_factorize
// gives (LG/2 of radix 4, then one of radix 2 when LG is odd), with stage<R, INV, true>'s arithmetic, so
// the results are the same.  LG = 0: not this path.
template <int LG>
struct P2Fast {
    static constexpr int N4 = LG / 2, N2 = LG & 1, NST = N4 + N2 > 0 ? N4 + N2 : 1;
};
// host: LG for a plan and nl lines if the fast path applies (8 <= LG <= 10), else 0
inline int p2_fast_lg(const Plan& p, int nl) {
    const int lg = p.lg1 - 1;
    if (p.lg1 <= 0 || lg < 8 || lg > 10) return 0;
    for (int s = 0; s < p.nst; ++s)
        if (p.rad[s] != (s < lg / 2 ? 4 : 2)) return 0;
    if ((int64_t)nl * ((1 << lg) >> ((lg & 1) ? 1 : 2)) > FFT_NT) return 0;
    return lg;
}
template <int LG>
__device__ __forceinline__ void fft_p2_twiddles(int nl, const double2* __restrict__ tw,
                                                double2 (&w)[P2Fast<LG>::NST][3]) {
    constexpr int n = 1 << LG;
    const int k = threadIdx.x;
    int lgNs = 0;
#pragma unroll
    for (int s = 0; s < P2Fast<LG>::NST; ++s) {
        const int LR = s < P2Fast<LG>::N4 ? 2 : 1;
        const int nR = n >> LR;
        const int kk = k < nl * nR ? k : 0;
        const int l = kk >> (LG - LR), j = kk - l * nR;
        const int jm = j & ((1 << lgNs) - 1);
        const int t0 = jm << (LG - lgNs - LR);
#pragma unroll
        for (int r = 1; r < 4; ++r)
            if (r < (1 << LR)) w[s][r - 1] = tw[(t0 * r) & (n - 1)];
        lgNs += LR;
    }
}
template <int R, bool INV, int LG>
__device__ __forceinline__ void stage_p2_one(const double2* __restrict__ src, double2* __restrict__ dst, int nl,
                                             int lgNs, const double2 (&wr)[3]) {
    constexpr int LR = R == 2 ? 1 : 2;
    constexpr int n = 1 << LG, nR = n >> LR;
    const int k = threadIdx.x;
    if (k >= nl * nR) return;
    const int l = k >> (LG - LR), j = k - l * nR;
    const int base = l * n;
    const int jm = j & ((1 << lgNs) - 1);
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        double2 x = src[base + j + r * nR];
        if (r > 0) {
            double2 w = wr[r - 1];
            if (INV) w.y = -w.y;
            x = cmul(x, w);
        }
        v[r] = x;
    }
    double2 o[R];
    if constexpr (R == 2) {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
    } else {
        const double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        const double2 s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
        const double2 jd13 = INV ? make_double2(-d13.y, d13.x) : make_double2(d13.y, -d13.x);
        o[0] = cadd(s02, s13);
        o[2] = csub(s02, s13);
        o[1] = cadd(d02, jd13);
        o[3] = csub(d02, jd13);
    }
    const int idxD = ((j >> lgNs) << (lgNs + LR)) + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + idxD + r * (1 << lgNs)] = o[r];
}
template <bool INV, int LG>
__device__ double2* lds_fft_p2(double2* a, double2* b, int nl, const double2 (&w)[P2Fast<LG>::NST][3]) {
    double2* src = a;
    double2* dst = b;
    int lgNs = 0;
#pragma unroll
    for (int s = 0; s < P2Fast<LG>::NST; ++s) {
        if (s < P2Fast<LG>::N4) {
            stage_p2_one<4, INV, LG>(src, dst, nl, lgNs, w[s]);
            lgNs += 2;
        } else {
            stage_p2_one<2, INV, LG>(src, dst, nl, lgNs, w[s]);
            lgNs += 1;
        }
        __syncthreads();
        double2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// host: radix sequence of n (false if it needs more than MAXST stages)
inline bool fft_factorize(int n, std::vector<int>& rad) {
    rad.clear();
    static const int order[] = {4, 2, 3, 5, 7, 11, 13};
    for (int r : order)
        while (n % r == 0) {
            rad.push_back(r);
            n /= r;
        }
    for (int p = 17; n > 1 && p <= n; p += 2)  // remaining primes: generic stages
        while (n % p == 0) {
            rad.push_back(p);
            n /= p;
        }
    return n == 1 && (int)rad.size() <= MAXST;
}

// host: W_n^t for t in [0, n), computed in long double
inline void fft_twiddles(int n, std::vector<double2>& tw) {
    tw.resize(n);
    for (int t = 0; t < n; ++t) {
        const long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
        tw[t] = make_double2((double)cosl(ang), (double)sinl(ang));
    }
}

}  // namespace
#endif
