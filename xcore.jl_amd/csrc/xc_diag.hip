// Diagnostic streaming kernel (no reference counterpart): the memory traffic of a stencil launch
// without its arithmetic, to measure the bandwidth HBM sustains for that mix of streams.  Each of
// nw outputs receives, at every cell (i, j) of out[0]'s shape, the sum of the nr inputs there;
// outputs k >= 1 are written one row down (the padded flux arrays' layout).  Mode bit 0: non-temporal
// stores (as the stencils use for launches larger than the last-level cache); bit 1: the aligned 16-byte
// form, outputs not shifted; bit 2: non-temporal loads.
#include "xc_internal.h"

namespace {

constexpr int MAXS = 3;
struct StreamArgs {
    const double* in[MAXS]; int64_t ld_in[MAXS];
    double* out[MAXS]; int64_t ld_out[MAXS];
    int Nz, Nx, nbz;
};

template <int NR, int NW, bool NTS, bool NTL>
__global__ __launch_bounds__(256) void k_stream(StreamArgs a) {
    const int t = blockIdx.x;
    const int iz = (t % a.nbz) * 64 + (threadIdx.x & 63);
    const int ix = (t / a.nbz) * 4 + (threadIdx.x >> 6);
    if (iz >= a.Nz || ix >= a.Nx) return;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const double* p = &a.in[r][iz + (int64_t)ix * a.ld_in[r]];
        s += NTL ? __builtin_nontemporal_load(p) : *p;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        double* p = &a.out[w][iz + (w ? 1 : 0) + (int64_t)ix * a.ld_out[w]];
        if constexpr (NTS) __builtin_nontemporal_store(s, p);
        else *p = s;
    }
}

// The best case of the same traffic: every stream aligned (outputs k >= 1 not shifted), two consecutive
// elements a lane as one 16-byte access, 128 rows a wave: what HBM sustains for NR reads and NW writes
// when every access is a whole, aligned line.
typedef double dbl2 __attribute__((ext_vector_type(2)));
template <int NR, int NW, bool NTS, bool NTL>
__global__ __launch_bounds__(256) void k_stream_vec(StreamArgs a) {
    const int t = blockIdx.x;
    const int iz = (t % a.nbz) * 128 + 2 * (threadIdx.x & 63);
    const int ix = (t / a.nbz) * 4 + (threadIdx.x >> 6);
    if (iz >= a.Nz || ix >= a.Nx) return;
    dbl2 s = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const dbl2* p = reinterpret_cast<const dbl2*>(a.in[r] + iz + (int64_t)ix * a.ld_in[r]);
        s += NTL ? __builtin_nontemporal_load(p) : *p;
    }
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        dbl2* p = reinterpret_cast<dbl2*>(a.out[w] + iz + (int64_t)ix * a.ld_out[w]);
        if constexpr (NTS) __builtin_nontemporal_store(s, p);
        else *p = s;
    }
}

template <int NR, int NW, bool NTL>
void launch_nr_nw_l(const StreamArgs& a, int mode, hipStream_t st) {
    const bool nts = mode & 1;
    if (mode & 2) {
        StreamArgs b = a;
        b.nbz = (a.Nz + 127) / 128;
        const dim3 grid((unsigned)(b.nbz * ((a.Nx + 3) / 4)));
        if (nts) hipLaunchKernelGGL((k_stream_vec<NR, NW, true, NTL>), grid, dim3(256), 0, st, b);
        else hipLaunchKernelGGL((k_stream_vec<NR, NW, false, NTL>), grid, dim3(256), 0, st, b);
        return;
    }
    const dim3 grid((unsigned)(a.nbz * ((a.Nx + 3) / 4)));
    if (nts) hipLaunchKernelGGL((k_stream<NR, NW, true, NTL>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_stream<NR, NW, false, NTL>), grid, dim3(256), 0, st, a);
}
template <int NR, int NW>
void launch_nr_nw(const StreamArgs& a, int mode, hipStream_t st) {
    if (mode & 4) launch_nr_nw_l<NR, NW, true>(a, mode, st);
    else launch_nr_nw_l<NR, NW, false>(a, mode, st);
}

}  // namespace

extern "C" int xc_diag_stream(xc_ctx* ctx, int nr, const xc_mat* in, int nw, const xc_mat* out, int nts) {
    XC_CHECK_CTX(ctx);
    if (nr < 0 || nr > MAXS || nw < 1 || nw > MAXS || (nr && !in) || !out)
        return xc_fail(ctx, XC_EINVAL_ARG, "diag_stream: 0..3 inputs and 1..3 outputs");
    const int64_t Nz = out[0].rows, Nx = out[0].cols;
    if (!xc_mat_ok(out[0]) || Nz < 1 || Nx < 1 || Nz > (1 << 30) || Nx > (1 << 30))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "diag_stream: out[0] must be a non-empty matrix");
    StreamArgs a{};
    for (int r = 0; r < nr; ++r) {
        if (!xc_mat_ok(in[r]) || in[r].rows < Nz || in[r].cols < Nx)
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diag_stream: input %d smaller than out[0]", r);
        a.in[r] = in[r].p; a.ld_in[r] = in[r].ld;
    }
    for (int w = 0; w < nw; ++w) {
        if (!xc_mat_ok(out[w]) || out[w].rows < Nz + ((w && !(nts & 2)) ? 1 : 0) || out[w].cols < Nx)
            return xc_fail(ctx, XC_EINVAL_SHAPE, "diag_stream: output %d too small (outputs 1.. need a row more)", w);
        a.out[w] = out[w].p; a.ld_out[w] = out[w].ld;
    }
    if (nts & 2) {  // aligned 16-byte form: even rows, 16-byte aligned columns, outputs not shifted
        bool ok = Nz % 2 == 0;
        for (int r = 0; r < nr; ++r) ok = ok && ((uintptr_t)in[r].p % 16 == 0) && in[r].ld % 2 == 0;
        for (int w = 0; w < nw; ++w) ok = ok && ((uintptr_t)out[w].p % 16 == 0) && out[w].ld % 2 == 0;
        if (!ok) return xc_fail(ctx, XC_EINVAL_ARG, "diag_stream: the aligned form needs even rows and 16-byte columns");
    }
    a.Nz = (int)Nz; a.Nx = (int)Nx; a.nbz = (int)((Nz + 63) / 64);
    if ((int64_t)a.nbz * ((Nx + 3) / 4) > (1ll << 31) - 1) return xc_fail(ctx, XC_EINVAL_SHAPE, "diag_stream: grid too large");
    hipStream_t st = ctx->stream;
    switch (nr * 4 + nw) {
        case 1: launch_nr_nw<0, 1>(a, nts, st); break;
        case 2: launch_nr_nw<0, 2>(a, nts, st); break;
        case 3: launch_nr_nw<0, 3>(a, nts, st); break;
        case 5: launch_nr_nw<1, 1>(a, nts, st); break;
        case 6: launch_nr_nw<1, 2>(a, nts, st); break;
        case 7: launch_nr_nw<1, 3>(a, nts, st); break;
        case 9: launch_nr_nw<2, 1>(a, nts, st); break;
        case 10: launch_nr_nw<2, 2>(a, nts, st); break;
        case 11: launch_nr_nw<2, 3>(a, nts, st); break;
        case 13: launch_nr_nw<3, 1>(a, nts, st); break;
        case 14: launch_nr_nw<3, 2>(a, nts, st); break;
        default: launch_nr_nw<3, 3>(a, nts, st); break;
    }
    return xc_finish(ctx, "diag_stream");
}
