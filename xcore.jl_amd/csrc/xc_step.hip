// xc_step.hip -- one outer time step of run! (src/run.jl:204-275) on the device,
// with the Picard residual of diagnose.jl:42-71, the reductions of
// record_history! (history.jl:101-222) and initialize! (run.jl:71-140).
//
// The whole sweep sequence (snapshot!, phsevo!, fluidmech!, noise!,
// update_phase_velocities!, update!, compute_resnorm) is enqueued on the
// context's stream; the host only reads back the handful of norms that decide
// whether the Picard loop continues (one small copy per sweep) and the five
// maxima of update_dt.
#include <math.h>

#include "xc_internal.h"

namespace {

constexpr int NT = 256;

// snapshot! (diagnose.jl:42-49): W, U, P, X, MFS into contiguous scratch
__global__ __launch_bounds__(NT) void k_snap(xc_mat W, xc_mat U, xc_mat P, xc_mat X, xc_mat MFS, double* dst,
                                             int Nz, int Nx) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1), nP = (int64_t)(Nz + 2) * (Nx + 2),
                  nC = (int64_t)Nz * Nx;
    auto g = [](const xc_mat& m, int64_t q) { return xc_mat_at(m, q); };
    if (k < nW) dst[k] = g(W, k);
    else if (k < nW + nU) dst[k] = g(U, k - nW);
    else if (k < nW + nU + nP) dst[k] = g(P, k - nW - nU);
    else if (k < nW + nU + nP + nC) dst[k] = g(X, k - nW - nU - nP);
    else if (k < nW + nU + nP + 2 * nC) dst[k] = g(MFS, k - nW - nU - nP - nC);
}

// sums of squares for compute_resnorm (diagnose.jl:57-71):
// [0] |X-Xo|^2 [1] |rho|^2 [2] |MFS-MFSo|^2 [3] |W-Wo|^2 [4] |W|^2 [5] |U-Uo|^2 [6] |U|^2 [7] |P-Po|^2 [8] |P|^2
__global__ __launch_bounds__(NT) void k_resnorm_part(xc_mat W, xc_mat U, xc_mat P, xc_mat X, xc_mat MFS, xc_mat rho,
                                                     const double* snap, int Nz, int Nx, double* part) {
    const int64_t nW = (int64_t)(Nz + 1) * (Nx + 2), nU = (int64_t)(Nz + 2) * (Nx + 1), nP = (int64_t)(Nz + 2) * (Nx + 2),
                  nC = (int64_t)Nz * Nx;
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    auto g = [](const xc_mat& m, int64_t q) { return xc_mat_at(m, q); };
    // element q of every block at once: the loads (clamped indices) are all in flight together
    const int64_t nmax = nW > nU ? (nW > nP ? nW : nP) : (nU > nP ? nU : nP);
    const int64_t oU = nW, oP = nW + nU, oX = nW + nU + nP, oM = oX + nC;
    for (int64_t q = (int64_t)blockIdx.x * NT + threadIdx.x; q < nmax; q += (int64_t)gridDim.x * NT) {
        const int64_t qw = q < nW ? q : nW - 1, qu = q < nU ? q : nU - 1, qp = q < nP ? q : nP - 1,
                      qc = q < nC ? q : nC - 1;
        const double w = g(W, qw), sw = snap[qw], u = g(U, qu), su = snap[oU + qu], pp = g(P, qp),
                     sp = snap[oP + qp], x = g(X, qc), sx = snap[oX + qc], rh = g(rho, qc), mf = g(MFS, qc),
                     sm = snap[oM + qc];
        if (q < nW) { const double d = w - sw; a[3] += d * d; a[4] += w * w; }
        if (q < nU) { const double d = u - su; a[5] += d * d; a[6] += u * u; }
        if (q < nP) { const double d = pp - sp; a[7] += d * d; a[8] += pp * pp; }
        if (q < nC) {
            const double d = x - sx, e = mf - sm;
            a[0] += d * d;
            a[1] += rh * rh;
            a[2] += e * e;
        }
    }
    // wave sums (DPP), then the NT / 64 wave sums of each quantity
    __shared__ double red[9][NT / 64];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        const double t = wave_total(a[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) t += red[threadIdx.x][w];
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// out[q] = sum of the nb partials of row q, one wave per row (blockDim.x = 64 nq)
__global__ __launch_bounds__(1024) void k_sum_rows(const double* part, int nb, int nq, double* out) {
    const int q = threadIdx.x >> 6;
    if (q >= nq) return;
    const double t = xc_rows_reduce<false>(part, nb, q);
    if ((threadIdx.x & 63) == 0) out[q] = t;
}

// history.jl:122-151 sums: [0] sum rho [1] sum M [2] sum X [3] sum Gx
// [4..11] top/bot line sums of qz_advn_X, qz_advn_M, qz_dffn_X, qz_dffn_M (interior columns)
constexpr int NHS = 12;
__global__ __launch_bounds__(NT) void k_hist_part(xc_mat rho, xc_mat M, xc_mat X, xc_mat Gx, xc_mat qaX, xc_mat qaM,
                                                  xc_mat qdX, xc_mat qdM, int Nz, int Nx, double* part) {
    double a[NHS] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t nC = (int64_t)Nz * Nx;
    auto g = [](const xc_mat& m, int r, int c) { return m.p[r + (int64_t)c * m.ld]; };
    for (int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x; k < nC; k += (int64_t)gridDim.x * NT) {
        const int r = (int)(k % Nz), c = (int)(k / Nz);
        a[0] += g(rho, r, c);
        a[1] += g(M, r, c);
        a[2] += g(X, r, c);
        a[3] += g(Gx, r, c);
        if (r == 0) {
            a[4] += g(qaX, 0, c + 1); a[5] += g(qaX, Nz, c + 1);
            a[6] += g(qaM, 0, c + 1); a[7] += g(qaM, Nz, c + 1);
            a[8] += g(qdX, 0, c + 1); a[9] += g(qdX, Nz, c + 1);
            a[10] += g(qdM, 0, c + 1); a[11] += g(qdM, Nz, c + 1);
        }
    }
    __shared__ double red[NHS][NT / 64];
#pragma unroll
    for (int q = 0; q < NHS; ++q) {
        const double t = wave_total(a[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = t;
    }
    __syncthreads();
    if (threadIdx.x < NHS) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) t += red[threadIdx.x][w];
        part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

// history.jl:184-221 field statistics.  For each of the NHF cell fields (order of xc_history_stats' doc):
// [0] min [1] max [2] sum [3] sum of squares [4] sum of (a - a[0]) [5] sum of (a - a[0])^2
// [6] sum of log(max(a, eps)).  The shifted sums give the n-1 standard deviation without the
// cancellation of the raw second moment.
constexpr int NHF = 19, NHQ = 7, HST_NB = 64;
struct HistFields { xc_mat f[NHF]; };

__device__ __forceinline__ double wave_max(double v) { return -wave_min(-v); }

__global__ __launch_bounds__(NT) void k_hstat_part(HistFields hf, int Nz, int Nx, double* part) {
    const xc_mat m = hf.f[blockIdx.y];
    const int64_t nC = (int64_t)Nz * Nx;
    const double K = m.p[0], ep = 2.220446049250313e-16;
    double mn = INFINITY, mx = -INFINITY, s1 = 0, s2 = 0, d1 = 0, d2 = 0, sl = 0;
    for (int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x; k < nC; k += (int64_t)gridDim.x * NT) {
        const double a = xc_mat_at(m, k);
        mn = fmin(mn, a);
        mx = fmax(mx, a);
        s1 += a;
        s2 += a * a;
        const double d = a - K;
        d1 += d;
        d2 += d * d;
        sl += log(fmax(a, ep));
    }
    double v[NHQ] = {wave_min(mn), wave_max(mx), wave_total(s1), wave_total(s2), wave_total(d1), wave_total(d2),
                     wave_total(sl)};
    __shared__ double red[NHQ][NT / 64];
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int q = 0; q < NHQ; ++q) red[q][threadIdx.x >> 6] = v[q];
    __syncthreads();
    if (threadIdx.x < NHQ) {
        const int q = threadIdx.x;
        double t = red[q][0];
#pragma unroll
        for (int w = 1; w < NT / 64; ++w)
            t = q == 0 ? fmin(t, red[q][w]) : q == 1 ? fmax(t, red[q][w]) : t + red[q][w];
        part[((int64_t)blockIdx.y * NHQ + q) * gridDim.x + blockIdx.x] = t;
    }
}

// the NHF x NHQ rows of partials (nb each) into out: one wave per row, min / max / sum by the row's kind
__global__ __launch_bounds__(64) void k_hstat_final(const double* part, int nb, double* out) {
    const int row = blockIdx.x, q = row % NHQ;
    double v;
    if (q == 0) v = xc_rows_reduce<true>(part, nb, row);
    else if (q == 1) {
        // max = -min(-a): negate on the way in
        const int lane = threadIdx.x & 63;
        double acc = INFINITY;
        for (int b = lane; b < nb; b += 64) acc = fmin(acc, -part[(int64_t)row * nb + b]);
        v = -wave_min(acc);
    } else v = xc_rows_reduce<false>(part, nb, row);
    if (threadIdx.x == 0) out[row] = v;
}

// ---- initialize! (run.jl:24-55, 90-107, 114-124) ----------------------------
// compute_bndtaperw! (run.jl:24-34) on (Nz+1, Nx+2) and compute_bndshape! (run.jl:47-55) on (Nz, Nx)
__global__ __launch_bounds__(NT) void k_init_shapes(xc_mat taper, xc_mat shape, int Nz, int Nx, double h, double D,
                                                    double l0h, double osgr, double bnd_w) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t nT = (int64_t)(Nz + 1) * (Nx + 2), nS = (int64_t)Nz * Nx;
    if (k < nT) {
        const int r = (int)(k % (Nz + 1)), c = (int)(k / (Nz + 1));
        const double z = (double)r * h;
        XM(taper, r, c) = 1.0 - (exp(-z / l0h) + exp(-(D - z) / l0h)) * (1.0 - osgr);
    } else if (k < nT + nS) {
        const int64_t q = k - nT;
        const int r = (int)(q % Nz), c = (int)(q / Nz);
        const double zc = ((double)(r + 1) - 0.5) * h;
        XM(shape, r, c) = exp((-zc + h / 2.0) / bnd_w);
    }
}

// per-block sums of rp (pass 0) or of (rp - mean)^2 with mean = tot[0] / n (pass 1)
__global__ __launch_bounds__(NT) void k_rp_part(const double* rp, int64_t n, const double* tot, int pass, double* part) {
    const double mean = pass ? tot[0] / (double)n : 0.0;
    double s = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x; k < n; k += (int64_t)gridDim.x * NT) {
        const double d = rp[k] - mean;
        s += pass ? d * d : rp[k];
    }
    __shared__ double red[NT / 64];
    const double t = wave_total(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        double u = 0.0;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) u += red[w];
        part[blockIdx.x] = u;
    }
}

// run.jl:98-107: rp normalised by its mean and n-1 standard deviation, the Gaussian bump gp,
// xin = x0 + (Da - x0) bndshape, x = xin (1 + dxr rp + dxg gp) (or x = x_init), m = 1 - x;
// eta, etaco and etas seeded with etam0 (run.jl:111-113)
__global__ __launch_bounds__(NT) void k_init_x(xc_mat x, xc_mat m, xc_mat etas, xc_mat eta, xc_mat etaco,
                                               const double* rp, const double* tot, const double* x_init,
                                               int64_t x_init_ld, xc_mat shape, int Nz, int Nx, double h,
                                               double L, double D, double x0, double Da, double dxr, double dxg,
                                               double etam0) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    const int64_t nS = (int64_t)Nz * Nx, nK = (int64_t)(Nz + 1) * (Nx + 1);
    if (k < nS) {
        const int r = (int)(k % Nz), c = (int)(k / Nz);
        double xv;
        if (x_init) {
            xv = x_init[r + (int64_t)c * x_init_ld];
        } else {
            const double n = (double)nS;
            const double mean = tot[0] / n, sd = sqrt(tot[1] / (n - 1.0));
            const double rv = (rp[k] - mean) / sd;
            const double xc = ((double)(c + 1) - 0.5) * h / L, zc = ((double)(r + 1) - 0.5) * h / D;
            const double ax = (xc - 0.5) / 0.125, az = (zc - 0.5) / 0.125;
            const double gp = exp(-(ax * ax)) * exp(-(az * az));
            const double xin = x0 + (Da - x0) * XM(shape, r, c);
            xv = xin * (1.0 + dxr * rv + dxg * gp);
        }
        XM(x, r, c) = xv;
        XM(m, r, c) = 1.0 - xv;
        XM(etas, r, c) = etam0;
        XM(eta, r, c) = etam0;
    }
    if (k < nK) {
        const int r = (int)(k % (Nz + 1)), c = (int)(k / (Nz + 1));
        XM(etaco, r, c) = etam0;
    }
}

// run.jl:117-118 X = rho x, M = rho m
__global__ __launch_bounds__(NT) void k_init_XM(xc_mat X, xc_mat M, xc_mat rho, xc_mat x, xc_mat m, int Nz, int Nx) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k >= (int64_t)Nz * Nx) return;
    const int r = (int)(k % Nz), c = (int)(k / Nz);
    const double rh = XM(rho, r, c);
    XM(X, r, c) = rh * XM(x, r, c);
    XM(M, r, c) = rh * XM(m, r, c);
}

// run.jl:124-126 Xo = Xoo = X, Mo = Moo = M, rhoo = rhooo = rho
__global__ __launch_bounds__(NT) void k_init_old(xc_mat Xo, xc_mat Xoo, xc_mat Mo, xc_mat Moo, xc_mat rhoo,
                                                 xc_mat rhooo, xc_mat X, xc_mat M, xc_mat rho, int Nz, int Nx) {
    const int64_t k = (int64_t)blockIdx.x * NT + threadIdx.x;
    if (k >= (int64_t)Nz * Nx) return;
    const int r = (int)(k % Nz), c = (int)(k / Nz);
    const double a = XM(X, r, c), b = XM(M, r, c), rh = XM(rho, r, c);
    XM(Xo, r, c) = a; XM(Xoo, r, c) = a;
    XM(Mo, r, c) = b; XM(Moo, r, c) = b;
    XM(rhoo, r, c) = rh; XM(rhooo, r, c) = rh;
}

constexpr int RED_BLOCKS = 512;

struct StepCtx {
    double* snap = nullptr;   // snapshot! of one grid: W, U, P, X, MFS
    size_t snap_n = 0;        // capacity (doubles)
    int snap_Nz = -1, snap_Nx = -1;  // grid of the snapshot held (-1: none)
    double* part = nullptr;   // reduction partials (resnorm, history sums, statistics, rp)
    double* out = nullptr;    // reduction results
    double* host = nullptr;   // pinned copy of out (mapped, coherent)
    double* host_dev = nullptr;  // the same memory as seen by kernels
    double* rp = nullptr;     // initialize!'s random perturbation
    size_t rp_n = 0;
    hipStream_t side = nullptr;  // run!'s second stream (noise!, update!'s lithostatic chain), default priority
    // the sweep's start and noise!'s end on it; the lithostatic chain's start and end (xc_update_side)
    hipEvent_t ev_go = nullptr, ev_noise = nullptr, ev_la = nullptr, ev_lb = nullptr;
};
// partials: the history sums (NHS x RED_BLOCKS) followed by the field statistics (NHF x NHQ x HST_NB); the
// residual norms (9 x RED_BLOCKS) and initialize!'s two sums use the front of the same buffer
constexpr size_t PART_N = (size_t)NHS * RED_BLOCKS + (size_t)NHF * NHQ * HST_NB;
constexpr int OUT_N = NHS + NHF * NHQ;  // 145

void step_free(xc_ctx*, void* p) {
    StepCtx* s = (StepCtx*)p;
    if (s->snap) (void)hipFree(s->snap);
    if (s->part) (void)hipFree(s->part);
    if (s->out) (void)hipFree(s->out);
    if (s->host) (void)hipHostFree(s->host);
    if (s->rp) (void)hipFree(s->rp);
    if (s->side) {
        (void)hipStreamSynchronize(s->side);
        (void)hipStreamDestroy(s->side);
    }
    if (s->ev_go) (void)hipEventDestroy(s->ev_go);
    if (s->ev_noise) (void)hipEventDestroy(s->ev_noise);
    if (s->ev_la) (void)hipEventDestroy(s->ev_la);
    if (s->ev_lb) (void)hipEventDestroy(s->ev_lb);
    delete s;
}
// the context's reduction buffers (the snapshot buffer is made by xc_snapshot only)
int step_ctx(xc_ctx* ctx, StepCtx** out) {
    StepCtx* s = (StepCtx*)ctx->ext[XC_EXT_STEP];
    if (!s) {
        s = new StepCtx();
        ctx->ext[XC_EXT_STEP] = s;
        ctx->ext_free[XC_EXT_STEP] = step_free;
        XC_TRY_HIP(ctx, hipMalloc(&s->part, PART_N * sizeof(double)));
        XC_TRY_HIP(ctx, hipMalloc(&s->out, OUT_N * sizeof(double)));
        XC_TRY_HIP(ctx, hipHostMalloc((void**)&s->host, OUT_N * sizeof(double),
                                      hipHostMallocMapped | hipHostMallocCoherent));
        XC_TRY_HIP(ctx, hipHostGetDevicePointer((void**)&s->host_dev, s->host, 0));
    }
    *out = s;
    return XC_OK;
}
// noise! of one Picard sweep on run!'s second stream, after the sweep's start (ev_go), its end marked by ev_noise
// (recorded on every path: the caller waits for it before it returns)
int noise_side(xc_ctx* ctx, StepCtx* sx, xc_noise* ns, const xc_phase* ph, const xc_grid* g, const xc_par* par,
               const xc_scales* sc, double dt, bool first) {
    XC_TRY_HIP(ctx, hipStreamWaitEvent(sx->side, sx->ev_go, 0));
    const hipStream_t main = ctx->stream;
    ctx->stream = sx->side;
    const int st = xc_noise_step(ctx, ns, ph, g, par, sc, dt, first ? 1 : 0);
    ctx->stream = main;
    XC_TRY_HIP(ctx, hipEventRecord(sx->ev_noise, sx->side));
    if (st) XC_TRY_HIP(ctx, hipStreamWaitEvent(ctx->stream, sx->ev_noise, 0));
    return st;
}
// run!'s second stream and its four events (made once per context)
int step_side(xc_ctx* ctx, StepCtx* s) {
    for (hipEvent_t* e : {&s->ev_go, &s->ev_noise, &s->ev_la, &s->ev_lb})
        if (!*e) XC_TRY_HIP(ctx, hipEventCreateWithFlags(e, hipEventDisableTiming));
    if (!s->side) XC_TRY_HIP(ctx, hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    return XC_OK;
}

// W, U, P, X, MFS, rho of one Nz x Nx grid (state.jl:101-122, 222-227 shapes)
int check_state(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, const char* who, int* Nz, int* Nx) {
    if (!ph || !fl) return xc_fail(ctx, XC_EINVAL_ARG, "%s: NULL state", who);
    const int64_t nz = ph->X.rows, nx = ph->X.cols;
    if (nz < 1 || nx < 1 || nz * nx > INT32_MAX / 4)
        return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: grid %lld x %lld", who, (long long)nz, (long long)nx);
    if (!xc_shape(ph->X, nz, nx) || !xc_shape(fl->MFS, nz, nx) || !xc_shape(fl->rho, nz, nx) ||
        !xc_shape(fl->W, nz + 1, nx + 2) || !xc_shape(fl->U, nz + 2, nx + 1) || !xc_shape(fl->P, nz + 2, nx + 2))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: W/U/P/X/MFS/rho shapes do not match the %lld x %lld grid", who,
                       (long long)nz, (long long)nx);
    *Nz = (int)nz;
    *Nx = (int)nx;
    return XC_OK;
}

int check_hist(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns, int Nz, int Nx,
               const char* who) {
    const xc_mat* cell[] = {&ph->M, &ph->Gx, &ph->x, &ph->V, &ph->vx, &ph->vm, &ph->Ra, &ph->Rc, &ph->ReD, &ph->Red,
                            &ph->ks, &ph->ke, &ph->kx, &fl->eta, &ph->etamix, &ph->etae, &ph->etas, &ph->etat};
    for (const xc_mat* m : cell)
        if (!xc_shape(*m, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: a cell field is not %d x %d", who, Nz, Nx);
    const xc_mat* wt[] = {&ph->qz_advn_X, &ph->qz_advn_M, &ph->qz_dffn_X, &ph->qz_dffn_M};
    for (const xc_mat* m : wt)
        if (!xc_shape(*m, Nz + 1, Nx + 2))
            return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: a z-face flux is not %d x %d", who, Nz + 1, Nx + 2);
    if (ns && (!xc_shape(ns->xie, Nz, Nx) || !xc_shape(ns->xix, Nz, Nx) || !xc_shape(ns->xis, Nz, Nx)))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "%s: noise speeds are not %d x %d", who, Nz, Nx);
    return XC_OK;
}

HistFields hist_fields(const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns) {
    HistFields hf;
    const xc_mat f[NHF] = {ph->x,  ph->V,  ph->vx, ph->vm,  ns->xie, ns->xix, ns->xis,   ph->Ra,   ph->Rc,  ph->ReD,
                           ph->Red, ph->ks, ph->ke, ph->kx, fl->eta, ph->etamix, ph->etae, ph->etas, ph->etat};
    for (int i = 0; i < NHF; ++i) hf.f[i] = f[i];
    return hf;
}

// the reductions of record_history! into out[0 .. OUT_N) on the device: the 12 sums of k_hist_part, then the
// field statistics; part holds PART_N doubles, every slot of it rewritten by this call
void enqueue_history(hipStream_t st, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns, int Nz, int Nx,
                     double* part, double* out) {
    hipLaunchKernelGGL(k_hist_part, dim3(RED_BLOCKS), dim3(NT), 0, st, fl->rho, ph->M, ph->X, ph->Gx, ph->qz_advn_X,
                       ph->qz_advn_M, ph->qz_dffn_X, ph->qz_dffn_M, Nz, Nx, part);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(NHS * 64), 0, st, part, RED_BLOCKS, NHS, out);
    if (!ns) return;
    double* p2 = part + (size_t)NHS * RED_BLOCKS;
    hipLaunchKernelGGL(k_hstat_part, dim3(HST_NB, NHF), dim3(NT), 0, st, hist_fields(ph, fl, ns), Nz, Nx, p2);
    hipLaunchKernelGGL(k_hstat_final, dim3(NHF * NHQ), dim3(64), 0, st, p2, HST_NB, out + NHS);
}

}  // namespace

// record_history!'s reductions of one state into out[0 .. XC_HIST_NSTATS) on stream st, with part (at least
// XC_HIST_PART_N doubles, every slot rewritten) for the partial sums; ns == NULL: the 12 sums only.  Checks the
// shapes first (errors into ctx, or the thread's create-error buffer when ctx is NULL); does not synchronize.
int xc_history_enqueue(xc_ctx* ctx, hipStream_t st, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns,
                       double* part, double* out, const char* who) {
    int Nz, Nx, r;
    if ((r = check_state(ctx, ph, fl, who, &Nz, &Nx))) return r;
    if ((r = check_hist(ctx, ph, fl, ns, Nz, Nx, who))) return r;
    enqueue_history(st, ph, fl, ns, Nz, Nx, part, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xc_fail(ctx, XC_EHIP, "%s: launch: %s", who, hipGetErrorString(e));
    return XC_OK;
}
static_assert(PART_N >= (size_t)XC_HIST_PART_N, "partials");
static_assert(OUT_N == XC_HIST_NSTATS, "history statistics count");

extern "C" {

int xc_snapshot(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl) {
    XC_CHECK_CTX(ctx);
    int Nz, Nx, st;
    if ((st = check_state(ctx, ph, fl, "snapshot!", &Nz, &Nx))) return st;
    const size_t n = (size_t)(Nz + 1) * (Nx + 2) + (size_t)(Nz + 2) * (Nx + 1) + (size_t)(Nz + 2) * (Nx + 2) +
                     2 * (size_t)Nz * Nx;
    StepCtx* s;
    if ((st = step_ctx(ctx, &s))) return st;
    if (s->snap_n < n) {
        if (s->snap) {
            XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
            XC_TRY_HIP(ctx, hipFree(s->snap));
            s->snap = nullptr;
            s->snap_n = 0;
            s->snap_Nz = s->snap_Nx = -1;
        }
        XC_TRY_HIP(ctx, hipMalloc(&s->snap, n * sizeof(double)));
        s->snap_n = n;
    }
    hipLaunchKernelGGL(k_snap, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, ctx->stream, fl->W, fl->U, fl->P,
                       ph->X, fl->MFS, s->snap, Nz, Nx);
    s->snap_Nz = Nz;
    s->snap_Nx = Nx;
    return xc_finish(ctx, "snapshot!");
}

int xc_compute_resnorm(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, double dt, double out[3]) {
    XC_CHECK_CTX(ctx);
    if (!out) return xc_fail(ctx, XC_EINVAL_ARG, "compute_resnorm: NULL out");
    int Nz, Nx, st;
    if ((st = check_state(ctx, ph, fl, "compute_resnorm", &Nz, &Nx))) return st;
    StepCtx* s = (StepCtx*)ctx->ext[XC_EXT_STEP];
    if (!s || !s->snap || s->snap_Nz < 0)
        return xc_fail(ctx, XC_EINVAL_ARG, "compute_resnorm: no snapshot! taken on this context");
    if (s->snap_Nz != Nz || s->snap_Nx != Nx)
        return xc_fail(ctx, XC_EINVAL_ARG, "compute_resnorm: the snapshot is of a %d x %d grid, the state %d x %d",
                       s->snap_Nz, s->snap_Nx, Nz, Nx);
    hipLaunchKernelGGL(k_resnorm_part, dim3(RED_BLOCKS), dim3(NT), 0, ctx->stream, fl->W, fl->U, fl->P, ph->X,
                       fl->MFS, fl->rho, s->snap, Nz, Nx, s->part);
    // the nine sums straight into the mapped host buffer
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(9 * 64), 0, ctx->stream, s->part, RED_BLOCKS, 9, s->host_dev);
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double* a = s->host;
    const double ep = 2.220446049250313e-16;
    const double nrho = sqrt(a[1]);
    const double rm = sqrt(a[0]) / (nrho + ep) + sqrt(a[2]) / (nrho / dt + ep);
    const double rp = sqrt(a[3]) / (sqrt(a[4]) + ep) + sqrt(a[5]) / (sqrt(a[6]) + ep) + sqrt(a[7]) / (sqrt(a[8]) + ep);
    out[0] = rm + rp;
    out[1] = rm;
    out[2] = rp;
    return XC_OK;
}

int xc_history_sums(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, double out[12]) {
    XC_CHECK_CTX(ctx);
    if (!out) return xc_fail(ctx, XC_EINVAL_ARG, "record_history!: NULL out");
    int st;
    StepCtx* s;
    if ((st = step_ctx(ctx, &s))) return st;
    if ((st = xc_history_enqueue(ctx, ctx->stream, ph, fl, nullptr, s->part, s->out, "record_history!"))) return st;
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->host, s->out, NHS * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < NHS; ++q) out[q] = s->host[q];
    return XC_OK;
}

int xc_history_stats(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns,
                     double out[XC_HIST_NSTATS]) {
    XC_CHECK_CTX(ctx);
    if (!out || !ns) return xc_fail(ctx, XC_EINVAL_ARG, "record_history!: NULL noise state or out");
    int st;
    StepCtx* s;
    if ((st = step_ctx(ctx, &s))) return st;
    if ((st = xc_history_enqueue(ctx, ctx->stream, ph, fl, ns, s->part, s->out, "record_history!"))) return st;
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->host, s->out, OUT_N * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < OUT_N; ++q) out[q] = s->host[q];
    return XC_OK;
}

int xc_initialize(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, xc_noise* ns, const xc_grid* g, const xc_par* par,
                  const xc_scales* sc, int xbc, int zbc, const xc_mat* x_init) {
    XC_CHECK_CTX(ctx);
    if (!ph || !fl || !ns || !g || !par || !sc) return xc_fail(ctx, XC_EINVAL_ARG, "initialize!: NULL argument");
    int Nz, Nx, st;
    if ((st = check_state(ctx, ph, fl, "initialize!", &Nz, &Nx))) return st;
    if (Nz != g->Nz || Nx != g->Nx) return xc_fail(ctx, XC_EINVAL_SHAPE, "initialize!: state is not the grid's size");
    if (!xc_shape(ph->x, Nz, Nx) || !xc_shape(ph->m, Nz, Nx) || !xc_shape(ph->etas, Nz, Nx) ||
        !xc_shape(fl->eta, Nz, Nx) || !xc_shape(fl->etaco, Nz + 1, Nx + 1) || !xc_shape(ph->bndshape, Nz, Nx) ||
        !xc_shape(ph->bndtaperw, Nz + 1, Nx + 2) || !xc_shape(ph->M, Nz, Nx) || !xc_shape(ph->Xo, Nz, Nx) ||
        !xc_shape(ph->Xoo, Nz, Nx) || !xc_shape(ph->Mo, Nz, Nx) || !xc_shape(ph->Moo, Nz, Nx) ||
        !xc_shape(fl->rhoo, Nz, Nx) || !xc_shape(fl->rhooo, Nz, Nx))
        return xc_fail(ctx, XC_EINVAL_SHAPE, "initialize!: a field does not match the %d x %d grid", Nz, Nx);
    if (x_init && !xc_shape(*x_init, Nz, Nx)) return xc_fail(ctx, XC_EINVAL_SHAPE, "initialize!: x_init shape");
    if (!(sc->l0h > 0) || !(sc->bnd_w > 0) || !(sc->dt0 > 0))
        return xc_fail(ctx, XC_EINVAL_ARG, "initialize!: scales l0h, bnd_w and dt0 must be positive");
    StepCtx* s;
    if ((st = step_ctx(ctx, &s))) return st;
    const int64_t nS = (int64_t)Nz * Nx, nT = (int64_t)(Nz + 1) * (Nx + 2), nK = (int64_t)(Nz + 1) * (Nx + 1);
    hipStream_t stm = ctx->stream;
    hipLaunchKernelGGL(k_init_shapes, dim3((unsigned)((nT + nS + NT - 1) / NT)), dim3(NT), 0, stm, ph->bndtaperw,
                       ph->bndshape, Nz, Nx, g->h, g->D, sc->l0h, par->open_sgr ? 1.0 : 0.0, sc->bnd_w);
    if (!x_init) {
        // run.jl:98-99: rp = randn(...) normalised; the generator is the library's Philox4x32-10 (seed par.seed,
        // stream 2, draw 0) in place of Julia's MersenneTwister
        if (s->rp_n < (size_t)nS) {
            if (s->rp) { XC_TRY_HIP(ctx, hipStreamSynchronize(stm)); XC_TRY_HIP(ctx, hipFree(s->rp)); s->rp = nullptr; }
            XC_TRY_HIP(ctx, hipMalloc(&s->rp, nS * sizeof(double)));
            s->rp_n = (size_t)nS;
        }
        xc_mat rpm{s->rp, Nz, Nx, Nz};
        if ((st = xc_normal_fill(ctx, rpm, par->seed, 2u, 0u))) return st;
        hipLaunchKernelGGL(k_rp_part, dim3(RED_BLOCKS), dim3(NT), 0, stm, s->rp, nS, s->out, 0, s->part);
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, stm, s->part, RED_BLOCKS, 1, s->out);
        hipLaunchKernelGGL(k_rp_part, dim3(RED_BLOCKS), dim3(NT), 0, stm, s->rp, nS, s->out, 1, s->part + RED_BLOCKS);
        hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(64), 0, stm, s->part + RED_BLOCKS, RED_BLOCKS, 1, s->out + 1);
    }
    const int64_t nI = nS > nK ? nS : nK;
    hipLaunchKernelGGL(k_init_x, dim3((unsigned)((nI + NT - 1) / NT)), dim3(NT), 0, stm, ph->x, ph->m, ph->etas,
                       fl->eta, fl->etaco, (const double*)s->rp, (const double*)s->out,
                       x_init ? (const double*)x_init->p : nullptr, x_init ? x_init->ld : 0, ph->bndshape, Nz, Nx,
                       g->h, g->L, g->D, par->x0, par->Da, par->dxr, par->dxg, par->etam0);
    XC_TRY_HIP(ctx, hipGetLastError());
    if ((st = xc_update(ctx, ph, fl, g, par, sc, xbc, zbc))) return st;
    const unsigned nbS = (unsigned)((nS + NT - 1) / NT);
    hipLaunchKernelGGL(k_init_XM, dim3(nbS), dim3(NT), 0, stm, ph->X, ph->M, fl->rho, ph->x, ph->m, Nz, Nx);
    XC_TRY_HIP(ctx, hipGetLastError());
    if ((st = xc_noise_step(ctx, ns, ph, g, par, sc, sc->dt0, 1))) return st;
    if ((st = xc_update_phase_velocities(ctx, ph, fl, ns, g, par, xbc, zbc))) return st;
    hipLaunchKernelGGL(k_init_old, dim3(nbS), dim3(NT), 0, stm, ph->Xo, ph->Xoo, ph->Mo, ph->Moo, fl->rhoo,
                       fl->rhooo, ph->X, ph->M, fl->rho, Nz, Nx);
    XC_TRY_HIP(ctx, hipGetLastError());
    if ((st = xc_store_noise(ctx, ns))) return st;
    return xc_finish(ctx, "initialize!");
}

int xc_run_step(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, xc_noise* ns, const xc_grid* g, const xc_par* par,
                const xc_scales* sc, const xc_step_opts* so, xc_step_state* ss, xc_step_stats* stats) {
    XC_CHECK_CTX(ctx);
    if (!ph || !fl || !ns || !g || !par || !sc || !so || !ss) return xc_fail(ctx, XC_EINVAL_ARG, "run!: NULL argument");
    int st;
    // the step counter, time and dt of *ss advance only when the whole step succeeds
    const int64_t step = ss->step + 1;
    double c[6];
    if ((st = xc_time_coefs(so->tint, step, c))) return xc_fail(ctx, XC_EINVAL_ARG, "time_coefs: TINT %d", so->tint);
    const double dt = ss->dt;
    if ((st = xc_store_previous(ctx, fl, ph))) return st;
    if ((st = xc_store_noise(ctx, ns))) return st;
    double resnorm = 1.0, resnorm0 = 1.0;
    int iter = 0, fm_iters = 0, fm_fallbacks = 0, fm_levels = 0, fm_first = 0, fm_learned = 0;
    double fm_worst = 0.0, fm_inertia = 0.0, fm_init_first = 0.0, fm_init_later = 0.0;
    const int fixed = so->fixed_sweeps;
    xc_fm_opts fo = so->fm;
    fo.dt = dt;
    fo.a1 = c[0]; fo.a2 = c[1]; fo.a3 = c[2]; fo.b1 = c[3]; fo.b2 = c[4]; fo.b3 = c[5];
    fo.xbc = so->xbc; fo.zbc = so->zbc; fo.scheme = so->scheme;
    StepCtx* sx;
    if ((st = step_ctx(ctx, &sx)) || (st = step_side(ctx, sx))) return st;
    while (fixed > 0 ? iter < fixed
                     : ((resnorm / resnorm0 >= so->rtol && resnorm >= so->atol && iter < so->maxit) || iter < 3)) {
        ++iter;
        // noise! reads what the previous sweep's update! wrote and its own state only, and snapshot!, phsevo! and
        // fluidmech! neither read nor write its fields: it runs on the second stream beside them (results
        // identical), from the start of the sweep (ev_go), enqueued after phsevo! so that the device starts on the
        // sweep's own chain first, and joined before update_phase_velocities!, which reads both
        XC_TRY_HIP(ctx, hipEventRecord(sx->ev_go, ctx->stream));
        if ((st = xc_snapshot(ctx, ph, fl))) return st;
        if ((st = xc_phsevo(ctx, ph, fl, g, par, sc, so->scheme, so->xbc, so->zbc, c, dt, par->alpha))) return st;
        if ((st = noise_side(ctx, sx, ns, ph, g, par, sc, dt, iter <= 1))) return st;
        // every later return leaves the context's stream ordered after the noise work
        auto join = [&]() -> int {
            const hipError_t e = hipStreamWaitEvent(ctx->stream, sx->ev_noise, 0);
            return e == hipSuccess ? XC_OK : xc_fail(ctx, XC_EHIP, "run!: %s", hipGetErrorString(e));
        };
        xc_solve_stats fs{};
        // the horizontal-mean correction's operator: probed in the step's first sweep, reused in the others
        if (so->fm.xmean == 0) fo.xmean = iter == 1 ? 1 : 2;
        st = xc_fluidmech(ctx, fl, ph, g, par, &fo, nullptr, &fs);
        fm_iters += fs.iters;
        fm_fallbacks += fs.fallback_used;
        fm_learned += fs.learned_steps;
        fm_levels = fs.mg_levels;
        fm_inertia = fs.inertia_ratio;
        if (iter == 1) {
            fm_first = fs.iters;
            fm_init_first = fs.init_rel_resid;
        } else if (fs.init_rel_resid > fm_init_later) {
            fm_init_later = fs.init_rel_resid;
        }
        if (fs.rel_resid > fm_worst) fm_worst = fs.rel_resid;
        const int sj = join();
        if (st) return st;
        if (sj) return sj;
        if ((st = xc_update_phase_velocities(ctx, ph, fl, ns, g, par, so->xbc, so->zbc))) return st;
        if ((st = xc_update_side(ctx, ph, fl, g, par, sc, so->xbc, so->zbc, sx->side, sx->ev_la, sx->ev_lb)))
            return st;
        double rn[3];
        if ((st = xc_compute_resnorm(ctx, ph, fl, dt, rn))) return st;
        if (!isfinite(rn[0])) return xc_fail(ctx, XC_ENOCONV, "run!: NaN Picard residual at step %lld sweep %d",
                                            (long long)step, iter);
        resnorm = rn[0];
        if (iter == 1 || resnorm > resnorm0) resnorm0 = resnorm + 1e-32;
    }
    double dtn;
    if ((st = xc_update_dt(ctx, ph, fl, g, par, dt, &dtn))) return st;
    ss->step = step;
    ss->time += dt;
    ss->dt = dtn;
    if (stats) {
        stats->sweeps = iter;
        stats->resnorm = resnorm;
        stats->stokes_iters = fm_iters;
        stats->stokes_worst_rel_resid = fm_worst;
        stats->stokes_fallbacks = fm_fallbacks;
        stats->stokes_learned_steps = fm_learned;
        stats->stokes_mg_levels = fm_levels;
        stats->stokes_inertia_ratio = fm_inertia;
        stats->stokes_iters_first = fm_first;
        stats->stokes_init_rel_first = fm_init_first;
        stats->stokes_init_rel_later = fm_init_later;
    }
    return XC_OK;
}

}  // extern "C"
