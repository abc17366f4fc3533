// xc_step.hip -- one outer time step of run! (src/run.jl:204-275) on the device,
// with the Picard residual of diagnose.jl:42-71 and the conservation sums of
// history.jl:101-181.
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
__global__ __launch_bounds__(NT) void k_hist_part(xc_mat rho, xc_mat M, xc_mat X, xc_mat Gx, xc_mat qaX, xc_mat qaM,
                                                  xc_mat qdX, xc_mat qdM, int Nz, int Nx, double* part) {
    double a[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
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
    __shared__ double red[12][NT];
    for (int q = 0; q < 12; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st)
            for (int q = 0; q < 12; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 12; ++q) part[(int64_t)q * gridDim.x + blockIdx.x] = red[q][0];
}

constexpr int RED_BLOCKS = 512;

struct StepCtx {
    double* snap = nullptr;
    size_t snap_n = 0;
    double* part = nullptr;
    double* out = nullptr;
    double* host = nullptr;
};
void step_free(xc_ctx*, void* p) {
    StepCtx* s = (StepCtx*)p;
    if (s->snap) (void)hipFree(s->snap);
    if (s->part) (void)hipFree(s->part);
    if (s->out) (void)hipFree(s->out);
    if (s->host) (void)hipHostFree(s->host);
    delete s;
}
int step_ctx(xc_ctx* ctx, size_t nsnap, StepCtx** out) {
    StepCtx* s = (StepCtx*)ctx->ext[XC_EXT_STEP];
    if (!s) {
        s = new StepCtx();
        ctx->ext[XC_EXT_STEP] = s;
        ctx->ext_free[XC_EXT_STEP] = step_free;
        XC_TRY_HIP(ctx, hipMalloc(&s->part, 16 * RED_BLOCKS * sizeof(double)));
        XC_TRY_HIP(ctx, hipMalloc(&s->out, 16 * sizeof(double)));
        XC_TRY_HIP(ctx, hipHostMalloc((void**)&s->host, 16 * sizeof(double), hipHostMallocDefault));
    }
    if (s->snap_n < nsnap) {
        if (s->snap) { XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream)); XC_TRY_HIP(ctx, hipFree(s->snap)); s->snap = nullptr; }
        XC_TRY_HIP(ctx, hipMalloc(&s->snap, nsnap * sizeof(double)));
        s->snap_n = nsnap;
    }
    *out = s;
    return XC_OK;
}

}  // namespace

extern "C" {

int xc_compute_resnorm(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, double dt, double out[3]);

int xc_snapshot(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl) {
    XC_CHECK_CTX(ctx);
    const int Nz = (int)ph->X.rows, Nx = (int)ph->X.cols;
    const size_t n = (size_t)(Nz + 1) * (Nx + 2) + (size_t)(Nz + 2) * (Nx + 1) + (size_t)(Nz + 2) * (Nx + 2) +
                     2 * (size_t)Nz * Nx;
    StepCtx* s;
    int st;
    if ((st = step_ctx(ctx, n, &s))) return st;
    hipLaunchKernelGGL(k_snap, dim3((unsigned)((n + NT - 1) / NT)), dim3(NT), 0, ctx->stream, fl->W, fl->U, fl->P,
                       ph->X, fl->MFS, s->snap, Nz, Nx);
    return xc_finish(ctx, "snapshot!");
}

int xc_compute_resnorm(xc_ctx* ctx, const xc_phase* ph, const xc_fluid* fl, double dt, double out[3]) {
    XC_CHECK_CTX(ctx);
    const int Nz = (int)ph->X.rows, Nx = (int)ph->X.cols;
    StepCtx* s = (StepCtx*)ctx->ext[XC_EXT_STEP];
    if (!s || !s->snap) return xc_fail(ctx, XC_EINVAL_ARG, "compute_resnorm: no snapshot taken");
    hipLaunchKernelGGL(k_resnorm_part, dim3(RED_BLOCKS), dim3(NT), 0, ctx->stream, fl->W, fl->U, fl->P, ph->X,
                       fl->MFS, fl->rho, s->snap, Nz, Nx, s->part);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(9 * 64), 0, ctx->stream, s->part, RED_BLOCKS, 9, s->out);
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->host, s->out, 9 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
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
    const int Nz = (int)ph->X.rows, Nx = (int)ph->X.cols;
    StepCtx* s;
    int st;
    if ((st = step_ctx(ctx, 1, &s))) return st;
    hipLaunchKernelGGL(k_hist_part, dim3(RED_BLOCKS), dim3(NT), 0, ctx->stream, fl->rho, ph->M, ph->X, ph->Gx,
                       ph->qz_advn_X, ph->qz_advn_M, ph->qz_dffn_X, ph->qz_dffn_M, Nz, Nx, s->part);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(12 * 64), 0, ctx->stream, s->part, RED_BLOCKS, 12, s->out);
    XC_TRY_HIP(ctx, hipMemcpyAsync(s->host, s->out, 12 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    XC_TRY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (int q = 0; q < 12; ++q) out[q] = s->host[q];
    return XC_OK;
}

int xc_run_step(xc_ctx* ctx, xc_phase* ph, xc_fluid* fl, xc_noise* ns, const xc_grid* g, const xc_par* par,
                const xc_scales* sc, const xc_step_opts* so, xc_step_state* ss, xc_step_stats* stats) {
    XC_CHECK_CTX(ctx);
    if (!ph || !fl || !ns || !g || !par || !sc || !so || !ss) return xc_fail(ctx, XC_EINVAL_ARG, "run!: NULL argument");
    int st;
    ss->step += 1;
    double c[6];
    if ((st = xc_time_coefs(so->tint, ss->step, c))) return xc_fail(ctx, XC_EINVAL_ARG, "time_coefs: TINT %d", so->tint);
    const double dt = ss->dt;
    if ((st = xc_store_previous(ctx, fl, ph))) return st;
    if ((st = xc_store_noise(ctx, ns))) return st;
    double resnorm = 1.0, resnorm0 = 1.0;
    int iter = 0, fm_iters = 0, fm_fallbacks = 0, fm_levels = 0, fm_first = 0;
    double fm_worst = 0.0, fm_inertia = 0.0, fm_init_first = 0.0, fm_init_later = 0.0;
    const int fixed = so->fixed_sweeps;
    xc_fm_opts fo = so->fm;
    fo.dt = dt;
    fo.a1 = c[0]; fo.a2 = c[1]; fo.a3 = c[2]; fo.b1 = c[3]; fo.b2 = c[4]; fo.b3 = c[5];
    fo.xbc = so->xbc; fo.zbc = so->zbc; fo.scheme = so->scheme;
    while (fixed > 0 ? iter < fixed
                     : ((resnorm / resnorm0 >= so->rtol && resnorm >= so->atol && iter < so->maxit) || iter < 3)) {
        ++iter;
        if ((st = xc_snapshot(ctx, ph, fl))) return st;
        if ((st = xc_phsevo(ctx, ph, fl, g, par, sc, so->scheme, so->xbc, so->zbc, c, dt, par->alpha))) return st;
        xc_solve_stats fs{};
        st = xc_fluidmech(ctx, fl, ph, g, par, &fo, nullptr, &fs);
        fm_iters += fs.iters;
        fm_fallbacks += fs.fallback_used;
        fm_levels = fs.mg_levels;
        fm_inertia = fs.inertia_ratio;
        if (iter == 1) {
            fm_first = fs.iters;
            fm_init_first = fs.init_rel_resid;
        } else if (fs.init_rel_resid > fm_init_later) {
            fm_init_later = fs.init_rel_resid;
        }
        if (fs.rel_resid > fm_worst) fm_worst = fs.rel_resid;
        if (st) return st;
        if ((st = xc_noise_step(ctx, ns, ph, g, par, sc, dt, iter <= 1))) return st;
        if ((st = xc_update_phase_velocities(ctx, ph, fl, ns, g, par, so->xbc, so->zbc))) return st;
        if ((st = xc_update(ctx, ph, fl, g, par, sc, so->xbc, so->zbc))) return st;
        double rn[3];
        if ((st = xc_compute_resnorm(ctx, ph, fl, dt, rn))) return st;
        if (!isfinite(rn[0])) return xc_fail(ctx, XC_ENOCONV, "run!: NaN Picard residual at step %lld sweep %d",
                                            (long long)ss->step, iter);
        resnorm = rn[0];
        if (iter == 1 || resnorm > resnorm0) resnorm0 = resnorm + 1e-32;
    }
    ss->time += dt;
    double dtn;
    if ((st = xc_update_dt(ctx, ph, fl, g, par, dt, &dtn))) return st;
    ss->dt = dtn;
    if (stats) {
        stats->sweeps = iter;
        stats->resnorm = resnorm;
        stats->stokes_iters = fm_iters;
        stats->stokes_worst_rel_resid = fm_worst;
        stats->stokes_fallbacks = fm_fallbacks;
        stats->stokes_mg_levels = fm_levels;
        stats->stokes_inertia_ratio = fm_inertia;
        stats->stokes_iters_first = fm_first;
        stats->stokes_init_rel_first = fm_init_first;
        stats->stokes_init_rel_later = fm_init_later;
    }
    return XC_OK;
}

}  // extern "C"
