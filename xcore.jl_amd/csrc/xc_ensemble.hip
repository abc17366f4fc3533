// xc_ensemble.hip -- ensembles of independent runs advanced together (no reference counterpart: each member is
// exactly one outer step of run!, src/run.jl:204-275, on its own grid, parameters, noise seed and state).
//
// Every member owns a context (its stream, its solver hierarchy, its FFT tables); a pool of host threads takes
// members from a shared counter and runs xc_run_step for each, so the members' kernels -- many of them small at
// 128^2 -- run concurrently on the device instead of one member after another.  The Picard loop and the Krylov
// solves stay adaptive per member (each synchronizes only its own stream).  A member's results are those of
// the same step run alone, bit for bit: the same kernels in the same order on the same data.
#include <atomic>
#include <thread>
#include <vector>

#include "xc_internal.h"

extern "C" int xc_ensemble_run_step(int n, xc_ensemble_member* members, int nthreads) {
    if (n < 0 || (n > 0 && !members)) return xc_fail(nullptr, XC_EINVAL_ARG, "ensemble: n %d", n);
    for (int i = 0; i < n; ++i)
        if (!members[i].ctx) return xc_fail(nullptr, XC_EINVAL_ARG, "ensemble: member %d has no context", i);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j)
            if (members[i].ctx == members[j].ctx)
                return xc_fail(nullptr, XC_EINVAL_ARG, "ensemble: members %d and %d share a context", j, i);
    // default: one thread per hardware queue (4 by default); more threads only contend for the same queues
    const int want = nthreads < 1 ? 4 : nthreads;
    const int nt = want > n ? (n > 0 ? n : 1) : want;
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) {
            xc_ensemble_member& m = members[i];
            const int dev = m.ctx->device;
            if (hipSetDevice(dev) != hipSuccess) {
                m.status = xc_fail(m.ctx, XC_EHIP, "ensemble: hipSetDevice(%d)", dev);
                continue;
            }
            m.status = xc_run_step(m.ctx, m.phase, m.fluid, m.ns, m.grid, m.par, m.scales, m.opts, m.state, m.stats);
        }
    };
    if (nt == 1) {
        work();
    } else {
        std::vector<std::thread> pool;
        pool.reserve((size_t)nt);
        for (int t = 0; t < nt; ++t) pool.emplace_back(work);
        for (auto& t : pool) t.join();
    }
    for (int i = 0; i < n; ++i)
        if (members[i].status != XC_OK) return members[i].status;
    return XC_OK;
}
