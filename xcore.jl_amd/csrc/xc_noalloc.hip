// xc_noalloc.hip -- record_history!'s reductions (history.jl:101-222) for callers that cannot allocate:
// code under HIP graph capture, tight loops, harnesses that forbid hipMalloc.
//
// The partial sums live in a file-scope device array instead of the context's buffers.  Each device holds its
// own copy of the array at its own address; the address is looked up once per device with
// hipGetSymbolAddress (at xc_create, or on the first call on a device) and only that address is handed to
// kernels -- the array's name in host code is not a device address.  Every call rewrites every partial it
// reads, so nothing has to be cleared between calls; calls that overlap on one device (two streams at once)
// race on the array.
#include <mutex>

#include "xc_internal.h"

namespace {

constexpr int MAX_DEV = 64;
// grids up to this many cells (the reductions index cells with 32-bit integers)
constexpr int64_t MAX_CELLS = 536870911;
// 12 x 512 partials of the history sums, then 19 x 7 x 64 of the field statistics: 14,656 doubles (114.5 KiB)
static __device__ __align__(256) double scratch[XC_HIST_PART_N];

std::mutex g_mu;
double* g_ptr[MAX_DEV] = {};

// the device address of `scratch` on the current device
int scratch_ptr(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return xc_fail(nullptr, XC_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEV) return xc_fail(nullptr, XC_EINVAL_ARG, "device %d beyond %d", dev, MAX_DEV);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return xc_fail(nullptr, XC_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                           hipGetErrorString(e));
        g_ptr[dev] = (double*)p;
    }
    *out = g_ptr[dev];
    return XC_OK;
}

}  // namespace

int xc_noalloc_init_device() {
    double* p;
    return scratch_ptr(&p);
}

extern "C" {

int xc_history_stats_noalloc(void* stream, const xc_phase* ph, const xc_fluid* fl, const xc_noise* ns,
                             double* out_dev) {
    if (!ph || !fl || !ns || !out_dev)
        return xc_fail(nullptr, XC_EINVAL_ARG, "xc_history_stats_noalloc: NULL argument");
    // the partials are a fixed 14,656 doubles whatever the grid; the kernels index a grid by 32-bit cell
    // counts, so larger grids are refused
    if (ph->X.rows < 1 || ph->X.cols < 1 || ph->X.rows * ph->X.cols > MAX_CELLS)
        return xc_fail(nullptr, XC_EINVAL_SHAPE,
                       "xc_history_stats_noalloc: grid %lld x %lld: 1 to %lld cells are supported",
                       (long long)ph->X.rows, (long long)ph->X.cols, (long long)MAX_CELLS);
    double* part;
    int st;
    if ((st = scratch_ptr(&part))) return st;
    return xc_history_enqueue(nullptr, (hipStream_t)stream, ph, fl, ns, part, out_dev, "xc_history_stats_noalloc");
}

}  // extern "C"
