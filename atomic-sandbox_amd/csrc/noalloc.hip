// lira_gram_noalloc: the tall-skinny Gram of lira_gram for callers that cannot allocate device memory (code under HIP
// graph capture, tight serving loops, harnesses that forbid hipMalloc). The per-chunk partial sums of the Gram live in
// a fixed device array, `scratch`, one copy per GPU for the life of the process; G goes straight to the caller's
// device buffer. The kernels are lira::gram's, so the result is bitwise that of lira_gram.
#include <hip/hip_runtime.h>

#include <mutex>

#include "lira.h"
#include "lira_internal.hpp"

namespace {

constexpr size_t SCRATCH_DOUBLES = LIRA_NOALLOC_SCRATCH_BYTES / sizeof(double);
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};

// The address of this GPU's copy of `scratch`, looked up once per device with that device current. Host code never
// passes the array's name itself to a kernel: on the host it is only a stand-in, not a device address.
double *scratch_on_current_device(lira_status *st) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        *st = LIRA_DEVICE;
        return nullptr;
    }
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *st = LIRA_DEVICE;
            return nullptr;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *st = LIRA_OK;
    return g_scratch[dev];
}

}  // namespace

extern "C" lira_status lira_gram_noalloc(int dtype, int n, double h, int m, const void *s_block, int q,
                                         const void *u_block, double *g_dev, void *stream) {
    if ((dtype != 4 && dtype != 8) || n < 1 || m < 1 || q < 1 || !s_block || !u_block || !g_dev || !(h > 0.0))
        return LIRA_INVALID;
    const long N = (long)n * n * n;
    // the partials of this shape must fit the fixed array (checked before anything touches the device)
    if (lira::gram_partials_needed(m, q, N) > SCRATCH_DOUBLES) return LIRA_INVALID;
    lira_status st = LIRA_OK;
    double *part = scratch_on_current_device(&st);
    if (st != LIRA_OK) return st;
    try {
        lira::Cols S{{{s_block, m}}, 1}, U{{{u_block, q}}, 1};
        hipStream_t s = static_cast<hipStream_t>(stream);
        if (dtype == 4) lira::gram<float>(S, U, N, h * h * h, part, SCRATCH_DOUBLES, g_dev, s);
        else lira::gram<double>(S, U, N, h * h * h, part, SCRATCH_DOUBLES, g_dev, s);
    } catch (const lira::InvalidError &) {
        return LIRA_INVALID;
    } catch (...) {
        return LIRA_DEVICE;
    }
    return LIRA_OK;
}
