// The device's streaming ceiling, measured by hand-written kernels: how fast this GPU reads, writes and copies a large
// buffer. The bandwidth-bound kernels of the library are judged against it (a stencil that reads one field and writes
// one can at best run at the copy rate). Each kernel moves 16-byte vectors, U of them in flight per thread, every
// workgroup covering a contiguous 256 x U x 16-byte span once (no grid-stride loop); stores are plain or nontemporal.
#include <cstdint>

#include "lira.h"
#include "lira_internal.hpp"

namespace lira {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

template <int U, bool NT> __global__ __launch_bounds__(256) void k_stream_copy(const f4 *src, f4 *dst, long n4) {
    const long base = (long)blockIdx.x * (256 * U) + threadIdx.x;
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long i = base + (long)u * 256;
        v[u] = i < n4 ? src[i] : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long i = base + (long)u * 256;
        if (i < n4) {
            if (NT) __builtin_nontemporal_store(v[u], dst + i);
            else dst[i] = v[u];
        }
    }
}

// reads every vector and writes one float per workgroup (the sum, so the loads are not dead)
template <int U> __global__ __launch_bounds__(256) void k_stream_read(const f4 *src, long n4, float *out) {
    __shared__ float sh[4];
    const long base = (long)blockIdx.x * (256 * U) + threadIdx.x;
    f4 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long i = base + (long)u * 256;
        v[u] = i < n4 ? src[i] : f4{0.f, 0.f, 0.f, 0.f};
    }
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < U; u++) s += (v[u].x + v[u].y) + (v[u].z + v[u].w);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

template <int U, bool NT> __global__ __launch_bounds__(256) void k_stream_write(f4 *dst, long n4, float value) {
    const long base = (long)blockIdx.x * (256 * U) + threadIdx.x;
    const f4 v{value, value, value, value};
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long i = base + (long)u * 256;
        if (i < n4) {
            if (NT) __builtin_nontemporal_store(v, dst + i);
            else dst[i] = v;
        }
    }
}

constexpr int STREAM_U = 4;  // 4 x 16 B in flight per thread

}  // namespace
}  // namespace lira

using namespace lira;

extern "C" lira_status lira_stream_bandwidth(int kind, int nontemporal, const void *src, void *dst,
                                             unsigned long long bytes, int reps, void *stream, double *gbs) {
    if (!gbs || reps < 1 || bytes < 16 || (bytes & 15) != 0 || kind < 0 || kind > 2) return LIRA_INVALID;
    if ((kind != 2 && !src) || (kind != 1 && !dst)) return LIRA_INVALID;
    if (((uintptr_t)src & 15) != 0 || ((uintptr_t)dst & 15) != 0) return LIRA_INVALID;
    const long n4 = (long)(bytes / 16);
    const long blocks = (n4 + 256L * STREAM_U - 1) / (256L * STREAM_U);
    if (blocks > 0x7fffffffL) return LIRA_INVALID;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    float *sums = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    lira_status status = LIRA_OK;
    auto launch = [&]() {
        const f4 *in = reinterpret_cast<const f4 *>(src);
        f4 *out = reinterpret_cast<f4 *>(dst);
        if (kind == 0) {
            if (nontemporal) hipLaunchKernelGGL((k_stream_copy<STREAM_U, true>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, n4);
            else hipLaunchKernelGGL((k_stream_copy<STREAM_U, false>), dim3((unsigned)blocks), dim3(256), 0, s, in, out, n4);
        } else if (kind == 1) {
            hipLaunchKernelGGL(k_stream_read<STREAM_U>, dim3((unsigned)blocks), dim3(256), 0, s, in, n4, sums);
        } else {
            if (nontemporal) hipLaunchKernelGGL((k_stream_write<STREAM_U, true>), dim3((unsigned)blocks), dim3(256), 0, s, out, n4, 1.0f);
            else hipLaunchKernelGGL((k_stream_write<STREAM_U, false>), dim3((unsigned)blocks), dim3(256), 0, s, out, n4, 1.0f);
        }
        return hipGetLastError();
    };
    if (hipMalloc((void **)&sums, (size_t)blocks * sizeof(float)) != hipSuccess) return LIRA_DEVICE;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) status = LIRA_DEVICE;
    // two untimed launches, then reps timed back to back on the stream
    for (int r = 0; r < 2 && status == LIRA_OK; r++)
        if (launch() != hipSuccess) status = LIRA_DEVICE;
    if (status == LIRA_OK && hipEventRecord(e0, s) != hipSuccess) status = LIRA_DEVICE;
    for (int r = 0; r < reps && status == LIRA_OK; r++)
        if (launch() != hipSuccess) status = LIRA_DEVICE;
    float ms = 0.f;
    if (status == LIRA_OK && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                              hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        status = LIRA_DEVICE;
    if (status == LIRA_OK) {
        const double moved = (double)bytes * (kind == 0 ? 2.0 : 1.0);
        *gbs = moved / ((double)ms / reps * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sums);
    return status;
}
