// stream_ceiling.hip -- plain HBM streams over a buffer the size of the grid,
// to measure what this part sustains for the access patterns of the two
// sweeps: a read-only stream (the bin sweep's 16 B per mode) and an in-place
// read + write stream (the rescale's 32 B per mode), each with and without
// the non-temporal hint.  Measurement infrastructure for bench.py; not part
// of the library's ABI (libkn_stream_ceiling.so).
//
// Each block streams a contiguous range (as the sweeps do), 512 threads,
// U 16-byte accesses per thread in flight.
#include <hip/hip_runtime.h>

#include "../csrc/kn_mem.h"

namespace {

template <int U, bool NT>
__global__ __launch_bounds__(512) void k_read(const double2 *__restrict__ F, long long n, double *__restrict__ sink)
{
    const long long m0 = n * blockIdx.x / gridDim.x, m1 = n * (blockIdx.x + 1) / gridDim.x;
    double s = 0.0;
    for (long long c = m0; c < m1; c += 512LL * U) {
        double2 f[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long m = c + 512LL * u + threadIdx.x;
            f[u] = m < m1 ? kn_mem::ld<NT>(F + m) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < U; u++) s += f[u].x + f[u].y;
    }
    if (s == 1234.5678) sink[0] = s;  // keeps the loads; never true for the benchmark's data
}

template <int U, bool NT>
__global__ __launch_bounds__(512) void k_scale(double2 *__restrict__ F, long long n, double sc)
{
    const long long m0 = n * blockIdx.x / gridDim.x, m1 = n * (blockIdx.x + 1) / gridDim.x;
    for (long long c = m0; c < m1; c += 512LL * U) {
        double2 f[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long m = c + 512LL * u + threadIdx.x;
            if (m < m1) f[u] = kn_mem::ld<NT>(F + m);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long m = c + 512LL * u + threadIdx.x;
            if (m < m1) {
                f[u].x *= sc;
                f[u].y *= sc;
                kn_mem::st<NT>(F + m, f[u]);
            }
        }
    }
}

int grid_for(const void *kernel, int slots)
{
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 512, 0) != hipSuccess) return -1;
    return (per_cu > 0 ? per_cu : 1) * cus * slots;
}

}  // namespace

// kind 0: read-only stream of n modes (16 B each) at F; kind 1: in-place
// read + write (F *= scale); nt 1 puts the non-temporal hint on every access.
// sink: one device double.  Launches on the null stream, returns at once
// (0, or -1 for bad arguments / a failed launch).
extern "C" int kn_stream_ceiling(void *F, long long n, int kind, int nt, int slots, double *sink)
{
    if (!F || n < 1 || slots < 1 || (kind != 0 && kind != 1)) return -1;
    const void *k = kind == 0 ? (nt ? (const void *)k_read<4, true> : (const void *)k_read<4, false>)
                              : (nt ? (const void *)k_scale<2, true> : (const void *)k_scale<2, false>);
    const int g = grid_for(k, slots);
    if (g < 1) return -1;
    if (kind == 0) {
        if (!sink) return -1;
        if (nt)
            hipLaunchKernelGGL((k_read<4, true>), dim3(g), dim3(512), 0, 0, (const double2 *)F, n, sink);
        else
            hipLaunchKernelGGL((k_read<4, false>), dim3(g), dim3(512), 0, 0, (const double2 *)F, n, sink);
    } else {
        if (nt)
            hipLaunchKernelGGL((k_scale<2, true>), dim3(g), dim3(512), 0, 0, (double2 *)F, n, 1.0);
        else
            hipLaunchKernelGGL((k_scale<2, false>), dim3(g), dim3(512), 0, 0, (double2 *)F, n, 1.0);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
