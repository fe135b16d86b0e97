// rescale.hip -- (c) per-mode rescale of the Fourier grid, and the synthetic
// Gaussian-field generator used by tests and benchmarks.
//
// Replaces the mode loop of add_nu_power_to_rhogrid (interface_gadget.c:163-188)
// with get_dnudcdm_powerspec (delta_pow.c:19-37) inlined:
//    k2 == 0 is left untouched; logk = log(sqrt(k2) * 2 pi / BoxSize);
//    clamp logk into [logkk[0], logkk[n-1]] (values below logkk[0] - log 2 are
//    clamped too; the reference also prints a warning for them);
//    smth = 1 + norm * (linear interpolation of ratio in logk);
//    NaN smth is fatal (code 5); F *= smth.
// MI355X design: one coalesced read+write sweep (32 B per mode).  The ratio
// table and a uniform bucket index over logk live in LDS, so the segment
// search is a bucket lookup plus a short forward probe instead of a binary
// search.  All arithmetic is FP64 and follows the reference's expression order.
#include <math.h>

#include "kn_hip.h"

namespace {

constexpr int kBlock = 256;
constexpr int kUnroll = 4;
constexpr int kBucketsPerBin = 4;

__device__ int g_nan_flag;

__global__ __launch_bounds__(kBlock) void k_rescale(double2 *__restrict__ F, int dims, int startslab,
                                                     long long rows, double boxsize,
                                                     const double *__restrict__ logkk_g,
                                                     const double *__restrict__ ratio_g, int nbins, double norm,
                                                     int nbuckets)
{
    extern __shared__ double sm[];
    double *lk = sm;               // nbins
    double *rt = lk + nbins;       // nbins
    int *bucket = (int *)(rt + nbins);  // nbuckets + 1
    const int tid = threadIdx.x;
    for (int t = tid; t < nbins; t += kBlock) {
        lk[t] = logkk_g[t];
        rt[t] = ratio_g[t];
    }
    __syncthreads();
    const double x0 = lk[0], x1 = lk[nbins - 1];
    const double inv_bw = nbins > 1 ? nbuckets / (x1 - x0) : 0.0;
    // bucket[q] = last segment start index i <= nbins-2 with lk[i] <= left edge of bucket q
    for (int q = tid; q <= nbuckets; q += kBlock) {
        const double edge = x0 + q / inv_bw;
        int lo = 0, hi = nbins - 1;
        while (hi > lo + 1) {
            const int mid = (lo + hi) >> 1;
            if (lk[mid] > edge) hi = mid;
            else lo = mid;
        }
        bucket[q] = lo;
    }
    __syncthreads();
    const double lower_clamp = x0 - log(2.0);
    const int rowlen = dims / 2 + 1;
    const long long total = rows * rowlen;
    const long long stride = (long long)gridDim.x * kBlock;
    bool saw_nan = false;
    for (long long m0 = (long long)blockIdx.x * kBlock + tid; m0 < total; m0 += stride * kUnroll) {
        double2 f[kUnroll];
        double smth[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const long long m = m0 + stride * u;
            if (m < total) f[u] = F[m];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const long long m = m0 + stride * u;
            smth[u] = 1.0;
            if (m >= total) continue;
            const long long row = m / rowlen;
            const int z = (int)(m - row * rowlen);
            const int x = (int)(row % dims), y = startslab + (int)(row / dims);
            const int kx = kn_kval(x, dims), ky = kn_kval(y, dims);
            const double k2 = (double)(kx * kx + ky * ky + z * z);
            if (k2 <= 0) continue;
            double kk = log(sqrt(k2) * 2 * M_PI / boxsize);
            // delta_pow.c:25-34 clamps
            if (kk < x0 && kk > lower_clamp) kk = x0;
            if (kk < lower_clamp) kk = x0;
            if (kk > x1) kk = x1;
            double interp;
            if (nbins == 1) {
                interp = rt[0];
            } else {
                int q = (int)((kk - x0) * inv_bw);
                q = q < 0 ? 0 : (q > nbuckets ? nbuckets : q);
                int i = bucket[q];
                while (i < nbins - 2 && lk[i + 1] <= kk) i++;
                const double xl = lk[i], h = lk[i + 1] - xl;
                interp = h > 0.0 ? rt[i] + (kk - xl) / h * (rt[i + 1] - rt[i]) : rt[i];
            }
            smth[u] = 1 + norm * interp;
            if (isnan(smth[u])) saw_nan = true;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const long long m = m0 + stride * u;
            if (m < total && smth[u] != 1.0) {
                f[u].x *= smth[u];
                f[u].y *= smth[u];
                F[m] = f[u];
            }
        }
    }
    if (saw_nan) atomicOr(&g_nan_flag, 1);
}

// ---- synthetic Gaussian random field -----------------------------------

__host__ __device__ inline unsigned long long mix64(unsigned long long z)
{
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Uniform in (0, 1] from the top 53 bits.
__host__ __device__ inline double unit53(unsigned long long r) { return ((r >> 11) + 1) * (1.0 / 9007199254740992.0); }

__device__ inline double bbks_power(double k, double amp, double ns, double gamma)
{
    const double q = k / gamma;
    const double t = log(1 + 2.34 * q) / (2.34 * q) *
                     pow(1 + 3.89 * q + pow(16.1 * q, 2) + pow(5.46 * q, 3) + pow(6.71 * q, 4), -0.25);
    return amp * pow(k, ns) * t * t;
}

__global__ __launch_bounds__(kBlock) void k_synth(double2 *__restrict__ F, int dims, int startslab, long long rows,
                                                   double kunit, double amp, double ns, double gamma,
                                                   unsigned long long seed)
{
    const int rowlen = dims / 2 + 1;
    const long long total = rows * rowlen;
    for (long long m = (long long)blockIdx.x * kBlock + threadIdx.x; m < total; m += (long long)gridDim.x * kBlock) {
        const long long row = m / rowlen;
        const int z = (int)(m - row * rowlen);
        const int x = (int)(row % dims), y = startslab + (int)(row / dims);
        const int kx = kn_kval(x, dims), ky = kn_kval(y, dims);
        const long long k2 = (long long)kx * kx + (long long)ky * ky + (long long)z * z;
        const long long gidx = ((long long)y * dims + x) * rowlen + z;  // index in the full grid
        double2 out;
        if (k2 == 0) {
            out = make_double2((double)dims * dims * dims, 0.0);
        } else {
            const double k = sqrt((double)k2) * kunit;
            const double sig = sqrt(bbks_power(k, amp, ns, gamma) / 2);
            const unsigned long long r1 = mix64(seed * 0x2545f4914f6cdd1dULL ^ (2 * (unsigned long long)gidx));
            const unsigned long long r2 = mix64(seed * 0x2545f4914f6cdd1dULL ^ (2 * (unsigned long long)gidx + 1));
            const double u1 = unit53(r1), u2 = unit53(r2);
            const double rad = sqrt(-2.0 * log(u1));
            double s, c;
            sincos(2 * M_PI * u2, &s, &c);
            out = make_double2(sig * rad * c, sig * rad * s);
        }
        F[m] = out;
    }
}

int launch_grid(long long work)
{
    long long g = (work + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    return (int)(g < 1 ? 1 : g);
}

}  // namespace

extern "C" int kn_rescale_device(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                                 const double *d_logkk, const double *d_ratio, int nbins, double norm, void *stream)
{
    if (nbins < 1 || dims < 2 || nslab < 1) return KN_ERR_ARG;
    const int nbuckets = kBucketsPerBin * nbins;
    const size_t smem = sizeof(double) * 2 * nbins + sizeof(int) * (nbuckets + 1);
    if (smem > 64 * 1024) return KN_ERR_ARG;
    const long long rows = (long long)nslab * dims;
    const int grid = launch_grid(rows * (dims / 2 + 1) / kUnroll);
    hipLaunchKernelGGL(k_rescale, dim3(grid), dim3(kBlock), smem, (hipStream_t)stream, (double2 *)d_field, dims,
                       startslab, rows, boxsize, d_logkk, d_ratio, nbins, norm, nbuckets);
    KN_CHECK(hipGetLastError());
    return 0;
}

// Host-side check of the NaN flag raised by k_rescale (clears it).
extern "C" int kn_rescale_nan_flag(void)
{
    int v = 0, z = 0;
    KN_CHECK(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_nan_flag), sizeof(int)));
    if (v) KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_nan_flag), &z, sizeof(int)));
    return v;
}

extern "C" int kn_synth_grid(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                             double amplitude, double ns, double gamma, unsigned long long seed)
{
    const long long rows = (long long)nslab * dims;
    hipLaunchKernelGGL(k_synth, dim3(launch_grid(rows * (dims / 2 + 1))), dim3(kBlock), 0, 0, (double2 *)d_field,
                       dims, startslab, rows, 2 * M_PI / boxsize, amplitude, ns, gamma, seed);
    KN_CHECK(hipGetLastError());
    KN_CHECK(hipDeviceSynchronize());
    return 0;
}

// ---- grid residency helpers for add_nu_power_to_rhogrid ------------------

// Device view of a caller grid: device/managed memory is used in place, host
// memory is staged into HBM (*staged = 1).
extern "C" fftw_complex *kn_grid_to_device(fftw_complex *grid, size_t nmodes, int *staged)
{
    if (kn_is_device_ptr(grid)) {
        *staged = 0;
        return grid;
    }
    fftw_complex *d = nullptr;
    KN_CHECK(hipMalloc(&d, sizeof(fftw_complex) * nmodes));
    KN_CHECK(hipMemcpy(d, grid, sizeof(fftw_complex) * nmodes, hipMemcpyHostToDevice));
    *staged = 1;
    return d;
}

extern "C" void kn_grid_release(fftw_complex *host, fftw_complex *dev, size_t nmodes, int staged, int copy_back)
{
    if (!staged) {
        KN_CHECK(hipDeviceSynchronize());
        return;
    }
    if (copy_back) KN_CHECK(hipMemcpy(host, dev, sizeof(fftw_complex) * nmodes, hipMemcpyDeviceToHost));
    KN_CHECK(hipFree(dev));
}

extern "C" void kn_rescale_from_host_table(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                                           const double *logkk, const double *ratio, int nbins, double norm)
{
    double *d_tab = nullptr;
    KN_CHECK(hipMalloc(&d_tab, sizeof(double) * 2 * (nbins > 0 ? nbins : 1)));
    KN_CHECK(hipMemcpy(d_tab, logkk, sizeof(double) * nbins, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(d_tab + nbins, ratio, sizeof(double) * nbins, hipMemcpyHostToDevice));
    const int rc = kn_rescale_device(d_field, dims, startslab, nslab, boxsize, d_tab, d_tab + nbins, nbins, norm, 0);
    if (rc) kn_terminate(rc, "rescale: bad table (nbins=%d)\n", nbins);
    KN_CHECK(hipDeviceSynchronize());
    KN_CHECK(hipFree(d_tab));
}
