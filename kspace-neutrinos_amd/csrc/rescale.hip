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
// MI355X design: one coalesced read+write sweep (32 B per mode, non-temporal
// loads and stores).  The ratio table and a uniform bucket index over logk
// live in LDS, so the segment search is a bucket lookup plus a short forward
// probe instead of a binary search.  All arithmetic is FP64 and follows the
// reference's expression order.
#include <math.h>
#include <string.h>

#include <vector>

#include "kn_hip.h"

#include "bin_kernels.h"
#include "rescale_kernels.h"

namespace {

constexpr int kBlock = 256;   // synthetic-field generator
constexpr int kResWaves = 8;  // rescale: 8 waves per block, 1 mode per lane per step, next step prefetched
constexpr int kResUnroll = 1;
constexpr int kResBlocksPerSlot = 16;

int *g_nan_flag = nullptr;    // device flag raised by the rescale on a NaN factor

int *nan_flag_ptr()
{
    if (!g_nan_flag) {
        KN_CHECK(hipMalloc(&g_nan_flag, sizeof(int)));
        KN_CHECK(hipMemset(g_nan_flag, 0, sizeof(int)));
    }
    return g_nan_flag;
}

template <bool GREEN>
int res_grid(long long rows, size_t smem)
{
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        KN_CHECK(hipGetDevice(&dev));
        KN_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    int per_cu = 0;
    const void *k = (const void *)kn_res::k_rescale_rows_pf<kResWaves, kResUnroll, GREEN>;
    if (smem > 64 * 1024) KN_CHECK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    KN_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kn_res::k_rescale_rows_pf<kResWaves, kResUnroll, GREEN>,
                                                          64 * kResWaves, smem));
    // sixteen blocks per resident slot: smaller contiguous row ranges per block
    // balance the sweep's tail across the channels (1024^3: 3.76 ms at one,
    // 3.43 at eight; with non-temporal loads and stores 3.35-3.40 at sixteen)
    const long long g = (long long)(per_cu > 0 ? per_cu : 1) * cus * kResBlocksPerSlot;
    return (int)(rows < g ? rows : g);
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
        const int kx = kn_res::kval(x, dims), ky = kn_res::kval(y, dims);
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
    const int nbuckets = kn_res::kBucketsPerBin * nbins;
    const size_t smem = kn_res::smem_bytes(nbins, nbuckets);
    if (smem > 160 * 1024) return KN_ERR_ARG;
    kn_res::ResArgs a;
    a.F = (double2 *)d_field;
    a.dims = dims;
    a.startslab = startslab;
    a.nbins = nbins;
    a.nbuckets = nbuckets;
    a.rows = (long long)nslab * dims;
    a.boxsize = boxsize;
    a.norm = norm;
    a.logkk = d_logkk;
    a.ratio = d_ratio;
    a.nan_flag = nan_flag_ptr();
    a.gw = nullptr;
    hipLaunchKernelGGL((kn_res::k_rescale_rows_pf<kResWaves, kResUnroll>), dim3(res_grid<false>(a.rows, smem)),
                       dim3(64 * kResWaves), smem, (hipStream_t)stream, a);
    KN_CHECK(hipGetLastError());
    return 0;
}

// The neutrino rescale and the PM Green's-function multiply that follows it
// in one read + write sweep (no reference counterpart; see the header).
extern "C" int kn_rescale_green_device(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                                       const double *d_logkk, const double *d_ratio, int nbins, double norm,
                                       double asmth, void *stream)
{
    if (nbins < 0 || dims < 2 || nslab < 1 || startslab < 0 || startslab + nslab > dims || !(boxsize > 0) ||
        !(asmth >= 0) || (nbins > 0 && (!d_logkk || !d_ratio))) {
        kn_set_last_error("kn_rescale_green_device: bad arguments");
        return KN_ERR_ARG;
    }
    const int nbuckets = kn_res::kBucketsPerBin * nbins;
    const size_t smem = kn_res::smem_bytes(nbins, nbuckets, dims);
    if (smem > 160 * 1024) {
        kn_set_last_error("kn_rescale_green_device: nbins=%d dims=%d exceed the LDS budget", nbins, dims);
        return KN_ERR_RANGE;
    }
    // the per-axis Green's factors of this (dims, r_s / BoxSize), kept for the
    // process: gw[t] = exp(-kt^2 asmth2) W(t)^-4
    static std::vector<double> iw_h, gw_h;
    static double *gw_d = nullptr;
    static int gw_dims = 0, gw_cap = 0;
    static double gw_asmth2 = -1.0;
    const double sc = 2 * M_PI * asmth / boxsize, asmth2 = sc * sc;
    if (gw_dims != dims || gw_asmth2 != asmth2) {
        kn_bin::build_invwindow(dims, iw_h);
        gw_h.resize(dims);
        for (int t = 0; t < dims; t++) {
            const double kt = kn_res::kval(t, dims), w2 = iw_h[t] * iw_h[t];
            gw_h[t] = exp(-kt * kt * asmth2) * (w2 * w2);
        }
        if (dims > gw_cap) {
            if (gw_d) KN_CHECK(hipFree(gw_d));
            KN_CHECK(hipMalloc(&gw_d, sizeof(double) * dims));
            gw_cap = dims;
        }
        KN_CHECK(hipMemcpy(gw_d, gw_h.data(), sizeof(double) * dims, hipMemcpyHostToDevice));
        gw_dims = dims;
        gw_asmth2 = asmth2;
    }
    kn_res::ResArgs a;
    a.F = (double2 *)d_field;
    a.dims = dims;
    a.startslab = startslab;
    a.nbins = nbins;
    a.nbuckets = nbuckets;
    a.rows = (long long)nslab * dims;
    a.boxsize = boxsize;
    a.norm = norm;
    a.logkk = d_logkk;
    a.ratio = d_ratio;
    a.nan_flag = nan_flag_ptr();
    a.gw = gw_d;
    hipLaunchKernelGGL((kn_res::k_rescale_rows_pf<kResWaves, kResUnroll, true>), dim3(res_grid<true>(a.rows, smem)),
                       dim3(64 * kResWaves), smem, (hipStream_t)stream, a);
    KN_CHECK(hipGetLastError());
    return 0;
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

// The rescale of add_nu_power_to_rhogrid with the ratio table on the host:
// the table goes through a pinned staging buffer into a device buffer, both
// kept for the process and grown when a longer table arrives, so a PM step
// allocates nothing; the copy, the sweep and the read-back of the NaN flag
// are ordered on the null stream and the one synchronisation is the flag's.
// Returns the flag (1: a NaN factor, terminate 5 in the caller).
extern "C" int kn_rescale_from_host_table(fftw_complex *d_field, int dims, int startslab, int nslab, double boxsize,
                                          const double *logkk, const double *ratio, int nbins, double norm)
{
    static double *d_tab = nullptr, *h_tab = nullptr;
    static int *h_flag = nullptr;
    static int cap = 0;
    if (nbins > cap) {
        if (d_tab) KN_CHECK(hipFree(d_tab));
        if (h_tab) KN_CHECK(hipHostFree(h_tab));
        KN_CHECK(hipMalloc(&d_tab, sizeof(double) * 2 * nbins));
        KN_CHECK(hipHostMalloc((void **)&h_tab, sizeof(double) * 2 * nbins, hipHostMallocDefault));
        cap = nbins;
    }
    if (!h_flag) KN_CHECK(hipHostMalloc((void **)&h_flag, sizeof(int), hipHostMallocDefault));
    memcpy(h_tab, logkk, sizeof(double) * nbins);
    memcpy(h_tab + nbins, ratio, sizeof(double) * nbins);
    KN_CHECK(hipMemcpyAsync(d_tab, h_tab, sizeof(double) * 2 * nbins, hipMemcpyHostToDevice, 0));
    const int rc = kn_rescale_device(d_field, dims, startslab, nslab, boxsize, d_tab, d_tab + nbins, nbins, norm, 0);
    if (rc) kn_terminate(rc, "rescale: bad table (nbins=%d)\n", nbins);
    KN_CHECK(hipMemcpyAsync(h_flag, nan_flag_ptr(), sizeof(int), hipMemcpyDeviceToHost, 0));
    KN_CHECK(hipStreamSynchronize(0));
    const int nan = *h_flag;
    if (nan) KN_CHECK(hipMemset(nan_flag_ptr(), 0, sizeof(int)));
    return nan;
}
