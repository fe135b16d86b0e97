// powerspectrum.hip -- (a) k-shell binning of the r2c PM grid into P(k).
//
// Replaces total_powerspectrum (powerspectrum.c:33-117, decl powerspectrum.h:30).
// Semantics kept from the reference:
//  - bin of a mode = floor(binsperunit * log(|k|)), binsperunit =
//    (nrbins-1)/log(sqrt(3) dims/2) (powerspectrum.c:40, :67);
//  - weight multiplicity * |F|^2 * W^-4, W^-1 the per-axis CiC inverse window
//    pi k/(n sin(pi k/n)) (powerspectrum.c:8-24, squared twice at :68);
//  - multiplicity 1 on the kz = 0 and kz = N/2 planes, 2 elsewhere; the zero
//    mode is skipped (powerspectrum.c:63-87);
//  - power /= |F0|^2, then /= count; keffs /= count; empty bins compacted
//    (powerspectrum.c:97-116).
//
// MI355X design: the pass streams the grid once from HBM (16 B per mode,
// coalesced non-temporal dwordx4 loads, transposed through LDS so that each
// lane bins a run of consecutive modes; see bin_kernels.h).  The bin index is
// integer-exact: the host evaluates the reference expression once per integer
// |k|^2 and stores, per bin, the smallest |k|^2 that reaches it; the kernel
// takes an fp32 estimate of the bin and corrects it against that threshold
// table, so counts are bit-exact with the reference's libm.  The wavefronts of
// a block add into one set of bin sums in LDS (ds_add_f64; the order of the
// adds between wavefronts is not fixed, so a sum can differ in its last bit
// from run to run), blocks write partial sums and a second kernel adds them in
// a fixed order.  Tables go to LDS when they fit, else are read from global
// memory, so any number of bins works.  Mode counts and <|k|> depend only on
// the geometry; they are computed once per (dims, nrbins, slab range) by a
// read-free kernel and cached (bit-exact, deterministic).
#include <math.h>
#include <string.h>
#include <vector>

#include "kn_hip.h"

#include "bin_kernels.h"

namespace {

using kn_bin::Pos;
using kn_bin::advance;
using kn_bin::find_bin;
using kn_bin::kval;

constexpr int kBlock = 256;  // geometry pass
constexpr int kWaves = kBlock / 64;
constexpr int kSweepSlots = 2;  // bin sweep blocks per resident slot (1.41-1.44 ms at 1024^3 against 1.45 at three)
constexpr size_t kLdsMax = 160 * 1024;

struct BinGeom {
    int dims, nrbins, startslab, nslab;
    int *d_thr;       // nrbins+1 thresholds on |k|^2
    double *d_iw;     // per-axis inverse window by storage index
    float half_bpu;   // 0.5 * binsperunit, for the fp32 estimate
    std::vector<long long> count;
    std::vector<double> keff_sum;
    int grid, ggrid;        // power-pass and geometry-pass grids
    int place;              // where the sweep keeps its tables (kn_bin::TablePlace)
    size_t smem;            // the sweep's LDS bytes
    double *d_partial;      // max(grid, ggrid) * nrbins
    unsigned *d_cpartial;   // ggrid * nrbins (geometry pass)
    double *d_sum;          // nrbins + 1 (last: |F0|^2)
    double *d_alpha;        // nrbins, adjoint coefficients (allocated on first use)
    int agrid;              // adjoint-sweep grid
};

std::vector<BinGeom *> g_geoms;

// Geometry pass (no grid reads): multiplicity counts and sum of multiplicity*|k|.
// GLOBAL: the per-block sums are added with global atomics into the block's
// partial rows and the thresholds read from global memory (any number of bins).
template <bool GLOBAL>
__global__ __launch_bounds__(kBlock) void k_bin_geometry(int dims, int startslab, long long rows, int nrbins,
                                                          const int *__restrict__ thr_g, float half_bpu,
                                                          unsigned *__restrict__ cpartial,
                                                          double *__restrict__ kpartial)
{
    extern __shared__ double smem[];
    double *khist;
    unsigned *chist;
    const int *thr;
    const int tid = threadIdx.x, wave = tid >> 6;
    if constexpr (GLOBAL) {
        khist = kpartial + (size_t)blockIdx.x * nrbins;
        chist = cpartial + (size_t)blockIdx.x * nrbins;
        thr = thr_g;
        for (int t = tid; t < nrbins; t += kBlock) {
            khist[t] = 0.0;
            chist[t] = 0u;
        }
    } else {
        khist = smem;                                          // kWaves * nrbins
        chist = (unsigned *)(khist + (size_t)kWaves * nrbins);  // kWaves * nrbins
        int *thr_l = (int *)(chist + (size_t)kWaves * nrbins);
        for (int t = tid; t < kWaves * nrbins; t += kBlock) {
            khist[t] = 0.0;
            chist[t] = 0u;
        }
        for (int t = tid; t <= nrbins; t += kBlock) thr_l[t] = thr_g[t];
        thr = thr_l;
    }
    const size_t hoff = GLOBAL ? 0 : (size_t)wave * nrbins;
    __syncthreads();
    const int rowlen = dims / 2 + 1, half = dims / 2;
    const long long r0 = rows * blockIdx.x / gridDim.x, r1 = rows * (blockIdx.x + 1) / gridDim.x;
    const long long m1 = r1 * rowlen;
    long long m = r0 * rowlen + tid;
    Pos p;
    {
        const long long rl = r0 + tid / rowlen;
        p.z = tid % rowlen;
        p.j = (int)(rl % dims);
        p.i = (int)(rl / dims);
    }
    for (; m < m1; m += kBlock) {
        const int kx = kval(startslab + p.i, dims), ky = kval(p.j, dims), kz = p.z;
        const int k2 = kx * kx + ky * ky + kz * kz;
        if (k2 != 0) {
            const int b = find_bin(k2, half_bpu, thr, nrbins);
            const unsigned c = (kz != 0 && kz != half) ? 2u : 1u;
            const double kk = sqrt((double)k2);
            atomicAdd(&chist[hoff + b], c);
            atomicAdd(&khist[hoff + b], c == 2u ? 2.0 * kk : kk);
        }
        advance(p, kBlock, rowlen, dims);
    }
    if constexpr (GLOBAL) return;
    __syncthreads();
    for (int b = tid; b < nrbins; b += kBlock) {
        double s = khist[b];
        unsigned c = chist[b];
        for (int w = 1; w < kWaves; w++) {
            s += khist[(size_t)w * nrbins + b];
            c += chist[(size_t)w * nrbins + b];
        }
        kpartial[(size_t)blockIdx.x * nrbins + b] = s;
        cpartial[(size_t)blockIdx.x * nrbins + b] = c;
    }
}

// Fixed-order sum of per-block partial histograms: 64 bins per block across
// lanes, 16 waves split the blocks, then a fixed-order combine in LDS.
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(64 * kRedWaves) void k_reduce_partials(const double *__restrict__ partial, int nblk,
                                                                     int nrbins, double *__restrict__ out)
{
    __shared__ double acc[kRedWaves][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int b = blockIdx.x * 64 + lane;
    const int g0 = nblk * wave / kRedWaves, g1 = nblk * (wave + 1) / kRedWaves;
    double s = 0.0;
    if (b < nrbins) {
        // eight loads in flight, added in block order
        constexpr int U = 8;
        int g = g0;
        for (; g + U <= g1; g += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; u++) v[u] = partial[(size_t)(g + u) * nrbins + b];
#pragma unroll
            for (int u = 0; u < U; u++) s += v[u];
        }
        for (; g < g1; g++) s += partial[(size_t)g * nrbins + b];
    }
    acc[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && b < nrbins) {
        double t = acc[0][lane];
        for (int w = 1; w < kRedWaves; w++) t += acc[w][lane];
        out[b] = t;
    }
}

__global__ void k_reduce_counts(const unsigned *__restrict__ cpartial, const double *__restrict__ kpartial,
                                int nblk, int nrbins, unsigned long long *__restrict__ count,
                                double *__restrict__ ksum)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nrbins) return;
    unsigned long long c = 0;
    double s = 0.0;
    for (int g = 0; g < nblk; g++) {
        c += cpartial[(size_t)g * nrbins + b];
        s += kpartial[(size_t)g * nrbins + b];
    }
    count[b] = c;
    ksum[b] = s;
}

// |F0|^2 from the rank holding slab 0 (powerspectrum.c:45-47), else 0.
__global__ void k_total_mass2(const double2 *__restrict__ F, int has_zero_mode, double *__restrict__ out)
{
    if (has_zero_mode) {
        const double2 f = F[0];
        *out = f.x * f.x + f.y * f.y;
    } else {
        *out = 0.0;
    }
}

int num_cus()
{
    static int n = 0;
    if (!n) {
        int dev = 0;
        KN_CHECK(hipGetDevice(&dev));
        KN_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    }
    return n;
}

size_t geometry_smem(int nrbins) { return (sizeof(double) + sizeof(unsigned)) * (size_t)kWaves * nrbins + sizeof(int) * (nrbins + 1); }

// Resident blocks per CU of a kernel at a given dynamic LDS size.
template <typename K>
int resident_blocks(K kernel, int threads, size_t smem)
{
    int n = 0;
    if (smem > 64 * 1024)
        KN_CHECK(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    KN_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, smem));
    return n > 0 ? n : 1;
}

BinGeom *get_geom(int dims, int nrbins, int startslab, int nslab)
{
    for (BinGeom *g : g_geoms)
        if (g->dims == dims && g->nrbins == nrbins && g->startslab == startslab && g->nslab == nslab) return g;
    if (dims < 2 || nrbins < 1 || nslab < 1 || startslab < 0 || startslab + nslab > dims)
        kn_terminate(KN_ERR_ARG, "total_powerspectrum: bad geometry dims=%d nrbins=%d slabs=[%d,%d)\n", dims,
                     nrbins, startslab, startslab + nslab);
    BinGeom *g = new BinGeom();
    g->dims = dims;
    g->nrbins = nrbins;
    g->startslab = startslab;
    g->nslab = nslab;
    std::vector<int> thr;
    const long long bad = kn_bin::build_thresholds(dims, nrbins, thr);
    if (bad) kn_terminate(KN_ERR_RANGE, "bin index of |k|^2 = %lld is not monotone or beyond %d bins\n", bad, nrbins);
    std::vector<double> iw;
    kn_bin::build_invwindow(dims, iw);
    g->half_bpu = (float)(0.5 * (nrbins - 1) / log(sqrt(3) * dims / 2.0));
    const long long rows = (long long)nslab * dims;
    // the first table placement whose LDS leaves room for two blocks per CU,
    // else the first that fits at all (kTablesGlobal always does)
    g->place = kn_bin::kTablesGlobal;
    for (int pl : {kn_bin::kTablesLds, kn_bin::kWindowGlobal})
        if (kn_bin::sweep_smem_bytes(pl, dims, nrbins) <= kLdsMax / 2) {
            g->place = pl;
            break;
        }
    if (g->place == kn_bin::kTablesGlobal)
        for (int pl : {kn_bin::kTablesLds, kn_bin::kWindowGlobal})
            if (kn_bin::sweep_smem_bytes(pl, dims, nrbins) <= kLdsMax) {
                g->place = pl;
                break;
            }
    g->smem = kn_bin::sweep_smem_bytes(g->place, dims, nrbins);
    const int per_cu = g->place == kn_bin::kTablesLds
                           ? resident_blocks(kn_bin::k_bin_sweep<kn_bin::kTablesLds>, 64 * kn_bin::kSweepWaves, g->smem)
                       : g->place == kn_bin::kWindowGlobal
                           ? resident_blocks(kn_bin::k_bin_sweep<kn_bin::kWindowGlobal>, 64 * kn_bin::kSweepWaves,
                                             g->smem)
                           : resident_blocks(kn_bin::k_bin_sweep<kn_bin::kTablesGlobal>, 64 * kn_bin::kSweepWaves,
                                             g->smem);
    const long long pgrid = (long long)per_cu * num_cus() * kSweepSlots;
    g->grid = (int)(rows < pgrid ? rows : pgrid);
    g->ggrid = (int)(rows < 4LL * num_cus() ? rows : 4LL * num_cus());
    const int pmax = g->grid > g->ggrid ? g->grid : g->ggrid;
    KN_CHECK(hipMalloc(&g->d_thr, sizeof(int) * (nrbins + 1)));
    KN_CHECK(hipMalloc(&g->d_iw, sizeof(double) * dims));
    KN_CHECK(hipMalloc(&g->d_partial, sizeof(double) * (size_t)pmax * nrbins));
    KN_CHECK(hipMalloc(&g->d_cpartial, sizeof(unsigned) * (size_t)g->ggrid * nrbins));
    KN_CHECK(hipMalloc(&g->d_sum, sizeof(double) * (nrbins + 1)));
    KN_CHECK(hipMemcpy(g->d_thr, thr.data(), sizeof(int) * (nrbins + 1), hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(g->d_iw, iw.data(), sizeof(double) * dims, hipMemcpyHostToDevice));
    // geometry pass, once per geometry
    unsigned long long *d_count;
    double *d_ksum;
    KN_CHECK(hipMalloc(&d_count, sizeof(unsigned long long) * nrbins));
    KN_CHECK(hipMalloc(&d_ksum, sizeof(double) * nrbins));
    const size_t gsmem = geometry_smem(nrbins);
    if (gsmem <= kLdsMax) {
        if (gsmem > 64 * 1024)
            KN_CHECK(hipFuncSetAttribute((const void *)k_bin_geometry<false>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)gsmem));
        hipLaunchKernelGGL(k_bin_geometry<false>, dim3(g->ggrid), dim3(kBlock), gsmem, 0, dims, startslab, rows,
                           nrbins, g->d_thr, g->half_bpu, g->d_cpartial, g->d_partial);
    } else {
        hipLaunchKernelGGL(k_bin_geometry<true>, dim3(g->ggrid), dim3(kBlock), 0, 0, dims, startslab, rows, nrbins,
                           g->d_thr, g->half_bpu, g->d_cpartial, g->d_partial);
    }
    KN_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_counts, dim3((nrbins + 255) / 256), dim3(256), 0, 0, g->d_cpartial, g->d_partial,
                       g->ggrid, nrbins, d_count, d_ksum);
    KN_CHECK(hipGetLastError());
    std::vector<unsigned long long> c(nrbins);
    g->count.resize(nrbins);
    g->keff_sum.resize(nrbins);
    KN_CHECK(hipMemcpy(c.data(), d_count, sizeof(unsigned long long) * nrbins, hipMemcpyDeviceToHost));
    KN_CHECK(hipMemcpy(g->keff_sum.data(), d_ksum, sizeof(double) * nrbins, hipMemcpyDeviceToHost));
    for (int b = 0; b < nrbins; b++) g->count[b] = (long long)c[b];
    KN_CHECK(hipFree(d_count));
    KN_CHECK(hipFree(d_ksum));
    g_geoms.push_back(g);
    return g;
}

// Power pass; the bin sums land in sum_out[nrbins] and |F0|^2 in mass_out
// (default: the geometry's own buffer, sums then |F0|^2).
void launch_power(BinGeom *g, const double2 *F, hipStream_t st, double *sum_out = nullptr, double *mass_out = nullptr)
{
    if (!sum_out) sum_out = g->d_sum;
    if (!mass_out) mass_out = g->d_sum + g->nrbins;
    kn_bin::BinArgs a;
    a.F = F;
    a.dims = g->dims;
    a.startslab = g->startslab;
    a.nrbins = g->nrbins;
    a.rows = (long long)g->nslab * g->dims;
    a.thr = g->d_thr;
    a.iw = g->d_iw;
    a.half_bpu = g->half_bpu;
    a.partial = g->d_partial;
    const dim3 grid(g->grid), block(64 * kn_bin::kSweepWaves);
    if (g->place == kn_bin::kTablesLds)
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kTablesLds>, grid, block, g->smem, st, a);
    else if (g->place == kn_bin::kWindowGlobal)
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kWindowGlobal>, grid, block, g->smem, st, a);
    else
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kTablesGlobal>, grid, block, g->smem, st, a);
    KN_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials, dim3((g->nrbins + 63) / 64), dim3(64 * kRedWaves), 0, st, g->d_partial,
                       g->grid, g->nrbins, sum_out);
    KN_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_total_mass2, dim3(1), dim3(1), 0, st, F, g->startslab == 0 ? 1 : 0, mass_out);
    KN_CHECK(hipGetLastError());
}

// Adjoint sweep of the binning: Fbar_m = s_m F_m with s_m = 2 c_m W_m^-4
// alpha[bin(m)] (c_m the multiplicity) and s_0 = beta for the zero mode.
// Reads 16 B and writes 16 B per mode; software-pipelined like the rescale sweep.
constexpr int kAdjWaves = 8;
constexpr int kAdjUnroll = 2;

// GLOBAL: the tables are read from global memory (through the caches) when
// they do not fit in LDS.
template <bool GLOBAL>
__global__ __launch_bounds__(64 * kAdjWaves) void k_bin_adjoint(const double2 *__restrict__ F,
                                                                double2 *__restrict__ Fb, int dims, int startslab,
                                                                long long rows, int nrbins,
                                                                const int *__restrict__ thr_g,
                                                                const double *__restrict__ iw_g,
                                                                const double *__restrict__ alpha_g, float half_bpu,
                                                                double beta)
{
    extern __shared__ double smem[];
    const double *alpha = alpha_g, *iw = iw_g;
    const int *thr = thr_g;
    constexpr int B = 64 * kAdjWaves;
    if constexpr (!GLOBAL) {
        double *alpha_l = smem;
        double *iw_l = alpha_l + nrbins;
        int *thr_l = (int *)(iw_l + dims);
        for (int t = threadIdx.x; t < nrbins; t += B) alpha_l[t] = alpha_g[t];
        for (int t = threadIdx.x; t < dims; t += B) iw_l[t] = iw_g[t];
        for (int t = threadIdx.x; t <= nrbins; t += B) thr_l[t] = thr_g[t];
        alpha = alpha_l;
        iw = iw_l;
        thr = thr_l;
        __syncthreads();
    }
    const int rowlen = dims / 2 + 1, half = dims / 2;
    const long long r0 = rows * blockIdx.x / gridDim.x, r1 = rows * (blockIdx.x + 1) / gridDim.x;
    const long long m1 = r1 * rowlen;
    constexpr long long S = (long long)B * kAdjUnroll;
    Pos p = kn_bin::pos_of(threadIdx.x, r0, rowlen, dims);
    long long m = r0 * rowlen + threadIdx.x;
    double2 cur[kAdjUnroll];
#pragma unroll
    for (int u = 0; u < kAdjUnroll; u++) {
        const long long mu = m + (long long)B * u;
        cur[u] = mu < m1 ? F[mu] : make_double2(0.0, 0.0);
    }
    for (; m < m1; m += S) {
        double2 nxt[kAdjUnroll];
#pragma unroll
        for (int u = 0; u < kAdjUnroll; u++) {
            const long long mu = m + S + (long long)B * u;
            nxt[u] = mu < m1 ? F[mu] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < kAdjUnroll; u++) {
            const long long mu = m + (long long)B * u;
            const Pos q = p;
            advance(p, B, rowlen, dims);
            if (mu >= m1) continue;
            const int kx = kval(startslab + q.i, dims), ky = kval(q.j, dims);
            const int k2 = kx * kx + ky * ky + q.z * q.z;
            double sc;
            if (k2 == 0) {
                sc = beta;
            } else {
                const int b = find_bin(k2, half_bpu, thr, nrbins);
                const double w = iw[startslab + q.i] * iw[q.j] * iw[q.z];
                const double w2 = w * w;
                sc = ((q.z != 0 && q.z != half) ? 4.0 : 2.0) * (w2 * w2) * alpha[b];
            }
            Fb[mu] = make_double2(sc * cur[u].x, sc * cur[u].y);
        }
#pragma unroll
        for (int u = 0; u < kAdjUnroll; u++) cur[u] = nxt[u];
    }
}

size_t adjoint_smem(int dims, int nrbins) { return sizeof(double) * (size_t)(nrbins + dims) + sizeof(int) * (nrbins + 1); }

}  // namespace

extern "C" int kn_bin_power_device(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                                   double *d_power, double *d_total_mass2, void *stream)
{
    BinGeom *g = get_geom(dims, nrbins, startslab, nslab);
    launch_power(g, (const double2 *)d_field, (hipStream_t)stream, d_power, d_total_mass2);
    return 0;
}

// ---- the bin sweep without allocation ---------------------------------
// kn_bin_power_device keeps its bin thresholds, inverse window and per-block
// partial sums in device buffers it allocates when it first meets a geometry.
// kn_bin_power_device_noalloc keeps them instead in `scratch`, a file-scope
// device array of kScratchBytes: every GPU the library runs on holds its own
// copy until the process ends, and every call on that GPU shares it (the
// tables of the last geometry stay there; the partial sums are rewritten
// whole by every sweep, so nothing has to start at zero).  Calls that overlap
// on one GPU race on it.
constexpr size_t kScratchBytes = (size_t)32 << 20;  // per GPU

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kScratchMaxDims = 65536, kScratchMaxBins = 1 << 18;
constexpr int kScratchMaxDevices = 64;

struct ScratchDev {
    unsigned char *ptr = nullptr;  // this device's copy of scratch (hipGetSymbolAddress)
    int dims = 0, nrbins = 0;      // the geometry whose tables the copy holds
};
ScratchDev g_scratch[kScratchMaxDevices];

// host copies of the last geometry's tables (no allocation per call)
int g_sthr[kScratchMaxBins + 1];
double g_siw[kScratchMaxDims];
int g_sdims = 0, g_snrbins = 0;

inline size_t round256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

// kn_bin_power_device (same arguments and outputs) with no allocation: the
// sweep's tables and partial sums live in scratch.  Limits: dims <= 65536,
// nrbins <= 262144, and the tables plus at least one block's partial sums
// within 32 MiB; the sweep's grid is capped so that its partial sums fit.
// Returns 0, 3002 for bad arguments, 3004 for sizes beyond those limits.
extern "C" int kn_bin_power_device_noalloc(const fftw_complex *d_field, int dims, int nrbins, int startslab,
                                           int nslab, double *d_power, double *d_total_mass2, void *stream)
{
    if (!d_field || !d_power || !d_total_mass2 || dims < 2 || nrbins < 1 || nslab < 1 || startslab < 0 ||
        startslab + nslab > dims) {
        kn_set_last_error("kn_bin_power_device_noalloc: bad geometry dims=%d nrbins=%d slabs=[%d,%d)", dims, nrbins,
                          startslab, startslab + nslab);
        return KN_ERR_ARG;
    }
    const size_t off_iw = round256(sizeof(int) * ((size_t)nrbins + 1));
    const size_t off_part = off_iw + round256(sizeof(double) * (size_t)dims);
    if (dims > kScratchMaxDims || nrbins > kScratchMaxBins || off_part + sizeof(double) * nrbins > kScratchBytes) {
        kn_set_last_error("kn_bin_power_device_noalloc: dims=%d nrbins=%d beyond the %zu-byte scratch (dims <= %d, "
                          "nrbins <= %d)", dims, nrbins, kScratchBytes, kScratchMaxDims, kScratchMaxBins);
        return KN_ERR_RANGE;
    }
    int dev = 0;
    KN_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchMaxDevices) {
        kn_set_last_error("kn_bin_power_device_noalloc: device %d beyond %d", dev, kScratchMaxDevices);
        return KN_ERR_RANGE;
    }
    ScratchDev &sd = g_scratch[dev];
    if (!sd.ptr) {
        const hipError_t e = hipGetSymbolAddress((void **)&sd.ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            kn_set_last_error("kn_bin_power_device_noalloc: hipGetSymbolAddress failed: %s", hipGetErrorString(e));
            sd.ptr = nullptr;
            return KN_ERR_HIP;
        }
    }
    hipStream_t st = (hipStream_t)stream;
    if (sd.dims != dims || sd.nrbins != nrbins) {
        if (g_sdims != dims || g_snrbins != nrbins) {
            const long long bad = kn_bin::build_thresholds(dims, nrbins, g_sthr);
            if (bad) {
                g_sdims = g_snrbins = 0;
                kn_set_last_error("kn_bin_power_device_noalloc: bin index of |k|^2 = %lld is not monotone or "
                                  "beyond %d bins", bad, nrbins);
                return KN_ERR_RANGE;
            }
            kn_bin::build_invwindow(dims, g_siw);
            g_sdims = dims;
            g_snrbins = nrbins;
        }
        KN_CHECK(hipMemcpyAsync(sd.ptr, g_sthr, sizeof(int) * ((size_t)nrbins + 1), hipMemcpyHostToDevice, st));
        KN_CHECK(hipMemcpyAsync(sd.ptr + off_iw, g_siw, sizeof(double) * dims, hipMemcpyHostToDevice, st));
        KN_CHECK(hipStreamSynchronize(st));  // the host tables may be rebuilt by the next call
        sd.dims = dims;
        sd.nrbins = nrbins;
    }
    // placement and grid as get_geom chooses them, the grid capped by the scratch
    int place = kn_bin::kTablesGlobal;
    for (int pl : {kn_bin::kTablesLds, kn_bin::kWindowGlobal})
        if (kn_bin::sweep_smem_bytes(pl, dims, nrbins) <= kLdsMax / 2) {
            place = pl;
            break;
        }
    if (place == kn_bin::kTablesGlobal)
        for (int pl : {kn_bin::kTablesLds, kn_bin::kWindowGlobal})
            if (kn_bin::sweep_smem_bytes(pl, dims, nrbins) <= kLdsMax) {
                place = pl;
                break;
            }
    const size_t smem = kn_bin::sweep_smem_bytes(place, dims, nrbins);
    const int per_cu = place == kn_bin::kTablesLds
                           ? resident_blocks(kn_bin::k_bin_sweep<kn_bin::kTablesLds>, 64 * kn_bin::kSweepWaves, smem)
                       : place == kn_bin::kWindowGlobal
                           ? resident_blocks(kn_bin::k_bin_sweep<kn_bin::kWindowGlobal>, 64 * kn_bin::kSweepWaves,
                                             smem)
                           : resident_blocks(kn_bin::k_bin_sweep<kn_bin::kTablesGlobal>, 64 * kn_bin::kSweepWaves,
                                             smem);
    const long long rows = (long long)nslab * dims;
    long long grid = (long long)per_cu * num_cus() * kSweepSlots;
    const long long fit = (long long)((kScratchBytes - off_part) / (sizeof(double) * (size_t)nrbins));
    if (grid > fit) grid = fit;
    if (grid > rows) grid = rows;
    kn_bin::BinArgs a;
    a.F = (const double2 *)d_field;
    a.dims = dims;
    a.startslab = startslab;
    a.nrbins = nrbins;
    a.rows = rows;
    a.thr = (const int *)sd.ptr;
    a.iw = (const double *)(sd.ptr + off_iw);
    a.half_bpu = (float)(0.5 * (nrbins - 1) / log(sqrt(3) * dims / 2.0));
    a.partial = (double *)(sd.ptr + off_part);
    const dim3 gr((unsigned)grid), block(64 * kn_bin::kSweepWaves);
    if (place == kn_bin::kTablesLds)
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kTablesLds>, gr, block, smem, st, a);
    else if (place == kn_bin::kWindowGlobal)
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kWindowGlobal>, gr, block, smem, st, a);
    else
        hipLaunchKernelGGL(kn_bin::k_bin_sweep<kn_bin::kTablesGlobal>, gr, block, smem, st, a);
    KN_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_partials, dim3((nrbins + 63) / 64), dim3(64 * kRedWaves), 0, st, a.partial,
                       (int)grid, nrbins, d_power);
    KN_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_total_mass2, dim3(1), dim3(1), 0, st, a.F, startslab == 0 ? 1 : 0, d_total_mass2);
    KN_CHECK(hipGetLastError());
    return 0;
}

// Geometry sums for callers that already hold the power sums (add_nu_power_to_rhogrid).
extern "C" void kn_bin_geometry_host(int dims, int nrbins, int startslab, int nslab, long long *count,
                                     double *keff_sum)
{
    BinGeom *g = get_geom(dims, nrbins, startslab, nslab);
    memcpy(count, g->count.data(), sizeof(long long) * nrbins);
    memcpy(keff_sum, g->keff_sum.data(), sizeof(double) * nrbins);
}

// Normalise and compact as powerspectrum.c:97-116.
extern "C" int kn_finish_powerspectrum(int nrbins, double total_mass2, double *power, long long *count,
                                       double *keffs)
{
    int nonzero = 0;
    kn_message(0, "total mass of the grid, |F(k = 0)|: %g\n", sqrt(total_mass2));
    for (int i = 0; i < nrbins; i++) {
        power[i] /= total_mass2;
        if (count[i]) {
            keffs[i] /= count[i];
            power[i] /= count[i];
        }
    }
    for (int i = 0; i < nrbins; i++) {
        if (count[i]) {
            if (nonzero < i) {
                power[nonzero] = power[i];
                keffs[nonzero] = keffs[i];
                count[nonzero] = count[i];
            }
            nonzero++;
        }
    }
    return nonzero;
}

extern "C" void kn_bin_power_host_sums(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                                       double *power, double *total_mass2)
{
    BinGeom *g = get_geom(dims, nrbins, startslab, nslab);
    std::vector<double> sums(nrbins + 1);
    launch_power(g, (const double2 *)d_field, 0);
    KN_CHECK(hipMemcpy(sums.data(), g->d_sum, sizeof(double) * (nrbins + 1), hipMemcpyDeviceToHost));
    memcpy(power, sums.data(), sizeof(double) * nrbins);
    *total_mass2 = sums[nrbins];
}

extern "C" int total_powerspectrum(const int dims, fftw_complex *outfield, const int nrbins, const int startslab,
                                   const int nslab, double *power, long long int *count, double *keffs,
                                   MPI_Comm comm)
{
    (void)comm;
    BinGeom *g = get_geom(dims, nrbins, startslab, nslab);
    const double2 *F = (const double2 *)outfield;
    double2 *staged = nullptr;
    if (!kn_is_device_ptr(outfield)) {
        const size_t bytes = sizeof(double2) * (size_t)nslab * dims * (dims / 2 + 1);
        KN_CHECK(hipMalloc(&staged, bytes));
        KN_CHECK(hipMemcpy(staged, outfield, bytes, hipMemcpyHostToDevice));
        F = staged;
    }
    launch_power(g, F, 0);
    std::vector<double> sums(nrbins + 1);
    KN_CHECK(hipMemcpy(sums.data(), g->d_sum, sizeof(double) * (nrbins + 1), hipMemcpyDeviceToHost));
    if (staged) KN_CHECK(hipFree(staged));
    memcpy(power, sums.data(), sizeof(double) * nrbins);
    memcpy(count, g->count.data(), sizeof(long long) * nrbins);
    memcpy(keffs, g->keff_sum.data(), sizeof(double) * nrbins);
    return kn_finish_powerspectrum(nrbins, sums[nrbins], power, count, keffs);
}

// Adjoint of total_powerspectrum with respect to the field (device arrays):
// gbar[nout] = dL/dpower[n] for the nout compacted bins of the forward on the
// same field (host array); d_field_bar receives dL/dF in the field's layout
// (Re and Im parts: dL/dRe F + i dL/dIm F).  Needs the zero mode (startslab 0).
extern "C" int kn_bin_power_adjoint(const fftw_complex *d_field, int dims, int nrbins, int startslab, int nslab,
                                    const double *gbar, int nout, fftw_complex *d_field_bar, void *stream)
{
    if (startslab != 0 || !gbar || !d_field || !d_field_bar) {
        kn_set_last_error("kn_bin_power_adjoint: needs startslab 0 (the zero mode) and non-null arrays");
        return KN_ERR_ARG;
    }
    BinGeom *g = get_geom(dims, nrbins, startslab, nslab);
    hipStream_t st = (hipStream_t)stream;
    launch_power(g, (const double2 *)d_field, st);
    std::vector<double> sums(nrbins + 1), alpha(nrbins, 0.0);
    KN_CHECK(hipMemcpyAsync(sums.data(), g->d_sum, sizeof(double) * (nrbins + 1), hipMemcpyDeviceToHost, st));
    KN_CHECK(hipStreamSynchronize(st));
    const double M = sums[nrbins];
    double gp = 0.0;
    int n = 0;
    for (int b = 0; b < nrbins; b++) {
        if (!g->count[b]) continue;
        if (n >= nout) break;
        const double P = sums[b] / M / (double)g->count[b];
        alpha[b] = gbar[n] / (M * (double)g->count[b]);
        gp += gbar[n] * P;
        n++;
    }
    int nonzero = 0;
    for (int b = 0; b < nrbins; b++) nonzero += g->count[b] ? 1 : 0;
    if (nonzero != nout) {
        kn_set_last_error("kn_bin_power_adjoint: nout=%d but the forward has %d non-empty bins", nout, nonzero);
        return KN_ERR_ARG;
    }
    const bool adj_global = adjoint_smem(dims, nrbins) > kLdsMax;
    const size_t asmem = adj_global ? 0 : adjoint_smem(dims, nrbins);
    if (!g->d_alpha) {
        KN_CHECK(hipMalloc(&g->d_alpha, sizeof(double) * nrbins));
        const long long rows = (long long)nslab * dims;
        const int per_cu = adj_global ? resident_blocks(k_bin_adjoint<true>, 64 * kAdjWaves, 0)
                                      : resident_blocks(k_bin_adjoint<false>, 64 * kAdjWaves, asmem);
        const long long ag = (long long)per_cu * num_cus();
        g->agrid = (int)(rows < ag ? rows : ag);
    }
    KN_CHECK(hipMemcpyAsync(g->d_alpha, alpha.data(), sizeof(double) * nrbins, hipMemcpyHostToDevice, st));
    if (adj_global)
        hipLaunchKernelGGL(k_bin_adjoint<true>, dim3(g->agrid), dim3(64 * kAdjWaves), 0, st, (const double2 *)d_field,
                           (double2 *)d_field_bar, dims, startslab, (long long)nslab * dims, nrbins, g->d_thr,
                           g->d_iw, g->d_alpha, g->half_bpu, -2.0 * gp / M);
    else
        hipLaunchKernelGGL(k_bin_adjoint<false>, dim3(g->agrid), dim3(64 * kAdjWaves), asmem, st,
                           (const double2 *)d_field, (double2 *)d_field_bar, dims, startslab, (long long)nslab * dims,
                           nrbins, g->d_thr, g->d_iw, g->d_alpha, g->half_bpu, -2.0 * gp / M);
    KN_CHECK(hipGetLastError());
    KN_CHECK(hipStreamSynchronize(st));
    return 0;
}
