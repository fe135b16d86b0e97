// delta_nu.hip -- (b) the free-streaming integral get_delta_nu on the GPU.
//
// Replaces the integral of get_delta_nu (delta_tot_table.c:551-605) for every
// species at once.  The host keeps what the reference computes once per call
// with the host's own hubble_function (the initial-condition piece,
// fslength(T0, a), delta_tot_table.c:522-535) and the per-species parameters
// (:539-548); the device computes, in order:
//   k_fs_first,  the 16 Na free-streaming lengths fslength(x_i, log a) on the
//   k_fs_rest    reference's uniform log-a grid, each a QAG-GK61 integral of
//                1/(a^2 H) (relerr 1e-6): the first estimate for two nodes per
//                wavefront, full QAG (one wavefront each) for those it leaves;
//   k_spline     natural cubic spline coefficients of that table (one
//                wavefront per row window: scans of the substitution recurrences);
//   k_rows_prep  the same for every delta_tot[k][0..Na-1] row, and a uniform
//                bucket index over the stored log a values, in one launch;
//   k_dnu_qag    one wavefront per (species, k): QAG-GK61 of the integrand
//                fsl/(a H) * specialJ * S_k over [log T0, log a] at the
//                species' tolerance, as the reference does per k.
// H(a) enters the device integrands as Chebyshev interpolants of 1/(a^2 H) and
// 1/(a H) in ln a, sampled from the host's hubble_function on fixed panels
// (built once, extended on demand).
#include <math.h>
#include <string.h>
#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <thread>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "dnu_device.h"
#include "kn_hip.h"

namespace {

using namespace kn_dnu;

constexpr int kWavesPerBlock = 4;

// ---------------------------------------------------------------- H table

struct HostHTab {
    double origin = 0.0;         // ln TimeTransfer the table is anchored on
    double x0 = 0.0;             // ln a of panel 0 (origin - 20); a panel edge falls on the origin
    double width = 1.0 / 256.0;  // panel width in ln a
    int npanel = 0;
    std::vector<double> coef;  // npanel panels; the split records follow them on the device
    std::vector<unsigned char> built;
    HSplits splits;
    double *d_coef = nullptr;
    size_t d_cap = 0;
    bool dirty = false;
};
HostHTab g_h;

// ln a covered: [origin - 20, origin + 8]
void h_reset(double origin)
{
    g_h.origin = origin;
    g_h.x0 = origin - 20.0;
    g_h.npanel = (int)(28.0 / g_h.width);
    g_h.coef.assign((size_t)g_h.npanel * kHStride, 0.0);
    g_h.built.assign(g_h.npanel, 0);
    g_h.splits = HSplits();
    g_h.dirty = true;
}

// Panel p of the process table (and its split record, if it needs one).
void h_build_panel(int p)
{
    double rec[kHSplitStride];
    const double lo = g_h.x0 + p * g_h.width;
    const double e = h_panel(lo, g_h.x0 + (p + 1) * g_h.width, [](double a) { return kn_hubble(a); },
                             &g_h.coef[(size_t)p * kHStride], rec);
    if (e > kHSplitTol) g_h.splits.offer(p, e, rec);
    g_h.built[p] = 1;
}

HTab device_htab()
{
    HTab H;
    H.coef = g_h.d_coef;
    H.x0 = g_h.x0;
    H.inv_w = 1.0 / g_h.width;
    H.npanel = g_h.npanel;
    H.nsplit = g_h.splits.n;
    for (int i = 0; i < kHSplitMax; i++) H.split_p[i] = i < g_h.splits.n ? g_h.splits.p[i] : -1;
    return H;
}

// ---------------------------------------------------------------- kernels

// fslength node `node` of nfs (one wavefront).
__device__ inline void fs_node_wave(const HTab &H, int nfs, double x0, double xa, double light, double *__restrict__ fx,
                                    double *__restrict__ fy, int *__restrict__ status, int node, QagWs &ws)
{
    const int lane = threadIdx.x & 63;
    // delta_tot_table.c:582: log T0 + i (log a - log T0) / (Nfs - 1.)
    const double xi = x0 + node * (xa - x0) / (nfs - 1.);
    double v = 0.0;
    if (xi < xa) {  // fslength returns 0 when logai >= logaf (:402)
        FslIntegrand f;
        f.H = H;
        double r;
        int ni;
        const int rc = wave_qag61(f, xi, xa, 1e-6, kLimit, ws, &r, &ni);
        if (rc && lane == 0) atomicOr(status, 1);
        v = light * r;
    }
    if (lane == 0) {
        fx[node] = xi;
        fy[node] = v;
    }
}

// The fslength nodes in two passes with the result of one QAG per node.
// Pass 1 (k_fs_first, no LDS): sixteen nodes per wavefront, four lanes
// each (gk61_q4), forming the first 61-point Gauss-Kronrod estimate over
// [x_i, log a] exactly as wave_qag61's first step does, and accepting it by
// the same test (QUADPACK returns that estimate when it already meets the
// tolerance); a node that needs bisection is appended to a list.  Pass 2
// (k_fs_rest) runs the full wavefront QAG for the listed nodes.  Node values
// are those of one wave_qag61 per node, bit for bit.  The 16 lanes of a
// quarter hold 16 neighbouring nodes ((log a - log T0) / (16 Na - 1) apart,
// far below a panel's width), so their points at one abscissa lie in one or
// two H panels and a coefficient load touches a few cache lines, not one per
// lane.  `push(node)` lists a node for pass 2.
template <class Push>
__device__ inline void fs_first_q4(const HTab &H, int nfs, double x0, double xa, double light,
                                   double *__restrict__ fx, double *__restrict__ fy, int *__restrict__ status,
                                   int wave, const Push &push)
{
    if (16 * wave >= nfs) return;  // wave-uniform
    const int lane = threadIdx.x & 63;
    const int node = 16 * wave + (lane & 15);
    // delta_tot_table.c:582: log T0 + i (log a - log T0) / (Nfs - 1.)
    const double xi = node < nfs ? x0 + node * (xa - x0) / (nfs - 1.) : xa;
    FslIntegrand f;
    f.H = H;
    const GK g = gk61_q4(f, xi < xa ? xi : xa, xa);
    if (lane < 16 && node < nfs) {
        double v = 0.0;
        if (xi < xa) {  // fslength returns 0 when logai >= logaf (:402)
            const double tol = 1e-6 * fabs(g.res);
            if (g.err <= 50 * DBL_EPSILON * g.resabs && g.err > tol) {
                atomicOr(status, 1);  // roundoff on the first estimate (wave_qag61 returns 18)
                v = light * g.res;
            } else if ((g.err <= tol && g.err != g.resasc) || g.err == 0.0) {
                v = light * g.res;
            } else {
                push(node);
                v = 0.0;
            }
        }
        fx[node] = xi;
        fy[node] = v;
    }
}

struct ListPush {
    int *list, *count;
    __device__ void operator()(int node) const { list[atomicAdd(count, 1)] = node; }
};

__global__ __launch_bounds__(256) void k_fs_first(HTab H, int nfs, double x0, double xa, double light,
                                                  double *__restrict__ fx, double *__restrict__ fy,
                                                  int *__restrict__ status, int *__restrict__ list,
                                                  int *__restrict__ count)
{
    const ListPush push = {list, count};
    fs_first_q4(H, nfs, x0, xa, light, fx, fy, status, (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), push);
}

// Pass 2: full QAG for the listed nodes (a bounded loop over the list).
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_fs_rest(HTab H, int nfs, double x0, double xa, double light,
                                                                  double *__restrict__ fx, double *__restrict__ fy,
                                                                  int *__restrict__ status,
                                                                  const int *__restrict__ list,
                                                                  const int *__restrict__ count)
{
    __shared__ QagWs ws[kWavesPerBlock];
    const int wave = threadIdx.x >> 6;
    const int n = *count;
    for (int t = blockIdx.x * kWavesPerBlock + wave; t < n; t += gridDim.x * kWavesPerBlock)
        fs_node_wave(H, nfs, x0, xa, light, fx, fy, status, list[t], ws[wave]);
}

constexpr int kFsRestBlocks = 64;

inline int fs_first_blocks(int nfs) { return (nfs + 63) / 64; }  // 4 waves x 16 nodes per block

// Rows of natural-spline coefficients (row r: x[0..n), y + r*ystride,
// c + r*cstride): one wavefront per (row, window of kKeep rows).
__device__ inline void spline_item_wave(const double *__restrict__ x, const double *__restrict__ y,
                                        double *__restrict__ c, int n, long long ystride, long long cstride,
                                        long long item)
{
    const int nwin = (n + kKeep - 1) / kKeep;
    const int r = (int)(item / nwin), w = (int)(item % nwin);
    const int lo = w * kKeep, hi = lo + kKeep < n ? lo + kKeep : n;
    SplineRhs rhs;
    rhs.x = x;
    rhs.y = y + r * ystride;
    tridiag_window_wave(x, n, rhs, lo, hi, c + r * cstride);
}

__global__ __launch_bounds__(256) void k_spline(const double *__restrict__ x, const double *__restrict__ y,
                                                double *__restrict__ c, int n, int nrows, long long ystride,
                                                long long cstride)
{
    const long long item = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (item >= (long long)((n + kKeep - 1) / kKeep) * nrows) return;  // wave-uniform
    spline_item_wave(x, y, c, n, ystride, cstride, item);
}

// Blocks of k_spline / k_adj_solve for nrows rows of n knots.
__host__ __device__ inline int spline_blocks(int n, int nrows)
{
    const long long items = (long long)((n + kKeep - 1) / kKeep) * nrows;
    return (int)((items + 3) / 4);
}

// Spline factors of the knots x[0..n) shared by all rows on them (fac,
// kFacStride doubles per window; none when n < 3) and the reciprocal knot
// spacings ih[0..n-1): block b's waves factor windows 4b .. 4b+3, its threads
// the spacings 256b + t.
__device__ inline void factor_block(const double *__restrict__ x, int n, double *__restrict__ fac,
                                    double *__restrict__ ih, int b)
{
    const int w = b * 4 + (threadIdx.x >> 6);
    if (n >= 3 && w < fac_windows(n)) tri_factor_store(x, n, w, fac);  // wave-uniform
    const int i = b * 256 + threadIdx.x;
    if (i < n - 1) {
        const double h = x[i + 1] - x[i];
        ih[i] = h != 0.0 ? 1.0 / h : 0.0;
    }
}

inline int factor_blocks(int n)
{
    const int bw = (fac_windows(n) + 3) / 4, bi = (n - 1 + 255) / 256;
    return bw > bi ? bw : bi;
}

__global__ __launch_bounds__(256) void k_factor(const double *__restrict__ x, int n, double *__restrict__ fac,
                                                double *__restrict__ ih)
{
    factor_block(x, n, fac, ih, blockIdx.x);
}

// doubles of factor storage for n knots: windows, then the spacings
inline size_t fac_doubles(int n) { return (size_t)fac_windows(n) * kFacStride + (size_t)n; }

// bucket[q] = largest i <= n-2 with x[i] <= x[0] + q / inv_bw
__device__ inline void bucket_at(const double *__restrict__ x, int n, double inv_bw, int *__restrict__ bucket, int q)
{
    const double edge = x[0] + q / inv_bw;
    int lo = 0, hi = n - 1;
    while (hi > lo + 1) {
        const int mid = (lo + hi) >> 1;
        if (x[mid] > edge) hi = mid;
        else lo = mid;
    }
    bucket[q] = lo;
}

// The delta_tot rows' splines (on the stored factors of the knots sx, see
// factor_block) and the bucket index in one launch: blocks [0, b_sp) four
// row windows each (none when na <= 2), the rest the bucket.  Rows of Y and
// C are `ystride` doubles apart.
// lds: kTriLds doubles per wavefront of the block
__device__ inline void rows_prep_block(const double *__restrict__ sx, const double *__restrict__ Y,
                                       double *__restrict__ C, int na, int nk, long long ystride, int b_sp, int nsb,
                                       double sinv_bw, int *__restrict__ bucket, const double *__restrict__ fac,
                                       int b, double *lds)
{
    if (b < b_sp) {
        const int nwin = fac_windows(na);
        const long long item = (long long)b * 4 + (threadIdx.x >> 6);
        if (item < (long long)nwin * nk) {  // wave-uniform
            const int r = (int)(item / nwin), w = (int)(item % nwin);
            const int lo = w * kKeep, hi = lo + kKeep < na ? lo + kKeep : na;
            tri_window_staged<false>(sx, na, Y + r * ystride, lo, hi, C + r * ystride, fac,
                                     lds + (threadIdx.x >> 6) * kTriLds);
        }
        return;
    }
    const int q = (b - b_sp) * 256 + threadIdx.x;
    if (q <= nsb) bucket_at(sx, na, sinv_bw, bucket, q);
}

// three wavefronts per SIMD (168 registers, a 20-byte spill): 0.91 ms against
// 1.06 ms at two for the 256-history group's rows
__global__ __launch_bounds__(256, 3) void k_rows_prep(const double *__restrict__ sx, const double *__restrict__ Y,
                                                   double *__restrict__ C, int na, int nk, long long ystride,
                                                   int b_sp, int nsb, double sinv_bw, int *__restrict__ bucket,
                                                   const double *__restrict__ fac)
{
    __shared__ double lds[4 * kTriLds];
    rows_prep_block(sx, Y, C, na, nk, ystride, b_sp, nsb, sinv_bw, bucket, fac, blockIdx.x, lds);
}

inline int rows_prep_blocks(int b_sp, int nsb) { return b_sp + (nsb + 256) / 256; }

__host__ __device__ inline int rows_prep_spline_blocks(int na, int nk) { return na > 2 ? spline_blocks(na, nk) : 0; }

struct SpeciesDev {
    double mnubykT, relerr;
    JPar J;
};

// Initial-condition piece (delta_tot_table.c:522-535) added in the QAG
// kernel when init != nullptr: specialJ(k fsl0 / m, qc = 0) init[k] lin,
// m = m/kT (1 when massless); then delta_nu = IC + prefac * integral.
struct IcArgs {
    const double *init;
    double fsl0, lin, prefac;
};

__device__ __forceinline__ double ic_piece(double kk, double init, const SpeciesDev &sp, const IcArgs &ic)
{
    const double m = sp.mnubykT > 0 ? sp.mnubykT : 1;
    JPar J0;
    J0.qc = 0.0;
    J0.nufrac_low = sp.J.nufrac_low;
    return specialJ(kk * ic.fsl0 / m, J0) * init * ic.lin;
}

// What one history's QAG items read and write (rows of Y and C `ystride`
// doubles apart).
struct QagArgs {
    HTab H;
    const double *fx, *fy, *fc;
    int nfs;
    const double *sx, *Y, *C;
    const int *sb;
    int ns, nsb;
    long long ystride;
    double sinv_bw;
    const double *wavenum;
    double x0, xa;
    double *out;
    int *nint_out, *status;
    double *ivl;
    IcArgs ic;
};

// Histories of at most this many stored steps have their log a values copied
// to LDS by the QAG kernels (48 KB; a block's four wavefronts share the copy,
// which every wavefront writes whole before it reads, so no block barrier).
constexpr int kStageMax = 6144;

// Item `item` (species sp, wavenumber k) by one wavefront; sxs: LDS for the
// stored log a values (na <= kStageMax), or null to read them from memory.
// The integrand of item (species sp, wavenumber index k), with the stored
// log a values read from sx (device memory or their LDS copy).
__device__ inline DnuIntegrand dnu_integrand(const QagArgs &a, const SpeciesDev &sp, int k, const double *sx)
{
    DnuIntegrand f;
    f.H = a.H;
    f.fy = a.fy;
    f.fc = a.fc;
    f.fk.x0 = a.x0;  // the fslength knots' own expression (fs_first_q4)
    f.fk.span = a.xa - a.x0;
    f.fk.den = a.nfs - 1.;
    f.nfs = a.nfs;
    f.fx0 = a.fx[0];
    f.fxhi = a.fx[a.nfs - 1];
    f.finv_h = (a.nfs - 1) / (a.fx[a.nfs - 1] - a.fx[0]);
    f.sx = sx;
    f.sxhi = a.sx[a.ns - 1];
    f.sy = a.Y + (size_t)k * a.ystride;
    f.sc = a.C + (size_t)k * a.ystride;
    f.sbucket = a.sb;
    f.ns = a.ns;
    f.nsb = a.nsb;
    f.linear = a.ns <= 2;
    f.sx0 = a.sx[0];
    f.sinv_bw = a.sinv_bw;
    f.k = a.wavenum[k];
    f.mnubykT = sp.mnubykT;
    f.k_by_m = f.k / sp.mnubykT;
    f.J = &sp.J;
    f.edom = a.status + 1;
    return f;
}

// The item's result, interval count and failure flag; the final subdivision
// (for the adjoint) by threads t0, t0 + stride, ...
__device__ inline void dnu_qag_store(const QagArgs &a, const SpeciesDev &sp, int item, int k, const QagWs &ws,
                                     double r, int ni, int rc, int t0, int stride)
{
    if (a.ivl)
        for (int t = t0; t < ni; t += stride) {
            a.ivl[((size_t)item * kLimit + t) * 2] = ws.lo[t];
            a.ivl[((size_t)item * kLimit + t) * 2 + 1] = ws.hi[t];
        }
    if (t0 == 0) {
        a.out[item] = a.ic.init ? ic_piece(a.wavenum[k], a.ic.init[k], sp, a.ic) + a.ic.prefac * r : r;
        a.nint_out[item] = ni;
        if (rc) atomicOr(a.status, 2);
    }
}

__device__ inline void dnu_qag_item(const QagArgs &a, const SpeciesDev &sp, int item, int k, QagWs &ws,
                                    double *sxs)
{
    const int lane = threadIdx.x & 63;
    const double *sx = a.sx;
    if (sxs) {
        for (int t = lane; t < a.ns; t += 64) sxs[t] = a.sx[t];
        __builtin_amdgcn_wave_barrier();  // every lane's entries written before any lane reads
        sx = sxs;
    }
    const DnuIntegrand f = dnu_integrand(a, sp, k, sx);
    double r;
    int ni;
    const int rc = wave_qag61(f, a.x0, a.xa, sp.relerr, kLimit, ws, &r, &ni);
    dnu_qag_store(a, sp, item, k, ws, r, ni, rc, lane, 64);
}

// One item by the W wavefronts of a block (block_qag61: speculative
// bisections on the wavefronts other than the first).
template <int W>
__device__ inline void dnu_qag_item_block(const QagArgs &a, const SpeciesDev &sp, int item, int k, QagBlock<W> &S,
                                          double *sxs)
{
    const double *sx = a.sx;
    if (sxs) {
        for (int t = threadIdx.x; t < a.ns; t += 64 * W) sxs[t] = a.sx[t];
        __syncthreads();
        sx = sxs;
    }
    const DnuIntegrand f = dnu_integrand(a, sp, k, sx);
    double r;
    int ni;
    const int rc = block_qag61<W>(f, a.x0, a.xa, sp.relerr, kLimit, S, &r, &ni);
    dnu_qag_store(a, sp, item, k, S.ws, r, ni, rc, threadIdx.x, 64 * W);
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void k_dnu_qag(QagArgs a, const SpeciesDev *__restrict__ sp,
                                                                  const int *__restrict__ item_species, int nk,
                                                                  int nitems)
{
    __shared__ QagWs ws[kWavesPerBlock];
    HIP_DYNAMIC_SHARED(double, sxs)  // a.ns doubles when a.ns <= kStageMax
    const int wave = threadIdx.x >> 6;
    const int item = blockIdx.x * kWavesPerBlock + wave;
    if (item >= nitems) return;
    dnu_qag_item(a, sp[item_species[item / nk]], item, item % nk, ws[wave], a.ns <= kStageMax ? sxs : nullptr);
}

// The same items, one block of kQagSpec wavefronts each, the highest k first
// (the longest bisection chains are at high k).
#ifndef KN_QAG_SPEC
#define KN_QAG_SPEC 4
#endif
constexpr int kQagSpec = KN_QAG_SPEC;
__global__ __launch_bounds__(64 * kQagSpec) void k_dnu_qag_spec(QagArgs a, const SpeciesDev *__restrict__ sp,
                                                                 const int *__restrict__ item_species, int nk,
                                                                 int nitems)
{
    __shared__ QagBlock<kQagSpec> S;
    HIP_DYNAMIC_SHARED(double, sxs)  // a.ns doubles when a.ns <= kStageMax
    const int nsp = nitems / nk, b = (int)blockIdx.x;
    const int k = nk - 1 - b / nsp, item = (b % nsp) * nk + k;
    dnu_qag_item_block<kQagSpec>(a, sp[item_species[item / nk]], item, k, S, a.ns <= kStageMax ? sxs : nullptr);
}

// Items in a given order: the first nlong (expected to need long bisection
// chains) one block of kQagSpec wavefronts each (dnu_qag_item_block), the
// rest kQagSpec per block, one wavefront each (dnu_qag_item).  A QAG's
// results do not depend on which of the two runs it.
union QagMixLds {
    QagBlock<kQagSpec> blk;
    QagWs ws[kQagSpec];
};
__global__ __launch_bounds__(64 * kQagSpec) void k_dnu_qag_mix(QagArgs a, const SpeciesDev *__restrict__ sp,
                                                                const int *__restrict__ item_species, int nk,
                                                                int nitems, const int *__restrict__ order, int nlong)
{
    __shared__ QagMixLds U;
    HIP_DYNAMIC_SHARED(double, sxs)  // a.ns doubles when a.ns <= kStageMax
    double *st = a.ns <= kStageMax ? sxs : nullptr;
    const int b = (int)blockIdx.x;
    if (b < nlong) {  // block-uniform
        const int item = order[b];
        dnu_qag_item_block<kQagSpec>(a, sp[item_species[item / nk]], item, item % nk, U.blk, st);
        return;
    }
    const int wave = threadIdx.x >> 6, j = nlong + (b - nlong) * kQagSpec + wave;
    if (j >= nitems) return;
    const int item = order[j];
    dnu_qag_item(a, sp[item_species[item / nk]], item, item % nk, U.ws[wave], st);
}

// Interval counts at or above which an item is expected to gain from the
// speculative block (about two bisections per pass there against one).
constexpr int kLongChain = 8;

// Adjoint of the initial-condition piece at k: dL/d delta_nu_init[k] and the
// piece's d/d(m/kT) term (zero when massless).
struct IcAdjArgs {
    const double *init;
    double fsl0, lin;
    double *init_bar, *mbar_ic;
};

__device__ __forceinline__ void ic_adjoint_at(int k, const double *__restrict__ wavenum,
                                              const double *__restrict__ gbar, const SpeciesDev &sp,
                                              const IcAdjArgs &a)
{
    const double m = sp.mnubykT > 0 ? sp.mnubykT : 1;
    JPar J0;
    J0.qc = 0.0;
    J0.nufrac_low = sp.J.nufrac_low;
    const double kk = wavenum[k] * a.fsl0 / m;
    double Jx, Jq;
    const double J = specialJ_d(kk, J0, Jx, Jq);
    a.init_bar[k] = gbar[k] * J * a.lin;
    a.mbar_ic[k] = sp.mnubykT > 0 ? gbar[k] * Jx * (-kk / m) * a.init[k] * a.lin : 0.0;
}

// Adjoint of one item's integral with its recorded subdivision held fixed:
// node weights onto delta_tot[k][.] (y) and the spline coefficients (c) are
// accumulated in LDS (ds_add_f64: the row belongs to this wave alone), plus
// the derivative in m/kT.  One wavefront per block, one item per block.
// Rows of Y, C, ybar and cbar are `ystride` doubles apart; entries
// [ns, ystride) of the item's ybar row are written as zeros.
struct AdjArgs {
    HTab H;
    const double *fx, *fy, *fc;
    int nfs;
    double x0, xa;  // the integration's limits, which set the fslength knots
    const double *sx, *Y, *C;
    const int *sb;
    int ns, nsb, nk;
    long long ystride;
    double sinv_bw;
    const double *wavenum;
    const SpeciesDev *sp;
    double vcrit;
    const double *ivl;
    const int *nint;
    const double *gbar;
    double prefac;
    double *ybar, *cbar, *mu_bar;  // mu_bar: [item][part]
    int nparts, plen;              // knots split in parts of plen (see adj_parts)
    IcAdjArgs ica;
};

// An item's knots [0, ns) are split in nparts parts of plen knots, one block
// each, so that a block's LDS rows stay small (more wavefronts per CU).
constexpr int kAdjPartKnots = 1024;
__host__ __device__ inline int adj_parts(int ns) { return (ns + kAdjPartKnots - 1) / kAdjPartKnots; }
__host__ __device__ inline int adj_plen(int ns) { return (ns + adj_parts(ns) - 1) / adj_parts(ns); }
// LDS of one block: ybar and cbar rows of plen doubles, the selected interval
// list and the queue of points to evaluate (kAdjQueue)
constexpr int kAdjQueue = 128;
inline size_t adj_lds_bytes(int ns)
{
    return sizeof(double) * 2 * (size_t)adj_plen(ns) + sizeof(int) * (kLimit + kAdjQueue);
}

// Block `blk` of one history's adjoint: (item, part) blocks [0, nk nparts),
// then the initial-condition piece, 64 k per block.  The part owning knots
// [K0, K1) evaluates the points of the recorded intervals that can reach
// them (a point in knot interval si feeds knots si and si+1: x within
// [sx[K0 - 1], sx[K1]]) and keeps the terms that land in [K0, K1); a point's
// d/d(m/kT) term goes to the part owning its si, so every term is counted
// once.  The points that pass are queued in LDS and evaluated 64 at a time,
// one per lane, so a part evaluates its own share of a long interval's
// points, not all of them.  `lds`: adj_lds_bytes(ns).
__device__ inline void dnu_adj_block(const AdjArgs &A, int blk, double *lds)
{
    const int nk = A.nk, ns = A.ns, nfs = A.nfs;
    const double *__restrict__ fx = A.fx, *__restrict__ fy = A.fy, *__restrict__ fc = A.fc, *__restrict__ sx = A.sx;
    const int *__restrict__ sb = A.sb;
    const double prefac = A.prefac, vcrit = A.vcrit, sinv_bw = A.sinv_bw;
    const int nsb = A.nsb;
    const HTab &H = A.H;
    const int P = A.nparts, L = A.plen;
    if (blk >= nk * P) {
        const int kk = (blk - nk * P) * 64 + (int)(threadIdx.x & 63);
        if (kk < nk) ic_adjoint_at(kk, A.wavenum, A.gbar, *A.sp, A.ica);
        return;
    }
    const int lane = threadIdx.x & 63, item = blk / P, part = blk % P, k = item % nk;
    const int K0 = part * L, K1 = K0 + L < ns ? K0 + L : ns;
    int *sel = (int *)(lds + 2 * L);
    for (int t = lane; t < 2 * L; t += 64) lds[t] = 0.0;
    const double *__restrict__ ivl = A.ivl;
    const int ni = A.nint[item];
    // the recorded intervals that reach this part's knots, in their order:
    // x in [sx[K0 - 1], sx[K1]] (clamped to the table)
    int nsel = 0;
    const double xlo = sx[K0 > 0 ? K0 - 1 : 0], xhi = sx[K1 < ns ? K1 : ns - 1];
    {
        for (int t0 = 0; t0 < ni; t0 += 64) {
            const int t = t0 + lane;
            int hit = 0;
            if (t < ni) {
                const double a = ivl[((size_t)item * kLimit + t) * 2], b = ivl[((size_t)item * kLimit + t) * 2 + 1];
                hit = (a <= xhi && b >= xlo) ? 1 : 0;
            }
            double sc = hit;  // inclusive prefix count over the wavefront
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double o = __shfl_up(sc, d, 64);
                if (lane >= d) sc += o;
            }
            if (hit) sel[nsel + (int)sc - 1] = t;
            nsel += (int)__shfl(sc, 63, 64);
        }
    }
    __syncthreads();
    const double *syk = A.Y + (size_t)k * A.ystride, *sck = A.C + (size_t)k * A.ystride;
    const bool linear = ns <= 2;
    const double kk = A.wavenum[k], mnubykT = A.sp->mnubykT, gb = A.gbar[item];
    const JPar &J = A.sp->J;
    const double fx0 = fx[0], finv_h = (nfs - 1) / (fx[nfs - 1] - fx[0]), sx0 = sx[0];
    UniformKnots fk;  // the fslength knots (fs_first_q4's expression)
    fk.x0 = A.x0;
    fk.span = A.xa - A.x0;
    fk.den = nfs - 1.;
    double mu_acc = 0.0;
    // point g of the selected intervals: interval sel[g / 61], node jj = g % 61
    // (jj < 30: c - h xgk[jj]; 30: c; jj > 30: c + h xgk[60 - jj])
    auto point = [&](int g, double &x, double &w) {
        const int t = sel[g / 61], jj = g % 61, jn = jj <= 30 ? jj : 60 - jj;
        const double a = ivl[((size_t)item * kLimit + t) * 2], b = ivl[((size_t)item * kLimit + t) * 2 + 1];
        const double c = 0.5 * (a + b), hl = 0.5 * (b - a);
        x = jj < 30 ? c - hl * c_xgk[jn] : (jj == 30 ? c : c + hl * c_xgk[jn]);
        w = c_wgk[jn] * hl * prefac * gb;
    };
    auto eval = [&](int g) {
        double x, w;
        point(g, x, w);
        // the pieces of delta_tot_table.c:492-500
        double xl, xr;
        const int fi = settle_uniform(fk, nfs, (int)((x - fx0) * finv_h), x, xl, xr);
        const double fsl = cspline_val(xl, xr, fy[fi], fy[fi + 1], fc[fi], fc[fi + 1], x);
        int q = (int)((x - sx0) * sinv_bw);
        q = q < 0 ? 0 : (q > nsb ? nsb : q);
        const int si = settle(sx, ns, sb[q], x);
        const double kx = kk * fsl / mnubykT;
        double Jx, Jq;
        const double Jv = specialJ_d(kx, J, Jx, Jq);
        const double base = fsl * inv_ah(H, x);  // fsl / (a H), as the forward
        const double S = linear ? linear_at(sx, syk, si, x) : cspline_at(sx, syk, sck, si, x);
        const bool own0 = si >= K0 && si < K1, own1 = si + 1 >= K0 && si + 1 < K1;
        if (own0)  // d/d(m/kT), qc = vcrit m/kT
            mu_acc += w * base * S * (Jx * (-kx / mnubykT) + Jq * vcrit);
        const double wj = w * base * Jv;
        if (linear) {
            const double t1 = (x - sx[si]) / (sx[si + 1] - sx[si]);
            if (own0) atomicAdd(&lds[si - K0], wj * (1.0 - t1));
            if (own1) atomicAdd(&lds[si + 1 - K0], wj * t1);
        } else {
            double wy0, wy1, wc0, wc1;
            cspline_grad(sx, si, x, wy0, wy1, wc0, wc1);
            if (own0) {
                atomicAdd(&lds[si - K0], wj * wy0);
                atomicAdd(&lds[L + si - K0], wj * wc0);
            }
            if (own1) {
                atomicAdd(&lds[si + 1 - K0], wj * wy1);
                atomicAdd(&lds[L + si + 1 - K0], wj * wc1);
            }
        }
    };
    int *queue = sel + kLimit;  // kAdjQueue point indices
    const int npts = nsel * 61;
    int nq = 0;  // queued points (wavefront-uniform), below 64 between rounds
    for (int g0 = 0; g0 < npts; g0 += 64) {
        const int g = g0 + lane;
        bool in = false;
        if (g < npts) {
            double x, w;
            point(g, x, w);
            in = x >= xlo && x <= xhi;
        }
        const unsigned long long m = __ballot(in);
        if (in) queue[nq + __popcll(m & ((1ull << lane) - 1))] = g;
        nq += __popcll(m);
        __builtin_amdgcn_wave_barrier();  // the queue is written before any lane reads it
        if (nq >= 64) {
            eval(queue[lane]);
            const int rest = nq - 64;
            const int v = lane < rest ? queue[64 + lane] : 0;
            __builtin_amdgcn_wave_barrier();  // every lane has read before any lane moves
            if (lane < rest) queue[lane] = v;
            __builtin_amdgcn_wave_barrier();
            nq = rest;
        }
    }
    if (lane < nq) eval(queue[lane]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) mu_acc += __shfl_xor(mu_acc, m, 64);
    __syncthreads();
    double *yrow = A.ybar + (size_t)k * A.ystride, *crow = A.cbar + (size_t)k * A.ystride;
    for (int t = lane; t < K1 - K0; t += 64) {
        yrow[K0 + t] = lds[t];
        crow[K0 + t] = lds[L + t];
    }
    if (part == P - 1)
        for (int t = ns + lane; t < A.ystride; t += 64) yrow[t] = 0.0;
    if (lane == 0) A.mu_bar[(size_t)item * P + part] = mu_acc;
}

__global__ __launch_bounds__(64) void k_dnu_adj(AdjArgs A)
{
    HIP_DYNAMIC_SHARED(double, lds)  // adj_lds_bytes(ns): part rows of ybar and cbar, interval list
    dnu_adj_block(A, blockIdx.x, lds);
}

// Per row of the adjoint, M z = cbar (the natural-spline matrix is
// symmetric, so its transpose is itself) by wavefronts as k_spline, and z
// added through the right-hand side 3((y_{i+1}-y_i)/h_i - (y_i-y_{i-1})/h_{i-1})
// onto ybar in the same pass.
// Rows `stride` doubles apart, on the stored factors of x; block `b` of
// spline_blocks(n, nrows).
// lds: kTriLds doubles per wavefront of the block
__device__ inline void adj_solve_block(const double *__restrict__ x, int n, int nrows, long long stride,
                                       const double *__restrict__ cbar, double *__restrict__ ybar,
                                       const double *__restrict__ fac, int b, double *lds)
{
    const int nwin = (n + kKeep - 1) / kKeep;
    const long long item = (long long)b * 4 + (threadIdx.x >> 6);
    if (item >= (long long)nwin * nrows) return;  // wave-uniform
    const int r = (int)(item / nwin), w = (int)(item % nwin);
    const int lo = w * kKeep, hi = lo + kKeep < n ? lo + kKeep : n;
    tri_window_staged<true>(x, n, cbar + (size_t)r * stride, lo, hi, ybar + (size_t)r * stride, fac,
                            lds + (threadIdx.x >> 6) * kTriLds);
}

__global__ __launch_bounds__(256) void k_adj_solve(const double *__restrict__ x, int n, int nrows, long long stride,
                                                   const double *__restrict__ cbar, double *__restrict__ ybar,
                                                   const double *__restrict__ fac)
{
    __shared__ double lds[4 * kTriLds];
    adj_solve_block(x, n, nrows, stride, cbar, ybar, fac, blockIdx.x, lds);
}

// ---------------------------------------------------------------- state

struct DnuDev {
    size_t cap_fs = 0, cap_y = 0, cap_x = 0, cap_k = 0, cap_items = 0, cap_b = 0;
    double *fx = nullptr, *fy = nullptr, *fc = nullptr;
    int *fs_list = nullptr;  // fslength nodes that need bisection, then their count at [cap_fs]
    double *sx = nullptr, *Y = nullptr, *C = nullptr, *wavenum = nullptr;
    int *bucket = nullptr;
    double *out = nullptr;
    int *nint = nullptr, *item_species = nullptr;
    SpeciesDev *sp = nullptr;
    int *status = nullptr;
    size_t cap_ivl = 0, cap_adj = 0, cap_adj_k = 0;
    double *ivl = nullptr;                                            // recorded subdivisions
    size_t cap_fac = 0;
    double *fac = nullptr;  // spline factors of scalefact (fac_doubles(Na))
    double *gbar = nullptr, *ybar = nullptr, *cbar = nullptr, *mubar = nullptr;  // adjoint
    // the delta_tot rows held in Y (row stride namax, the host table's), and a
    // host copy of what Y holds: each call uploads a row only from its first
    // entry that differs from that copy (an appended or overwritten step is
    // one or two columns), whoever wrote the host table
    int ystride = 0, have_nk = 0, have_na = 0;
    std::vector<double> shadow;
    std::vector<int> first;
    // each item's final interval count in the last call, which orders the
    // next call's QAG items (the same k bins are integrated every step)
    std::vector<int> prev_nint;
    int *order = nullptr;
};
std::unordered_map<const void *, DnuDev *> g_dev;
int g_last_max_intervals = 0, g_last_items = 0;
long long g_last_intervals = 0;
// optional HIP-event timing of the QAG launch (k_dnu_qag) of each get_delta_nu
int g_qag_timing = 0;
hipEvent_t g_qag_ev[2];
float g_qag_ms = 0.0f;
int g_last_upload_cols = 0;

template <typename T>
void grow(T *&p, size_t &cap, size_t need)
{
    if (need <= cap) return;
    if (p) KN_CHECK(hipFree(p));
    KN_CHECK(hipMalloc(&p, sizeof(T) * need));
    cap = need;
}

bool g_consts_ready = false;

void upload_constants()
{
    if (g_consts_ready) return;
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_xgk), kn_xgk61, sizeof(kn_xgk61)));
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_wgk), kn_wgk61, sizeof(kn_wgk61)));
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_wg), kn_wg30, sizeof(kn_wg30)));
    g_consts_ready = true;
}

int blocks_for(long long waves) { return (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock); }

// ---------------------------------------------------------------- host staging
// The per-call inputs go through pinned host memory (grown as needed, kept
// for the process): host threads pack the delta_tot rows (stride namax on the
// host) to stride Na there, and one copy stream uploads them while the
// fslength kernels, which need none of them, already run.

// Threads that split a packing loop with the caller (KN_HOST_THREADS, default
// min(8, hardware threads), counting the caller); a forked child packs alone.
class HostPool {
public:
    HostPool()
    {
        int n = (int)std::thread::hardware_concurrency();
        n = n < 1 ? 1 : (n > 8 ? 8 : n);
        if (const char *e = getenv("KN_HOST_THREADS")) n = atoi(e) > 0 ? atoi(e) : 1;
        pid_ = getpid();
        for (int i = 1; i < n; i++) th_.emplace_back([this] { work(); });
    }
    // f(0) .. f(parts - 1), on the pool and the calling thread
    void run(int parts, const std::function<void(int)> &f)
    {
        // a forked child has none of the parent's threads: it packs alone
        if (th_.empty() || parts <= 1 || getpid() != pid_) {
            for (int p = 0; p < parts; p++) f(p);
            return;
        }
        std::unique_lock<std::mutex> lk(m_);
        job_ = &f;
        nparts_ = parts;
        next_ = 0;
        pending_ = parts;
        gen_++;
        cv_work_.notify_all();
        drain(lk);
        cv_done_.wait(lk, [this] { return pending_ == 0; });
        job_ = nullptr;
    }

private:
    void drain(std::unique_lock<std::mutex> &lk)
    {
        while (next_ < nparts_) {
            const int p = next_++;
            const std::function<void(int)> *f = job_;
            lk.unlock();
            (*f)(p);
            lk.lock();
            if (--pending_ == 0) cv_done_.notify_all();
        }
    }
    void work()
    {
        unsigned long seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_work_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
            drain(lk);
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_work_, cv_done_;
    const std::function<void(int)> *job_ = nullptr;
    int nparts_ = 0, next_ = 0, pending_ = 0;
    unsigned long gen_ = 0;
    bool stop_ = false;
    pid_t pid_ = 0;
};

// Never destroyed: its threads wait on a condition variable until the process
// ends (a forked child, which has none of them, must not tear it down either).
HostPool &host_pool()
{
    static HostPool *p = new HostPool();
    return *p;
}

char *g_stage = nullptr;
size_t g_stage_cap = 0;
hipStream_t g_copy = nullptr;
hipEvent_t g_copied = nullptr;

char *pinned_stage(size_t bytes)
{
    if (bytes > g_stage_cap) {
        if (g_stage) KN_CHECK(hipHostFree(g_stage));
        KN_CHECK(hipHostMalloc((void **)&g_stage, bytes, 0));
        g_stage_cap = bytes;
    }
    if (!g_copy) {
        KN_CHECK(hipStreamCreateWithFlags(&g_copy, hipStreamNonBlocking));
        KN_CHECK(hipEventCreateWithFlags(&g_copied, hipEventDisableTiming));
    }
    return g_stage;
}

inline size_t align64(size_t n) { return (n + 63) & ~(size_t)63; }

}  // namespace

// ------------------------------------------------------------ background

// The background object: parameters, the device tables and host copies of
// them, and the same numbers as a host BgDev (host pointers) and a device
// BgDev (device pointers).
struct kn_background {
    kn_background_params p;
    BgDev host, dev;
    std::vector<double> tab;  // [species][loga, rho, c2][kRhoTab]
    double *d_tab = nullptr;
};

namespace {

struct BgBuild {
    int nt;  // tables to build
    int sp[kNuSpecies];
    double lo[kNuSpecies], hi[kNuSpecies];
    double mnu[kNuSpecies], kT, conv;
    double *loga[kNuSpecies], *rho[kNuSpecies], *c2[kNuSpecies];
};

// rho_nu_init's integrals (omega_nu_single.c:141-150), one wavefront per
// table point: QAG-GK61 of q^2 eps f0 over [0, 500 kT] at relative error
// 1e-9, divided by a^4 and converted to g cm^-3.
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_bg_rho(BgBuild g, int *status)
{
    __shared__ QagWs ws[kWavesPerBlock];
    const int wave = threadIdx.x >> 6;
    const int item = blockIdx.x * kWavesPerBlock + wave;
    if (item >= g.nt * kRhoTab) return;  // a whole wavefront; no block barrier follows
    const int t = item / kRhoTab, i = item - t * kRhoTab;
    const double la = g.lo[t] + i * (g.hi[t] - g.lo[t]) / (kRhoTab - 1);
    FdEnergy f;
    f.am = g.mnu[t] * exp(la);
    f.kT = g.kT;
    double r;
    int n;
    const int rc = wave_qag61(f, 0.0, 500 * g.kT, 1e-9, KN_GSL_VAL, ws[wave], &r, &n);
    if ((threadIdx.x & 63) == 0) {
        if (rc) atomicOr(status, 1);
        g.loga[t][i] = la;
        g.rho[t][i] = r / pow(exp(la), 4) * g.conv;
    }
}

// The tables' natural cubic splines (gsl_interp_init of omega_nu_single.c:151),
// one wavefront per table (block t): the wavefront stages the table in LDS,
// lane 0 runs the elimination there, and the coefficients go out coalesced.
__global__ __launch_bounds__(64) void k_bg_spline(BgBuild g)
{
    __shared__ double xs[kRhoTab], ys[kRhoTab], cs[kRhoTab], work[3 * kRhoTab];
    const int t = blockIdx.x;
    for (int i = threadIdx.x; i < kRhoTab; i += 64) {
        xs[i] = g.loga[t][i];
        ys[i] = g.rho[t][i];
    }
    __syncthreads();
    if (threadIdx.x == 0) bg_spline_init(xs, ys, kRhoTab, cs, work);
    __syncthreads();
    for (int i = threadIdx.x; i < kRhoTab; i += 64) g.c2[t][i] = cs[i];
}

// Omega_nu(a) (which 0) or H(a) (which 1) at n points.
__global__ __launch_bounds__(256) void k_bg_eval(BgDev b, int n, const double *__restrict__ a, double *__restrict__ out,
                                                 int which)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = which ? bg_hubble(b, a[i]) : bg_omega_nu(b, a[i]);
}

// One H table to build from a background: panels of width w from x0, split
// where needed from panel split_from on (panels below it lie below the
// integrals' lower limit and are never read).
struct BgJob {
    BgDev b;
    double x0, w;
    int npanel, split_from;
    double *coef;  // npanel * kHStride, then kHSplitMax * kHSplitStride
    double *miss;  // npanel
    int *hdr;      // nsplit, then the split panels' indices
};

// One lane per panel (blockIdx.y: the table): both functions' fits on
// [x0 + p w, x0 + (p + 1) w] and their miss against H.
__global__ __launch_bounds__(64) void k_bg_panels(const BgJob *__restrict__ jobs)
{
    const BgJob &J = jobs[blockIdx.y];
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= J.npanel) return;
    const BgHub hub{J.b};
    const double lo = J.x0 + p * J.w, hi = J.x0 + (p + 1) * J.w;
    double *c = J.coef + (size_t)p * kHStride;
    h_fit(lo, hi, hub, c);
    J.miss[p] = h_fit_err(c, lo, hi, hub);
}

// One wavefront per table: the panels that miss are gathered (a strided scan,
// up to kCand of them); lane 0 keeps the kHSplitMax that miss most, offered
// in panel order as the host's HSplits::offer does; the wavefront then splits
// each kept panel (h_split_wave) into the records after the regular panels.
constexpr int kCand = 64;
__global__ __launch_bounds__(64) void k_bg_splits(const BgJob *__restrict__ jobs)
{
    __shared__ int ncand, sp_n;
    __shared__ int cand[kCand], sp_p[kHSplitMax];
    __shared__ HFit2Lds L;
    const BgJob &J = jobs[blockIdx.x];
    const int lane = threadIdx.x;
    if (lane == 0) ncand = 0;
    __syncthreads();
    for (int q = J.split_from + lane; q < J.npanel; q += 64)
        if (J.miss[q] > kHSplitTol) {
            const int i = atomicAdd(&ncand, 1);
            if (i < kCand) cand[i] = q;
        }
    __syncthreads();
    if (lane == 0) {
        const bool all = ncand <= kCand;  // else every panel is offered again, in order
        const int nc = all ? ncand : J.npanel - J.split_from;
        if (all)
            for (int i = 1; i < nc; i++)  // panel order
                for (int j = i; j > 0 && cand[j - 1] > cand[j]; j--) {
                    const int t = cand[j];
                    cand[j] = cand[j - 1];
                    cand[j - 1] = t;
                }
        int n = 0, p[kHSplitMax];
        double err[kHSplitMax];
        for (int c = 0; c < nc; c++) {
            const int q = all ? cand[c] : J.split_from + c;
            const double e = J.miss[q];
            if (!(e > kHSplitTol)) continue;
            int slot = n;
            if (n == kHSplitMax) {
                slot = 0;
                for (int i = 1; i < n; i++)
                    if (err[i] < err[slot]) slot = i;
                if (err[slot] >= e) continue;
            } else {
                n++;
            }
            p[slot] = q;
            err[slot] = e;
        }
        sp_n = n;
        J.hdr[0] = n;
        for (int i = 0; i < kHSplitMax; i++) {
            if (i < n) sp_p[i] = p[i];
            J.hdr[1 + i] = i < n ? p[i] : -1;
        }
    }
    __syncthreads();
    const BgHub hub{J.b};
    for (int s = 0; s < sp_n; s++) {
        const int q = sp_p[s];
        h_split_wave(J.x0 + q * J.w, J.x0 + (q + 1) * J.w, hub,
                     J.coef + (size_t)J.npanel * kHStride + (size_t)s * kHSplitStride, L);
    }
}

// A table to build: the background, the panels, the device buffer (npanel *
// kHStride + kHSplitMax * kHSplitStride doubles) and where its description goes.
struct BgTableReq {
    const kn_background *bg;
    double x0, w;
    int npanel, split_from;
    double *coef;
    HTab *out;
};

// All requested tables in two launches and one read-back (synchronous).
void bg_htables(const std::vector<BgTableReq> &reqs, hipStream_t st)
{
    constexpr size_t kMaxY = 65535;  // tables per launch (grid y)
    for (size_t base = 0; base < reqs.size(); base += kMaxY) {
        const size_t n = std::min(kMaxY, reqs.size() - base);
        size_t panels = 0;
        int maxp = 0;
        for (size_t i = 0; i < n; i++) {
            panels += (size_t)reqs[base + i].npanel;
            maxp = std::max(maxp, reqs[base + i].npanel);
        }
        BgJob *d_jobs;
        double *miss;
        int *hdr;
        KN_CHECK(hipMalloc(&d_jobs, sizeof(BgJob) * n));
        KN_CHECK(hipMalloc(&miss, sizeof(double) * std::max<size_t>(panels, 1)));
        KN_CHECK(hipMalloc(&hdr, sizeof(int) * (1 + kHSplitMax) * n));
        std::vector<BgJob> jobs(n);
        size_t off = 0;
        for (size_t i = 0; i < n; i++) {
            const BgTableReq &r = reqs[base + i];
            jobs[i].b = r.bg->dev;
            jobs[i].x0 = r.x0;
            jobs[i].w = r.w;
            jobs[i].npanel = r.npanel;
            jobs[i].split_from = r.split_from;
            jobs[i].coef = r.coef;
            jobs[i].miss = miss + off;
            jobs[i].hdr = hdr + (1 + kHSplitMax) * i;
            off += (size_t)r.npanel;
        }
        KN_CHECK(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(BgJob) * n, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_bg_panels, dim3((maxp + 63) / 64, (unsigned)n), dim3(64), 0, st, (const BgJob *)d_jobs);
        KN_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_bg_splits, dim3((unsigned)n), dim3(64), 0, st, (const BgJob *)d_jobs);
        KN_CHECK(hipGetLastError());
        std::vector<int> h((1 + kHSplitMax) * n);
        KN_CHECK(hipMemcpyAsync(h.data(), hdr, sizeof(int) * h.size(), hipMemcpyDeviceToHost, st));
        KN_CHECK(hipStreamSynchronize(st));
        KN_CHECK(hipFree(d_jobs));
        KN_CHECK(hipFree(miss));
        KN_CHECK(hipFree(hdr));
        for (size_t i = 0; i < n; i++) {
            const BgTableReq &r = reqs[base + i];
            HTab &H = *r.out;
            const int *hi = h.data() + (1 + kHSplitMax) * i;
            H.coef = r.coef;
            H.x0 = r.x0;
            H.inv_w = 1.0 / r.w;
            H.npanel = r.npanel;
            H.nsplit = hi[0];
            for (int k = 0; k < kHSplitMax; k++) H.split_p[k] = hi[1 + k];
        }
    }
}

const kn_background *g_bg = nullptr;  // the process's H (kn_set_background)
const kn_background *g_h_bg = nullptr;  // the background the per-call table was built from

}  // namespace

extern "C" double kn_background_hubble_ctx(double a, void *ctx)
{
    return bg_hubble(((const kn_background *)ctx)->host, a);
}

extern "C" int kn_background_create(kn_background **out, const kn_background_params *p)
{
    if (!out || !p || !(p->Hubble > 0) || !(p->HubbleParam > 0) || !(p->tcmb0 > 0) || !(p->a0 > 0) ||
        !(p->a0 <= 1)) {
        kn_set_last_error("kn_background_create: bad arguments");
        return KN_ERR_ARG;
    }
    for (int i = 0; i < NUSPECIES; i++)
        if (!(p->MNu[i] >= 0)) {
            kn_set_last_error("kn_background_create: neutrino mass %d is %g", i, p->MNu[i]);
            return KN_ERR_ARG;
        }
    upload_constants();
    kn_background *bg = new kn_background();
    bg->p = *p;
    BgDev &b = bg->host;
    memset(&b, 0, sizeof b);
    b.Omega0 = p->Omega0;
    b.OmegaLambda = p->OmegaLambda;
    b.Hubble = p->Hubble;
    b.kBtnu = kn_kbtnu(p->tcmb0);
    b.rhocrit = kn_rhocrit(p->HubbleParam);
    b.conv = kn_rho_conversion();
    b.og = kn_omegag0(p->tcmb0, b.rhocrit);
    b.radiation = p->radiation ? 1 : 0;
    kn_nu_degeneracies(p->MNu, b.deg);
    BgBuild g;
    memset(&g, 0, sizeof g);
    g.kT = b.kBtnu;
    g.conv = b.conv;
    for (int i = 0; i < kNuSpecies; i++) {
        b.mnu[i] = p->MNu[i];
        double lo, hi;
        if (b.deg[i] && kn_rho_table_range(p->a0, p->MNu[i], b.kBtnu, &lo, &hi)) {
            b.tab[i] = 1;
            g.sp[g.nt] = i;
            g.lo[g.nt] = lo;
            g.hi[g.nt] = hi;
            g.mnu[g.nt] = p->MNu[i];
            g.nt++;
        }
    }
    bg->tab.assign((size_t)kNuSpecies * 3 * kRhoTab, 0.0);
    BgDev d = b;
    if (g.nt) {
        KN_CHECK(hipMalloc(&bg->d_tab, sizeof(double) * bg->tab.size()));
        int *status;
        KN_CHECK(hipMalloc(&status, sizeof(int)));
        KN_CHECK(hipMemset(status, 0, sizeof(int)));
        for (int t = 0; t < g.nt; t++) {
            double *base = bg->d_tab + (size_t)g.sp[t] * 3 * kRhoTab;
            g.loga[t] = base;
            g.rho[t] = base + kRhoTab;
            g.c2[t] = base + 2 * kRhoTab;
        }
        hipLaunchKernelGGL(k_bg_rho, dim3(blocks_for((long long)g.nt * kRhoTab)), dim3(64 * kWavesPerBlock), 0, 0, g,
                           status);
        KN_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_bg_spline, dim3(g.nt), dim3(64), 0, 0, g);
        KN_CHECK(hipGetLastError());
        int st = 0;
        KN_CHECK(hipMemcpy(&st, status, sizeof(int), hipMemcpyDeviceToHost));
        KN_CHECK(hipMemcpy(bg->tab.data(), bg->d_tab, sizeof(double) * bg->tab.size(), hipMemcpyDeviceToHost));
        KN_CHECK(hipFree(status));
        if (st) {
            KN_CHECK(hipFree(bg->d_tab));
            delete bg;
            kn_set_last_error("kn_background_create: quadrature failure in a neutrino density table");
            return 2001;
        }
        for (int i = 0; i < kNuSpecies; i++)
            if (b.tab[i]) {
                const size_t o = (size_t)i * 3 * kRhoTab;
                b.loga[i] = bg->tab.data() + o;
                b.rho[i] = b.loga[i] + kRhoTab;
                b.c2[i] = b.loga[i] + 2 * kRhoTab;
                d.loga[i] = bg->d_tab + o;
                d.rho[i] = d.loga[i] + kRhoTab;
                d.c2[i] = d.loga[i] + 2 * kRhoTab;
            }
    }
    // Omega_nu(1) once, on the host, for both copies
    b.onu1 = bg_omega_nu(b, 1.0);
    d.onu1 = b.onu1;
    bg->dev = d;
    *out = bg;
    return 0;
}

extern "C" void kn_background_destroy(kn_background *bg)
{
    if (!bg) return;
    if (g_bg == bg) kn_set_background(nullptr);
    if (g_h_bg == bg) g_h_bg = nullptr;
    if (bg->d_tab) KN_CHECK(hipFree(bg->d_tab));
    delete bg;
}

static int bg_eval(const kn_background *bg, int n, const double *d_a, double *d_out, void *stream, int which)
{
    if (!bg || n < 0 || (n > 0 && (!d_a || !d_out))) {
        kn_set_last_error("kn_background: bad arguments");
        return KN_ERR_ARG;
    }
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_bg_eval, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, bg->dev, n, d_a, d_out,
                       which);
    KN_CHECK(hipGetLastError());
    return 0;
}

extern "C" int kn_background_omega_nu(const kn_background *bg, int n, const double *d_a, double *d_out, void *stream)
{
    return bg_eval(bg, n, d_a, d_out, stream, 0);
}

extern "C" int kn_background_hubble(const kn_background *bg, int n, const double *d_a, double *d_out, void *stream)
{
    return bg_eval(bg, n, d_a, d_out, stream, 1);
}

extern "C" int kn_background_rho_table(const kn_background *bg, int i, double *loga, double *rhonu)
{
    if (!bg || i < 0 || i >= kNuSpecies) return -KN_ERR_ARG;
    if (!bg->host.tab[i]) return 0;
    const double *t = bg->tab.data() + (size_t)i * 3 * kRhoTab;
    if (loga) memcpy(loga, t, sizeof(double) * kRhoTab);
    if (rhonu) memcpy(rhonu, t + kRhoTab, sizeof(double) * kRhoTab);
    return kRhoTab;
}

extern "C" int kn_set_background(const kn_background *bg)
{
    g_bg = bg;
    kn_set_hubble_ctx(bg ? kn_background_hubble_ctx : nullptr, (void *)bg);
    return 0;
}


// ------------------------------------------------------------ host API

// Host part: build the panels covering [xlo, xhi] (no device work).  The
// table is anchored so that a panel edge falls exactly on origin (the
// integration's lower limit ln TimeTransfer): discontinuities of the host's
// H below that limit never enter a panel that the integrals use.
extern "C" void kn_htable_build_host(double origin, double xlo, double xhi)
{
    if (g_h.coef.empty() || origin != g_h.origin) h_reset(origin);
    const int p0 = (int)floor((xlo - g_h.x0) / g_h.width), p1 = (int)floor((xhi - g_h.x0) / g_h.width);
    if (p0 < 0 || p1 >= g_h.npanel)
        kn_terminate(KN_ERR_RANGE, "scale factor range [%g, %g] outside the Hubble table\n", exp(xlo), exp(xhi));
    for (int p = p0; p <= p1; p++)
        if (!g_h.built[p]) {
            h_build_panel(p);
            g_h.dirty = true;
        }
}

extern "C" double kn_htable_eval_host(double x);

// Drop the table when the host's H no longer matches it (a new cosmology):
// 17 points spread over [xlo, xhi], each within 1e-12 of the host's H (the
// panels fit it to ~1e-14, split panels included).
static void h_validate(double xlo, double xhi)
{
    constexpr int kProbe = 17;
    for (int i = 0; i < kProbe; i++) {
        const double x = xlo + (xhi - xlo) * i / (kProbe - 1.0);
        const int p = (int)floor((x - g_h.x0) / g_h.width);
        if (p < 0 || p >= g_h.npanel || !g_h.built[p]) continue;
        const double h = kn_hubble(exp(x));
        if (fabs(kn_htable_eval_host(x) / h - 1) > 1e-12) {
            h_reset(g_h.origin);
            return;
        }
    }
}

// The whole process table from the background, on the device, with host
// copies of its panels and split records (kn_htable_eval_host reads them).
static void h_from_background(double origin)
{
    h_reset(origin);
    const size_t nreg = g_h.coef.size(), need = nreg + (size_t)kHSplitMax * kHSplitStride;
    if (g_h.d_cap < need) {
        if (g_h.d_coef) KN_CHECK(hipFree(g_h.d_coef));
        KN_CHECK(hipMalloc(&g_h.d_coef, sizeof(double) * need));
        g_h.d_cap = need;
    }
    HTab H;
    const int p_origin = (int)llround((origin - g_h.x0) / g_h.width);
    bg_htables({BgTableReq{g_bg, g_h.x0, g_h.width, g_h.npanel, p_origin, g_h.d_coef, &H}}, 0);
    KN_CHECK(hipMemcpy(g_h.coef.data(), g_h.d_coef, sizeof(double) * nreg, hipMemcpyDeviceToHost));
    g_h.splits.n = H.nsplit;
    for (int i = 0; i < H.nsplit; i++) {
        g_h.splits.p[i] = H.split_p[i];
        g_h.splits.err[i] = 0.0;
        KN_CHECK(hipMemcpy(g_h.splits.rec[i], g_h.d_coef + nreg + (size_t)i * kHSplitStride,
                           sizeof(double) * kHSplitStride, hipMemcpyDeviceToHost));
    }
    g_h.built.assign(g_h.npanel, 1);
    g_h.dirty = false;
    g_h_bg = g_bg;
}

extern "C" void kn_htable_ensure(double origin, double xlo, double xhi)
{
    if (g_bg) {  // the H of a background: built whole on the device, no host sampling
        if (g_h_bg != g_bg || g_h.coef.empty() || origin != g_h.origin) h_from_background(origin);
        const int p0 = (int)floor((xlo - g_h.x0) / g_h.width), p1 = (int)floor((xhi - g_h.x0) / g_h.width);
        if (p0 < 0 || p1 >= g_h.npanel)
            kn_terminate(KN_ERR_RANGE, "scale factor range [%g, %g] outside the Hubble table\n", exp(xlo), exp(xhi));
        return;
    }
    if (g_h_bg) {  // built from a background that is no longer the process's H
        g_h.coef.clear();
        g_h_bg = nullptr;
    }
    if (!g_h.coef.empty() && origin == g_h.origin) h_validate(xlo, xhi);
    kn_htable_build_host(origin, xlo, xhi);
    const size_t nreg = g_h.coef.size(), need = nreg + (size_t)kHSplitMax * kHSplitStride;
    if (g_h.d_cap < need) {
        if (g_h.d_coef) KN_CHECK(hipFree(g_h.d_coef));
        KN_CHECK(hipMalloc(&g_h.d_coef, sizeof(double) * need));
        g_h.d_cap = need;
        g_h.dirty = true;
    }
    if (g_h.dirty) {
        KN_CHECK(hipMemcpy(g_h.d_coef, g_h.coef.data(), sizeof(double) * nreg, hipMemcpyHostToDevice));
        if (g_h.splits.n)
            KN_CHECK(hipMemcpy(g_h.d_coef + nreg, g_h.splits.rec, sizeof(double) * kHSplitStride * g_h.splits.n,
                               hipMemcpyHostToDevice));
        g_h.dirty = false;
    }
}

// H(e^x) from the table's 1/(a H) panels
extern "C" double kn_htable_eval_host(double x)
{
    const double u = (x - g_h.x0) / g_h.width;
    int p = (int)floor(u);
    p = p < 0 ? 0 : (p >= g_h.npanel ? g_h.npanel - 1 : p);
    for (int s = 0; s < g_h.splits.n; s++)
        if (p == g_h.splits.p[s]) return exp(-x) / h_split_eval((const double *)g_h.splits.rec[s], x, 1);
    return exp(-x) / h_clenshaw(g_h.coef.data() + (size_t)p * kHStride + kHnc, 2.0 * (u - p) - 1.0);
}

extern "C" int kn_dnu_last_max_intervals(void) { return g_last_max_intervals; }

// HIP-event timing of get_delta_nu's QAG launch on (1) or off (0); the last
// call's time in ms (0 when off).
extern "C" int kn_dnu_qag_timing(int on)
{
    if (on && !g_qag_timing) {
        KN_CHECK(hipEventCreate(&g_qag_ev[0]));
        KN_CHECK(hipEventCreate(&g_qag_ev[1]));
    } else if (!on && g_qag_timing) {
        KN_CHECK(hipEventDestroy(g_qag_ev[0]));
        KN_CHECK(hipEventDestroy(g_qag_ev[1]));
    }
    g_qag_timing = on ? 1 : 0;
    g_qag_ms = 0.0f;
    return 0;
}

extern "C" double kn_dnu_last_qag_ms(void) { return g_qag_timing ? (double)g_qag_ms : 0.0; }

// QAG items of the last get_delta_nu and their final subinterval count: an
// item ending with n intervals evaluated the 61-point rule 2n - 1 times.
extern "C" long long kn_dnu_last_intervals(int *items)
{
    if (items) *items = g_last_items;
    return g_last_intervals;
}

// Columns of the delta_tot history the last get_delta_nu uploaded (the rest
// was already on the device).
extern "C" int kn_dnu_last_upload_columns(void) { return g_last_upload_cols; }

extern "C" void kn_dnu_release(const _delta_tot_table *d)
{
    auto it = g_dev.find(d);
    if (it == g_dev.end()) return;
    DnuDev *v = it->second;
    void *ptrs[] = {v->fs_list, v->fx,  v->fy,   v->fc,   v->sx,   v->Y,    v->C,    v->wavenum, v->bucket, v->out,
                    v->nint, v->item_species, v->sp, v->status, v->ivl, v->gbar, v->ybar, v->cbar,
                    v->mubar, v->fac, v->order};
    for (void *p : ptrs)
        if (p) KN_CHECK(hipFree(p));
    delete v;
    g_dev.erase(it);
}

// Returns 0, or 2001 when the QAG or the interpolation fails: fatal
// (terminate 2001, as the reference's GSL error handler) unless report is
// set, when the code is returned with the message in kn_last_error.
static int dnu_solve(const _delta_tot_table *d, double a, const double *wavenum, int nspecies,
                     const kn_species *sp, double *out, bool record, bool report = false);

extern "C" void kn_dnu_solve(const _delta_tot_table *d, double a, const double *wavenum, int nspecies,
                             const kn_species *sp, double *out)
{
    dnu_solve(d, a, wavenum, nspecies, sp, out, false);
}

static int dnu_solve(const _delta_tot_table *d, double a, const double *wavenum, int nspecies,
                     const kn_species *sp, double *out, bool record, bool report)
{
    const int Na = d->ia, nk = d->nk;
    const double x0 = log(d->TimeTransfer), xa = log(a);
    const double nfl0 = d->omnu->hybnu.nufrac_low[0];
    // initial-condition piece, on the host exactly as delta_tot_table.c:522-535
    const double fsl_A0a = fslength(x0, xa, d->light);
    const double deriv_prefac = kn_deriv_prefac(d);
    for (int s = 0; s < nspecies; s++)
        for (int k = 0; k < nk; k++) {
            const double m = sp[s].mnubykT > 0 ? sp[s].mnubykT : 1;
            const double specJ = specialJ(wavenum[k] * fsl_A0a / m, 0, nfl0);
            out[(size_t)s * nk + k] = specJ * d->delta_nu_init[k] * (1. + deriv_prefac * fsl_A0a);
        }
    std::vector<int> need;
    for (int s = 0; s < nspecies; s++)
        if (!sp[s].ic_only) need.push_back(s);
    if (need.empty()) return 0;
    if (Na < 2) return 0;

    upload_constants();
    kn_htable_ensure(x0, x0, xa);
    DnuDev *&dev = g_dev[d];
    if (!dev) dev = new DnuDev();
    const int nfs = 16 * Na;
    const int nsb = 4 * Na;
    const int nitems = (int)need.size() * nk;
    if (dev->cap_fs < (size_t)nfs) {  // fx, fy, fc, fs_list
        for (double **p : {&dev->fx, &dev->fy, &dev->fc}) {
            if (*p) KN_CHECK(hipFree(*p));
            KN_CHECK(hipMalloc(p, sizeof(double) * nfs));
        }
        if (dev->fs_list) KN_CHECK(hipFree(dev->fs_list));
        KN_CHECK(hipMalloc(&dev->fs_list, sizeof(int) * ((size_t)nfs + 1)));
        dev->cap_fs = nfs;
    }
    grow(dev->sx, dev->cap_x, (size_t)Na);
    const int namax = d->namax;
    if (dev->ystride != namax || dev->have_nk != nk) {  // a new layout: nothing of the history is on the device
        dev->ystride = namax;
        dev->have_nk = nk;
        dev->have_na = 0;
        dev->shadow.assign((size_t)nk * namax, 0.0);
        dev->first.assign(nk, 0);
    }
    if (dev->cap_y < (size_t)nk * namax) {  // Y, C (their contents follow the shadow: start again)
        for (double **p : {&dev->Y, &dev->C}) {
            if (*p) KN_CHECK(hipFree(*p));
            KN_CHECK(hipMalloc(p, sizeof(double) * (size_t)nk * namax));
        }
        dev->cap_y = (size_t)nk * namax;
        dev->have_na = 0;
    }
    grow(dev->wavenum, dev->cap_k, (size_t)nk);
    grow(dev->bucket, dev->cap_b, (size_t)nsb + 1);
    grow(dev->fac, dev->cap_fac, fac_doubles(Na));
    if (dev->cap_items < (size_t)nitems) {
        if (dev->out) KN_CHECK(hipFree(dev->out));
        if (dev->nint) KN_CHECK(hipFree(dev->nint));
        KN_CHECK(hipMalloc(&dev->out, sizeof(double) * nitems));
        KN_CHECK(hipMalloc(&dev->nint, sizeof(int) * nitems));
        if (dev->order) KN_CHECK(hipFree(dev->order));
        KN_CHECK(hipMalloc(&dev->order, sizeof(int) * nitems));
        dev->cap_items = nitems;
    }
    if (!dev->sp) {
        KN_CHECK(hipMalloc(&dev->sp, sizeof(SpeciesDev) * NUSPECIES));
        KN_CHECK(hipMalloc(&dev->item_species, sizeof(int) * NUSPECIES));
        KN_CHECK(hipMalloc(&dev->status, sizeof(int) * 2));
    }
    if (record && dev->cap_ivl < (size_t)nitems) {
        if (dev->ivl) KN_CHECK(hipFree(dev->ivl));
        KN_CHECK(hipMalloc(&dev->ivl, sizeof(double) * (size_t)nitems * kLimit * 2));
        dev->cap_ivl = nitems;
    }
    // the fslength nodes and their spline need only (T0, a): they start
    // first, and run while the host finds and packs the missing columns
    const HTab H = device_htab();
    KN_CHECK(hipMemsetAsync(dev->status, 0, sizeof(int) * 2, 0));
    int *fs_count = dev->fs_list + dev->cap_fs;
    KN_CHECK(hipMemsetAsync(fs_count, 0, sizeof(int), 0));
    hipLaunchKernelGGL(k_fs_first, dim3(fs_first_blocks(nfs)), dim3(256), 0, 0, H, nfs, x0, xa, d->light, dev->fx,
                       dev->fy, dev->status, dev->fs_list, fs_count);
    hipLaunchKernelGGL(k_fs_rest, dim3(kFsRestBlocks), dim3(64 * kWavesPerBlock), 0, 0, H, nfs, x0, xa, d->light,
                       dev->fx, dev->fy, dev->status, dev->fs_list, fs_count);
    hipLaunchKernelGGL(k_spline, dim3(spline_blocks(nfs, 1)), dim3(256), 0, 0, dev->fx, dev->fy, dev->fc, nfs, 1, 0LL,
                       0LL);
    KN_CHECK(hipGetLastError());
    // which columns of the history the device lacks: per row, the first entry
    // that differs from the shadow (entries past the shadow's count differ)
    {
        const int have = dev->have_na < Na ? dev->have_na : Na;
        const int parts = (size_t)nk * Na < (1u << 16) ? 1 : 8;
        host_pool().run(parts, [&](int p) {
            const int k0 = (int)((long long)nk * p / parts), k1 = (int)((long long)nk * (p + 1) / parts);
            for (int k = k0; k < k1; k++) {
                const double *src = d->delta_tot[k];
                double *sh = &dev->shadow[(size_t)k * namax];
                int f = 0;
                while (f < have) {  // 32-entry blocks by memcmp, then the entry within the first differing block
                    const int n = have - f < 32 ? have - f : 32;
                    if (memcmp(src + f, sh + f, sizeof(double) * n)) {
                        while (!memcmp(src + f, sh + f, sizeof(double))) f++;
                        break;
                    }
                    f += n;
                }
                memcpy(sh + f, src + f, sizeof(double) * (Na - f));
                dev->first[k] = f;
            }
        });
    }
    int col0 = Na;
    for (int k = 0; k < nk; k++) col0 = dev->first[k] < col0 ? dev->first[k] : col0;
    dev->have_na = Na;
    const int ncol = Na - col0;  // columns [col0, Na) of every row go up
    g_last_upload_cols = ncol;
    // pinned staging: inputs [Y columns col0..Na | scalefact | wavenum | species | species index], then outputs
    const size_t nY = (size_t)nk * ncol;
    const size_t b_sp = sizeof(double) * (nY + Na + nk);
    const size_t b_in = align64(b_sp + sizeof(SpeciesDev) * NUSPECIES + sizeof(int) * NUSPECIES + sizeof(int) * nitems);
    char *stage = pinned_stage(b_in + align64(sizeof(double) * nitems) + sizeof(int) * ((size_t)nitems + 2));
    double *stY = (double *)stage, *stX = stY + nY, *stK = stX + Na;
    SpeciesDev *sd = (SpeciesDev *)(stage + b_sp);
    int *ispec = (int *)(sd + NUSPECIES);
    int *ord = ispec + NUSPECIES;
    double *res = (double *)(stage + b_in);
    int *nint = (int *)(stage + b_in + align64(sizeof(double) * nitems)), *status = nint + nitems;

    // meanwhile the missing columns are packed (ncol per row) and uploaded on
    // the copy stream into rows of stride namax, in chunks of rows: a chunk's
    // upload runs while the next is packed
    if (ncol > 0) {
        const int chunks = nY < (1u << 16) ? 1 : 4, parts = nY < (1u << 16) ? 1 : 8;
        for (int c = 0; c < chunks; c++) {
            const int c0 = (int)((long long)nk * c / chunks), c1 = (int)((long long)nk * (c + 1) / chunks);
            host_pool().run(parts, [&](int p) {
                const int k0 = c0 + (int)((long long)(c1 - c0) * p / parts);
                const int k1 = c0 + (int)((long long)(c1 - c0) * (p + 1) / parts);
                for (int k = k0; k < k1; k++)
                    memcpy(stY + (size_t)k * ncol, &dev->shadow[(size_t)k * namax + col0], sizeof(double) * ncol);
            });
            if (c1 > c0)
                KN_CHECK(hipMemcpy2DAsync(dev->Y + (size_t)c0 * namax + col0, sizeof(double) * namax,
                                          stY + (size_t)c0 * ncol, sizeof(double) * ncol, sizeof(double) * ncol,
                                          c1 - c0, hipMemcpyHostToDevice, g_copy));
        }
    }
    memcpy(stX, d->scalefact, sizeof(double) * Na);
    memcpy(stK, wavenum, sizeof(double) * nk);
    for (size_t i = 0; i < need.size(); i++) {
        const kn_species &q = sp[need[i]];
        sd[i].mnubykT = q.mnubykT;
        sd[i].relerr = q.relerr;
        sd[i].J.qc = q.qc;
        sd[i].J.nufrac_low = nfl0;
        sd[i].J.en[0] = 1.0;
        for (int n = 1; n < 20; n++) sd[i].J.en[n] = exp(-n * q.qc);
        ispec[i] = (int)i;
    }
    KN_CHECK(hipMemcpyAsync(dev->sx, stX, sizeof(double) * Na, hipMemcpyHostToDevice, g_copy));
    KN_CHECK(hipMemcpyAsync(dev->wavenum, stK, sizeof(double) * nk, hipMemcpyHostToDevice, g_copy));
    KN_CHECK(hipMemcpyAsync(dev->sp, sd, sizeof(SpeciesDev) * need.size(), hipMemcpyHostToDevice, g_copy));
    KN_CHECK(hipMemcpyAsync(dev->item_species, ispec, sizeof(int) * need.size(), hipMemcpyHostToDevice, g_copy));
    // the QAG items' order: those that needed kLongChain intervals or more in
    // the last call first (longest first), then the rest in item order
    int nlong = 0;
    {
        const bool pred = dev->prev_nint.size() == (size_t)nitems;
        for (int i = 0; i < nitems; i++)
            if (pred && dev->prev_nint[i] >= kLongChain) ord[nlong++] = i;
        std::stable_sort(ord, ord + nlong, [&](int x, int y) { return dev->prev_nint[x] > dev->prev_nint[y]; });
        int j = nlong;
        for (int i = 0; i < nitems; i++)
            if (!(pred && dev->prev_nint[i] >= kLongChain)) ord[j++] = i;
    }
    KN_CHECK(hipMemcpyAsync(dev->order, ord, sizeof(int) * nitems, hipMemcpyHostToDevice, g_copy));
    KN_CHECK(hipEventRecord(g_copied, g_copy));
    KN_CHECK(hipStreamWaitEvent(0, g_copied, 0));

    const double sinv_bw = nsb / (d->scalefact[Na - 1] - d->scalefact[0]);
    {
        const int bsp = rows_prep_spline_blocks(Na, nk);
        hipLaunchKernelGGL(k_factor, dim3(factor_blocks(Na)), dim3(256), 0, 0, dev->sx, Na, dev->fac,
                           dev->fac + (size_t)fac_windows(Na) * kFacStride);
        hipLaunchKernelGGL(k_rows_prep, dim3(rows_prep_blocks(bsp, nsb)), dim3(256), 0, 0, dev->sx, dev->Y, dev->C, Na,
                           nk, (long long)namax, bsp, nsb, sinv_bw, dev->bucket, dev->fac);
    }
    QagArgs qa;
    qa.H = H;
    qa.fx = dev->fx;
    qa.fy = dev->fy;
    qa.fc = dev->fc;
    qa.nfs = nfs;
    qa.sx = dev->sx;
    qa.Y = dev->Y;
    qa.C = dev->C;
    qa.sb = dev->bucket;
    qa.ns = Na;
    qa.nsb = nsb;
    qa.ystride = namax;
    qa.sinv_bw = sinv_bw;
    qa.wavenum = dev->wavenum;
    qa.x0 = x0;
    qa.xa = xa;
    qa.out = dev->out;
    qa.nint_out = dev->nint;
    qa.status = dev->status;
    qa.ivl = record ? dev->ivl : nullptr;
    qa.ic.init = nullptr;
    qa.ic.fsl0 = qa.ic.lin = qa.ic.prefac = 0.0;
    if (g_qag_timing) KN_CHECK(hipEventRecord(g_qag_ev[0], 0));
    hipLaunchKernelGGL(k_dnu_qag_mix, dim3(nlong + (nitems - nlong + kQagSpec - 1) / kQagSpec), dim3(64 * kQagSpec),
                       Na <= kStageMax ? sizeof(double) * Na : 0, 0, qa, dev->sp, dev->item_species, nk, nitems,
                       dev->order, nlong);
    KN_CHECK(hipGetLastError());
    if (g_qag_timing) KN_CHECK(hipEventRecord(g_qag_ev[1], 0));
    KN_CHECK(hipMemcpyAsync(res, dev->out, sizeof(double) * nitems, hipMemcpyDeviceToHost, 0));
    KN_CHECK(hipMemcpyAsync(nint, dev->nint, sizeof(int) * nitems, hipMemcpyDeviceToHost, 0));
    KN_CHECK(hipMemcpyAsync(status, dev->status, sizeof(int) * 2, hipMemcpyDeviceToHost, 0));
    KN_CHECK(hipStreamSynchronize(0));
    if (status[1] || status[0]) {
        if (status[1])
            kn_set_last_error("GSL_ERROR-equivalent: interpolation outside the stored table (GSL_EDOM)");
        else
            kn_set_last_error("GSL_ERROR-equivalent: QAG failed in get_delta_nu (status %d)", status[0]);
        if (report) return 2001;
        kn_terminate(2001, "%s\n", kn_last_error());
    }
    if (g_qag_timing) KN_CHECK(hipEventElapsedTime(&g_qag_ms, g_qag_ev[0], g_qag_ev[1]));
    g_last_max_intervals = 0;
    g_last_intervals = 0;
    g_last_items = nitems;
    dev->prev_nint.assign(nint, nint + nitems);
    for (int i = 0; i < nitems; i++) {
        g_last_max_intervals = nint[i] > g_last_max_intervals ? nint[i] : g_last_max_intervals;
        g_last_intervals += nint[i];
    }
    for (size_t i = 0; i < need.size(); i++)
        for (int k = 0; k < nk; k++)
            out[(size_t)need[i] * nk + k] += d->delta_nu_prefac * res[i * nk + k];
    return 0;
}

// d specialJ / dx of the rational fit (delta_tot_table.c:417-428), the branch
// the initial-condition piece always takes (qc = 0).
static double fit_dJdx(double x)
{
    if (x <= 0.) return 0.0;
    const double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x7 = x4 * x3, x8 = x4 * x4;
    const double p = exp(4.1811 * log(x));
    const double N = 1. + 0.0168 * x2 + 0.0407 * x4;
    const double D = 1. + 2.1734 * x2 + 1.6787 * p + 0.1467 * x8;
    const double Nx = 0.0336 * x + 0.1628 * x3;
    const double Dx = 4.3468 * x + 1.6787 * 4.1811 * p / x + 1.1736 * x7;
    return (Nx * D - N * Dx) / (D * D);
}

// Adjoint of get_delta_nu for one mass (delta_tot_table.c:507-611): the
// forward above with its final QAG subdivisions recorded, then k_dnu_adj and
// the two row passes on the device; the initial-condition piece on the host.
extern "C" int kn_get_delta_nu_adjoint(const _delta_tot_table *d, double a, const double *wavenum, double mnu,
                                       const double *gbar, double *delta_nu, double *dtot_bar, double *init_bar,
                                       double *mnu_bar)
{
    if (!d || !wavenum || !gbar || d->ia < 1 || d->nk < 1) return KN_ERR_ARG;
    kn_species s;
    kn_species_for(d, a, mnu, &s);
    const int nk = d->nk, Na = d->ia;
    std::vector<double> out(nk);
    const int rc = dnu_solve(d, a, wavenum, 1, &s, out.data(), true, true);
    if (rc) return rc;
    if (delta_nu) memcpy(delta_nu, out.data(), sizeof(double) * nk);
    // initial-condition piece (delta_tot_table.c:522-535)
    const double fsl_A0a = fslength(log(d->TimeTransfer), log(a), d->light);
    const double lin = 1. + kn_deriv_prefac(d) * fsl_A0a;
    const double m = s.mnubykT > 0 ? s.mnubykT : 1;
    const double nfl0 = d->omnu->hybnu.nufrac_low[0];
    double mb = 0.0;
    for (int k = 0; k < nk; k++) {
        const double kk = wavenum[k] * fsl_A0a / m;
        if (init_bar) init_bar[k] = gbar[k] * specialJ(kk, 0, nfl0) * lin;
        if (s.mnubykT > 0) mb += gbar[k] * fit_dJdx(kk) * (-kk / m) * d->delta_nu_init[k] * lin;
    }
    if (dtot_bar && (s.ic_only || Na < 2)) memset(dtot_bar, 0, sizeof(double) * (size_t)nk * Na);
    if (!s.ic_only && Na >= 2) {
        DnuDev *dev = g_dev[d];
        const int namax = dev->ystride;  // rows of Y, C, ybar and cbar are namax doubles apart
        const size_t ny = (size_t)nk * Na, nyd = (size_t)nk * namax;
        if (dev->cap_adj < nyd) {
            for (double **p : {&dev->ybar, &dev->cbar}) {
                if (*p) KN_CHECK(hipFree(*p));
                KN_CHECK(hipMalloc(p, sizeof(double) * nyd));
            }
            dev->cap_adj = nyd;
        }
        const int P = adj_parts(Na);
        if (dev->cap_adj_k < (size_t)nk * P) {  // gbar (nk), mubar (nk x parts)
            for (double **p : {&dev->gbar, &dev->mubar}) {
                if (*p) KN_CHECK(hipFree(*p));
                KN_CHECK(hipMalloc(p, sizeof(double) * nk * P));
            }
            dev->cap_adj_k = (size_t)nk * P;
        }
        KN_CHECK(hipMemcpy(dev->gbar, gbar, sizeof(double) * nk, hipMemcpyHostToDevice));
        const size_t lds = adj_lds_bytes(Na);
        const int nfs = 16 * Na, nsb = 4 * Na;
        const double sinv_bw = nsb / (d->scalefact[Na - 1] - d->scalefact[0]);
        const double vcrit = s.qc > 0 ? d->omnu->hybnu.vcrit : 0.0;
        AdjArgs A;
        A.H = device_htab();
        A.fx = dev->fx;
        A.fy = dev->fy;
        A.fc = dev->fc;
        A.nfs = nfs;
        A.x0 = log(d->TimeTransfer);
        A.xa = log(a);
        A.sx = dev->sx;
        A.Y = dev->Y;
        A.C = dev->C;
        A.sb = dev->bucket;
        A.ns = Na;
        A.nsb = nsb;
        A.nk = nk;
        A.ystride = namax;
        A.sinv_bw = sinv_bw;
        A.wavenum = dev->wavenum;
        A.sp = dev->sp;
        A.vcrit = vcrit;
        A.ivl = dev->ivl;
        A.nint = dev->nint;
        A.gbar = dev->gbar;
        A.prefac = d->delta_nu_prefac;
        A.ybar = dev->ybar;
        A.cbar = dev->cbar;
        A.mu_bar = dev->mubar;
        A.nparts = P;
        A.plen = adj_plen(Na);
        A.ica = IcAdjArgs{nullptr, 0.0, 0.0, nullptr, nullptr};
        hipLaunchKernelGGL(k_dnu_adj, dim3(nk * P), dim3(64), lds, 0, A);
        KN_CHECK(hipGetLastError());
        if (Na > 2) {
            hipLaunchKernelGGL(k_adj_solve, dim3(spline_blocks(Na, nk)), dim3(256), 0, 0, dev->sx, Na, nk,
                               (long long)namax, dev->cbar, dev->ybar, dev->fac);
            KN_CHECK(hipGetLastError());
        }
        // results through the pinned stage: [mu_bar (nk x parts) | dL/d delta_tot (nk Na)]
        const size_t nmu = (size_t)nk * P;
        double *mub = (double *)pinned_stage(sizeof(double) * (nmu + (dtot_bar ? ny : 0))), *yb = mub + nmu;
        KN_CHECK(hipMemcpyAsync(mub, dev->mubar, sizeof(double) * nmu, hipMemcpyDeviceToHost, 0));
        // dL/d delta_tot comes back in chunks: the host copies one out while the next is in flight
        static hipEvent_t done[4];
        const int chunks = dtot_bar ? (ny < (1u << 16) ? 1 : 4) : 0;
        for (int c = 0; c < chunks; c++) {  // chunks of rows, each row's [0, Na) to stride Na
            if (!done[c]) KN_CHECK(hipEventCreateWithFlags(&done[c], hipEventDisableTiming));
            const int r0 = (int)((long long)nk * c / chunks), r1 = (int)((long long)nk * (c + 1) / chunks);
            if (r1 > r0)
                KN_CHECK(hipMemcpy2DAsync(yb + (size_t)r0 * Na, sizeof(double) * Na, dev->ybar + (size_t)r0 * namax,
                                          sizeof(double) * namax, sizeof(double) * Na, r1 - r0, hipMemcpyDeviceToHost,
                                          0));
            KN_CHECK(hipEventRecord(done[c], 0));
        }
        KN_CHECK(chunks ? hipEventSynchronize(done[0]) : hipStreamSynchronize(0));  // mu_bar went first
        for (size_t i = 0; i < nmu; i++) mb += mub[i];  // k-major, parts in order
        for (int c = 0; c < chunks; c++) {
            KN_CHECK(hipEventSynchronize(done[c]));
            const size_t i0 = (size_t)((long long)nk * c / chunks) * Na, i1 = (size_t)((long long)nk * (c + 1) / chunks) * Na;
            const int parts = chunks > 1 ? 8 : 1;
            host_pool().run(parts, [&](int p) {
                const size_t j0 = i0 + (i1 - i0) * p / parts, j1 = i0 + (i1 - i0) * (p + 1) / parts;
                memcpy(dtot_bar + j0, yb + j0, sizeof(double) * (j1 - j0));
            });
        }
    }
    if (mnu_bar) *mnu_bar = mb / d->omnu->kBtnu;
    return 0;
}

// ================================================================ groups
//
// Many get_delta_nu integrals (one kn_batch_history each: its own H(a), mass,
// stored steps and k bins) solved with the per-call kernels' device code, in
// one of two ways: merged launches (default), one grid per phase covering
// every history (a block finds its history in the phase's block prefix), or
// each history's per-call launch sequence on one of several HIP streams
// forked from the caller's stream (longest histories first, each to the
// least-loaded stream).  Histories live in device memory in the caller's
// layout (delta_tot rows of stride na_max, which the kernels read and write
// in place).  Every solve records the final QAG subdivisions, which the
// adjoint then holds fixed.

namespace {

constexpr int kGroupStreams = 8;
constexpr double kGroupHWidth = 1.0 / 256.0;  // H panel width (the per-call table's)

// delta_nu[k] = IC(k) of an IC-only history (delta_tot_table.c:522-535,
// :543 and :551: hybrid with all neutrinos in particles, massless, or one
// stored step).
__global__ void k_group_ic(const double *__restrict__ wavenum, const double *__restrict__ init,
                           const SpeciesDev *__restrict__ sp, int nk, IcArgs ic, double *__restrict__ out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    out[k] = ic_piece(wavenum[k], init[k], *sp, ic);
}

// Adjoint of the initial-condition piece alone (IC-only histories).
__global__ void k_group_ic_adj(const double *__restrict__ wavenum, const double *__restrict__ gbar,
                               const SpeciesDev *__restrict__ sp, int nk, IcAdjArgs ica)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    ic_adjoint_at(k, wavenum, gbar, *sp, ica);
}

struct GroupMeta {
    long long off_k, off_mu;
    int nk, integ, nparts;
    double kBtnu;
};

// dL/d mnu per history, one wavefront each: lane l sums k = l, l + 64, ...
// (the IC term, then the integral's), then a fixed butterfly (deterministic).
__global__ __launch_bounds__(64) void k_group_mass_bar(const GroupMeta *__restrict__ meta,
                                                       const double *__restrict__ mbar_ic,
                                                       const double *__restrict__ mubar, double *__restrict__ out)
{
    const int h = blockIdx.x;
    const GroupMeta g = meta[h];
    double s = 0.0;
    for (int k = threadIdx.x; k < g.nk; k += 64) {
        s += mbar_ic[g.off_k + k];
        if (g.integ)
            for (int p = 0; p < g.nparts; p++) s += mubar[g.off_mu + (long long)k * g.nparts + p];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (threadIdx.x == 0) out[h] = s / g.kBtnu;
}

__global__ void k_group_zero(double *__restrict__ p, long long n)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

struct HubCtx {
    double (*fn)(double, void *);
    void *ctx;
};

double fsl_integrand_ctx(double x, void *p)
{
    const HubCtx *h = (const HubCtx *)p;
    const double a = exp(x);
    return 1. / a / (a * h->fn(a, h->ctx));
}

struct GroupHist {
    kn_batch_history spec;
    int na, integ, htab, stream;
    double x0, xa, xmax, fsl0, lin, vcrit, deriv_prefac;
    long long off_k, off_y, off_s, off_fs, off_b, off_fac, off_mu;
};

// ---- merged launches: one grid per phase for every history of the group.
// Block b of a phase belongs to history h with pre[h] <= b < pre[h + 1]
// (pre: the phase's exclusive prefix of per-history block counts) and is
// that history's block b - pre[h] of the per-call kernel's grid, running the
// same device code on the history's arrays (descriptor below).

// Per-history arrays and scalars (device copy; rows of Y, C, dbar and cbar
// are na_max doubles apart).
struct GroupDesc {
    HTab H;
    double *sx, *Y;
    const double *wavenum, *init, *gbar;
    double *C, *fx, *fy, *fc, *out, *ivl, *dbar, *cbar, *mubar, *ibar, *mbic, *fac;
    int *bucket, *nint, *status;
    const SpeciesDev *sp;
    // the stored steps and what they set: live in device memory (the device
    // append updates them), read by every launch
    int nk, na, na_max, nfs, nsb, bsp, integ, ic_only;
    double x0, xa, a, light, sinv_bw, fsl0, lin, prefac, vcrit;
    // constants of the history that the device append derives from
    int hybrid;
    double xmax, nu_crit_time, nufrac_low, vcrit_spec, mnu, kBtnu, deriv_prefac;
};

// phases (rows of the prefix table); the adjoint's (item, part) blocks are
// split in classes of part length (LDS rows of 2 adj_plen(na) doubles):
// class c holds adj_plen(na) <= 128 << c
enum { kPhFs, kPhFactor, kPhRows, kPhFsSpline, kPhQag, kPhIc, kPhSolve, kPhAdj0 };
constexpr int kAdjClasses = 4;
constexpr int kPhases = kPhAdj0 + kAdjClasses;
constexpr int kMergedRestBlocks = 1280;

// The hardware deals the blocks of a grid to the 8 XCDs in turn (block b to
// XCD b mod 8), each XCD with its own L2.  Merged launches renumber their
// blocks so that each XCD runs one contiguous range of logical blocks, and so
// of histories: a history's tables are then read into one L2, not all eight.
// A bijection of [0, gridDim.x): XCD x takes logical blocks
// [x (q + 1), (x + 1)(q + 1)) for x < r, the next q-long ranges after, with
// n = 8 q + r blocks.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block()
{
    const int n = gridDim.x, b = blockIdx.x;
    const int x = b % kXcds, i = b / kXcds, q = n / kXcds, r = n % kXcds;
    return x < r ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

__device__ __forceinline__ int hist_of(const int *__restrict__ pre, int nhist, int blk)
{
    int lo = 0, hi = nhist;  // pre[lo] <= blk < pre[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= blk) lo = mid;
        else hi = mid;
    }
    return lo;
}

// pass-2 list entries: (history << 32) | node
struct GroupPush {
    long long *list;
    int *count;
    int h;
    __device__ void operator()(int node) const { list[atomicAdd(count, 1)] = ((long long)h << 32) | node; }
};

__global__ __launch_bounds__(256) void k_m_fs_first(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                    int nhist, long long *__restrict__ list, int *__restrict__ count)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    if (!d.integ) return;
    const GroupPush push = {list, count, h};
    fs_first_q4(d.H, d.nfs, d.x0, d.xa, d.light, d.fx, d.fy, d.status,
                (int)(((blk - pre[h]) * blockDim.x + threadIdx.x) >> 6), push);
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void k_m_fs_rest(const GroupDesc *__restrict__ D,
                                                                    const long long *__restrict__ list,
                                                                    const int *__restrict__ count)
{
    __shared__ QagWs ws[kWavesPerBlock];
    const int wave = threadIdx.x >> 6;
    const int n = *count;
    for (int t = blockIdx.x * kWavesPerBlock + wave; t < n; t += gridDim.x * kWavesPerBlock) {
        const long long e = list[t];
        const GroupDesc &d = D[(int)(e >> 32)];
        fs_node_wave(d.H, d.nfs, d.x0, d.xa, d.light, d.fx, d.fy, d.status, (int)(e & 0xffffffffLL), ws[wave]);
    }
}

__global__ __launch_bounds__(256) void k_m_factor(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                  int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    if (!d.integ) return;
    factor_block(d.sx, d.na, d.fac, d.fac + (size_t)fac_windows(d.na) * kFacStride, blk - pre[h]);
}

__global__ __launch_bounds__(256, 3) void k_m_rows_prep(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                     int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    if (!d.integ) return;
    __shared__ double lds[4 * kTriLds];
    rows_prep_block(d.sx, d.Y, d.C, d.na, d.nk, (long long)d.na_max, d.bsp, d.nsb, d.sinv_bw, d.bucket, d.fac,
                    blk - pre[h], lds);
}

__global__ __launch_bounds__(256) void k_m_fs_spline(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                     int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    const long long item = (long long)(blk - pre[h]) * 4 + (threadIdx.x >> 6);
    if (!d.integ || item >= (d.nfs + kKeep - 1) / kKeep) return;  // wave-uniform
    spline_item_wave(d.fx, d.fy, d.fc, d.nfs, 0LL, 0LL, item);
}

__global__ __launch_bounds__(64 * kWavesPerBlock) void k_m_dnu_qag(const GroupDesc *__restrict__ D,
                                                                    const int *__restrict__ pre, int nhist)
{
    __shared__ QagWs ws[kWavesPerBlock];
    HIP_DYNAMIC_SHARED(double, sxs)  // min(largest na, kStageMax) doubles
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    const int wave = threadIdx.x >> 6;
    const int k = (blk - pre[h]) * kWavesPerBlock + wave;
    if (!d.integ || k >= d.nk) return;
    QagArgs a;
    a.H = d.H;
    a.fx = d.fx;
    a.fy = d.fy;
    a.fc = d.fc;
    a.nfs = d.nfs;
    a.sx = d.sx;
    a.Y = d.Y;
    a.C = d.C;
    a.sb = d.bucket;
    a.ns = d.na;
    a.nsb = d.nsb;
    a.ystride = d.na_max;
    a.sinv_bw = d.sinv_bw;
    a.wavenum = d.wavenum;
    a.x0 = d.x0;
    a.xa = d.xa;
    a.out = d.out;
    a.nint_out = d.nint;
    a.status = d.status;
    a.ivl = d.ivl;
    a.ic.init = d.init;
    a.ic.fsl0 = d.fsl0;
    a.ic.lin = d.lin;
    a.ic.prefac = d.prefac;
    dnu_qag_item(a, *d.sp, k, k, ws[wave], d.na <= kStageMax ? sxs : nullptr);
}

// initial-condition-only histories, 64 k per block
__global__ __launch_bounds__(64) void k_m_ic(const GroupDesc *__restrict__ D, const int *__restrict__ pre, int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    const int k = (blk - pre[h]) * 64 + threadIdx.x;
    if (d.integ || k >= d.nk) return;
    IcArgs ic;
    ic.init = d.init;
    ic.fsl0 = d.fsl0;
    ic.lin = d.lin;
    ic.prefac = d.prefac;
    d.out[k] = ic_piece(d.wavenum[k], d.init[k], *d.sp, ic);
}

__device__ inline IcAdjArgs group_ica(const GroupDesc &d)
{
    IcAdjArgs ica;
    ica.init = d.init;
    ica.fsl0 = d.fsl0;
    ica.lin = d.lin;
    ica.init_bar = d.ibar;
    ica.mbar_ic = d.mbic;
    return ica;
}

// adjoint of the initial-condition-only histories: the piece's gradients and
// zero rows of dL/d delta_tot, 64 k (rows) per block
__global__ __launch_bounds__(64) void k_m_ic_adj(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                 int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    const int k0 = (blk - pre[h]) * 64;
    if (d.integ) return;
    if (k0 + (int)threadIdx.x < d.nk) ic_adjoint_at(k0 + threadIdx.x, d.wavenum, d.gbar, *d.sp, group_ica(d));
    for (int r = 0; r < 64 && k0 + r < d.nk; r++) {
        double *row = d.dbar + (size_t)(k0 + r) * d.na_max;
        for (int t = threadIdx.x; t < d.na_max; t += 64) row[t] = 0.0;
    }
}

__global__ __launch_bounds__(64) void k_m_dnu_adj(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                  int nhist)
{
    HIP_DYNAMIC_SHARED(double, lds)  // adj_lds_bytes of the class's longest part
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    if (!d.integ) return;
    AdjArgs A;
    A.H = d.H;
    A.fx = d.fx;
    A.fy = d.fy;
    A.fc = d.fc;
    A.nfs = d.nfs;
    A.x0 = d.x0;
    A.xa = d.xa;
    A.sx = d.sx;
    A.Y = d.Y;
    A.C = d.C;
    A.sb = d.bucket;
    A.ns = d.na;
    A.nsb = d.nsb;
    A.nk = d.nk;
    A.ystride = d.na_max;
    A.sinv_bw = d.sinv_bw;
    A.wavenum = d.wavenum;
    A.sp = d.sp;
    A.vcrit = d.vcrit;
    A.ivl = d.ivl;
    A.nint = d.nint;
    A.gbar = d.gbar;
    A.prefac = d.prefac;
    A.ybar = d.dbar;
    A.cbar = d.cbar;
    A.mu_bar = d.mubar;
    A.nparts = adj_parts(d.na);
    A.plen = adj_plen(d.na);
    A.ica = group_ica(d);
    dnu_adj_block(A, blk - pre[h], lds);
}

__global__ __launch_bounds__(256) void k_m_adj_solve(const GroupDesc *__restrict__ D, const int *__restrict__ pre,
                                                     int nhist)
{
    const int blk = xcd_block();
    const int h = hist_of(pre, nhist, blk);
    const GroupDesc &d = D[h];
    if (!d.integ) return;
    __shared__ double lds[4 * kTriLds];
    adj_solve_block(d.sx, d.na, d.nk, (long long)d.na_max, d.cbar, d.dbar, d.fac, blk - pre[h], lds);
}

// One more stored step per history, on the device (one wavefront per
// history).  mode bit 0: check the new log a la[h] (capacity, the H table's
// range, strictly after the last stored step; 3004 / 3002 as the host's
// append) and compute fslength(log T0, la) by the wavefront QAG over the
// history's H panels (delta_tot_table.c:394-407), into err[h] and fsl[h];
// bit 1: where err[h] is 0, store la and the history's new delta_tot column
// (rows[off_k + k]) at column na and update what get_delta_nu derives from a
// and the stored steps (delta_tot_table.c:513-555: hybrid qc and tolerance,
// the IC-only cases, the IC factor 1 + deriv_prefac fsl) in the descriptor,
// the species record and the mass-gradient metadata.  Mode 3 does both in one
// launch and also ORs the code into the sticky sticky[h] (the capturable
// step); the host's append runs mode 1, reads err, then mode 2.
__global__ __launch_bounds__(64) void k_group_step(GroupDesc *__restrict__ D, SpeciesDev *__restrict__ sp,
                                                   GroupMeta *__restrict__ meta, const double *__restrict__ la_in,
                                                   const double *__restrict__ rows, double *__restrict__ fsl,
                                                   int *__restrict__ err, int *__restrict__ sticky, int mode)
{
    __shared__ QagWs ws;
    const int h = blockIdx.x, lane = threadIdx.x & 63;
    GroupDesc &d = D[h];
    const double la = la_in[h];
    int code = 0;
    double f = 0.0;
    if (mode & 1) {
        const int na = d.na;
        if (na >= d.na_max || !(la <= d.xmax) || !(la > d.x0)) code = KN_ERR_RANGE;
        else if (!(la > d.sx[na - 1])) code = KN_ERR_ARG;
        if (!code) {  // wave-uniform
            FslIntegrand fi;
            fi.H = d.H;
            double r = 0.0;
            int ni = 0;
            if (wave_qag61(fi, d.x0, la, 1e-6, kLimit, ws, &r, &ni)) code = 2001;
            f = d.light * r;
        }
        if (lane == 0) {
            err[h] = code;
            fsl[h] = f;
            if ((mode & 2) && code) sticky[h] = code;
        }
    } else {
        code = err[h];
        f = fsl[h];
    }
    if (!(mode & 2) || code) return;
    const int na = d.na, nk = d.nk;
    const long long off = meta[h].off_k;
    for (int k = lane; k < nk; k += 64) d.Y[(size_t)k * d.na_max + na] = rows[off + k];
    const int n1 = na + 1;
    const double a = exp(la);
    const double partnu = (d.hybrid && a > d.nu_crit_time) ? d.nufrac_low : 0.0;
    double qc = 0.0, relerr = 1e-6;
    int ic_only = 0;
    if (partnu > 0) {
        if (1 - partnu < 1e-3) ic_only = 1;
        qc = d.vcrit_spec * (d.mnu / d.kBtnu);
        relerr /= (1. + 1e-5 - partnu);
    }
    if (!(n1 > 1 && d.mnu > 0)) ic_only = 1;
    const int integ = !ic_only && d.mnu > 0 && n1 >= 2;
    __syncthreads();  // every lane has read the descriptor before lane 0 rewrites it
    if (lane < 20) sp[h].J.en[lane] = lane == 0 ? 1.0 : exp(-lane * qc);
    if (lane == 0) {
        d.sx[na] = la;
        d.na = n1;
        d.nfs = 16 * n1;
        d.nsb = 4 * n1;
        d.bsp = rows_prep_spline_blocks(n1, nk);
        d.integ = integ;
        d.ic_only = ic_only;
        d.xa = la;
        d.a = a;
        d.sinv_bw = 4.0 * n1 / (la - d.x0);
        d.fsl0 = f;
        d.lin = 1. + d.deriv_prefac * f;
        d.vcrit = qc > 0 ? d.vcrit_spec : 0.0;
        sp[h].relerr = relerr;
        sp[h].J.qc = qc;
        meta[h].integ = integ;
        meta[h].nparts = adj_parts(n1);
    }
}

}  // namespace

struct kn_group {
    int nhist = 0, nstream = 0, cap_err = 0;  // nstream 0: merged launches
    long long nk_tot = 0, y_tot = 0, s_tot = 0, fs_tot = 0, b_tot = 0, fac_tot = 0, mu_tot = 0;
    std::vector<GroupHist> h;
    std::vector<HTab> htabs;
    std::vector<double *> hcoef;
    std::vector<int> order;
    double *sx = nullptr, *Y = nullptr, *C = nullptr, *dbar = nullptr, *cbar = nullptr, *fx = nullptr,
           *fy = nullptr, *fc = nullptr, *wavenum = nullptr, *init = nullptr, *out = nullptr, *gbar = nullptr,
           *ibar = nullptr, *mbic = nullptr, *mubar = nullptr, *mbar = nullptr, *ivl = nullptr, *fac = nullptr;
    int *bucket = nullptr, *nint = nullptr, *status = nullptr, *zeros = nullptr, *spec0 = nullptr, *fs_list = nullptr;
    SpeciesDev *sp = nullptr;
    GroupMeta *meta = nullptr;
    // merged launches
    GroupDesc *desc = nullptr;
    int *pre = nullptr;      // kPhases x (nhist + 1)
    long long *mlist = nullptr;  // pass-2 fslength list (fs_tot entries)
    std::vector<GroupDesc> desc_h;
    size_t qag_lds = 0;  // staged log a values of the longest stageable history
    int tot[kPhases] = {};
    size_t adj_lds[kAdjClasses] = {};
    hipStream_t st[kGroupStreams] = {};
    hipEvent_t fork = nullptr, join[kGroupStreams] = {};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int have_graph = 0, graph_appends = 0;
    long long gen = 0, graph_gen = -1;
    // launches sized for na_max (the device append leaves them valid)
    int cap_plan = 0;
    // the device changed the stored steps since the host mirror was read
    int host_stale = 0;
    // device append: fslength(log T0, la) and the check code per history, and
    // the codes of capturable appends, kept until kn_group_status reports them
    double *step_fsl = nullptr;
    int *step_err = nullptr, *step_sticky = nullptr;
};

namespace {

SpeciesDev species_dev(const kn_batch_history &s)
{
    SpeciesDev d;
    d.mnubykT = s.mnu / s.kBtnu;
    d.relerr = s.relerr;
    d.J.qc = s.qc;
    d.J.nufrac_low = s.nufrac_low;
    d.J.en[0] = 1.0;
    for (int n = 1; n < 20; n++) d.J.en[n] = exp(-n * s.qc);
    return d;
}

// fslength(log T0, log xa) with the history's own H (delta_tot_table.c:394-407),
// and the initial-condition factor 1 + deriv_prefac fsl (:522-535).
int group_ic_factors(const GroupHist &q, double xa, double *fsl0, double *lin)
{
    HubCtx hc = {q.spec.hubble, q.spec.hubble_ctx};
    double v = 0.0, e = 0.0;
    if (q.x0 < xa && kn_qag61(fsl_integrand_ctx, &hc, q.x0, xa, 0, 1e-6, KN_GSL_VAL, &v, &e)) {
        kn_set_last_error("kn_group: QAG failed in fslength(%g, %g)", q.x0, xa);
        return 2001;
    }
    *fsl0 = q.spec.light * v;
    *lin = 1. + q.deriv_prefac * *fsl0;
    return 0;
}

int group_refresh_ic(GroupHist &q) { return group_ic_factors(q, q.xa, &q.fsl0, &q.lin); }

void group_set_integ(GroupHist &q) { q.integ = !q.spec.ic_only && q.spec.mnu > 0 && q.na >= 2; }

// What get_delta_nu derives from a, the stored step count and the mass
// (delta_tot_table.c:513-555): with particle neutrinos active (hybrid and
// a > nu_crit_time, fraction nufrac_low[0]) the integral is skipped when
// they are nearly all the neutrinos, otherwise qc = vcrit m/kT and the
// tolerance is loosened; one stored step or no mass leaves the
// initial-condition piece alone.
void group_derive(GroupHist &q)
{
    kn_batch_history &s = q.spec;
    const double partnu = (s.hybrid && s.a > s.nu_crit_time) ? s.nufrac_low : 0.0;
    s.qc = 0.0;
    s.relerr = 1e-6;
    s.ic_only = 0;
    if (partnu > 0) {
        if (1 - partnu < 1e-3) s.ic_only = 1;
        s.qc = s.vcrit * (s.mnu / s.kBtnu);
        s.relerr /= (1. + 1e-5 - partnu);
    }
    if (!(q.na > 1 && s.mnu > 0)) s.ic_only = 1;
    q.vcrit = s.qc > 0 ? s.vcrit : 0.0;
    group_set_integ(q);
}

// Longest-first issue order, each history to the least-loaded stream.
void group_schedule(kn_group *g)
{
    std::vector<double> cost(g->nhist);
    for (int i = 0; i < g->nhist; i++) {
        const GroupHist &q = g->h[i];
        cost[i] = q.integ ? (double)(q.spec.nk + 16) * q.na : 1.0;
    }
    g->order.resize(g->nhist);
    for (int i = 0; i < g->nhist; i++) g->order[i] = i;
    std::stable_sort(g->order.begin(), g->order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    if (g->nstream < 1) return;  // merged launches use no streams of their own
    std::vector<double> load(g->nstream, 0.0);
    for (int i : g->order) {
        int best = 0;
        for (int s = 1; s < g->nstream; s++)
            if (load[s] < load[best]) best = s;
        g->h[i].stream = best;
        load[best] += cost[i];
    }
}

void group_meta_upload(kn_group *g)
{
    std::vector<GroupMeta> m(g->nhist);
    for (int i = 0; i < g->nhist; i++) {
        m[i].off_k = g->h[i].off_k;
        m[i].off_mu = g->h[i].off_mu;
        m[i].nparts = adj_parts(g->h[i].na);
        m[i].nk = g->h[i].spec.nk;
        m[i].integ = g->h[i].integ;
        m[i].kBtnu = g->h[i].spec.kBtnu;
    }
    KN_CHECK(hipMemcpy(g->meta, m.data(), sizeof(GroupMeta) * g->nhist, hipMemcpyHostToDevice));
}

inline double group_sinv_bw(const GroupHist &q)
{
    // any positive bucket width gives the same segments (the lookup walks from
    // its bucket); this one spans [log T0, log a]
    return 4.0 * q.na / (q.xa - q.x0);
}

// Descriptors and per-phase block prefixes for the stored steps (after
// create and append; launches read them from device memory, so a captured
// graph needs no host data).
void group_plan(kn_group *g)
{
    const int n = g->nhist;
    g->desc_h.resize(n);
    std::vector<int> pre((size_t)kPhases * (n + 1), 0);
    std::vector<long long> acc(kPhases, 0);
    size_t lds[kAdjClasses] = {};
    for (int i = 0; i < n; i++) {
        const GroupHist &q = g->h[i];
        GroupDesc &d = g->desc_h[i];
        d.H = g->htabs[q.htab];
        d.sx = g->sx + q.off_s;
        d.Y = g->Y + q.off_y;
        d.wavenum = g->wavenum + q.off_k;
        d.init = g->init + q.off_k;
        d.gbar = g->gbar + q.off_k;
        d.C = g->C + q.off_y;
        d.fx = g->fx + q.off_fs;
        d.fy = g->fy + q.off_fs;
        d.fc = g->fc + q.off_fs;
        d.out = g->out + q.off_k;
        d.ivl = g->ivl + (size_t)q.off_k * kLimit * 2;
        d.dbar = g->dbar + q.off_y;
        d.cbar = g->cbar + q.off_y;
        d.mubar = g->mubar + q.off_mu;
        d.ibar = g->ibar + q.off_k;
        d.mbic = g->mbic + q.off_k;
        d.fac = g->fac + q.off_fac;
        d.bucket = g->bucket + q.off_b;
        d.nint = g->nint + q.off_k;
        d.status = g->status + 2 * i;
        d.sp = g->sp + i;
        d.nk = q.spec.nk;
        d.na = q.na;
        d.na_max = q.spec.na_max;
        d.nfs = 16 * q.na;
        d.nsb = 4 * q.na;
        d.bsp = rows_prep_spline_blocks(q.na, q.spec.nk);
        d.integ = q.integ;
        d.ic_only = q.spec.ic_only;
        d.x0 = q.x0;
        d.xa = q.xa;
        d.a = q.spec.a;
        d.light = q.spec.light;
        d.sinv_bw = group_sinv_bw(q);
        d.fsl0 = q.fsl0;
        d.lin = q.lin;
        d.prefac = q.spec.delta_nu_prefac;
        d.vcrit = q.vcrit;
        d.hybrid = q.spec.hybrid;
        d.xmax = q.xmax;
        d.nu_crit_time = q.spec.nu_crit_time;
        d.nufrac_low = q.spec.nufrac_low;
        d.vcrit_spec = q.spec.vcrit;
        d.mnu = q.spec.mnu;
        d.kBtnu = q.spec.kBtnu;
        d.deriv_prefac = q.deriv_prefac;
        // launch sizes: for the stored steps, or (capacity plan) for na_max
        // with both the integral's and the IC-only phases, the kernels reading
        // which one applies and how much of the grid is live from d
        const int np = g->cap_plan ? q.spec.na_max : q.na;
        const bool p_integ = g->cap_plan || q.integ, p_ic = g->cap_plan || !q.integ;
        long long cnt[kPhases] = {};
        if (p_integ) {
            cnt[kPhFs] = fs_first_blocks(16 * np);
            cnt[kPhFactor] = factor_blocks(np);
            cnt[kPhRows] = rows_prep_blocks(rows_prep_spline_blocks(np, d.nk), 4 * np);
            cnt[kPhFsSpline] = spline_blocks(16 * np, 1);
            cnt[kPhQag] = blocks_for(d.nk);
            cnt[kPhSolve] = np > 2 ? spline_blocks(np, d.nk) : 0;
            // part length: adj_plen is not monotone in the step count (1024 at
            // 1024 knots, 513 at 1025), so the capacity plan takes the longest
            // part any count up to na_max has, min(na_max, kAdjPartKnots)
            const int nl = g->cap_plan ? (np < kAdjPartKnots ? np : kAdjPartKnots) : np;
            int c = 0;
            while (c < kAdjClasses - 1 && adj_plen(nl) > (128 << c)) c++;
            cnt[kPhAdj0 + c] = (long long)d.nk * adj_parts(np) + (d.nk + 63) / 64;
            const size_t b = adj_lds_bytes(nl);
            lds[c] = b > lds[c] ? b : lds[c];
        }
        if (p_ic) cnt[kPhIc] = (d.nk + 63) / 64;
        for (int p = 0; p < kPhases; p++) {
            acc[p] += cnt[p];
            if (acc[p] > 0x7fffffffLL) kn_terminate(KN_ERR_RANGE, "kn_group: %lld blocks in one phase\n", acc[p]);
            pre[(size_t)p * (n + 1) + i + 1] = (int)acc[p];
        }
    }
    for (int p = 0; p < kPhases; p++) g->tot[p] = (int)acc[p];
    g->qag_lds = 0;
    for (int i = 0; i < n; i++) {
        const int np = g->cap_plan ? (g->h[i].spec.na_max < kStageMax ? g->h[i].spec.na_max : kStageMax) : g->h[i].na;
        if ((g->cap_plan || g->h[i].integ) && np <= kStageMax && sizeof(double) * np > g->qag_lds)
            g->qag_lds = sizeof(double) * np;
    }
    for (int c = 0; c < kAdjClasses; c++) g->adj_lds[c] = lds[c];
    KN_CHECK(hipMemcpy(g->desc, g->desc_h.data(), sizeof(GroupDesc) * n, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(g->pre, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice));
    group_meta_upload(g);
}

// Forward of one history on stream s, recording its subdivisions: the
// preparations in one launch, the fslength spline, then QAG with the
// initial-condition piece added per k (k_group_ic alone when IC-only).
void group_forward_one(kn_group *g, const GroupHist &q, int i, hipStream_t s)
{
    const int nk = q.spec.nk, na = q.na, na_max = q.spec.na_max;
    if (!q.integ) {
        IcArgs ic;
        ic.init = g->init + q.off_k;
        ic.fsl0 = q.fsl0;
        ic.lin = q.lin;
        ic.prefac = q.spec.delta_nu_prefac;
        hipLaunchKernelGGL(k_group_ic, dim3((nk + 63) / 64), dim3(64), 0, s, g->wavenum + q.off_k, g->init + q.off_k,
                           g->sp + i, nk, ic, g->out + q.off_k);
        return;
    }
    const int nfs = 16 * na;
    double *fx = g->fx + q.off_fs, *fy = g->fy + q.off_fs, *fc = g->fc + q.off_fs;
    int *fs_list = g->fs_list + q.off_fs, *fs_count = g->status + 2 * g->nhist + i;
    const GroupDesc &d = g->desc_h[i];
    hipLaunchKernelGGL(k_fs_first, dim3(fs_first_blocks(nfs)), dim3(256), 0, s, d.H, nfs, q.x0, q.xa, q.spec.light, fx,
                       fy, d.status, fs_list, fs_count);
    hipLaunchKernelGGL(k_fs_rest, dim3(kFsRestBlocks), dim3(64 * kWavesPerBlock), 0, s, d.H, nfs, q.x0, q.xa,
                       q.spec.light, fx, fy, d.status, fs_list, fs_count);
    hipLaunchKernelGGL(k_factor, dim3(factor_blocks(na)), dim3(256), 0, s, d.sx, na, d.fac,
                       d.fac + (size_t)fac_windows(na) * kFacStride);
    hipLaunchKernelGGL(k_rows_prep, dim3(rows_prep_blocks(d.bsp, d.nsb)), dim3(256), 0, s, d.sx, d.Y, d.C, na, nk,
                       (long long)na_max, d.bsp, d.nsb, d.sinv_bw, d.bucket, d.fac);
    hipLaunchKernelGGL(k_spline, dim3(spline_blocks(nfs, 1)), dim3(256), 0, s, fx, fy, fc, nfs, 1, 0LL, 0LL);
    QagArgs a;
    a.H = d.H;
    a.fx = fx;
    a.fy = fy;
    a.fc = fc;
    a.nfs = nfs;
    a.sx = d.sx;
    a.Y = d.Y;
    a.C = d.C;
    a.sb = d.bucket;
    a.ns = na;
    a.nsb = d.nsb;
    a.ystride = na_max;
    a.sinv_bw = d.sinv_bw;
    a.wavenum = d.wavenum;
    a.x0 = q.x0;
    a.xa = q.xa;
    a.out = d.out;
    a.nint_out = d.nint;
    a.status = d.status;
    a.ivl = d.ivl;
    a.ic.init = d.init;
    a.ic.fsl0 = q.fsl0;
    a.ic.lin = q.lin;
    a.ic.prefac = q.spec.delta_nu_prefac;
    hipLaunchKernelGGL(k_dnu_qag_spec, dim3(nk), dim3(64 * kQagSpec), na <= kStageMax ? sizeof(double) * na : 0, s,
                       a, g->sp + i, g->spec0, nk, nk);
}

// Adjoint of one history's last forward on stream s.
void group_adjoint_one(kn_group *g, const GroupHist &q, int i, hipStream_t s)
{
    const int nk = q.spec.nk, na = q.na, na_max = q.spec.na_max;
    const long long nblk = (long long)nk * na_max;
    double *dbar = g->dbar + q.off_y;
    IcAdjArgs ica;
    ica.init = g->init + q.off_k;
    ica.fsl0 = q.fsl0;
    ica.lin = q.lin;
    ica.init_bar = g->ibar + q.off_k;
    ica.mbar_ic = g->mbic + q.off_k;
    if (!q.integ) {
        hipLaunchKernelGGL(k_group_zero, dim3((int)((nblk + 255) / 256)), dim3(256), 0, s, dbar, nblk);
        hipLaunchKernelGGL(k_group_ic_adj, dim3((nk + 63) / 64), dim3(64), 0, s, g->wavenum + q.off_k,
                           g->gbar + q.off_k, g->sp + i, nk, ica);
        return;
    }
    const GroupDesc &d = g->desc_h[i];
    AdjArgs A;
    A.H = d.H;
    A.fx = d.fx;
    A.fy = d.fy;
    A.fc = d.fc;
    A.nfs = d.nfs;
    A.x0 = d.x0;
    A.xa = d.xa;
    A.sx = d.sx;
    A.Y = d.Y;
    A.C = d.C;
    A.sb = d.bucket;
    A.ns = na;
    A.nsb = d.nsb;
    A.nk = nk;
    A.ystride = na_max;
    A.sinv_bw = d.sinv_bw;
    A.wavenum = d.wavenum;
    A.sp = d.sp;
    A.vcrit = q.vcrit;
    A.ivl = d.ivl;
    A.nint = d.nint;
    A.gbar = d.gbar;
    A.prefac = q.spec.delta_nu_prefac;
    A.ybar = dbar;  // rows written whole: [0, na) the item's terms, [na, na_max) zeros
    A.cbar = d.cbar;
    A.mu_bar = d.mubar;
    A.nparts = adj_parts(na);
    A.plen = adj_plen(na);
    A.ica = ica;
    // nk x parts item blocks, then the initial-condition piece's blocks
    hipLaunchKernelGGL(k_dnu_adj, dim3(nk * A.nparts + (nk + 63) / 64), dim3(64), adj_lds_bytes(na), s, A);
    if (na > 2)
        hipLaunchKernelGGL(k_adj_solve, dim3(spline_blocks(na, nk)), dim3(256), 0, s, d.sx, na, nk, (long long)na_max,
                           d.cbar, dbar, d.fac);
}

// The whole group in one grid per phase, on stream s.
void group_launch_merged(kn_group *g, hipStream_t s, bool forward, bool adjoint)
{
    const int n = g->nhist;
    auto pre = [&](int ph) { return g->pre + (size_t)ph * (n + 1); };
    if (forward) {
        if (g->tot[kPhFs]) {
            hipLaunchKernelGGL(k_m_fs_first, dim3(g->tot[kPhFs]), dim3(256), 0, s, g->desc, pre(kPhFs), n, g->mlist,
                               g->status + 3 * n);
            hipLaunchKernelGGL(k_m_fs_rest, dim3(kMergedRestBlocks), dim3(64 * kWavesPerBlock), 0, s, g->desc,
                               g->mlist, g->status + 3 * n);
        }
        if (g->tot[kPhFactor])
            hipLaunchKernelGGL(k_m_factor, dim3(g->tot[kPhFactor]), dim3(256), 0, s, g->desc, pre(kPhFactor), n);
        if (g->tot[kPhRows])
            hipLaunchKernelGGL(k_m_rows_prep, dim3(g->tot[kPhRows]), dim3(256), 0, s, g->desc, pre(kPhRows), n);
        if (g->tot[kPhFsSpline])
            hipLaunchKernelGGL(k_m_fs_spline, dim3(g->tot[kPhFsSpline]), dim3(256), 0, s, g->desc, pre(kPhFsSpline), n);
        if (g->tot[kPhQag])
            hipLaunchKernelGGL(k_m_dnu_qag, dim3(g->tot[kPhQag]), dim3(64 * kWavesPerBlock), g->qag_lds, s,
                               g->desc, pre(kPhQag), n);
        if (g->tot[kPhIc]) hipLaunchKernelGGL(k_m_ic, dim3(g->tot[kPhIc]), dim3(64), 0, s, g->desc, pre(kPhIc), n);
    }
    if (adjoint) {
        for (int c = 0; c < kAdjClasses; c++)
            if (g->tot[kPhAdj0 + c])
                hipLaunchKernelGGL(k_m_dnu_adj, dim3(g->tot[kPhAdj0 + c]), dim3(64), g->adj_lds[c], s, g->desc,
                                   pre(kPhAdj0 + c), n);
        if (g->tot[kPhSolve])
            hipLaunchKernelGGL(k_m_adj_solve, dim3(g->tot[kPhSolve]), dim3(256), 0, s, g->desc, pre(kPhSolve), n);
        if (g->tot[kPhIc])
            hipLaunchKernelGGL(k_m_ic_adj, dim3(g->tot[kPhIc]), dim3(64), 0, s, g->desc, pre(kPhIc), n);
    }
}

// Forward and/or adjoint of every history, asynchronous on origin (and
// capturable): merged launches (nstream 0), or each history's launch sequence
// on one of nstream streams forked from origin and joined back.
void group_launch(kn_group *g, hipStream_t origin, bool forward, bool adjoint)
{
    if (forward)
        KN_CHECK(hipMemcpyAsync(g->status, g->zeros, sizeof(int) * (3 * (size_t)g->nhist + 1), hipMemcpyDeviceToDevice,
                                origin));
    if (g->nstream == 0) {
        group_launch_merged(g, origin, forward, adjoint);
        KN_CHECK(hipGetLastError());
        if (adjoint) {
            hipLaunchKernelGGL(k_group_mass_bar, dim3(g->nhist), dim3(64), 0, origin, g->meta, g->mbic, g->mubar,
                               g->mbar);
            KN_CHECK(hipGetLastError());
        }
        return;
    }
    KN_CHECK(hipEventRecord(g->fork, origin));
    for (int s = 0; s < g->nstream; s++) KN_CHECK(hipStreamWaitEvent(g->st[s], g->fork, 0));
    for (int i : g->order) {
        const GroupHist &q = g->h[i];
        if (forward) group_forward_one(g, q, i, g->st[q.stream]);
        if (adjoint) group_adjoint_one(g, q, i, g->st[q.stream]);
    }
    KN_CHECK(hipGetLastError());
    for (int s = 0; s < g->nstream; s++) {
        KN_CHECK(hipEventRecord(g->join[s], g->st[s]));
        KN_CHECK(hipStreamWaitEvent(origin, g->join[s], 0));
    }
    if (adjoint) {
        hipLaunchKernelGGL(k_group_mass_bar, dim3(g->nhist), dim3(64), 0, origin, g->meta, g->mbic, g->mubar,
                           g->mbar);
        KN_CHECK(hipGetLastError());
    }
}

// The host mirror of what the device append changes (stored steps, upper
// limit, derived tolerance / qc / IC-only state, IC factors), read back after
// device appends; every host call that uses them reads it first.
void group_pull(kn_group *g)
{
    if (!g->host_stale) return;
    const int n = g->nhist;
    KN_CHECK(hipDeviceSynchronize());
    KN_CHECK(hipMemcpy(g->desc_h.data(), g->desc, sizeof(GroupDesc) * n, hipMemcpyDeviceToHost));
    std::vector<SpeciesDev> sd(n);
    KN_CHECK(hipMemcpy(sd.data(), g->sp, sizeof(SpeciesDev) * n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) {
        GroupHist &q = g->h[i];
        const GroupDesc &d = g->desc_h[i];
        q.na = d.na;
        q.xa = d.xa;
        q.spec.a = d.a;
        q.fsl0 = d.fsl0;
        q.lin = d.lin;
        q.vcrit = d.vcrit;
        q.integ = d.integ;
        q.spec.ic_only = d.ic_only;
        q.spec.qc = sd[i].J.qc;
        q.spec.relerr = sd[i].relerr;
    }
    g->host_stale = 0;
}

int group_status(kn_group *g)
{
    if (g->cap_err) {
        kn_set_last_error("kn_group: capacity (na_max or the H table range) exceeded");
        return KN_ERR_RANGE;
    }
    if (g->cap_plan) {  // codes of device appends (sticky until reported here)
        std::vector<int> se(g->nhist);
        KN_CHECK(hipMemcpy(se.data(), g->step_sticky, sizeof(int) * se.size(), hipMemcpyDeviceToHost));
        for (int i = 0; i < g->nhist; i++)
            if (se[i]) {
                KN_CHECK(hipMemset(g->step_sticky, 0, sizeof(int) * se.size()));
                kn_set_last_error("kn_group: history %d: device append rejected (%s)", i,
                                  se[i] == KN_ERR_RANGE ? "capacity or H table range"
                                  : se[i] == KN_ERR_ARG ? "log a does not follow the last stored step"
                                                        : "QAG failed in fslength");
                return se[i];
            }
    }
    std::vector<int> st(2 * (size_t)g->nhist);
    KN_CHECK(hipMemcpy(st.data(), g->status, sizeof(int) * st.size(), hipMemcpyDeviceToHost));
    for (int i = 0; i < g->nhist; i++) {
        if (st[2 * i + 1]) {
            kn_set_last_error("kn_group: history %d: delta_nu integrand evaluated outside a stored table (GSL_EDOM)", i);
            return 2001;
        }
        if (st[2 * i]) {  // GSL error -> terminate(2001) in the reference
            kn_set_last_error("kn_group: history %d: QAG failed (%s)", i, (st[2 * i] & 1) ? "fslength" : "delta_nu integral");
            return 2001;
        }
    }
    return 0;
}

template <typename T>
T *group_alloc(size_t n)
{
    T *p = nullptr;
    KN_CHECK(hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)));
    return p;
}

}  // namespace

extern "C" void kn_group_destroy(kn_group *g)
{
    if (!g) return;
    if (g->have_graph) {
        KN_CHECK(hipGraphExecDestroy(g->exec));
        KN_CHECK(hipGraphDestroy(g->graph));
    }
    void *ptrs[] = {g->sx,   g->Y,    g->C,      g->dbar, g->cbar,   g->fx,     g->fy,    g->fc,
                    g->wavenum, g->init, g->out,  g->gbar,   g->ibar, g->mbic,   g->mubar,  g->mbar,  g->ivl,
                    g->bucket, g->nint, g->status, g->zeros, g->spec0, g->sp,    g->fs_list, g->meta, g->desc,
                    g->pre,  g->mlist, g->fac,  g->step_fsl, g->step_err, g->step_sticky};
    for (void *p : ptrs)
        if (p) KN_CHECK(hipFree(p));
    for (double *p : g->hcoef) KN_CHECK(hipFree(p));
    for (int s = 0; s < kGroupStreams; s++) {
        if (g->st[s]) KN_CHECK(hipStreamDestroy(g->st[s]));
        if (g->join[s]) KN_CHECK(hipEventDestroy(g->join[s]));
    }
    if (g->fork) KN_CHECK(hipEventDestroy(g->fork));
    delete g;
}

extern "C" int kn_group_create(kn_group **out, int nhist, const kn_batch_history *hist)
{
    if (!out || nhist < 1 || !hist) return KN_ERR_ARG;
    for (int i = 0; i < nhist; i++) {
        const kn_batch_history &s = hist[i];
        if (s.nk < 1 || s.na < 1 || s.na_max < s.na || !s.hubble || !(s.a > s.time_transfer) || !(s.kBtnu > 0)) {
            kn_set_last_error("kn_group_create: history %d: bad sizes, range or missing hubble", i);
            return KN_ERR_ARG;
        }
    }
    upload_constants();
    kn_group *g = new kn_group();
    g->nhist = nhist;
    g->h.resize(nhist);
    std::vector<BgTableReq> bg_reqs;  // H tables built from backgrounds, on the device
    std::vector<int> bg_tab;          // their places in g->htabs
    for (int i = 0; i < nhist; i++) {
        GroupHist &q = g->h[i];
        q.spec = hist[i];
        q.na = q.spec.na;
        q.x0 = log(q.spec.time_transfer);
        q.xa = log(q.spec.a);
        q.xmax = log(q.spec.a_max > q.spec.a ? q.spec.a_max : q.spec.a);
        group_derive(q);
        q.off_k = g->nk_tot;
        q.off_y = g->y_tot;
        q.off_s = g->s_tot;
        q.off_fs = g->fs_tot;
        q.off_b = g->b_tot;
        q.off_fac = g->fac_tot;
        g->fac_tot += (long long)fac_doubles(q.spec.na_max);
        q.off_mu = g->mu_tot;
        g->mu_tot += (long long)q.spec.nk * adj_parts(q.spec.na_max);
        g->nk_tot += q.spec.nk;
        g->y_tot += (long long)q.spec.nk * q.spec.na_max;
        g->s_tot += q.spec.na_max;
        g->fs_tot += 16LL * q.spec.na_max;
        g->b_tot += 4LL * q.spec.na_max + 1;
        {
            const double T0 = q.spec.time_transfer;  // delta_tot_table.c:522-535
            q.deriv_prefac = T0 * (q.spec.hubble(T0, q.spec.hubble_ctx) / q.spec.light) * T0;
        }
        const int rc = group_refresh_ic(q);
        if (rc) {
            kn_group_destroy(g);
            return rc;
        }
        // one H table per (H, T0): panel edges on log T0, up to log max(a, a_max)
        int found = -1;
        for (int j = 0; j < i && found < 0; j++) {
            const GroupHist &p = g->h[j];
            if (p.spec.hubble == q.spec.hubble && p.spec.hubble_ctx == q.spec.hubble_ctx && p.x0 == q.x0 &&
                p.xmax >= q.xmax)
                found = p.htab;
        }
        if (found < 0) {
            const int npanel = (int)ceil((q.xmax - q.x0) / kGroupHWidth) + 1;
            HTab H;
            double *d;
            if (q.spec.hubble == kn_background_hubble_ctx) {  // built on the device, all together below
                d = group_alloc<double>((size_t)npanel * kHStride + (size_t)kHSplitMax * kHSplitStride);
                bg_reqs.push_back(BgTableReq{(const kn_background *)q.spec.hubble_ctx, q.x0, kGroupHWidth, npanel, 0,
                                             d, nullptr});
                bg_tab.push_back((int)g->htabs.size());
            } else {
                std::vector<double> coef;
                double (*hub)(double, void *) = q.spec.hubble;
                void *ctx = q.spec.hubble_ctx;
                h_table(q.x0, kGroupHWidth, npanel, [hub, ctx](double a) { return hub(a, ctx); }, coef, H);
                d = group_alloc<double>(coef.size());
                KN_CHECK(hipMemcpy(d, coef.data(), sizeof(double) * coef.size(), hipMemcpyHostToDevice));
            }
            H.coef = d;
            g->hcoef.push_back(d);
            g->htabs.push_back(H);
            found = (int)g->htabs.size() - 1;
        }
        q.htab = found;
    }
    if (!bg_reqs.empty()) {
        for (size_t j = 0; j < bg_reqs.size(); j++) bg_reqs[j].out = &g->htabs[bg_tab[j]];
        bg_htables(bg_reqs, 0);
    }
    const size_t ny = (size_t)g->y_tot, nk = (size_t)g->nk_tot;
    g->sx = group_alloc<double>(g->s_tot);
    for (double **p : {&g->Y, &g->C, &g->dbar, &g->cbar}) *p = group_alloc<double>(ny);
    for (double **p : {&g->fx, &g->fy, &g->fc}) *p = group_alloc<double>(g->fs_tot);
    for (double **p : {&g->wavenum, &g->init, &g->out, &g->gbar, &g->ibar, &g->mbic})
        *p = group_alloc<double>(nk);
    g->mubar = group_alloc<double>(g->mu_tot);  // [k][part] per history
    g->mbar = group_alloc<double>(nhist);
    g->fac = group_alloc<double>(g->fac_tot);
    g->ivl = group_alloc<double>(nk * kLimit * 2);
    g->bucket = group_alloc<int>(g->b_tot);
    g->nint = group_alloc<int>(nk);
    // status: [2 h] QAG, [2 h + 1] domain, [2 nhist + h] fs list count, [3 nhist] merged fs list count
    g->status = group_alloc<int>(3 * (size_t)nhist + 1);
    g->zeros = group_alloc<int>(3 * (size_t)nhist + 1);
    g->fs_list = group_alloc<int>(g->fs_tot);
    g->mlist = group_alloc<long long>(g->fs_tot);
    g->spec0 = group_alloc<int>(1);
    g->step_fsl = group_alloc<double>(nhist);
    g->step_err = group_alloc<int>(nhist);
    g->step_sticky = group_alloc<int>(nhist);
    KN_CHECK(hipMemset(g->step_sticky, 0, sizeof(int) * nhist));
    g->desc = group_alloc<GroupDesc>(nhist);
    g->pre = group_alloc<int>((size_t)kPhases * (nhist + 1));
    KN_CHECK(hipMemset(g->zeros, 0, sizeof(int) * (3 * (size_t)nhist + 1)));
    KN_CHECK(hipMemset(g->status, 0, sizeof(int) * (3 * (size_t)nhist + 1)));
    KN_CHECK(hipMemset(g->spec0, 0, sizeof(int)));
    g->sp = group_alloc<SpeciesDev>(nhist);
    {
        std::vector<SpeciesDev> sd(nhist);
        for (int i = 0; i < nhist; i++) sd[i] = species_dev(g->h[i].spec);
        KN_CHECK(hipMemcpy(g->sp, sd.data(), sizeof(SpeciesDev) * nhist, hipMemcpyHostToDevice));
    }
    g->meta = group_alloc<GroupMeta>(nhist);
    group_plan(g);

    for (int s = 0; s < kGroupStreams; s++) {
        KN_CHECK(hipStreamCreateWithFlags(&g->st[s], hipStreamNonBlocking));
        KN_CHECK(hipEventCreateWithFlags(&g->join[s], hipEventDisableTiming));
    }
    KN_CHECK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
    group_schedule(g);
    *out = g;
    return 0;
}

// Same views and layout as kn_batch_views / kn_batch_layout.
extern "C" int kn_group_views(kn_group *g, double **scalefact, double **delta_tot, double **wavenum,
                              double **delta_nu_init, double **delta_nu, double **gbar, double **dtot_bar,
                              double **init_bar, double **mnu_bar)
{
    if (scalefact) *scalefact = g->sx;
    if (delta_tot) *delta_tot = g->Y;
    if (wavenum) *wavenum = g->wavenum;
    if (delta_nu_init) *delta_nu_init = g->init;
    if (delta_nu) *delta_nu = g->out;
    if (gbar) *gbar = g->gbar;
    if (dtot_bar) *dtot_bar = g->dbar;
    if (init_bar) *init_bar = g->ibar;
    if (mnu_bar) *mnu_bar = g->mbar;
    return 0;
}

extern "C" int kn_group_layout(const kn_group *g, long long *off_k, long long *off_y, long long *off_s)
{
    for (int i = 0; i < g->nhist; i++) {
        if (off_k) off_k[i] = g->h[i].off_k;
        if (off_y) off_y[i] = g->h[i].off_y;
        if (off_s) off_s[i] = g->h[i].off_s;
    }
    if (off_k) off_k[g->nhist] = g->nk_tot;
    if (off_y) off_y[g->nhist] = g->y_tot;
    if (off_s) off_s[g->nhist] = g->s_tot;
    return 0;
}

extern "C" int kn_group_solve(kn_group *g, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    group_launch(g, st, true, false);
    KN_CHECK(hipStreamSynchronize(st));
    return group_status(g);
}

extern "C" int kn_group_adjoint(kn_group *g, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    group_launch(g, st, false, true);
    KN_CHECK(hipStreamSynchronize(st));
    return group_status(g);
}

// One more stored step per history: d_loga[nhist] and d_rows[items] in device
// memory.  The device checks every history first (capacity, the H table
// range, a strictly increasing log a: the reference's spline would reject
// equal or decreasing knots, GSL_EINVAL -> 2001) and computes the new
// fslength(log T0, log a) for the initial-condition factor; the host reads the
// codes, and only when all are 0 does the device store the new column and
// update each history's derived state (k_group_step).  Launch sizes follow
// the new stored steps (unless the group is on its capacity plan).
extern "C" int kn_group_append(kn_group *g, const double *d_loga, const double *d_rows, void *stream)
{
    if (!g || !d_loga || !d_rows) return KN_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int n = g->nhist;
    group_pull(g);
    hipLaunchKernelGGL(k_group_step, dim3(n), dim3(64), 0, st, g->desc, g->sp, g->meta, d_loga, d_rows, g->step_fsl,
                       g->step_err, g->step_sticky, 1);
    KN_CHECK(hipGetLastError());
    std::vector<int> err(n);
    KN_CHECK(hipMemcpyAsync(err.data(), g->step_err, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    KN_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < n; i++)
        if (err[i]) {
            if (err[i] == KN_ERR_RANGE) {
                g->cap_err = 1;
                kn_set_last_error("kn_group_append: history %d: capacity (na_max %d) or H table range exceeded", i,
                                  g->h[i].spec.na_max);
            } else if (err[i] == KN_ERR_ARG) {
                kn_set_last_error("kn_group_append: history %d: log a does not follow the last stored step", i);
            } else {
                kn_set_last_error("kn_group_append: history %d: QAG failed in fslength", i);
            }
            return err[i];
        }
    hipLaunchKernelGGL(k_group_step, dim3(n), dim3(64), 0, st, g->desc, g->sp, g->meta, d_loga, d_rows, g->step_fsl,
                       g->step_err, g->step_sticky, 2);
    KN_CHECK(hipGetLastError());
    KN_CHECK(hipStreamSynchronize(st));
    g->host_stale = 1;
    group_pull(g);
    if (!g->cap_plan) {
        group_plan(g);
        group_schedule(g);
        g->gen++;
    }
    return 0;
}

// The same step with no host synchronisation, for groups on their capacity
// plan (so launch sizes need not follow it): check, fslength and store in
// one launch per history; a history whose check fails is left as it was and
// its code is kept for kn_group_status.
extern "C" int kn_group_append_async(kn_group *g, const double *d_loga, const double *d_rows, void *stream)
{
    if (!g || !d_loga || !d_rows || !g->cap_plan) {
        kn_set_last_error("kn_group_append_async: needs device arrays and a group on its capacity plan");
        return KN_ERR_ARG;
    }
    hipLaunchKernelGGL(k_group_step, dim3(g->nhist), dim3(64), 0, (hipStream_t)stream, g->desc, g->sp, g->meta, d_loga,
                       d_rows, g->step_fsl, g->step_err, g->step_sticky, 3);
    KN_CHECK(hipGetLastError());
    g->host_stale = 1;
    return 0;
}

// Launch sizes for na_max (on = 1) or for the stored steps (0, the default).
// On the capacity plan every history's grids cover na_max stored steps and
// both the integral's and the IC-only phases, and the kernels read from
// device memory how much applies, so appends (kn_group_append_async, or an
// append captured in the graph) leave the launches valid; it runs merged
// launches only.
extern "C" int kn_group_set_capacity_plan(kn_group *g, int on)
{
    if (!g || (on != 0 && on != 1)) return KN_ERR_ARG;
    group_pull(g);
    g->cap_plan = on;
    if (on) g->nstream = 0;
    group_plan(g);
    group_schedule(g);
    g->gen++;
    return 0;
}

// Capture [append +] solve [+ adjoint] on stream (not the null stream) into
// one graph.  Without append (d_loga, d_rows NULL) launch sizes follow the
// stored steps, so the graph is replayable until the next append (new
// masses and new delta_tot/wavenum/init values in the device arrays are read
// at replay).  With d_loga and d_rows the group goes on its capacity plan and
// every replay first appends the step found in those arrays at replay time
// (kn_group_append_async's launch), so one graph serves every step until a
// history is full.
extern "C" int kn_group_graph_capture(kn_group *g, int with_adjoint, const double *d_loga, const double *d_rows,
                                      void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (!g || !st || (!d_loga) != (!d_rows)) {
        kn_set_last_error("kn_group_graph_capture: needs a non-null stream, and d_loga and d_rows together");
        return KN_ERR_ARG;
    }
    const bool append = d_loga != nullptr;
    if (append && !g->cap_plan) {
        const int rc = kn_group_set_capacity_plan(g, 1);
        if (rc) return rc;
    }
    group_pull(g);
    if (g->have_graph) {
        KN_CHECK(hipGraphExecDestroy(g->exec));
        KN_CHECK(hipGraphDestroy(g->graph));
        g->have_graph = 0;
    }
    KN_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    if (append)
        hipLaunchKernelGGL(k_group_step, dim3(g->nhist), dim3(64), 0, st, g->desc, g->sp, g->meta, d_loga, d_rows,
                           g->step_fsl, g->step_err, g->step_sticky, 3);
    group_launch(g, st, true, with_adjoint != 0);
    KN_CHECK(hipStreamEndCapture(st, &g->graph));
    KN_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->have_graph = 1;
    g->graph_appends = append;
    g->graph_gen = g->gen;
    return 0;
}

extern "C" int kn_group_graph_launch(kn_group *g, void *stream)
{
    if (!g->have_graph || g->graph_gen != g->gen) {
        kn_set_last_error("kn_group_graph_launch: no graph for the current stored steps (capture after append)");
        return KN_ERR_ARG;
    }
    KN_CHECK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    if (g->graph_appends) g->host_stale = 1;
    return 0;
}

extern "C" int kn_group_status(kn_group *g) { return group_status(g); }

// History h's final QAG subdivisions of the last solve (the ones the adjoint
// holds fixed), in the interval-list order of QUADPACK's workspace:
// ivl[(k * 200 + t) * 2 + {0, 1}] for t < nint[k] (host arrays of nk * 400
// doubles and nk ints).  Items that only take the initial-condition piece
// have nint 0.
extern "C" int kn_group_subdivision(kn_group *g, int h, double *ivl, int *nint)
{
    if (!g || h < 0 || h >= g->nhist || !ivl || !nint) return KN_ERR_ARG;
    group_pull(g);
    const GroupHist &q = g->h[h];
    const int nk = q.spec.nk;
    if (!q.integ) {
        memset(nint, 0, sizeof(int) * nk);
        return 0;
    }
    KN_CHECK(hipMemcpy(ivl, g->ivl + (size_t)q.off_k * kLimit * 2, sizeof(double) * (size_t)nk * kLimit * 2,
                       hipMemcpyDeviceToHost));
    KN_CHECK(hipMemcpy(nint, g->nint + q.off_k, sizeof(int) * nk, hipMemcpyDeviceToHost));
    return 0;
}

// New neutrino mass (eV) of history h; qc, the tolerance and the IC-only
// state follow it as get_delta_nu derives them.  When only the device species
// record changes a captured graph stays valid; when the history starts or
// stops integrating (a massless history given a mass) the launch plan changes
// and a captured graph is stale.
extern "C" int kn_group_set_mass(kn_group *g, int h, double mnu)
{
    if (h < 0 || h >= g->nhist || !(mnu > 0)) return KN_ERR_ARG;
    group_pull(g);
    GroupHist &q = g->h[h];
    const int integ0 = q.integ;
    const double vcrit0 = q.vcrit;
    q.spec.mnu = mnu;
    group_derive(q);
    const SpeciesDev sd = species_dev(q.spec);
    KN_CHECK(hipMemcpy(g->sp + h, &sd, sizeof sd, hipMemcpyHostToDevice));
    group_plan(g);  // the descriptors carry the mass for the device append
    if (!g->cap_plan && (q.integ != integ0 || q.vcrit != vcrit0)) {
        group_schedule(g);
        g->gen++;
    }
    return 0;
}

// History h's stored steps to a host table and back (the table's own layout:
// rows delta_tot[k], scalefact, delta_nu_init, wavenum), so that the
// reference's state files (save_all_nu_state / read_all_nu_state,
// delta_tot_table.c:254-352) and flat blocks (get_nu_state / set_nu_state,
// interface_common.c:150-180) checkpoint a device-resident group.  Export
// also gives delta_nu_last the last solve's delta_nu.
extern "C" int kn_group_export(kn_group *g, int h, _delta_tot_table *d)
{
    if (!g || !d || h < 0 || h >= g->nhist) return KN_ERR_ARG;
    group_pull(g);
    const GroupHist &q = g->h[h];
    const int nk = q.spec.nk, na = q.na, na_max = q.spec.na_max;
    if (d->nk_allocated < nk || d->namax < na || !d->scalefact || !d->delta_tot || !d->delta_nu_init ||
        !d->wavenum) {
        kn_set_last_error("kn_group_export: the table holds fewer than %d k or %d steps", nk, na);
        return KN_ERR_ARG;
    }
    KN_CHECK(hipDeviceSynchronize());
    std::vector<double> rows((size_t)nk * na);
    KN_CHECK(hipMemcpy2D(rows.data(), sizeof(double) * na, g->Y + q.off_y, sizeof(double) * na_max,
                         sizeof(double) * na, nk, hipMemcpyDeviceToHost));
    for (int k = 0; k < nk; k++) memcpy(d->delta_tot[k], rows.data() + (size_t)k * na, sizeof(double) * na);
    KN_CHECK(hipMemcpy(d->scalefact, g->sx + q.off_s, sizeof(double) * na, hipMemcpyDeviceToHost));
    KN_CHECK(hipMemcpy(d->delta_nu_init, g->init + q.off_k, sizeof(double) * nk, hipMemcpyDeviceToHost));
    KN_CHECK(hipMemcpy(d->wavenum, g->wavenum + q.off_k, sizeof(double) * nk, hipMemcpyDeviceToHost));
    if (d->delta_nu_last)
        KN_CHECK(hipMemcpy(d->delta_nu_last, g->out + q.off_k, sizeof(double) * nk, hipMemcpyDeviceToHost));
    d->nk = nk;
    d->ia = na;
    return 0;
}

// A host table's stored steps into history h (same nk, at most na_max steps),
// integrated to a from then on: the history's derived state (tolerance, qc,
// IC-only, the initial-condition factors) follows as after an append.
extern "C" int kn_group_import(kn_group *g, int h, const _delta_tot_table *d, double a)
{
    if (!g || !d || h < 0 || h >= g->nhist) return KN_ERR_ARG;
    group_pull(g);
    GroupHist &q = g->h[h];
    const int nk = q.spec.nk, na = d->ia, na_max = q.spec.na_max;
    if (d->nk != nk || na < 1 || na > na_max || !d->scalefact || !d->delta_tot || !d->delta_nu_init ||
        !d->wavenum) {
        kn_set_last_error("kn_group_import: history %d: the table has %d k and %d steps (the history %d k, at most "
                          "%d steps)", h, d->nk, na, nk, na_max);
        return KN_ERR_ARG;
    }
    const double xa = log(a);
    for (int i = 1; i < na; i++)
        if (!(d->scalefact[i] > d->scalefact[i - 1])) {
            kn_set_last_error("kn_group_import: history %d: stored log a not increasing at step %d", h, i);
            return KN_ERR_ARG;
        }
    if (!(a > q.spec.time_transfer) || !(xa >= d->scalefact[na - 1]) || !(xa <= q.xmax)) {
        kn_set_last_error("kn_group_import: history %d: a = %g before the last stored step or outside the H table",
                          h, a);
        return KN_ERR_RANGE;
    }
    KN_CHECK(hipDeviceSynchronize());
    std::vector<double> rows((size_t)nk * na);
    for (int k = 0; k < nk; k++) memcpy(rows.data() + (size_t)k * na, d->delta_tot[k], sizeof(double) * na);
    KN_CHECK(hipMemcpy2D(g->Y + q.off_y, sizeof(double) * na_max, rows.data(), sizeof(double) * na,
                         sizeof(double) * na, nk, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(g->sx + q.off_s, d->scalefact, sizeof(double) * na, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(g->init + q.off_k, d->delta_nu_init, sizeof(double) * nk, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(g->wavenum + q.off_k, d->wavenum, sizeof(double) * nk, hipMemcpyHostToDevice));
    const GroupHist keep = q;
    q.na = na;
    q.spec.na = na;
    q.spec.a = a;
    q.xa = xa;
    group_derive(q);
    group_set_integ(q);
    const int rc = group_refresh_ic(q);
    if (rc) {
        q = keep;
        return rc;
    }
    const SpeciesDev sd = species_dev(q.spec);
    KN_CHECK(hipMemcpy(g->sp + h, &sd, sizeof sd, hipMemcpyHostToDevice));
    group_plan(g);
    if (!g->cap_plan) {
        group_schedule(g);
        g->gen++;
    }
    return 0;
}

// Pick the launch strategy by timing the forward: merged launches (0) or
// per-history launch sequences on 1, 2, 4 or 8 streams.
extern "C" int kn_group_autotune(kn_group *g, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    group_pull(g);
    if (g->cap_plan) return group_status(g);  // merged launches only
    const int cand[] = {0, 1, 2, 4, 8};
    float best = 1e30f;
    int bs = g->nstream;
    hipEvent_t e0, e1;
    KN_CHECK(hipEventCreate(&e0));
    KN_CHECK(hipEventCreate(&e1));
    for (int c : cand) {
        g->nstream = c;
        group_schedule(g);
        group_launch(g, st, true, false);
        KN_CHECK(hipEventRecord(e0, st));
        group_launch(g, st, true, false);
        KN_CHECK(hipEventRecord(e1, st));
        KN_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        KN_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bs = c;
        }
    }
    KN_CHECK(hipEventDestroy(e0));
    KN_CHECK(hipEventDestroy(e1));
    g->nstream = bs;
    group_schedule(g);
    g->gen++;  // the stream assignment changed: a captured graph is stale
    return group_status(g);
}

extern "C" int kn_group_streams(const kn_group *g) { return g->nstream; }

extern "C" int kn_group_set_streams(kn_group *g, int nstream)
{
    if (!g || nstream < 0 || nstream > kGroupStreams || (g->cap_plan && nstream)) return KN_ERR_ARG;
    group_pull(g);
    g->nstream = nstream;
    if (nstream) group_schedule(g);
    g->gen++;  // a captured graph is stale
    return 0;
}

// The parameters get_delta_nu (delta_tot_table.c:507-555) derives from a table,
// a and mnu, as one group history (na = na_max = ia; hubble = kn_hubble_ctx).
extern "C" double kn_hubble_ctx(double a, void *) { return kn_hubble(a); }

extern "C" int kn_batch_history_from_table(const _delta_tot_table *d, double a, double mnu, kn_batch_history *out)
{
    if (!d || !out || d->ia < 1) return KN_ERR_ARG;
    kn_batch_history hh;
    memset(&hh, 0, sizeof hh);
    hh.nk = d->nk;
    hh.na = d->ia;
    hh.na_max = d->ia;
    hh.a = a;
    hh.a_max = a;
    hh.time_transfer = d->TimeTransfer;
    hh.mnu = mnu;
    hh.kBtnu = d->omnu->kBtnu;
    hh.light = d->light;
    hh.delta_nu_prefac = d->delta_nu_prefac;
    hh.nufrac_low = d->omnu->hybnu.nufrac_low[0];
    hh.vcrit = d->omnu->hybnu.vcrit;
    hh.hybrid = d->omnu->hybnu.enabled;
    hh.nu_crit_time = d->omnu->hybnu.nu_crit_time;
    kn_species s;
    kn_species_for(d, a, mnu, &s);
    hh.qc = s.qc;
    hh.relerr = s.relerr;
    hh.ic_only = s.ic_only;
    hh.hubble = g_bg ? kn_background_hubble_ctx : kn_hubble_ctx;
    hh.hubble_ctx = (void *)g_bg;
    *out = hh;
    return 0;
}
