// batch.hip -- many free-streaming integrals in one launch set, and the adjoint.
//
// A batch holds H histories, each with its own cosmology (Hubble function),
// neutrino mass, stored steps Na_h and k bins nk_h; each history is what one
// get_delta_nu call of the reference integrates (delta_tot_table.c:507-611).
// Everything per history lives in device memory (HistDev): sizes, offsets and
// parameters are read by the kernels at run time, so one captured hipGraph
// stays valid while Na_h grows by appended steps.
//
// Forward, per launch set:
//   k_b_plan       per-history work counts and exclusive prefix sums (device);
//   k_b_fs         fslength nodes, one thread per node: the 61-point Gauss-Kronrod
//                  estimate in GSL's summation order; nodes whose first estimate
//                  is not accepted are queued for k_b_fs_qag (full QAG, one wavefront);
//   k_b_spline_*   natural-spline coefficients (fslength tables, delta_tot rows);
//   k_b_bucket     bucket index over each history's stored log a;
//   k_b_qag        persistent wavefronts pull (history, k) items from a queue
//                  ordered by cost; each runs QAG-GK61 and records its final
//                  subdivision; the initial-condition piece is added per item.
// Adjoint (kn_batch_adjoint): with the recorded subdivision fixed, the result
// is linear in delta_tot through the spline; k_b_adj_nodes accumulates the
// node weights onto y and on the spline coefficients c, k_b_adj_rows maps
// c-bar back through the transposed natural-spline solve, and the mass
// derivative uses the analytic specialJ derivatives.
#include <math.h>
#include <string.h>
#include <algorithm>
#include <numeric>
#include <vector>

#include "dnu_device.h"
#include "kn_hip.h"

using namespace kn_dnu;


namespace {

constexpr int kBW = 4;        // wavefronts per block in the wavefront kernels
constexpr int kChunkB = 32;   // spline rows per thread
constexpr double kHWidth = 1.0 / 128.0;
constexpr int kFsPerStep = 16;

struct HistDev {
    int nk, na, na_max, ic_only;
    long long off_k, off_y, off_s, off_fs, off_b;
    double x0, xa, light, prefac, mnubykT, relerr, deriv_prefac, vcrit, kBtnu;
    JPar J;
    HTab H;
};

// counts/prefix kinds
enum { W_FS = 0, W_FSCH, W_SCH, W_B, W_NKIND };

struct BatchDev {
    HistDev *hist;
    long long *prefix;   // W_NKIND * (nhist + 1)
    int *ctr;            // work counters: [0] fs fallback list size, [1] fs fallback next, [2] qag next
    int *fs_fallback;    // node ids (global fs index) needing a full QAG
    int *status;         // [0] qag failure, [1] domain error, [2] capacity (sticky)
};

__device__ __forceinline__ int find_hist(const long long *prefix, int nhist, long long idx)
{
    int lo = 0, hi = nhist;  // prefix[lo] <= idx < prefix[hi]
    while (hi > lo + 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= idx) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Work-queue pop for a whole wavefront: lane 0 takes the next slot and the
// value is made wave-uniform (scalar), so everything indexed by it is too.
__device__ __forceinline__ int next_slot(int *ctr, int lane)
{
    int v = 0;
    if (lane == 0) v = atomicAdd(ctr, 1);
    return __builtin_amdgcn_readfirstlane(v);
}

__global__ void k_b_plan(BatchDev B, int nhist)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        long long acc[W_NKIND] = {0, 0, 0, 0};
        for (int h = 0; h < nhist; h++) {
            const HistDev &d = B.hist[h];
            const int na = d.na;
            const bool integ = !d.ic_only && na >= 2;
            // node 0 of the fslength table (fsl_A0a) is needed by the initial-condition piece always
            const long long c[W_NKIND] = {integ ? (long long)kFsPerStep * na : 1,
                                          integ ? (long long)((kFsPerStep * na + kChunkB - 1) / kChunkB) : 0,
                                          integ && na > 2 ? (long long)d.nk * ((na + kChunkB - 1) / kChunkB) : 0,
                                          integ ? (long long)(4 * na + 1) : 0};
            for (int w = 0; w < W_NKIND; w++) {
                B.prefix[(size_t)w * (nhist + 1) + h] = acc[w];
                acc[w] += c[w];
            }
        }
        for (int w = 0; w < W_NKIND; w++) B.prefix[(size_t)w * (nhist + 1) + nhist] = acc[w];
        B.ctr[0] = 0;
        B.ctr[1] = 0;
        B.ctr[2] = 0;
        B.status[0] = 0;
        B.status[1] = 0;
    }
}

// One 61-point Gauss-Kronrod estimate by one thread, GSL qk summation order.
__device__ inline void gk61_thread(const FslIntegrand &f, double a, double b, double &res, double &err,
                                   double &resabs, double &resasc)
{
    double lo[31], hi[31];
    const double c = 0.5 * (a + b), h = 0.5 * (b - a);
    const double fc = f(c);
    double sg = 0, sk = fc * c_wgk[30], sa = fabs(sk);
#pragma unroll
    for (int j = 0; j < 15; j++) {
        const int q = 2 * j + 1;
        const double dx = h * c_xgk[q];
        const double f1 = f(c - dx), f2 = f(c + dx);
        lo[q] = f1;
        hi[q] = f2;
        sg += c_wg[j] * (f1 + f2);
        sk += c_wgk[q] * (f1 + f2);
        sa += c_wgk[q] * (fabs(f1) + fabs(f2));
    }
#pragma unroll
    for (int j = 0; j < 15; j++) {
        const int q = 2 * j;
        const double dx = h * c_xgk[q];
        const double f1 = f(c - dx), f2 = f(c + dx);
        lo[q] = f1;
        hi[q] = f2;
        sk += c_wgk[q] * (f1 + f2);
        sa += c_wgk[q] * (fabs(f1) + fabs(f2));
    }
    const double mean = 0.5 * sk;
    double sasc = c_wgk[30] * fabs(fc - mean);
#pragma unroll
    for (int j = 0; j < 30; j++) sasc += c_wgk[j] * (fabs(lo[j] - mean) + fabs(hi[j] - mean));
    res = sk * h;
    resabs = sa * fabs(h);
    resasc = sasc * fabs(h);
    err = gk_rescale((sk - sg) * h, resabs, resasc);
}

__global__ __launch_bounds__(256) void k_b_fs(BatchDev B, int nhist, double *__restrict__ fx, double *__restrict__ fy)
{
    const long long *pre = B.prefix + (size_t)W_FS * (nhist + 1);
    const long long total = pre[nhist];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = B.hist[h];
        const int nfs = kFsPerStep * d.na;
        const int i = (int)(idx - pre[h]);
        // delta_tot_table.c:582
        const double xi = d.x0 + i * (d.xa - d.x0) / (nfs - 1.);
        double v = 0.0;
        if (xi < d.xa) {
            FslIntegrand f;
            f.H = d.H;
            double r, e, ra, rs;
            gk61_thread(f, xi, d.xa, r, e, ra, rs);
            const double tol = 1e-6 * fabs(r);
            if (e <= 50 * DBL_EPSILON * ra && e > tol) {
                atomicOr(&B.status[0], 1);                     // GSL_EROUND on the first estimate
            } else if (!((e <= tol && e != rs) || e == 0.0)) {
                const int slot = atomicAdd(&B.ctr[0], 1);      // needs bisection: full QAG later
                B.fs_fallback[slot] = (int)idx;
            }
            v = d.light * r;
        }
        fx[d.off_fs + i] = xi;
        fy[d.off_fs + i] = v;
    }
}

__global__ __launch_bounds__(64 * kBW, 2) void k_b_fs_qag(BatchDev B, const HistDev *__restrict__ hist, int nhist,
                                                           const double *__restrict__ fx, double *__restrict__ fy)
{
    __shared__ QagWs ws[kBW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long *pre = B.prefix + (size_t)W_FS * (nhist + 1);
    const int nlist = B.ctr[0];
    for (;;) {
        const int slot = next_slot(&B.ctr[1], lane);
        if (slot >= nlist) break;
        const long long idx = B.fs_fallback[slot];
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = hist[h];
        const int i = (int)(idx - pre[h]);
        FslIntegrand f;
        f.H = d.H;
        double r;
        int ni;
        const int rc = wave_qag61(f, fx[d.off_fs + i], d.xa, 1e-6, kLimit, ws[wave], &r, &ni);
        if (lane == 0) {
            if (rc) atomicOr(&B.status[0], 1);
            fy[d.off_fs + i] = d.light * r;
        }
    }
}

__global__ void k_b_spline_fs(BatchDev B, int nhist, const double *__restrict__ fx, const double *__restrict__ fy,
                              double *__restrict__ fc)
{
    const long long *pre = B.prefix + (size_t)W_FSCH * (nhist + 1);
    const long long total = pre[nhist];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = B.hist[h];
        const int n = kFsPerStep * d.na, ch = (int)(idx - pre[h]);
        const int lo = ch * kChunkB, hi = lo + kChunkB < n ? lo + kChunkB : n;
        spline_window(fx + d.off_fs, fy + d.off_fs, n, lo, hi, fc + d.off_fs);
    }
}

__global__ void k_b_spline_rows(BatchDev B, int nhist, const double *__restrict__ sx, const double *__restrict__ Y,
                                double *__restrict__ Cc)
{
    const long long *pre = B.prefix + (size_t)W_SCH * (nhist + 1);
    const long long total = pre[nhist];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = B.hist[h];
        const int nch = (d.na + kChunkB - 1) / kChunkB;
        const long long t = idx - pre[h];
        const int k = (int)(t / nch), ch = (int)(t % nch);
        const int lo = ch * kChunkB, hi = lo + kChunkB < d.na ? lo + kChunkB : d.na;
        const size_t row = (size_t)d.off_y + (size_t)k * d.na_max;
        spline_window(sx + d.off_s, Y + row, d.na, lo, hi, Cc + row);
    }
}

__global__ void k_b_bucket(BatchDev B, int nhist, const double *__restrict__ sx, int *__restrict__ bucket)
{
    const long long *pre = B.prefix + (size_t)W_B * (nhist + 1);
    const long long total = pre[nhist];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = B.hist[h];
        const int q = (int)(idx - pre[h]), nb = 4 * d.na;
        const double *x = sx + d.off_s;
        const double inv_bw = nb / (x[d.na - 1] - x[0]);
        const double edge = x[0] + q / inv_bw;
        int lo = 0, hi = d.na - 1;
        while (hi > lo + 1) {
            const int mid = (lo + hi) >> 1;
            if (x[mid] > edge) hi = mid;
            else lo = mid;
        }
        bucket[d.off_b + q] = lo;
    }
}

struct QagArgs {
    const double *fx, *fy, *fc, *sx, *Y, *C, *wavenum, *init;
    const int *bucket, *order, *item_hist;
    double *out, *ivl;  // ivl: per item kLimit (lo, hi) pairs of the final subdivision
    int *nint;
    int nitems;
};

// The integrand of delta_tot_table.c:492-500 for one (history, k) item: the
// same arithmetic as DnuIntegrand, with the history's H table and specialJ
// parameters read through the (wave-uniform) descriptor pointer.
struct BatchIntegrand {
    const HistDev *hd;
    const double *fx, *fy, *fc;  // fslength spline, uniform nodes
    const double *sx, *sy, *sc;  // delta_tot spline of this k
    const int *sbucket;
    int nfs, ns, nsb, linear;
    double fx0, finv_h, sx0, sinv_bw, k, mnubykT;
    int *edom;
    __device__ __forceinline__ double operator()(double x) const
    {
        if (x < fx[0] || x > fx[nfs - 1] || x < sx[0] || x > sx[ns - 1]) {
            atomicOr(edom, 1);  // GSL_EDOM in gsl_interp_eval -> terminate(2001)
            return 0.0;
        }
        const int fi = settle(fx, nfs, (int)((x - fx0) * finv_h), x);
        const double fsl = cspline_at(fx, fy, fc, fi, x);
        int q = (int)((x - sx0) * sinv_bw);
        q = q < 0 ? 0 : (q > nsb ? nsb : q);
        const int si = settle(sx, ns, sbucket[q], x);
        const double dtot = linear ? linear_at(sx, sy, si, x) : cspline_at(sx, sy, sc, si, x);
        const double specJ = specialJ(k * fsl / mnubykT, hd->J);
        return fsl * inv_ah(hd->H, x) * specJ * dtot;  // fsl / (a H) specJ dtot
    }
};

__device__ __forceinline__ BatchIntegrand make_integrand(const HistDev *hd, const QagArgs &q, int k, int *edom)
{
    const HistDev &d = *hd;
    BatchIntegrand f;
    const int nfs = kFsPerStep * d.na;
    f.hd = hd;
    f.fx = q.fx + d.off_fs;
    f.fy = q.fy + d.off_fs;
    f.fc = q.fc + d.off_fs;
    f.nfs = nfs;
    f.fx0 = f.fx[0];
    f.finv_h = (nfs - 1) / (f.fx[nfs - 1] - f.fx[0]);
    f.sx = q.sx + d.off_s;
    f.sy = q.Y + d.off_y + (size_t)k * d.na_max;
    f.sc = q.C + d.off_y + (size_t)k * d.na_max;
    f.sbucket = q.bucket + d.off_b;
    f.ns = d.na;
    f.nsb = 4 * d.na;
    f.linear = d.na <= 2;
    f.sx0 = f.sx[0];
    f.sinv_bw = f.nsb / (f.sx[d.na - 1] - f.sx[0]);
    f.k = q.wavenum[d.off_k + k];
    f.mnubykT = d.mnubykT;
    f.edom = edom;
    return f;
}

// IC piece of delta_tot_table.c:533-534 with fsl_A0a = the fslength node 0.
__device__ __forceinline__ double ic_piece(const HistDev &d, double kk, double init, double fsl0)
{
    JPar J0;
    J0.qc = 0.0;
    J0.nufrac_low = d.J.nufrac_low;
    const double m = d.mnubykT > 0 ? d.mnubykT : 1;
    return specialJ(kk * fsl0 / m, J0) * init * (1. + d.deriv_prefac * fsl0);
}

// At most 256 VGPRs (two waves per SIMD): the queue kernels must stay within
// the arch register file.
__global__ __launch_bounds__(64 * kBW, 2) void k_b_qag(BatchDev B, const HistDev *__restrict__ hist, int nhist,
                                                      QagArgs q)
{
    __shared__ QagWs ws[kBW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (;;) {
        const int slot = next_slot(&B.ctr[2], lane);
        if (slot >= q.nitems) break;
        const int item = q.order[slot];
        const int h = q.item_hist[item];
        const HistDev &d = hist[h];
        const int k = (int)(item - d.off_k);
        const bool integ = !d.ic_only && d.na >= 2;
        const double fsl0 = q.fy[d.off_fs];
        double r = 0.0;
        int ni = 0;
        if (integ) {
            const BatchIntegrand f = make_integrand(&hist[h], q, k, &B.status[1]);
            const int rc = wave_qag61(f, d.x0, d.xa, d.relerr, kLimit, ws[wave], &r, &ni);
            if (lane == 0 && rc) atomicOr(&B.status[0], 2);
            // record the final subdivision for the adjoint
            for (int t = lane; t < ni; t += 64) {
                q.ivl[((size_t)item * kLimit + t) * 2] = ws[wave].lo[t];
                q.ivl[((size_t)item * kLimit + t) * 2 + 1] = ws[wave].hi[t];
            }
        }
        if (lane == 0) {
            const double ic = ic_piece(d, q.wavenum[item], q.init[item], fsl0);
            q.out[item] = integ ? ic + d.prefac * r : ic;
            q.nint[item] = ni;
        }
    }
}

// ---------------------------------------------------------------- adjoint

struct AdjArgs {
    const double *gbar;   // per item
    double *ybar, *cbar;  // like Y (row stride na_max)
    double *init_bar;     // per item
    double *mu_bar;       // per item: d(result)/d(mnu/kT) * gbar
};

__global__ __launch_bounds__(64 * kBW, 2) void k_b_adj_nodes(BatchDev B, const HistDev *__restrict__ hist,
                                                            int nhist, QagArgs q, AdjArgs g)
{
    const int lane = threadIdx.x & 63;
    const int hf = lane >> 5, j = lane & 31;
    for (;;) {
        const int slot = next_slot(&B.ctr[2], lane);
        if (slot >= q.nitems) break;
        const int item = q.order[slot];
        const int h = q.item_hist[item];
        const HistDev &d = hist[h];
        const int k = (int)(item - d.off_k);
        const double gb = g.gbar[item];
        const bool integ = !d.ic_only && d.na >= 2;
        double mu_acc = 0.0;
        if (integ) {
            const BatchIntegrand f = make_integrand(&hist[h], q, k, &B.status[1]);
            const int ni = q.nint[item];
            double *yb = g.ybar + d.off_y + (size_t)k * d.na_max;
            double *cb = g.cbar + d.off_y + (size_t)k * d.na_max;
            for (int t0 = 0; t0 < ni; t0 += 2) {
                const int t = t0 + hf;
                if (t < ni && j <= 30) {
                    const double a = q.ivl[((size_t)item * kLimit + t) * 2];
                    const double b = q.ivl[((size_t)item * kLimit + t) * 2 + 1];
                    const double c = 0.5 * (a + b), hl = 0.5 * (b - a);
                    const int npt = j < 30 ? 2 : 1;
                    for (int side = 0; side < npt; side++) {
                        const double x = j < 30 ? (side ? c + hl * c_xgk[j] : c - hl * c_xgk[j]) : c;
                        const double w = c_wgk[j] * hl * d.prefac * gb;
                        // integrand pieces (delta_tot_table.c:492-500) without S_k
                        const int fi = settle(f.fx, f.nfs, (int)((x - f.fx0) * f.finv_h), x);
                        const double fsl = cspline_at(f.fx, f.fy, f.fc, fi, x);
                        int qq = (int)((x - f.sx0) * f.sinv_bw);
                        qq = qq < 0 ? 0 : (qq > f.nsb ? f.nsb : qq);
                        const int si = settle(f.sx, f.ns, f.sbucket[qq], x);
                        const double kx = f.k * fsl / f.mnubykT;
                        double Jx, Jq;
                        const double J = specialJ_d(kx, d.J, Jx, Jq);
                        const double base = fsl * inv_ah(d.H, x);  // fsl / (a H)
                        const double S = f.linear ? linear_at(f.sx, f.sy, si, x) : cspline_at(f.sx, f.sy, f.sc, si, x);
                        // d/dmu of J(k fsl/mu, qc = vcrit mu)
                        mu_acc += w * base * S * (Jx * (-kx / f.mnubykT) + Jq * d.vcrit);
                        const double wj = w * base * J;
                        if (f.linear) {
                            const double t1 = (x - f.sx[si]) / (f.sx[si + 1] - f.sx[si]);
                            atomicAdd(&yb[si], wj * (1.0 - t1));
                            atomicAdd(&yb[si + 1], wj * t1);
                        } else {
                            double wy0, wy1, wc0, wc1;
                            cspline_grad(f.sx, si, x, wy0, wy1, wc0, wc1);
                            atomicAdd(&yb[si], wj * wy0);
                            atomicAdd(&yb[si + 1], wj * wy1);
                            atomicAdd(&cb[si], wj * wc0);
                            atomicAdd(&cb[si + 1], wj * wc1);
                        }
                    }
                }
            }
        }
        // wave sum of the mass derivative
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) mu_acc += __shfl_xor(mu_acc, m, 64);
        if (lane == 0) {
            const double fsl0 = q.fy[d.off_fs];
            double ibar = 0.0, mbar = mu_acc;
            {
                JPar J0;
                J0.qc = 0.0;
                J0.nufrac_low = d.J.nufrac_low;
                const double m = d.mnubykT > 0 ? d.mnubykT : 1;
                const double kk = q.wavenum[item] * fsl0 / m;
                double Jx, Jq;
                const double J = specialJ_d(kk, J0, Jx, Jq);
                const double lin = 1. + d.deriv_prefac * fsl0;
                ibar = gb * J * lin;
                if (d.mnubykT > 0) mbar += gb * Jx * (-kk / m) * q.init[item] * lin;
            }
            g.init_bar[item] = ibar;
            g.mu_bar[item] = mbar;
        }
    }
}

// Per row: z = M^-1 cbar (transposed natural-spline solve, M symmetric),
// then ybar_j += sum_i z_i dg_i/dy_j with g_i = 3((y_{i+1}-y_i)/h_i - (y_i-y_{i-1})/h_{i-1}).
__global__ void k_b_adj_rows(BatchDev B, int nhist, const double *__restrict__ sx, const double *__restrict__ cbar,
                             double *__restrict__ zbuf, double *__restrict__ ybar, int phase)
{
    const long long *pre = B.prefix + (size_t)W_SCH * (nhist + 1);
    const long long total = pre[nhist];
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += (long long)gridDim.x * blockDim.x) {
        const int h = find_hist(pre, nhist, idx);
        const HistDev &d = B.hist[h];
        const int nch = (d.na + kChunkB - 1) / kChunkB;
        const long long t = idx - pre[h];
        const int k = (int)(t / nch), ch = (int)(t % nch);
        const int lo = ch * kChunkB, hi = lo + kChunkB < d.na ? lo + kChunkB : d.na;
        const size_t row = (size_t)d.off_y + (size_t)k * d.na_max;
        const double *x = sx + d.off_s;
        if (phase == 0) {
            tridiag_window(x, cbar + row, d.na, lo, hi, zbuf + row);
        } else {
            const double *z = zbuf + row;
            const int n = d.na;
            for (int jj = lo; jj < hi; jj++) {
                double acc = 0.0;
                // z_i nonzero only for interior i in [1, n-2]
                if (jj - 1 >= 1 && jj - 1 <= n - 2) acc += z[jj - 1] * 3.0 / (x[jj] - x[jj - 1]);
                if (jj >= 1 && jj <= n - 2) acc -= z[jj] * (3.0 / (x[jj + 1] - x[jj]) + 3.0 / (x[jj] - x[jj - 1]));
                if (jj + 1 >= 1 && jj + 1 <= n - 2) acc += z[jj + 1] * 3.0 / (x[jj + 1] - x[jj]);
                ybar[row + jj] += acc;
            }
        }
    }
}

__global__ void k_b_zero(double *p, long long n)
{
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

__global__ void k_b_reset_queue(BatchDev B)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) B.ctr[2] = 0;
}

// mnu_bar[h] = sum over the history's items of mu_bar / kBtnu
__global__ void k_b_mass_bar(BatchDev B, int nhist, const double *__restrict__ mu_bar, double *__restrict__ out)
{
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= nhist) return;
    const HistDev &d = B.hist[h];
    double s = 0.0;
    for (int k = 0; k < d.nk; k++) s += mu_bar[d.off_k + k];
    out[h] = s / d.kBtnu;
}

// Append one stored step per history: scalefact[na] = loga[h], delta_tot[k][na] = rows,
// na += 1, and the integration's upper limit moves to the new log a.
__global__ void k_b_append(BatchDev B, int nhist, const double *__restrict__ loga, const double *__restrict__ rows,
                           double *__restrict__ sx, double *__restrict__ Y)
{
    const int h = blockIdx.x;
    if (h >= nhist) return;
    HistDev &d = B.hist[h];
    const int na = d.na;
    // capacity, and the Hubble table's range (built up to a_max)
    if (na >= d.na_max || loga[h] > d.H.x0 + d.H.npanel / d.H.inv_w) {
        if (threadIdx.x == 0) atomicOr(&B.status[2], 1);
        return;
    }
    for (int k = threadIdx.x; k < d.nk; k += blockDim.x) Y[d.off_y + (size_t)k * d.na_max + na] = rows[d.off_k + k];
    __syncthreads();
    if (threadIdx.x == 0) {
        sx[d.off_s + na] = loga[h];
        d.xa = loga[h];
        d.na = na + 1;
    }
}


}  // namespace

// The process's hubble_function in the per-history callback form.
extern "C" double kn_hubble_ctx(double a, void *) { return kn_hubble(a); }

struct kn_batch {
    int nhist;
    std::vector<kn_batch_history> hist;
    std::vector<HistDev> hd;
    long long nk_tot, y_tot, s_tot, fs_tot, b_tot;
    BatchDev B;
    double *hcoef;
    double *fx, *fy, *fc, *sx, *Y, *C, *wavenum, *init, *out, *ivl, *ybar, *cbar, *zbuf, *ibar, *mubar, *mbar, *gbar;
    int *bucket, *order, *item_hist, *nint;
    int grid_threads, grid_waves, grid_rows;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int have_graph;
};

namespace {

void build_htable(const kn_batch_history &h, std::vector<double> &coef, double &x0, int &npanel)
{
    x0 = log(h.time_transfer);
    const double x1 = log(h.a_max > h.a ? h.a_max : h.a);
    npanel = (int)ceil((x1 - x0) / kHWidth) + 1;
    coef.assign((size_t)npanel * kHStride, 0.0);
    for (int p = 0; p < npanel; p++)
        h_cheb_panel(x0 + p * kHWidth, kHWidth, [&](double a) { return h.hubble(a, h.hubble_ctx); },
                     &coef[(size_t)p * kHStride]);
}

void upload_consts()
{
    static bool done = false;
    if (done) return;
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_xgk), kn_xgk61, sizeof(kn_xgk61)));
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_wgk), kn_wgk61, sizeof(kn_wgk61)));
    KN_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_wg), kn_wg30, sizeof(kn_wg30)));
    done = true;
}

template <typename T>
T *dalloc(size_t n)
{
    T *p = nullptr;
    KN_CHECK(hipMalloc(&p, sizeof(T) * (n > 0 ? n : 1)));
    return p;
}

// Resident blocks per CU x CUs, or -1 when the runtime cannot answer the query.
int occupancy_grid(const void *kernel, const char *name, int threads, int per_cu_cap)
{
    int dev = 0, cus = 0, per_cu = 0;
    KN_CHECK(hipGetDevice(&dev));
    KN_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0);
    if (e != hipSuccess) {
        // a kernel whose occupancy the runtime cannot compute is never launched
        (void)hipGetLastError();
        kn_set_last_error("occupancy query for %s failed: %s", name, hipGetErrorString(e));
        return -1;
    }
    if (per_cu < 1) per_cu = 1;
    if (per_cu_cap > 0 && per_cu > per_cu_cap) per_cu = per_cu_cap;
    return per_cu * cus;
}

void launch_forward(kn_batch *b, hipStream_t st)
{
    const int H = b->nhist;
    hipLaunchKernelGGL(k_b_plan, dim3(1), dim3(64), 0, st, b->B, H);
    hipLaunchKernelGGL(k_b_fs, dim3(b->grid_threads), dim3(256), 0, st, b->B, H, b->fx, b->fy);
    hipLaunchKernelGGL(k_b_fs_qag, dim3(b->grid_waves), dim3(64 * kBW), 0, st, b->B, b->B.hist, H, b->fx, b->fy);
    hipLaunchKernelGGL(k_b_spline_fs, dim3(b->grid_rows), dim3(128), 0, st, b->B, H, b->fx, b->fy, b->fc);
    hipLaunchKernelGGL(k_b_spline_rows, dim3(b->grid_rows), dim3(128), 0, st, b->B, H, b->sx, b->Y, b->C);
    hipLaunchKernelGGL(k_b_bucket, dim3(b->grid_rows), dim3(256), 0, st, b->B, H, b->sx, b->bucket);
    QagArgs q;
    q.fx = b->fx; q.fy = b->fy; q.fc = b->fc; q.sx = b->sx; q.Y = b->Y; q.C = b->C;
    q.wavenum = b->wavenum; q.init = b->init; q.bucket = b->bucket; q.order = b->order; q.item_hist = b->item_hist;
    q.out = b->out; q.ivl = b->ivl; q.nint = b->nint; q.nitems = (int)b->nk_tot;
    hipLaunchKernelGGL(k_b_qag, dim3(b->grid_waves), dim3(64 * kBW), 0, st, b->B, b->B.hist, H, q);
    KN_CHECK(hipGetLastError());
}

void launch_adjoint(kn_batch *b, hipStream_t st)
{
    const int H = b->nhist;
    const int zb = (int)std::min<long long>(1024, (b->y_tot + 255) / 256);
    hipLaunchKernelGGL(k_b_zero, dim3(zb), dim3(256), 0, st, b->ybar, b->y_tot);
    hipLaunchKernelGGL(k_b_zero, dim3(zb), dim3(256), 0, st, b->cbar, b->y_tot);
    hipLaunchKernelGGL(k_b_reset_queue, dim3(1), dim3(64), 0, st, b->B);
    QagArgs q;
    q.fx = b->fx; q.fy = b->fy; q.fc = b->fc; q.sx = b->sx; q.Y = b->Y; q.C = b->C;
    q.wavenum = b->wavenum; q.init = b->init; q.bucket = b->bucket; q.order = b->order; q.item_hist = b->item_hist;
    q.out = b->out; q.ivl = b->ivl; q.nint = b->nint; q.nitems = (int)b->nk_tot;
    AdjArgs g;
    g.gbar = b->gbar; g.ybar = b->ybar; g.cbar = b->cbar; g.init_bar = b->ibar; g.mu_bar = b->mubar;
    hipLaunchKernelGGL(k_b_adj_nodes, dim3(b->grid_waves), dim3(64 * kBW), 0, st, b->B, b->B.hist, H, q, g);
    hipLaunchKernelGGL(k_b_adj_rows, dim3(b->grid_rows), dim3(128), 0, st, b->B, H, b->sx, b->cbar, b->zbuf, b->ybar, 0);
    hipLaunchKernelGGL(k_b_adj_rows, dim3(b->grid_rows), dim3(128), 0, st, b->B, H, b->sx, b->cbar, b->zbuf, b->ybar, 1);
    hipLaunchKernelGGL(k_b_mass_bar, dim3((H + 127) / 128), dim3(128), 0, st, b->B, H, b->mubar, b->mbar);
    KN_CHECK(hipGetLastError());
}

int check_status(kn_batch *b)
{
    int st[3];
    KN_CHECK(hipMemcpy(st, b->B.status, sizeof st, hipMemcpyDeviceToHost));
    if (st[2]) {
        kn_set_last_error("kn_batch: capacity (na_max or the H table range) exceeded");
        return KN_ERR_RANGE;
    }
    if (st[1]) {
        kn_set_last_error("kn_batch: delta_nu integrand evaluated outside a stored table (GSL_EDOM)");
        return 2001;
    }
    if (st[0]) {  // GSL error -> terminate(2001) in the reference
        kn_set_last_error("kn_batch: QAG failed (%s)", (st[0] & 1) ? "fslength" : "delta_nu integral");
        return 2001;
    }
    return 0;
}

}  // namespace

extern "C" int kn_batch_create(kn_batch **out, int nhist, const kn_batch_history *hist)
{
#ifndef KN_BATCH_ENABLE
    // Switched off on gfx950: a single-history solve + adjoint did not complete
    // on the device (no fault, no progress), and the cause is not established.
    // The kernels run only in the CPU SIMT emulation build (KN_BATCH_ENABLE).
    (void)out;
    (void)nhist;
    (void)hist;
    kn_set_last_error("kn_batch is switched off on the GPU: device run did not complete, cause not established");
    return KN_ERR_DISABLED;
#endif
    if (nhist < 1 || !hist) return KN_ERR_ARG;
    upload_consts();
    kn_batch *b = new kn_batch();
    b->nhist = nhist;
    b->hist.assign(hist, hist + nhist);
    b->hd.resize(nhist);
    long long ok = 0, oy = 0, os = 0, ofs = 0, ob = 0, oc = 0;
    std::vector<std::vector<double>> coefs(nhist);
    std::vector<double> hx0(nhist);
    std::vector<int> hnp(nhist);
    for (int h = 0; h < nhist; h++) {
        const kn_batch_history &s = hist[h];
        if (s.nk < 1 || s.na < 1 || s.na_max < s.na || !s.hubble) {
            delete b;
            return KN_ERR_ARG;
        }
        HistDev &d = b->hd[h];
        memset(&d, 0, sizeof d);
        d.nk = s.nk;
        d.na = s.na;
        d.na_max = s.na_max;
        d.ic_only = s.ic_only || !(s.mnu > 0);  // delta_tot_table.c:549: massless -> IC piece only
        d.off_k = ok;
        d.off_y = oy;
        d.off_s = os;
        d.off_fs = ofs;
        d.off_b = ob;
        d.x0 = log(s.time_transfer);
        d.xa = log(s.a);
        d.light = s.light;
        d.prefac = s.delta_nu_prefac;
        d.mnubykT = s.mnu / s.kBtnu;
        d.relerr = s.relerr;
        d.vcrit = s.vcrit;
        d.kBtnu = s.kBtnu;
        d.deriv_prefac = s.time_transfer * (s.hubble(s.time_transfer, s.hubble_ctx) / s.light) * s.time_transfer;
        d.J.qc = s.qc;
        d.J.nufrac_low = s.nufrac_low;
        d.J.en[0] = 1.0;
        for (int n = 1; n < 20; n++) d.J.en[n] = exp(-n * s.qc);
        build_htable(s, coefs[h], hx0[h], hnp[h]);
        ok += s.nk;
        oy += (long long)s.nk * s.na_max;
        os += s.na_max;
        ofs += (long long)kFsPerStep * s.na_max;
        ob += 4LL * s.na_max + 1;
        oc += (long long)coefs[h].size();
    }
    b->nk_tot = ok;
    b->y_tot = oy;
    b->s_tot = os;
    b->fs_tot = ofs;
    b->b_tot = ob;
    b->hcoef = dalloc<double>(oc);
    {
        long long off = 0;
        for (int h = 0; h < nhist; h++) {
            KN_CHECK(hipMemcpy(b->hcoef + off, coefs[h].data(), sizeof(double) * coefs[h].size(), hipMemcpyHostToDevice));
            b->hd[h].H.coef = b->hcoef + off;
            b->hd[h].H.x0 = hx0[h];
            b->hd[h].H.inv_w = 1.0 / kHWidth;
            b->hd[h].H.npanel = hnp[h];
            off += (long long)coefs[h].size();
        }
    }
    b->B.hist = dalloc<HistDev>(nhist);
    KN_CHECK(hipMemcpy(b->B.hist, b->hd.data(), sizeof(HistDev) * nhist, hipMemcpyHostToDevice));
    b->B.prefix = dalloc<long long>((size_t)W_NKIND * (nhist + 1));
    b->B.ctr = dalloc<int>(4);
    b->B.fs_fallback = dalloc<int>(ofs);
    b->B.status = dalloc<int>(4);
    KN_CHECK(hipMemset(b->B.status, 0, 4 * sizeof(int)));
    b->fx = dalloc<double>(ofs);
    b->fy = dalloc<double>(ofs);
    b->fc = dalloc<double>(ofs);
    b->sx = dalloc<double>(os);
    b->Y = dalloc<double>(oy);
    b->C = dalloc<double>(oy);
    b->ybar = dalloc<double>(oy);
    b->cbar = dalloc<double>(oy);
    b->zbuf = dalloc<double>(oy);
    b->wavenum = dalloc<double>(ok);
    b->init = dalloc<double>(ok);
    b->out = dalloc<double>(ok);
    b->gbar = dalloc<double>(ok);
    b->ibar = dalloc<double>(ok);
    b->mubar = dalloc<double>(ok);
    b->mbar = dalloc<double>(nhist);
    b->nint = dalloc<int>(ok);
    b->ivl = dalloc<double>((size_t)ok * kLimit * 2);
    b->bucket = dalloc<int>(ob);
    // work order: most expensive items first (cost ~ stored steps), history of each item
    std::vector<int> order(ok), ih(ok);
    for (int h = 0; h < nhist; h++)
        for (int k = 0; k < hist[h].nk; k++) ih[b->hd[h].off_k + k] = h;
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return hist[ih[x]].na_max > hist[ih[y]].na_max; });
    b->order = dalloc<int>(ok);
    b->item_hist = dalloc<int>(ok);
    KN_CHECK(hipMemcpy(b->order, order.data(), sizeof(int) * ok, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(b->item_hist, ih.data(), sizeof(int) * ok, hipMemcpyHostToDevice));
    const int g_fs = occupancy_grid((const void *)k_b_fs, "k_b_fs", 256, 0);
    const int g_q = occupancy_grid((const void *)k_b_qag, "k_b_qag", 64 * kBW, 0);
    const int g_a = occupancy_grid((const void *)k_b_adj_nodes, "k_b_adj_nodes", 64 * kBW, 0);
    const int g_f = occupancy_grid((const void *)k_b_fs_qag, "k_b_fs_qag", 64 * kBW, 0);
    const int g_r = occupancy_grid((const void *)k_b_spline_rows, "k_b_spline_rows", 128, 0);
    if (g_fs < 0 || g_q < 0 || g_a < 0 || g_f < 0 || g_r < 0) {
        kn_batch_destroy(b);
        return KN_ERR_HIP;
    }
    b->grid_threads = g_fs * 2;
    b->grid_waves = std::min(std::min(g_q, g_a), g_f);
    b->grid_rows = g_r * 2;
    b->have_graph = 0;
    *out = b;
    return 0;
}

extern "C" void kn_batch_destroy(kn_batch *b)
{
    if (!b) return;
    if (b->have_graph) {
        KN_CHECK(hipGraphExecDestroy(b->exec));
        KN_CHECK(hipGraphDestroy(b->graph));
    }
    void *ptrs[] = {b->hcoef, b->B.hist, b->B.prefix, b->B.ctr, b->B.fs_fallback, b->B.status, b->fx, b->fy, b->fc,
                    b->sx, b->Y, b->C, b->ybar, b->cbar, b->zbuf, b->wavenum, b->init, b->out, b->gbar, b->ibar,
                    b->mubar, b->mbar, b->nint, b->ivl, b->bucket, b->order, b->item_hist};
    for (void *p : ptrs)
        if (p) KN_CHECK(hipFree(p));
    delete b;
}

// Device views of the batch's inputs and outputs (row stride of delta_tot is
// na_max of the history).  Any pointer argument may be NULL.
extern "C" int kn_batch_views(kn_batch *b, double **scalefact, double **delta_tot, double **wavenum,
                              double **delta_nu_init, double **delta_nu, double **gbar, double **dtot_bar,
                              double **init_bar, double **mnu_bar)
{
    if (scalefact) *scalefact = b->sx;
    if (delta_tot) *delta_tot = b->Y;
    if (wavenum) *wavenum = b->wavenum;
    if (delta_nu_init) *delta_nu_init = b->init;
    if (delta_nu) *delta_nu = b->out;
    if (gbar) *gbar = b->gbar;
    if (dtot_bar) *dtot_bar = b->ybar;
    if (init_bar) *init_bar = b->ibar;
    if (mnu_bar) *mnu_bar = b->mbar;
    return 0;
}

// Offsets of history h: into the per-k arrays, the delta_tot block and scalefact.
extern "C" int kn_batch_layout(const kn_batch *b, long long *off_k, long long *off_y, long long *off_s)
{
    for (int h = 0; h < b->nhist; h++) {
        if (off_k) off_k[h] = b->hd[h].off_k;
        if (off_y) off_y[h] = b->hd[h].off_y;
        if (off_s) off_s[h] = b->hd[h].off_s;
    }
    if (off_k) off_k[b->nhist] = b->nk_tot;
    if (off_y) off_y[b->nhist] = b->y_tot;
    if (off_s) off_s[b->nhist] = b->s_tot;
    return 0;
}

extern "C" int kn_batch_solve(kn_batch *b, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    launch_forward(b, st);
    KN_CHECK(hipStreamSynchronize(st));
    return check_status(b);
}

extern "C" int kn_batch_adjoint(kn_batch *b, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    launch_adjoint(b, st);
    KN_CHECK(hipStreamSynchronize(st));
    return check_status(b);
}

// Append one stored step per history (device arrays: new log a per history,
// new delta_tot row per item); na grows on the device.
extern "C" int kn_batch_append(kn_batch *b, const double *d_loga, const double *d_rows, void *stream)
{
    hipLaunchKernelGGL(k_b_append, dim3(b->nhist), dim3(64), 0, (hipStream_t)stream, b->B, b->nhist, d_loga, d_rows,
                       b->sx, b->Y);
    KN_CHECK(hipGetLastError());
    return 0;
}

// Capture [append] + forward [+ adjoint] into one graph on stream.  Append reads
// its inputs from the pointers given here at every replay.
extern "C" int kn_batch_graph_capture(kn_batch *b, int with_adjoint, const double *d_loga, const double *d_rows,
                                      void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    if (!st) return KN_ERR_ARG;  // the null stream cannot be captured
    if (b->have_graph) {
        KN_CHECK(hipGraphExecDestroy(b->exec));
        KN_CHECK(hipGraphDestroy(b->graph));
        b->have_graph = 0;
    }
    KN_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    if (d_loga && d_rows)
        hipLaunchKernelGGL(k_b_append, dim3(b->nhist), dim3(64), 0, st, b->B, b->nhist, d_loga, d_rows, b->sx, b->Y);
    launch_forward(b, st);
    if (with_adjoint) launch_adjoint(b, st);
    KN_CHECK(hipStreamEndCapture(st, &b->graph));
    KN_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->have_graph = 1;
    return 0;
}

extern "C" int kn_batch_graph_launch(kn_batch *b, void *stream)
{
    if (!b->have_graph) return KN_ERR_ARG;
    KN_CHECK(hipGraphLaunch(b->exec, (hipStream_t)stream));
    return 0;
}

extern "C" int kn_batch_status(kn_batch *b) { return check_status(b); }

// New neutrino mass for history h (qc follows it for a hybrid history).  Only
// the mass fields of the device descriptor are written: na and log a, which
// kn_batch_append advances on the device, are left alone.
extern "C" int kn_batch_set_mass(kn_batch *b, int h, double mnu)
{
    if (h < 0 || h >= b->nhist || !(mnu > 0)) return KN_ERR_ARG;
    HistDev &d = b->hd[h];
    const kn_batch_history &s = b->hist[h];
    d.ic_only = s.ic_only;
    d.mnubykT = mnu / s.kBtnu;
    if (s.qc > 0) d.J.qc = s.vcrit * d.mnubykT;
    for (int n = 1; n < 20; n++) d.J.en[n] = exp(-n * d.J.qc);
    b->hist[h].mnu = mnu;
    HistDev *dd = b->B.hist + h;
    KN_CHECK(hipMemcpy((char *)dd + offsetof(HistDev, ic_only), &d.ic_only, sizeof(int), hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy((char *)dd + offsetof(HistDev, mnubykT), &d.mnubykT, sizeof(double), hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy((char *)dd + offsetof(HistDev, J), &d.J, sizeof(JPar), hipMemcpyHostToDevice));
    return 0;
}

// Launch-shape tuning: time the forward at a few persistent-grid sizes and keep
// the fastest (the work split itself is fixed by the algorithm).
extern "C" int kn_batch_autotune(kn_batch *b, void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    const int base_w = b->grid_waves, base_r = b->grid_rows;
    const int cand[][2] = {{1, 1}, {2, 1}, {1, 2}, {2, 2}};
    float best = 1e30f;
    int bw = base_w, br = base_r;
    hipEvent_t e0, e1;
    KN_CHECK(hipEventCreate(&e0));
    KN_CHECK(hipEventCreate(&e1));
    for (auto &c : cand) {
        b->grid_waves = base_w * c[0];
        b->grid_rows = base_r * c[1];
        launch_forward(b, st);
        KN_CHECK(hipEventRecord(e0, st));
        launch_forward(b, st);
        KN_CHECK(hipEventRecord(e1, st));
        KN_CHECK(hipEventSynchronize(e1));
        float ms = 0;
        KN_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bw = b->grid_waves;
            br = b->grid_rows;
        }
    }
    KN_CHECK(hipEventDestroy(e0));
    KN_CHECK(hipEventDestroy(e1));
    b->grid_waves = bw;
    b->grid_rows = br;
    return check_status(b);
}

// ---------------------------------------------------------------- plain adjoint

// Adjoint of get_delta_nu for one table and one species (additive entry
// point): gbar[nk] is dL/d delta_nu; outputs dL/d delta_tot[k][0..ia) (row
// stride ia), dL/d delta_nu_init[k] and dL/d mnu.  delta_nu (optional)
// receives the forward result the gradient belongs to.  The QAG subdivision is
// that of this forward evaluation.
// The parameters get_delta_nu (delta_tot_table.c:507-555) derives from a table,
// a and mnu, as one batch history (na = na_max = ia; hubble = kn_hubble_ctx).
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
    hh.relerr = 1e-6;
    {
        const double partnu = particle_nu_fraction(&d->omnu->hybnu, a, 0);
        if (partnu > 0) {
            if (1 - partnu < 1e-3) hh.ic_only = 1;
            hh.qc = d->omnu->hybnu.vcrit * (mnu / d->omnu->kBtnu);
            hh.vcrit = d->omnu->hybnu.vcrit;
            hh.relerr /= (1. + 1e-5 - partnu);
        }
        if (!(d->ia > 1 && mnu > 0)) hh.ic_only = 1;
    }
    hh.hubble = kn_hubble_ctx;
    *out = hh;
    return 0;
}

// kn_get_delta_nu_adjoint computed through a one-history batch (same arguments).
extern "C" int kn_batch_get_delta_nu_adjoint(const _delta_tot_table *d, double a, const double *wavenum, double mnu,
                                             const double *gbar, double *delta_nu, double *dtot_bar, double *init_bar,
                                             double *mnu_bar)
{
    kn_batch_history hh;
    int rc = kn_batch_history_from_table(d, a, mnu, &hh);
    if (rc) return rc;
    kn_batch *b = nullptr;
    rc = kn_batch_create(&b, 1, &hh);
    if (rc) return rc;
    const int nk = d->nk, na = d->ia;
    std::vector<double> Yh((size_t)nk * na);
    for (int k = 0; k < nk; k++) memcpy(&Yh[(size_t)k * na], d->delta_tot[k], sizeof(double) * na);
    KN_CHECK(hipMemcpy(b->Y, Yh.data(), sizeof(double) * Yh.size(), hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(b->sx, d->scalefact, sizeof(double) * na, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(b->wavenum, wavenum, sizeof(double) * nk, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(b->init, d->delta_nu_init, sizeof(double) * nk, hipMemcpyHostToDevice));
    KN_CHECK(hipMemcpy(b->gbar, gbar, sizeof(double) * nk, hipMemcpyHostToDevice));
    rc = kn_batch_solve(b, nullptr);
    if (!rc) rc = kn_batch_adjoint(b, nullptr);
    if (!rc) {
        if (delta_nu) KN_CHECK(hipMemcpy(delta_nu, b->out, sizeof(double) * nk, hipMemcpyDeviceToHost));
        if (dtot_bar) KN_CHECK(hipMemcpy(dtot_bar, b->ybar, sizeof(double) * (size_t)nk * na, hipMemcpyDeviceToHost));
        if (init_bar) KN_CHECK(hipMemcpy(init_bar, b->ibar, sizeof(double) * nk, hipMemcpyDeviceToHost));
        if (mnu_bar) KN_CHECK(hipMemcpy(mnu_bar, b->mbar, sizeof(double), hipMemcpyDeviceToHost));
    }
    kn_batch_destroy(b);
    return rc;
}
