// rescale_kernels.h -- device kernels of the per-mode rescale sweep (c).
//
// Included by rescale.hip.  Arithmetic follows interface_gadget.c:163-188 and
// delta_pow.c:19-37 (see rescale.hip); the templates only vary the launch
// shape and the lane-to-mode mapping.
#ifndef KN_RESCALE_KERNELS_H
#define KN_RESCALE_KERNELS_H

#include <hip/hip_runtime.h>
#include <math.h>

#include "kn_mem.h"

namespace kn_res {

__host__ __device__ inline int kval(int n, int dims) { return n <= dims / 2 ? n : n - dims; }

struct ResArgs {
    double2 *F;
    int dims, startslab, nbins, nbuckets;
    long long rows;
    double boxsize, norm;
    const double *logkk, *ratio;
    int *nan_flag;
    // Green's function of the PM potential (GREEN sweeps): per axis index t,
    // gw[t] = exp(-kt^2 asmth2) W(t)^-4 (dims doubles; kt the integer
    // wavenumber, W the CiC window, asmth2 = (2 pi r_s / BoxSize)^2)
    const double *gw;
};

constexpr int kBucketsPerBin = 4;

__host__ __device__ inline size_t smem_bytes(int nbins, int nbuckets, int green_dims = 0)
{
    return sizeof(double) * (2 * (size_t)nbins + (size_t)green_dims) +
           sizeof(int) * (size_t)(nbuckets + 1 + ((nbuckets + 1) & 1));
}

struct Tab {
    const double *lk, *rt;
    const int *bucket;
    int nbins, nbuckets;
    double x0, x1, inv_bw, lower_clamp, norm, log_kunit;
};

// Ratio table and bucket index into LDS (all threads of the block).
__device__ __forceinline__ Tab load_tab(const ResArgs &a, int nthreads)
{
    extern __shared__ double sm[];
    double *lk = sm, *rt = sm + a.nbins;
    int *bucket = (int *)(rt + a.nbins);
    for (int t = threadIdx.x; t < a.nbins; t += nthreads) {
        lk[t] = a.logkk[t];
        rt[t] = a.ratio[t];
    }
    __syncthreads();
    Tab T;
    T.lk = lk;
    T.rt = rt;
    T.bucket = bucket;
    T.nbins = a.nbins;
    T.nbuckets = a.nbuckets;
    T.x0 = lk[0];
    T.x1 = lk[a.nbins - 1];
    T.inv_bw = a.nbins > 1 ? a.nbuckets / (T.x1 - T.x0) : 0.0;
    T.lower_clamp = T.x0 - log(2.0);
    T.norm = a.norm;
    T.log_kunit = log(2 * M_PI / a.boxsize);
    // bucket[q]: the GSL segment index (largest i <= n-2 with lk[i] <= x) at the bucket's left edge
    for (int q = threadIdx.x; q <= a.nbuckets; q += nthreads) {
        const double edge = T.x0 + q / T.inv_bw;
        int lo = 0, hi = a.nbins - 1;
        while (hi > lo + 1) {
            const int mid = (lo + hi) >> 1;
            if (lk[mid] > edge) hi = mid;
            else lo = mid;
        }
        bucket[q] = lo;
    }
    __syncthreads();
    return T;
}

// 1 + get_dnudcdm_powerspec(log(|k|)) for k2 > 0, with |k| = sqrt(k2) 2 pi / BoxSize.
// log|k| is evaluated as 0.5 log(k2) + log(2 pi / BoxSize): the same value as
// the reference's log(sqrt(k2)*2*M_PI/BoxSize) to about one ulp, with one
// transcendental per mode instead of sqrt, two divides and a log.
__device__ __forceinline__ double smth_of(int k2i, const Tab &T)
{
    double kk = 0.5 * log((double)k2i) + T.log_kunit;
    if (kk < T.x0 && kk > T.lower_clamp) kk = T.x0;
    if (kk < T.lower_clamp) kk = T.x0;
    if (kk > T.x1) kk = T.x1;
    double interp;
    if (T.nbins == 1) {
        interp = T.rt[0];
    } else {
        int q = (int)((kk - T.x0) * T.inv_bw);
        q = q < 0 ? 0 : (q > T.nbuckets ? T.nbuckets : q);
        int i = T.bucket[q];
        while (i < T.nbins - 2 && T.lk[i + 1] <= kk) i++;
        const double xl = T.lk[i], h = T.lk[i + 1] - xl;
        interp = h > 0.0 ? T.rt[i] + (kk - xl) / h * (T.rt[i + 1] - T.rt[i]) : T.rt[i];
    }
    return 1 + T.norm * interp;
}

struct Pos {
    int i, j, z;
};

__device__ __forceinline__ void advance(Pos &p, int step, int rowlen, int dims)
{
    p.z += step;
    while (p.z >= rowlen) {
        p.z -= rowlen;
        if (++p.j == dims) {
            p.j = 0;
            p.i++;
        }
    }
}

__device__ __forceinline__ Pos pos_of(long long off, long long row0, int rowlen, int dims)
{
    Pos p;
    const long long rl = row0 + off / rowlen;
    p.z = (int)(off % rowlen);
    p.j = (int)(rl % dims);
    p.i = (int)(rl / dims);
    return p;
}

// The PM Green's function with CiC deconvolution at k2 > 0, as the potential
// pass of a PM step applies it after the neutrino rescale (Gadget-2's
// pmforce_periodic, where the hook of 0002-Add-hooks-to-kspace-neutrino-code.patch:116-125
// calls add_nu_power_to_rhogrid): -exp(-k2 asmth2) / k2 times ff^4, ff the
// product of the three per-axis inverse windows.  exp(-k2 asmth2) factors
// over the axes, so with gw (ResArgs) it is -gw[x] gw[y] gw[z] / k2: no
// exponential per mode, and 1/k2 from the hardware reciprocal refined by one
// Newton step (a few units in the last place from the direct form).
__device__ __forceinline__ double green_of(int k2, double gx, double gy, double gz)
{
    const double d = (double)k2;
    const double r = __builtin_amdgcn_rcp(d);
    const double inv = fma(r, fma(-d, r, 1.0), r);
    return -((gx * gy) * gz) * inv;
}

// The rescale sweep: each block a contiguous row range, lane l of the block
// takes modes l, l + B, ... (coalesced), the next UNROLL modes loaded before
// the current ones are rescaled and stored; loads and stores carry the
// non-temporal hint (each mode is read and written once per sweep).
// GREEN: the same sweep also applies the Green's function (F s_nu, then times
// G, the two roundings of the two passes it replaces) and zeroes the k = 0
// mode; with nbins 0 it applies the Green's function alone.
template <int WAVES, int UNROLL, bool GREEN = false>
__global__ __launch_bounds__(64 * WAVES) void k_rescale_rows_pf(ResArgs a)
{
    constexpr int B = 64 * WAVES;
    Tab T;
    const double *gw = nullptr;
    if (!GREEN || a.nbins > 0) T = load_tab(a, B);
    if constexpr (GREEN) {
        extern __shared__ double sm[];
        int *after_buckets = (int *)(sm + 2 * (size_t)a.nbins) + a.nbuckets + 1 + ((a.nbuckets + 1) & 1);
        double *gw_l = (double *)after_buckets;  // 8-byte aligned
        for (int t = threadIdx.x; t < a.dims; t += B) gw_l[t] = a.gw[t];
        __syncthreads();
        gw = gw_l;
    }
    const int rowlen = a.dims / 2 + 1;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m1 = r1 * rowlen;
    const long long S = (long long)B * UNROLL;
    Pos p = pos_of(threadIdx.x, r0, rowlen, a.dims);
    bool saw_nan = false;
    long long m = r0 * rowlen + threadIdx.x;
    double2 cur[UNROLL];
#pragma unroll
    for (int u = 0; u < UNROLL; u++) {
        const long long mu = m + (long long)B * u;
        cur[u] = mu < m1 ? kn_mem::ld<true>(a.F + mu) : make_double2(0.0, 0.0);
    }
    for (; m < m1; m += S) {
        double2 nxt[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const long long mu = m + S + (long long)B * u;
            nxt[u] = mu < m1 ? kn_mem::ld<true>(a.F + mu) : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) {
            const long long mu = m + (long long)B * u;
            const Pos q = p;
            advance(p, B, rowlen, a.dims);
            if (mu >= m1) continue;
            const int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
            const int k2 = kx * kx + ky * ky + q.z * q.z;
            if (k2 <= 0) {
                if (GREEN) kn_mem::st<true>(a.F + mu, make_double2(0.0, 0.0));
                continue;
            }
            double2 f = cur[u];
            if (!GREEN || a.nbins > 0) {
                const double s = smth_of(k2, T);
                saw_nan |= isnan(s);
                f.x *= s;
                f.y *= s;
            }
            if constexpr (GREEN) {
                const double g = green_of(k2, gw[a.startslab + q.i], gw[q.j], gw[q.z]);
                f.x *= g;
                f.y *= g;
            }
            kn_mem::st<true>(a.F + mu, f);
        }
#pragma unroll
        for (int u = 0; u < UNROLL; u++) cur[u] = nxt[u];
    }
    if (saw_nan) atomicOr(a.nan_flag, 1);
}

}  // namespace kn_res

#endif
