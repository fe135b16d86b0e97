// bin_kernels.h -- device kernels of the k-shell binning sweep (a).
//
// Included by powerspectrum.hip.  The kernels are templates over the launch
// shape so the shape can be tuned without touching the arithmetic; the
// arithmetic itself follows powerspectrum.c:56-89 (see powerspectrum.hip).
#ifndef KN_BIN_KERNELS_H
#define KN_BIN_KERNELS_H

#include <hip/hip_runtime.h>
#include <math.h>
#include <vector>

#include "kn_mem.h"

namespace kn_bin {

// KVAL of powerspectrum.c:27
__host__ __device__ inline int kval(int n, int dims) { return n <= dims / 2 ? n : n - dims; }

// Per-bin smallest |k|^2 reaching the bin under the reference expression
// floor(binsperunit * log(sqrt(k2))) (powerspectrum.c:40, :67).  Returns 0,
// or the first k2 where the expression is not monotone / exceeds nrbins-1.
// thr holds nrbins + 1 ints.
inline long long build_thresholds(int dims, int nrbins, int *thr)
{
    const double binsperunit = (nrbins - 1) / log(sqrt(3) * dims / 2.0);
    const long long kmax2 = 3LL * (dims / 2) * (dims / 2);
    for (int b = 0; b <= nrbins; b++) thr[b] = 0x7fffffff;
    thr[0] = 0;
    int prev = 0;
    for (long long k2 = 1; k2 <= kmax2; k2++) {
        const int b = (int)floor(binsperunit * log(sqrt((double)k2)));
        if (b < prev || b >= nrbins) return k2;
        while (prev < b) thr[++prev] = (int)k2;
    }
    return 0;
}

inline long long build_thresholds(int dims, int nrbins, std::vector<int> &thr)
{
    thr.resize(nrbins + 1);
    return build_thresholds(dims, nrbins, thr.data());
}

// Per-axis inverse CiC window by storage index (powerspectrum.c:8-12); iw holds dims doubles.
inline void build_invwindow(int dims, double *iw)
{
    for (int n = 0; n < dims; n++) {
        const int kx = kval(n, dims);
        iw[n] = kx ? M_PI * kx / (dims * sin(M_PI * kx / (double)dims)) : 1.0;
    }
}

inline void build_invwindow(int dims, std::vector<double> &iw)
{
    iw.resize(dims);
    build_invwindow(dims, iw.data());
}

__device__ __forceinline__ int find_bin(int k2, float half_bpu, const int *__restrict__ thr, int nrbins)
{
    int b = (int)(__logf((float)k2) * half_bpu);
    b = b < 0 ? 0 : (b > nrbins - 1 ? nrbins - 1 : b);
    while (b > 0 && k2 < thr[b]) b--;
    while (b < nrbins - 1 && k2 >= thr[b + 1]) b++;
    return b;
}

struct Pos {
    int i, j, z;  // slab offset, row, z of a flat mode offset
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

// multiplicity * |F|^2 * ((iwx iwy iwz)^2)^2 in the reference's rounding order
__device__ __forceinline__ double mode_weight(double2 f, double iwxy, double iwz, int kz, int half)
{
    const double w = __dmul_rn(iwxy, iwz);
    const double w2 = __dmul_rn(w, w);
    const double w4 = __dmul_rn(w2, w2);
    const double p2 = __dadd_rn(__dmul_rn(f.x, f.x), __dmul_rn(f.y, f.y));
    double v = __dmul_rn(p2, w4);
    if (kz != 0 && kz != half) v *= 2.0;
    return v;
}

struct BinArgs {
    const double2 *F;
    int dims, startslab, nrbins;
    long long rows;
    const int *thr;
    const double *iw;
    float half_bpu;
    double *partial;  // gridDim.x * nrbins
};

// A step of n modes along the flattened [slab][row][z] order, split so that
// advancing a position costs two compares whatever n is.
struct Step {
    int di, dj, dz;
};

__host__ __device__ inline Step step_of(long long n, int rowlen, int dims)
{
    const long long r = n / rowlen;
    Step s;
    s.dz = (int)(n % rowlen);
    s.dj = (int)(r % dims);
    s.di = (int)(r / dims);
    return s;
}

__device__ __forceinline__ void advance(Pos &p, const Step &s, int rowlen, int dims)
{
    p.z += s.dz;
    p.j += s.dj;
    p.i += s.di;
    if (p.z >= rowlen) {
        p.z -= rowlen;
        p.j++;
    }
    if (p.j >= dims) {
        p.j -= dims;
        p.i++;
    }
}

// Where the sweep keeps its tables: the bin thresholds, the per-axis inverse
// window and the bin sums all in LDS (the fast form); the window read from
// global memory (through the caches) when LDS cannot hold it next to the
// others; or everything in global memory, the bin sums then added with
// global atomics into the block's own partial row (any number of bins).
enum TablePlace { kTablesLds = 0, kWindowGlobal = 1, kTablesGlobal = 2 };

constexpr int kSweepWaves = 8;  // wavefronts per block
constexpr int kSweepRun = 4;    // consecutive modes per lane (one staged chunk: 64 * kSweepRun modes per wave)

// LDS bytes of the sweep: the staging buffers, then the tables the placement keeps there.
__host__ __device__ constexpr size_t sweep_smem_bytes(int place, int dims, int nrbins)
{
    return sizeof(double2) * (size_t)kSweepWaves * 64 * kSweepRun +
           (place == kTablesGlobal
                ? 0
                : sizeof(double) * (size_t)nrbins + sizeof(int) * (size_t)(nrbins + 1 + ((nrbins + 1) & 1)) +
                      (place == kTablesLds ? sizeof(double) * (size_t)dims : 0));
}

// Slot of mode m (0 <= m < 64 * kSweepRun) of a wavefront's staging buffer.
// Lanes write consecutive modes (m = 64u + l) and read back runs (m = 4l + r).
// A 16-byte access pass serves 8 or 16 lanes; for the reads the lanes of a
// pass differ in lane bits l1..l3, so (l1 ^ l3, l2) is XOR-ed into the slot's
// two low bits (one-to-one for both pass widths); the writes of a pass keep
// distinct low bits since the XOR term is constant across each aligned group
// of 8 lanes.  The slot stays inside m's aligned group of 4.
__device__ __forceinline__ int stage_slot(int m)
{
    return m ^ ((((m >> 3) ^ (m >> 5)) & 1) | (((m >> 4) & 1) << 1));
}

// The bin sweep (powerspectrum.c:56-89).  Block b takes a contiguous range of
// rows; wavefront w of the block takes chunks of 256 consecutive modes (w,
// w + 8, ...).  Load u of a chunk is one 1 KB contiguous dwordx4 load across
// the wave, non-temporal (each mode is read once per sweep); the next chunk's
// loads are issued before the current one is binned.  The chunk passes
// through the wave's LDS staging buffer so that lane l holds modes 4l .. 4l+3
// of it: along a row |k|^2 grows with z, so after a run's first mode its bin
// is found incrementally (the next threshold kept in a register, at most one
// step before a full search), equal-bin runs are summed in registers and
// flushed with one atomic add per run.  The block's wavefronts share one set
// of bin sums, so the order of the adds between wavefronts (and the last bit
// of a sum) can differ from run to run; the block writes its sums to its
// partial row, which k_reduce_partials adds in a fixed order.
template <int PLACE>
__global__ __launch_bounds__(64 * kSweepWaves) void k_bin_sweep(BinArgs a)
{
    constexpr int WAVES = kSweepWaves, RUN = kSweepRun;
    extern __shared__ double smem[];
    double2 *stage_all = (double2 *)smem;  // WAVES * 64 * RUN modes, 16-byte aligned
    double *hist;
    const double *iw;
    const int *thr;
    if constexpr (PLACE == kTablesGlobal) {
        hist = a.partial + (size_t)blockIdx.x * a.nrbins;
        iw = a.iw;
        thr = a.thr;
        for (int t = threadIdx.x; t < a.nrbins; t += 64 * WAVES) hist[t] = 0.0;
    } else {
        hist = (double *)(stage_all + (size_t)WAVES * 64 * RUN);
        int *thr_l = (int *)(hist + a.nrbins);
        for (int t = threadIdx.x; t < a.nrbins; t += 64 * WAVES) hist[t] = 0.0;
        for (int t = threadIdx.x; t <= a.nrbins; t += 64 * WAVES) thr_l[t] = a.thr[t];
        thr = thr_l;
        if constexpr (PLACE == kTablesLds) {
            double *iw_l = (double *)(thr_l + a.nrbins + 1 + ((a.nrbins + 1) & 1));  // 8-byte aligned
            for (int t = threadIdx.x; t < a.dims; t += 64 * WAVES) iw_l[t] = a.iw[t];
            iw = iw_l;
        } else {
            iw = a.iw;
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double2 *stage = stage_all + (size_t)wave * 64 * RUN;
    const int rowlen = a.dims / 2 + 1, half = a.dims / 2;
    const long long r0 = a.rows * blockIdx.x / gridDim.x, r1 = a.rows * (blockIdx.x + 1) / gridDim.x;
    const long long m0 = r0 * rowlen, m1 = r1 * rowlen;
    constexpr int CH = 64 * RUN;
    const long long S = (long long)CH * WAVES;
    const Step sS = step_of(S, rowlen, a.dims);
    long long c = m0 + (long long)wave * CH;
    Pos p = pos_of((long long)wave * CH + (long long)lane * RUN, r0, rowlen, a.dims);
    double2 cur[RUN];
#pragma unroll
    for (int u = 0; u < RUN; u++) {
        const long long mu = c + 64 * u + lane;
        cur[u] = mu < m1 ? kn_mem::ld<true>(a.F + mu) : make_double2(0.0, 0.0);
    }
    for (; c < m1; c += S) {
        double2 nxt[RUN];
#pragma unroll
        for (int u = 0; u < RUN; u++) {
            const long long mu = c + S + 64 * u + lane;
            nxt[u] = mu < m1 ? kn_mem::ld<true>(a.F + mu) : make_double2(0.0, 0.0);
        }
        // transpose through LDS: lane l wrote modes l, 64 + l, ...; reads 4l .. 4l + 3.
        // LDS operations of one wavefront complete in order, so the fences only
        // keep the compiler from moving the reads above the writes (and the next
        // chunk's writes above these reads).
#pragma unroll
        for (int u = 0; u < RUN; u++) stage[stage_slot(64 * u + lane)] = cur[u];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        double2 f[RUN];
#pragma unroll
        for (int r = 0; r < RUN; r++) f[r] = stage[stage_slot(RUN * lane + r)];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const long long mrun = c + (long long)RUN * lane;
        Pos q = p;
        int curb = -1, b = -1, bnext = 0;
        double acc = 0.0;
        // the row's kx^2 + ky^2 and window product, at the run start and at a new row
        int kx = kval(a.startslab + q.i, a.dims), ky = kval(q.j, a.dims);
        int kxy2 = kx * kx + ky * ky;
        double iwxy = __dmul_rn(iw[a.startslab + q.i], iw[q.j]);
#pragma unroll
        for (int r = 0; r < RUN; r++) {
            if (mrun + r >= m1) break;
            if (r && ++q.z == rowlen) {
                q.z = 0;
                if (++q.j == a.dims) {
                    q.j = 0;
                    q.i++;
                }
                kx = kval(a.startslab + q.i, a.dims);
                ky = kval(q.j, a.dims);
                kxy2 = kx * kx + ky * ky;
                iwxy = __dmul_rn(iw[a.startslab + q.i], iw[q.j]);
            }
            const int k2 = kxy2 + q.z * q.z;
            if (k2 == 0) continue;
            if (b < 0 || q.z == 0) {
                b = find_bin(k2, a.half_bpu, thr, a.nrbins);
                bnext = thr[b + 1];
            } else if (k2 >= bnext) {
                const int t2 = thr[b + 2];  // b + 1 <= nrbins - 1 here, so b + 2 <= nrbins
                if (k2 < t2) {
                    b++;
                    bnext = t2;
                } else {
                    b = find_bin(k2, a.half_bpu, thr, a.nrbins);
                    bnext = thr[b + 1];
                }
            }
            const double v = mode_weight(f[r], iwxy, iw[q.z], q.z, half);
            if (b != curb) {
                if (curb >= 0) atomicAdd(&hist[curb], acc);
                curb = b;
                acc = v;
            } else {
                acc += v;
            }
        }
        if (curb >= 0) atomicAdd(&hist[curb], acc);
#pragma unroll
        for (int u = 0; u < RUN; u++) cur[u] = nxt[u];
        advance(p, sS, rowlen, a.dims);
    }
    if constexpr (PLACE != kTablesGlobal) {
        __syncthreads();
        for (int t = threadIdx.x; t < a.nrbins; t += 64 * WAVES)
            a.partial[(size_t)blockIdx.x * a.nrbins + t] = hist[t];
    }
}

}  // namespace kn_bin

#endif
